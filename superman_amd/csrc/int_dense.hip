// int_dense.hip -- exact Gray-code Ryser for integer matrices on gfx950.
//
// The reference computes every permanent in FP64, sparse ones through
// SparseDefinitions::* with a zero-count product skip
// (Algorithms/GPU/Sparse/SparseKernelDefinitions.cuh:34-509); on integer
// inputs its result is only as exact as FP64 rounding allows (a dense 0/1
// 24 x 24 already loses the last digits).  Here integer matrices run exactly:
//
//   y_i(S) = 2 x_i(S) = 2 a_{i,n-1} - sum_j a_ij + 2 sum_{j in S} a_ij   (int32)
//   T      = sum_k (-1)^k prod_i y_i(G(k))                   (mod 2^128)
//   perm   = (-1)^(n-1) T / 2^(n-1)
//
// The host guarantees |T| < 2^127 (Bregman-Minc bound for 0/1 matrices, the
// row-sum product otherwise), so the mod-2^128 sum is T itself.
//
// Walk: as dense_f64.hip -- lane chunks of 2^L aligned Gray steps, 2-step
// blocks, flip columns scalar-loaded from +2A / -2A planes (a column is N
// SGPRs here), only column L-1 flips in a per-lane direction.  Per step:
// N v_add_u32, an all-rows-nonzero test, and -- only on lanes whose term is
// nonzero -- the 128-bit product.  Sparse 0/1 matrices have a zero row in
// almost every subset, so most steps cost the adds and the test alone.
#include "spm_device.h"
#include "spm_kernels.h"

#include <type_traits>

#ifndef SPM_INT_ONLY_N
#define SPM_INT_ONLY_N 0
#endif

namespace spm {

namespace ints {

constexpr uint32_t kFoldSteps = 1;  // (no folding needed: sums are exact)

typedef __attribute__((address_space(4))) const int const_i32;

__device__ __forceinline__ int opaque_zero()
{
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    return z;
}

__device__ __forceinline__ const const_i32* step_ptr(const int* p)
{
    return (const const_i32*)(p + opaque_zero());
}

__device__ __forceinline__ void fence() { __builtin_amdgcn_sched_barrier(0); }

template <int N>
struct Column {
    int v[N];
};

template <int N>
__device__ __forceinline__ void load_column(Column<N>& c, const int* p)
{
    const const_i32* __restrict__ q = step_ptr(p);
#pragma unroll
    for (int r = 0; r < N; ++r) c.v[r] = q[r];
}

template <int N>
__device__ __forceinline__ void add_column(int (&y)[N], const Column<N>& c)
{
#pragma unroll
    for (int r = 0; r < N; ++r) y[r] += c.v[r];
}

// y += s * column with a per-lane s = +-1 (as a sign mask m = 0 or -1:
// s*v = (v ^ m) - m).
template <int N>
__device__ __forceinline__ void add_column_signed(int (&y)[N], const Column<N>& c, int m)
{
#pragma unroll
    for (int r = 0; r < N; ++r) y[r] += (c.v[r] ^ m) - m;
}

// acc +/-= prod_r y_r (mod 2^128), on the lanes where every y_r != 0.
template <int N>
__device__ __forceinline__ void add_term(const int (&y)[N], bool negative, unsigned __int128& acc)
{
    bool nz = true;
#pragma unroll
    for (int r = 0; r < N; ++r) nz = nz && (y[r] != 0);
    if (nz) {
        // Pairs of rows multiply exactly in 64 bits (|y| < 2^31); the
        // running product is kept mod 2^128.
        unsigned __int128 p = (unsigned __int128)(__int128)y[0];
#pragma unroll
        for (int r = 1; r + 1 < N; r += 2) {
            const long long q = (long long)y[r] * (long long)y[r + 1];
            p *= (unsigned __int128)(__int128)q;
        }
        if constexpr (N % 2 == 0) p *= (unsigned __int128)(__int128)y[N - 1];
        acc = negative ? acc - p : acc + p;
    }
}

template <int N>
__global__ void __launch_bounds__(kDenseBlock)
int_dense_kernel(const int* __restrict__ M,        // [2][N][N]: +2A then -2A, column-major
                 const int* __restrict__ y0,       // y(empty set), N values
                 unsigned long long* __restrict__ partial,  // [gridDim.x][2]: T mod 2^128 (lo, hi)
                 uint64_t task_begin, uint64_t task_end, uint64_t num_chunks, int L, int accumulate)
{
    __shared__ unsigned long long red[2 * (kDenseBlock / kWave)];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = (uint64_t)blockIdx.x * (kDenseBlock / kWave) + (threadIdx.x / kWave);
    const uint64_t nwave = (uint64_t)gridDim.x * (kDenseBlock / kWave);
    const int* Mneg = M + N * N;
    const bool idle = (uint64_t)lane >= num_chunks;  // see dense_c128.hip
    unsigned __int128 acc = 0;

    for (uint64_t task = task_begin + wave; task < task_end; task += nwave) {
        const uint64_t h = task * kWave + lane;
        int y[N];
        {
            const const_i32* __restrict__ ys = step_ptr(y0);
#pragma unroll
            for (int r = 0; r < N; ++r) y[r] = ys[r];
        }
        const uint64_t k0 = h << L;
        const uint64_t g0 = k0 ^ (k0 >> 1);
        for (int j = (L > 0 ? L - 1 : 0); j < N - 1; ++j) {
            const int b = -(int)((g0 >> j) & 1u);  // 0 or -1: mask of "column j in the start set"
            const const_i32* __restrict__ col = step_ptr(M + j * N);
#pragma unroll
            for (int r = 0; r < N; ++r) y[r] += col[r] & b;
        }
        const int lane_mask = -(int)(h & 1u);  // direction of column L-1: -1 = remove

        if constexpr (N == 1) {
            add_term<N>(y, false, acc);
        } else if constexpr (N == 2) {  // L = 1
            add_term<N>(y, false, acc);
            Column<N> c;
            load_column(c, M);
            add_column_signed(y, c, lane_mask);
            add_term<N>(y, true, acc);
        } else {
            const uint32_t nblk = 1u << (L - 1);
            const uint32_t half = nblk >> 1;
            auto a_col = [&](uint32_t m) __attribute__((always_inline)) {
                const int tz = __builtin_ctz(m);
                return (((m >> (tz + 1)) & 1u) ? Mneg : M) + (tz + 1) * N;
            };
            Column<N> ca, cb;
            add_term<N>(y, false, acc);  // k = 0
            load_column(cb, M);
            add_column(y, cb);           // k = 1: +column 0
            load_column(ca, a_col(1));
            fence();
            add_term<N>(y, true, acc);
            fence();
            auto block = [&](uint32_t m, auto SPECIAL) __attribute__((always_inline)) {
                if constexpr (decltype(SPECIAL)::value) add_column_signed(y, ca, lane_mask);
                else add_column(y, ca);
                load_column(cb, (m & 1u) ? Mneg : M);
                fence();
                add_term<N>(y, false, acc);
                fence();
                add_column(y, cb);
                load_column(ca, a_col(m + 1));  // m+1 = nblk: column L <= N-1, a valid address
                fence();
                add_term<N>(y, true, acc);
                fence();
            };
#pragma nounroll
            for (uint32_t m = 1; m < half; ++m) block(m, std::false_type{});
            block(half, std::true_type{});
#pragma nounroll
            for (uint32_t m = half + 1; m < nblk; ++m) block(m, std::false_type{});
        }
    }
    if (idle) acc = 0;

    // Block sum mod 2^128 (fixed order: wave butterfly, then waves in order).
    unsigned long long lo = (unsigned long long)acc, hi = (unsigned long long)(acc >> 64);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long olo = __shfl_xor(lo, off, kWave);
        const unsigned long long ohi = __shfl_xor(hi, off, kWave);
        const unsigned long long s = lo + olo;
        hi = hi + ohi + (s < lo ? 1ull : 0ull);
        lo = s;
    }
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[2 * w] = lo;
        red[2 * w + 1] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned __int128 t = 0;
        for (int i = 0; i < kDenseBlock / kWave; ++i)
            t += ((unsigned __int128)red[2 * i + 1] << 64) | red[2 * i];
        unsigned long long* out = partial + 2 * blockIdx.x;
        if (accumulate) t += ((unsigned __int128)out[1] << 64) | out[0];
        out[0] = (unsigned long long)t;
        out[1] = (unsigned long long)(t >> 64);
    }
}

// Sum of the per-block records mod 2^128 (one block, fixed order).
__global__ void __launch_bounds__(256)
reduce_i128_kernel(const unsigned long long* __restrict__ partial, int nrec, unsigned long long* __restrict__ out)
{
    __shared__ unsigned long long red[2 * 256];
    unsigned __int128 t = 0;
    for (int i = threadIdx.x; i < nrec; i += blockDim.x)
        t += ((unsigned __int128)partial[2 * i + 1] << 64) | partial[2 * i];
    red[2 * threadIdx.x] = (unsigned long long)t;
    red[2 * threadIdx.x + 1] = (unsigned long long)(t >> 64);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned __int128 s = 0;
        for (int i = 0; i < (int)blockDim.x; ++i) s += ((unsigned __int128)red[2 * i + 1] << 64) | red[2 * i];
        out[0] = (unsigned long long)s;
        out[1] = (unsigned long long)(s >> 64);
    }
}

template <int N>
const void* int_ptr()
{
    if constexpr (SPM_INT_ONLY_N != 0 && N != SPM_INT_ONLY_N) return nullptr;
    else return reinterpret_cast<const void*>(&int_dense_kernel<N>);
}

template <int... Is>
const void* int_lookup(int n, std::integer_sequence<int, Is...>)
{
    static const void* const tab[] = {nullptr, int_ptr<Is + 1>()...};
    return tab[n];
}

}  // namespace ints

const void* int_dense_kernel_ptr(int n)
{
    if (n < 1 || n > kMaxN) return nullptr;
    return ints::int_lookup(n, std::make_integer_sequence<int, kMaxN>{});
}

const void* reduce_i128_kernel_ptr() { return reinterpret_cast<const void*>(&ints::reduce_i128_kernel); }

}  // namespace spm
