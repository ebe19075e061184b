// walk_int.h -- exact Gray-code Ryser walk for integer matrices on gfx950:
// device code shared by the single-matrix kernel (int_dense.hip) and the
// batched kernel (batch.hip).
//
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
#ifndef SPM_WALK_INT_H
#define SPM_WALK_INT_H

#include "spm_device.h"
#include "spm_kernels.h"

#include <type_traits>

namespace spm {
namespace ints {

constexpr uint32_t kFoldSteps = 1;  // (no folding needed: sums are exact)

typedef __attribute__((address_space(4))) const int const_i32;

// The pointer passes through an empty asm that claims to rewrite it, so a
// step's scalar loads are neither hoisted out of the walk nor merged across
// steps, and the asm costs no instruction (an offset from an asm would be
// sign-extended and added in 64 bits at every load).
__device__ __forceinline__ const const_i32* step_ptr(const int* p)
{
    asm volatile("" : "+s"(p));
    return (const const_i32*)p;
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

// One lane's chunk: the 2^L Gray steps starting at k = h 2^L, added into
// acc (mod 2^128).
template <int N>
__device__ __forceinline__ void walk_int(const int* __restrict__ M,   // [2][N][N]: +2A, -2A
                                         const int* __restrict__ y0,  // y(empty set)
                                         uint64_t h, int L, unsigned __int128& acc)
{
    const int* Mneg = M + N * N;
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

}  // namespace ints
}  // namespace spm

#endif
