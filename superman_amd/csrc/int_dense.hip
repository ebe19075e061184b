// int_dense.hip -- single-matrix exact integer Ryser kernel (gfx950).  The
// per-chunk walk (and the arithmetic it uses) is in walk_int.h; this file
// holds the grid-strided task loop over one matrix's chunks, the block sum
// mod 2^128 and the final reduction kernel.
#include "walk_int.h"

#ifndef SPM_INT_ONLY_N
#define SPM_INT_ONLY_N 0
#endif

namespace spm {

namespace ints {

// Orders 36..40 ask the register allocator for 4 waves per SIMD: their
// natural allocation (131-137 VGPRs) leaves 3, and at 128 the spills fall
// outside the step loops.
template <int N>
constexpr int kIntWaves = (N >= 36 && N <= 40) ? 4 : 1;

template <int N>
__global__ void __launch_bounds__(kDenseBlock) __attribute__((amdgpu_waves_per_eu(kIntWaves<N>)))
int_dense_kernel(const int* __restrict__ M,        // [2][N][N]: +2A then -2A, column-major
                 const int* __restrict__ y0,       // y(empty set), N values
                 unsigned long long* __restrict__ partial,  // [gridDim.x][2]: T mod 2^128 (lo, hi)
                 uint64_t task_begin, uint64_t task_end, uint64_t num_chunks, int L, int accumulate)
{
    __shared__ unsigned long long red[2 * (kDenseBlock / kWave)];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = (uint64_t)blockIdx.x * (kDenseBlock / kWave) + (threadIdx.x / kWave);
    const uint64_t nwave = (uint64_t)gridDim.x * (kDenseBlock / kWave);
    const bool idle = (uint64_t)lane >= num_chunks;  // see dense_c128.hip
    unsigned __int128 acc = 0;

    for (uint64_t task = task_begin + wave; task < task_end; task += nwave) {
        const uint64_t h = task * kWave + lane;
        walk_int<N>(M, y0, h, L, acc);
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
