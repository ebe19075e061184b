// int_small.hip -- single-matrix exact Ryser for integer matrices with small
// row sums (every sum_j |a_ij| < 2^B, B = bit width of n; all 0/1 matrices),
// walked in FP64 with exact grouped products (walk_intfp.h).  Same task
// loop, block sum mod 2^128 and per-block records as int_dense.hip, whose
// final reduction kernel (reduce_i128_kernel) it shares.
#include "walk_intfp.h"

#ifndef SPM_INTFP_ONLY_N
#define SPM_INTFP_ONLY_N 0
#endif

namespace spm {

// Orders 37..40 ask the register allocator for 4 waves per SIMD: their
// natural allocation (128-133 VGPRs) leaves 3, and at 128 the spills fall in
// the task loop (once per chunk), not in the steps.
template <int N>
constexpr int kIntFpWaves = (N >= 37 && N <= 40) ? 4 : 1;

template <int N>
__global__ void __launch_bounds__(kDenseBlock) __attribute__((amdgpu_waves_per_eu(kIntFpWaves<N>)))
int_small_kernel(const double* __restrict__ M,            // [2][N][N]: +2A, -2A column-major
                 const double* __restrict__ y0,           // y(empty set), N values
                 unsigned long long* __restrict__ partial,  // [gridDim.x][2]: T mod 2^128 (lo, hi)
                 uint64_t task_begin, uint64_t task_end, uint64_t num_chunks, int L, int accumulate)
{
    __shared__ unsigned long long red[2 * (kDenseBlock / kWave)];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = (uint64_t)blockIdx.x * (kDenseBlock / kWave) + (threadIdx.x / kWave);
    const uint64_t nwave = (uint64_t)gridDim.x * (kDenseBlock / kWave);
    unsigned __int128 acc = 0;
    for (uint64_t task = task_begin + wave; task < task_end; task += nwave) {
        const uint64_t h = task * kWave + lane;
        const unsigned __int128 t = real::walk_intfp<N>(M, y0, h, L);
        if (h < num_chunks) acc += t;  // lanes past the end walk a dummy chunk
    }
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
        for (int i = 0; i < kDenseBlock / kWave; ++i) t += ((unsigned __int128)red[2 * i + 1] << 64) | red[2 * i];
        unsigned long long* out = partial + 2 * blockIdx.x;
        if (accumulate) t += ((unsigned __int128)out[1] << 64) | out[0];
        out[0] = (unsigned long long)t;
        out[1] = (unsigned long long)(t >> 64);
    }
}

template <int N>
const void* int_small_ptr()
{
    if constexpr (SPM_INTFP_ONLY_N != 0 && N != SPM_INTFP_ONLY_N) return nullptr;
    else return reinterpret_cast<const void*>(&int_small_kernel<N>);
}

template <int... Is>
const void* int_small_lookup(int n, std::integer_sequence<int, Is...>)
{
    static const void* const tab[] = {nullptr, int_small_ptr<Is + 1>()...};
    return tab[n];
}

const void* int_small_kernel_ptr(int n)
{
    if (n < 1 || n > kMaxN) return nullptr;
    return int_small_lookup(n, std::make_integer_sequence<int, kMaxN>{});
}

int int_small_bits(int n)
{
    int b = 0;
    while ((1 << b) <= n) ++b;
    return b;
}

}  // namespace spm
