// dense_c128.hip -- single-matrix dense complex128 Ryser kernel (gfx950).
//
// Replaces the reference's complex GPU kernels
// (Algorithms/Complex/gpuComputeComplex.cu:39-197, xRegisterMSharedComplex
// and its NOV-specialised twin).  The per-chunk walk is in walk_c128.h; this
// file holds the grid-strided task loop over one matrix's chunks and the
// block reduction into per-block records.
#include "walk_c128.h"

// SPM_C128_ONLY_N (sweeps only): instantiate a single order.
#ifndef SPM_C128_ONLY_N
#define SPM_C128_ONLY_N 0
#endif

namespace spm {

namespace cplx {

// Minimum waves per SIMD asked of the register allocator, per order: 3 for
// 29..35, where 168 VGPRs keep the 2-step walk free of spills (the one-off
// chunk-start code may spill a few values) and the compiler's own choice
// (up to 200 VGPRs) leaves 2 -- measured 15% faster at n = 32; at 36 the
// walk itself would spill; 1 = the compiler's choice (forcing 4 at n = 24
// measured slower).
// SPM_C128_WAVES_PER_EU (sweeps only) overrides it for every order.
#ifndef SPM_C128_WAVES_PER_EU
#define SPM_C128_WAVES_PER_EU 0
#endif
template <int N>
constexpr int kC128Waves = SPM_C128_WAVES_PER_EU > 0 ? SPM_C128_WAVES_PER_EU : (N >= 29 && N <= 35) ? 3 : 1;

template <int N, bool COMP>
__global__ void __launch_bounds__(kDenseBlock) __attribute__((amdgpu_waves_per_eu(kC128Waves<N>)))
dense_c128_kernel(const double* __restrict__ M,   // [2][N][2N]: +A then -A, column-major, (re, im)
                  const double* __restrict__ x0,  // 2N doubles
                  double* __restrict__ partial,   // [gridDim.x][8]: re hi lo abs, im hi lo 0, 0, 0
                  uint64_t task_begin, uint64_t task_end, uint64_t num_chunks, int L, int accumulate)
{
    __shared__ double red[3 * (kDenseBlock / kWave)];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = (uint64_t)blockIdx.x * (kDenseBlock / kWave) + (threadIdx.x / kWave);
    const uint64_t nwave = (uint64_t)gridDim.x * (kDenseBlock / kWave);

    // Only a grid with fewer than 64 chunks has idle lanes, and then each
    // lane has exactly one task: such lanes' sums are dropped at the end.
    const bool idle = (uint64_t)lane >= num_chunks;
    double re_hi = 0.0, re_lo = 0.0, im_hi = 0.0, im_lo = 0.0, abs_sum = 0.0;

    for (uint64_t task = task_begin + wave; task < task_end; task += nwave) {
        const uint64_t h = task * kWave + lane;
        walk_c128<N, COMP>(M, x0, h, L, re_hi, re_lo, im_hi, im_lo, abs_sum);
    }
    if (idle) re_hi = re_lo = im_hi = im_lo = abs_sum = 0.0;

    double zero_abs = 0.0;
    block_reduce_dd(re_hi, re_lo, abs_sum, red);
    __syncthreads();
    block_reduce_dd(im_hi, im_lo, zero_abs, red);
    if (threadIdx.x == 0) {
        double* out = partial + 8 * blockIdx.x;
        if (accumulate) {
            dd_add(re_hi, re_lo, out[0], out[1]);
            abs_sum += out[2];
            dd_add(im_hi, im_lo, out[3], out[4]);
        }
        out[0] = re_hi;
        out[1] = re_lo;
        out[2] = abs_sum;
        out[3] = im_hi;
        out[4] = im_lo;
        out[5] = 0.0;
        out[6] = 0.0;
        out[7] = 0.0;
    }
}

template <int N, bool COMP>
const void* c128_ptr()
{
    if constexpr (SPM_C128_ONLY_N != 0 && N != SPM_C128_ONLY_N) return nullptr;
    else return reinterpret_cast<const void*>(&dense_c128_kernel<N, COMP>);
}

template <int... Is>
const void* c128_lookup(int n, bool comp, std::integer_sequence<int, Is...>)
{
    static const void* const plain[] = {nullptr, c128_ptr<Is + 1, false>()...};
    static const void* const compd[] = {nullptr, c128_ptr<Is + 1, true>()...};
    return comp ? compd[n] : plain[n];
}

}  // namespace cplx

const void* dense_c128_kernel_ptr(int n, bool compensated)
{
    if (n < 1 || n > kMaxComplexN) return nullptr;
    return cplx::c128_lookup(n, compensated, std::make_integer_sequence<int, kMaxComplexN>{});
}

}  // namespace spm
