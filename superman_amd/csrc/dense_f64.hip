// dense_f64.hip -- single-matrix dense real FP64 Ryser kernel (gfx950).
//
// Replaces the reference's dense GPU kernels
// (Algorithms/GPU/Dense/DenseKernelDefinitions.cuh:37-552: xRegisterMShared,
// xRegisterMSharedKahan and their placement variants).  The per-chunk walk is
// in walk_f64.h; this file holds the grid-strided task loop over one
// matrix's chunks, the block reduction into per-block records and the final
// reduction kernel.
#include "walk_f64.h"

namespace spm {

// The grid-strided task loop over one matrix's chunks and the block
// reduction into per-block records, for the accumulation policy Term.
template <int N, int U, class Term>
__device__ __forceinline__ void dense_chunks(const double* __restrict__ M, const double* __restrict__ x0,
                                             double* __restrict__ partial, uint64_t task_begin,
                                             uint64_t task_end, uint64_t num_chunks, int L, int accumulate)
{
    constexpr bool COMP = Term::kCompensated;
    __shared__ double red[3 * (kDenseBlock / kWave)];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = (uint64_t)blockIdx.x * (kDenseBlock / kWave) + (threadIdx.x / kWave);
    const uint64_t nwave = (uint64_t)gridDim.x * (kDenseBlock / kWave);

    double acc_hi = 0.0, acc_lo = 0.0, acc_abs = 0.0;
    for (uint64_t task = task_begin + wave; task < task_end; task += nwave) {
        const uint64_t h = task * kWave + lane;
        const bool active = h < num_chunks;  // lanes past the end walk a dummy chunk
        Term term;
        real::walk_gray<N, U>(M, x0, h, L, term);
        if (active) {
            if constexpr (COMP) {
                dd_add(acc_hi, acc_lo, term.t_hi, term.t_lo);
                acc_abs += term.t_abs;
            } else {
                acc_hi += term.t_hi;
            }
        }
    }
    block_reduce_dd(acc_hi, acc_lo, acc_abs, red);
    if (threadIdx.x == 0) {
        double* out = partial + 4 * blockIdx.x;
        if (accumulate) {  // later slice of a long run: add onto this block's record
            dd_add(acc_hi, acc_lo, out[0], out[1]);
            acc_abs += out[2];
        }
        out[0] = acc_hi;
        out[1] = acc_lo;
        out[2] = acc_abs;
        out[3] = 0.0;
    }
}

template <int N, int U, bool COMP>
__global__ void __launch_bounds__(kDenseBlock) __attribute__((amdgpu_waves_per_eu(real::kDenseWaves<N>)))
dense_f64_kernel(const double* __restrict__ M,   // [2][N][N]: +A then -A, column-major
                 const double* __restrict__ x0,  // x(empty set), N values
                 double* __restrict__ partial,   // [gridDim.x][4]: hi, lo, abs, 0
                 uint64_t task_begin, uint64_t task_end, uint64_t num_chunks, int L,
                 int accumulate)
{
    dense_chunks<N, U, real::FpTerm<N, COMP>>(M, x0, partial, task_begin, task_end, num_chunks, L, accumulate);
}

// FP32-product variant for nonnegative, row-scaled matrices (same records
// as the compensated kernel).
template <int N, int U>
__global__ void __launch_bounds__(kDenseBlock) __attribute__((amdgpu_waves_per_eu(real::kDenseWaves<N>)))
dense_f32p_kernel(const double* __restrict__ M, const double* __restrict__ x0, double* __restrict__ partial,
                  uint64_t task_begin, uint64_t task_end, uint64_t num_chunks, int L, int accumulate)
{
    dense_chunks<N, U, real::F32ProdTerm<N>>(M, x0, partial, task_begin, task_end, num_chunks, L, accumulate);
}

// Final reduction of the per-block partials (one block; fixed order).
// `stride` doubles per record, `comps` double-double components per record.
__global__ void __launch_bounds__(256)
reduce_partials_kernel(const double* __restrict__ partial, int nrec, int stride, int comps,
                       double* __restrict__ out)
{
    __shared__ double red[3 * (256 / kWave)];
    for (int q = 0; q < comps; ++q) {
        double h = 0.0, l = 0.0, a = 0.0;
        for (int i = threadIdx.x; i < nrec; i += blockDim.x) {
            const double* p = partial + (size_t)i * stride + 3 * q;
            dd_add(h, l, p[0], p[1]);
            a += p[2];
        }
        block_reduce_dd(h, l, a, red);
        if (threadIdx.x == 0) {
            out[3 * q + 0] = h;
            out[3 * q + 1] = l;
            out[3 * q + 2] = a;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Instantiation table: N = 1..kMaxN, U = 3 for N >= 5 (blocks of 8 steps),
// U = 0 for the tiny orders.
// ---------------------------------------------------------------------------
template <int N, bool COMP>
const void* dense_f64_ptr()
{
    constexpr int U = real::kDenseUnroll<N>;
    if constexpr (SPM_DENSE_ONLY_N != 0 && N != SPM_DENSE_ONLY_N) return nullptr;
    else return reinterpret_cast<const void*>(&dense_f64_kernel<N, U, COMP>);
}

template <int N>
const void* dense_f32p_ptr()
{
    constexpr int U = real::kDenseUnroll<N>;
    if constexpr (SPM_DENSE_ONLY_N != 0 && N != SPM_DENSE_ONLY_N) return nullptr;
    else return reinterpret_cast<const void*>(&dense_f32p_kernel<N, U>);
}

template <int... Is>
const void* dense_f32p_lookup(int n, std::integer_sequence<int, Is...>)
{
    static const void* const table[] = {nullptr, dense_f32p_ptr<Is + 1>()...};
    return table[n];
}

template <int... Is>
const void* dense_f64_lookup(int n, bool comp, std::integer_sequence<int, Is...>)
{
    static const void* const plain[] = {nullptr, dense_f64_ptr<Is + 1, false>()...};
    static const void* const compd[] = {nullptr, dense_f64_ptr<Is + 1, true>()...};
    return comp ? compd[n] : plain[n];
}

const void* dense_f64_kernel_ptr(int n, bool compensated)
{
    if (n < 1 || n > kMaxN) return nullptr;
    return dense_f64_lookup(n, compensated, std::make_integer_sequence<int, kMaxN>{});
}

const void* dense_f32p_kernel_ptr(int n)
{
    if (n < 1 || n > kMaxN) return nullptr;
    return dense_f32p_lookup(n, std::make_integer_sequence<int, kMaxN>{});
}

int dense_f64_unroll_bits(int n)
{
    return (n >= SPM_DENSE_UNROLL + 2 && n <= SPM_DENSE_PREFETCH_MAX) ? SPM_DENSE_UNROLL : 0;
}

const void* reduce_partials_kernel_ptr() { return reinterpret_cast<const void*>(&reduce_partials_kernel); }

}  // namespace spm
