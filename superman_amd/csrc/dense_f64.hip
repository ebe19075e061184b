// dense_f64.hip -- Gray-code (Nijenhuis-Wilf) Ryser for dense real FP64
// matrices on CDNA4 (gfx950).
//
// Replaces the reference's dense GPU kernels
// (Algorithms/GPU/Dense/DenseKernelDefinitions.cuh:37-552: xRegisterMShared,
// xRegisterMSharedKahan and their placement variants) with one design:
//
//   perm(A) = 2 (-1)^(n-1) sum_{k=0}^{2^(n-1)-1} (-1)^k prod_i x_i(G(k)),
//   x_i(S)  = a_{i,n-1} - rowsum_i / 2 + sum_{j in S} a_ij,  G(k) = k ^ (k>>1).
//
// Work decomposition.  Lane h (a "chunk") owns the contiguous, aligned Gray
// range k in [h 2^L, (h+1) 2^L).  Inside an aligned chunk the flipped column
// c = ctz(k) and the term sign (-1)^k depend only on the low L bits, so all
// 64 lanes of a wavefront walk the SAME column sequence in lockstep: the
// column index, its +/- direction and the loop control are wave-uniform and
// live in SGPRs, and the column values are fetched with scalar loads and fed
// to v_add_f64 as SGPR operands.  (The reference keeps the matrix in shared
// memory; on CDNA a broadcast LDS read costs LDS cycles per lane and would
// cap the kernel near half the FP64 VALU rate.)  Only one step per chunk (the
// flip of column L-1, whose direction is bit 0 of h) has a per-lane sign.
//
// The walk is unrolled by 2^U steps: steps j = 1..2^U-1 of a block flip the
// compile-time columns ctz(j) < U with compile-time signs (except column U-1,
// whose sign is bit 0 of the block index, uniform); step j = 0 flips the
// runtime column ctz(m) + U.  Per subset the kernel issues n adds, n-2
// multiplies and 1 FMA (the last factor fused with the accumulation), which
// is the 2n-1 VALU floor of the formula.
//
// x is rebuilt exactly from the matrix at the start of every chunk (the same
// ascending-column summation the reference uses for its chunk start,
// DenseKernelDefinitions.cuh:230-240), which bounds incremental drift to 2^L
// steps.  Precision modes:
//   PLAIN (reference "dd"):    per-block plain sums folded into one double;
//   COMP  (reference "kahan"): 2^U-term block sums folded with TwoSum into a
//          per-lane double-double, plus sum |term| for the error estimate;
//          the lo word is kept through every reduction (the reference drops
//          its Kahan error term, DenseKernelDefinitions.cuh:263).
#include "spm_device.h"
#include "spm_kernels.h"

#include <type_traits>

namespace spm {

template <int J, int END, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (J < END) {
        f(std::integral_constant<int, J>{});
        static_for<J + 1, END>(f);
    }
}

// Product of the first N-1 rows, in four independent chains for ILP.
template <int N>
__device__ __forceinline__ double head_product(const double (&x)[N])
{
    constexpr int H = N - 1;
    if constexpr (H == 0) {
        return 1.0;
    } else if constexpr (H <= 4) {
        double p = x[0];
#pragma unroll
        for (int r = 1; r < H; ++r) p *= x[r];
        return p;
    } else {
        double c0 = x[0], c1 = x[1], c2 = x[2], c3 = x[3];
#pragma unroll
        for (int r = 4; r < H; ++r) {
            if ((r & 3) == 0) c0 *= x[r];
            else if ((r & 3) == 1) c1 *= x[r];
            else if ((r & 3) == 2) c2 *= x[r];
            else c3 *= x[r];
        }
        return (c0 * c1) * (c2 * c3);
    }
}

template <int N, bool COMP>
__device__ __forceinline__ void add_term(const double (&x)[N], bool negative, double& blk, double& blk_abs)
{
    const double p = head_product<N>(x);
    blk = __builtin_fma(negative ? -p : p, x[N - 1], blk);
    if constexpr (COMP) blk_abs = __builtin_fma(__builtin_fabs(p), __builtin_fabs(x[N - 1]), blk_abs);
}

// The column base is passed through an empty asm that claims to rewrite it,
// so the scalar loads of a step cannot be hoisted out of the walk (hoisting
// all columns would overflow the SGPR file and spill to VGPR lanes, costing
// v_readlane VALU issues every step).
// The result is a constant-address-space (read-only, wave-uniform) pointer,
// which makes every read of the column a scalar load into SGPRs.
typedef __attribute__((address_space(4))) const double const_f64;

__device__ __forceinline__ const const_f64* step_column(const double* p)
{
    asm volatile("" : "+s"(p));
    return (const const_f64*)p;
}

// x[r] += s * col[r] with s = +-1 (exact: same rounding as the add/sub).
template <int N>
__device__ __forceinline__ void flip_signed(double (&x)[N], const double* __restrict__ colp, double s)
{
    const const_f64* __restrict__ col = step_column(colp);
#pragma unroll
    for (int r = 0; r < N; ++r) x[r] = __builtin_fma(s, col[r], x[r]);
}

template <int N, bool NEG>
__device__ __forceinline__ void flip_const(double (&x)[N], const double* __restrict__ colp)
{
    const const_f64* __restrict__ col = step_column(colp);
#pragma unroll
    for (int r = 0; r < N; ++r) x[r] = NEG ? x[r] - col[r] : x[r] + col[r];
}

template <int N, int U, bool COMP>
__global__ void __launch_bounds__(kDenseBlock)
dense_f64_kernel(const double* __restrict__ M,   // column-major: column c at M + c*N
                 const double* __restrict__ x0,  // x(empty set), N values
                 double* __restrict__ partial,   // [gridDim.x][4]: hi, lo, abs, 0
                 uint64_t task_begin, uint64_t task_end, uint64_t num_chunks, int L,
                 int accumulate)
{
    __shared__ double red[3 * (kDenseBlock / kWave)];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = (uint64_t)blockIdx.x * (kDenseBlock / kWave) + (threadIdx.x / kWave);
    const uint64_t nwave = (uint64_t)gridDim.x * (kDenseBlock / kWave);

    double acc_hi = 0.0, acc_lo = 0.0, acc_abs = 0.0;

    for (uint64_t task = task_begin + wave; task < task_end; task += nwave) {
        const uint64_t h = task * kWave + lane;
        const bool active = h < num_chunks;  // lanes past the end walk a dummy chunk

        // x at the chunk start G(h 2^L): columns >= L-1 may be set.
        double x[N];
#pragma unroll
        for (int r = 0; r < N; ++r) x[r] = x0[r];
        const uint64_t k0 = h << L;
        const uint64_t g0 = k0 ^ (k0 >> 1);
        for (int j = (L > 0 ? L - 1 : 0); j < N - 1; ++j) {
            const double b = ((g0 >> j) & 1u) ? 1.0 : 0.0;
            const double* __restrict__ col = M + j * N;
#pragma unroll
            for (int r = 0; r < N; ++r) x[r] = __builtin_fma(b, col[r], x[r]);
        }

        const double lane_sign = (h & 1u) ? -1.0 : 1.0;  // direction of column L-1
        const uint32_t nblk = 1u << (L - U);
        double t_hi = 0.0, t_lo = 0.0, t_abs = 0.0;
        for (uint32_t m = 0; m < nblk; ++m) {
            double blk = 0.0, blk_abs = 0.0;
            // step j = 0: runtime column ctz(m) + U (none for the chunk start)
            if (m != 0) {
                const int tz = __builtin_ctz(m);
                const int c = tz + U;
                const double s = (c + 1 < L) ? (((m >> (tz + 1)) & 1u) ? -1.0 : 1.0) : lane_sign;
                flip_signed<N>(x, M + c * N, s);
            }
            add_term<N, COMP>(x, U == 0 && (m & 1u), blk, blk_abs);
            if constexpr (U > 0) {
                // bit U of the step index.  Held in a VGPR: a VOP3 op takes one
                // SGPR source, and that slot belongs to the column value.
                double s_mid = (m & 1u) ? -1.0 : 1.0;
                asm("" : "+v"(s_mid));
                static_for<1, (1 << U)>([&](auto J) {
                    constexpr int j = decltype(J)::value;
                    constexpr int c = __builtin_ctz(j);
                    if constexpr (c + 1 < U) {
                        flip_const<N, ((j >> (c + 1)) & 1) != 0>(x, M + c * N);
                    } else {
                        flip_signed<N>(x, M + c * N, s_mid);
                    }
                    add_term<N, COMP>(x, (j & 1) != 0, blk, blk_abs);
                });
            }
            if constexpr (COMP) {
                dd_add_d(t_hi, t_lo, blk);
                t_abs += blk_abs;
            } else {
                t_hi += blk;
            }
        }
        if (active) {
            if constexpr (COMP) {
                dd_add(acc_hi, acc_lo, t_hi, t_lo);
                acc_abs += t_abs;
            } else {
                acc_hi += t_hi;
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
    constexpr int U = (N >= 5) ? 3 : 0;
    return reinterpret_cast<const void*>(&dense_f64_kernel<N, U, COMP>);
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

int dense_f64_unroll_bits(int n) { return n >= 5 ? 3 : 0; }

const void* reduce_partials_kernel_ptr() { return reinterpret_cast<const void*>(&reduce_partials_kernel); }

}  // namespace spm
