// walk_f64.h -- Gray-code (Nijenhuis-Wilf) Ryser walk for dense real FP64
// matrices on CDNA4 (gfx950): device code shared by the single-matrix kernel
// (dense_f64.hip) and the batched kernel (batch.hip).
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
#ifndef SPM_WALK_F64_H
#define SPM_WALK_F64_H

#include "spm_device.h"
#include "spm_kernels.h"

#include <type_traits>

// Tuning knobs (build-time).  SPM_DENSE_WAVES_PER_EU (sweeps only) caps
// VGPRs so that this many waves fit per SIMD for every order (0 = the
// per-order table kDenseWaves); SPM_DENSE_UNROLL is U, the log2 of the steps
// per unrolled block.
#ifndef SPM_DENSE_WAVES_PER_EU
#define SPM_DENSE_WAVES_PER_EU 0
#endif
#ifndef SPM_DENSE_UNROLL
#define SPM_DENSE_UNROLL 1
#endif
// Steps whose terms are summed in plain FP64 before each compensated fold.
#ifndef SPM_DENSE_FOLD_STEPS
#define SPM_DENSE_FOLD_STEPS 16
#endif
// SPM_DENSE_ONLY_N (sweeps only): instantiate a single order.
#ifndef SPM_DENSE_ONLY_N
#define SPM_DENSE_ONLY_N 0
#endif

namespace spm {
namespace real {

constexpr uint32_t kFoldSteps = SPM_DENSE_FOLD_STEPS;

template <int J, int END, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (J < END) {
        f(std::integral_constant<int, J>{});
        static_for<J + 1, END>(f);
    }
}

// Product of the first N-1 rows, in four independent chains for ILP.
template <int N, class T>
__device__ __forceinline__ T head_product(const T (&x)[N])
{
    constexpr int H = N - 1;
    if constexpr (H == 0) {
        return T(1);
    } else if constexpr (H <= 4) {
        T p = x[0];
#pragma unroll
        for (int r = 1; r < H; ++r) p *= x[r];
        return p;
    } else {
        T c0 = x[0], c1 = x[1], c2 = x[2], c3 = x[3];
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

// A matrix column held in SGPRs (wave-uniform values from scalar loads).
// The pipelined walk keeps a prefetched column live across a step's
// multiplies; a full column needs 2N SGPRs, which fits the 102-SGPR file up
// to N = 40.  Larger orders prefetch a 32-value head and load the tail when
// the column is applied.
__device__ __forceinline__ void step_fence()
{
    __builtin_amdgcn_sched_barrier(0);
}

#ifndef SPM_DENSE_PREFETCH_MAX
#define SPM_DENSE_PREFETCH_MAX 40
#endif
template <int N>
constexpr int kColumnHead = N <= SPM_DENSE_PREFETCH_MAX ? N : 32;

// Minimum waves per SIMD asked of the register allocator: 5 for the orders
// 29..36, whose natural allocation (97..111 VGPRs) leaves 4 and where 96
// VGPRs keep the 2-step walk free of spills (measured 3-5% faster at n = 30,
// 32, 36; at n = 40 occupancy 5 measured 1.7% slower than 4); 1 = the
// compiler's choice.
template <int N>
constexpr int kDenseWaves = SPM_DENSE_WAVES_PER_EU > 0 ? SPM_DENSE_WAVES_PER_EU : (N >= 29 && N <= 36) ? 5 : 1;

// Unroll depth per order: the pipelined 2-step blocks up to the order whose
// prefetched column still fits the SGPR file; single steps beyond that and
// for the tiny orders.
template <int N>
constexpr int kDenseUnroll = (N >= SPM_DENSE_UNROLL + 2 && N <= SPM_DENSE_PREFETCH_MAX) ? SPM_DENSE_UNROLL : 0;

template <int N, bool TAIL = (kColumnHead<N> < N)>
struct Column {
    double v[kColumnHead<N>];
};
template <int N>
struct Column<N, true> {
    double v[kColumnHead<N>];
    const double* p;  // column base, for the tail loads
};

template <int N>
__device__ __forceinline__ void load_column(Column<N>& c, const double* p)
{
    const const_f64* __restrict__ q = step_column(p);
#pragma unroll
    for (int r = 0; r < kColumnHead<N>; ++r) c.v[r] = q[r];
    if constexpr (kColumnHead<N> < N) c.p = p;
}

template <int N>
__device__ __forceinline__ void add_column(double (&x)[N], const Column<N>& c)
{
    constexpr int H = kColumnHead<N>;
#pragma unroll
    for (int r = 0; r < H; ++r) x[r] += c.v[r];
    if constexpr (H < N) {
        step_fence();  // the head's SGPRs are free before the tail loads
        const const_f64* __restrict__ q = step_column(c.p);
#pragma unroll
        for (int r = H; r < N; ++r) x[r] += q[r];
    }
}

template <int N>
__device__ __forceinline__ void fma_column(double (&x)[N], const Column<N>& c, double s)
{
    constexpr int H = kColumnHead<N>;
#pragma unroll
    for (int r = 0; r < H; ++r) x[r] = __builtin_fma(s, c.v[r], x[r]);
    if constexpr (H < N) {
        step_fence();
        const const_f64* __restrict__ q = step_column(c.p);
#pragma unroll
        for (int r = H; r < N; ++r) x[r] = __builtin_fma(s, q[r], x[r]);
    }
}

// The accumulation policy of a walk: add(x, negative) takes the term
// (-1)^k prod_i x_i of the current subset, fold() is called every
// kFoldSteps steps and at the end of the chunk.
//
// Compensated/plain FP64 sums: block sums in plain FP64 for kFoldSteps
// steps, folded into the per-lane (compensated) sum.
template <int N, bool COMP>
struct FpTerm {
    static constexpr bool kCompensated = COMP;
    double t_hi = 0.0, t_lo = 0.0, t_abs = 0.0;
    double blk = 0.0, blk_abs = 0.0;
    __device__ __forceinline__ void add(const double (&x)[N], bool negative)
    {
        add_term<N, COMP>(x, negative, blk, blk_abs);
    }
    __device__ __forceinline__ void fold()
    {
        if constexpr (COMP) {
            dd_add_d(t_hi, t_lo, blk);
            t_abs += blk_abs;
        } else {
            t_hi += blk;
        }
        blk = 0.0;
        blk_abs = 0.0;
    }
};

// FP32 products for nonnegative matrices whose rows the host scaled by
// powers of two so that every |x_i| < 2^(c+1) with (c+1) N <= 127: x stays
// FP64 (exact updates), the first N-1 factors are rounded to FP32 and
// multiplied in FP32 (no overflow by the scaling), and the FP32 product
// times x_(N-1) is accumulated in the compensated FP64 sums.  Per term the
// relative error is below (2N-3) 2^-24 to first order (N-1 roundings of x,
// N-2 of the product), unless the FP32 product falls below 2^-126.
template <int N>
struct F32ProdTerm {
    static constexpr bool kCompensated = true;
    double t_hi = 0.0, t_lo = 0.0, t_abs = 0.0;
    double blk = 0.0, blk_abs = 0.0;
    __device__ __forceinline__ void add(const double (&x)[N], bool negative)
    {
        float xf[N];
#pragma unroll
        for (int r = 0; r + 1 < N; ++r) xf[r] = (float)x[r];
        xf[N - 1] = 0.0f;  // not a factor of the head product
        const double p = (double)head_product<N>(xf);
        blk = __builtin_fma(negative ? -p : p, x[N - 1], blk);
        blk_abs = __builtin_fma(__builtin_fabs(p), __builtin_fabs(x[N - 1]), blk_abs);
    }
    __device__ __forceinline__ void fold()
    {
        dd_add_d(t_hi, t_lo, blk);
        t_abs += blk_abs;
        blk = 0.0;
        blk_abs = 0.0;
    }
};

// One lane's chunk: the 2^L Gray steps starting at k = h 2^L, every term
// handed to `term`.
template <int N, int U, class Term>
__device__ __forceinline__ void walk_gray(const double* __restrict__ M,   // [2][N][N]: +A, -A
                                          const double* __restrict__ x0,  // x(empty set)
                                          uint64_t h, int L, Term& term)
{
    // x at the chunk start G(h 2^L): columns >= L-1 may be set.
    // (Loads go through step_column too: kept as loop invariants, x0 alone
    // would pin 2N SGPRs for the whole task loop.)
    double x[N];
    {
        const const_f64* __restrict__ xs = step_column(x0);
#pragma unroll
        for (int r = 0; r < N; ++r) x[r] = xs[r];
    }
    const uint64_t k0 = h << L;
    const uint64_t g0 = k0 ^ (k0 >> 1);
    for (int j = (L > 0 ? L - 1 : 0); j < N - 1; ++j) {
        const double b = ((g0 >> j) & 1u) ? 1.0 : 0.0;
        const const_f64* __restrict__ col = step_column(M + j * N);
#pragma unroll
        for (int r = 0; r < N; ++r) x[r] = __builtin_fma(b, col[r], x[r]);
    }

    const double lane_sign = (h & 1u) ? -1.0 : 1.0;  // direction of column L-1
    const uint32_t nblk = 1u << (L - U);
    auto fold = [&]() __attribute__((always_inline)) { term.fold(); };
    if constexpr (U == 1) {
        // Blocks of 2 steps: k = 2m flips column ctz(m)+1, k = 2m+1 flips
        // column 0.  A uniform direction selects the +A or -A plane, so a
        // flip is N plain v_add_f64 with SGPR operands; only block
        // m = nblk/2 (column L-1) has a per-lane direction and uses FMA.
        // Software pipeline: the next column's scalar loads are issued as
        // soon as the current flip has consumed its SGPRs, and the step's
        // N-1 multiplies cover their latency; the fences stop the
        // scheduler from hoisting loads further (SGPR file overflow).
        const double* Mneg = M + N * N;
        auto a_col = [&](uint32_t m) __attribute__((always_inline)) {  // column of step k = 2m, uniform direction
            const int tz = __builtin_ctz(m);
            return (((m >> (tz + 1)) & 1u) ? Mneg : M) + (tz + 1) * N;
        };
        const uint32_t half = nblk >> 1;
        constexpr uint32_t G = kFoldSteps / 2;
        Column<N> ca, cb;
        term.add(x, false);  // k = 0: the chunk start
        load_column(cb, M);                          // k = 1: +column 0
        add_column(x, cb);
        load_column(ca, a_col(1));
        step_fence();
        term.add(x, true);
        step_fence();
        // One block; SPECIAL marks block `half`, whose column L-1 flips in a
        // per-lane direction.  Kept branch-free: the loop is split there.
        auto block = [&](uint32_t m, auto SPECIAL) __attribute__((always_inline)) {
            if constexpr (decltype(SPECIAL)::value) fma_column(x, ca, lane_sign);
            else add_column(x, ca);
            load_column(cb, (m & 1u) ? Mneg : M);  // step 2m+1: column 0
            step_fence();
            term.add(x, false);
            step_fence();
            add_column(x, cb);
            load_column(ca, a_col(m + 1));  // m+1 = nblk loads column L <= N-1: harmless
            step_fence();
            term.add(x, true);
            step_fence();
        };
        auto run = [&](uint32_t m, uint32_t end) __attribute__((always_inline)) {
            while (m < end) {
                const uint32_t mend = min(end, (m | (G - 1)) + 1);
#pragma nounroll
                for (; m < mend; ++m) block(m, std::false_type{});
                fold();
            }
        };
        run(1, half);
        block(half, std::true_type{});
        run(half + 1, nblk);
        fold();
    } else {
        // Unpipelined walk (U = 0: tiny orders; U >= 2: tuning builds).
        // Block m starts with the flip of the runtime column ctz(m) + U;
        // its other steps flip compile-time columns.
        constexpr uint32_t G = (kFoldSteps >> U) > 0 ? (kFoldSteps >> U) : 1;
        auto block_tail = [&](uint32_t m) __attribute__((always_inline)) {
            if constexpr (U > 0) {
                // bit U of the step index.  Held in a VGPR: a VOP3 op takes
                // one SGPR source, and that slot belongs to the column value.
                double s_mid = (m & 1u) ? -1.0 : 1.0;
                asm("" : "+v"(s_mid));
                static_for<1, (1 << U)>([&](auto J) __attribute__((always_inline)) {
                    constexpr int j = decltype(J)::value;
                    constexpr int c = __builtin_ctz(j);
                    if constexpr (c + 1 < U) {
                        flip_const<N, ((j >> (c + 1)) & 1) != 0>(x, M + c * N);
                    } else {
                        flip_signed<N>(x, M + c * N, s_mid);
                    }
                    term.add(x, (j & 1) != 0);
                });
            }
        };
        term.add(x, false);  // block 0: the chunk start
        block_tail(0u);
        uint32_t m = 1;
        while (m < nblk) {
            const uint32_t mend = min(nblk, (m | (G - 1)) + 1);
#pragma nounroll
            for (; m < mend; ++m) {
                const int tz = __builtin_ctz(m);
                const int c = tz + U;
                const double s = (c + 1 < L) ? (((m >> (tz + 1)) & 1u) ? -1.0 : 1.0) : lane_sign;
                flip_signed<N>(x, M + c * N, s);
                term.add(x, U == 0 && (m & 1u));
                block_tail(m);
            }
            fold();
        }
        fold();
    }
}

// The compensated (COMP) or plain FP64 sum of one lane's chunk: t_hi + t_lo
// and sum |term| (t_abs, COMP only).
template <int N, int U, bool COMP>
__device__ __forceinline__ void walk_f64(const double* __restrict__ M, const double* __restrict__ x0, uint64_t h,
                                         int L, double& t_hi, double& t_lo, double& t_abs)
{
    FpTerm<N, COMP> term;
    walk_gray<N, U>(M, x0, h, L, term);
    t_hi = term.t_hi;
    t_lo = term.t_lo;
    t_abs = term.t_abs;
}

}  // namespace real
}  // namespace spm

#endif
