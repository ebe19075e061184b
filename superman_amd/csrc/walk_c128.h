// walk_c128.h -- Gray-code (Nijenhuis-Wilf) Ryser walk for dense complex128
// matrices on CDNA4 (gfx950): device code shared by the single-matrix kernel
// (dense_c128.hip) and the batched kernel (batch.hip).
//
// Replaces the reference's complex GPU kernels
// (Algorithms/Complex/gpuComputeComplex.cu:39-197, xRegisterMSharedComplex
// and its NOV-specialised twin) with the decomposition of dense_f64.hip:
// lane h owns the aligned Gray chunk [h 2^L, (h+1) 2^L), all lanes of a wave
// flip the same column in lockstep, the column comes in through scalar loads
// from a +A or -A plane chosen by the (uniform) flip direction, and only the
// flip of column L-1 has a per-lane direction.
//
// x is complex: re/im interleaved in 2N doubles (4N VGPRs).  A column is 2N
// doubles (4N dwords: 128 SGPRs at N = 32), more than the SGPR file can hold
// next to the loop state, so each flip takes its column in parts of at most
// 32 doubles, one part per phase of rows, and loads the next part while
// the product runs over the rows the previous part updated (gray_step).
// Every lambda of the walk is force-inlined: an outlined one takes x by
// reference and sends it to scratch memory.
//
// Per subset: 2N adds (the flip), 4(N-2) mul/FMA for the product of N-1
// rows (complex product as mul + FMA per part), 4 FMAs to accumulate the
// last factor times the product, = 6N - 4 VALU ops for 8N algorithmic flops.
// The reference uses cuCmul/cuCadd with no compensation
// (gpuComputeComplex.cu:98-114); here both the real and imaginary sums are
// compensated (TwoSum fold every kFoldSteps steps) in the COMP variant.
#ifndef SPM_WALK_C128_H
#define SPM_WALK_C128_H

#include "spm_device.h"
#include "spm_kernels.h"

#include <type_traits>

namespace spm {
namespace cplx {

constexpr uint32_t kFoldSteps = 16;
constexpr int kChunk = 8;  // doubles per streamed column chunk (one 64-byte scalar load)

typedef __attribute__((address_space(4))) const double const_f64;

// A column address passed through an empty asm that claims to rewrite it:
// the column's scalar loads stay at their use (hoisted as loop invariants
// they would pin the SGPR file), and the asm costs no instruction.  Constant
// offsets are added after it, so that they fold into the loads' immediates.
__device__ __forceinline__ const const_f64* step_ptr(const double* p)
{
    asm volatile("" : "+s"(p));
    return (const const_f64*)p;
}

__device__ __forceinline__ void fence() { __builtin_amdgcn_sched_barrier(0); }

template <int J, int END, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (J < END) {
        f(std::integral_constant<int, J>{});
        static_for<J + 1, END>(f);
    }
}

template <int N>
constexpr int kChunks = (2 * N + kChunk - 1) / kChunk;

struct Chunk {
    double v[kChunk];
};

// chunk K of a column (doubles [8K, 8K+8) clipped to 2N)
template <int N, int K>
__device__ __forceinline__ void load_chunk(Chunk& c, const double* col)
{
    const const_f64* __restrict__ q = step_ptr(col) + K * kChunk;
#pragma unroll
    for (int i = 0; i < kChunk; ++i)
        if (K * kChunk + i < 2 * N) c.v[i] = q[i];
}

// x = column (MODE 0), x += column (1) or x = fma(s, column, x) (2), the
// column streamed chunk by chunk (chunk-start rebuild and x(empty)).
template <int N, int MODE>
__device__ __forceinline__ void stream_column(double (&x)[2 * N], const double* col, double s)
{
    static_for<0, kChunks<N>>([&](auto KK) __attribute__((always_inline)) {
        constexpr int k = decltype(KK)::value;
        Chunk c;
        load_chunk<N, k>(c, col);
#pragma unroll
        for (int i = 0; i < kChunk; ++i) {
            constexpr int base = k * kChunk;
            if (base + i < 2 * N) {
                if constexpr (MODE == 0) x[base + i] = c.v[i];
                else if constexpr (MODE == 1) x[base + i] += c.v[i];
                else x[base + i] = __builtin_fma(s, c.v[i], x[base + i]);
            }
        }
        fence();
    });
}

// The column of a flip is consumed in kPhases<N> parts of at most
// kPartMax doubles (2 kPhaseRows<N> doubles = the rows of one phase).  Part
// p+1 is loaded while the product runs over the rows that part p just
// updated, and the next flip's part 0 while the product finishes: every
// wait for a scalar load follows a product segment, so no load is waited
// on right after it is issued (scalar loads complete out of order, so each
// wait is for all of them).
// SPM_C128_PART_MAX (sweeps only) overrides the part size.
#ifndef SPM_C128_PART_MAX
#define SPM_C128_PART_MAX 32
#endif
constexpr int kPartMax = SPM_C128_PART_MAX;  // doubles (64 SGPRs at 32)
template <int N>
constexpr int kPhases = (2 * N + kPartMax - 1) / kPartMax;
template <int N>
constexpr int kPhaseRows = (N + kPhases<N> - 1) / kPhases<N>;

template <int N>
struct Part {
    double v[2 * kPhaseRows<N>];
};

// Part P of a column: doubles [2 P R, 2 min(N, (P+1) R)), R = kPhaseRows<N>.
template <int N, int P>
__device__ __forceinline__ void load_part(Part<N>& c, const double* col)
{
    constexpr int b = 2 * P * kPhaseRows<N>;
    const const_f64* __restrict__ q = step_ptr(col) + b;
#pragma unroll
    for (int i = 0; i < 2 * kPhaseRows<N>; ++i)
        if (b + i < 2 * N) c.v[i] = q[i];
}

// (ar, ai) *= (br, bi): mul + FMA per part.
__device__ __forceinline__ void cmul(double& ar, double& ai, double br, double bi)
{
    const double t = ai * bi;
    const double u = ai * br;
    const double r = __builtin_fma(ar, br, -t);
    const double i = __builtin_fma(ar, bi, u);
    ar = r;
    ai = i;
}

// Product of rows 0..N-2 in four chains (one for N <= 5), advanced over the
// row range [R0, R1) so that it can be split between column parts.
struct Prod {
    double r[4], i[4];
};

template <int N, int R0, int R1>
__device__ __forceinline__ void prod_rows(const double (&x)[2 * N], Prod& p)
{
    constexpr int H = N - 1;
    static_for<R0, (R1 < H ? R1 : H)>([&](auto RR) __attribute__((always_inline)) {
        constexpr int r = decltype(RR)::value;
        if constexpr (H <= 4) {
            if constexpr (r == 0) {
                p.r[0] = x[0];
                p.i[0] = x[1];
            } else {
                cmul(p.r[0], p.i[0], x[2 * r], x[2 * r + 1]);
            }
        } else if constexpr (r < 4) {
            p.r[r] = x[2 * r];
            p.i[r] = x[2 * r + 1];
        } else {
            cmul(p.r[r & 3], p.i[r & 3], x[2 * r], x[2 * r + 1]);
        }
    });
}

// blk += sign * (product of rows 0..N-2) * x[N-1].
template <int N, bool COMP>
__device__ __forceinline__ void finish_term(const double (&x)[2 * N], Prod& p, bool negative, double& br,
                                            double& bi, double& babs)
{
    constexpr int H = N - 1;
    double pr, pi;
    if constexpr (H == 0) {
        pr = 1.0;
        pi = 0.0;
    } else if constexpr (H <= 4) {
        pr = p.r[0];
        pi = p.i[0];
    } else {
        cmul(p.r[0], p.i[0], p.r[1], p.i[1]);
        cmul(p.r[2], p.i[2], p.r[3], p.i[3]);
        cmul(p.r[0], p.i[0], p.r[2], p.i[2]);
        pr = p.r[0];
        pi = p.i[0];
    }
    const double lr = x[2 * N - 2], li = x[2 * N - 1];
    const double sr = negative ? -pr : pr, si = negative ? -pi : pi;
    br = __builtin_fma(sr, lr, br);
    br = __builtin_fma(-si, li, br);
    bi = __builtin_fma(sr, li, bi);
    bi = __builtin_fma(si, lr, bi);
    if constexpr (COMP) {
        // |term| bounded by its L1 norm, |p|_1 |x|_1
        const double pa = __builtin_fabs(pr) + __builtin_fabs(pi);
        const double la = __builtin_fabs(lr) + __builtin_fabs(li);
        babs = __builtin_fma(pa, la, babs);
    }
}

template <int N, bool COMP>
__device__ __forceinline__ void add_term(const double (&x)[2 * N], bool negative, double& br, double& bi,
                                         double& babs)
{
    Prod p;
    prod_rows<N, 0, N>(x, p);
    finish_term<N, COMP>(x, p, negative, br, bi, babs);
}

// One Gray step: x += column `col` (x = fma(s, column, x) when FMA: the
// per-lane direction of the special flip), whose part 0 is in `buf`, then
// the term of the new subset.  Leaves part 0 of `next` loading into `buf`.
template <int N, bool COMP, bool FMA>
__device__ __forceinline__ void gray_step(double (&x)[2 * N], Part<N>& buf, const double* col, double s,
                                          const double* next, bool negative, double& br, double& bi,
                                          double& babs)
{
    Prod p;
    static_for<0, kPhases<N>>([&](auto PP) __attribute__((always_inline)) {
        constexpr int ph = decltype(PP)::value;
        constexpr int r0 = ph * kPhaseRows<N>;
        constexpr int r1 = (r0 + kPhaseRows<N> < N) ? r0 + kPhaseRows<N> : N;
#pragma unroll
        for (int i = 0; i < 2 * (r1 - r0); ++i) {
            if constexpr (FMA) x[2 * r0 + i] = __builtin_fma(s, buf.v[i], x[2 * r0 + i]);
            else x[2 * r0 + i] += buf.v[i];
        }
        fence();
        if constexpr (ph + 1 < kPhases<N>) load_part<N, ph + 1>(buf, col);
        else load_part<N, 0>(buf, next);
        fence();
        prod_rows<N, r0, r1>(x, p);
        if constexpr (ph + 1 == kPhases<N>) finish_term<N, COMP>(x, p, negative, br, bi, babs);
        fence();
    });
}

// One lane's chunk: the 2^L Gray steps starting at k = h 2^L, folded into the
// caller's compensated sums (re_hi + re_lo, im_hi + im_lo; abs_sum bounds
// sum |term| in the L1 norm, COMP only).
template <int N, bool COMP>
__device__ __forceinline__ void walk_c128(const double* __restrict__ M,   // [2][N][2N]: +A, -A
                                          const double* __restrict__ x0,  // 2N doubles
                                          uint64_t h, int L, double& re_hi, double& re_lo, double& im_hi,
                                          double& im_lo, double& abs_sum)
{
    constexpr int C = 2 * N;  // doubles per column
    const double* Mneg = M + (size_t)N * C;
    double x[C];
    stream_column<N, 0>(x, x0, 0.0);
    const uint64_t k0 = h << L;
    const uint64_t g0 = k0 ^ (k0 >> 1);
    for (int j = (L > 0 ? L - 1 : 0); j < N - 1; ++j) {
        const double b = ((g0 >> j) & 1u) ? 1.0 : 0.0;
        stream_column<N, 2>(x, M + (size_t)j * C, b);
    }

    const double lane_sign = (h & 1u) ? -1.0 : 1.0;
    double br = 0.0, bi = 0.0, babs = 0.0;
    auto fold = [&]() __attribute__((always_inline)) {
        if constexpr (COMP) {
            dd_add_d(re_hi, re_lo, br);
            dd_add_d(im_hi, im_lo, bi);
            abs_sum += babs;
        } else {
            re_hi += br;
            im_hi += bi;
        }
        br = bi = babs = 0.0;
    };

    if constexpr (N == 1) {  // L = 0: the single subset
        add_term<N, COMP>(x, false, br, bi, babs);
    } else if constexpr (N == 2) {  // L = 1: steps k = 0, 1; column 0 flips by lane
        add_term<N, COMP>(x, false, br, bi, babs);
        Part<N> buf;
        load_part<N, 0>(buf, M);
        gray_step<N, COMP, true>(x, buf, M, lane_sign, M, true, br, bi, babs);
    } else {
        // 2-step blocks as in dense_f64 (L >= 2): k = 2m flips column
        // ctz(m)+1, k = 2m+1 flips column 0; block nblk/2 (column L-1)
        // flips by lane.
        const uint32_t nblk = 1u << (L - 1);
        const uint32_t half = nblk >> 1;
        auto a_col = [&](uint32_t m) __attribute__((always_inline)) {
            const int tz = __builtin_ctz(m);
            return (((m >> (tz + 1)) & 1u) ? Mneg : M) + (size_t)(tz + 1) * C;
        };
        Part<N> buf;
        add_term<N, COMP>(x, false, br, bi, babs);  // k = 0
        load_part<N, 0>(buf, M);
        const double* pa = a_col(1);
        fence();
        gray_step<N, COMP, false>(x, buf, M, 1.0, pa, true, br, bi, babs);  // k = 1: +column 0
        auto block = [&](uint32_t m, auto SPECIAL) __attribute__((always_inline)) {
            const double* pb = (m & 1u) ? Mneg : M;
            gray_step<N, COMP, decltype(SPECIAL)::value>(x, buf, pa, lane_sign, pb, false, br, bi, babs);
            pa = a_col(m + 1);  // m+1 = nblk: column L <= N-1, a valid address
            gray_step<N, COMP, false>(x, buf, pb, 1.0, pa, true, br, bi, babs);
        };
        constexpr uint32_t G = kFoldSteps / 2;
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
    }
    fold();
}

}  // namespace cplx
}  // namespace spm

#endif
