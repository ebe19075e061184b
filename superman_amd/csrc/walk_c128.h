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
// next to the loop state, so each flip streams its column in 8-double
// chunks with one chunk of lookahead, and the first chunk of the next column
// is prefetched across the current step's products.  Every lambda of the
// walk is force-inlined: an outlined one takes x by reference and sends it to
// scratch memory.
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

// A wave-uniform zero the compiler cannot see through (a volatile scalar
// move).  Offsetting a column address by it keeps that column's scalar loads
// at their use: hoisted as loop invariants they would pin the SGPR file.
__device__ __forceinline__ int opaque_zero()
{
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    return z;
}

__device__ __forceinline__ const const_f64* step_ptr(const double* p)
{
    return (const const_f64*)(p + opaque_zero());
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

// chunk k of a column (doubles [16k, 16k+16) clipped to 2N)
template <int N, int K>
__device__ __forceinline__ void load_chunk(Chunk& c, const double* col)
{
    const const_f64* __restrict__ q = step_ptr(col + K * kChunk);
#pragma unroll
    for (int i = 0; i < kChunk; ++i)
        if (K * kChunk + i < 2 * N) c.v[i] = q[i];
}

// x += s * column (FMA = per-lane direction) or x += column, with chunk 0
// already loaded in c0.
template <int N, bool FMA>
__device__ __forceinline__ void apply_column(double (&x)[2 * N], const Chunk& c0, const double* col, double s)
{
    Chunk cur = c0;
    static_for<0, kChunks<N>>([&](auto KK) __attribute__((always_inline)) {
        constexpr int k = decltype(KK)::value;
        Chunk nxt;
        if constexpr (k + 1 < kChunks<N>) load_chunk<N, k + 1>(nxt, col);
        fence();
#pragma unroll
        for (int i = 0; i < kChunk; ++i) {
            constexpr int base = k * kChunk;
            if (base + i < 2 * N) {
                if constexpr (FMA) x[base + i] = __builtin_fma(s, cur.v[i], x[base + i]);
                else x[base + i] += cur.v[i];
            }
        }
        fence();
        if constexpr (k + 1 < kChunks<N>) cur = nxt;
    });
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

// Product of rows 0..N-2 (four chains), then blk += sign * p * x[N-1].
template <int N, bool COMP>
__device__ __forceinline__ void add_term(const double (&x)[2 * N], bool negative, double& br, double& bi,
                                         double& babs)
{
    constexpr int H = N - 1;
    double pr, pi;
    if constexpr (H == 0) {
        pr = 1.0;
        pi = 0.0;
    } else if constexpr (H <= 4) {
        pr = x[0];
        pi = x[1];
#pragma unroll
        for (int r = 1; r < H; ++r) cmul(pr, pi, x[2 * r], x[2 * r + 1]);
    } else {
        double c0r = x[0], c0i = x[1], c1r = x[2], c1i = x[3];
        double c2r = x[4], c2i = x[5], c3r = x[6], c3i = x[7];
#pragma unroll
        for (int r = 4; r < H; ++r) {
            if ((r & 3) == 0) cmul(c0r, c0i, x[2 * r], x[2 * r + 1]);
            else if ((r & 3) == 1) cmul(c1r, c1i, x[2 * r], x[2 * r + 1]);
            else if ((r & 3) == 2) cmul(c2r, c2i, x[2 * r], x[2 * r + 1]);
            else cmul(c3r, c3i, x[2 * r], x[2 * r + 1]);
        }
        cmul(c0r, c0i, c1r, c1i);
        cmul(c2r, c2i, c3r, c3i);
        cmul(c0r, c0i, c2r, c2i);
        pr = c0r;
        pi = c0i;
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
        Chunk c;
        load_chunk<N, 0>(c, M);
        apply_column<N, true>(x, c, M, lane_sign);
        add_term<N, COMP>(x, true, br, bi, babs);
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
        Chunk ca, cb;
        const double* pa = a_col(1);
        add_term<N, COMP>(x, false, br, bi, babs);  // k = 0
        load_chunk<N, 0>(cb, M);
        apply_column<N, false>(x, cb, M, 1.0);      // k = 1: +column 0
        load_chunk<N, 0>(ca, pa);
        fence();
        add_term<N, COMP>(x, true, br, bi, babs);
        fence();
        auto block = [&](uint32_t m, auto SPECIAL) __attribute__((always_inline)) {
            if constexpr (decltype(SPECIAL)::value) apply_column<N, true>(x, ca, pa, lane_sign);
            else apply_column<N, false>(x, ca, pa, 1.0);
            const double* pb = (m & 1u) ? Mneg : M;
            load_chunk<N, 0>(cb, pb);
            fence();
            add_term<N, COMP>(x, false, br, bi, babs);
            fence();
            apply_column<N, false>(x, cb, pb, 1.0);
            pa = a_col(m + 1);  // m+1 = nblk: column L <= N-1, a valid address
            load_chunk<N, 0>(ca, pa);
            fence();
            add_term<N, COMP>(x, true, br, bi, babs);
            fence();
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
