// walk_intfp.h -- exact Gray-code Ryser for integer matrices with small
// row sums, walked in FP64 (gfx950 device code; used by int_small.hip and
// the batched kernel).
//
// For an integer matrix, y_i(S) = 2 x_i(S) = 2 a_{i,n-1} - sum_j a_ij +
// 2 sum_{j in S} a_ij is an integer with |y_i| <= R_i = sum_j |a_ij|.  When
// every R_i < 2^B with B = bit width of N (all 0/1 matrices qualify), y and
// the +2A / -2A planes are small integers that FP64 holds exactly, so the
// dense walk of walk_f64.h (scalar-loaded columns, full-rate v_add_f64)
// updates y exactly.  A term prod_i y_i is formed in groups of G = 51 / B
// rows -- each group product is an exact FP64 integer below 2^51 -- and the
// NG groups are converted to int64 (the 1.5 * 2^52 addition trick) and
// multiplied mod 2^128 only when no group is zero.  T = sum (-1)^k prod y is
// kept mod 2^128; the host proves |T| < 2^127 beforehand, exactly as for
// the int32 walk (walk_int.h), which remains the path for larger entries.
//
// Per subset: N adds, N - NG multiplies, NG compares; on nonzero terms
// about 3 NG + 6 (NG - 1) + 4 integer ops.  The int32 walk spends a 64-bit
// multiply per row pair and a 128 x 64-bit multiply per pair.
#ifndef SPM_WALK_INTFP_H
#define SPM_WALK_INTFP_H

#include "walk_f64.h"

namespace spm {
namespace real {

// Bit width of N: a 0/1 row sums to at most N < 2^kIntFpBits<N>.
template <int N>
constexpr int kIntFpBits = N < 2 ? 1 : 1 + kIntFpBits<N / 2>;
template <int N>
constexpr int kIntFpGroup = 51 / kIntFpBits<N>;
template <int N>
constexpr int kIntFpGroups = (N + kIntFpGroup<N> - 1) / kIntFpGroup<N>;

// Exact int64 of an integer-valued double |v| < 2^51.
__device__ __forceinline__ long long exact_int64(double v)
{
    return __double_as_longlong(v + 6755399441055744.0) - 0x4338000000000000LL;
}

template <int N>
struct SmallIntTerm {
    static constexpr int G = kIntFpGroup<N>;
    static constexpr int NG = kIntFpGroups<N>;
    unsigned __int128 acc = 0;

    __device__ __forceinline__ void add(const double (&x)[N], bool negative)
    {
        double g[NG];
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const int r0 = q * G;
            const int r1 = (r0 + G < N) ? r0 + G : N;
            double p = x[r0];
#pragma unroll
            for (int r = r0 + 1; r < r1; ++r) p *= x[r];
            g[q] = p;
        }
        bool nz = true;
#pragma unroll
        for (int q = 0; q < NG; ++q) nz = nz && (g[q] != 0.0);
        if (nz) {
            unsigned __int128 p = (unsigned __int128)(__int128)exact_int64(g[0]);
#pragma unroll
            for (int q = 1; q < NG; ++q) p *= (unsigned __int128)(__int128)exact_int64(g[q]);
            acc = negative ? acc - p : acc + p;
        }
    }
    __device__ __forceinline__ void fold() {}
};

// T mod 2^128 over one lane's chunk.  M = [2][N][N] doubles (+2A, -2A),
// y0 = y(empty set) as doubles.
template <int N>
__device__ __forceinline__ unsigned __int128 walk_intfp(const double* __restrict__ M, const double* __restrict__ y0,
                                                        uint64_t h, int L)
{
    SmallIntTerm<N> term;
    walk_gray<N, kDenseUnroll<N>>(M, y0, h, L, term);
    return term.acc;
}

}  // namespace real
}  // namespace spm

#endif
