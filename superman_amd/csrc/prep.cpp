// prep.cpp -- host preparation of one matrix for the Gray-walk kernels and
// the conversion of their reduced sums into results; shared by the
// single-matrix plans (capi.cpp) and the batch engine (batch.cpp).
#include "superman_amd.h"

#include "spm_host.h"
#include "spm_kernels.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace spm {

// dpSingleGPU.cuh:62-82 / gpuComputeComplex.cu:212-241: x(empty) =
// a[i][n-1] - rowsum_i / 2 with the row sum taken left to right (re and im
// separately), and the column-major matrix with a negated copy (the kernels
// pick the plane by flip direction).  `w` = doubles per entry (1 real, 2
// complex interleaved); planes = 2 w n^2 doubles, x0 = w n doubles.
void dense_prepare(const double* a, int n, int w, double* planes, double* x0)
{
    const size_t col = (size_t)w * n;
    const size_t plane = col * n;
    for (int i = 0; i < n; ++i) {
        for (int c = 0; c < w; ++c) {
            double rowSum = 0;
            for (int j = 0; j < n; ++j) rowSum += a[(size_t)w * (i * n + j) + c];
            x0[(size_t)w * i + c] = a[(size_t)w * (i * n + n - 1) + c] - (rowSum / 2);
            for (int j = 0; j < n; ++j) {
                const double v = a[(size_t)w * (i * n + j) + c];
                planes[j * col + (size_t)w * i + c] = v;
                planes[plane + j * col + (size_t)w * i + c] = -v;
            }
        }
    }
}

// Integer entry checks: |a| <= 2^30 and absolute row sums < 2^30 (so |y|
// fits int32), y0 = y(empty set), and log2 of the bound on
// |T| = 2^(n-1-e) |perm| (e = rows carried halved, int_row_weights):
// Bregman-Minc for 0/1 matrices, the product of absolute row sums
// otherwise.  The mod-2^128 sum is T itself only when |T| < 2^127 (checked
// with margin for the log arithmetic).
int int_prepare(const int64_t* a, int n, int* y0, double* log2_bound, const unsigned char* w)
{
    if (!a || n < 1 || n > kMaxN) return SPM_EINVAL;
    bool binary = true;
    for (int e = 0; e < n * n; ++e) binary = binary && (a[e] == 0 || a[e] == 1);
    int halved = 0;
    for (int i = 0; w && i < n; ++i) halved += w[i] == 1 ? 1 : 0;
    double lb = n - 1 - halved;
    for (int i = 0; i < n; ++i) {
        int64_t rs = 0, ra = 0;
        for (int j = 0; j < n; ++j) {
            const int64_t v = a[i * n + j];
            if (v > (int64_t(1) << 30) || v < -(int64_t(1) << 30)) return SPM_ERANGE;
            rs += v;
            ra += v < 0 ? -v : v;
        }
        if (ra >= (int64_t(1) << 30)) return SPM_ERANGE;
        y0[i] = (int)(2 * a[i * n + n - 1] - rs);
        if (w && w[i] == 1) y0[i] /= 2;  // an even row sum: y0 is even
        if (ra == 0) lb = -1e300;  // a zero row: perm = 0
        else lb += binary ? std::lgamma((double)ra + 1.0) / ((double)ra * std::log(2.0)) : std::log2((double)ra);
    }
    *log2_bound = lb;
    return lb >= 126.0 ? SPM_ERANGE : SPM_OK;
}

// Row weights of the exact walks: y_i(S) = 2 x_i(S) has the parity of the
// row sum r_i, so for a row with an even sum every y_i is even and the walks
// may carry y_i / 2 instead (weight 1: planes +-a_ij, y0 / 2; weight 2
// otherwise: +-2 a_ij).  Each halved row divides T by 2, so perm =
// (-1)^(n-1) T / 2^(n-1-e) with e halved rows, and the exact 128-bit range
// grows by e bits (half the rows of a typical 0/1 matrix; every row of the
// 4-per-row config-3 inputs).  Returns e.
int int_row_weights(const int64_t* a, int n, unsigned char* w)
{
    int e = 0;
    for (int i = 0; i < n; ++i) {
        int64_t r = 0;
        for (int j = 0; j < n; ++j) r += a[(size_t)i * n + j];
        w[i] = (r & 1) == 0 ? 1 : 2;
        e += w[i] == 1 ? 1 : 0;
    }
    return e;
}

static int64_t row_weight(const unsigned char* w, int i) { return w ? w[i] : 2; }

// +wA and -wA (w = 2 unless a row is halved), column-major int32 (2 n^2 values).
void int_planes(const int64_t* a, int n, int* planes, const unsigned char* w)
{
    const size_t nn = (size_t)n * n;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            planes[(size_t)j * n + i] = (int)(row_weight(w, i) * a[i * n + j]);
            planes[nn + (size_t)j * n + i] = (int)(-row_weight(w, i) * a[i * n + j]);
        }
}

// Columns in ascending order of their nonzero count, ties by index (the
// order of IO::ascendingColSort, IO/IO.h:694-738): the permanent does not
// change, and the Gray walk flips its low columns most often (column 0
// every other step), so the sparse walk does the least work per step.
std::vector<int64_t> ascending_col_sort(const int64_t* a, int n)
{
    std::vector<int> deg(n, 0), order(n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) deg[j] += a[(size_t)i * n + j] != 0 ? 1 : 0;
    for (int j = 0; j < n; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return deg[x] < deg[y]; });
    std::vector<int64_t> b((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) b[(size_t)i * n + j] = a[(size_t)i * n + order[j]];
    return b;
}

// Compressed columns of +wA / -wA as the records of spm_kernels.h.
void int_sparse_blob(const int64_t* a, int n, std::vector<int>& blob, const unsigned char* w)
{
    std::vector<std::vector<int>> rows(n), vals(n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            if (a[(size_t)i * n + j] != 0) {
                rows[j].push_back(i);
                vals[j].push_back((int)(row_weight(w, i) * a[(size_t)i * n + j]));
            }
    int extra = 0;
    for (int j = 0; j < n; ++j) {
        const int k = (int)rows[j].size();
        extra += k > kSparseSlots ? (k - 1) / kSparseSlots : 0;
    }
    const int nrec = 2 * n + 2 * extra;
    blob.assign((size_t)nrec * kSparseRec, 0);
    int next_free = 2 * n;
    for (int j = 0; j < n; ++j)
        for (int plane = 0; plane < 2; ++plane) {
            const int k = (int)rows[j].size();
            int rec = 2 * j + plane;
            for (int e0 = 0;; e0 += kSparseSlots) {
                int* r = blob.data() + (size_t)rec * kSparseRec;
                const int cnt = std::min(k, e0 + kSparseSlots) - e0;
                for (int s = 0; s < cnt; ++s) {
                    r[2 + s] = rows[j][e0 + s];
                    r[2 + kSparseSlots + s] = plane ? -vals[j][e0 + s] : vals[j][e0 + s];
                }
                r[1] = std::max(cnt, 1);  // a zero column: one entry of value 0 on row 0 (slots are zeroed)
                if (e0 + kSparseSlots >= k) {
                    r[0] = -1;
                    break;
                }
                r[0] = next_free;
                rec = next_free++;
            }
        }
}

// Fraction of Gray subsets whose term has no zero row, the terms the
// zero-skipping walks must multiply out.  y_i(S) = 2 a_{i,n-1} - r_i +
// 2 sum_{j in S} a_ij has the parity of the row sum r_i, so a row with an odd
// sum is never zero; a row with an even sum and k entries in the walked
// columns 0..n-2 is zero for about C(k, k/2) / 2^k of the subsets (exact for
// 0/1 rows, an upper estimate for larger entries), independently of the
// other rows to a first approximation.
double int_nonzero_term_fraction(const int64_t* a, int n)
{
    double q = 1.0;
    for (int i = 0; i < n; ++i) {
        int64_t r = 0;
        int k = 0;
        for (int j = 0; j < n; ++j) {
            r += a[(size_t)i * n + j];
            k += (j < n - 1 && a[(size_t)i * n + j] != 0) ? 1 : 0;
        }
        if ((r & 1) != 0 || k == 0) continue;
        double p = 1.0;  // C(k, floor(k/2)) / 2^k
        for (int t = 1; t <= k; ++t) p *= (t <= k / 2 ? double(k - k / 2 + t) / t : 1.0) * 0.5;
        q *= 1.0 - std::min(1.0, p);
    }
    return q;
}

// Seconds per Gray subset of the exact integer walks on MI355X, fitted to
// measured kernel times on 0/1 matrices of orders 28-40 with 3, 4, 6 and 8
// entries per row (within 46 %, 32 % from order 36).  Every walk pays, on
// the fraction q of terms with no zero row, for the 128-bit product:
//  - compressed-column walk: per entry of the flipped column (p = entries
//    per row), more at small orders where chunks are short and their
//    starts weigh more, plus the queued products -- twice as dear at
//    orders 39-41, whose queues are capped below a wave's 64 lanes;
//  - FP64 walk (every absolute row sum below 2^b) and int32 walk: per row.
// With odd row sums (q = 1) the sparse walk loses its advantage: sparse
// 0/1 order 40 with 3 entries per row takes 7.4 s there against 2.5 s in
// the FP64 walk, and 0.21 s with 4 entries per row (q near 0).
IntWalkCosts int_walk_costs(const int64_t* a, int n)
{
    IntWalkCosts c;
    size_t nnz = 0;
    for (size_t e = 0; e < (size_t)n * n; ++e) nnz += a[e] != 0 ? 1 : 0;
    c.per_row = double(nnz) / n;
    c.q = int_nonzero_term_fraction(a, n);
    const double capped = (n >= 39 && n <= 41) ? 2.0 : 1.0;
    c.sparse = 1.13e-13 * c.per_row * (1.0 + 1.44 * std::ldexp(1.0, 28 - n)) + 1.72e-13 * capped * n * c.q;
    c.fp64 = int_small_ok(a, n) ? 6.75e-14 * n + 4.7e-14 * n * c.q : 1e300;
    c.int32 = 9.6e-14 * n + 2.19e-13 * n * c.q;
    return c;
}

// Where the sparse walk (columns in ascending_col_sort order) is the faster
// one: from order 26 (below, a permanent is too small a job for its
// per-chunk start) when the cost model above puts it below the dense exact
// walk that applies (measured on 0/1 matrices with 4 entries per row: 2.0x
// the FP64 walk at order 28, 4.3x at 32, 6.7x at 40, 6.2x on the config-3
// order-40 matrix; with larger entries 2.3x over the int32 walk on
// order-30 curtis54 leaves).
bool int_sparse_preferred(const int64_t* a, int n)
{
    if (n < 26) return false;
    const IntWalkCosts c = int_walk_costs(a, n);
    return c.sparse < std::min(c.fp64, c.int32);
}

// Rows that are scalar multiples of integer vectors: perm(A) = prod_i s_i *
// perm(Q) with a_ij = s_i q_ij exactly (multilinearity in the rows).  s_i
// is the row's smallest absolute entry; every q_ij = a_ij / s_i must be an
// integer of magnitude <= 2^20 that reproduces a_ij exactly, and the
// product of the s_i (kept as a double-double) must stay normal.  False for
// any other row (or an all-zero row, whose permanent the caller gets
// anyway), and for integer matrices (every s_i would be 1 only by accident
// of their entries; those take the exact path directly).
bool row_scaled_integer(const double* a, int n, std::vector<int64_t>& q, double* scale_hi, double* scale_lo)
{
    q.assign((size_t)n * n, 0);
    double sh = 1.0, sl = 0.0;
    bool all_one = true;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) {
            const double v = std::fabs(a[(size_t)i * n + j]);
            if (!std::isfinite(v)) return false;
            if (v != 0.0 && (s == 0.0 || v < s)) s = v;
        }
        if (s == 0.0) return false;
        for (int j = 0; j < n; ++j) {
            const double v = a[(size_t)i * n + j];
            if (v == 0.0) continue;
            const double t = v / s;
            const double r = std::nearbyint(t);
            if (r != t || std::fabs(r) > 1048576.0 || r * s != v) return false;
            q[(size_t)i * n + j] = (int64_t)r;
        }
        all_one = all_one && s == 1.0;
        // (sh, sl) *= s exactly to double-double precision (TwoProd by FMA)
        const double p = sh * s;
        const double e = std::fma(sh, s, -p) + sl * s;
        sh = p + e;
        sl = e - (sh - p);
    }
    if (all_one || !std::isnormal(sh)) return false;
    *scale_hi = sh;
    *scale_lo = sl;
    return true;
}

// value = (scale_hi + scale_lo) * T for an exact 128-bit T, rounded once
// from double-double; fills the result as a (non-integer) real value.
void scaled_int_result(double scale_hi, double scale_lo, const spm_result& t, spm_result* r)
{
    const __int128 T = ((__int128)t.exact_hi << 64) | (unsigned __int128)(uint64_t)t.exact_lo;
    const double th = (double)T;
    const double tl = (double)(T - (__int128)th);
    const double p = scale_hi * th;
    const double e = std::fma(scale_hi, th, -p) + (scale_hi * tl + scale_lo * th);
    const double hi = p + e;
    const double lo = e - (hi - p);
    *r = t;
    r->is_exact = 0;
    r->exact_lo = r->exact_hi = 0;
    r->re = hi;
    r->hi_re = hi;
    r->lo_re = lo;
    r->abs_sum = 0.0;
    r->err_est = std::ldexp(std::fabs(hi), -52);  // the final rounding and the double-double scale
}

// True when every absolute row sum is below 2^int_small_bits(n), so that
// the FP64 integer walk (int_small_kernel, batch kind 4) is exact.
bool int_small_ok(const int64_t* a, int n)
{
    const int64_t lim = int64_t(1) << int_small_bits(n);
    for (int i = 0; i < n; ++i) {
        int64_t ra = 0;
        for (int j = 0; j < n; ++j) ra += a[i * n + j] < 0 ? -a[i * n + j] : a[i * n + j];
        if (ra >= lim) return false;
    }
    return true;
}

// +wA and -wA, column-major doubles (2 n^2 values).
void int_small_planes(const int64_t* a, int n, double* planes, const unsigned char* w)
{
    const size_t nn = (size_t)n * n;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            planes[(size_t)j * n + i] = (double)(row_weight(w, i) * a[i * n + j]);
            planes[nn + (size_t)j * n + i] = (double)(-row_weight(w, i) * a[i * n + j]);
        }
}

// rec = (hi, lo, abs) for real, (re hi, re lo, abs, im hi, im lo) for
// complex: the compensated sum over all 2^(n-1) terms.
void dense_fill_result(int n, bool complex_kind, bool compensated, const double* rec, spm_result* r)
{
    const double f = (double)(4 * (n % 2) - 2);  // 2 (-1)^(n-1), dpSingleGPU.cuh:218
    r->hi_re = f * rec[0];
    r->lo_re = f * rec[1];
    r->re = f * (rec[0] + rec[1]);
    r->abs_sum = 2.0 * rec[2];
    if (complex_kind) {
        r->hi_im = f * rec[3];
        r->lo_im = f * rec[4];
        r->im = f * (rec[3] + rec[4]);
    }
    if (!compensated) r->err_est = -1.0;
    else r->err_est = complex_kind ? complex_err_est(n, r->abs_sum) : dense_err_est(n, r->abs_sum);
    r->subsets = uint64_t(1) << (n - 1);
    r->n = n;
}

// FP32-product runs: every half row sum r_i/2 of the scaled matrix lies in
// [2^c, 2^(c+1)), so |x_i| < 2^(c+1) and a product of n-1 factors stays
// below 2^127 (FP32 range) when (c+1) n <= 127.
int f32prod_scale_exponent(int n)
{
    return 127 / n - 1;
}

// scaled = D a with D = diag(2^-e_i), exact (powers of two) unless a scaled
// entry falls below the FP64 normal range; returns sum e_i, perm(a) =
// 2^(sum e_i) perm(scaled).  Zero rows stay zero (e_i = 0).
int f32prod_scale(const double* a, int n, double* scaled)
{
    const int c = f32prod_scale_exponent(n);
    int total = 0;
    for (int i = 0; i < n; ++i) {
        double r = 0.0;
        for (int j = 0; j < n; ++j) r += a[(size_t)i * n + j];
        int e = 0;
        if (r > 0.0) e = std::ilogb(0.5 * r) - c;
        for (int j = 0; j < n; ++j) scaled[(size_t)i * n + j] = std::ldexp(a[(size_t)i * n + j], -e);
        total += e;
    }
    return total;
}

// Error estimate and scale of an FP32-product result.  A term whose FP32
// product falls below 2^-126 loses relative accuracy: its absolute error is
// below 2^-149 |x_last| < 2^(c+1-149), over 2^(n-1) terms (times the final
// factor 2).
void f32prod_finish(int n, int log2_scale, spm_result* r)
{
    const int c = f32prod_scale_exponent(n);
    r->err_est = f32prod_err_est(n, r->abs_sum) + std::ldexp(1.0, n + c + 1 - 149);
    r->re = std::ldexp(r->re, log2_scale);
    r->hi_re = std::ldexp(r->hi_re, log2_scale);
    r->lo_re = std::ldexp(r->lo_re, log2_scale);
    r->abs_sum = std::ldexp(r->abs_sum, log2_scale);
    r->err_est = std::ldexp(r->err_est, log2_scale);
}

// T mod 2^128 -> perm = (-1)^(n-1) T / 2^(n-1-e) (e halved rows; with
// every row halved, e = n, perm = 2 (-1)^(n-1) T).  T must be a multiple of
// 2^(n-1-e); a violation means a corrupted run and is reported as a HIP
// error.
int int_fill_result(int n, unsigned long long lo, unsigned long long hi, spm_result* r, int e)
{
    const unsigned __int128 T = ((unsigned __int128)hi << 64) | lo;
    const int shift = n - 1 - e;
    __int128 perm;
    if (shift >= 0) {
        const unsigned __int128 mask = shift >= 1 ? (((unsigned __int128)1 << shift) - 1) : 0;
        if (T & mask) return SPM_EHIP;
        perm = ((__int128)T) >> shift;
    } else {
        perm = (__int128)(T << 1);
    }
    if ((n - 1) & 1) perm = -perm;
    r->exact_lo = (int64_t)(uint64_t)(unsigned __int128)perm;
    r->exact_hi = (int64_t)(uint64_t)(((unsigned __int128)perm) >> 64);
    r->is_exact = 1;
    r->re = (double)perm;
    r->hi_re = r->re;
    r->lo_re = (double)(perm - (__int128)r->re);
    r->err_est = 0.0;
    r->subsets = uint64_t(1) << (n - 1);
    r->n = n;
    return SPM_OK;
}

}  // namespace spm
