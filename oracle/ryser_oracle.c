/*
 * ryser_oracle.c -- CPU oracle for the exact-permanent hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file gives the reference answers for tests
 * and the CPU baseline: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it (through oracle/oracle.py).  The product library (superman_amd/csrc)
 * never links, loads or calls anything here.
 *
 * Contents
 *   orc_dense_ref    restatement of the reference CPU dense Gray-code Ryser,
 *                    Algorithms/CPU/Dense/dpNaivePerman.h:45-142 (x init
 *                    :53-65, transpose :67-74, OpenMP chunking :96-98, Gray
 *                    walk :100-133, omp atomic :135-136, final factor :141).
 *   orc_sparse_ref   restatement of Algorithms/CPU/Sparse/spNaivePerman.h:45-172
 *                    on the matrix after IO::ascendingColSort (IO/IO.h:694-738)
 *                    and the SparseMatrix CSC/CSR build
 *                    (Structures/SparseMatrix.h:60-102), which is what
 *                    Permanent::computePermanent (CommonFiles/Permanent.h:64-91)
 *                    feeds it in CPU mode.
 *   orc_complex_ref  restatement of Algorithms/Complex/cpuComputeComplex.h:34-142.
 *   orc_dense_ref_sample / orc_complex_ref_sample
 *                    the same two walks over a prefix [1, 1+count) of the
 *                    Gray indices (bounded CPU-baseline samples).
 *   orc_dense_dd     double-double Ryser (x, products and sums carried as
 *                    unevaluated double pairs); accuracy oracle for FP64.
 *   orc_complex_dd   the same for complex matrices.
 *   orc_int_exact    exact Ryser for integer matrices: y = 2x is integral,
 *                    T = sum_k (-1)^k prod_i y_i is accumulated in 128-bit
 *                    two's-complement (mod 2^128), perm = (-1)^(n-1) T / 2^(n-1).
 *
 * The restatements keep the reference's arithmetic order so that with one
 * thread they reproduce its output bit for bit (the reference binary is only
 * deterministic at thread_count=1: its omp atomic / critical adds are
 * unordered).  Build with -ffp-contract=off so no FMA changes that order.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

#define ORC_MAX_N 63

/* ------------------------------------------------------------------------ */
/* Dense real: dpNaivePerman.h:45-142                                        */
/* ------------------------------------------------------------------------ */
double orc_dense_ref(const double *mat, int nov, int threads)
{
    if (nov < 1 || nov > ORC_MAX_N || threads < 1) return NAN;
    double x[ORC_MAX_N];
    double product = 1;
    for (int i = 0; i < nov; ++i) {                       /* :53-63 */
        double rowSum = 0;
        for (int j = 0; j < nov; ++j) rowSum += mat[i * nov + j];
        x[i] = mat[i * nov + (nov - 1)] - (rowSum / 2);
        product *= x[i];
    }
    double productSum = product;                          /* :64 */

    double *mt = (double *)malloc(sizeof(double) * nov * nov);   /* :66-74 */
    for (int i = 0; i < nov; ++i)
        for (int j = 0; j < nov; ++j) mt[j * nov + i] = mat[i * nov + j];

    const long long start = 1, end = 1LL << (nov - 1);  /* :76-77 */

#pragma omp parallel num_threads(threads)
    {
        int tid = omp_get_thread_num();
        double myResult = 0;
        double myX[ORC_MAX_N];
        memcpy(myX, x, sizeof(double) * nov);

        long long chunk = (end - start + threads - 1) / threads;       /* :96 */
        long long myStart = start + tid * chunk;
        long long myEnd = start + (tid + 1) * chunk;
        if (myEnd > end) myEnd = end;

        long long gray = (myStart - 1) ^ ((myStart - 1) >> 1);         /* :100 */
        for (int j = 0; j < nov - 1; ++j)
            if ((gray >> j) & 1LL)
                for (int i = 0; i < nov; ++i) myX[i] += mt[j * nov + i];

        int productSign = (myStart & 1LL) ? -1 : 1;                   /* :114 */
        for (long long i = myStart; i < myEnd; ++i) {                  /* :116 */
            int col = __builtin_ctzll(i);
            gray ^= (1LL << col);
            double added = ((1LL << col) & gray) ? 1 : -1;
            double myProduct = 1;
            for (int r = 0; r < nov; ++r) {
                myX[r] += added * mt[col * nov + r];
                myProduct *= myX[r];
            }
            myResult += productSign * myProduct;
            productSign *= -1;
        }
#pragma omp atomic
        productSum += myResult;                                        /* :135 */
    }
    free(mt);
    return (4 * (nov % 2) - 2) * productSum;                         /* :141 */
}

/* Bounded sample of the same walk, for the CPU baseline: Gray indices
 * [1, 1 + count) of the matrix, split over `threads` exactly as
 * dpNaivePerman.h:96-98 splits [1, 2^(n-1)).  Returns the (unscaled) partial
 * sum, so the work cannot be optimised away. */
double orc_dense_ref_sample(const double *mat, int nov, int threads, long long count)
{
    if (nov < 2 || nov > ORC_MAX_N || threads < 1 || count < 1) return NAN;
    double x[ORC_MAX_N];
    for (int i = 0; i < nov; ++i) {
        double rowSum = 0;
        for (int j = 0; j < nov; ++j) rowSum += mat[i * nov + j];
        x[i] = mat[i * nov + (nov - 1)] - (rowSum / 2);
    }
    double *mt = (double *)malloc(sizeof(double) * nov * nov);
    for (int i = 0; i < nov; ++i)
        for (int j = 0; j < nov; ++j) mt[j * nov + i] = mat[i * nov + j];
    const long long start = 1;
    long long end = 1LL << (nov - 1);
    if (start + count < end) end = start + count;
    double productSum = 0;
#pragma omp parallel num_threads(threads)
    {
        int tid = omp_get_thread_num();
        double myResult = 0;
        double myX[ORC_MAX_N];
        memcpy(myX, x, sizeof(double) * nov);
        long long chunk = (end - start + threads - 1) / threads;
        long long myStart = start + tid * chunk;
        long long myEnd = start + (tid + 1) * chunk;
        if (myEnd > end) myEnd = end;
        long long gray = (myStart - 1) ^ ((myStart - 1) >> 1);
        for (int j = 0; j < nov - 1; ++j)
            if ((gray >> j) & 1LL)
                for (int i = 0; i < nov; ++i) myX[i] += mt[j * nov + i];
        int productSign = (myStart & 1LL) ? -1 : 1;
        for (long long i = myStart; i < myEnd; ++i) {
            int col = __builtin_ctzll(i);
            gray ^= (1LL << col);
            double added = ((1LL << col) & gray) ? 1 : -1;
            double myProduct = 1;
            for (int r = 0; r < nov; ++r) {
                myX[r] += added * mt[col * nov + r];
                myProduct *= myX[r];
            }
            myResult += productSign * myProduct;
            productSign *= -1;
        }
#pragma omp atomic
        productSum += myResult;
    }
    free(mt);
    return productSum;
}

/* ------------------------------------------------------------------------ */
/* Sparse real: ascendingColSort + SparseMatrix + spNaivePerman              */
/* ------------------------------------------------------------------------ */

/* IO/IO.h:694-738: repeatedly take the (first) column of least nnz count. */
void orc_ascending_col_sort(const double *mat, int nov, double *out)
{
    int deg[ORC_MAX_N];
    for (int j = 0; j < nov; ++j) {
        int d = 0;
        for (int i = 0; i < nov; ++i) d += (mat[i * nov + j] != 0);
        deg[j] = d;
    }
    for (int oc = 0; oc < nov; ++oc) {
        int mind = 0x7fffffff, chosen = 0;
        for (int j = 0; j < nov; ++j)
            if (deg[j] < mind) { mind = deg[j]; chosen = j; }
        deg[chosen] = 0x7fffffff;
        for (int i = 0; i < nov; ++i) out[i * nov + oc] = mat[i * nov + chosen];
    }
}

double orc_sparse_ref(const double *mat_in, int nov, int threads)
{
    if (nov < 1 || nov > ORC_MAX_N || threads < 1) return NAN;
    double *mat = (double *)malloc(sizeof(double) * nov * nov);
    orc_ascending_col_sort(mat_in, nov, mat);

    /* Structures/SparseMatrix.h:60-102 */
    int nnz = 0;
    for (int e = 0; e < nov * nov; ++e) nnz += (mat[e] != 0);
    int *cptrs = (int *)malloc(sizeof(int) * (nov + 1));
    int *rptrs = (int *)malloc(sizeof(int) * (nov + 1));
    int *rows = (int *)malloc(sizeof(int) * (nnz + 1));
    double *cvals = (double *)malloc(sizeof(double) * (nnz + 1));
    double *rvals = (double *)malloc(sizeof(double) * (nnz + 1));
    int rn = 0, cn = 0;
    for (int i = 0; i < nov; ++i) {
        rptrs[i] = rn;
        cptrs[i] = cn;
        for (int j = 0; j < nov; ++j) {
            if (mat[i * nov + j] != 0) rvals[rn++] = mat[i * nov + j];
            if (mat[j * nov + i] != 0) { rows[cn] = j; cvals[cn++] = mat[j * nov + i]; }
        }
    }
    rptrs[nov] = rn;
    cptrs[nov] = cn;

    double x[ORC_MAX_N];                                   /* spNaivePerman.h:61-72 */
    double product = 1;
    for (int i = 0; i < nov; ++i) {
        double rowSum = 0;
        for (int p = rptrs[i]; p < rptrs[i + 1]; ++p) rowSum += rvals[p];
        x[i] = mat[i * nov + (nov - 1)] - (rowSum / 2);
        product *= x[i];
    }
    double productSum = product;
    const long long start = 1, end = 1LL << (nov - 1);

#pragma omp parallel num_threads(threads)
    {
        int tid = omp_get_thread_num();
        double myResult = 0;
        double myX[ORC_MAX_N];
        memcpy(myX, x, sizeof(double) * nov);
        long long chunk = (end - start + threads - 1) / threads;
        long long myStart = start + tid * chunk;
        long long myEnd = start + (tid + 1) * chunk;
        if (myEnd > end) myEnd = end;

        long long gray = (myStart - 1) ^ ((myStart - 1) >> 1);
        for (int j = 0; j < nov - 1; ++j)
            if ((gray >> j) & 1LL)
                for (int p = cptrs[j]; p < cptrs[j + 1]; ++p) myX[rows[p]] += cvals[p];

        int zeroNumber = 0;                                 /* :108-115 */
        for (int i = 0; i < nov; ++i) zeroNumber += (myX[i] == 0);

        int productSign = (myStart & 1LL) ? -1 : 1;
        for (long long i = myStart; i < myEnd; ++i) {       /* :120-165 */
            int col = __builtin_ctzll(i);
            gray ^= (1LL << col);
            double added = ((1LL << col) & gray) ? 1 : -1;
            for (int p = cptrs[col]; p < cptrs[col + 1]; ++p) {
                int r = rows[p];
                double v = myX[r];
                if (v == 0) --zeroNumber;
                v += added * cvals[p];
                if (v == 0) ++zeroNumber;
                myX[r] = v;
            }
            if (zeroNumber == 0) {
                double myProduct = 1;
                for (int r = 0; r < nov; ++r) myProduct *= myX[r];
                myResult += productSign * myProduct;
            }
            productSign *= -1;
        }
#pragma omp atomic
        productSum += myResult;
    }
    free(mat); free(cptrs); free(rptrs); free(rows); free(cvals); free(rvals);
    return (4 * (nov % 2) - 2) * productSum;
}

/* ------------------------------------------------------------------------ */
/* Complex: cpuComputeComplex.h:34-142.  a is row-major interleaved (re,im). */
/* The (+-1, 0) factors of the reference multiply exactly, so they are       */
/* written as sign flips; complex products keep (ac - bd, ad + bc).          */
/* ------------------------------------------------------------------------ */
void orc_complex_ref(const double *a, int nov, int threads, double *out)
{
    out[0] = out[1] = NAN;
    if (nov < 1 || nov > ORC_MAX_N || threads < 1) return;
    double xr[ORC_MAX_N], xi[ORC_MAX_N];
    double pr = 1.0, pi = 0.0;
    for (int i = 0; i < nov; ++i) {                        /* :41-51 */
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < nov; ++j) { sr += a[2 * (i * nov + j)]; si += a[2 * (i * nov + j) + 1]; }
        xr[i] = a[2 * (i * nov + nov - 1)] - sr / 2.0;
        xi[i] = a[2 * (i * nov + nov - 1) + 1] - si / 2.0;
        double tr = pr * xr[i] - pi * xi[i];
        double ti = pr * xi[i] + pi * xr[i];
        pr = tr; pi = ti;
    }
    double sumr = pr, sumi = pi;
    double *mt = (double *)malloc(sizeof(double) * 2 * nov * nov);
    for (int i = 0; i < nov; ++i)
        for (int j = 0; j < nov; ++j) {
            mt[2 * (j * nov + i)] = a[2 * (i * nov + j)];
            mt[2 * (j * nov + i) + 1] = a[2 * (i * nov + j) + 1];
        }
    const long long start = 1, end = 1LL << (nov - 1);

#pragma omp parallel num_threads(threads)
    {
        double rr = 0.0, ri = 0.0;
        double mxr[ORC_MAX_N], mxi[ORC_MAX_N];
        memcpy(mxr, xr, sizeof(double) * nov);
        memcpy(mxi, xi, sizeof(double) * nov);
        long long chunk = (end - start + threads - 1) / threads;   /* :84-87 */
        int tid = omp_get_thread_num();
        long long myStart = start + tid * chunk;
        long long myEnd = start + (tid + 1) * chunk;
        if (myEnd > end) myEnd = end;

        long long gray = (myStart - 1) ^ ((myStart - 1) >> 1);
        for (int j = 0; j < nov - 1; ++j)
            if ((gray >> j) & 1LL)
                for (int i = 0; i < nov; ++i) { mxr[i] += mt[2 * (j * nov + i)]; mxi[i] += mt[2 * (j * nov + i) + 1]; }

        double sgn = (myStart & 1LL) ? -1.0 : 1.0;
        for (long long i = myStart; i < myEnd; ++i) {      /* :103-119 */
            int col = __builtin_ctzll(i);
            gray ^= (1LL << col);
            double added = ((1LL << col) & gray) ? 1.0 : -1.0;
            double qr = 1.0, qi = 0.0;
            for (int r = 0; r < nov; ++r) {
                mxr[r] += added * mt[2 * (col * nov + r)];
                mxi[r] += added * mt[2 * (col * nov + r) + 1];
                double tr = qr * mxr[r] - qi * mxi[r];
                double ti = qr * mxi[r] + qi * mxr[r];
                qr = tr; qi = ti;
            }
            rr += sgn * qr;
            ri += sgn * qi;
            sgn = -sgn;
        }
#pragma omp critical
        {
            sumr += rr;
            sumi += ri;
        }
    }
    free(mt);
    double f = (double)(4 * (nov % 2) - 2);
    out[0] = f * sumr;
    out[1] = f * sumi;
}

/* Bounded sample of the complex walk, for the CPU baseline: Gray indices
 * [1, 1 + count) split over `threads` as cpuComputeComplex.h:84-87 splits
 * [1, 2^(n-1)), the same arithmetic as orc_complex_ref.  out = the unscaled
 * partial sum (re, im), so the work cannot be optimised away. */
void orc_complex_ref_sample(const double *a, int nov, int threads, long long count, double *out)
{
    out[0] = out[1] = NAN;
    if (nov < 2 || nov > ORC_MAX_N || threads < 1 || count < 1) return;
    double xr[ORC_MAX_N], xi[ORC_MAX_N];
    for (int i = 0; i < nov; ++i) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < nov; ++j) { sr += a[2 * (i * nov + j)]; si += a[2 * (i * nov + j) + 1]; }
        xr[i] = a[2 * (i * nov + nov - 1)] - sr / 2.0;
        xi[i] = a[2 * (i * nov + nov - 1) + 1] - si / 2.0;
    }
    double *mt = (double *)malloc(sizeof(double) * 2 * nov * nov);
    for (int i = 0; i < nov; ++i)
        for (int j = 0; j < nov; ++j) {
            mt[2 * (j * nov + i)] = a[2 * (i * nov + j)];
            mt[2 * (j * nov + i) + 1] = a[2 * (i * nov + j) + 1];
        }
    const long long start = 1;
    long long end = 1LL << (nov - 1);
    if (start + count < end) end = start + count;
    double sumr = 0.0, sumi = 0.0;
#pragma omp parallel num_threads(threads)
    {
        double rr = 0.0, ri = 0.0;
        double mxr[ORC_MAX_N], mxi[ORC_MAX_N];
        memcpy(mxr, xr, sizeof(double) * nov);
        memcpy(mxi, xi, sizeof(double) * nov);
        long long chunk = (end - start + threads - 1) / threads;
        int tid = omp_get_thread_num();
        long long myStart = start + tid * chunk;
        long long myEnd = start + (tid + 1) * chunk;
        if (myEnd > end) myEnd = end;
        long long gray = (myStart - 1) ^ ((myStart - 1) >> 1);
        for (int j = 0; j < nov - 1; ++j)
            if ((gray >> j) & 1LL)
                for (int i = 0; i < nov; ++i) { mxr[i] += mt[2 * (j * nov + i)]; mxi[i] += mt[2 * (j * nov + i) + 1]; }
        double sgn = (myStart & 1LL) ? -1.0 : 1.0;
        for (long long i = myStart; i < myEnd; ++i) {
            int col = __builtin_ctzll(i);
            gray ^= (1LL << col);
            double added = ((1LL << col) & gray) ? 1.0 : -1.0;
            double qr = 1.0, qi = 0.0;
            for (int r = 0; r < nov; ++r) {
                mxr[r] += added * mt[2 * (col * nov + r)];
                mxi[r] += added * mt[2 * (col * nov + r) + 1];
                double tr = qr * mxr[r] - qi * mxi[r];
                double ti = qr * mxi[r] + qi * mxr[r];
                qr = tr; qi = ti;
            }
            rr += sgn * qr;
            ri += sgn * qi;
            sgn = -sgn;
        }
#pragma omp critical
        {
            sumr += rr;
            sumi += ri;
        }
    }
    free(mt);
    out[0] = sumr;
    out[1] = sumi;
}

/* ------------------------------------------------------------------------ */
/* Double-double arithmetic (Dekker / Knuth error-free transformations).     */
/* ------------------------------------------------------------------------ */
typedef struct { double hi, lo; } dd;

static inline dd dd_two_sum(double a, double b)
{
    double s = a + b, bb = s - a;
    dd r = { s, (a - (s - bb)) + (b - bb) };
    return r;
}
static inline dd dd_fast_two_sum(double a, double b)
{
    double s = a + b;
    dd r = { s, b - (s - a) };
    return r;
}
static inline dd dd_add(dd a, dd b)
{
    dd s = dd_two_sum(a.hi, b.hi), t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return dd_fast_two_sum(s.hi, s.lo);
}
static inline dd dd_add_d(dd a, double b)
{
    dd s = dd_two_sum(a.hi, b);
    s.lo += a.lo;
    return dd_fast_two_sum(s.hi, s.lo);
}
static inline dd dd_mul(dd a, dd b)
{
    double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return dd_fast_two_sum(p, e);
}
static inline dd dd_neg(dd a) { dd r = { -a.hi, -a.lo }; return r; }

/* Shared chunking for the dd / exact oracles: thread t takes Gray indices
 * [start + t*chunk, ...) exactly as dpNaivePerman.h:96-98 does; per-thread
 * partial sums are reduced in thread order, so the result is deterministic. */
static void orc_chunk(int threads, int tid, long long start, long long end,
                      long long *s, long long *e)
{
    long long chunk = (end - start + threads - 1) / threads;
    *s = start + tid * chunk;
    *e = start + (tid + 1) * chunk;
    if (*e > end) *e = end;
    if (*s > end) *s = end;
}

__attribute__((target_clones("fma", "default")))
void orc_dense_dd(const double *mat, int nov, int threads, double *out_hi, double *out_lo)
{
    *out_hi = *out_lo = NAN;
    if (nov < 1 || nov > ORC_MAX_N || threads < 1) return;
    dd x[ORC_MAX_N];
    dd prod = { 1.0, 0.0 };
    for (int i = 0; i < nov; ++i) {
        dd rs = { 0.0, 0.0 };
        for (int j = 0; j < nov; ++j) rs = dd_add_d(rs, mat[i * nov + j]);
        dd half = { -0.5 * rs.hi, -0.5 * rs.lo };
        x[i] = dd_add_d(half, mat[i * nov + nov - 1]);
        prod = dd_mul(prod, x[i]);
    }
    dd *part = (dd *)calloc((size_t)threads, sizeof(dd));
    const long long start = 1, end = 1LL << (nov - 1);
#pragma omp parallel num_threads(threads)
    {
        int tid = omp_get_thread_num();
        long long s, e;
        orc_chunk(threads, tid, start, end, &s, &e);
        dd mx[ORC_MAX_N];
        memcpy(mx, x, sizeof(dd) * nov);
        long long gray = (s - 1) ^ ((s - 1) >> 1);
        for (int j = 0; j < nov - 1; ++j)
            if ((gray >> j) & 1LL)
                for (int i = 0; i < nov; ++i) mx[i] = dd_add_d(mx[i], mat[i * nov + j]);
        dd acc = { 0.0, 0.0 };
        for (long long k = s; k < e; ++k) {
            int col = __builtin_ctzll(k);
            gray ^= (1LL << col);
            int add = ((1LL << col) & gray) != 0;
            dd p = { 1.0, 0.0 };
            for (int r = 0; r < nov; ++r) {
                double v = mat[r * nov + col];
                mx[r] = dd_add_d(mx[r], add ? v : -v);
                p = dd_mul(p, mx[r]);
            }
            acc = dd_add(acc, (k & 1LL) ? dd_neg(p) : p);
        }
        part[tid] = acc;
    }
    dd sum = prod;
    for (int t = 0; t < threads; ++t) sum = dd_add(sum, part[t]);
    free(part);
    double f = (double)(4 * (nov % 2) - 2);
    *out_hi = f * sum.hi;
    *out_lo = f * sum.lo;
}

__attribute__((target_clones("fma", "default")))
void orc_complex_dd(const double *a, int nov, int threads, double *out)
{
    for (int q = 0; q < 4; ++q) out[q] = NAN;
    if (nov < 1 || nov > ORC_MAX_N || threads < 1) return;
    dd xr[ORC_MAX_N], xi[ORC_MAX_N];
    dd pr = { 1.0, 0.0 }, pi = { 0.0, 0.0 };
    for (int i = 0; i < nov; ++i) {
        dd sr = { 0.0, 0.0 }, si = { 0.0, 0.0 };
        for (int j = 0; j < nov; ++j) {
            sr = dd_add_d(sr, a[2 * (i * nov + j)]);
            si = dd_add_d(si, a[2 * (i * nov + j) + 1]);
        }
        dd hr = { -0.5 * sr.hi, -0.5 * sr.lo }, hi = { -0.5 * si.hi, -0.5 * si.lo };
        xr[i] = dd_add_d(hr, a[2 * (i * nov + nov - 1)]);
        xi[i] = dd_add_d(hi, a[2 * (i * nov + nov - 1) + 1]);
        dd tr = dd_add(dd_mul(pr, xr[i]), dd_neg(dd_mul(pi, xi[i])));
        dd ti = dd_add(dd_mul(pr, xi[i]), dd_mul(pi, xr[i]));
        pr = tr; pi = ti;
    }
    dd *part = (dd *)calloc((size_t)threads * 2, sizeof(dd));
    const long long start = 1, end = 1LL << (nov - 1);
#pragma omp parallel num_threads(threads)
    {
        int tid = omp_get_thread_num();
        long long s, e;
        orc_chunk(threads, tid, start, end, &s, &e);
        dd mr[ORC_MAX_N], mi[ORC_MAX_N];
        memcpy(mr, xr, sizeof(dd) * nov);
        memcpy(mi, xi, sizeof(dd) * nov);
        long long gray = (s - 1) ^ ((s - 1) >> 1);
        for (int j = 0; j < nov - 1; ++j)
            if ((gray >> j) & 1LL)
                for (int i = 0; i < nov; ++i) {
                    mr[i] = dd_add_d(mr[i], a[2 * (i * nov + j)]);
                    mi[i] = dd_add_d(mi[i], a[2 * (i * nov + j) + 1]);
                }
        dd ar = { 0.0, 0.0 }, ai = { 0.0, 0.0 };
        for (long long k = s; k < e; ++k) {
            int col = __builtin_ctzll(k);
            gray ^= (1LL << col);
            int add = ((1LL << col) & gray) != 0;
            dd qr = { 1.0, 0.0 }, qi = { 0.0, 0.0 };
            for (int r = 0; r < nov; ++r) {
                double vr = a[2 * (r * nov + col)], vi = a[2 * (r * nov + col) + 1];
                mr[r] = dd_add_d(mr[r], add ? vr : -vr);
                mi[r] = dd_add_d(mi[r], add ? vi : -vi);
                dd tr = dd_add(dd_mul(qr, mr[r]), dd_neg(dd_mul(qi, mi[r])));
                dd ti = dd_add(dd_mul(qr, mi[r]), dd_mul(qi, mr[r]));
                qr = tr; qi = ti;
            }
            if (k & 1LL) { ar = dd_add(ar, dd_neg(qr)); ai = dd_add(ai, dd_neg(qi)); }
            else { ar = dd_add(ar, qr); ai = dd_add(ai, qi); }
        }
        part[2 * tid] = ar;
        part[2 * tid + 1] = ai;
    }
    dd sr = pr, si = pi;
    for (int t = 0; t < threads; ++t) { sr = dd_add(sr, part[2 * t]); si = dd_add(si, part[2 * t + 1]); }
    free(part);
    double f = (double)(4 * (nov % 2) - 2);
    out[0] = f * sr.hi; out[1] = f * sr.lo;
    out[2] = f * si.hi; out[3] = f * si.lo;
}

/* ------------------------------------------------------------------------ */
/* Exact integer Ryser.                                                      */
/*   y_i(S) = 2 a_{i,n-1} - sum_j a_ij + 2 sum_{j in S} a_ij   (= 2 x_i(S))   */
/*   T = sum_k (-1)^k prod_i y_i(G(k))          (mod 2^128)                  */
/*   perm = 2 (-1)^(n-1) sum_k (-1)^k prod x_i = (-1)^(n-1) T / 2^(n-1)       */
/* Entries must satisfy |a| <= 2^40.  Return codes:                          */
/*   0 exact; 1 the bound 2^(n-1) prod_i sum_j |a_ij| reaches 2^126, so the   */
/*   value is only known mod 2^128; -1 bad input; -2 T not divisible by       */
/*   2^(n-1) (cannot happen for a correct walk: used as a self check).        */
/* out[0], out[1] = low and high 64-bit words of perm (two's complement).    */
/* ------------------------------------------------------------------------ */
/* log2 of a bound on |T| = 2^(n-1) |perm(A)|.  For 0/1 matrices the
 * Bregman-Minc bound perm <= prod_i (r_i!)^(1/r_i); otherwise
 * |perm| <= prod_i sum_j |a_ij|. */
double orc_int_log2_bound(const int64_t *a, int nov)
{
    int binary = 1;
    for (int e = 0; e < nov * nov; ++e)
        if (a[e] != 0 && a[e] != 1) binary = 0;
    double lb = nov - 1;
    for (int i = 0; i < nov; ++i) {
        double rs = 0;
        for (int j = 0; j < nov; ++j) rs += (double)llabs(a[i * nov + j]);
        if (rs == 0) return -INFINITY;
        lb += binary ? lgamma(rs + 1.0) / (rs * log(2.0)) : log2(rs);
    }
    return lb;
}

int orc_int_exact(const int64_t *a, int nov, int threads, int64_t *out)
{
    out[0] = out[1] = 0;
    if (nov < 1 || nov > ORC_MAX_N || threads < 1) return -1;
    const int64_t lim = (int64_t)1 << 40;
    for (int e = 0; e < nov * nov; ++e)
        if (a[e] > lim || a[e] < -lim) return -1;

    int64_t y0[ORC_MAX_N];
    for (int i = 0; i < nov; ++i) {
        int64_t rs = 0;
        for (int j = 0; j < nov; ++j) rs += a[i * nov + j];
        y0[i] = 2 * a[i * nov + nov - 1] - rs;
    }
    u128 *part = (u128 *)calloc((size_t)threads, sizeof(u128));
    const long long start = 0, end = 1LL << (nov - 1);
#pragma omp parallel num_threads(threads)
    {
        int tid = omp_get_thread_num();
        long long s, e;
        orc_chunk(threads, tid, start, end, &s, &e);
        int64_t y[ORC_MAX_N];
        memcpy(y, y0, sizeof(int64_t) * nov);
        u128 acc = 0;
        if (s < e) {
            long long gray = s ^ (s >> 1);
            int zeros = 0;
            for (int j = 0; j < nov - 1; ++j)
                if ((gray >> j) & 1LL)
                    for (int i = 0; i < nov; ++i) y[i] += 2 * a[i * nov + j];
            for (int i = 0; i < nov; ++i) zeros += (y[i] == 0);
            for (long long k = s; k < e; ++k) {
                if (k != s) {
                    int col = __builtin_ctzll(k);
                    gray ^= (1LL << col);
                    int add = ((1LL << col) & gray) != 0;
                    for (int r = 0; r < nov; ++r) {
                        int64_t d = 2 * a[r * nov + col];
                        if (d == 0) continue;
                        zeros -= (y[r] == 0);
                        y[r] += add ? d : -d;
                        zeros += (y[r] == 0);
                    }
                }
                if (zeros == 0) {
                    u128 p = 1;
                    for (int r = 0; r < nov; ++r) p *= (u128)(i128)y[r];
                    acc += (k & 1LL) ? (u128)0 - p : p;
                }
            }
        }
        part[tid] = acc;
    }
    u128 T = 0;
    for (int t = 0; t < threads; ++t) T += part[t];
    free(part);

    u128 mask = (nov - 1 >= 1) ? ((((u128)1) << (nov - 1)) - 1) : 0;
    if (T & mask) return -2;
    i128 perm = ((i128)T) >> (nov - 1);
    if ((nov - 1) & 1) perm = -perm;
    out[0] = (int64_t)(uint64_t)((u128)perm);
    out[1] = (int64_t)(uint64_t)(((u128)perm) >> 64);
    return orc_int_log2_bound(a, nov) >= 126.0 ? 1 : 0;
}

/* Version of the restated reference walks that the Python side checks. */
int orc_abi_version(void) { return 1; }
