// decompose.cpp -- structural preprocessing of sparse matrices before the
// Gray-walk kernels, with the reference's semantics:
//
//   * fine Dulmage-Mendelsohn pruning: an entry that lies in no perfect
//     matching is dropped (IO/IO.h:477-508, IO/dm.cpp:372-501);
//   * connected components of the bipartite row/column graph, each a block
//     whose permanents multiply (IO/dm.h:43-107, main.cpp:73-111);
//   * DecomposePerman (CommonFiles/DecomposePerman.h:119-557) on each block
//     larger than 30: Laplace expansion along a degree-1 row/column
//     (compress1), column merging along a degree-2 row/column (compress2),
//     and the split of a degree-3/4 row into two degree-2 halves whose
//     permanents add (compress34); the leaves' permanents add up to the
//     block's.
//
// None of these changes the permanent; they change what the kernels see
// (the sparse 0/1 n=40 inputs become a handful of blocks of order ~18).
// The search orders (row i before column i, lowest index first; the last
// zero column standing in for a missing fourth nonzero) follow the
// reference so that the leaves are the reference's leaves.
#include "superman_amd.h"

#include "spm_host.h"
#include "spm_kernels.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <queue>
#include <vector>

namespace spm {
namespace {

struct Mat {
    int n = 0;
    std::vector<double> a;  // row-major n x n
    double& at(int i, int j) { return a[(size_t)i * n + j]; }
    double at(int i, int j) const { return a[(size_t)i * n + j]; }
};

int row_nnz(const Mat& m, int i)
{
    int k = 0;
    for (int j = 0; j < m.n; ++j) k += m.at(i, j) != 0.0;
    return k;
}

int col_nnz(const Mat& m, int j)
{
    int k = 0;
    for (int i = 0; i < m.n; ++i) k += m.at(i, j) != 0.0;
    return k;
}

// ---- fine Dulmage-Mendelsohn pruning ---------------------------------------

// Maximum bipartite matching by augmenting paths (iterative DFS).
int max_matching(const std::vector<std::vector<int>>& adj, int n, std::vector<int>& rmatch, std::vector<int>& cmatch)
{
    rmatch.assign(n, -1);
    cmatch.assign(n, -1);
    int size = 0;
    std::vector<int> seen(n, -1), parent_row(n), it(n);
    for (int r0 = 0; r0 < n; ++r0) {
        // DFS over rows along alternating paths; stack of rows
        std::vector<int> stack{r0};
        std::fill(it.begin(), it.end(), 0);
        int found_col = -1;
        while (!stack.empty() && found_col < 0) {
            const int r = stack.back();
            if (it[r] >= (int)adj[r].size()) {
                stack.pop_back();
                continue;
            }
            const int c = adj[r][it[r]++];
            if (seen[c] == r0) continue;
            seen[c] = r0;
            parent_row[c] = r;
            if (cmatch[c] < 0) found_col = c;
            else stack.push_back(cmatch[c]);
        }
        if (found_col < 0) continue;
        // flip the path ending at found_col
        int c = found_col;
        while (c >= 0) {
            const int r = parent_row[c];
            const int next = rmatch[r];
            rmatch[r] = c;
            cmatch[c] = r;
            c = next;
        }
        ++size;
    }
    return size;
}

// Strongly connected components (Tarjan, iterative) of the row digraph.
std::vector<int> scc_rows(const std::vector<std::vector<int>>& g, int n)
{
    std::vector<int> index(n, -1), low(n, 0), comp(n, -1), st;
    std::vector<char> on(n, 0);
    std::vector<std::pair<int, int>> call;
    int counter = 0, ncomp = 0;
    for (int s = 0; s < n; ++s) {
        if (index[s] >= 0) continue;
        call.push_back({s, 0});
        while (!call.empty()) {
            int v = call.back().first;
            int& k = call.back().second;
            if (k == 0) {
                index[v] = low[v] = counter++;
                st.push_back(v);
                on[v] = 1;
            }
            if (k < (int)g[v].size()) {
                const int w = g[v][k++];
                if (index[w] < 0) call.push_back({w, 0});
                else if (on[w]) low[v] = std::min(low[v], index[w]);
                continue;
            }
            if (low[v] == index[v]) {
                int w;
                do {
                    w = st.back();
                    st.pop_back();
                    on[w] = 0;
                    comp[w] = ncomp;
                } while (w != v);
                ++ncomp;
            }
            call.pop_back();
            if (!call.empty()) {
                const int u = call.back().first;
                low[u] = std::min(low[u], low[v]);
            }
        }
    }
    return comp;
}

// Drops every entry outside all perfect matchings; false if there is no
// perfect matching (the permanent is 0).
bool dm_prune(Mat& m)
{
    const int n = m.n;
    std::vector<std::vector<int>> adj(n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (m.at(i, j) != 0.0) adj[i].push_back(j);
    std::vector<int> rmatch, cmatch;
    if (max_matching(adj, n, rmatch, cmatch) != n) return false;
    // row i -> row matched to column j, for every entry (i, j)
    std::vector<std::vector<int>> g(n);
    for (int i = 0; i < n; ++i)
        for (int j : adj[i]) g[i].push_back(cmatch[j]);
    const std::vector<int> comp = scc_rows(g, n);
    for (int i = 0; i < n; ++i)
        for (int j : adj[i])
            if (comp[i] != comp[cmatch[j]]) m.at(i, j) = 0.0;
    return true;
}

// ---- connected components (breadth first, rows before columns) -------------

void components(const Mat& m, std::vector<std::vector<int>>& rows, std::vector<std::vector<int>>& cols)
{
    const int n = m.n;
    std::vector<char> seen(2 * n, 0);
    for (int v = 0; v < 2 * n; ++v) {
        if (seen[v]) continue;
        std::vector<int> rv, cv;
        std::queue<int> q;
        seen[v] = 1;
        q.push(v);
        while (!q.empty()) {
            const int u = q.front();
            q.pop();
            if (u < n) {
                rv.push_back(u);
                for (int c = 0; c < n; ++c)
                    if (m.at(u, c) != 0.0 && !seen[n + c]) {
                        seen[n + c] = 1;
                        q.push(n + c);
                    }
            } else {
                const int c = u - n;
                cv.push_back(c);
                for (int r = 0; r < n; ++r)
                    if (m.at(r, c) != 0.0 && !seen[r]) {
                        seen[r] = 1;
                        q.push(r);
                    }
            }
        }
        rows.push_back(rv);
        cols.push_back(cv);
    }
}

// ---- DecomposePerman ----------------------------------------------------------

// The first row or column (row i before column i) with exactly `deg`
// nonzeros: returns 0 none, 1 row, 2 column; index in *k.
int find_degree(const Mat& m, int deg, int* k)
{
    for (int i = 0; i < m.n; ++i) {
        if (row_nnz(m, i) == deg) {
            *k = i;
            return 1;
        }
        if (col_nnz(m, i) == deg) {
            *k = i;
            return 2;
        }
    }
    return 0;
}

Mat transposed(const Mat& m)
{
    Mat t;
    t.n = m.n;
    t.a.resize(m.a.size());
    for (int i = 0; i < m.n; ++i)
        for (int j = 0; j < m.n; ++j) t.at(j, i) = m.at(i, j);
    return t;
}

// m without row r and column c.
Mat minor_of(const Mat& m, int r, int c)
{
    Mat o;
    o.n = m.n - 1;
    o.a.resize((size_t)o.n * o.n);
    for (int i = 0, oi = 0; i < m.n; ++i) {
        if (i == r) continue;
        for (int j = 0, oj = 0; j < m.n; ++j) {
            if (j == c) continue;
            o.at(oi, oj++) = m.at(i, j);
        }
        ++oi;
    }
    return o;
}

// Degree-1 row/column: perm = v perm(minor), v folded into the minor's first row.
bool compress1(Mat& m)
{
    int k = 0;
    const int kind = find_degree(m, 1, &k);
    if (kind == 0) return false;
    int r = k, c = k;
    if (kind == 1) {
        for (int j = 0; j < m.n; ++j)
            if (m.at(r, j) != 0.0) {
                c = j;
                break;
            }
    } else {
        for (int i = 0; i < m.n; ++i)
            if (m.at(i, c) != 0.0) {
                r = i;
                break;
            }
    }
    const double v = m.at(r, c);
    m = minor_of(m, r, c);
    for (int j = 0; j < m.n; ++j) m.at(0, j) *= v;
    return true;
}

// A row r whose only nonzeros sit in columns `keep` and `drop`: remove row r
// and column `drop`, and replace column `keep` by
// a_r,drop * col_keep + a_r,keep * col_drop (expansion along row r).
Mat merge_row_pair(const Mat& m, int r, int keep, int drop)
{
    Mat o = minor_of(m, r, drop);
    const int ko = keep < drop ? keep : keep - 1;
    for (int i = 0, oi = 0; i < m.n; ++i) {
        if (i == r) continue;
        o.at(oi++, ko) = m.at(i, keep) * m.at(r, drop) + m.at(i, drop) * m.at(r, keep);
    }
    return o;
}

bool compress2(Mat& m)
{
    int k = 0;
    const int kind = find_degree(m, 2, &k);
    if (kind == 0) return false;
    if (kind == 2) {  // a column: the same merge on the transposed matrix
        Mat t = transposed(m);
        int f = -1, s = -1;
        for (int j = 0; j < t.n; ++j)
            if (t.at(k, j) != 0.0) {
                if (f < 0) f = j;
                else {
                    s = j;
                    break;
                }
            }
        m = transposed(merge_row_pair(t, k, f, s));
        return true;
    }
    int f = -1, s = -1;
    for (int j = 0; j < m.n; ++j)
        if (m.at(k, j) != 0.0) {
            if (f < 0) f = j;
            else {
                s = j;
                break;
            }
        }
    m = merge_row_pair(m, k, f, s);
    return true;
}

// A row (or, transposed, a column) with 3 or 4 nonzeros: split it into two
// rows holding two nonzeros each; perm = perm(m1) + perm(m2) with each half
// merged as in compress2 (a missing 4th nonzero is the last zero column).
bool compress34(Mat& m, Mat& m2, int deg)
{
    int k = 0;
    const int kind = find_degree(m, deg, &k);
    if (kind == 0) return false;
    const Mat t = kind == 2 ? transposed(m) : m;
    int nz[4] = {-1, -1, -1, -1}, cnt = 0, zero = -1;
    for (int j = 0; j < t.n; ++j) {
        if (t.at(k, j) != 0.0) nz[cnt++] = j;
        else zero = j;
    }
    if (nz[3] < 0) nz[3] = zero;
    // each half: row k keeps only its pair of columns
    Mat h1 = t, h2 = t;
    for (int j = 0; j < t.n; ++j) {
        if (j != nz[0] && j != nz[1]) h1.at(k, j) = 0.0;
        if (j != nz[2] && j != nz[3]) h2.at(k, j) = 0.0;
    }
    m = merge_row_pair(h1, k, nz[0], nz[1]);
    m2 = merge_row_pair(h2, k, nz[2], nz[3]);
    return true;
}

bool rank_deficient(const Mat& m)
{
    for (int i = 0; i < m.n; ++i)
        if (row_nnz(m, i) == 0 || col_nnz(m, i) == 0) return true;
    return false;
}

int min_degree(const Mat& m)
{
    int d = m.n;
    for (int i = 0; i < m.n; ++i) d = std::min({d, row_nnz(m, i), col_nnz(m, i)});
    return d;
}

constexpr int kCompressAbove = 30;  // DecomposePerman.h:128,156
constexpr int kCompressAboveSmall = 8;  // SPM_DECOMP_SMALL: exercises every compression on small inputs

void recurse(Mat m, int above, std::vector<Mat>& leaves)
{
    const int d = min_degree(m);
    if (d < 5 && m.n > above) {
        if (d == 1) {
            compress1(m);
            recurse(std::move(m), above, leaves);
        } else if (d == 2) {
            compress2(m);
            recurse(std::move(m), above, leaves);
        } else if (d == 3 || d == 4) {
            Mat m2;
            compress34(m, m2, d);
            recurse(std::move(m), above, leaves);
            recurse(std::move(m2), above, leaves);
        }
        // d == 0: a zero row or column, the block contributes nothing
    } else {
        leaves.push_back(std::move(m));
    }
}

void decompose_block(Mat m, bool codegen, int above, std::vector<Mat>& leaves)
{
    bool again = true;
    while (again && m.n > above) {
        again = compress1(m) || compress2(m);
        if (rank_deficient(m)) return;
    }
    if (codegen) leaves.push_back(std::move(m));
    else recurse(std::move(m), above, leaves);
}

}  // namespace

// Sinkhorn-style balancing of one leaf (the reference's IO::scale,
// IO/IO.h:109-161, applied per leaf by DecomposePerman::addQueue,
// CommonFiles/DecomposePerman.h:189-193): `iterations` rounds of row then
// column passes, each setting a scale to target / (scaled line sum) where
// that sum is nonzero, then a scaled in place.  The arithmetic follows the
// reference's so that the scales, and the scaled leaf, are the same
// numbers: line sums accumulate (a_ij * r_i) * c_j left to right, and the
// final pass visits (i, j) in row order, multiplying a_ij by r_i and a_ji
// by c_i -- so an entry below the diagonal gets its column scale first.
void sinkhorn_scale(double* a, int n, int iterations, double target, double* rs, double* cs)
{
    for (int i = 0; i < n; ++i) rs[i] = cs[i] = 1.0;
    for (int it = 0; it < iterations; ++it) {
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) s += a[(size_t)i * n + j] * rs[i] * cs[j];
            if (s != 0.0) rs[i] = target / s;
        }
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int i = 0; i < n; ++i) s += a[(size_t)i * n + j] * rs[i] * cs[j];
            if (s != 0.0) cs[j] = target / s;
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            a[(size_t)i * n + j] *= rs[i];
            a[(size_t)j * n + i] *= cs[i];
        }
}

}  // namespace spm

using namespace spm;

struct spm_decomp {
    int components = 0;
    bool zero = false;                 // structurally zero: the permanent is 0
    std::vector<int> leaf_comp;        // component of each leaf
    std::vector<Mat> leaves;
};

extern "C" int spm_decompose(const double* a, int n, int options, spm_decomp** out)
{
    if (!out) return SPM_EINVAL;
    *out = nullptr;
    if (!a || n < 1 || n > 4096 || (options & ~SPM_DECOMP_ALL_FLAGS)) return SPM_EINVAL;
    for (size_t e = 0; e < (size_t)n * n; ++e)
        if (!std::isfinite(a[e])) return SPM_EINVAL;
    std::unique_ptr<spm_decomp> d(new spm_decomp);
    Mat m;
    m.n = n;
    m.a.assign(a, a + (size_t)n * n);
    if ((options & SPM_DECOMP_DM) && !dm_prune(m)) {
        d->zero = true;
        *out = d.release();
        return SPM_OK;
    }
    std::vector<Mat> blocks;
    const bool codegen = (options & SPM_DECOMP_CODEGEN) != 0;
    if ((options & SPM_DECOMP_COMPONENTS) && !codegen) {
        std::vector<std::vector<int>> rows, cols;
        components(m, rows, cols);
        for (size_t k = 0; k < rows.size(); ++k) {
            if (rows[k].size() != cols[k].size()) {  // main.cpp:88-92
                d->zero = true;
                *out = d.release();
                return SPM_OK;
            }
            Mat b;
            b.n = (int)rows[k].size();
            b.a.resize((size_t)b.n * b.n);
            for (int i = 0; i < b.n; ++i)
                for (int j = 0; j < b.n; ++j) b.at(i, j) = m.at(rows[k][i], cols[k][j]);
            blocks.push_back(std::move(b));
        }
    } else {
        blocks.push_back(m);
    }
    d->components = (int)blocks.size();
    for (size_t k = 0; k < blocks.size(); ++k) {
        const size_t before = d->leaves.size();
        const int above = (options & SPM_DECOMP_SMALL) ? kCompressAboveSmall : kCompressAbove;
        if (options & SPM_DECOMP_COMPRESS) decompose_block(std::move(blocks[k]), codegen, above, d->leaves);
        else d->leaves.push_back(std::move(blocks[k]));
        if (d->leaves.size() == before) {  // a block with permanent 0
            d->zero = true;
            d->leaves.clear();
            d->leaf_comp.clear();
            break;
        }
        d->leaf_comp.resize(d->leaves.size(), (int)k);
    }
    *out = d.release();
    return SPM_OK;
}

extern "C" int spm_decomp_stats(const spm_decomp* d, int* components, int* leaves, double* avg_n, double* avg_nnz,
                                int* structurally_zero)
{
    if (!d) return SPM_EINVAL;
    if (components) *components = d->components;
    if (leaves) *leaves = (int)d->leaves.size();
    double sn = 0, sz = 0;
    for (const Mat& m : d->leaves) {
        sn += m.n;
        for (double v : m.a) sz += v != 0.0;
    }
    const double k = d->leaves.empty() ? 1.0 : (double)d->leaves.size();
    if (avg_n) *avg_n = sn / k;
    if (avg_nnz) *avg_nnz = sz / k;
    if (structurally_zero) *structurally_zero = d->zero ? 1 : 0;
    return SPM_OK;
}

extern "C" int spm_decomp_leaf(const spm_decomp* d, int leaf, int* component, int* n, double* out)
{
    if (!d || leaf < 0 || leaf >= (int)d->leaves.size()) return SPM_EINVAL;
    const Mat& m = d->leaves[leaf];
    if (component) *component = d->leaf_comp[leaf];
    if (n) *n = m.n;
    if (out) std::memcpy(out, m.a.data(), sizeof(double) * m.a.size());
    return SPM_OK;
}

extern "C" int spm_decomp_destroy(spm_decomp* d)
{
    delete d;
    return SPM_OK;
}

// ---- the permanent through the decomposition ----------------------------------

extern "C" int spm_sinkhorn_scale(double* a, int n, int iterations, double target, double* row_scale,
                                  double* col_scale)
{
    if (!a || !row_scale || !col_scale || n < 1 || iterations < 0 || !(target > 0.0) || !std::isfinite(target))
        return SPM_EINVAL;
    for (size_t e = 0; e < (size_t)n * n; ++e)
        if (!std::isfinite(a[e])) return SPM_EINVAL;
    sinkhorn_scale(a, n, iterations, target, row_scale, col_scale);
    return SPM_OK;
}

// The decomposed permanent; scaling_iterations > 0 balances every leaf
// first and divides its permanent by the scales afterwards (the reference's
// DecomposePerman with Settings.scaling, DecomposePerman.h:97-107).
static int permanent_decomposed(const double* a, int n, int options, int precision, int scaling_iterations,
                                double scale_into, int device, void* stream, spm_result* out)
{
    if (!out) return SPM_EINVAL;
    std::memset(out, 0, sizeof *out);
    const double t0 = now_seconds();
    spm_decomp* d = nullptr;
    int st = spm_decompose(a, n, options, &d);
    if (st != SPM_OK) {
        out->status = st;
        return st;
    }
    std::unique_ptr<spm_decomp> guard(d);
    out->n = n;
    if (d->zero) {
        out->is_exact = 1;
        out->status = SPM_OK;
        out->seconds = now_seconds() - t0;
        return SPM_OK;
    }
    // Leaves: exact for integer matrices (while the exact range allows),
    // else FP64 with the requested accumulation; a component's sum stays
    // exact while every leaf of the whole matrix is.
    bool exact = true;
    std::vector<__int128> comp_exact(d->components, 0);
    std::vector<double> comp_val(d->components, 0.0), comp_err(d->components, 0.0);
    double kernel_ms = 0, abs_bound = 0;
    int launches = 0;
    for (size_t k = 0; k < d->leaves.size(); ++k) {
        const Mat& m = d->leaves[k];
        if (m.n > kMaxN) {
            out->status = SPM_EINVAL;  // DecomposePerman.h:185-188
            return SPM_EINVAL;
        }
        // each leaf by the automatic path choice (exact integer walks or
        // generated kernels, row-scaled integers, generated FP64 kernels,
        // the dense FP64 walk)
        const double* leaf = m.a.data();
        std::vector<double> scaled, rs, cs;
        if (scaling_iterations > 0) {
            scaled = m.a;
            rs.resize(m.n);
            cs.resize(m.n);
            sinkhorn_scale(scaled.data(), m.n, scaling_iterations, scale_into, rs.data(), cs.data());
            leaf = scaled.data();
        }
        spm_result r;
        st = spm_permanent_auto(leaf, m.n, precision, device, stream, &r, nullptr);
        if (st != SPM_OK) {
            out->status = st;
            return st;
        }
        if (scaling_iterations > 0) {  // perm(leaf) = perm(scaled) / prod r_i c_i, divided in the reference's order
            for (int i = 0; i < m.n; ++i) {
                r.re /= rs[i];
                r.re /= cs[i];
                if (r.err_est > 0) r.err_est /= std::fabs(rs[i] * cs[i]);
            }
            r.is_exact = 0;
        }
        if (!r.is_exact) exact = false;
        const int c = d->leaf_comp[k];
        if (r.is_exact) {
            const __int128 v = ((__int128)r.exact_hi << 64) | (unsigned __int128)(uint64_t)r.exact_lo;
            if (__builtin_add_overflow(comp_exact[c], v, &comp_exact[c])) exact = false;
        }
        comp_val[c] += r.re;
        comp_err[c] += r.err_est > 0 ? r.err_est : 0.0;
        kernel_ms += r.kernel_ms;
        launches += r.launches;
    }
    __int128 total = 1;
    double value = 1.0;
    for (int c = 0; c < d->components; ++c) {
        if (exact && __builtin_mul_overflow(total, comp_exact[c], &total)) exact = false;
        value *= comp_val[c];
        abs_bound = abs_bound * std::fabs(comp_val[c]) + comp_err[c] * std::fabs(value / (comp_val[c] != 0 ? comp_val[c] : 1.0));
    }
    if (exact) {
        out->is_exact = 1;
        out->exact_lo = (int64_t)(uint64_t)(unsigned __int128)total;
        out->exact_hi = (int64_t)(uint64_t)((unsigned __int128)total >> 64);
        out->re = (double)total;
        out->hi_re = out->re;
        out->lo_re = (double)(total - (__int128)out->re);
        out->err_est = 0.0;
    } else {
        out->re = value;
        out->hi_re = value;
        out->err_est = abs_bound;
    }
    out->status = SPM_OK;
    out->kernel_ms = kernel_ms;
    out->launches = launches;
    out->subsets = 0;
    for (const Mat& m : d->leaves) out->subsets += uint64_t(1) << (m.n - 1);
    out->seconds = now_seconds() - t0;
    return SPM_OK;
}

extern "C" int spm_permanent_decomposed(const double* a, int n, int options, int precision, int device, void* stream,
                                        spm_result* out)
{
    return permanent_decomposed(a, n, options, precision, 0, 1.0, device, stream, out);
}

extern "C" int spm_permanent_decomposed_scaled(const double* a, int n, int options, int precision,
                                               int scaling_iterations, double scale_into, int device, void* stream,
                                               spm_result* out)
{
    if (scaling_iterations < 0 || !(scale_into > 0.0) || !std::isfinite(scale_into)) {
        if (out) {
            std::memset(out, 0, sizeof *out);
            out->status = SPM_EINVAL;
        }
        return SPM_EINVAL;
    }
    return permanent_decomposed(a, n, options, precision, scaling_iterations, scale_into, device, stream, out);
}
