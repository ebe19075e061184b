// codegen.cpp -- SpaRyser: per-pattern kernels generated and compiled
// in-process with hiprtc.
//
// Replaces the reference's code generation path
// (Algorithms/GPU/CodeGeneration/KernelGenerator.cuh:114-176,
// UTOrderedKernelCodeGen.cuh:34-583, NaiveKernelCodeGen.cuh:34-365) and its
// driver KernelGenSingleGPU (KernelGenSingleGPU.cuh:47-197).  The reference
// writes generated CUDA into its source tree, exits, has a wrapper rebuild
// the whole program, and re-runs it (wrapper.cpp:271-344); here the source
// is compiled by hiprtc inside the call and the code object is cached by a
// hash of the pattern and values.
//
// What is generated, for a matrix whose entries are integers:
//   * the rows after IO::UTOrder (IO/IO.h:926-1002) for the register-efficient
//     variant (column order = ascending remaining in-degree, so the columns
//     flipped most often -- column 0 every other step -- are the sparsest),
//     or the input order for the naive variant;
//   * y = 2x (x on rows with even sums, whose y is always even) kept
//     entirely in VGPRs (at most 63 int32 per lane: the
//     reference's register-area cost model, UTOrderedKernelCodeGen.cuh:68-132,
//     trades registers against global-memory x for CUDA's 255-register
//     limit, which the 512-entry gfx950 file does not need);
//   * for every column c and direction, the exact row updates
//     y_r += +-2 a_rc as integer literals (the reference prints FP literals
//     with std::to_string -- 6 decimals -- which is lossy for general values),
//     with the zero-row count of each lane kept incrementally;
//   * the same wave-uniform Gray walk as the dense kernels: lane chunks of
//     2^L steps, blocks of 2^U steps whose first flip is a runtime column
//     (a uniform switch) and whose other flips are compile-time columns;
//   * the term prod_r y_r (128-bit) only on lanes whose zero count is 0, and
//     128-bit accumulation (mod 2^128; the host proves |T| < 2^127).
// Compiled code objects are kept per (device, hash) in the process and on
// disk (codegen_get_kernel), so a pattern compiles once per machine.
#include "superman_amd.h"

#include "spm_host.h"
#include "spm_kernels.h"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

namespace spm {

// ---------------------------------------------------------------------------
// IO::UTOrder (IO/IO.h:926-1002): repeatedly take the first column of least
// remaining in-degree; rows are numbered in the order they are first touched
// and, once numbered, no longer count towards the in-degrees.
// ---------------------------------------------------------------------------
void ut_order_mask(const unsigned char* nz, int n, std::vector<int>& rowPerm, std::vector<int>& colPerm)
{
    rowPerm.assign(n, 0);
    colPerm.assign(n, 0);
    for (int i = 0; i < n; ++i) rowPerm[i] = colPerm[i] = i;
    const unsigned kDone = 0xffffffffu;
    std::vector<unsigned> indeg(n, 0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) indeg[j] += nz[i * n + j] != 0;
    std::vector<bool> marked(n, false);
    int nextRow = 0;
    for (int step = 0; step < n; ++step) {
        unsigned mind = kDone;
        int chosen = 0;
        for (int j = 0; j < n; ++j)
            if (indeg[j] < mind) {
                mind = indeg[j];
                chosen = j;
            }
        colPerm[chosen] = step;
        indeg[chosen] = kDone;
        for (int i = 0; i < n; ++i) {
            if (nz[i * n + chosen] == 0 || marked[i]) continue;
            marked[i] = true;
            rowPerm[i] = nextRow++;
            for (int j = 0; j < n; ++j)
                if (indeg[j] != kDone && nz[i * n + j] != 0) --indeg[j];
        }
    }
    for (int i = 0; i < n; ++i)
        if (!marked[i]) rowPerm[i] = nextRow++;
}

void ut_order(const int64_t* a, int n, std::vector<int>& rowPerm, std::vector<int>& colPerm)
{
    std::vector<unsigned char> nz((size_t)n * n);
    for (size_t e = 0; e < nz.size(); ++e) nz[e] = a[e] != 0;
    ut_order_mask(nz.data(), n, rowPerm, colPerm);
}

// newMat[rowPerm[i]][colPerm[j]] = mat[i][j] (IO::applyPermutations, IO/IO.h:1095-1112)
std::vector<int64_t> apply_perm(const int64_t* a, int n, const std::vector<int>& rp, const std::vector<int>& cp)
{
    std::vector<int64_t> b((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) b[(size_t)rp[i] * n + cp[j]] = a[i * n + j];
    return b;
}

// log2 of the steps per generated block.  A block's first flip goes
// through a chain of bit tests on the block index, its 2^U - 1 others are
// straight-line code with the column and direction known at compile time,
// so longer blocks run faster and compile longer (2^U copies of the term
// product; about 2.5e-3 2^U n seconds of hiprtc).  Measured on order-40 0/1
// patterns (config 3 / rows of 4 and 5 entries): U = 3 127.6 / 198.1 ms,
// 0.9 s compile; U = 4 100.0 / 187.1 ms, 1.6 s; U = 5 77.1 / 175.8 ms,
// 3.3 s.  Integer kernels: U = 3 up to order 30 (a permanent there takes
// milliseconds, the compile dominates), 4 up to 45, 5 beyond (one walk
// takes tens of seconds).  Kernels whose terms are rarest (gen_compact_term)
// test them once per block, with one copy of the term product, and afford
// U = 5 from order 31 (config 3: 58 ms with a 0.31 s compile, against 86 ms
// and 0.69 s testing each step with the product in a loop, 100 ms and
// 1.58 s inlined at U = 4; rows of 6 entries at order 40: 272 -> 223 ms,
// 1.72 -> 0.40 s).  With more terms the block replays cost more than they
// save (rows of 8 entries, about 3e-4 of the subsets with no zero row:
// 606 -> 1099 ms).  Real kernels form their term from cached
// suffix products, so their code stays small: U = 5 from order 21 (order
// 40, N(0,1) on 4 entries per row: 136 / 102 / 85 ms at U = 3 / 4 / 5,
// 0.25 s compile each).
int gen_unroll_bits_int(int n, bool compact) { return n <= 30 ? 3 : ((compact || n > 45) ? 5 : 4); }
int gen_unroll_bits_real(int n) { return n <= 20 ? 3 : 5; }

// ---------------------------------------------------------------------------
// Source emitter (integer entries).
// ---------------------------------------------------------------------------
namespace {


// Waves per SIMD asked of the register allocator (0: the compiler's
// choice).  Orders 29..40: the natural allocation (100-127 VGPRs) leaves 4;
// at 96 VGPRs the compiler spills only around the term product and the
// final reduction (config-3 order-40 pattern: 155 -> 141 ms).  Orders
// 41..48: 129-130 VGPRs leave 3; 128 spill at most 8 bytes (order-42
// pattern: 706 -> 590 ms); kernels testing their terms once per block
// (RARE) hold one product and take 5 there too (order 44, 6 entries per
// row: 2158 -> 1892 ms; no gain at 6).  Orders up to 28 reach 5 without it.
int gen_waves_per_simd(int n, bool rare)
{
    if (n >= 29 && n <= 40) return 5;
    if (n >= 41 && n <= 48) return rare ? 5 : 4;
    return 0;
}

// Real kernels hold x as FP64 (2 VGPRs per row); 0 = the compiler's choice.
int gen_real_waves_per_simd(int n)
{
    (void)n;
    return 0;
}

std::string lit(int64_t v)
{
    return std::to_string(v);  // integers are exact in decimal
}

// y_r += D, keeping Z = number of zero rows among the rows that can be zero
// (zr[r]; a row with an odd sum never is); wb = the walk's increments (row
// weight x entry).
void emit_flip_const(std::ostringstream& o, const std::vector<int64_t>& wb, const std::vector<char>& zr, int n,
                     int c, int sgn, const char* ind)
{
    for (int r = 0; r < n; ++r) {
        const int64_t d = wb[(size_t)r * n + c] * sgn;
        if (d == 0) continue;
        o << ind << "y" << r << " += " << lit(d) << ";";
        if (zr[r]) o << " Z += (y" << r << " == 0) - (y" << r << " == " << lit(d) << ");";
        o << "\n";
    }
}

// per-lane direction: D = (d ^ mask) - mask with mask = 0 (add) or -1
// (remove) in each lane
void emit_flip_lane(std::ostringstream& o, const std::vector<int64_t>& wb, const std::vector<char>& zr, int n, int c,
                    const char* mask, const char* ind)
{
    for (int r = 0; r < n; ++r) {
        const int64_t d = wb[(size_t)r * n + c];
        if (d == 0) continue;
        o << ind << "{ const int D = (" << lit(d) << " ^ " << mask << ") - " << mask << "; y" << r << " += D;";
        if (zr[r]) o << " Z += (y" << r << " == 0) - (y" << r << " == D);";
        o << " }\n";
    }
}

// The term of the current subset, on the lanes where no row is zero.  The
// test is a wave-wide ballot with a cold-branch hint: on sparse patterns
// almost no wave holds a nonzero term, and a uniform, not-taken branch costs
// far less per step than a lane-divergent region around the product, an
// inlined 128-bit chain over row pairs.  FOLD (steps of kernels whose terms
// are rarest, gen_compact_term) only folds Z into the block's least zero
// count Zm; SIGN_VAR names a run-time sign (true: subtract).
void emit_term(std::ostringstream& o, int n, bool negative, const char* ind, bool fold = false, bool first = false,
               const char* sign_var = nullptr)
{
    if (fold) {
        o << ind << (first ? "Zm = Z;\n" : "Zm = min(Zm, Z);\n");
        return;
    }
    o << ind << "if (__builtin_expect(__builtin_amdgcn_ballot_w64(Z == 0) != 0, 0)) {\n";
    o << ind << "  if (Z == 0) {\n";
    o << ind << "    unsigned __int128 p = (unsigned __int128)(__int128)y0;\n";
    int r = 1;
    for (; r + 1 < n; r += 2)
        o << ind << "    p *= (unsigned __int128)(__int128)((long long)y" << r << " * (long long)y" << r + 1 << ");\n";
    if (r < n) o << ind << "    p *= (unsigned __int128)(__int128)y" << r << ";\n";
    if (sign_var) o << ind << "    acc = " << sign_var << " ? acc - p : acc + p;\n";
    else o << ind << "    acc " << (negative ? "-" : "+") << "= p;\n";
    o << ind << "  }\n";
    o << ind << "}\n";
}

// Compact kernels test no term per step: each step folds Z into the block's
// least zero count Zm (one VALU op, no branch waiting on a compare).  At the
// end of block M, when some lane's Zm is 0, the wave walks the block back
// from its last step (undoing flips 2^U - 1 .. 1 through flipc, a switch
// over columns 0..U-1), forming each subset's term on the way, and then
// forward again to restore its rows.
void emit_block_replay(std::ostringstream& o, int U, const char* M, const char* ind)
{
    const unsigned last = (1u << U) - 1;
    o << ind << "if (__builtin_expect(__builtin_amdgcn_ballot_w64(Zm == 0) != 0, 0)) {\n";
    o << ind << "#pragma nounroll\n";
    o << ind << "  for (unsigned j = " << last << "u;; --j) {\n";
    o << ind << "    termc((j & 1u) != 0);\n";
    o << ind << "    if (j == 0u) break;\n";
    o << ind << "    const unsigned c = __builtin_ctz(j);\n";
    o << ind << "    const unsigned bit = c == " << U - 1 << "u ? (" << M << " & 1u) : ((j >> (c + 1)) & 1u);\n";
    o << ind << "    flipc(c, ~(-(int)bit));\n";
    o << ind << "  }\n";
    o << ind << "#pragma nounroll\n";
    o << ind << "  for (unsigned j = 1u; j <= " << last << "u; ++j) {\n";
    o << ind << "    const unsigned c = __builtin_ctz(j);\n";
    o << ind << "    const unsigned bit = c == " << U - 1 << "u ? (" << M << " & 1u) : ((j >> (c + 1)) & 1u);\n";
    o << ind << "    flipc(c, -(int)bit);\n";
    o << ind << "  }\n";
    o << ind << "}\n";
}

}  // namespace

std::string emit_int_kernel(const std::vector<int64_t>& b, int n, const std::string& name)
{
    const bool compact = gen_compact_term(b.data(), n);
    const int U = gen_unroll_bits_int(n, compact);
    // y = 2x, halved on rows with even sums (int_row_weights): wb holds the
    // walk's increments w_r b_rc
    std::vector<unsigned char> w(n);
    int_row_weights(b.data(), n, w.data());
    std::vector<int64_t> wb(b.size()), y0(n);
    // a row with an odd sum keeps y = 2x odd (2 a_{r,n-1} - r_r + 2 sum_S
    // a_rj), so it is never zero and stays out of Z: its flips are one add
    std::vector<char> zr(n);
    for (int i = 0; i < n; ++i) {
        int64_t rs = 0;
        for (int j = 0; j < n; ++j) {
            rs += b[(size_t)i * n + j];
            wb[(size_t)i * n + j] = w[i] * b[(size_t)i * n + j];
        }
        y0[i] = 2 * b[(size_t)i * n + n - 1] - rs;
        if (w[i] == 1) y0[i] /= 2;
        zr[i] = (rs & 1) == 0 ? 1 : 0;
    }
    std::ostringstream o;
    o << "// generated: SpaRyser kernel for one " << n << "x" << n << " integer pattern\n";
    o << "extern \"C\" __global__ void __launch_bounds__(256)";
    if (const int w = gen_waves_per_simd(n, compact)) o << " __attribute__((amdgpu_waves_per_eu(" << w << ")))";
    o << "\n";
    o << name << "(unsigned long long* __restrict__ partial, unsigned long long task_begin,\n"
      << "    unsigned long long task_end, unsigned long long num_chunks, int L, int accumulate)\n{\n";
    o << "    __shared__ unsigned long long red[8];\n"
         "    const unsigned lane = threadIdx.x & 63u;\n"
         "    const unsigned long long wave = (unsigned long long)blockIdx.x * 4 + (threadIdx.x >> 6);\n"
         "    const unsigned long long nwave = (unsigned long long)gridDim.x * 4;\n"
         "    const bool idle = (unsigned long long)lane >= num_chunks;\n"
         "    unsigned __int128 acc = 0;\n"
         "    for (unsigned long long task = task_begin + wave; task < task_end; task += nwave) {\n"
         "        const unsigned long long h = task * 64 + lane;\n";
    o << "        int ";
    for (int i = 0; i < n; ++i) o << "y" << i << " = " << lit(y0[i]) << (i + 1 < n ? ", " : ";\n");
    o << "        const unsigned long long k0 = h << L;\n"
         "        const unsigned long long g0 = k0 ^ (k0 >> 1);\n";
    // chunk-start rebuild: columns j >= L-1 whose bit is set in G(h 2^L)
    for (int j = 0; j < n - 1; ++j) {
        bool any = false;
        for (int r = 0; r < n; ++r) any = any || b[(size_t)r * n + j] != 0;
        if (!any) continue;
        o << "        { const int mk = -(int)(((g0 >> " << j << ") & 1ull) & (unsigned)(" << j << " >= L - 1));";
        for (int r = 0; r < n; ++r) {
            const int64_t d = wb[(size_t)r * n + j];
            if (d != 0) o << " y" << r << " += " << lit(d) << " & mk;";
        }
        o << " }\n";
    }
    o << "        int Z = 0";
    for (int i = 0; i < n; ++i)
        if (zr[i]) o << " + (y" << i << " == 0)";
    o << ";\n";
    o << "        const int lm = -(int)(h & 1ull);  // direction of column L-1: -1 = remove\n";
    o << "        const unsigned nblk = 1u << (L - " << U << ");\n"
         "        const unsigned half = nblk >> 1;\n";
    if (compact) {
        o << "        int Zm;  // the block's least Z\n";
        o << "        auto flipc = [&](unsigned c, int mk) __attribute__((always_inline)) {\n"
             "            switch (c) {\n";
        for (int c = 0; c < U; ++c) {
            o << "            case " << c << "u:\n";
            emit_flip_lane(o, wb, zr, n, c, "mk", "                ");
            o << "                break;\n";
        }
        o << "            default: break;\n"
             "            }\n"
             "        };\n";
        o << "        auto termc = [&](bool neg) __attribute__((always_inline)) {\n";
        emit_term(o, n, false, "            ", false, false, "neg");
        o << "        };\n";
    }
    // steps j = 1 .. 2^U - 1 of block m: compile-time columns ctz(j); the
    // direction of column U-1 is bit 0 of m
    auto block_tail = [&](const char* ind) {
        for (int j = 1; j < (1 << U); ++j) {
            const int c = __builtin_ctz(j);
            if (c + 1 < U) {
                const int sgn = ((j >> (c + 1)) & 1) ? -1 : 1;
                emit_flip_const(o, wb, zr, n, c, sgn, ind);
            } else {
                o << ind << "if (m & 1u) {\n";
                emit_flip_const(o, wb, zr, n, c, -1, (std::string(ind) + "    ").c_str());
                o << ind << "} else {\n";
                emit_flip_const(o, wb, zr, n, c, 1, (std::string(ind) + "    ").c_str());
                o << ind << "}\n";
            }
            emit_term(o, n, (j & 1) != 0, ind, compact);
        }
        if (compact) emit_block_replay(o, U, "m", ind);
    };
    o << "        {\n            const unsigned m = 0;\n";
    emit_term(o, n, false, "            ", compact, true);
    block_tail("            ");
    o << "        }\n";
    // block m first flips column U + ctz(m), in the block's direction
    // (bit ctz(m)+1 of m) or, at the middle block, each lane's: one mask,
    // so each column's flip is one branch-free path, found by testing the
    // bits of m from the lowest (half the blocks stop at the first test)
    o << "        for (unsigned m = 1; m < nblk; ++m) {\n"
         "            const int lmm = (m == half) ? lm : -(int)((m >> (__builtin_ctz(m) + 1)) & 1u);\n";
    for (int c = U; c < n - 1 && c - U < 32; ++c) {
        o << "            " << (c == U ? "if" : "} else if") << " (m & " << (1u << (c - U)) << "u) {\n";
        emit_flip_lane(o, wb, zr, n, c, "lmm", "                ");
    }
    if (U < n - 1) o << "            }\n";
    emit_term(o, n, false, "            ", compact, true);
    block_tail("            ");
    o << "        }\n";
    o << "    }\n"
         "    if (idle) acc = 0;\n"
         "    unsigned long long lo = (unsigned long long)acc, hi = (unsigned long long)(acc >> 64);\n"
         "    for (int off = 32; off >= 1; off >>= 1) {\n"
         "        const unsigned long long olo = __shfl_xor(lo, off, 64);\n"
         "        const unsigned long long ohi = __shfl_xor(hi, off, 64);\n"
         "        const unsigned long long s = lo + olo;\n"
         "        hi = hi + ohi + (s < lo ? 1ull : 0ull);\n"
         "        lo = s;\n"
         "    }\n"
         "    const int w = threadIdx.x >> 6;\n"
         "    if ((threadIdx.x & 63u) == 0) { red[2 * w] = lo; red[2 * w + 1] = hi; }\n"
         "    __syncthreads();\n"
         "    if (threadIdx.x == 0) {\n"
         "        unsigned __int128 t = 0;\n"
         "        for (int i = 0; i < 4; ++i) t += ((unsigned __int128)red[2 * i + 1] << 64) | red[2 * i];\n"
         "        unsigned long long* out = partial + 2 * blockIdx.x;\n"
         "        if (accumulate) t += ((unsigned __int128)out[1] << 64) | out[0];\n"
         "        out[0] = (unsigned long long)t;\n"
         "        out[1] = (unsigned long long)(t >> 64);\n"
         "    }\n"
         "}\n";
    return o.str();
}

// ---------------------------------------------------------------------------
// Source emitter (real FP64 entries).
//
// The reference's generated kernels compute in FP64 with decimal literals
// printed by std::to_string (6 decimals, UTOrderedKernelCodeGen.cuh:292-295),
// exact only for dyadic values with few digits.  Here every literal is the
// hex-float spelling of the matrix entry (%a), so the kernel adds exactly the
// entry's binary64 value.  Exact zeros of x_i(S) need exact cancellation, so
// generic real matrices get no zero-count skip (integer-valued and row-scaled
// integer matrices take the exact integer kernels instead); what they get is
// a product that does not recompute the rows a flip cannot have changed:
//
//   R(c)  = the largest row any of columns 0..c touches (rows in UT order
//           are numbered as columns first touch them, so R grows slowly);
//   S_j   = prod_{r > R(j)} x_r, j = 0..J-1 (cached "suffix" products);
//   term  = (prod_{r <= R(0)} x_r) * S_0.
//
// Flipping column c changes rows <= R(c) only: S_j for j >= c stays valid,
// S_{c-1} .. S_0 are rebuilt from S_c (one level each), and a column c >= J
// rebuilds S_{J-1} from its rows first.  J is chosen per pattern to minimise
// the expected multiplies per step (column c flips every 2^(c+1) steps).
// ---------------------------------------------------------------------------
namespace {

std::string hexlit(double v)
{
    char buf[48];
    std::snprintf(buf, sizeof buf, "%a", v);
    return buf;
}

struct RealLevels {
    std::vector<int> R;  // R[c], c = 0..n-2 (-1: no row yet)
    int J = 1;
};

RealLevels real_levels(const std::vector<double>& b, int n)
{
    RealLevels lv;
    lv.R.assign(std::max(1, n - 1), -1);
    int cur = -1;
    for (int c = 0; c + 1 < n; ++c) {
        for (int r = 0; r < n; ++r)
            if (b[(size_t)r * n + c] != 0.0) cur = std::max(cur, r);
        lv.R[c] = cur;
    }
    if (n < 2) return lv;
    // expected multiplies per step for J levels (column c flips with
    // probability 2^-(c+1) per step; columns >= J rebuild the top level)
    auto cost = [&](int J) {
        const double R0 = lv.R[0];
        double e = R0 + 1.0;
        for (int c = 1; c < J && c < n - 1; ++c) e += std::ldexp(lv.R[c] - R0 + c, -(c + 1));
        const int top = lv.R[std::min(J, n - 1) - 1];
        e += std::ldexp((n - 1 - top) + (top - R0) + J - 1, -J);
        return e;
    };
    double best = 1e300;
    for (int J = 1; J <= std::min(8, n - 1); ++J) {
        const double e = cost(J);
        if (e < best - 1e-9) {
            best = e;
            lv.J = J;
        }
    }
    return lv;
}

// Product of x_lo .. x_hi (inclusive) as an expression in up to four
// independent chains ("1.0" when empty).
std::string prod_expr(int lo, int hi)
{
    if (hi < lo) return "1.0";
    const int k = hi - lo + 1;
    const int chains = k >= 8 ? 4 : k >= 4 ? 2 : 1;
    std::vector<std::string> ch(chains);
    for (int i = 0; i < k; ++i) {
        std::string& c = ch[i % chains];
        c += (c.empty() ? "" : " * ") + std::string("x") + std::to_string(lo + i);
    }
    std::string e;
    for (int i = 0; i < chains; ++i) e += (i ? " * (" : "(") + ch[i] + ")";
    return e;
}

// sgn 0: the per-lane sign in the variable `sign` (x += sign * d, exact for
// sign = +-1)
void emit_real_flip(std::ostringstream& o, const std::vector<double>& b, int n, int c, int sgn, const char* ind,
                    const char* sign = "sd")
{
    for (int r = 0; r < n; ++r) {
        const double d = b[(size_t)r * n + c];
        if (d == 0.0) continue;
        if (sgn == 0)
            o << ind << "x" << r << " = __builtin_fma(" << sign << ", " << hexlit(d) << ", x" << r << ");\n";
        else
            o << ind << "x" << r << " += " << hexlit(sgn > 0 ? d : -d) << ";\n";
    }
}

// S_{from-1} .. S_0 from S_from; from == J rebuilds S_{J-1} from its rows.
void emit_real_levels(std::ostringstream& o, const RealLevels& lv, int n, int from, const char* ind)
{
    const int J = lv.J;
    if (from >= J) {
        o << ind << "S" << J - 1 << " = " << prod_expr(lv.R[J - 1] + 1, n - 1) << ";\n";
        from = J - 1;
    }
    for (int j = from - 1; j >= 0; --j)
        o << ind << "S" << j << " = " << prod_expr(lv.R[j] + 1, lv.R[j + 1]) << " * S" << j + 1 << ";\n";
}

void emit_real_term(std::ostringstream& o, const RealLevels& lv, bool negative, const char* ind)
{
    o << ind << "{ const double P = " << prod_expr(0, lv.R[0]) << "; blk = __builtin_fma(" << (negative ? "-P" : "P")
      << ", S0, blk); ba = __builtin_fma(__builtin_fabs(P), __builtin_fabs(S0), ba); }\n";
}

}  // namespace

std::string emit_real_kernel(const std::vector<double>& b, int n, bool compensated, const std::string& name)
{
    const int U = gen_unroll_bits_real(n);
    const RealLevels lv = real_levels(b, n);
    std::vector<double> x0(n);
    for (int i = 0; i < n; ++i) {  // x(empty set) with the left-to-right row sum (dpSingleGPU.cuh:62-74)
        double rs = 0.0;
        for (int j = 0; j < n; ++j) rs += b[(size_t)i * n + j];
        x0[i] = b[(size_t)i * n + n - 1] - rs / 2;
    }
    std::ostringstream o;
    o << "// generated: SpaRyser kernel for one " << n << "x" << n << " real matrix (FP64, "
      << (compensated ? "compensated" : "plain") << " sums, " << lv.J << " suffix-product levels)\n";
    o << "static __device__ __forceinline__ void ts(double a, double b, double& s, double& e)\n"
         "{ s = a + b; const double bb = s - a; e = (a - (s - bb)) + (b - bb); }\n"
         "static __device__ __forceinline__ void ddadd(double& ah, double& al, double bh, double bl)\n"
         "{ double s, e; ts(ah, bh, s, e); e += al + bl; ah = s + e; al = e - (ah - s); }\n";
    o << "extern \"C\" __global__ void __launch_bounds__(256)";
    if (const int w = gen_real_waves_per_simd(n)) o << " __attribute__((amdgpu_waves_per_eu(" << w << ")))";
    o << "\n" << name
      << "(double* __restrict__ partial, unsigned long long task_begin,\n"
         "    unsigned long long task_end, unsigned long long num_chunks, int L, int accumulate)\n{\n";
    o << "    __shared__ double red[12];\n"
         "    const unsigned lane = threadIdx.x & 63u;\n"
         "    const unsigned long long wave = (unsigned long long)blockIdx.x * 4 + (threadIdx.x >> 6);\n"
         "    const unsigned long long nwave = (unsigned long long)gridDim.x * 4;\n"
         "    double a_hi = 0.0, a_lo = 0.0, a_abs = 0.0;\n"
         "    for (unsigned long long task = task_begin + wave; task < task_end; task += nwave) {\n"
         "        const unsigned long long h = task * 64 + lane;\n";
    o << "        double ";
    for (int i = 0; i < n; ++i) o << "x" << i << " = " << hexlit(x0[i]) << (i + 1 < n ? ", " : ";\n");
    o << "        const unsigned long long k0 = h << L;\n"
         "        const unsigned long long g0 = k0 ^ (k0 >> 1);\n";
    // chunk-start rebuild in ascending column order (DenseKernelDefinitions.cuh:230-240)
    for (int j = 0; j < n - 1; ++j) {
        bool any = false;
        for (int r = 0; r < n; ++r) any = any || b[(size_t)r * n + j] != 0.0;
        if (!any) continue;
        o << "        { const double bj = ((((g0 >> " << j << ") & 1ull) != 0ull) && (" << j
          << " >= L - 1)) ? 1.0 : 0.0;";
        for (int r = 0; r < n; ++r) {
            const double d = b[(size_t)r * n + j];
            if (d != 0.0) o << " x" << r << " = __builtin_fma(bj, " << hexlit(d) << ", x" << r << ");";
        }
        o << " }\n";
    }
    o << "        double ";
    for (int j = 0; j < lv.J; ++j) o << "S" << j << (j + 1 < lv.J ? ", " : ";\n");
    emit_real_levels(o, lv, n, lv.J, "        ");
    o << "        const double sd = (h & 1ull) ? -1.0 : 1.0;  // direction of column L-1\n"
         "        double t_hi = 0.0, t_lo = 0.0, t_abs = 0.0, blk = 0.0, ba = 0.0;\n";
    o << "        const unsigned nblk = 1u << (L - " << U << ");\n"
         "        const unsigned half = nblk >> 1;\n";
    const std::string fold = compensated
        ? "{ double s_, e_; ts(t_hi, blk, s_, e_); t_hi = s_; t_lo += e_; t_abs += ba; blk = 0.0; ba = 0.0; }\n"
        : "{ t_hi += blk; t_abs += ba; blk = 0.0; ba = 0.0; }\n";
    auto flip = [&](int c, int sgn, const char* ind) {
        emit_real_flip(o, b, n, c, sgn, ind);
        if (c > 0) emit_real_levels(o, lv, n, c, ind);
    };
    auto block_tail = [&](const char* ind) {
        for (int j = 1; j < (1 << U); ++j) {
            const int c = __builtin_ctz(j);
            if (c + 1 < U) {
                flip(c, ((j >> (c + 1)) & 1) ? -1 : 1, ind);
            } else {
                o << ind << "if (m & 1u) {\n";
                flip(c, -1, (std::string(ind) + "    ").c_str());
                o << ind << "} else {\n";
                flip(c, 1, (std::string(ind) + "    ").c_str());
                o << ind << "}\n";
            }
            emit_real_term(o, lv, (j & 1) != 0, ind);
        }
        o << ind << fold;
    };
    o << "        {\n            const unsigned m = 0;\n";
    emit_real_term(o, lv, false, "            ");
    block_tail("            ");
    o << "        }\n";
    // as in the integer kernels: one per-lane sign and a chain of bit tests
    o << "        for (unsigned m = 1; m < nblk; ++m) {\n"
         "            const double sdm = (m == half) ? sd : (((m >> (__builtin_ctz(m) + 1)) & 1u) ? -1.0 : 1.0);\n";
    for (int c = U; c < n - 1 && c - U < 32; ++c) {
        o << "            " << (c == U ? "if" : "} else if") << " (m & " << (1u << (c - U)) << "u) {\n";
        emit_real_flip(o, b, n, c, 0, "                ", "sdm");
        if (c > 0) emit_real_levels(o, lv, n, c, "                ");
    }
    if (U < n - 1) o << "            }\n";
    emit_real_term(o, lv, false, "            ");
    block_tail("            ");
    o << "        }\n";
    o << "        if (h < num_chunks) { ddadd(a_hi, a_lo, t_hi, t_lo); a_abs += t_abs; }\n"
         "    }\n"
         "    for (int off = 32; off >= 1; off >>= 1) {\n"
         "        const double oh = __shfl_xor(a_hi, off, 64), ol = __shfl_xor(a_lo, off, 64);\n"
         "        const double oa = __shfl_xor(a_abs, off, 64);\n"
         "        ddadd(a_hi, a_lo, oh, ol);\n"
         "        a_abs += oa;\n"
         "    }\n"
         "    const int w = threadIdx.x >> 6;\n"
         "    if (lane == 0) { red[3 * w] = a_hi; red[3 * w + 1] = a_lo; red[3 * w + 2] = a_abs; }\n"
         "    __syncthreads();\n"
         "    if (threadIdx.x == 0) {\n"
         "        double hh = red[0], ll = red[1], aa = red[2];\n"
         "        for (int i = 1; i < 4; ++i) { ddadd(hh, ll, red[3 * i], red[3 * i + 1]); aa += red[3 * i + 2]; }\n"
         "        double* out = partial + 4 * blockIdx.x;\n"
         "        if (accumulate) { ddadd(hh, ll, out[0], out[1]); aa += out[2]; }\n"
         "        out[0] = hh; out[1] = ll; out[2] = aa; out[3] = 0.0;\n"
         "    }\n"
         "}\n";
    return o.str();
}

int real_levels_count(const std::vector<double>& b, int n) { return real_levels(b, n).J; }

// FNV-1a over the values' bit patterns and the generator options.
uint64_t real_pattern_hash(const double* b, int n, int options, int precision)
{
    uint64_t h = 1469598103934665603ull ^ 0x5a5a5a5a5a5a5a5aull;
    auto mix = [&](uint64_t v) {
        for (int k = 0; k < 8; ++k) {
            h ^= (v >> (8 * k)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    mix((uint64_t)n);
    mix((uint64_t)options);
    mix((uint64_t)precision);
    mix((uint64_t)gen_unroll_bits_real(n));
    for (int e = 0; e < n * n; ++e) {
        uint64_t w;
        std::memcpy(&w, b + e, 8);
        mix(w);
    }
    return h;
}

int codegen_real_unroll_bits(int n) { return gen_unroll_bits_real(n); }
bool gen_compact_term(const int64_t* b, int n) { return int_nonzero_term_fraction(b, n) < kGenCompactTermQ; }
int codegen_int_unroll_bits(const int64_t* b, int n) { return gen_unroll_bits_int(n, gen_compact_term(b, n)); }

// FNV-1a over the pattern, values and generator options.
uint64_t pattern_hash(const int64_t* b, int n, int options)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) {
        for (int k = 0; k < 8; ++k) {
            h ^= (v >> (8 * k)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    mix((uint64_t)n);
    mix((uint64_t)options);
    mix((uint64_t)codegen_int_unroll_bits(b, n));
    mix((uint64_t)gen_compact_term(b, n));
    for (int e = 0; e < n * n; ++e) mix((uint64_t)b[e]);
    return h;
}

// ---------------------------------------------------------------------------
// hiprtc compile and per-device module cache
// ---------------------------------------------------------------------------
namespace {
struct CachedKernel {
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
    double compile_s = 0;
};
std::mutex g_cg_mu;
std::map<std::pair<int, uint64_t>, CachedKernel> g_cg_cache;
const char* const kRtcOpts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
constexpr int kRtcOptCount = (int)(sizeof kRtcOpts / sizeof kRtcOpts[0]);
}  // namespace

int rtc_compile(const std::string& src, const std::string& name, std::vector<char>& code, std::string& log)
{
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), (name + ".hip").c_str(), 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return SPM_ECOMPILE;
    const hiprtcResult rc = hiprtcCompileProgram(prog, kRtcOptCount, const_cast<const char**>(kRtcOpts));
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    log.assign(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    if (rc != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return SPM_ECOMPILE;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return SPM_OK;
}

int codegen_compile_only(const std::string& src, const std::string& name, double* seconds, size_t* code_bytes)
{
    const double t0 = now_seconds();
    std::vector<char> code;
    std::string log;
    const int st = rtc_compile(src, name, code, log);
    if (st != SPM_OK) std::fprintf(stderr, "superman_amd: hiprtc failed for %s:\n%s\n", name.c_str(), log.c_str());
    if (seconds) *seconds = now_seconds() - t0;
    if (code_bytes) *code_bytes = code.size();
    return st;
}

// ---------------------------------------------------------------------------
// On-disk code objects.  Directory: $SPM_KERNEL_CACHE ("" or "0" turns the
// disk cache off), else $XDG_CACHE_HOME/superman_amd, else
// $HOME/.cache/superman_amd.  A file is <kernel name>-<FNV-1a of the source
// and the compile options>.co, so a changed generator or option never
// reuses an old object; it is written under a temporary name and renamed,
// so a reader sees a whole file or none, and one that fails to load is
// compiled again.  Any filesystem failure only means no disk cache.
// ---------------------------------------------------------------------------
namespace {

std::string disk_cache_dir()
{
    const char* e = std::getenv("SPM_KERNEL_CACHE");
    if (e) return (e[0] == 0 || (e[0] == '0' && e[1] == 0)) ? std::string() : std::string(e);
    if (const char* x = std::getenv("XDG_CACHE_HOME"); x && x[0]) return std::string(x) + "/superman_amd";
    if (const char* h = std::getenv("HOME"); h && h[0]) return std::string(h) + "/.cache/superman_amd";
    return std::string();
}

std::string disk_cache_path(const std::string& src, const std::string& name)
{
    const std::string dir = disk_cache_dir();
    if (dir.empty()) return std::string();
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const char* s) {
        for (const char* c = s; *c; ++c) {
            h ^= (unsigned char)*c;
            h *= 1099511628211ull;
        }
        h ^= 0xff;  // separator
        h *= 1099511628211ull;
    };
    mix(src.c_str());
    for (int k = 0; k < kRtcOptCount; ++k) mix(kRtcOpts[k]);
    char tail[40];
    std::snprintf(tail, sizeof tail, "-%016llx.co", (unsigned long long)h);
    return dir + "/" + name + tail;
}

bool disk_load(const std::string& path, std::vector<char>& code)
{
    if (path.empty()) return false;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    bool ok = sz > 0;
    if (ok) {
        code.resize((size_t)sz);
        ok = std::fread(code.data(), 1, (size_t)sz, f) == (size_t)sz;
    }
    std::fclose(f);
    return ok;
}

void disk_store(const std::string& path, const std::vector<char>& code)
{
    if (path.empty() || code.empty()) return;
    const std::string dir = path.substr(0, path.rfind('/'));
    for (size_t p = 1; p <= dir.size(); ++p)  // mkdir -p
        if (p == dir.size() || dir[p] == '/') (void)::mkdir(dir.substr(0, p).c_str(), 0755);
    char tmp[32];
    std::snprintf(tmp, sizeof tmp, ".tmp%d", (int)::getpid());
    const std::string t = path + tmp;
    FILE* f = std::fopen(t.c_str(), "wb");
    if (!f) return;
    const bool ok = std::fwrite(code.data(), 1, code.size(), f) == code.size();
    if (std::fclose(f) != 0 || !ok || std::rename(t.c_str(), path.c_str()) != 0) (void)std::remove(t.c_str());
}

}  // namespace

bool codegen_is_cached(int device, uint64_t hash, const std::string& src, const std::string& name)
{
    {
        std::lock_guard<std::mutex> lk(g_cg_mu);
        if (g_cg_cache.count({device, hash})) return true;
    }
    const std::string path = disk_cache_path(src, name);
    struct stat st;
    return !path.empty() && ::stat(path.c_str(), &st) == 0 && st.st_size > 0;
}

int codegen_get_kernel(int device, uint64_t hash, const std::string& src, const std::string& name, hipFunction_t* fn,
                       double* compile_s, bool* cached)
{
    std::lock_guard<std::mutex> lk(g_cg_mu);
    auto it = g_cg_cache.find({device, hash});
    if (it != g_cg_cache.end()) {
        *fn = it->second.fn;
        *compile_s = it->second.compile_s;
        *cached = true;
        return SPM_OK;
    }
    const double t0 = now_seconds();
    const std::string path = disk_cache_path(src, name);
    std::vector<char> code;
    CachedKernel ck;
    if (disk_load(path, code) && hipModuleLoadData(&ck.module, code.data()) == hipSuccess) {
        if (hipModuleGetFunction(&ck.fn, ck.module, name.c_str()) == hipSuccess) {
            ck.compile_s = now_seconds() - t0;  // a load, not a compile
            g_cg_cache[{device, hash}] = ck;
            *fn = ck.fn;
            *compile_s = ck.compile_s;
            *cached = false;
            return SPM_OK;
        }
        (void)hipModuleUnload(ck.module);
        ck = CachedKernel();
    }
    code.clear();
    std::string log;
    int st = rtc_compile(src, name, code, log);
    if (st != SPM_OK) {
        std::fprintf(stderr, "superman_amd: hiprtc failed for %s:\n%s\n", name.c_str(), log.c_str());
        return st;
    }
    disk_store(path, code);
    if (hipModuleLoadData(&ck.module, code.data()) != hipSuccess) return SPM_EHIP;
    if (hipModuleGetFunction(&ck.fn, ck.module, name.c_str()) != hipSuccess) return SPM_EHIP;
    ck.compile_s = now_seconds() - t0;
    g_cg_cache[{device, hash}] = ck;
    *fn = ck.fn;
    *compile_s = ck.compile_s;
    *cached = false;
    return SPM_OK;
}

}  // namespace spm
