// capi.cpp -- the C ABI (include/superman_amd.h) for the dense FP64 path:
// host-side preparation, launch geometry, device buffers, launches, results.
//
// Host preparation is the reference's own (dpSingleGPU.cuh:62-82): x(empty)
// with the row sum taken left to right in FP64, and the column-major copy of
// the matrix.  Everything per-subset runs on the GPU.
#include "superman_amd.h"

#include "spm_host.h"
#include "spm_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#ifndef SPM_CHUNK_AMORTIZE
#define SPM_CHUNK_AMORTIZE 1024  // chunk work >= this x its x-rebuild work
#endif


namespace spm {

// ---------------------------------------------------------------------------
// Device properties and occupancy, cached per (device, kernel).
// ---------------------------------------------------------------------------
namespace {
std::mutex g_mu;
struct OccEntry {
    int device;
    const void* kern;
    int block;
    int bpc;
};
std::vector<OccEntry> g_occ;
int g_cus[64] = {0};
}  // namespace

int device_cus(int device, int* cus)
{
    if (device < 0 || device >= 64) return SPM_EINVAL;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_cus[device] > 0) {
            *cus = g_cus[device];
            return SPM_OK;
        }
    }
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || v <= 0)
        return SPM_EHIP;
    std::lock_guard<std::mutex> lk(g_mu);
    g_cus[device] = v;
    *cus = v;
    return SPM_OK;
}

int blocks_per_cu(int device, const void* kern, int block, int* bpc)
{
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for (const auto& e : g_occ)
            if (e.device == device && e.kern == kern && e.block == block) {
                *bpc = e.bpc;
                return SPM_OK;
            }
    }
    int v = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, kern, block, 0) != hipSuccess || v <= 0) return SPM_EHIP;
    std::lock_guard<std::mutex> lk(g_mu);
    g_occ.push_back({device, kern, block, v});
    *bpc = v;
    return SPM_OK;
}

// Scoped hipSetDevice that restores the caller's current device.
DeviceGuard::DeviceGuard(int device)
{
    if (hipGetDevice(&prev_) != hipSuccess) prev_ = -1;
    ok_ = hipSetDevice(device) == hipSuccess;
}
DeviceGuard::~DeviceGuard()
{
    if (prev_ >= 0) (void)hipSetDevice(prev_);
}

double now_seconds()
{
    using namespace std::chrono;
    return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

// Chunk size 2^L per lane.  Small enough that every wave slot gets work
// (tasks >= resident waves), large enough that rebuilding x at a chunk start
// (~n (n-L) FMAs plus the task bookkeeping) is <~0.1% of the chunk's
// 2^L (2n-1) ops (measured at n=40: 2^10-step chunks cost 3% more time).
// Longer chunks still gain a little on the dense real walk (measured
// against 2^14 steps: order 40 with 2^17 1.4% faster, 35 and 36 with 2^15
// and 2^16 2.1% and 1.8%, 37-38 flat) but each factor then drifts over 8x
// more incremental updates: the order-41 transposed pair moved from within
// 1e-10 to 2.0e-10 apart, so the chunk stays.  Shorter chunks to fill the
// last round of tasks when the resident waves are not a power of two (3 or
// 5 per SIMD) measured slower at every order tried (real 30-36, complex
// 26-34): a partial round runs at lower occupancy and costs less than its
// share.
// Where the wave count sets L (small orders), a chunk start -- n columns
// through dependent scalar loads -- still outweighs a few more idle wave
// slots, so the dense walks take longer chunks there (`walk`): real +2
// while that L is at most 5, +1 up to 8; complex +1 up to 5 (measured:
// real n=22 24.1 -> 15.5 us, n=24 35.3 -> 27.2 us, n=26 78.5 -> 73.8 us;
// complex n=22 36.8 -> 32.8 us, n=24 70.8 -> 67.6 us; unchanged or worse
// beyond).  The exact integer walks and the generated kernels keep the
// wave-count chunk: the same rule gained only at order 22 for them and lost
// 5-45% at orders 20, 26 and 28.
void gray_geometry(int n, int unroll_bits, int64_t resident_waves, int* L, uint64_t* chunks, uint64_t* tasks,
                   GrayWalk walk)
{
    const int B = n - 1;
    // L >= 1 whenever there is a column to flip: the kernel takes the term
    // sign (-1)^k from the low L bits, so a chunk must start at an even k;
    // with unrolling, L > U so that block m's bit U exists inside the chunk.
    const int lo = unroll_bits > 0 ? unroll_bits + 1 : (B >= 1 ? 1 : 0);
    int La = lo;
    while (La < B && std::ldexp(1.0, La) * (2.0 * n - 1.0) < double(SPM_CHUNK_AMORTIZE) * n * (double)(n - La)) ++La;
    int lw = 0;
    while ((int64_t(1) << lw) < std::max<int64_t>(resident_waves, 1)) ++lw;
    const int Lp = B - 6 - lw;
    int l = std::max(lo, std::min(La, Lp));
    if (Lp < La) {
        int extra = 0;
        if (walk == GrayWalk::kReal) extra = Lp <= 5 ? 2 : Lp <= 8 ? 1 : 0;
        else if (walk == GrayWalk::kComplex) extra = Lp <= 5 ? 1 : 0;
        l = std::min(La, l + extra);
    }
    l = std::min(l, B);
    *L = l;
    *chunks = uint64_t(1) << (B - l);
    *tasks = (*chunks + kWaveLanes - 1) / kWaveLanes;
}

}  // namespace spm

using namespace spm;

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------
struct spm_plan {
    int kind = 0;  // 1 = dense f64, 2 = dense c128, 3 = exact integer, 4 = generated SpaRyser (integer),
                   // 5 = generated SpaRyser (real FP64)
    int n = 0;
    int precision = 0;
    int device = 0;
    int grid = 0;
    int L = 0;
    uint64_t chunks = 0, tasks = 0;
    int slices = 1;
    int max_grid = 0;      // resident blocks on the device (dense kinds)
    int partial_cap = 0;   // per-block records allocated in d_partial
    int rec = 4;    // doubles per per-block partial record
    int comps = 1;  // double-double components per record (1 real, 2 complex)
    const void* kern = nullptr;
    double* d_M = nullptr;   // dense kinds; the integer kind stores int32 here
    double* d_x0 = nullptr;
    double log2_bound = 0;   // integer kind: log2 of the bound on |T|
    int halved = 0;          // integer kinds: rows carried halved (int_row_weights)
    int log2_scale = 0;      // FP32-product plans: perm = 2^log2_scale perm(scaled)
    double* d_partial = nullptr;
    double* d_out = nullptr;
    double* h_out = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string kname;
    hipFunction_t fn = nullptr;  // generated kernel (kind 4), owned by the module cache
    uint64_t hash = 0;
    double compile_s = 0;
    bool cached = false;
};

static void plan_free(spm_plan* p)
{
    if (!p) return;
    DeviceGuard g(p->device);
    if (p->d_M) (void)hipFree(p->d_M);
    if (p->d_x0) (void)hipFree(p->d_x0);
    if (p->d_partial) (void)hipFree(p->d_partial);
    if (p->d_out) (void)hipFree(p->d_out);
    if (p->h_out) (void)hipHostFree(p->h_out);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    delete p;
}

#define SPM_TRY(expr)                                  \
    do {                                               \
        if ((expr) != hipSuccess) {                    \
            plan_free(p);                              \
            return SPM_EHIP;                           \
        }                                              \
    } while (0)

// Grid and launch slices of a dense plan for its geometry (L, tasks): as many
// blocks as the tasks fill, at most the resident ones; very long runs
// (~> 4 s of VALU work at 64 lanes x 256 CUs x 2.4 GHz) are split into
// several launches that accumulate into the same per-block records.
namespace spm {
void dense_launch_shape(int kind, int n, int precision, uint64_t tasks, int max_grid, int* grid, int* slices)
{
    const uint64_t need_blocks = (tasks + (kDenseBlock / kWaveLanes) - 1) / (kDenseBlock / kWaveLanes);
    *grid = (int)std::min<uint64_t>(uint64_t(max_grid), need_blocks);
    const double ops = precision == SPM_PREC_F32_PRODUCT ? 3.0 * n : kind == 2 ? 6.0 * n : 2.0 * n;
    const double est_s = std::ldexp(1.0, n - 1) * ops / 3.9e13;
    *slices = (int)std::min<double>(std::max(1.0, std::ceil(est_s / 4.0)), (double)tasks);
}
}  // namespace spm

static void dense_grid(spm_plan* p)
{
    dense_launch_shape(p->kind, p->n, p->precision, p->tasks, p->max_grid, &p->grid, &p->slices);
}

// Dense plans, real (kind 1) and complex (kind 2).  `a` is row-major n x n,
// doubles or interleaved (re, im) pairs.
static int make_dense_plan(int kind, const double* a, int n, int precision, int device, spm_plan** out)
{
    if (!out) return SPM_EINVAL;
    *out = nullptr;
    const int maxn = kind == 1 ? kMaxN : kMaxComplexN;
    const int w = kind == 1 ? 1 : 2;  // doubles per entry
    if (!a || n < 1 || n > maxn) return SPM_EINVAL;
    const bool f32p = precision == SPM_PREC_F32_PRODUCT;
    if (precision != SPM_PREC_PLAIN && precision != SPM_PREC_COMPENSATED && !(kind == 1 && f32p))
        return SPM_EINVAL;
    for (int e = 0; e < w * n * n; ++e)
        if (!std::isfinite(a[e]) || (f32p && a[e] < 0.0)) return SPM_EINVAL;
    int cus = 0;
    int st = device_cus(device, &cus);
    if (st != SPM_OK) return st;
    DeviceGuard guard(device);
    if (!guard.ok()) return SPM_EHIP;

    spm_plan* p = new spm_plan;
    p->kind = kind;
    p->n = n;
    p->precision = precision;
    p->device = device;
    const bool comp = precision == SPM_PREC_COMPENSATED;
    p->kern = f32p ? dense_f32p_kernel_ptr(n) : kind == 1 ? dense_f64_kernel_ptr(n, comp) : dense_c128_kernel_ptr(n, comp);
    p->rec = kind == 1 ? 4 : 8;
    p->comps = kind == 1 ? 1 : 2;
    if (!p->kern) {
        plan_free(p);
        return SPM_EINVAL;
    }
    int bpc = 0;
    st = blocks_per_cu(device, p->kern, kDenseBlock, &bpc);
    if (st != SPM_OK) {
        plan_free(p);
        return st;
    }
    const int64_t waves = int64_t(bpc) * cus * (kDenseBlock / kWaveLanes);
    const int ub = kind == 1 ? dense_f64_unroll_bits(n) : (n >= 3 ? 1 : 0);
    gray_geometry(n, ub, waves, &p->L, &p->chunks, &p->tasks, kind == 1 ? GrayWalk::kReal : GrayWalk::kComplex);
    p->max_grid = bpc * cus;
    dense_grid(p);
    char kn[96];
    if (f32p)
        std::snprintf(kn, sizeof kn, "dense_f32p_kernel<%d, %d>", n, ub);
    else if (kind == 1)
        std::snprintf(kn, sizeof kn, "dense_f64_kernel<%d, %d, %s>", n, ub, comp ? "true" : "false");
    else
        std::snprintf(kn, sizeof kn, "dense_c128_kernel<%d, %s>", n, comp ? "true" : "false");
    p->kname = kn;

    // Host preparation (dpSingleGPU.cuh:62-82; gpuComputeComplex.cu:212-241).
    const size_t col = (size_t)w * n;   // doubles per column
    const size_t plane = col * n;
    std::vector<double> colmaj(2 * plane), x0(col), scaled;
    if (f32p) {
        scaled.resize((size_t)n * n);
        p->log2_scale = f32prod_scale(a, n, scaled.data());
        a = scaled.data();
    }
    dense_prepare(a, n, w, colmaj.data(), x0.data());
    SPM_TRY(hipMalloc(&p->d_M, sizeof(double) * 2 * plane));
    SPM_TRY(hipMalloc(&p->d_x0, sizeof(double) * col));
    SPM_TRY(hipMalloc(&p->d_partial, sizeof(double) * p->rec * p->grid));
    p->partial_cap = p->grid;
    SPM_TRY(hipMalloc(&p->d_out, sizeof(double) * 6));
    SPM_TRY(hipHostMalloc(&p->h_out, sizeof(double) * 6, hipHostMallocDefault));
    SPM_TRY(hipMemcpy(p->d_M, colmaj.data(), sizeof(double) * 2 * plane, hipMemcpyHostToDevice));
    SPM_TRY(hipMemcpy(p->d_x0, x0.data(), sizeof(double) * col, hipMemcpyHostToDevice));
    SPM_TRY(hipEventCreate(&p->ev0));
    SPM_TRY(hipEventCreate(&p->ev1));
    *out = p;
    return SPM_OK;
}

// Integer checks and y(empty set) (prep.cpp), into a vector; w = row
// weights (nullptr: no halved rows).
static int int_prepare(const int64_t* a, int n, std::vector<int>& y0, double* log2_bound,
                       const unsigned char* w = nullptr)
{
    if (!a || n < 1 || n > kMaxN) return SPM_EINVAL;
    y0.assign(n, 0);
    return spm::int_prepare(a, n, y0.data(), log2_bound, w);
}

// Exact integer plans (kind 3): y = 2x in int32, T mod 2^128.
// walk: SPM_INT_WALK_AUTO picks the FP64 walk when every absolute row sum is
// below 2^int_small_bits(n) (int_small_ok), else the int32 walk.
static int make_int_plan(const int64_t* a, int n, int walk, int device, spm_plan** out)
{
    if (!out) return SPM_EINVAL;
    *out = nullptr;
    if (walk != SPM_INT_WALK_AUTO && walk != SPM_INT_WALK_INT32 && walk != SPM_INT_WALK_FP64 &&
        walk != SPM_INT_WALK_SPARSE)
        return SPM_EINVAL;
    if (!a || n < 1 || n > kMaxN) return SPM_EINVAL;
    // rows with even sums carried halved (int_row_weights); the column sort
    // below keeps every row sum
    std::vector<unsigned char> w(n);
    const int halved = int_row_weights(a, n, w.data());
    std::vector<int> y0;
    double lb = 0;
    int rc = int_prepare(a, n, y0, &lb, w.data());
    if (rc != SPM_OK) return rc;
    const bool small = int_small_ok(a, n);
    if (walk == SPM_INT_WALK_FP64 && !small) return SPM_ERANGE;
    const bool sp = walk == SPM_INT_WALK_SPARSE || (walk == SPM_INT_WALK_AUTO && int_sparse_preferred(a, n));
    const bool fp = !sp && (walk == SPM_INT_WALK_FP64 || (walk == SPM_INT_WALK_AUTO && small));
    std::vector<int64_t> sorted;
    if (sp) {  // sparsest columns first: flipped most often (ascending_col_sort)
        sorted = ascending_col_sort(a, n);
        a = sorted.data();
        rc = int_prepare(a, n, y0, &lb, w.data());
        if (rc != SPM_OK) return rc;
    }
    int cus = 0;
    int st = device_cus(device, &cus);
    if (st != SPM_OK) return st;
    DeviceGuard guard(device);
    if (!guard.ok()) return SPM_EHIP;

    spm_plan* p = new spm_plan;
    p->kind = 3;
    p->n = n;
    p->precision = -1;
    p->device = device;
    p->log2_bound = lb;
    p->halved = halved;
    p->kern = sp ? int_sparse_kernel_ptr(n) : fp ? int_small_kernel_ptr(n) : int_dense_kernel_ptr(n);
    p->rec = 2;
    p->comps = 0;
    if (!p->kern) {
        plan_free(p);
        return SPM_EINVAL;
    }
    int bpc = 0;
    st = blocks_per_cu(device, p->kern, kDenseBlock, &bpc);
    if (st != SPM_OK) {
        plan_free(p);
        return st;
    }
    const int64_t waves = int64_t(bpc) * cus * (kDenseBlock / kWaveLanes);
    gray_geometry(n, fp ? dense_f64_unroll_bits(n) : (n >= 3 ? 1 : 0), waves, &p->L, &p->chunks, &p->tasks);
    const uint64_t need_blocks = (p->tasks + (kDenseBlock / kWaveLanes) - 1) / (kDenseBlock / kWaveLanes);
    p->grid = (int)std::min<uint64_t>(uint64_t(bpc) * cus, need_blocks);
    const double est_s = std::ldexp(1.0, n - 1) * ((fp ? 5.0 : 12.0) * n) / 3.9e13;
    p->slices = (int)std::min<double>(std::max(1.0, std::ceil(est_s / 4.0)), (double)p->tasks);
    char kn[96];
    std::snprintf(kn, sizeof kn, sp ? "int_sparse_kernel<%d>" : fp ? "int_small_kernel<%d>" : "int_dense_kernel<%d>",
                  n);
    p->kname = kn;
    const size_t nn = (size_t)n * n;
    if (sp) {
        std::vector<int> blob;
        int_sparse_blob(a, n, blob, w.data());
        SPM_TRY(hipMalloc(&p->d_M, sizeof(int) * blob.size()));
        std::vector<int> ys(y0.begin(), y0.end());
        ys.push_back(small ? 1 : 0);  // products in FP64 groups (int_sparse_kernel)
        SPM_TRY(hipMalloc(&p->d_x0, sizeof(int) * ys.size()));
        SPM_TRY(hipMemcpy(p->d_M, blob.data(), sizeof(int) * blob.size(), hipMemcpyHostToDevice));
        SPM_TRY(hipMemcpy(p->d_x0, ys.data(), sizeof(int) * ys.size(), hipMemcpyHostToDevice));
    } else if (fp) {
        std::vector<double> colmaj(2 * nn), yd(y0.begin(), y0.end());
        int_small_planes(a, n, colmaj.data(), w.data());
        SPM_TRY(hipMalloc(&p->d_M, sizeof(double) * 2 * nn));
        SPM_TRY(hipMalloc(&p->d_x0, sizeof(double) * n));
        SPM_TRY(hipMemcpy(p->d_M, colmaj.data(), sizeof(double) * 2 * nn, hipMemcpyHostToDevice));
        SPM_TRY(hipMemcpy(p->d_x0, yd.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    } else {
        std::vector<int> colmaj(2 * nn);
        int_planes(a, n, colmaj.data(), w.data());
        SPM_TRY(hipMalloc(&p->d_M, sizeof(int) * 2 * nn));
        SPM_TRY(hipMalloc(&p->d_x0, sizeof(int) * n));
        SPM_TRY(hipMemcpy(p->d_M, colmaj.data(), sizeof(int) * 2 * nn, hipMemcpyHostToDevice));
        SPM_TRY(hipMemcpy(p->d_x0, y0.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    }
    SPM_TRY(hipMalloc(&p->d_partial, sizeof(unsigned long long) * 2 * p->grid));
    SPM_TRY(hipMalloc(&p->d_out, sizeof(double) * 6));
    SPM_TRY(hipHostMalloc(&p->h_out, sizeof(double) * 6, hipHostMallocDefault));
    SPM_TRY(hipEventCreate(&p->ev0));
    SPM_TRY(hipEventCreate(&p->ev1));
    *out = p;
    return SPM_OK;
}

extern "C" int spm_plan_dense_f64(const double* a, int n, int precision, int device, spm_plan** out)
{
    return make_dense_plan(1, a, n, precision, device, out);
}

extern "C" int spm_plan_int(const int64_t* a, int n, int device, spm_plan** out)
{
    return make_int_plan(a, n, SPM_INT_WALK_AUTO, device, out);
}

extern "C" int spm_plan_int_walk(const int64_t* a, int n, int walk, int device, spm_plan** out)
{
    return make_int_plan(a, n, walk, device, out);
}

// Generated SpaRyser plans (kind 4).
static std::vector<int64_t> codegen_matrix(const int64_t* a, int n, int options)
{
    if (options & SPM_CG_REGISTER_EFFICIENT) {
        std::vector<int> rp, cp;
        ut_order(a, n, rp, cp);
        return apply_perm(a, n, rp, cp);
    }
    return std::vector<int64_t>(a, a + (size_t)n * n);
}

static std::string codegen_name(uint64_t h)
{
    char nm[40];
    std::snprintf(nm, sizeof nm, "spm_gen_%016llx", (unsigned long long)h);
    return nm;
}

extern "C" int spm_codegen_source(const int64_t* a, int n, int options, char* buf, size_t buflen, size_t* needed)
{
    if (!a || n < kCodegenMinN || n > kMaxN) return SPM_EINVAL;
    std::vector<int> y0;
    double lb = 0;
    const std::vector<int64_t> b = codegen_matrix(a, n, options);
    int rc = int_prepare(b.data(), n, y0, &lb);
    if (rc != SPM_OK && rc != SPM_ERANGE) return rc;
    const std::string src = emit_int_kernel(b, n, codegen_name(pattern_hash(b.data(), n, options)));
    if (needed) *needed = src.size() + 1;
    if (buf && buflen > 0) {
        const size_t k = std::min(buflen - 1, src.size());
        std::memcpy(buf, src.data(), k);
        buf[k] = 0;
    }
    return SPM_OK;
}

extern "C" int spm_codegen_compile_only(const int64_t* a, int n, int options, double* seconds, size_t* code_bytes)
{
    if (!a || n < kCodegenMinN || n > kMaxN) return SPM_EINVAL;
    const std::vector<int64_t> b = codegen_matrix(a, n, options);
    const uint64_t h = pattern_hash(b.data(), n, options);
    return codegen_compile_only(emit_int_kernel(b, n, codegen_name(h)), codegen_name(h), seconds, code_bytes);
}

extern "C" int spm_ut_order(const double* a, int n, int* row_perm, int* col_perm)
{
    if (!a || !row_perm || !col_perm || n < 1 || n > 4096) return SPM_EINVAL;
    std::vector<unsigned char> nz((size_t)n * n);
    for (size_t e = 0; e < nz.size(); ++e) nz[e] = a[e] != 0;
    std::vector<int> rp, cp;
    ut_order_mask(nz.data(), n, rp, cp);
    std::memcpy(row_perm, rp.data(), sizeof(int) * n);
    std::memcpy(col_perm, cp.data(), sizeof(int) * n);
    return SPM_OK;
}

extern "C" int spm_plan_codegen(const int64_t* a, int n, int options, int device, spm_plan** out)
{
    if (!out) return SPM_EINVAL;
    *out = nullptr;
    if (!a || n < kCodegenMinN || n > kMaxN) return SPM_EINVAL;
    const std::vector<int64_t> b = codegen_matrix(a, n, options);
    std::vector<unsigned char> w(n);
    const int halved = int_row_weights(b.data(), n, w.data());  // the emitted kernel halves the same rows
    std::vector<int> y0;
    double lb = 0;
    int rc = int_prepare(b.data(), n, y0, &lb, w.data());
    if (rc != SPM_OK) return rc;
    int cus = 0;
    rc = device_cus(device, &cus);
    if (rc != SPM_OK) return rc;
    DeviceGuard guard(device);
    if (!guard.ok()) return SPM_EHIP;

    spm_plan* p = new spm_plan;
    p->kind = 4;
    p->n = n;
    p->precision = -1;
    p->device = device;
    p->log2_bound = lb;
    p->halved = halved;
    p->rec = 2;
    p->hash = pattern_hash(b.data(), n, options);
    p->kname = codegen_name(p->hash);
    rc = codegen_get_kernel(device, p->hash, emit_int_kernel(b, n, p->kname), p->kname, &p->fn, &p->compile_s,
                            &p->cached);
    if (rc != SPM_OK) {
        plan_free(p);
        return rc;
    }
    int bpc = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, p->fn, kDenseBlock, 0) != hipSuccess || bpc <= 0) {
        plan_free(p);
        return SPM_EHIP;
    }
    const int64_t waves = int64_t(bpc) * cus * (kDenseBlock / kWaveLanes);
    gray_geometry(n, codegen_int_unroll_bits(b.data(), n), waves, &p->L, &p->chunks, &p->tasks);
    const uint64_t need_blocks = (p->tasks + (kDenseBlock / kWaveLanes) - 1) / (kDenseBlock / kWaveLanes);
    p->grid = (int)std::min<uint64_t>(uint64_t(bpc) * cus, need_blocks);
    const double est_s = std::ldexp(1.0, n - 1) * 16.0 / 3.9e13;
    p->slices = (int)std::min<double>(std::max(1.0, std::ceil(est_s / 4.0)), (double)p->tasks);
    SPM_TRY(hipMalloc(&p->d_partial, sizeof(unsigned long long) * 2 * p->grid));
    SPM_TRY(hipMalloc(&p->d_out, sizeof(double) * 6));
    SPM_TRY(hipHostMalloc(&p->h_out, sizeof(double) * 6, hipHostMallocDefault));
    SPM_TRY(hipEventCreate(&p->ev0));
    SPM_TRY(hipEventCreate(&p->ev1));
    *out = p;
    return SPM_OK;
}

// Generated SpaRyser plans for real FP64 matrices (kind 5): the same
// per-block records and final reduction as the dense kernels.
static std::vector<double> codegen_matrix_real(const double* a, int n, int options)
{
    if (options & SPM_CG_REGISTER_EFFICIENT) {
        std::vector<unsigned char> nz((size_t)n * n);
        for (size_t e = 0; e < nz.size(); ++e) nz[e] = a[e] != 0.0;
        std::vector<int> rp, cp;
        ut_order_mask(nz.data(), n, rp, cp);
        std::vector<double> b((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) b[(size_t)rp[i] * n + cp[j]] = a[(size_t)i * n + j];
        return b;
    }
    return std::vector<double>(a, a + (size_t)n * n);
}

static std::string codegen_real_name(uint64_t h)
{
    char nm[40];
    std::snprintf(nm, sizeof nm, "spm_genf_%016llx", (unsigned long long)h);
    return nm;
}

static int real_codegen_args(const double* a, int n, int precision)
{
    if (!a || n < kCodegenMinN || n > kMaxN) return SPM_EINVAL;
    if (precision != SPM_PREC_PLAIN && precision != SPM_PREC_COMPENSATED) return SPM_EINVAL;
    for (int e = 0; e < n * n; ++e)
        if (!std::isfinite(a[e])) return SPM_EINVAL;
    return SPM_OK;
}

extern "C" int spm_codegen_source_real(const double* a, int n, int options, int precision, char* buf, size_t buflen,
                                       size_t* needed)
{
    const int rc = real_codegen_args(a, n, precision);
    if (rc != SPM_OK) return rc;
    const std::vector<double> b = codegen_matrix_real(a, n, options);
    const std::string src = emit_real_kernel(b, n, precision == SPM_PREC_COMPENSATED,
                                             codegen_real_name(real_pattern_hash(b.data(), n, options, precision)));
    if (needed) *needed = src.size() + 1;
    if (buf && buflen > 0) {
        const size_t k = std::min(buflen - 1, src.size());
        std::memcpy(buf, src.data(), k);
        buf[k] = 0;
    }
    return SPM_OK;
}

extern "C" int spm_plan_codegen_real(const double* a, int n, int options, int precision, int device, spm_plan** out)
{
    if (!out) return SPM_EINVAL;
    *out = nullptr;
    int rc = real_codegen_args(a, n, precision);
    if (rc != SPM_OK) return rc;
    const std::vector<double> b = codegen_matrix_real(a, n, options);
    int cus = 0;
    rc = device_cus(device, &cus);
    if (rc != SPM_OK) return rc;
    DeviceGuard guard(device);
    if (!guard.ok()) return SPM_EHIP;

    spm_plan* p = new spm_plan;
    p->kind = 5;
    p->n = n;
    p->precision = precision;
    p->device = device;
    p->rec = 4;
    p->comps = 1;
    p->hash = real_pattern_hash(b.data(), n, options, precision);
    p->kname = codegen_real_name(p->hash);
    rc = codegen_get_kernel(device, p->hash,
                            emit_real_kernel(b, n, precision == SPM_PREC_COMPENSATED, p->kname), p->kname, &p->fn,
                            &p->compile_s, &p->cached);
    if (rc != SPM_OK) {
        plan_free(p);
        return rc;
    }
    int bpc = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, p->fn, kDenseBlock, 0) != hipSuccess || bpc <= 0) {
        plan_free(p);
        return SPM_EHIP;
    }
    const int64_t waves = int64_t(bpc) * cus * (kDenseBlock / kWaveLanes);
    gray_geometry(n, codegen_real_unroll_bits(n), waves, &p->L, &p->chunks, &p->tasks);
    p->max_grid = bpc * cus;
    dense_grid(p);
    SPM_TRY(hipMalloc(&p->d_partial, sizeof(double) * p->rec * p->grid));
    p->partial_cap = p->grid;
    SPM_TRY(hipMalloc(&p->d_out, sizeof(double) * 6));
    SPM_TRY(hipHostMalloc(&p->h_out, sizeof(double) * 6, hipHostMallocDefault));
    SPM_TRY(hipEventCreate(&p->ev0));
    SPM_TRY(hipEventCreate(&p->ev1));
    *out = p;
    return SPM_OK;
}

extern "C" int spm_plan_codegen_info(const spm_plan* p, double* compile_seconds, int* cached, uint64_t* hash)
{
    if (!p || (p->kind != 4 && p->kind != 5)) return SPM_EINVAL;
    if (compile_seconds) *compile_seconds = p->compile_s;
    if (cached) *cached = p->cached ? 1 : 0;
    if (hash) *hash = p->hash;
    return SPM_OK;
}

extern "C" int spm_plan_dense_c128(const double* a, int n, int precision, int device, spm_plan** out)
{
    return make_dense_plan(2, a, n, precision, device, out);
}

#undef SPM_TRY

static int dense_run(spm_plan* p, hipStream_t st, spm_result* r)
{
    const double t0 = now_seconds();
    DeviceGuard guard(p->device);
    if (!guard.ok()) return SPM_EHIP;
    if (hipEventRecord(p->ev0, st) != hipSuccess) return SPM_EHIP;
    const uint64_t per = (p->tasks + p->slices - 1) / p->slices;
    int launches = 0;
    for (int s = 0; s < p->slices; ++s) {
        uint64_t tb = per * s, te = std::min<uint64_t>(p->tasks, per * (s + 1));
        if (tb >= te) break;
        int acc = s > 0 ? 1 : 0;
        if (p->kind == 5) {
            void* args[] = {&p->d_partial, &tb, &te, &p->chunks, &p->L, &acc};
            if (hipModuleLaunchKernel(p->fn, p->grid, 1, 1, kDenseBlock, 1, 1, 0, st, args, nullptr) != hipSuccess)
                return SPM_EHIP;
        } else {
            void* args[] = {&p->d_M, &p->d_x0, &p->d_partial, &tb, &te, &p->chunks, &p->L, &acc};
            if (hipLaunchKernel(p->kern, dim3(p->grid), dim3(kDenseBlock), args, 0, st) != hipSuccess)
                return SPM_EHIP;
        }
        ++launches;
    }
    {
        int nrec = p->grid, stride = p->rec, comps = p->comps;
        void* args[] = {&p->d_partial, &nrec, &stride, &comps, &p->d_out};
        if (hipLaunchKernel(reduce_partials_kernel_ptr(), dim3(1), dim3(256), args, 0, st) != hipSuccess)
            return SPM_EHIP;
    }
    if (hipEventRecord(p->ev1, st) != hipSuccess) return SPM_EHIP;
    if (hipMemcpyAsync(p->h_out, p->d_out, sizeof(double) * 3 * p->comps, hipMemcpyDeviceToHost, st) != hipSuccess)
        return SPM_EHIP;
    if (hipStreamSynchronize(st) != hipSuccess) return SPM_EHIP;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p->ev0, p->ev1) != hipSuccess) ms = -1.f;

    std::memset(r, 0, sizeof *r);
    const int n = p->n;
    dense_fill_result(n, p->comps == 2, p->precision != SPM_PREC_PLAIN, p->h_out, r);
    if (p->precision == SPM_PREC_F32_PRODUCT) f32prod_finish(n, p->log2_scale, r);
    r->status = SPM_OK;
    r->kernel_ms = ms;
    r->subsets = uint64_t(1) << (n - 1);
    r->n = n;
    r->launches = launches;
    r->seconds = now_seconds() - t0;
    return SPM_OK;
}

static int int_run(spm_plan* p, hipStream_t st, spm_result* r)
{
    const double t0 = now_seconds();
    DeviceGuard guard(p->device);
    if (!guard.ok()) return SPM_EHIP;
    if (hipEventRecord(p->ev0, st) != hipSuccess) return SPM_EHIP;
    const uint64_t per = (p->tasks + p->slices - 1) / p->slices;
    int launches = 0;
    for (int s = 0; s < p->slices; ++s) {
        uint64_t tb = per * s, te = std::min<uint64_t>(p->tasks, per * (s + 1));
        if (tb >= te) break;
        int acc = s > 0 ? 1 : 0;
        if (p->kind == 4) {
            void* args[] = {&p->d_partial, &tb, &te, &p->chunks, &p->L, &acc};
            if (hipModuleLaunchKernel(p->fn, p->grid, 1, 1, kDenseBlock, 1, 1, 0, st, args, nullptr) != hipSuccess)
                return SPM_EHIP;
        } else {
            void* args[] = {&p->d_M, &p->d_x0, &p->d_partial, &tb, &te, &p->chunks, &p->L, &acc};
            if (hipLaunchKernel(p->kern, dim3(p->grid), dim3(kDenseBlock), args, 0, st) != hipSuccess)
                return SPM_EHIP;
        }
        ++launches;
    }
    {
        int nrec = p->grid;
        void* args[] = {&p->d_partial, &nrec, &p->d_out};
        if (hipLaunchKernel(reduce_i128_kernel_ptr(), dim3(1), dim3(256), args, 0, st) != hipSuccess)
            return SPM_EHIP;
    }
    if (hipEventRecord(p->ev1, st) != hipSuccess) return SPM_EHIP;
    if (hipMemcpyAsync(p->h_out, p->d_out, 16, hipMemcpyDeviceToHost, st) != hipSuccess) return SPM_EHIP;
    if (hipStreamSynchronize(st) != hipSuccess) return SPM_EHIP;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p->ev0, p->ev1) != hipSuccess) ms = -1.f;
    std::memset(r, 0, sizeof *r);
    unsigned long long w[2];
    std::memcpy(w, p->h_out, 16);
    const int rc = int_fill_result(p->n, w[0], w[1], r, p->halved);
    if (rc != SPM_OK) return rc;
    r->status = SPM_OK;
    r->kernel_ms = ms;
    r->subsets = uint64_t(1) << (p->n - 1);
    r->n = p->n;
    r->launches = launches;
    r->seconds = now_seconds() - t0;
    return SPM_OK;
}

extern "C" int spm_plan_run(spm_plan* p, void* stream, spm_result* out)
{
    if (!p || !out) return SPM_EINVAL;
    if (p->kind == 1 || p->kind == 2 || p->kind == 5) return dense_run(p, (hipStream_t)stream, out);
    if (p->kind == 3 || p->kind == 4) return int_run(p, (hipStream_t)stream, out);
    return SPM_EINVAL;
}

extern "C" int spm_plan_destroy(spm_plan* p)
{
    plan_free(p);
    return SPM_OK;
}

extern "C" int spm_plan_info(const spm_plan* p, int* grid, int* block, int* log2_chunk, uint64_t* tasks)
{
    if (!p) return SPM_EINVAL;
    if (grid) *grid = p->grid;
    if (block) *block = kDenseBlock;
    if (log2_chunk) *log2_chunk = p->L;
    if (tasks) *tasks = p->tasks;
    return SPM_OK;
}

extern "C" const char* spm_plan_kernel_name(const spm_plan* p) { return p ? p->kname.c_str() : ""; }

extern "C" int spm_plan_set_log2_chunk(spm_plan* p, int L)
{
    if (!p || (p->kind != 1 && p->kind != 2)) return SPM_EINVAL;
    const int n = p->n, B = n - 1;
    const int ub = p->kind == 1 ? dense_f64_unroll_bits(n) : (n >= 3 ? 1 : 0);
    const int lo = ub > 0 ? ub + 1 : (B >= 1 ? 1 : 0);
    if (L < lo || L > B) return SPM_EINVAL;
    p->L = L;
    p->chunks = uint64_t(1) << (B - L);
    p->tasks = (p->chunks + kWaveLanes - 1) / kWaveLanes;
    dense_grid(p);
    if (p->grid > p->partial_cap) {
        DeviceGuard guard(p->device);
        if (!guard.ok()) return SPM_EHIP;
        double* np = nullptr;
        if (hipMalloc(&np, sizeof(double) * p->rec * p->grid) != hipSuccess) return SPM_ENOMEM;
        if (p->d_partial) (void)hipFree(p->d_partial);
        p->d_partial = np;
        p->partial_cap = p->grid;
    }
    return SPM_OK;
}

static int dense_oneshot(int kind, const double* a, int n, int precision, int device, void* stream,
                         spm_result* out)
{
    if (!out) return SPM_EINVAL;
    const double t0 = now_seconds();
    spm_plan* p = nullptr;
    int st = make_dense_plan(kind, a, n, precision, device, &p);
    if (st != SPM_OK) {
        std::memset(out, 0, sizeof *out);
        out->status = st;
        return st;
    }
    st = dense_run(p, (hipStream_t)stream, out);
    plan_free(p);
    if (st == SPM_OK) out->seconds = now_seconds() - t0;
    else out->status = st;
    return st;
}

extern "C" int spm_dense_f64(const double* a, int n, int precision, int device, void* stream, spm_result* out)
{
    return dense_oneshot(1, a, n, precision, device, stream, out);
}

extern "C" int spm_dense_c128(const double* a, int n, int precision, int device, void* stream, spm_result* out)
{
    return dense_oneshot(2, a, n, precision, device, stream, out);
}

extern "C" int spm_sparse_codegen(const int64_t* a, int n, int options, int device, void* stream, spm_result* out)
{
    if (!out) return SPM_EINVAL;
    const double t0 = now_seconds();
    spm_plan* p = nullptr;
    int st = spm_plan_codegen(a, n, options, device, &p);
    if (st != SPM_OK) {
        std::memset(out, 0, sizeof *out);
        out->status = st;
        return st;
    }
    st = int_run(p, (hipStream_t)stream, out);
    plan_free(p);
    if (st == SPM_OK) out->seconds = now_seconds() - t0;
    else out->status = st;
    return st;
}

extern "C" int spm_sparse_codegen_real(const double* a, int n, int options, int precision, int device, void* stream,
                                       spm_result* out)
{
    if (!out) return SPM_EINVAL;
    const double t0 = now_seconds();
    spm_plan* p = nullptr;
    int st = spm_plan_codegen_real(a, n, options, precision, device, &p);
    if (st != SPM_OK) {
        std::memset(out, 0, sizeof *out);
        out->status = st;
        return st;
    }
    st = dense_run(p, (hipStream_t)stream, out);
    plan_free(p);
    if (st == SPM_OK) out->seconds = now_seconds() - t0;
    else out->status = st;
    return st;
}

// ---------------------------------------------------------------------------
// Automatic path choice for one matrix.  The reference's rule
// (AlgorithmSelector.h:52-67) picks between its dense kernel and code
// generation by nnz per row and component size, with costs of its own GPUs
// and a recompile of the whole program per pattern.  Here every candidate's
// time is estimated from rates measured on MI355X (per Gray subset) plus the
// hiprtc compile where a generated kernel is not yet in the module cache,
// and the cheapest runs:
//   SPM_PATH_INT        integer matrix: exact walk (compressed columns for
//                       sparse ones, else FP64 or int32 walk)
//   SPM_PATH_GEN_INT    integer matrix: exact generated kernel
//   SPM_PATH_SCALED_INT rows s_i * integer vectors: exact integer permanent
//                       times prod s_i (double-double), one rounding
//   SPM_PATH_GEN_REAL   real matrix: generated FP64 kernel
//   SPM_PATH_DENSE      real matrix: dense FP64 walk
// ---------------------------------------------------------------------------
namespace {

struct Shape {
    int n = 0;
    size_t nnz = 0;
    bool integral = false;  // every entry an integer of magnitude <= 2^30
    IntWalkCosts walks;     // integral matrices: the exact walks' per-subset costs
};

Shape shape_of(const double* a, int n)
{
    Shape s;
    s.n = n;
    s.integral = true;
    for (size_t e = 0; e < (size_t)n * n; ++e) {
        const double v = a[e];
        s.nnz += v != 0.0 ? 1 : 0;
        if (!(std::fabs(v) <= 1073741824.0) || v != std::floor(v)) s.integral = false;
    }
    if (s.integral) {
        std::vector<int64_t> q((size_t)n * n);
        for (size_t e = 0; e < q.size(); ++e) q[e] = (int64_t)a[e];
        s.walks = int_walk_costs(q.data(), n);
    }
    return s;
}

double subsets(int n) { return std::ldexp(1.0, n - 1); }
double per_row(const Shape& s) { return double(s.nnz) / s.n; }

// Seconds, from the kernels' measured rates on MI355X at orders 32-40
// (fixed launch costs of a few tens of microseconds included; below order
// ~30 every path is dominated by the ~0.15 ms a call costs on the host and
// the dense or exact walks win by not compiling).  Generated kernels: per
// subset, the flipped column's entries and the product (the real kernels'
// cached suffix products stop helping as rows fill: order 36, 3 entries per
// row 20 ms, dense 295 ms against 71 ms for the dense walk), and for
// integer kernels the 128-bit product on the fraction q of terms with no
// zero row (order 36, 3 entries per row, odd row sums: 291 ms; 4 per row:
// 9 ms).  Per-subset rates of the current generated kernels at orders 36
// and 40 (0/1 and N(0,1) values on 4, 6 and 8 entries per row): integer
// 1.6-2.3e-13 s, 4.9-6.7e-13, 1.1-1.6e-12 (as (entries per row)^2.75);
// real 1.6-1.8e-13, 2.4-2.8e-13, 3.5-3.8e-13 (the quadratic term keeps
// dense patterns on the dense walk).  Kernels whose terms are rarest (q
// below kGenCompactTermQ: one term test per block) run 0.7x those integer
// rates (order 40: 4 per row 51-58 ms against 78-86 ms with a test per
// step, 6 per row 223 against 272 ms).  Compiles: first-call minus cached
// times; integer kernels about 2.5e-3 2^U n s (U = log2 of the block length
// for order n), 2e-4 2^U n s where terms are rarest, real ones 0.25-0.35 s.
double est_dense(const Shape& s) { return 3e-5 + subsets(s.n) * 2.0 * s.n / 3.5e13; }
double est_int_walk(const Shape& s)
{
    const IntWalkCosts& c = s.walks;
    const double walk = s.n >= 26 ? std::min(c.sparse, std::min(c.fp64, c.int32)) : std::min(c.fp64, c.int32);
    return 3e-5 + subsets(s.n) * walk;
}
double est_gen_int(const Shape& s)
{
    const double pr = per_row(s);
    const double rare = s.walks.q < kGenCompactTermQ ? 0.7 : 1.0;
    return 3e-5 + subsets(s.n) * rare * (4.3e-15 * std::pow(pr, 2.75) + 2.3e-13 * s.n * s.walks.q);
}
double est_gen_real(const Shape& s)
{
    const double pr = per_row(s);
    return 3e-5 + subsets(s.n) * (3.0e-14 * pr + 4e-15 * pr * pr);
}
double est_compile(const Shape& s)
{
    const bool compact = s.walks.q < kGenCompactTermQ;  // one term test per block: far less code
    return 0.1 + (compact ? 2e-4 : 2.5e-3) * std::ldexp(1.0, gen_unroll_bits_int(s.n, compact)) * s.n;
}
double est_compile_real(const Shape& s) { return 0.15 + 1.2e-3 * double(s.nnz); }

}  // namespace

// Whether the register-efficient generated kernel for the integer matrix
// `a` is already compiled (in this process or in the on-disk cache): then it
// costs a load, not a compile.
static bool gen_kernel_cached(const double* a, int n, bool integral, int precision, int device)
{
    (void)precision;
    if (!integral) return false;
    const int opt = SPM_CG_REGISTER_EFFICIENT;
    std::vector<int64_t> q((size_t)n * n);
    for (size_t e = 0; e < q.size(); ++e) q[e] = (int64_t)a[e];
    const std::vector<int64_t> b = codegen_matrix(q.data(), n, opt);
    const uint64_t h = pattern_hash(b.data(), n, opt);
    const std::string name = codegen_name(h);
    return codegen_is_cached(device, h, emit_int_kernel(b, n, name), name);
}

// A cached kernel still costs a module load (from disk: milliseconds), so
// below order ~30, where every path takes well under a millisecond, the
// choice never depends on what happens to be cached.  Only the exact paths
// look at the caches (use_cache): their result is the same whichever runs,
// while an FP64 result would change its last bits with the state of a cache.
constexpr double kCachedLoadSeconds = 5e-3;

static int choose_path(const double* a, int n, int precision, int device, bool use_cache, int* path)
{
    const Shape s = shape_of(a, n);
    const bool can_gen = n >= kCodegenMinN;
    if (s.integral) {
        const double g = est_gen_int(s), w = est_int_walk(s);
        bool gen = can_gen && g + est_compile(s) < w;
        if (!gen && use_cache && can_gen && g + kCachedLoadSeconds < w)
            gen = gen_kernel_cached(a, n, true, precision, device);
        *path = gen ? SPM_PATH_GEN_INT : SPM_PATH_INT;
        return SPM_OK;
    }
    std::vector<int64_t> q;
    double sh = 0, sl = 0;
    if (row_scaled_integer(a, n, q, &sh, &sl)) {
        *path = SPM_PATH_SCALED_INT;
        return SPM_OK;
    }
    *path = can_gen && est_gen_real(s) + est_compile_real(s) < est_dense(s) ? SPM_PATH_GEN_REAL : SPM_PATH_DENSE;
    return SPM_OK;
}

extern "C" int spm_choose_path(const double* a, int n, int* path)
{
    if (!a || !path || n < 1 || n > kMaxN) return SPM_EINVAL;
    for (size_t e = 0; e < (size_t)n * n; ++e)
        if (!std::isfinite(a[e])) return SPM_EINVAL;
    return choose_path(a, n, SPM_PREC_COMPENSATED, 0, false, path);
}

extern "C" int spm_path_estimate(const double* a, int n, int path, double* seconds)
{
    if (!a || !seconds || n < 1 || n > kMaxN) return SPM_EINVAL;
    const Shape s = shape_of(a, n);
    switch (path) {
    case SPM_PATH_INT: *seconds = est_int_walk(s); break;
    case SPM_PATH_GEN_INT: *seconds = est_gen_int(s) + est_compile(s); break;
    case SPM_PATH_SCALED_INT: *seconds = est_int_walk(s); break;
    case SPM_PATH_GEN_REAL: *seconds = est_gen_real(s) + est_compile_real(s); break;
    case SPM_PATH_DENSE: *seconds = est_dense(s); break;
    default: return SPM_EINVAL;
    }
    return SPM_OK;
}

// Integer matrix q by the chosen integer path; SPM_ERANGE when the exact
// range does not hold.
static int run_int_path(const int64_t* q, int n, bool gen, int device, void* stream, spm_result* out)
{
    if (gen) return spm_sparse_codegen(q, n, SPM_CG_REGISTER_EFFICIENT, device, stream, out);
    return spm_int_exact(q, n, device, stream, out);
}

extern "C" int spm_permanent_auto(const double* a, int n, int precision, int device, void* stream, spm_result* out,
                                  int* path_taken)
{
    if (!out) return SPM_EINVAL;
    std::memset(out, 0, sizeof *out);
    if (!a || n < 1 || n > kMaxN) return out->status = SPM_EINVAL;
    if (precision != SPM_PREC_PLAIN && precision != SPM_PREC_COMPENSATED) return out->status = SPM_EINVAL;
    for (size_t e = 0; e < (size_t)n * n; ++e)
        if (!std::isfinite(a[e])) return out->status = SPM_EINVAL;
    const double t0 = now_seconds();
    int path = SPM_PATH_DENSE;
    int st = choose_path(a, n, precision, device, true, &path);
    if (st != SPM_OK) return out->status = st;
    if (path == SPM_PATH_INT || path == SPM_PATH_GEN_INT) {
        std::vector<int64_t> q((size_t)n * n);
        for (size_t e = 0; e < q.size(); ++e) q[e] = (int64_t)a[e];
        st = run_int_path(q.data(), n, path == SPM_PATH_GEN_INT, device, stream, out);
        if (st != SPM_ERANGE) {
            if (path_taken) *path_taken = path;
            if (st == SPM_OK) out->seconds = now_seconds() - t0;
            return st;
        }
        path = SPM_PATH_DENSE;  // beyond the exact range: FP64
    } else if (path == SPM_PATH_SCALED_INT) {
        std::vector<int64_t> q;
        double sh = 0, sl = 0;
        row_scaled_integer(a, n, q, &sh, &sl);
        std::vector<double> qd(q.begin(), q.end());
        int qp = SPM_PATH_INT;
        choose_path(qd.data(), n, precision, device, true, &qp);
        spm_result t;
        st = run_int_path(q.data(), n, qp == SPM_PATH_GEN_INT, device, stream, &t);
        if (st == SPM_OK) {
            scaled_int_result(sh, sl, t, out);
            if (path_taken) *path_taken = path;
            out->seconds = now_seconds() - t0;
            return SPM_OK;
        }
        if (st != SPM_ERANGE) return out->status = st;
        Shape sp = shape_of(a, n);
        path = n >= kCodegenMinN && est_gen_real(sp) + est_compile_real(sp) < est_dense(sp) ? SPM_PATH_GEN_REAL
                                                                                           : SPM_PATH_DENSE;
    }
    if (path == SPM_PATH_GEN_REAL)
        st = spm_sparse_codegen_real(a, n, SPM_CG_REGISTER_EFFICIENT, precision, device, stream, out);
    else
        st = spm_dense_f64(a, n, precision, device, stream, out);
    if (path_taken) *path_taken = path;
    if (st == SPM_OK) out->seconds = now_seconds() - t0;
    return st;
}

extern "C" int spm_int_exact(const int64_t* a, int n, int device, void* stream, spm_result* out)
{
    if (!out) return SPM_EINVAL;
    const double t0 = now_seconds();
    spm_plan* p = nullptr;
    int st = make_int_plan(a, n, SPM_INT_WALK_AUTO, device, &p);
    if (st != SPM_OK) {
        std::memset(out, 0, sizeof *out);
        out->status = st;
        return st;
    }
    st = int_run(p, (hipStream_t)stream, out);
    plan_free(p);
    if (st == SPM_OK) out->seconds = now_seconds() - t0;
    else out->status = st;
    return st;
}

// ---------------------------------------------------------------------------
// Misc
// ---------------------------------------------------------------------------
extern "C" const char* spm_version(void) { return "superman_amd 0.1 (gfx950)"; }

extern "C" const char* spm_strerror(int s)
{
    switch (s) {
    case SPM_OK: return "ok";
    case SPM_EINVAL: return "invalid argument";
    case SPM_EHIP: return "HIP runtime error";
    case SPM_ENOMEM: return "out of device memory or over budget";
    case SPM_ERANGE: return "exact result not representable (bound exceeded)";
    case SPM_ECOMPILE: return "generated kernel failed to compile";
    case SPM_ENOTINT: return "matrix has non-integer entries";
    case SPM_ECANCELLED: return "cancelled";
    case SPM_EAGAIN: return "admission refused, retry later";
    case SPM_ETIMEDOUT: return "deadline exceeded";
    case SPM_EFAULTINJ: return "injected fault";
    default: return "unknown status";
    }
}

extern "C" int spm_device_count(int* count)
{
    if (!count) return SPM_EINVAL;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *count = 0;
        return e == hipErrorNoDevice ? SPM_OK : SPM_EHIP;
    }
    *count = c;
    return SPM_OK;
}

extern "C" int spm_gray_geometry(int n, int64_t resident_waves, int* log2_chunk, uint64_t* chunks, uint64_t* tasks,
                                 int* unroll_bits)
{
    if (n < 1 || n > kMaxN || !log2_chunk || !chunks || !tasks) return SPM_EINVAL;
    gray_geometry(n, dense_f64_unroll_bits(n), resident_waves, log2_chunk, chunks, tasks, GrayWalk::kReal);
    if (unroll_bits) *unroll_bits = dense_f64_unroll_bits(n);
    return SPM_OK;
}

extern "C" int spm_device_synchronize(int device)
{
    DeviceGuard g(device);
    if (!g.ok()) return SPM_EHIP;
    return hipDeviceSynchronize() == hipSuccess ? SPM_OK : SPM_EHIP;
}

extern "C" int spm_device_info(int device, int* compute_units, int* clock_khz, char* name, int name_len)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return SPM_EHIP;
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (clock_khz) *clock_khz = prop.clockRate;
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    return SPM_OK;
}
