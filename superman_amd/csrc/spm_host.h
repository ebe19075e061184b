// spm_host.h -- host-side helpers shared by the drivers behind the C ABI.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "superman_amd.h"

namespace spm {

constexpr int kWaveLanes = 64;

int device_cus(int device, int* cus);
int blocks_per_cu(int device, const void* kern, int block, int* bpc);
double now_seconds();

// Gray-walk geometry shared by the dense kernels: 2^L steps per lane chunk,
// 2^(n-1-L) chunks, grouped 64 per wave task.
// Which walk a geometry is for: the dense walks take longer chunks than the
// wave count alone would give at small orders.
enum class GrayWalk { kOther, kReal, kComplex };
void gray_geometry(int n, int unroll_bits, int64_t resident_waves, int* L, uint64_t* chunks, uint64_t* tasks,
                   GrayWalk walk = GrayWalk::kOther);
// Grid and launch slices of a dense walk over `tasks` wave tasks (kind 1
// real, 2 complex, 5 generated real): as many blocks as the tasks fill, at
// most max_grid (the resident ones); runs longer than ~4 s of VALU work are
// split into launches that accumulate into the same per-block records.
void dense_launch_shape(int kind, int n, int precision, uint64_t tasks, int max_grid, int* grid, int* slices);

// Absolute error estimate (not a guaranteed bound) for a compensated FP64
// real run: each term carries ~(n-1) product roundings plus one rounding
// per factor; abs_sum = 2 sum |term|.  The drift of a factor over the up to
// 2^L incremental updates since its chunk's exact start is not counted: a
// guaranteed bound is n (2^L + 2n + 1) u prod_i R_i (R_i = sum_j |a_ij|),
// far above the errors seen.
inline double dense_err_est(int n, double abs_sum)
{
    const double u = 1.1102230246251565e-16;  // 2^-53
    return 2.0 * n * u * abs_sum;
}

// The same estimate for complex runs: abs_sum = 2 sum |p|_1 |x_last|_1 bounds the
// terms' moduli; each term carries ~2(n-1) roundings from its complex
// products plus its factors' incremental updates.
inline double complex_err_est(int n, double abs_sum)
{
    const double u = 1.1102230246251565e-16;
    return 4.0 * n * u * abs_sum;
}

// FP32-product runs: each term's first-order relative error is below
// (2n-3) 2^-24 (dense_f32p_kernel), on top of the FP64 accumulation's.
inline double f32prod_err_est(int n, double abs_sum)
{
    const double u32 = 5.9604644775390625e-08;  // 2^-24
    return (2.0 * n * u32 + dense_err_est(n, 1.0)) * abs_sum;
}

// prep.cpp -- per-matrix host preparation and result conversion
void dense_prepare(const double* a, int n, int w, double* planes, double* x0);
// w: per-row weights from int_row_weights (1 = row carried halved, 2 =
// not); nullptr = every row 2 (the batch engine's layout).
int int_row_weights(const int64_t* a, int n, unsigned char* w);
int int_prepare(const int64_t* a, int n, int* y0, double* log2_bound, const unsigned char* w = nullptr);
void int_planes(const int64_t* a, int n, int* planes, const unsigned char* w = nullptr);
void int_sparse_blob(const int64_t* a, int n, std::vector<int>& blob, const unsigned char* w = nullptr);
std::vector<int64_t> ascending_col_sort(const int64_t* a, int n);
bool int_sparse_preferred(const int64_t* a, int n);
double int_nonzero_term_fraction(const int64_t* a, int n);
// Seconds per Gray subset of the exact integer walks (prep.cpp), and the
// inputs of the estimate: entries per row and the fraction q of terms with
// no zero row.
struct IntWalkCosts {
    double sparse = 0, fp64 = 0, int32 = 0, per_row = 0, q = 1;
};
IntWalkCosts int_walk_costs(const int64_t* a, int n);
bool int_small_ok(const int64_t* a, int n);
bool row_scaled_integer(const double* a, int n, std::vector<int64_t>& q, double* scale_hi, double* scale_lo);
void scaled_int_result(double scale_hi, double scale_lo, const spm_result& t, spm_result* r);
void int_small_planes(const int64_t* a, int n, double* planes, const unsigned char* w = nullptr);
void dense_fill_result(int n, bool complex_kind, bool compensated, const double* rec, spm_result* r);
int f32prod_scale_exponent(int n);
int f32prod_scale(const double* a, int n, double* scaled);
void f32prod_finish(int n, int log2_scale, spm_result* r);
int int_fill_result(int n, unsigned long long lo, unsigned long long hi, spm_result* r, int halved_rows = 0);

// codegen.cpp -- SpaRyser per-pattern kernels
void ut_order_mask(const unsigned char* nz, int n, std::vector<int>& rowPerm, std::vector<int>& colPerm);
void ut_order(const int64_t* a, int n, std::vector<int>& rowPerm, std::vector<int>& colPerm);
int codegen_compile_only(const std::string& src, const std::string& name, double* seconds, size_t* code_bytes);
std::vector<int64_t> apply_perm(const int64_t* a, int n, const std::vector<int>& rp, const std::vector<int>& cp);
std::string emit_int_kernel(const std::vector<int64_t>& b, int n, const std::string& name);
std::string emit_real_kernel(const std::vector<double>& b, int n, bool compensated, const std::string& name);
int real_levels_count(const std::vector<double>& b, int n);
uint64_t real_pattern_hash(const double* b, int n, int options, int precision);
// log2 of a generated kernel's block length: real kernels by order, integer
// kernels by order and whether their terms are rare enough to be tested once
// per block (gen_compact_term: the row-independence estimate of the fraction
// of terms with no zero row below kGenCompactTermQ; the true fraction runs
// up to 100x higher on uniform patterns)
constexpr double kGenCompactTermQ = 5e-7;
int gen_unroll_bits_int(int n, bool compact);
bool gen_compact_term(const int64_t* b, int n);
int codegen_int_unroll_bits(const int64_t* b, int n);
int codegen_real_unroll_bits(int n);
uint64_t pattern_hash(const int64_t* b, int n, int options);
int codegen_get_kernel(int device, uint64_t hash, const std::string& src, const std::string& name, hipFunction_t* fn,
                       double* compile_s, bool* cached);
// Whether codegen_get_kernel would find the kernel compiled (this process,
// or the on-disk cache) -- host only.
bool codegen_is_cached(int device, uint64_t hash, const std::string& src, const std::string& name);

class DeviceGuard {
public:
    explicit DeviceGuard(int device);
    ~DeviceGuard();
    bool ok() const { return ok_; }
    DeviceGuard(const DeviceGuard&) = delete;
    DeviceGuard& operator=(const DeviceGuard&) = delete;

private:
    int prev_ = -1;
    bool ok_ = false;
};

}  // namespace spm
