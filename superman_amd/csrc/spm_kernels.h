// spm_kernels.h -- host-visible handles to the device kernels.
//
// Each kernel family exposes a lookup returning the __global__ function's
// host stub address, launched with hipLaunchKernel by the drivers.
#pragma once

#include <stdint.h>

namespace spm {

constexpr int kMaxN = 63;          // 64-bit Gray index (reference README.md:75)
constexpr int kDenseBlock = 256;   // threads per block of the dense kernels

// dense_f64.hip
//   args: (const double* M /* [2][n][n]: +A, -A column-major */, const double* x0, double* partial,
//          uint64_t task_begin, uint64_t task_end, uint64_t num_chunks, int L,
//          int accumulate)
const void* dense_f64_kernel_ptr(int n, bool compensated);
// FP32-product variant (rows pre-scaled by f32prod_scale); same arguments.
const void* dense_f32p_kernel_ptr(int n);
int dense_f64_unroll_bits(int n);
//   args: (const double* partial, int nrec, int stride, int comps, double* out)
const void* reduce_partials_kernel_ptr();

// dense_c128.hip -- same arguments as the real kernel; M is [2][n][2n]
// (+A, -A; column-major, interleaved re/im), x0 is 2n doubles, and each
// partial record is 8 doubles (re hi, lo, abs, im hi, lo, 0, 0, 0).
constexpr int kMaxComplexN = 63;
const void* dense_c128_kernel_ptr(int n, bool compensated);

// int_dense.hip -- exact integer walk.
//   args: (const int* M /* [2][n][n]: +2A, -2A column-major */, const int* y0,
//          unsigned long long* partial /* [grid][2] */, uint64_t task_begin,
//          uint64_t task_end, uint64_t num_chunks, int L, int accumulate)
const void* int_dense_kernel_ptr(int n);
//   args: (const unsigned long long* partial, int nrec, unsigned long long* out /* [2] */)
const void* reduce_i128_kernel_ptr();

// int_small.hip -- exact walk in FP64 for integer matrices whose absolute
// row sums are all below 2^int_small_bits(n) (every 0/1 matrix); same
// arguments as int_dense_kernel except that M is [2][n][n] doubles (+2A,
// -2A) and y0 is n doubles.
const void* int_small_kernel_ptr(int n);
int int_small_bits(int n);

// int_sparse.hip -- exact walk over compressed columns (sparse integer
// matrices); same arguments as int_dense_kernel except that M is the record
// array below and y0 is n + 1 int32 values: y(empty set), then 1 when every
// absolute row sum is below 2^b, b = bit width of n (products may then be
// formed in FP64 groups, as int_small_kernel does), else 0.  A column's
// entries in one plane (+2A or -2A) sit in 16-int records (64 bytes, one
// scalar load):
//   [0] next record of the column (or -1), [1] entries in this record
//   (1..7), [2..8] rows, [9..15] values
// Record 2j + p (p = 0: +2A, 1: -2A) is the first of column j; longer
// columns continue in records after the first 2n (a record that continues
// is full).  A zero column has one entry of value 0 on row 0, so that every
// record's first slot is in use.  Unused slots hold row 0 and value 0 and
// are never read.
const void* int_sparse_kernel_ptr(int n);
constexpr int kSparseRec = 16;   // ints per record
constexpr int kSparseSlots = 7;  // entries per record

// Generated SpaRyser kernels walk 2^3-step blocks with L > 3.
constexpr int kCodegenMinN = 5;

}  // namespace spm
