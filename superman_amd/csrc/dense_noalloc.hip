// dense_noalloc.hip -- spm_dense_f64_noalloc: the dense real FP64 permanent
// of spm_dense_f64 with no allocation per call.
//
// spm_dense_f64 builds a plan per call: device buffers for the +A/-A planes,
// x(empty set), the per-block records and the reduced record, pinned host
// memory for the read-back and two events.  This variant keeps all of that
// device memory in one file-scope array, `scratch`, of a fixed size, takes
// the host-side copies on the stack and reads the result back into the
// stack, so a call makes no hipMalloc / hipFree / hipHostMalloc.  The kernels,
// the launch geometry and the host preparation are the ones spm_dense_f64
// uses, so both return bit-identical results for the same input.
//
// Each device holds its own copy of `scratch` at its own address; that
// address is looked up once per device (hipGetSymbolAddress with the device
// current) at the first call on it and kept, and only that address is passed
// to kernels and copies (the array's name in host code is a host-side
// stand-in, not a device address).  Every call on a device shares that
// device's array, so calls that overlap on one device race.  No part of the
// array needs to start at zero: the first launch of a run overwrites every
// per-block record it later accumulates into.
#include "superman_amd.h"

#include "spm_host.h"
#include "spm_kernels.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

namespace spm {
namespace {

// Layout of `scratch`, in doubles.  Offsets are multiples of 32 doubles
// (256 bytes).
constexpr size_t kPlanesOff = 0;                          // [2][n][n] +A, -A column-major
constexpr size_t kX0Off = 7968;                           // x(empty set), n values
constexpr size_t kOutOff = kX0Off + 64;                   // reduced record: hi, lo, abs
constexpr size_t kPartialOff = kOutOff + 32;              // per-block records, 4 doubles each
constexpr int kMaxBlocks = 4096;                          // per-block records held
constexpr size_t kScratchDoubles = kPartialOff + 4 * (size_t)kMaxBlocks;  // 24,448 doubles = 191 KiB
static_assert(2 * kMaxN * kMaxN <= (int)kX0Off, "planes of the largest order fit");
static_assert(kMaxN <= 64, "x(empty set) fits");
static_assert(kPlanesOff % 32 == 0 && kX0Off % 32 == 0 && kOutOff % 32 == 0 && kPartialOff % 32 == 0,
              "256-byte aligned regions");

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;

struct DeviceState {
    double* base = nullptr;       // device address of `scratch` on this device
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int bpc[kMaxN + 1][3] = {};   // resident blocks per CU by order and kernel (plain, compensated, FP32-product)
};

std::mutex g_mu;
DeviceState g_dev[kMaxDevices];

// Set-up of the calling thread's current device (once per device): the
// scratch array's device address and the two timing events.
int device_setup(int device, DeviceState** out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState& d = g_dev[device];
    if (!d.base) {
        double* ptr = nullptr;
        if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return SPM_EHIP;
        if (!d.ev0 && hipEventCreate(&d.ev0) != hipSuccess) return SPM_EHIP;
        if (!d.ev1 && hipEventCreate(&d.ev1) != hipSuccess) return SPM_EHIP;
        d.base = ptr;
    }
    *out = &d;
    return SPM_OK;
}

int resident_blocks(DeviceState& d, int n, int variant, const void* kern, int* bpc)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (d.bpc[n][variant] <= 0) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, kern, kDenseBlock, 0) != hipSuccess || v <= 0)
            return SPM_EHIP;
        d.bpc[n][variant] = v;
    }
    *bpc = d.bpc[n][variant];
    return SPM_OK;
}

}  // namespace
}  // namespace spm

using namespace spm;

extern "C" int spm_dense_f64_noalloc(const double* a, int n, int precision, int device, void* stream,
                                     spm_result* out)
{
    if (!out) return SPM_EINVAL;
    std::memset(out, 0, sizeof *out);
    const double t0 = now_seconds();
    if (!a || n < 1 || n > kMaxN || device < 0 || device >= kMaxDevices) return out->status = SPM_EINVAL;
    const bool f32p = precision == SPM_PREC_F32_PRODUCT;
    const bool comp = precision == SPM_PREC_COMPENSATED;
    if (!f32p && !comp && precision != SPM_PREC_PLAIN) return out->status = SPM_EINVAL;
    for (int e = 0; e < n * n; ++e)
        if (!std::isfinite(a[e]) || (f32p && a[e] < 0.0)) return out->status = SPM_EINVAL;
    int cus = 0;
    int st = device_cus(device, &cus);
    if (st != SPM_OK) return out->status = st;
    DeviceGuard guard(device);
    if (!guard.ok()) return out->status = SPM_EHIP;
    DeviceState* d = nullptr;
    if ((st = device_setup(device, &d)) != SPM_OK) return out->status = st;

    // The geometry of spm_dense_f64 (make_dense_plan + dense_grid in capi.cpp).
    const void* kern = f32p ? dense_f32p_kernel_ptr(n) : dense_f64_kernel_ptr(n, comp);
    if (!kern) return out->status = SPM_EINVAL;
    int bpc = 0;
    if ((st = resident_blocks(*d, n, f32p ? 2 : comp ? 1 : 0, kern, &bpc)) != SPM_OK) return out->status = st;
    const int64_t waves = int64_t(bpc) * cus * (kDenseBlock / kWaveLanes);
    int L = 0;
    uint64_t chunks = 0, tasks = 0;
    gray_geometry(n, dense_f64_unroll_bits(n), waves, &L, &chunks, &tasks, GrayWalk::kReal);
    int grid = 0, slices = 1;
    dense_launch_shape(1, n, precision, tasks, bpc * cus, &grid, &slices);
    if (grid > kMaxBlocks) return out->status = SPM_ENOMEM;  // more per-block records than the array holds

    // Host preparation on the stack (dpSingleGPU.cuh:62-82), as spm_dense_f64.
    double planes[2 * kMaxN * kMaxN];
    double x0[kMaxN];
    double scaled[kMaxN * kMaxN];
    int log2_scale = 0;
    if (f32p) {
        log2_scale = f32prod_scale(a, n, scaled);
        a = scaled;
    }
    dense_prepare(a, n, 1, planes, x0);

    hipStream_t s = (hipStream_t)stream;
    double* dM = d->base + kPlanesOff;
    double* dx0 = d->base + kX0Off;
    double* dout = d->base + kOutOff;
    double* dpart = d->base + kPartialOff;
    if (hipMemcpyAsync(dM, planes, sizeof(double) * 2 * n * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dx0, x0, sizeof(double) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipEventRecord(d->ev0, s) != hipSuccess)
        return out->status = SPM_EHIP;
    const uint64_t per = (tasks + slices - 1) / slices;
    int launches = 0;
    for (int k = 0; k < slices; ++k) {
        uint64_t tb = per * k, te = std::min<uint64_t>(tasks, per * (k + 1));
        if (tb >= te) break;
        int acc = k > 0 ? 1 : 0;
        void* args[] = {&dM, &dx0, &dpart, &tb, &te, &chunks, &L, &acc};
        if (hipLaunchKernel(kern, dim3(grid), dim3(kDenseBlock), args, 0, s) != hipSuccess)
            return out->status = SPM_EHIP;
        ++launches;
    }
    {
        int nrec = grid, stride = 4, comps = 1;
        void* args[] = {&dpart, &nrec, &stride, &comps, &dout};
        if (hipLaunchKernel(reduce_partials_kernel_ptr(), dim3(1), dim3(256), args, 0, s) != hipSuccess)
            return out->status = SPM_EHIP;
    }
    double rec[3];
    if (hipEventRecord(d->ev1, s) != hipSuccess ||
        hipMemcpyAsync(rec, dout, sizeof rec, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return out->status = SPM_EHIP;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, d->ev0, d->ev1) != hipSuccess) ms = -1.f;

    dense_fill_result(n, false, precision != SPM_PREC_PLAIN, rec, out);
    if (f32p) f32prod_finish(n, log2_scale, out);
    out->status = SPM_OK;
    out->kernel_ms = ms;
    out->subsets = uint64_t(1) << (n - 1);
    out->n = n;
    out->launches = launches;
    out->seconds = now_seconds() - t0;
    return SPM_OK;
}
