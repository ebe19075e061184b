// spm_device.h -- device-side helpers shared by the Ryser kernels (gfx950).
//
// * error-free transformations (TwoSum, Fast2Sum) and double-double adds used
//   for the compensated per-lane sums and for every cross-lane / cross-block
//   reduction;
// * the wavefront (64-lane) reduction of a double-double triple.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spm {

constexpr int kWave = 64;

// TwoSum (Knuth): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Fast2Sum (Dekker), valid when |a| >= |b| or a == 0.
__device__ __forceinline__ void fast_two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    e = b - (s - a);
}

// (ah, al) += (bh, bl) in double-double.
__device__ __forceinline__ void dd_add(double& ah, double& al, double bh, double bl)
{
    double s, e;
    two_sum(ah, bh, s, e);
    e += al + bl;
    fast_two_sum(s, e, ah, al);
}

// Accumulate a plain double into a compensated (hi, lo) pair.
__device__ __forceinline__ void dd_add_d(double& ah, double& al, double b)
{
    double s, e;
    two_sum(ah, b, s, e);
    ah = s;
    al += e;
}

// Butterfly all-reduce of (hi, lo, abs) over the 64 lanes of a wavefront.
// Every lane ends with the same value; the combination order is fixed, so
// the result is reproducible.
__device__ __forceinline__ void wave_reduce_dd(double& hi, double& lo, double& ab)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double oh = __shfl_xor(hi, off, kWave);
        double ol = __shfl_xor(lo, off, kWave);
        double oa = __shfl_xor(ab, off, kWave);
        dd_add(hi, lo, oh, ol);
        ab += oa;
    }
}

// Reduce (hi, lo, abs) over a block of up to 1024 threads; thread 0 returns
// the block total.  `scratch` must hold 3 * (blockDim.x / 64) doubles.
__device__ __forceinline__ void block_reduce_dd(double& hi, double& lo, double& ab, double* scratch)
{
    wave_reduce_dd(hi, lo, ab);
    const int w = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        scratch[3 * w + 0] = hi;
        scratch[3 * w + 1] = lo;
        scratch[3 * w + 2] = ab;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double h = scratch[0], l = scratch[1], a = scratch[2];
        for (int i = 1; i < nw; ++i) {
            dd_add(h, l, scratch[3 * i], scratch[3 * i + 1]);
            a += scratch[3 * i + 2];
        }
        hi = h;
        lo = l;
        ab = a;
    }
}

}  // namespace spm
