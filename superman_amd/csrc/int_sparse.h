// int_sparse.h -- single-matrix exact Ryser kernel for sparse integer
// matrices in compressed-column form (gfx950).  The per-chunk walk is in
// walk_sparse.h; this file holds the grid-strided task loop over one
// matrix's chunks (each wave with its queue of deferred products in LDS),
// the final flush of the queues and the block sum mod 2^128 into per-block
// records, which int_dense.hip's reduce_i128_kernel sums.  The orders are
// instantiated in several translation units (int_sparse*.hip), which the
// build compiles in parallel: each order's kernel takes 10-30 s to compile.
#ifndef SPM_INT_SPARSE_H
#define SPM_INT_SPARSE_H

#include "walk_sparse.h"

#ifndef SPM_SPARSE_ONLY_N
#define SPM_SPARSE_ONLY_N 0
#endif

namespace spm {
namespace sparse {

// Orders whose queues are sized for 4 blocks per CU (kQCap), and the
// orders with paired blocks (132-140 VGPRs at orders 26-31 otherwise, the
// peak in the queue's flush), ask the register allocator for 4 waves per
// SIMD.
template <int N>
constexpr int kSparseWaves = (kQCap<N> < 64 || (N >= kSparsePairMinN && N <= kSparsePairMaxN)) ? 4 : 1;

template <int N>
__global__ void __launch_bounds__(kDenseBlock) __attribute__((amdgpu_waves_per_eu(kSparseWaves<N>)))
int_sparse_kernel(const int* __restrict__ blob,                // compressed columns (walk_sparse.h)
                  const int* __restrict__ y0,                  // y(empty set), N values
                  unsigned long long* __restrict__ partial,    // [gridDim.x][2]: T mod 2^128 (lo, hi)
                  uint64_t task_begin, uint64_t task_end, uint64_t num_chunks, int L, int accumulate)
{
    __shared__ unsigned long long red[2 * (kDenseBlock / kWave)];
    __shared__ int queue[(kDenseBlock / kWave) * kQCap<N> * kQStride<N>];  // per-wave product queues
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = (uint64_t)blockIdx.x * (kDenseBlock / kWave) + (threadIdx.x / kWave);
    const uint64_t nwave = (uint64_t)gridDim.x * (kDenseBlock / kWave);
    unsigned __int128 acc = 0;

    Queue Q{queue + (threadIdx.x / kWave) * kQCap<N> * kQStride<N>, 0u};
    const bool small = ((const ints::const_i32*)y0)[N] != 0;
    for (uint64_t task = task_begin + wave; task < task_end; task += nwave) {
        const uint64_t h = task * kWave + lane;
        walk_sparse<N>(blob, y0, h, L, h < num_chunks, Q, lane, acc);  // lanes past the end walk a dummy chunk
    }
    flush<N>(Q, small, lane, acc);

    // Block sum mod 2^128 (fixed order: wave butterfly, then waves in order).
    unsigned long long lo = (unsigned long long)acc, hi = (unsigned long long)(acc >> 64);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long olo = __shfl_xor(lo, off, kWave);
        const unsigned long long ohi = __shfl_xor(hi, off, kWave);
        const unsigned long long s = lo + olo;
        hi = hi + ohi + (s < lo ? 1ull : 0ull);
        lo = s;
    }
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[2 * w] = lo;
        red[2 * w + 1] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned __int128 t = 0;
        for (int i = 0; i < kDenseBlock / kWave; ++i) t += ((unsigned __int128)red[2 * i + 1] << 64) | red[2 * i];
        unsigned long long* out = partial + 2 * blockIdx.x;
        if (accumulate) t += ((unsigned __int128)out[1] << 64) | out[0];
        out[0] = (unsigned long long)t;
        out[1] = (unsigned long long)(t >> 64);
    }
}

template <int N>
const void* sparse_ptr()
{
    if constexpr (SPM_SPARSE_ONLY_N != 0 && N != SPM_SPARSE_ONLY_N) return nullptr;
    else return reinterpret_cast<const void*>(&int_sparse_kernel<N>);
}

// The kernel of order n for LO <= n <= LO + sizeof...(Is) - 1, else null.
template <int LO, int... Is>
const void* sparse_range(int n, std::integer_sequence<int, Is...>)
{
    static const void* const tab[] = {sparse_ptr<LO + Is>()...};
    return (n >= LO && n < LO + (int)sizeof...(Is)) ? tab[n - LO] : nullptr;
}

template <int LO, int HI>
const void* sparse_range_ptr(int n)
{
    return sparse_range<LO>(n, std::make_integer_sequence<int, HI - LO + 1>{});
}

}  // namespace sparse

// Orders per translation unit (int_sparse_<LO>.hip holds LO..HI).
const void* int_sparse_ptr_17(int n);  // 17..25
const void* int_sparse_ptr_26(int n);  // 26..31
const void* int_sparse_ptr_32(int n);  // 32..36
const void* int_sparse_ptr_37(int n);  // 37..41
const void* int_sparse_ptr_42(int n);  // 42..45
const void* int_sparse_ptr_46(int n);  // 46..48
const void* int_sparse_ptr_49(int n);  // 49..56
const void* int_sparse_ptr_57(int n);  // 57..63

}  // namespace spm

#endif
