// batch.hip -- many permanents (real FP64, complex128, exact integer; orders
// 1..kBatchMaxN) in one persistent launch fed by a device work queue.
//
// The reference runs one matrix per launch sequence (dpSingleGPU.cuh:148-199
// and its twins: a power-of-two chunk loop with a device synchronisation per
// launch).  Here a batch of matrices sits in one device arena; each matrix is
// split into work items of 64 lane chunks (one wavefront's task: lane h of
// item t walks the aligned Gray range [(64 t + h) 2^L, (64 t + h + 1) 2^L) of
// that matrix, exactly as in the single-matrix kernels, walk_*.h).  The grid
// is sized to the resident wave count; every wave pulls the next item index
// from a device counter (one atomic per item), dispatches on the item's kind
// and order (wave-uniform) and writes the item's wave-reduced sum to its own
// record.  A wave leaves when the counter passes the item count, so every
// wave reaches the exit.  The wave that completes a matrix's last item (a
// per-matrix count) reduces that matrix's item records in a fixed order and
// publishes the result to host memory: the result of a matrix depends only
// on its data and its chunk size 2^L, never on the batch it rides in or on
// scheduling.
//
// Arena layout (byte offsets in the header, all 64-byte aligned):
//   header | request descriptors | item -> request table | matrix data
// Records: item_rec[item][8] (device), result_host[request][8] (doubles, in
// host-mapped memory, published with a per-request flag so the host can
// hand a result out as soon as its last item is done):
//   real    : hi, lo, abs
//   complex : re hi, re lo, abs, im hi, im lo
//   integer : T mod 2^128 as (lo, hi) 64-bit words (both integer walks:
//             int32 for general entries, FP64 when every absolute row sum
//             is below 2^(bit width of n), e.g. 0/1 matrices)
// and in every record word 7 = batch_tag(seq, item or request index); in a
// result record word 6 = the number of item records the reduction found
// without this batch's tag and word 5 = record_check() of the other seven
// words.
//
// Memory ordering.  Nothing here depends on a cache being written back or
// invalidated at the right moment:
//   - the batch reaches the device arena through batch_stage_kernel, the
//     first kernel of the batch's launch sequence (a kernel, not a DMA
//     copy): kernel-to-kernel visibility on one stream is all the walk
//     kernel relies on;
//   - item records are written and read with device-scope atomic stores and
//     loads, and the per-request count that hands a request to its reducing
//     wave is an acquire-release device-scope atomic;
//   - a result record is stored to host memory with system-scope stores,
//     then its flag with a system-scope release store, so a host that reads
//     the flag with acquire semantics sees the whole record.
#include "spm_batch.h"
#include "walk_c128.h"
#include "walk_f64.h"
#include "walk_int.h"
#include "walk_intfp.h"

#include <utility>

namespace spm {

namespace batch {

// The item of this wave.  Only the lane's chunk index h is per-lane; the
// descriptor fields are passed through readfirstlane so that the compiler
// keeps them (and the column addresses built from them) in SGPRs.
struct Item {
    const uint8_t* M;
    const uint8_t* x0;
    uint64_t h;
    int L;
};

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

__device__ __forceinline__ uint64_t uniform(uint64_t v)
{
    return ((uint64_t)uniform((uint32_t)(v >> 32)) << 32) | uniform((uint32_t)v);
}

// Item-record words: device-scope atomic (coherent) stores and loads.
__device__ __forceinline__ void store_agent(double* p, double v)
{
    __hip_atomic_store((unsigned long long*)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double load_agent(const double* p)
{
    return __longlong_as_double(
        (long long)__hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

template <int N>
__device__ __forceinline__ void item_real(const Item& it, double (&o)[8])
{
    const double* M = (const double*)it.M;
    const double* x0 = (const double*)it.x0;
    double hi, lo, ab;
    real::walk_f64<N, real::kDenseUnroll<N>, true>(M, x0, it.h, it.L, hi, lo, ab);
    o[0] = hi;
    o[1] = lo;
    o[2] = ab;
}

template <int N>
__device__ __forceinline__ void item_complex(const Item& it, double (&o)[8])
{
    const double* M = (const double*)it.M;
    const double* x0 = (const double*)it.x0;
    double rh = 0.0, rl = 0.0, ih = 0.0, il = 0.0, ab = 0.0;
    cplx::walk_c128<N, true>(M, x0, it.h, it.L, rh, rl, ih, il, ab);
    o[0] = rh;
    o[1] = rl;
    o[2] = ab;
    o[3] = ih;
    o[4] = il;
}

template <int N>
__device__ __forceinline__ void item_int(const Item& it, double (&o)[8])
{
    const int* M = (const int*)it.M;
    const int* y0 = (const int*)it.x0;
    unsigned __int128 acc = 0;
    ints::walk_int<N>(M, y0, it.h, it.L, acc);
    o[0] = __longlong_as_double((long long)(unsigned long long)acc);
    o[1] = __longlong_as_double((long long)(unsigned long long)(acc >> 64));
}

template <int N>
__device__ __forceinline__ void item_int_small(const Item& it, double (&o)[8])
{
    const unsigned __int128 acc = real::walk_intfp<N>((const double*)it.M, (const double*)it.x0, it.h, it.L);
    o[0] = __longlong_as_double((long long)(unsigned long long)acc);
    o[1] = __longlong_as_double((long long)(unsigned long long)(acc >> 64));
}

template <int KIND, int N>
__device__ __forceinline__ void run_item(const Item& it, double (&o)[8])
{
    if constexpr (KIND == kBatchReal) item_real<N>(it, o);
    else if constexpr (KIND == kBatchComplex) item_complex<N>(it, o);
    else if constexpr (KIND == kBatchInt) item_int<N>(it, o);
    else item_int_small<N>(it, o);
}

// Order dispatch: a chain of uniform compares, one inlined walk per order.
template <int KIND, int... Is>
__device__ __forceinline__ void dispatch(int n, const Item& it, double (&o)[8], std::integer_sequence<int, Is...>)
{
    (void)((n == Is + 1 ? (run_item<KIND, Is + 1>(it, o), true) : false) || ...);
}

__device__ __forceinline__ void reduce_i128_wave(unsigned long long& lo, unsigned long long& hi)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long olo = __shfl_xor(lo, off, kWave);
        const unsigned long long ohi = __shfl_xor(hi, off, kWave);
        const unsigned long long s = lo + olo;
        hi = hi + ohi + (s < lo ? 1ull : 0ull);
        lo = s;
    }
}

// Combine the lanes' sums of one item (or of one request's items) in a fixed
// butterfly order; every lane ends with the total.
__device__ __forceinline__ void wave_combine(int kind, double (&o)[8])
{
    if (kind == kBatchInt || kind == kBatchIntSmall) {
        unsigned long long lo = (unsigned long long)__double_as_longlong(o[0]);
        unsigned long long hi = (unsigned long long)__double_as_longlong(o[1]);
        reduce_i128_wave(lo, hi);
        o[0] = __longlong_as_double((long long)lo);
        o[1] = __longlong_as_double((long long)hi);
    } else {
        wave_reduce_dd(o[0], o[1], o[2]);
        if (kind == kBatchComplex) {
            double z = 0.0;
            wave_reduce_dd(o[3], o[4], z);
        }
    }
}

// A request's item records summed in a fixed order (lane l takes items
// l, l+64, ..., then the butterfly); every lane writes the same total.
// Word kBatchTag of each item record must hold the item's tag for this
// batch; records that do not are counted into word kBatchStale of the
// result (0 when every record was this batch's), and the result's own tag
// goes into its word kBatchTag.
__device__ __forceinline__ void reduce_request(int kind, const double* __restrict__ recs, uint32_t item_base,
                                               uint32_t items, uint32_t seq, uint32_t r, uint32_t lane,
                                               double* __restrict__ out)
{
    double o[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    unsigned __int128 t = 0;
    uint32_t stale = 0;
    const bool is_int = kind == kBatchInt || kind == kBatchIntSmall;
    for (uint32_t j = lane; j < items; j += kWave) {
        const double* p = recs + (size_t)j * kBatchRec;
        double rec[kBatchRec];
#pragma unroll
        for (int c = 0; c < kBatchRec; ++c) rec[c] = load_agent(p + c);
        stale += __double_as_longlong(rec[kBatchTag]) != (long long)batch_tag(seq, item_base + j) ? 1u : 0u;
        if (is_int) {
            t += ((unsigned __int128)(unsigned long long)__double_as_longlong(rec[1]) << 64) |
                 (unsigned long long)__double_as_longlong(rec[0]);
        } else {
            dd_add(o[0], o[1], rec[0], rec[1]);
            o[2] += rec[2];
            if (kind == kBatchComplex) dd_add(o[3], o[4], rec[3], rec[4]);
        }
    }
    if (is_int) {
        o[0] = __longlong_as_double((long long)(unsigned long long)t);
        o[1] = __longlong_as_double((long long)(unsigned long long)(t >> 64));
    }
    wave_combine(kind, o);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) stale += __shfl_xor(stale, off, kWave);
    o[kBatchStale] = (double)stale;
    o[kBatchTag] = __longlong_as_double((long long)batch_tag(seq, r));
    uint64_t w[kBatchRec];
#pragma unroll
    for (int c = 0; c < kBatchRec; ++c) w[c] = (uint64_t)__double_as_longlong(o[c]);
    o[kBatchCheck] = __longlong_as_double((long long)record_check(w));
#pragma unroll
    for (int c = 0; c < kBatchRec; ++c)
        __hip_atomic_store((unsigned long long*)(out + c), (unsigned long long)__double_as_longlong(o[c]),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace batch

// KINDS: bit mask of the kinds this instantiation walks (1 real, 2 complex,
// 4 integer int32 walk, 8 integer FP64 walk).  The fused kernel (15) walks
// header range 0 with counter 0; the real (1), complex (2) and integer (12)
// kernels walk range and counter 1, 2 and 3.
template <int KINDS>
__global__ void __launch_bounds__(kDenseBlock)
batch_walk_kernel(const uint8_t* __restrict__ arena, double* __restrict__ item_rec, unsigned int* __restrict__ counter,
                  unsigned int* __restrict__ done, double* __restrict__ result_host,
                  unsigned int* __restrict__ flag_host)
{
    using namespace batch;
    const BatchHeader* hd = (const BatchHeader*)arena;
    constexpr int q = KINDS == 15 ? 0 : KINDS == 1 ? 1 : KINDS == 2 ? 2 : 3;
    const uint32_t seq = uniform(hd->seq);
    const uint32_t begin = uniform(hd->range_begin[q]);
    const uint32_t end = uniform(hd->range_end[q]);
    const BatchRequestDesc* reqs = (const BatchRequestDesc*)(arena + hd->desc_off);
    const uint32_t* item_req = (const uint32_t*)(arena + hd->item_off);
    const uint32_t lane = threadIdx.x & (kWave - 1);

    // The loop body has no lane-divergent control flow at its own level: the
    // work counter is bumped by every lane (lane 0 adds 1, the others 0) and
    // every lane stores the (identical) wave total.  A lane-0-only branch
    // here lets the compiler's control-flow structurizer split the loop so
    // that the item index is read from a partial wave.
    uint32_t i = begin + uniform(atomicAdd(counter + q, lane == 0 ? 1u : 0u));
    while (i < end) {
        const BatchRequestDesc* d = reqs + uniform(item_req[i]);
        Item it;
        it.M = arena + uniform(d->mat_off);
        it.x0 = arena + uniform(d->x0_off);
        it.L = (int)uniform((uint32_t)d->L);
        it.h = (uint64_t)(i - uniform(d->item_base)) * kWave + lane;
        const uint64_t num_chunks = uniform(d->num_chunks);
        const int kind = (int)uniform(d->kind);
        const int n = (int)uniform(d->n);
        double o[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        constexpr auto orders = std::make_integer_sequence<int, kBatchMaxN>{};
        if constexpr ((KINDS & 1) != 0) if (kind == kBatchReal) dispatch<kBatchReal>(n, it, o, orders);
        if constexpr ((KINDS & 2) != 0) if (kind == kBatchComplex) dispatch<kBatchComplex>(n, it, o, orders);
        if constexpr ((KINDS & 4) != 0) if (kind == kBatchInt) dispatch<kBatchInt>(n, it, o, orders);
        if constexpr ((KINDS & 8) != 0) if (kind == kBatchIntSmall) dispatch<kBatchIntSmall>(n, it, o, orders);
        const bool idle = it.h >= num_chunks;  // lanes past the matrix's last chunk
#pragma unroll
        for (int c = 0; c < 8; ++c) o[c] = idle ? 0.0 : o[c];
        wave_combine(kind, o);
        o[kBatchTag] = __longlong_as_double((long long)batch_tag(seq, i));
        double* rec = item_rec + (size_t)i * kBatchRec;
#pragma unroll
        for (int c = 0; c < kBatchRec; ++c) store_agent(rec + c, o[c]);
        // The wave that completes a request's last item reduces its records
        // (release/acquire through the per-request count).
        const uint32_t r = uniform(item_req[i]);
        const uint32_t items = uniform(d->items);
        const uint32_t prev = uniform(
            __hip_atomic_fetch_add(done + r, lane == 0 ? 1u : 0u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT));
        if (prev + 1 == items) {
            const uint32_t base = uniform(d->item_base);
            reduce_request(kind, item_rec + (size_t)base * kBatchRec, base, items, seq, r, lane,
                           result_host + (size_t)r * kBatchRec);
            // publish: the record's stores are ordered before the flag's
            __hip_atomic_store(flag_host + r, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        i = begin + uniform(atomicAdd(counter + q, lane == 0 ? 1u : 0u));
    }
}

// Copies the staged batch from host-mapped (coherent, uncached) memory into
// the device arena and zeroes the four work counters and the per-request
// item counts.  The byte count is the staged header's; every region of the
// arena is 64-byte aligned, so the copy moves 16-byte words.
__global__ void __launch_bounds__(256) batch_stage_kernel(const uint8_t* __restrict__ stage,
                                                          uint8_t* __restrict__ arena, unsigned int* counter,
                                                          unsigned int* done, unsigned int n_done)
{
    const uint64_t words = ((const BatchHeader*)stage)->bytes / 16;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint4* __restrict__ src = (const uint4*)stage;
    uint4* __restrict__ dst = (uint4*)arena;
    for (uint64_t w = t; w < words; w += stride) dst[w] = src[w];
    if (t < 4) counter[t] = 0u;
    for (uint64_t j = t; j < n_done; j += stride) done[j] = 0u;
}

const void* batch_walk_kernel_ptr(int kinds)
{
    switch (kinds) {
    case 15: return reinterpret_cast<const void*>(&batch_walk_kernel<15>);
    case 1: return reinterpret_cast<const void*>(&batch_walk_kernel<1>);
    case 2: return reinterpret_cast<const void*>(&batch_walk_kernel<2>);
    case 12: return reinterpret_cast<const void*>(&batch_walk_kernel<12>);
    default: return nullptr;
    }
}
const void* batch_stage_kernel_ptr() { return reinterpret_cast<const void*>(&batch_stage_kernel); }

}  // namespace spm
