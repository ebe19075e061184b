// spm_batch.h -- device-visible layout of a batch arena, shared by the
// persistent batch kernels (batch.hip) and the host engine (batch_engine.cpp).
#pragma once

#include <stdint.h>

namespace spm {

constexpr int kBatchMaxN = 28;      // orders handled by the batched kernel
constexpr int kBatchReal = 1;       // compensated FP64
constexpr int kBatchComplex = 2;    // compensated complex128
constexpr int kBatchInt = 3;        // exact integer (mod 2^128), int32 walk
constexpr int kBatchIntSmall = 4;   // exact integer, small row sums: FP64 walk
constexpr int kBatchRec = 8;        // doubles per item / result record
constexpr int kBatchCheck = 5;      // result word: record_check() of the other seven words
constexpr int kBatchStale = 6;      // result word: item records of another batch seen by the reduction
constexpr int kBatchTag = 7;        // item / result word: batch_tag(seq, item or request index)

// Tag of item or request `idx` in the batch numbered `seq`, kept in word
// kBatchTag of its record so that a reader can tell this batch's record
// from an older one at the same place.
#ifdef __HIPCC__
__host__ __device__
#endif
inline uint64_t batch_tag(uint32_t seq, uint32_t idx)
{
    return ((uint64_t)seq << 32) | idx;
}

// Check word of a result record: an FNV-1a style mix of its other words'
// bit patterns, so that a reader can tell a whole record from one whose
// stores have not all arrived (the flag's release store orders them; the
// host checks anyway).
#ifdef __HIPCC__
__host__ __device__
#endif
inline uint64_t record_check(const uint64_t* w)
{
    uint64_t h = 0xcbf29ce484222325ull;
    for (int c = 0; c < kBatchRec; ++c)
        if (c != kBatchCheck) h = (h ^ w[c]) * 0x100000001b3ull;
    return h;
}
constexpr int kBatchAlign = 64;     // byte alignment of every arena region

struct BatchHeader {                // 64 bytes at arena offset 0
    uint32_t requests;
    uint32_t items;
    uint32_t desc_off;              // BatchRequestDesc[requests]
    uint32_t item_off;              // uint32_t[items]: item -> request
    uint32_t range_begin[4];        // item range walked by kernel q: q = 0 all kinds
    uint32_t range_end[4];          // (fused kernel); q = 1 real, 2 complex, 3 integer
    uint32_t seq;                   // batch number, written to a request's flag when done
    uint32_t pad;
    uint64_t bytes;                 // staged bytes: the upload kernel copies [0, bytes)
};

struct BatchRequestDesc {           // 64 bytes
    uint32_t kind;
    uint32_t n;
    int32_t L;                      // lane chunk = 2^L Gray steps
    uint32_t items;                 // ceil(num_chunks / 64)
    uint64_t num_chunks;            // 2^(n-1-L)
    uint64_t mat_off;               // +A/-A planes (f64, c128) or +2A/-2A (int32)
    uint64_t x0_off;                // x(empty set) / y(empty set)
    uint32_t item_base;             // first item of this request
    uint32_t pad[5];
};

static_assert(sizeof(BatchHeader) == 64, "header layout");
static_assert(sizeof(BatchRequestDesc) == 64, "descriptor layout");

// batch.hip
//   walk:  (const uint8_t* arena, double* item_rec, unsigned int* counter /* [4] */,
//           unsigned int* done /* [requests] */, double* result_host /* [requests][8] */,
//           unsigned int* flag_host /* [requests] */)
//          result_host / flag_host are host-mapped: a request's result record
//          is published there, then its flag is set to the header's seq.
//          `kinds` = 15 (every kind in one kernel), 1 (real), 2 (complex)
//          or 12 (both integer walks)
//   stage: (const uint8_t* stage /* host-mapped, coherent */, uint8_t* arena,
//           unsigned int* counter, unsigned int* done, unsigned int n_done)
//          copies the staged batch (header.bytes bytes, a multiple of 16)
//          into the device arena and zeroes the work counters and the
//          per-request item counts; launched with kStageBlocks x 256.
const void* batch_walk_kernel_ptr(int kinds);
const void* batch_stage_kernel_ptr();
constexpr int kStageBlocks = 128;

}  // namespace spm
