// walk_sparse.h -- exact Gray-code Ryser walk for sparse integer matrices on
// gfx950, with the matrix's columns read at run time.
//
// Replaces the reference's runtime-CSC sparse kernels
// (Algorithms/GPU/Sparse/SparseKernelDefinitions.cuh:34-509 driven by
// spSingleGPU.cuh:44-234): per Gray step only the rows of the flipped
// column change, a count of zero rows is kept, and the product over all
// rows is formed only when no row is zero -- here deferred through a
// per-wave queue, so that 64 queued products are formed together.
//
// As in the dense walks here, every lane of a wavefront walks its own
// aligned chunk of 2^L steps and all lanes flip the same column in
// lockstep, so a column's rows and values are wave-uniform: a column comes
// in as 64-byte records (spm_kernels.h), one scalar load each, prefetched
// one step ahead, and a row index selects the register of y by VGPR
// indexing (the index is an SGPR), so y stays in VGPRs.  An indexed access
// reaches within one register tuple of at most 32 registers, so past order
// 32 the rows are two arrays ("banks", split at sparse_bank_rows), each
// indexed on its own; a record lists its bank-a entries from slot 0 up and
// its bank-b entries from slot 6 down, so the bank of every slot is known
// at compile time.  The slots in use are found by uniform tests that stop
// at the first empty slot.  Arithmetic is exact (y = 2x in int32, T mod
// 2^128, as walk_int.h).
#ifndef SPM_WALK_SPARSE_H
#define SPM_WALK_SPARSE_H

#include "spm_device.h"
#include "spm_kernels.h"
#include "walk_int.h"
#include "walk_intfp.h"

namespace spm {
namespace sparse {

using ints::const_i32;
using ints::fence;
using ints::step_ptr;

struct Rec {
    int v[kSparseRec];
};

__device__ __forceinline__ void load_rec(Rec& r, const int* recs, int idx)
{
    const const_i32* __restrict__ q = step_ptr(recs + idx * kSparseRec);
#pragma unroll
    for (int i = 0; i < kSparseRec; ++i) r.v[i] = q[i];
}

// The record of column j in plane p (0: +2A, 1: -2A).
__device__ __forceinline__ int rec_index(int j, int p) { return 2 * j + p; }

// Row sums of the walk: rows below sparse_bank_rows(N) in ya, the rest in
// yb, two separate arrays (one aggregate would not be kept in registers).
// Only compile-time indexes reach both; run-time (wave-uniform) row indexes
// come with their bank from the record layout.
template <int N>
constexpr int kRowsA = sparse_bank_rows(N);
template <int N>
constexpr int kRowsB = N - kRowsA<N>;
template <int N>
using BankA = int[kRowsA<N>];
template <int N>
using BankB = int[kRowsB<N> > 0 ? kRowsB<N> : 1];

// y[row] += value (MODE 0: as stored; 1: negated on lanes with m = -1,
// the per-lane direction; 2: kept on lanes with m = -1 only, chunk-start
// membership), keeping Z = number of zero rows.
template <int MODE, int R>
__device__ __forceinline__ void update_row(int (&y)[R], int& Z, int row, int v, int m)
{
    if constexpr (MODE == 1) v = (v ^ m) - m;
    else if constexpr (MODE == 2) v &= m;
    const int old = y[row];
    const int now = old + v;
    y[row] = now;
    Z += (now == 0) - (old == 0);
}

// Slots U, U+1, ... (bank a) while U < used; nested, so a record costs one
// scalar test per entry plus the first failing one.
template <int U, int MODE, int R>
__device__ __forceinline__ void slots_up(int (&y)[R], int& Z, const Rec& r, int used, int m)
{
    if constexpr (U < kSparseSlots) {
        if (U < used) {
            update_row<MODE>(y, Z, r.v[1 + U], r.v[1 + kSparseSlots + U], m);
            slots_up<U + 1, MODE>(y, Z, r, used, m);
        }
    }
}

// Slots 6-U, 5-U, ... (bank b) while U < used.
template <int U, int MODE, int R>
__device__ __forceinline__ void slots_down(int (&y)[R], int& Z, const Rec& r, int used, int m)
{
    if constexpr (U < kSparseSlots) {
        if (U < used) {
            constexpr int s = kSparseSlots - 1 - U;
            update_row<MODE>(y, Z, r.v[1 + s], r.v[1 + kSparseSlots + s], m);
            slots_down<U + 1, MODE>(y, Z, r, used, m);
        }
    }
}

// The entries of one record: the count word is the bank-a count, plus the
// bank-b count times 2^16 when N > 32.
template <int N, int MODE>
__device__ __forceinline__ void apply_slots(BankA<N>& ya, BankB<N>& yb, int& Z, const Rec& r, int m)
{
    const int cnt = r.v[kSparseRec - 1];
    if constexpr (kRowsB<N> == 0) {
        slots_up<0, MODE>(ya, Z, r, cnt, m);
    } else {
        slots_up<0, MODE>(ya, Z, r, cnt & 0xffff, m);
        slots_down<0, MODE>(yb, Z, r, cnt >> 16, m);
    }
}

// A whole column from its first record (already loaded); longer columns
// continue in further records, loaded here (rare for sparse inputs).
template <int N, int MODE>
__device__ __forceinline__ void apply_column(BankA<N>& ya, BankB<N>& yb, int& Z, const Rec& first, const int* recs,
                                             int m)
{
    apply_slots<N, MODE>(ya, yb, Z, first, m);
    int nx = first.v[0];
    while (nx >= 0) {
        Rec r;
        load_rec(r, recs, nx);
        apply_slots<N, MODE>(ya, yb, Z, r, m);
        nx = r.v[0];
    }
}

// Exact product of rows 0..N-1 mod 2^128.  SMALL (every |y_r| < 2^b, b =
// bit width of N) forms it as walk_intfp.h does -- FP64 groups of exact
// integers below 2^51, then int64 -- with far fewer multiplies than the
// int64 row pairs of the general case.
template <int N>
__device__ __forceinline__ unsigned __int128 row_product(const int (&y)[N], bool small)
{
    unsigned __int128 p;
    if (small) {
        constexpr int G = real::kIntFpGroup<N>;
        constexpr int NG = real::kIntFpGroups<N>;
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const int r0 = q * G;
            const int r1 = (r0 + G < N) ? r0 + G : N;
            double g = (double)y[r0];
#pragma unroll
            for (int r = r0 + 1; r < r1; ++r) g *= (double)y[r];
            const unsigned __int128 w = (unsigned __int128)(__int128)real::exact_int64(g);
            p = q == 0 ? w : p * w;
        }
    } else {
        p = (unsigned __int128)(__int128)y[0];
#pragma unroll
        for (int r = 1; r + 1 < N; r += 2) {
            const long long q = (long long)y[r] * (long long)y[r + 1];
            p *= (unsigned __int128)(__int128)q;
        }
        if constexpr (N % 2 == 0) p *= (unsigned __int128)(__int128)y[N - 1];
    }
    return p;
}

// Deferred products.  A term is nonzero on a few lanes of a wave at a time
// (every row nonzero: 0.2-1.5 % of subsets of the sparse matrices
// measured), so forming it in place would make the whole wave wait on a
// 64-lane product for one lane.  Instead a lane with a nonzero term copies
// its row sums and sign into the wave's LDS queue, and when the queue is
// full every lane forms one queued product.  Entries are kQStride ints (N
// rows, the sign), an odd stride so that 64 lanes reading one entry each
// touch distinct banks.  A queue holds 64 entries, except at orders 39..41,
// where 64 would hold the kernel to 3 blocks of 4 waves per CU although
// its registers allow 4 (at 128 VGPRs the spills fall around the flush's
// products): there it holds what fits in a sixteenth of the CU's 160 KiB
// LDS less the block's reduction records.
template <int N>
constexpr int kQStride = (N + 1) | 1;
constexpr int kCuLdsBytes = 160 * 1024;
constexpr int kQBlocksPerCu = 4;
constexpr int kQBlockReserve = 512;  // the block's reduction records, rounded to the allocation unit
template <int N>
constexpr uint32_t kQCap =
    (N >= 39 && N <= 41)
        ? (uint32_t)((kCuLdsBytes / kQBlocksPerCu - kQBlockReserve) / ((kDenseBlock / kWave) * kQStride<N> * 4))
        : 64u;
template <int N>
constexpr bool kQFitsFourBlocks =
    (kDenseBlock / kWave) * (int)kQCap<N> * kQStride<N> * 4 + kQBlockReserve <= kCuLdsBytes / kQBlocksPerCu;

struct Queue {
    int* q;          // this wave's kQCap x kQStride ints
    uint32_t count;  // wave-uniform
};

template <int N>
__device__ __forceinline__ void flush(Queue& Q, bool small, uint32_t lane, unsigned __int128& acc)
{
    static_assert(kQCap<N> >= 1 && kQCap<N> <= 64, "one queued product per lane");
    static_assert(kQCap<N> == 64 || kQFitsFourBlocks<N>, "capped queues are sized for 4 blocks per CU");
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane < Q.count) {
        const int* e = Q.q + lane * kQStride<N>;
        int y[N];
#pragma unroll
        for (int r = 0; r < N; ++r) y[r] = e[r];
        const unsigned __int128 p = row_product<N>(y, small);
        acc = e[N] != 0 ? acc - p : acc + p;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    Q.count = 0;
}

// Copies this lane's row sums and sign into queue entry `slot`.
template <int N>
__device__ __forceinline__ void enqueue(const BankA<N>& ya, const BankB<N>& yb, bool negative, Queue& Q,
                                        uint32_t slot)
{
    int* e = Q.q + slot * kQStride<N>;
#pragma unroll
    for (int r = 0; r < kRowsA<N>; ++r) e[r] = ya[r];
#pragma unroll
    for (int r = 0; r < kRowsB<N>; ++r) e[kRowsA<N> + r] = yb[r];
    e[N] = negative ? 1 : 0;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Queues +/- prod_r y_r for the active lanes (the wave mask `act`) with no
// zero row: one compare into a lane mask and a scalar test per step.
template <int N>
__device__ __forceinline__ void add_term(const BankA<N>& ya, const BankB<N>& yb, int Z, bool negative, bool small,
                                         uint64_t act, Queue& Q, uint32_t lane, unsigned __int128& acc)
{
    uint64_t need = __builtin_amdgcn_ballot_w64(Z == 0) & act;
    if constexpr (kQCap<N> == 64) {
        if (__builtin_expect(need != 0, 0)) {
            const uint32_t cnt = (uint32_t)__builtin_popcountll(need);
            if (Q.count + cnt > kQCap<N>) flush<N>(Q, small, lane, acc);
            if ((need >> lane) & 1ull) enqueue<N>(ya, yb, negative, Q, Q.count + lanes_below(need));
            Q.count += cnt;
        }
    } else {
        // A ballot may name more lanes than a queue holds: take the lowest
        // lanes that fit, flush, and go on with the rest.
        while (__builtin_expect(need != 0, 0)) {
            if (Q.count != 0 && Q.count + (uint32_t)__builtin_popcountll(need) > kQCap<N>)
                flush<N>(Q, small, lane, acc);
            uint64_t take = need;
            while ((uint32_t)__builtin_popcountll(take) > kQCap<N> - Q.count)
                take &= ~(1ull << (63 - __builtin_clzll(take)));
            if ((take >> lane) & 1ull) enqueue<N>(ya, yb, negative, Q, Q.count + lanes_below(take));
            Q.count += (uint32_t)__builtin_popcountll(take);
            need &= ~take;
            if (need != 0) flush<N>(Q, small, lane, acc);
        }
    }
}

// One lane's chunk: the 2^L Gray steps starting at k = h 2^L, its nonzero
// terms queued (inactive lanes -- past the matrix's last chunk -- queue
// nothing); flushed products go into acc (mod 2^128).  Steps as in
// walk_int.h: k = 2m flips column ctz(m)+1 in a uniform direction, k = 2m+1
// flips column 0, block nblk/2 flips column L-1 in a per-lane direction.
template <int N>
__device__ __forceinline__ void walk_sparse(const int* __restrict__ recs, const int* __restrict__ y0, uint64_t h,
                                            int L, bool active, Queue& Q, uint32_t lane, unsigned __int128& acc)
{
    BankA<N> ya;
    BankB<N> yb;
    {
        const const_i32* __restrict__ ys = step_ptr(y0);
#pragma unroll
        for (int r = 0; r < kRowsA<N>; ++r) ya[r] = ys[r];
#pragma unroll
        for (int r = 0; r < kRowsB<N>; ++r) yb[r] = ys[kRowsA<N> + r];
    }
    const bool small = ((const const_i32*)y0)[N] != 0;
    const uint64_t act = __builtin_amdgcn_ballot_w64(active);
    int Z = 0;
    const uint64_t k0 = h << L;
    const uint64_t g0 = k0 ^ (k0 >> 1);
    for (int j = (L > 0 ? L - 1 : 0); j < N - 1; ++j) {
        Rec r;
        load_rec(r, recs, rec_index(j, 0));
        apply_column<N, 2>(ya, yb, Z, r, recs, -(int)((g0 >> j) & 1u));
    }
    Z = 0;
#pragma unroll
    for (int r = 0; r < kRowsA<N>; ++r) Z += ya[r] == 0 ? 1 : 0;
#pragma unroll
    for (int r = 0; r < kRowsB<N>; ++r) Z += yb[r] == 0 ? 1 : 0;
    const int lane_mask = -(int)(h & 1u);  // direction of column L-1: -1 = remove

    if constexpr (N == 1) {
        add_term<N>(ya, yb, Z, false, small, act, Q, lane, acc);
    } else if constexpr (N == 2) {  // L = 1
        add_term<N>(ya, yb, Z, false, small, act, Q, lane, acc);
        Rec r;
        load_rec(r, recs, rec_index(0, 0));
        apply_column<N, 1>(ya, yb, Z, r, recs, lane_mask);
        add_term<N>(ya, yb, Z, true, small, act, Q, lane, acc);
    } else {
        const uint32_t nblk = 1u << (L - 1);
        const uint32_t half = nblk >> 1;
        auto a_rec = [&](uint32_t m) __attribute__((always_inline)) {  // record of step k = 2m
            const int tz = __builtin_ctz(m);
            return rec_index(tz + 1, (int)((m >> (tz + 1)) & 1u));
        };
        Rec ra, rb;
        add_term<N>(ya, yb, Z, false, small, act, Q, lane, acc);  // k = 0
        load_rec(rb, recs, rec_index(0, 0));
        apply_column<N, 0>(ya, yb, Z, rb, recs, 0);  // k = 1: +column 0
        load_rec(ra, recs, a_rec(1));
        fence();
        add_term<N>(ya, yb, Z, true, small, act, Q, lane, acc);
        fence();
        auto block = [&](uint32_t m, auto SPECIAL) __attribute__((always_inline)) {
            if constexpr (decltype(SPECIAL)::value) apply_column<N, 1>(ya, yb, Z, ra, recs, lane_mask);
            else apply_column<N, 0>(ya, yb, Z, ra, recs, 0);
            load_rec(rb, recs, rec_index(0, (int)(m & 1u)));  // step 2m+1: column 0
            fence();
            add_term<N>(ya, yb, Z, false, small, act, Q, lane, acc);
            fence();
            apply_column<N, 0>(ya, yb, Z, rb, recs, 0);
            load_rec(ra, recs, a_rec(m + 1));  // m+1 = nblk: column L <= N-1, a valid record
            fence();
            add_term<N>(ya, yb, Z, true, small, act, Q, lane, acc);
            fence();
        };
#pragma nounroll
        for (uint32_t m = 1; m < half; ++m) block(m, std::false_type{});
        block(half, std::true_type{});
#pragma nounroll
        for (uint32_t m = half + 1; m < nblk; ++m) block(m, std::false_type{});
    }
}

}  // namespace sparse
}  // namespace spm

#endif
