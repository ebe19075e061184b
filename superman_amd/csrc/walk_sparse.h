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
// one step ahead.  The row sums y stay in VGPRs at fixed registers, row r
// in v(2 + r) (Rows below), and an entry of the flipped column updates its
// row through VGPR indexing with the row as the index: one indexing region
// reads the old value's zero test, adds in place and tests the new value
// (row_entry), so an entry costs 2 scalar and 5 vector instructions and
// one uniform slot test.  The slots in use are found by tests that stop at
// the first empty slot.  Arithmetic is exact (y = 2x in int32, T mod
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
    // the offset after the asm: a constant record index folds into the loads
    const const_i32* __restrict__ q = step_ptr(recs) + idx * kSparseRec;
#pragma unroll
    for (int i = 0; i < kSparseRec; ++i) r.v[i] = q[i];
}

// The record of column j in plane p (0: +2A, 1: -2A).
__device__ __forceinline__ int rec_index(int j, int p) { return 2 * j + p; }

// Row sums of the walk, pinned at v2, v3, ...: a first register tuple at
// v2 holds rows 0..31 (fewer below order 32) and, past order 32, a second
// one at v34 the rest, so that one wave-uniform index reaches every row
// (row r is v(2 + r)).  Tuples come in the register-class sizes 1-12, 16
// and 32; rows past N in the last tuple are padding, never indexed.
constexpr int tuple_regs(int r) { return r <= 12 ? r : (r <= 16 ? 16 : 32); }
template <int N>
constexpr int kTupA = tuple_regs(N < 32 ? N : 32);
template <int N>
constexpr int kTupB = N > 32 ? tuple_regs(N - 32) : 0;

template <int K>
struct Tuple {
    typedef int T __attribute__((ext_vector_type(K > 0 ? K : 1)));
};

template <int N>
struct Rows {
    typename Tuple<kTupA<N>>::T a;
    typename Tuple<kTupB<N>>::T b;  // unused up to order 32
};

template <int N, int R>
__device__ __forceinline__ int row_get(const Rows<N>& y)
{
    if constexpr (R < 32) return y.a[R];
    else return y.b[R - 32];
}

template <int N, int R>
__device__ __forceinline__ void row_set(Rows<N>& y, int v)
{
    if constexpr (R < 32) y.a[R] = v;
    else y.b[R - 32] = v;
}

template <int J, int END, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (J < END) {
        f(std::integral_constant<int, J>{});
        static_for<J + 1, END>(f);
    }
}

// y[row] += w with Z (the count of zero rows) kept: in one indexing region
// (source 1 and destination relative to the row) the old value is tested
// into an SGPR pair, the sum formed in place and the new value tested into
// VCC; Z then takes the two tests as borrow and carry.  Nothing in the
// region writes a VGPR other than the indexed row, and a compare's SGPR
// destination is not relative to the index (checked on gfx950).  At least
// two instructions separate each VALU write of an SGPR pair from its read
// as carry-in, the wait gfx950 needs between them.
// M0 is listed as clobbered (the region's index lives there); clang warns
// that M0 is reserved, which is the point: nothing here keeps a value in it.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
#define SPM_ROW_TEXT                                        \
    "s_set_gpr_idx_on %[row], gpr_idx(SRC1,DST)\n\t"       \
    "v_cmp_eq_u32_e64 %[old], 0, v2\n\t"                   \
    "v_add_u32_e32 v2, %[w], v2\n\t"                       \
    "v_cmp_eq_u32_e32 vcc, 0, v2\n\t"                      \
    "s_set_gpr_idx_off\n\t"                                \
    "v_subbrev_co_u32_e64 %[t], %[junk], 0, %[z], %[old]\n\t" \
    "v_addc_co_u32_e32 %[z], vcc, 0, %[t], vcc"
#define SPM_ROW_OUTS [z] "+v"(Z), [t] "=&v"(t), [old] "=&s"(old), [junk] "=&s"(junk)
#define SPM_ROW_ENTRY1(CA)                                                                                  \
    do {                                                                                                    \
        if constexpr (VV)                                                                                   \
            asm volatile(SPM_ROW_TEXT : "+" CA(a), SPM_ROW_OUTS : [row] "s"(row), [w] "v"(w) : "vcc", "m0"); \
        else                                                                                                \
            asm volatile(SPM_ROW_TEXT : "+" CA(a), SPM_ROW_OUTS : [row] "s"(row), [w] "s"(w) : "vcc", "m0"); \
    } while (0)
#define SPM_ROW_ENTRY2(CB)                                                                                 \
    do {                                                                                                   \
        if constexpr (VV)                                                                                  \
            asm volatile(SPM_ROW_TEXT : "+{v[2:33]}"(a), "+" CB(b), SPM_ROW_OUTS : [row] "s"(row), [w] "v"(w) \
                         : "vcc", "m0");                                                                   \
        else                                                                                               \
            asm volatile(SPM_ROW_TEXT : "+{v[2:33]}"(a), "+" CB(b), SPM_ROW_OUTS : [row] "s"(row), [w] "s"(w) \
                         : "vcc", "m0");                                                                   \
    } while (0)

// One row entry for each row layout: the tuples' registers are named in
// the constraints (v2 upward, then v34 upward), which asm cannot compute.
template <int KA, int KB, bool VV>
__device__ __forceinline__ void row_entry(typename Tuple<KA>::T& a, typename Tuple<KB>::T& b, int& Z, int row, int w)
{
    int t;
    uint64_t old, junk;
    if constexpr (KB == 0) {
        if constexpr (KA == 1) SPM_ROW_ENTRY1("{v2}");
        else if constexpr (KA == 2) SPM_ROW_ENTRY1("{v[2:3]}");
        else if constexpr (KA == 3) SPM_ROW_ENTRY1("{v[2:4]}");
        else if constexpr (KA == 4) SPM_ROW_ENTRY1("{v[2:5]}");
        else if constexpr (KA == 5) SPM_ROW_ENTRY1("{v[2:6]}");
        else if constexpr (KA == 6) SPM_ROW_ENTRY1("{v[2:7]}");
        else if constexpr (KA == 7) SPM_ROW_ENTRY1("{v[2:8]}");
        else if constexpr (KA == 8) SPM_ROW_ENTRY1("{v[2:9]}");
        else if constexpr (KA == 9) SPM_ROW_ENTRY1("{v[2:10]}");
        else if constexpr (KA == 10) SPM_ROW_ENTRY1("{v[2:11]}");
        else if constexpr (KA == 11) SPM_ROW_ENTRY1("{v[2:12]}");
        else if constexpr (KA == 12) SPM_ROW_ENTRY1("{v[2:13]}");
        else if constexpr (KA == 16) SPM_ROW_ENTRY1("{v[2:17]}");
        else if constexpr (KA == 32) SPM_ROW_ENTRY1("{v[2:33]}");
        else static_assert(KA < 0, "row tuple size");
    } else {
        static_assert(KA == 32, "rows past 32 follow a full first tuple");
        if constexpr (KB == 1) SPM_ROW_ENTRY2("{v34}");
        else if constexpr (KB == 2) SPM_ROW_ENTRY2("{v[34:35]}");
        else if constexpr (KB == 3) SPM_ROW_ENTRY2("{v[34:36]}");
        else if constexpr (KB == 4) SPM_ROW_ENTRY2("{v[34:37]}");
        else if constexpr (KB == 5) SPM_ROW_ENTRY2("{v[34:38]}");
        else if constexpr (KB == 6) SPM_ROW_ENTRY2("{v[34:39]}");
        else if constexpr (KB == 7) SPM_ROW_ENTRY2("{v[34:40]}");
        else if constexpr (KB == 8) SPM_ROW_ENTRY2("{v[34:41]}");
        else if constexpr (KB == 9) SPM_ROW_ENTRY2("{v[34:42]}");
        else if constexpr (KB == 10) SPM_ROW_ENTRY2("{v[34:43]}");
        else if constexpr (KB == 11) SPM_ROW_ENTRY2("{v[34:44]}");
        else if constexpr (KB == 12) SPM_ROW_ENTRY2("{v[34:45]}");
        else if constexpr (KB == 16) SPM_ROW_ENTRY2("{v[34:49]}");
        else if constexpr (KB == 32) SPM_ROW_ENTRY2("{v[34:65]}");
        else static_assert(KB < 0, "row tuple size");
    }
}

#undef SPM_ROW_ENTRY1
#undef SPM_ROW_ENTRY2
#undef SPM_ROW_OUTS
#undef SPM_ROW_TEXT
#pragma clang diagnostic pop

// One entry of a column: MODE 0 adds the record's value (an SGPR); 1 adds
// it negated on lanes with m = -1 (the per-lane direction); 2 adds it on
// lanes with m = -1 only (chunk-start membership).
template <int N, int MODE>
__device__ __forceinline__ void entry(Rows<N>& y, int& Z, int row, int v, int m)
{
    if constexpr (MODE == 0) row_entry<kTupA<N>, kTupB<N>, false>(y.a, y.b, Z, row, v);
    else if constexpr (MODE == 1) row_entry<kTupA<N>, kTupB<N>, true>(y.a, y.b, Z, row, (v ^ m) - m);
    else row_entry<kTupA<N>, kTupB<N>, true>(y.a, y.b, Z, row, v & m);
}

// Slots U, U+1, ... of a record while U < count, nested so that a record
// costs one uniform test per entry after the first (slot 0 is always in
// use) plus the first failing one; true when all slots were in use and the
// column continues in another record.
template <int U, int N, int MODE>
__device__ __forceinline__ bool apply_slots(Rows<N>& y, int& Z, const Rec& r, int cnt, int m)
{
    if constexpr (U == kSparseSlots) {
        return r.v[0] >= 0;
    } else {
        if (U == 0 || U < cnt) {
            entry<N, MODE>(y, Z, r.v[2 + U], r.v[2 + kSparseSlots + U], m);
            return apply_slots<U + 1, N, MODE>(y, Z, r, cnt, m);
        }
        return false;
    }
}

// A whole column from its first record (already loaded); longer columns
// continue in further records, loaded here (rare for sparse inputs).
template <int N, int MODE>
__device__ __forceinline__ void apply_column(Rows<N>& y, int& Z, const Rec& first, const int* recs, int m)
{
    if (__builtin_expect(apply_slots<0, N, MODE>(y, Z, first, first.v[1], m), 0)) {
        int nx = first.v[0];
        for (;;) {
            Rec r;
            load_rec(r, recs, nx);
            if (!apply_slots<0, N, MODE>(y, Z, r, r.v[1], m)) break;
            nx = r.v[0];
        }
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
__device__ __forceinline__ void enqueue(const Rows<N>& y, bool negative, Queue& Q, uint32_t slot)
{
    int* e = Q.q + slot * kQStride<N>;
    static_for<0, N>([&](auto R) __attribute__((always_inline)) { e[decltype(R)::value] = row_get<N, decltype(R)::value>(y); });
    e[N] = negative ? 1 : 0;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Queues +/- prod_r y_r for the lanes with no zero row (a lane past the
// matrix's last chunk carries a Z that never reaches 0): one compare into a
// lane mask and a scalar test per step.
template <int N>
__device__ __forceinline__ void add_term(const Rows<N>& y, int Z, bool negative, bool small, Queue& Q, uint32_t lane,
                                         unsigned __int128& acc)
{
    // The step's test branches on a ballot whose mask is not used again, so
    // the compare goes straight to the branch; the rare taken branch takes
    // the mask anew (Z <= 0 names the same lanes, Z never being negative,
    // and is not merged with the test): 249.5 -> 228.4 ms at order 40.
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(Z == 0) == 0, 1)) return;
    uint64_t need = __builtin_amdgcn_ballot_w64(Z <= 0);
    if constexpr (kQCap<N> == 64) {
        {
            const uint32_t cnt = (uint32_t)__builtin_popcountll(need);
            if (Q.count + cnt > kQCap<N>) flush<N>(Q, small, lane, acc);
            if ((need >> lane) & 1ull) enqueue<N>(y, negative, Q, Q.count + lanes_below(need));
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
            if ((take >> lane) & 1ull) enqueue<N>(y, negative, Q, Q.count + lanes_below(take));
            Q.count += (uint32_t)__builtin_popcountll(take);
            need &= ~take;
            if (need != 0) flush<N>(Q, small, lane, acc);
        }
    }
}

// Over these orders (from where the automatic choice takes this walk to
// where a permanent takes minutes) blocks run in (even, odd) pairs, which
// makes column 0's record a constant and halves the loop control (352 ->
// 296 ms on the config-3 order-40 matrix); elsewhere one block per
// iteration keeps the kernels' code, and their compile, half as long.
constexpr int kSparsePairMinN = 26;
constexpr int kSparsePairMaxN = 48;

// Lanes past the matrix's last chunk walk a dummy chunk with this added to
// their zero count, so that they never queue a term.
constexpr int kInactiveZ = 1 << 16;

// One lane's chunk: the 2^L Gray steps starting at k = h 2^L, its nonzero
// terms queued (inactive lanes -- past the matrix's last chunk -- queue
// nothing); flushed products go into acc (mod 2^128).  Steps as in
// walk_int.h: k = 2m flips column ctz(m)+1 in a uniform direction, k = 2m+1
// flips column 0, block nblk/2 flips column L-1 in a per-lane direction.
template <int N>
__device__ __forceinline__ void walk_sparse(const int* __restrict__ recs, const int* __restrict__ y0, uint64_t h,
                                            int L, bool active, Queue& Q, uint32_t lane, unsigned __int128& acc)
{
    Rows<N> y;
    {
        const const_i32* __restrict__ ys = step_ptr(y0);
        static_for<0, N>([&](auto R) __attribute__((always_inline)) {
            row_set<N, decltype(R)::value>(y, ys[decltype(R)::value]);
        });
    }
    const bool small = ((const const_i32*)y0)[N] != 0;
    int Z = 0;
    const uint64_t k0 = h << L;
    const uint64_t g0 = k0 ^ (k0 >> 1);
    for (int j = (L > 0 ? L - 1 : 0); j < N - 1; ++j) {
        Rec r;
        load_rec(r, recs, rec_index(j, 0));
        apply_column<N, 2>(y, Z, r, recs, -(int)((g0 >> j) & 1u));
    }
    Z = active ? 0 : kInactiveZ;
    static_for<0, N>([&](auto R) __attribute__((always_inline)) {
        Z += row_get<N, decltype(R)::value>(y) == 0 ? 1 : 0;
    });
    const int lane_mask = -(int)(h & 1u);  // direction of column L-1: -1 = remove

    if constexpr (N == 1) {
        add_term<N>(y, Z, false, small, Q, lane, acc);
    } else if constexpr (N == 2) {  // L = 1
        add_term<N>(y, Z, false, small, Q, lane, acc);
        Rec r;
        load_rec(r, recs, rec_index(0, 0));
        apply_column<N, 1>(y, Z, r, recs, lane_mask);
        add_term<N>(y, Z, true, small, Q, lane, acc);
    } else {
        const uint32_t nblk = 1u << (L - 1);
        const uint32_t half = nblk >> 1;
        auto a_rec = [&](uint32_t m) __attribute__((always_inline)) {  // record of step k = 2m
            const int tz = __builtin_ctz(m);
            return rec_index(tz + 1, (int)((m >> (tz + 1)) & 1u));
        };
        Rec ra, rb;
        add_term<N>(y, Z, false, small, Q, lane, acc);  // k = 0
        load_rec(rb, recs, rec_index(0, 0));
        apply_column<N, 0>(y, Z, rb, recs, 0);  // k = 1: +column 0
        load_rec(ra, recs, a_rec(1));
        fence();
        add_term<N>(y, Z, true, small, Q, lane, acc);
        fence();
        // Block m: step 2m flips the record in ra (column ctz(m)+1), step
        // 2m+1 column 0 in plane m & 1 (PLANE, -1 when not known at compile
        // time).
        auto block = [&](uint32_t m, auto SPECIAL, auto PLANE) __attribute__((always_inline)) {
            constexpr int plane = decltype(PLANE)::value;
            if constexpr (decltype(SPECIAL)::value) apply_column<N, 1>(y, Z, ra, recs, lane_mask);
            else apply_column<N, 0>(y, Z, ra, recs, 0);
            load_rec(rb, recs, rec_index(0, plane >= 0 ? plane : (int)(m & 1u)));  // step 2m+1: column 0
            fence();
            add_term<N>(y, Z, false, small, Q, lane, acc);
            fence();
            apply_column<N, 0>(y, Z, rb, recs, 0);
            load_rec(ra, recs, a_rec(m + 1));  // m+1 = nblk: column L <= N-1, a valid record
            fence();
            add_term<N>(y, Z, true, small, Q, lane, acc);
            fence();
        };
        using Even = std::integral_constant<int, 0>;
        using Odd = std::integral_constant<int, 1>;
        using Any = std::integral_constant<int, -1>;
        if constexpr (N < kSparsePairMinN || N > kSparsePairMaxN) {  // one block per iteration
#pragma nounroll
            for (uint32_t m = 1; m < half; ++m) block(m, std::false_type{}, Any{});
            block(half, std::true_type{}, Any{});
#pragma nounroll
            for (uint32_t m = half + 1; m < nblk; ++m) block(m, std::false_type{}, Any{});
        } else if (L >= 3) {
            // half = 2^(L-2) is even: blocks 1 and half+1 alone, the rest in
            // (even, odd) pairs, so that column 0's plane is a constant
            block(1u, std::false_type{}, Odd{});
#pragma nounroll
            for (uint32_t m = 2; m < half; m += 2) {
                block(m, std::false_type{}, Even{});
                block(m + 1, std::false_type{}, Odd{});
            }
            block(half, std::true_type{}, Even{});
            block(half + 1, std::false_type{}, Odd{});
#pragma nounroll
            for (uint32_t m = half + 2; m < nblk; m += 2) {
                block(m, std::false_type{}, Even{});
                block(m + 1, std::false_type{}, Odd{});
            }
        } else {  // L = 2: one block, the special one
            block(half, std::true_type{}, Any{});
        }
    }
}

}  // namespace sparse
}  // namespace spm

#endif
