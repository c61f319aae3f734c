// Stable LSD radix sort of (key, index) pairs, shared by the Barnes-Hut
// tree (Morton keys) and the SPH neighbour grid (cell keys).  Device code
// only; each translation unit that includes it gets its own copy of the
// kernels.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "rg_internal.h"

namespace {

constexpr int kSortBS = 256;
// Keys per thread in a sort tile.  Each pass is a latency chain (ticket,
// digit-count scan, loads, look-back, scatter) whose length does not depend
// on the tile count, so a small sort wants many small tiles in flight and a
// large one wants fewer tiles (less per-tile scan and look-back work).  The
// switch is where 8-per-thread tiles would no longer cover the chip's 256
// CUs.  Measured for the whole Morton phase (bounding box, keys, 8 passes):
// 0.113 ms with 4 per thread against 0.122 ms with 8 at 120k keys, and
// 0.339 against 0.26 ms at 10^6 keys.  16 per thread from the size where
// 8-per-thread tiles would need 1.5 workgroups per CU: sort phase per BH
// step 0.193 against 0.166 ms with 8 at 524,288 keys, 0.238 against 0.265
// at 10^6, 0.384 against 0.436 at 2 x 10^6.
constexpr int kSortItemsSmall = 4;
constexpr int kSortItemsLarge = 8;
constexpr int kSortItemsHuge = 16;
constexpr int64_t kSortLargeFrom = 256ll * kSortBS * kSortItemsLarge;  // 524,288 keys
constexpr int64_t kSortHugeFrom = 384ll * kSortBS * kSortItemsLarge;   // 786,432 keys
inline int sort_items_for(int64_t n)
{
    if (rg::g_sort_items) return rg::g_sort_items;
    if (n < kSortLargeFrom) return kSortItemsSmall;
    return n < kSortHugeFrom ? kSortItemsLarge : kSortItemsHuge;
}
// Threads per sort tile: 256, or 1024 (16 waves, 4 keys each: a 4096-key
// tile with the per-thread work of a 256-thread one).  Measured sort phase
// per BH step (ms, 256-thread default tiles / 1024 x 4): 30k 0.079 / 0.092,
// 60k 0.090 / 0.100, 120k 0.104 / 0.109, 250k 0.132 / 0.127, 524,288 0.168 /
// 0.161, 10^6 0.241 (256 x 16, the same 4096-key tiles) / 0.208: below about
// 200k keys a tile's latency chain grows with its waves more than the
// shorter look-back saves; above, the 16 waves load and rank a tile's keys
// in parallel where 256 threads take 8 or 16 rounds.
constexpr int kSortBSWide = 1024;
constexpr int64_t kSortWideFrom = 200000;
inline int sort_block_for(int64_t n)
{
    if (rg::g_sort_block) return rg::g_sort_block;
    return n >= kSortWideFrom ? kSortBSWide : kSortBS;
}
inline int sort_tiles(int64_t n, int items, int bs = kSortBS)
{
    const int64_t tile = static_cast<int64_t>(bs) * items;
    return static_cast<int>((n + tile - 1) / tile);
}
// Radix digit width.  9-bit digits (7 passes for a 63-bit key) measured
// slower than 8-bit ones (8 passes): 15.3 against 12.1 us per pass at 120k.
constexpr int kDigitBits = 8;
constexpr int kBuckets = 1 << kDigitBits;         // 256
constexpr unsigned kDigitMask = kBuckets - 1u;
constexpr int kMaxPasses = (63 + kDigitBits - 1) / kDigitBits;  // 8 for a 63-bit key

// Stable LSD radix sort, one kernel per kDigitBits-bit digit ("one sweep"):
// each tile of kSortTile keys takes a ticket
// (tiles start in ticket order), counts its digits, publishes the counts and
// looks back over earlier tiles' published counts (decoupled look-back) for
// its exclusive offsets.  A look-back word packs a status (1 = this tile's
// count, 2 = inclusive prefix) with a 30-bit count, so a single relaxed
// agent-scope atomic carries both.  Every wait is on a tile that started
// earlier and publishes without waiting, and is bounded by spin_limit
// re-polls (rg::g_spin_limit, 2^20 by default; 0 gives up at once).  A wait
// that gives up adds 1 to stats[3] instead of hanging: the pass's order is
// then wrong (in bounds, but with repeated and missing entries), so every
// kernel that consumes it checks stats[3] and does nothing, and the host
// reports the failure.
constexpr unsigned kFlagAgg = 1u << 30;
constexpr unsigned kFlagInc = 2u << 30;
constexpr unsigned kCountMask = (1u << 30) - 1u;

// The status and count fields of a look-back word: 32-bit words as above;
// 64-bit words keep the status in the top two bits and a 62-bit count (two
// 31-bit counts side by side add as one while neither total reaches 2^31).
template <typename W>
struct LbWord {
    static constexpr W agg = kFlagAgg;
    static constexpr W inc = kFlagInc;
    static constexpr W mask = kCountMask;
};
template <>
struct LbWord<unsigned long long> {
    static constexpr unsigned long long agg = 1ull << 62;
    static constexpr unsigned long long inc = 2ull << 62;
    static constexpr unsigned long long mask = (1ull << 62) - 1ull;
};

// Exclusive prefix over earlier tiles of one look-back word column.  The
// words of the kLookWin tiles before the current position are loaded
// together (independent loads, one memory round trip) and consumed in tile
// order: counts are added until an inclusive word ends the walk; a word not
// yet published stops the window there, and the next window starts at it.
// The sum is the same as one word at a time; the round trips, each a
// cross-XCD atomic load, drop by up to kLookWin.  Measured per tile at
// 120k keys (118 tiles): 4.63 us of look-back one word at a time, 3.74
// with 8, 4.14 with 16, 5.62 with 32 (the wider windows issue more
// device-scope loads than they save round trips).
constexpr int kLookWin = 8;

template <int WIN = kLookWin, typename W = unsigned>
__device__ __forceinline__ W look_back(const W *__restrict__ lb, int tile, int stride,
                                       unsigned long long *__restrict__ stats, int spin_limit)
{
    constexpr int kLookWin = WIN;
    W excl = 0;
    int j = tile - 1;
    int spins = 0;
    if (spin_limit <= 0) {  // host-forced give-up (tests of the failure path)
        atomicAdd(&stats[3], 1ull);
        return 0;
    }
    while (j >= 0) {
        W v[kLookWin];
#pragma unroll
        for (int q = 0; q < kLookWin; ++q)
            v[q] = j - q >= 0 ? __hip_atomic_load(&lb[(j - q) * stride], __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : LbWord<W>::inc;  // before tile 0: an inclusive zero
        int used = 0;
        bool done = false;
#pragma unroll
        for (int q = 0; q < kLookWin; ++q) {
            if (done || used < q) continue;  // stopped earlier in this window
            const W f = v[q] & ~LbWord<W>::mask;
            if (f == 0) continue;            // not published: retry from here
            excl += v[q] & LbWord<W>::mask;
            ++used;
            if (f == LbWord<W>::inc) done = true;
        }
        if (done) break;
        j -= used;
        if (used < kLookWin) {  // stopped at an unpublished word
            if (++spins > spin_limit) {
                atomicAdd(&stats[3], 1ull);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return excl;
}

// TICKET: tiles take their index from an atomic ticket, so a tile only ever
// waits on tiles that started before it, whatever order the workgroups are
// dispatched in.  Without it a tile is its workgroup index, which lets it
// load its keys at once (one memory round trip earlier); that is used only
// when every tile of the pass fits on the GPU at the same time
// (kTicketlessTiles), so no tile can wait on one that has no slot.
//
// BS threads per tile (256 or 1024); threads 0..255 (the first four waves)
// own one digit each for the scans and the look-back.
//
// DIG: the digit of the key at input position i is dig[i] (a bucket id of
// the sample sort) instead of bits [shift, shift + 8) of the key.
template <int ITEMS, bool TICKET, int BS = kSortBS, bool DIG = false>
__global__ __launch_bounds__(BS) void k_onesweep(
    const uint64_t *__restrict__ kin, const int *__restrict__ iin, uint64_t *__restrict__ kout,
    int *__restrict__ iout, int n, int shift, const unsigned *__restrict__ ghist_p,
    unsigned *__restrict__ lookback_p, int *__restrict__ ticket_p,
    unsigned long long *__restrict__ stats, int spin_limit, const uint8_t *__restrict__ dig)
{
    static_assert(BS >= kBuckets && BS % 64 == 0, "a tile needs a thread per digit");
    constexpr int kWaves = BS / 64;
    constexpr int kOwnWaves = kBuckets / 64;   // waves whose threads own a digit each
    constexpr int kRounds = ITEMS;             // each wave: kRounds rounds of 64 keys
    constexpr int kSortTile = BS * ITEMS;
    constexpr int DPT = 1;                     // digits owned per owner thread
    __shared__ unsigned s_base[kWaves][kBuckets];
    __shared__ int s_tile;
    const int t = threadIdx.x;
    const int w = t >> 6;
    const int lane = t & 63;
    const bool owner = t < kBuckets;
    if (TICKET && t == 0) s_tile = atomicAdd(ticket_p, 1);
    for (int k = t; k < kWaves * kBuckets; k += BS) (&s_base[0][0])[k] = 0u;
    uint64_t key[kRounds];
    int idv[kRounds];
    unsigned dg[DIG ? kRounds : 1];
    auto load_keys = [&](int tl) {
        const int seg = tl * kSortTile + w * (kRounds * 64);
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const int i = seg + r * 64 + lane;
            if (i < n) {
                key[r] = kin[i];
                idv[r] = iin ? iin[i] : i;
                if (DIG) dg[DIG ? r : 0] = dig[i];
            } else {
                key[r] = 0;
                idv[r] = -1;
                if (DIG) dg[DIG ? r : 0] = 0u;
            }
        }
    };
    auto digit = [&](int r) -> unsigned {
        if (DIG) return dg[DIG ? r : 0];
        return static_cast<unsigned>((key[r] >> shift) & kDigitMask);
    };
    if (!TICKET) load_keys(blockIdx.x);
    // Exclusive scan of this pass's global digit counts with no block
    // barrier: every wave reads all of them, reduces the ranges of the
    // threads in the waves before it and scans its own range.
    unsigned gh[DPT] = {0u};
    unsigned goff[DPT] = {0u};
    if (w < kOwnWaves) {  // uniform per wave
        unsigned before = 0, gsum = 0;
#pragma unroll
        for (int ww = 0; ww < kOwnWaves; ++ww) {
            unsigned v[DPT];
            unsigned sum = 0;
#pragma unroll
            for (int q = 0; q < DPT; ++q) {
                v[q] = ghist_p[(ww * 64 + lane) * DPT + q];
                sum += v[q];
            }
            if (ww < w) {
                before += __ockl_wfred_add_u32(sum);
            } else if (ww == w) {
                gsum = sum;
#pragma unroll
                for (int q = 0; q < DPT; ++q) gh[q] = v[q];
            }
        }
        unsigned incl = gsum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned u = __shfl_up(incl, off, 64);
            if (lane >= off) incl += u;
        }
        unsigned run = before + incl - gsum;
#pragma unroll
        for (int q = 0; q < DPT; ++q) {
            goff[q] = run;
            run += gh[q];
        }
    }
    __syncthreads();  // s_tile and the zeroed counts
    const int tile = TICKET ? s_tile : static_cast<int>(blockIdx.x);
    if (TICKET) load_keys(tile);
    const int seg = tile * kSortTile + w * (kRounds * 64);
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
        if (seg + r * 64 + lane < n) atomicAdd(&s_base[w][digit(r)], 1u);
    __syncthreads();
    // thread t < 256 owns digit t: publish its count, then look back
    if (owner) {
        unsigned cnt[DPT];
#pragma unroll
        for (int q = 0; q < DPT; ++q) {
            const int d = t * DPT + q;
            cnt[q] = 0;
#pragma unroll
            for (int ww = 0; ww < kWaves; ++ww) cnt[q] += s_base[ww][d];
            __hip_atomic_store(&lookback_p[tile * kBuckets + d],
                               (tile == 0 ? kFlagInc : kFlagAgg) | cnt[q], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        unsigned excl[DPT];
#pragma unroll
        for (int q = 0; q < DPT; ++q) {
            const int d = t * DPT + q;
            excl[q] = 0;
            if (tile > 0) {
                excl[q] = look_back(lookback_p + d, tile, kBuckets, stats, spin_limit);
                __hip_atomic_store(&lookback_p[tile * kBuckets + d],
                                   kFlagInc | (excl[q] + cnt[q]), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
#pragma unroll
        for (int q = 0; q < DPT; ++q) {
            const int d = t * DPT + q;
            unsigned run = goff[q] + excl[q];
#pragma unroll
            for (int ww = 0; ww < kWaves; ++ww) {
                const unsigned c = s_base[ww][d];
                s_base[ww][d] = run;
                run += c;
            }
        }
    }
    __syncthreads();
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int i = seg + r * 64 + lane;
        const bool valid = i < n;
        const unsigned d = digit(r);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kDigitBits; ++b) {
            const unsigned long long bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const unsigned rank = __popcll(peers & lt_mask);
        const unsigned base = s_base[w][d];
        if (valid) {
            const unsigned dst = min(base + rank, static_cast<unsigned>(n - 1));
            kout[dst] = key[r];
            iout[dst] = idv[r];
            if (rank == 0) s_base[w][d] = base + static_cast<unsigned>(__popcll(peers));
        }
    }
}

// One pass over nblk tiles of kSortBS * items keys (items as sort_items_for
// chose it when the tiles were counted).
// Tiles that are surely resident together: one 256-thread workgroup per CU
// of MI355X's 256 (a onesweep workgroup needs 4 KB of LDS and few registers,
// so the real limit is several times higher).
constexpr int kTicketlessTiles = 256;

// dig != nullptr: the digits are dig[i] (k_onesweep's DIG form).
inline void launch_onesweep(int items, int bs, int nblk, hipStream_t st, const uint64_t *kin,
                            const int *iin, uint64_t *kout, int *iout, int n, int shift,
                            const unsigned *ghist_p, unsigned *lookback_p, int *ticket_p,
                            unsigned long long *stats, const uint8_t *dig = nullptr)
{
    const bool ticket = rg::g_sort_ticket || nblk > kTicketlessTiles;
#define RG_ONESWEEP(IT, TK, B, DG)                                                            \
    k_onesweep<IT, TK, B, DG><<<nblk, B, 0, st>>>(kin, iin, kout, iout, n, shift, ghist_p,    \
                                                  lookback_p, ticket_p, stats,                 \
                                                  rg::g_spin_limit, dig)
#define RG_ONESWEEP_D(IT, TK, B)                   \
    do {                                           \
        if (dig) RG_ONESWEEP(IT, TK, B, true);     \
        else RG_ONESWEEP(IT, TK, B, false);        \
    } while (0)
    if (bs == kSortBSWide) {  // 4 keys per thread
        if (ticket) RG_ONESWEEP_D(kSortItemsSmall, true, kSortBSWide);
        else RG_ONESWEEP_D(kSortItemsSmall, false, kSortBSWide);
        return;
    }
    if (items == kSortItemsSmall) {
        if (ticket) RG_ONESWEEP_D(kSortItemsSmall, true, kSortBS);
        else RG_ONESWEEP_D(kSortItemsSmall, false, kSortBS);
    } else if (items == kSortItemsHuge) {
        if (ticket) RG_ONESWEEP_D(kSortItemsHuge, true, kSortBS);
        else RG_ONESWEEP_D(kSortItemsHuge, false, kSortBS);
    } else {
        if (ticket) RG_ONESWEEP_D(kSortItemsLarge, true, kSortBS);
        else RG_ONESWEEP_D(kSortItemsLarge, false, kSortBS);
    }
#undef RG_ONESWEEP_D
#undef RG_ONESWEEP
}

// ------------------------------------------------------------ sample sort
// The second half of a sample sort.  A stable pass over bucket ids (the
// DIG form of k_onesweep) leaves the keys grouped by bucket, each bucket a
// contiguous range whose length is hist[b], in input order within the
// bucket; the buckets are key ranges in increasing order.  One block per
// bucket then sorts its range by key, stably (LSD radix, 8-bit digits), so
// the whole sequence ends in (key, input position) order: the stable order.
//
// A bucket of up to kBucketCap keys is sorted in LDS: the keys stay in
// registers, each pass ranks them (per-wave digit counts, a digit-major
// scan, ballot ranks in position order) and exchanges them through LDS.
// Only the digits below the highest bit in which the bucket's smallest and
// largest keys differ are sorted; the bits above are equal in every key.
// A larger bucket (the splitters came from a different distribution) is
// sorted in global memory by the same block, chunk by chunk, through the
// spare buffers (kalt, ialt), and counted in stats[6].
constexpr int kBucketBS = 1024;
constexpr int kBucketWaves = kBucketBS / 64;
constexpr int kBucketRounds = 8;                       // keys per thread, at most
constexpr int kBucketCap = kBucketBS * kBucketRounds;  // 8192 keys in LDS
constexpr int kMaxBuckets = kBuckets;                  // bucket ids are one digit

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long u = __shfl_xor(v, o, 64);
        v = u < v ? u : v;
    }
    return v;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

// Threads 0..255 each own one digit: s_base[w][d] holds wave w's count of
// digit d; on return it holds the exclusive base of (d, w) in digit-major,
// wave-minor order, starting at run0[d] (absolute bases) or, with run0 null,
// at the exclusive scan of the digit totals.  run0, when given, advances by
// the digit totals.  Contains block barriers; every thread calls it.
__device__ __forceinline__ void bucket_bases(unsigned (*s_base)[kBuckets], unsigned *run0,
                                             unsigned *s_ws)
{
    const int t = threadIdx.x;
    const int lane = t & 63;
    unsigned tot = 0, incl = 0;
    if (t < kBuckets) {
#pragma unroll
        for (int ww = 0; ww < kBucketWaves; ++ww) tot += s_base[ww][t];
        if (!run0) {
            incl = tot;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const unsigned u = __shfl_up(incl, off, 64);
                if (lane >= off) incl += u;
            }
            if (lane == 63) s_ws[t >> 6] = incl;
        }
    }
    __syncthreads();
    if (t < kBuckets) {
        unsigned run;
        if (run0) {
            run = run0[t];
            run0[t] = run + tot;
        } else {
            unsigned before = 0;
            for (int ww = 0; ww < (t >> 6); ++ww) before += s_ws[ww];
            run = before + incl - tot;
        }
#pragma unroll
        for (int ww = 0; ww < kBucketWaves; ++ww) {
            const unsigned c = s_base[ww][t];
            s_base[ww][t] = run;
            run += c;
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(kBucketBS) void k_bucket_sort(
    uint64_t *__restrict__ keys, int *__restrict__ idx, uint64_t *__restrict__ kalt,
    int *__restrict__ ialt, const unsigned *__restrict__ hist, int key_bits,
    unsigned long long *__restrict__ stats)
{
    if (stats[3] != 0ull) return;  // the bucket pass gave up: its order is invalid
    __shared__ uint64_t s_key[kBucketCap];
    __shared__ int s_idx[kBucketCap];
    __shared__ unsigned s_base[kBucketWaves][kBuckets];
    __shared__ unsigned s_run[kBuckets];
    __shared__ unsigned s_ws[kBucketWaves];
    __shared__ unsigned long long s_mm[2][kBucketWaves];
    __shared__ unsigned s_off;
    const int t = threadIdx.x;
    const int w = t >> 6;
    const int lane = t & 63;
    const int b = blockIdx.x;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;

    // this bucket's range: the counts of the buckets before it
    {
        unsigned v = (t < b) ? hist[t] : 0u;  // b < kMaxBuckets <= kBucketBS
        v = __ockl_wfred_add_u32(v);
        if (lane == 0) s_ws[w] = v;
        __syncthreads();
        if (t == 0) {
            unsigned o = 0;
            for (int ww = 0; ww < kBucketWaves; ++ww) o += s_ws[ww];
            s_off = o;
        }
        __syncthreads();
    }
    const int off = static_cast<int>(s_off);
    const int cnt = static_cast<int>(hist[b]);
    if (cnt <= 1) return;

    // Rank the valid keys of one chunk (positions w * ru * 64 + r * 64 +
    // lane, r < ru, below m) by digit d = (key >> shift) & 255 and hand each
    // (key, index, destination) to put; bases come from bucket_bases.
    // (Loops over rounds run to kBucketRounds, unrolled, so the key
    // registers are indexed statically; rounds past ru are skipped.)
    auto rank_chunk = [&](const uint64_t (&kr)[kBucketRounds], const int (&ir)[kBucketRounds],
                          int ru, int m, int shift, unsigned *run0, auto &&put) {
        for (int k = t; k < kBucketWaves * kBuckets; k += kBucketBS) (&s_base[0][0])[k] = 0u;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kBucketRounds; ++r) {
            const int q = (w * ru + r) * 64 + lane;
            if (r < ru && q < m) atomicAdd(&s_base[w][(kr[r] >> shift) & kDigitMask], 1u);
        }
        __syncthreads();
        bucket_bases(s_base, run0, s_ws);
#pragma unroll
        for (int r = 0; r < kBucketRounds; ++r) {
            if (r >= ru) break;  // uniform
            const int q = (w * ru + r) * 64 + lane;
            const bool valid = q < m;
            const unsigned d = static_cast<unsigned>((kr[r] >> shift) & kDigitMask);
            unsigned long long peers = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < kDigitBits; ++bit) {
                const unsigned long long bb = __ballot((d >> bit) & 1u);
                peers &= ((d >> bit) & 1u) ? bb : ~bb;
            }
            const unsigned rank = __popcll(peers & lt_mask);
            const unsigned base = s_base[w][d];
            if (valid) {
                put(base + rank, kr[r], ir[r]);
                if (rank == 0) s_base[w][d] = base + static_cast<unsigned>(__popcll(peers));
            }
        }
        __syncthreads();
    };

    uint64_t key[kBucketRounds];
    int idv[kBucketRounds];
    if (cnt <= kBucketCap) {
        const int ru = (cnt + kBucketBS - 1) / kBucketBS;  // rounds in use, uniform
        unsigned long long kmin = ~0ull, kmax = 0ull;
#pragma unroll
        for (int r = 0; r < kBucketRounds; ++r) {
            const int q = (w * ru + r) * 64 + lane;
            if (r < ru && q < cnt) {
                key[r] = keys[off + q];
                idv[r] = idx[off + q];
                kmin = key[r] < kmin ? key[r] : kmin;
                kmax = key[r] > kmax ? key[r] : kmax;
            } else {
                key[r] = 0ull;
                idv[r] = 0;
            }
        }
        kmin = wave_min_u64(kmin);
        kmax = wave_max_u64(kmax);
        if (lane == 0) {
            s_mm[0][w] = kmin;
            s_mm[1][w] = kmax;
        }
        __syncthreads();
        unsigned long long lo = ~0ull, hi = 0ull;
        for (int ww = 0; ww < kBucketWaves; ++ww) {
            lo = s_mm[0][ww] < lo ? s_mm[0][ww] : lo;
            hi = s_mm[1][ww] > hi ? s_mm[1][ww] : hi;
        }
        const unsigned long long diff = lo ^ hi;
        if (diff == 0ull) return;  // equal keys: input order is the stable order
        const int top = 64 - __clzll(static_cast<long long>(diff));
        const int passes = (top + kDigitBits - 1) / kDigitBits;
        for (int p = 0; p < passes; ++p) {
            rank_chunk(key, idv, ru, cnt, kDigitBits * p, nullptr,
                       [&](unsigned dst, uint64_t k, int i) {
                           s_key[dst] = k;
                           s_idx[dst] = i;
                       });
            // back to registers in position order (rank_chunk ended on a barrier)
#pragma unroll
            for (int r = 0; r < kBucketRounds; ++r) {
                const int q = (w * ru + r) * 64 + lane;
                if (r < ru && q < cnt) {
                    key[r] = s_key[q];
                    idv[r] = s_idx[q];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kBucketRounds; ++r) {
            const int q = (w * ru + r) * 64 + lane;
            if (r < ru && q < cnt) {
                keys[off + q] = key[r];
                idx[off + q] = idv[r];
            }
        }
        return;
    }

    // Oversize bucket: passes over global memory, chunk by chunk in order,
    // ping-ponging between the buffers; digit bases from the whole bucket's
    // digit counts.
    if (t == 0) atomicAdd(&stats[6], 1ull);
    const int passes = (key_bits + kDigitBits - 1) / kDigitBits;
    uint64_t *ksrc = keys, *kdst = kalt;
    int *isrc = idx, *idst = ialt;
    for (int p = 0; p < passes; ++p) {
        const int shift = kDigitBits * p;
        if (t < kBuckets) s_run[t] = 0u;
        __syncthreads();
        for (int q = t; q < cnt; q += kBucketBS)
            atomicAdd(&s_run[(ksrc[off + q] >> shift) & kDigitMask], 1u);
        __syncthreads();
        // s_run: exclusive scan of the digit totals (owners, one wave scan each)
        unsigned tot = 0, incl = 0;
        if (t < kBuckets) {
            tot = s_run[t];
            incl = tot;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned u = __shfl_up(incl, o, 64);
                if (lane >= o) incl += u;
            }
            if (lane == 63) s_ws[w] = incl;
        }
        __syncthreads();
        if (t < kBuckets) {
            unsigned before = 0;
            for (int ww = 0; ww < w; ++ww) before += s_ws[ww];
            s_run[t] = before + incl - tot;
        }
        __syncthreads();
        for (int c0 = 0; c0 < cnt; c0 += kBucketCap) {
            const int m = min(kBucketCap, cnt - c0);
            const int ru = (m + kBucketBS - 1) / kBucketBS;
#pragma unroll
            for (int r = 0; r < kBucketRounds; ++r) {
                const int q = (w * ru + r) * 64 + lane;
                if (r < ru && q < m) {
                    key[r] = ksrc[off + c0 + q];
                    idv[r] = isrc[off + c0 + q];
                } else {
                    key[r] = 0ull;
                    idv[r] = 0;
                }
            }
            rank_chunk(key, idv, ru, m, shift, s_run, [&](unsigned dst, uint64_t k, int i) {
                kdst[off + dst] = k;
                idst[off + dst] = i;
            });
        }
        uint64_t *kt = ksrc;
        ksrc = kdst;
        kdst = kt;
        int *it = isrc;
        isrc = idst;
        idst = it;
    }
    if (ksrc != keys) {  // odd pass count: the result is in the spare buffers
        for (int q = t; q < cnt; q += kBucketBS) {
            keys[off + q] = ksrc[off + q];
            idx[off + q] = isrc[off + q];
        }
    }
}

}  // namespace
