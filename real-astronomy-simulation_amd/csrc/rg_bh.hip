// LBVH Barnes-Hut treecode, entirely on the device.
//
// Reference: core/solvers/barnes_hut.py (paths relative to the reference
// checkout).  The reference computes Morton keys and a stable argsort on the
// host in numpy, then builds a Karras radix tree, aggregates every node by
// per-ancestor FP64 atomics and walks a per-particle stack.  Here:
//
//   1. k_bbox_*    FP64 bounding box (exact min/max, fixed-order reduction)
//   2. k_morton    keys bit-identical to morton_codes (barnes_hut.py:38-46)
//   3. k_onesweep  stable LSD radix sort of (key, index), 8-bit digits (8
//                  passes for 63-bit keys), one kernel per digit with
//                  decoupled look-back; the permutation equals
//                  np.argsort(kind="stable") (:247)
//   4. k_build     Karras build with the reference's exact delta (incl. the
//                  index tie-break) using hardware clz (:50-121)
//   5. k_agg_*     leaf init + deterministic aggregation (:124-177), every
//                  node computed exactly once and bitwise reproducibly:
//                  treelets of <= 256 leaves bottom-up in LDS (the second
//                  child to finish combines both), then every node above
//                  them on its own as the sum of its run of treelets, in a
//                  fixed order (k_rank_preorder orders the treelets).
//   6. k_walk*     stackless walk in right-first DFS order: the node
//                  sequence the reference's stack walk pops (:180-217).
//                  FP64 follows right-child/escape links over node-id
//                  records; FP32 scans records stored in that preorder,
//                  moving to the next slot to open a node and skipping the
//                  subtree to accept it.
//
// Node ids follow the reference: leaves 0..n-1 in sorted order, internal
// node i has id i+n, root id n.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "rg_internal.h"
#include "rg_sort.h"

// Wavefront min-reduction from the ROCm device library (ockl).
extern "C" __device__ int __ockl_wfred_min_i32(int);

namespace {

constexpr int kBoxBS = 256;
constexpr int kWalkBSDefault = 256;

int g_walk_bs = kWalkBSDefault;
// FP32 walk: -1 auto, 0 one lane, 1 two lanes, 2 wave-coherent, 3 wave with two
// slots per step, 4 / 5 wave-coherent split over two / four waves per particle
// group, 6 wave-coherent over far-field chunks taken by four waves
int g_walk_mode = -1;
// chunked walk shape: nearest boundary distance, growth 2^shift (0: by n),
// XCD run length, waves per group (0: by n)
int g_chunk_dmin = 256;
int g_chunk_shift = 0;
int g_chunk_run = 32;
int g_chunk_waves = 0;
int g_chunk_read_ahead = -1;  // walk 7's read-ahead form: -1 (default) or 1 on, 0 off

struct alignas(16) WalkRec64 {
    double x, y, z, m;   // centre of mass, mass
    double size;         // max-axis AABB extent
    int right;
    int escape;
};

// --------------------------------------------------------------- bbox
// Block-wide min / max / sum of the 7 box values (lo xyz, hi xyz, non-finite
// count): wave butterflies, then the 4 wave results combined in wave order
// by every thread (so all threads hold the same result).  Min and max are
// exact in any order; the count is a sum of integers held exactly.
template <int BS = kBoxBS>
__device__ __forceinline__ void block_box_reduce(double v[7], double (*s_w)[7])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            v[a] = fmin(v[a], __shfl_xor(v[a], off, 64));
            v[3 + a] = fmax(v[3 + a], __shfl_xor(v[3 + a], off, 64));
        }
        v[6] += __shfl_xor(v[6], off, 64);
    }
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < 7; ++c) s_w[w][c] = v[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 7; ++c) v[c] = s_w[0][c];
    for (int ww = 1; ww < BS / 64; ++ww) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            v[a] = fmin(v[a], s_w[ww][a]);
            v[3 + a] = fmax(v[3 + a], s_w[ww][3 + a]);
        }
        v[6] += s_w[ww][6];
    }
}

// Statistics words of a compute (device, 8 x u64): [0] nodes visited by the
// walk, [1] the most one particle visited, [2] non-finite coordinates,
// [3] look-back waits that gave up (radix sort, treelet ranking: the order
// and tree are then invalid), [4] FP64 walks that ran out of bounds (a
// corrupt tree), [6] sample-sort buckets sorted in global memory.
constexpr int kStatWords = 8;

// A look-back wait of this compute gave up (stats[3]): every kernel after
// the sort returns at once, so nothing indexes through the invalid order.
__device__ __forceinline__ bool tree_aborted(const unsigned long long *__restrict__ stats)
{
    return stats[3] != 0ull;
}

// Resets, by block 0 of the first pass of a compute, for the Morton pass
// and the sort and tree passes after it: the digit histograms, the sort's
// tile tickets, the tree's counters (rank tickets at [2], treelets at [3])
// and the statistics (all but [2], which the Morton pass writes).
__device__ __forceinline__ void reset_counters(unsigned long long *__restrict__ stats,
                                               unsigned *__restrict__ ghist,
                                               int *__restrict__ tickets,
                                               int *__restrict__ counts)
{
    const int t = threadIdx.x;
    if (ghist)
        for (int k = t; k < kMaxPasses * kBuckets; k += blockDim.x) ghist[k] = 0u;
    if (tickets && t < kMaxPasses) tickets[t] = 0;
    if (counts && t < 4) counts[t] = 0;
    if (stats && t == 0) {
        stats[0] = 0ull;
        stats[1] = 0ull;
        stats[3] = 0ull;
        stats[4] = 0ull;
        stats[6] = 0ull;
    }
}

// Per-block partial box of the finite coordinates and count of non-finite
// ones (plus the counter resets in block 0).
__global__ __launch_bounds__(kBoxBS) void k_bbox_partial(const double *__restrict__ pos, int n,
                                                          double *__restrict__ part,
                                                          unsigned long long *__restrict__ stats,
                                                          unsigned *__restrict__ ghist,
                                                          int *__restrict__ tickets,
                                                          int *__restrict__ counts)
{
    __shared__ double s_w[kBoxBS / 64][7];
    const int t = threadIdx.x;
    if (blockIdx.x == 0) reset_counters(stats, ghist, tickets, counts);
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (int i = blockIdx.x * kBoxBS + t; i < n; i += gridDim.x * kBoxBS) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double x = pos[3 * i + a];
            if (isfinite(x)) {
                v[a] = fmin(v[a], x);
                v[3 + a] = fmax(v[3 + a], x);
            } else {
                v[6] += 1.0;
            }
        }
    }
    block_box_reduce(v, s_w);
    if (t == 0) {
        const int nb = gridDim.x;
#pragma unroll
        for (int c = 0; c < 7; ++c) part[c * nb + blockIdx.x] = v[c];
    }
}

// The opening kick and drift of a KDK step (leapfrog.py:11-22: v += a h,
// x += v dt, per component, the same two roundings as k_kick_drift) with the
// partial box of the new positions, one thread per particle -- the drifted
// positions are boxed while in registers instead of read back by a
// separate pass.  Block 0 also resets the counters, as k_bbox_partial.
constexpr int kKdBS = 256;  // 1024-thread blocks left half the CUs idle at 120k
__global__ __launch_bounds__(kKdBS) void k_kick_drift_box(
    double *__restrict__ pos, double *__restrict__ vel, const double *__restrict__ acc, int n,
    double dt_half, double dt, double *__restrict__ part, unsigned long long *__restrict__ stats,
    unsigned *__restrict__ ghist, int *__restrict__ tickets, int *__restrict__ counts)
{
    __shared__ double s_w[kKdBS / 64][7];
    const int t = threadIdx.x;
    if (blockIdx.x == 0) reset_counters(stats, ghist, tickets, counts);
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (int i = blockIdx.x * kKdBS + t; i < n; i += gridDim.x * kKdBS) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int k = 3 * i + a;
            const double vk = vel[k] + acc[k] * dt_half;
            vel[k] = vk;
            const double x = pos[k] + vk * dt;
            pos[k] = x;
            if (isfinite(x)) {
                v[a] = fmin(v[a], x);
                v[3 + a] = fmax(v[3 + a], x);
            } else {
                v[6] += 1.0;
            }
        }
    }
    block_box_reduce<kKdBS>(v, s_w);
    if (t == 0) {
        const int nb = gridDim.x;
#pragma unroll
        for (int c = 0; c < 7; ++c) part[c * nb + blockIdx.x] = v[c];
    }
}

// The box from the nb partials, reduced by each block of the consumer (no
// separate final launch): lo[3] and scale = (2^bits - 1) / (span *
// 1.0000001), span = max over axes of (hi - lo), 1.0 if not positive
// (barnes_hut.py:40-43).  Returns the non-finite count.
__device__ double box_from_partials(const double *__restrict__ part, int nb, int bits,
                                    double lo[3], double *scale, double (*s_w)[7])
{
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (int b = threadIdx.x; b < nb; b += kBoxBS) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            v[a] = fmin(v[a], part[a * nb + b]);
            v[3 + a] = fmax(v[3 + a], part[(3 + a) * nb + b]);
        }
        v[6] += part[6 * nb + b];
    }
    block_box_reduce(v, s_w);
    double span = v[3] - v[0];
    for (int a = 1; a < 3; ++a) {
        const double sa = v[3 + a] - v[a];
        if (sa > span) span = sa;
    }
    if (!(span > 0.0)) span = 1.0;
    const double qmax = static_cast<double>((1ull << bits) - 1ull);
    lo[0] = v[0];
    lo[1] = v[1];
    lo[2] = v[2];
    *scale = qmax / (span * 1.0000001);
    return v[6];
}

// ------------------------------------------------------------- morton
__device__ __forceinline__ uint64_t part1by2(uint64_t x)
{
    // barnes_hut.py:27-35: spread the low 21 bits to every third bit.
    x &= 0x1FFFFFull;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// Keys, plus (when passes > 0) the global histogram of every radix digit
// for the radix sort and the reset of the sort's look-back flags.  Every
// block first reduces the bbox partials to the box; block 0 stores it and
// the non-finite count (stats[2]).
__global__ __launch_bounds__(kBoxBS) void k_morton(const double *__restrict__ pos, int n, int bits,
                                                   const double *__restrict__ part, int nb,
                                                   double *__restrict__ box,
                                                   unsigned long long *__restrict__ stats,
                                                   uint64_t *__restrict__ keys, int passes,
                                                   unsigned *__restrict__ ghist,
                                                   unsigned *__restrict__ lookback, int lb_words,
                                                   int2 *__restrict__ tl_of)
{
    __shared__ unsigned s_h[kMaxPasses * kBuckets];
    __shared__ double s_w[kBoxBS / 64][7];
    for (int k = threadIdx.x; k < passes * kBuckets; k += kBoxBS) s_h[k] = 0u;
    double lo[3], scale;
    const double bad = box_from_partials(part, nb, bits, lo, &scale, s_w);  // has a block barrier
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        box[0] = lo[0];
        box[1] = lo[1];
        box[2] = lo[2];
        box[3] = scale;
        if (stats) stats[2] = static_cast<unsigned long long>(bad);
    }
    const double qmax = static_cast<double>((1ull << bits) - 1ull);
    for (int i = blockIdx.x * kBoxBS + threadIdx.x; i < n; i += gridDim.x * kBoxBS) {
        uint64_t key = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            // (pos - lo) * scale, clipped to [0, qmax], truncated (numpy astype).
            double t = (pos[3 * i + a] - lo[a]) * scale;
            if (!(t >= 0.0)) t = 0.0;
            if (t > qmax) t = qmax;
            key |= part1by2(static_cast<uint64_t>(t)) << a;
        }
        keys[i] = key;
        // set again by the build: the treelet starting here, and the count of
        // internal nodes whose leaf range ends here
        if (tl_of) tl_of[i] = make_int2(-1, 0);
        for (int p = 0; p < passes; ++p)
            atomicAdd(&s_h[p * kBuckets + ((key >> (kDigitBits * p)) & kDigitMask)], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < passes * kBuckets; k += kBoxBS)
        if (s_h[k]) atomicAdd(&ghist[k], s_h[k]);
    for (int k = blockIdx.x * kBoxBS + threadIdx.x; k < lb_words; k += gridDim.x * kBoxBS)
        lookback[k] = 0u;
}

// Sample-sort form of k_morton: keys, and each key's bucket among the
// nbuckets key ranges that nbuckets - 1 splitters cut, with the histogram of
// bucket ids (the one digit the bucket pass sorts by).  The splitters are
// the composites (key, index) at every stride-th position of the previous
// compute's sorted keys and order: key values are places in the box, and
// the count of particles below a place changes only by the particles that
// cross it, so each bucket stays near stride keys while particles move
// little between computes (the current keys of sampled particles would not:
// one that crosses a coarse octree boundary jumps far in the order, and its
// neighbours' buckets merge).  A bucket's share only affects speed, never
// the order: k_bucket_sort sorts a bucket of any size, and every block
// sorts the splitters in LDS first, so any values work.  Bucket of particle
// i: the number of splitters <= (key_i, i).
__device__ __forceinline__ uint64_t morton_key_at(const double *__restrict__ pos, int i,
                                                  const double lo[3], double scale, double qmax)
{
    uint64_t key = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double t = (pos[3 * i + a] - lo[a]) * scale;
        if (!(t >= 0.0)) t = 0.0;
        if (t > qmax) t = qmax;
        key |= part1by2(static_cast<uint64_t>(t)) << a;
    }
    return key;
}

__global__ __launch_bounds__(kBoxBS) void k_morton_sample(
    const double *__restrict__ pos, int n, int bits, const double *__restrict__ part, int nb,
    double *__restrict__ box, unsigned long long *__restrict__ stats, uint64_t *__restrict__ keys,
    const uint64_t *__restrict__ prev_keys, const int *__restrict__ prev_idx, int stride,
    int nbuckets, uint8_t *__restrict__ bkt,
    unsigned *__restrict__ ghist, unsigned *__restrict__ lookback, int lb_words,
    int2 *__restrict__ tl_of)
{
    static_assert(kBoxBS == kBuckets, "one thread per splitter slot and per bucket");
    __shared__ unsigned s_h[kBuckets];
    __shared__ uint64_t s_sk[kBuckets];
    __shared__ int s_sp[kBuckets];
    __shared__ double s_w[kBoxBS / 64][7];
    const int t = threadIdx.x;
    s_h[t] = 0u;
    double lo[3], scale;
    const double bad = box_from_partials(part, nb, bits, lo, &scale, s_w);  // has a block barrier
    if (blockIdx.x == 0 && t == 0) {
        box[0] = lo[0];
        box[1] = lo[1];
        box[2] = lo[2];
        box[3] = scale;
        if (stats) stats[2] = static_cast<unsigned long long>(bad);
    }
    const double qmax = static_cast<double>((1ull << bits) - 1ull);
    // splitter t (t < nbuckets - 1); the slots past them sort last
    uint64_t sk = ~0ull;
    int sp = 0x7fffffff;
    if (t < nbuckets - 1) {
        sk = prev_keys[(t + 1) * stride];
        sp = prev_idx[(t + 1) * stride];
    }
    s_sk[t] = sk;
    s_sp[t] = sp;
    __syncthreads();
    // Taken from a sorted result, the splitters are in order already; only
    // an aborted previous compute leaves them out of order, and then the
    // block sorts them (bitonic, ascending).
    bool down = false;
    if (t + 1 < kBuckets) {
        const uint64_t nk = s_sk[t + 1];
        down = nk < sk || (nk == sk && s_sp[t + 1] < sp);
    }
    if (__syncthreads_or(down)) {
        for (int k = 2; k <= kBuckets; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                const int u = t ^ j;
                if (u > t) {
                    const uint64_t ak = s_sk[t], bk = s_sk[u];
                    const int ap = s_sp[t], bp = s_sp[u];
                    const bool b_less = bk < ak || (bk == ak && bp < ap);
                    const bool up = (t & k) == 0;
                    if (b_less == up) {
                        s_sk[t] = bk;
                        s_sk[u] = ak;
                        s_sp[t] = bp;
                        s_sp[u] = ap;
                    }
                }
                __syncthreads();
            }
        }
    }
    for (int i = blockIdx.x * kBoxBS + t; i < n; i += gridDim.x * kBoxBS) {
        const uint64_t key = morton_key_at(pos, i, lo, scale, qmax);
        keys[i] = key;
        // set again by the build: the treelet starting here, and the count of
        // internal nodes whose leaf range ends here
        if (tl_of) tl_of[i] = make_int2(-1, 0);
        // splitters <= (key, i): at most nbuckets - 1 <= 255 of the 256 slots
        int c = 0;
#pragma unroll
        for (int step = kBuckets / 2; step > 0; step >>= 1) {
            const uint64_t k2 = s_sk[c + step - 1];
            if (k2 < key || (k2 == key && s_sp[c + step - 1] <= i)) c += step;
        }
        bkt[i] = static_cast<uint8_t>(c);
        atomicAdd(&s_h[c], 1u);
    }
    __syncthreads();
    if (t < nbuckets && s_h[t]) atomicAdd(&ghist[t], s_h[t]);
    for (int k = blockIdx.x * kBoxBS + t; k < lb_words; k += gridDim.x * kBoxBS)
        lookback[k] = 0u;
}

// The opening kick + drift fused into the box pass (rg_bh_step).
struct KickDrift {
    double *pos;
    double *vel;
    const double *acc;
    double dt_half, dt;
};

// Splitters and bucket ids for the sample-sort form of the key pass.
struct SampleSort {
    const uint64_t *prev_keys;  // the previous compute's sorted keys
    const int *prev_idx;        // and order (not the buffers this pass writes)
    int stride;                 // splitter spacing in them
    int nbuckets;               // <= kMaxBuckets
    uint8_t *bkt;               // bucket id per particle (out)
};

int launch_bbox_morton(const double *pos, int n, int bits, double *part, double *box,
                       unsigned long long *stats, uint64_t *keys, int passes, unsigned *ghist,
                       int *tickets, unsigned *lookback, int lb_words, int *counts, int2 *tl_of,
                       hipStream_t st, const KickDrift *kd = nullptr,
                       const SampleSort *ss = nullptr)
{
    // at most 1024 partials (the part buffer's size) either way
    int nb;
    if (kd) {
        nb = min((n + kKdBS - 1) / kKdBS, 1024);
        k_kick_drift_box<<<nb, kKdBS, 0, st>>>(kd->pos, kd->vel, kd->acc, n, kd->dt_half, kd->dt,
                                               part, stats, ghist, tickets, counts);
    } else {
        nb = min((n + kBoxBS - 1) / kBoxBS, 1024);
        k_bbox_partial<<<nb, kBoxBS, 0, st>>>(pos, n, part, stats, ghist, tickets, counts);
    }
    RG_CHECK_LAUNCH();
    // each block flushes passes x kBuckets global atomics: keep the grid modest
    const int mg = min((n + kBoxBS - 1) / kBoxBS, passes > 0 ? 256 : 2048);
    if (ss)
        k_morton_sample<<<mg, kBoxBS, 0, st>>>(pos, n, bits, part, nb, box, stats, keys,
                                               ss->prev_keys, ss->prev_idx, ss->stride,
                                               ss->nbuckets, ss->bkt,
                                               ghist, lookback, lb_words, tl_of);
    else
        k_morton<<<mg, kBoxBS, 0, st>>>(pos, n, bits, part, nb, box, stats, keys, passes, ghist,
                                        lookback, lb_words, tl_of);
    RG_CHECK_LAUNCH();
    return RG_OK;
}

size_t bbox_part_doubles(int n)
{
    const int nb = min((n + kBoxBS - 1) / kBoxBS, 1024);
    return 7 * static_cast<size_t>(nb) + 8;
}

// --------------------------------------------------------- radix sort
// ------------------------------------------------------------- build
__device__ __forceinline__ int delta_dev(int64_t i, int64_t j, int64_t n, uint64_t ci,
                                         const uint64_t *__restrict__ code)
{
    // barnes_hut.py:71-81.  _clz64 counts from bit 62 (= clzll - 1 for a
    // 63-bit value); the equal-key tie-break 63 + _clz_i32(i ^ j) counts
    // from bit 30 (= 62 + clz32).
    if (j < 0 || j >= n) return -1;
    const uint64_t x = ci ^ code[j];
    if (x == 0) return 62 + __clz(static_cast<int>(i ^ j));
    return __clzll(static_cast<long long>(x)) - 1;
}

constexpr int kTreelet = 256;  // max leaves of a treelet (one workgroup, LDS)

// The Karras build, plus the aggregation plan and the preorder counts.
// Treelets are the maximal subtrees with at most kTreelet leaves (a treelet
// root has at most kTreelet leaves and its parent more); they partition the
// leaves.  The thread of internal node i knows the leaf counts of i and of
// both children, so it marks each child that is a treelet root:
// tl_of[first leaf of a treelet].x = its root.  It also counts itself at
// the last leaf of its range, tl_of[hi].y (the ranking pass scans these
// counts into preorder slots).  Both fields were cleared by the key pass.
// Also clears the ranking pass's look-back words.
__global__ void k_build(const uint64_t *__restrict__ code, int n, int *__restrict__ left,
                        int *__restrict__ right, int *__restrict__ parent,
                        int2 *__restrict__ range, int *__restrict__ left_at_split,
                        int2 *__restrict__ tl_of, unsigned long long *__restrict__ rank_lb,
                        int rank_tiles, const unsigned long long *__restrict__ stats)
{
    if (tree_aborted(stats)) return;
    const int64_t nn = n;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < rank_tiles;
         k += gridDim.x * blockDim.x)
        rank_lb[k] = 0ull;
    for (int ii = blockIdx.x * blockDim.x + threadIdx.x; ii < n - 1;
         ii += gridDim.x * blockDim.x) {
        const int64_t i = ii;
        const uint64_t ci = code[i];
        // barnes_hut.py:92-121, step for step.
        const int64_t d = delta_dev(i, i + 1, nn, ci, code) > delta_dev(i, i - 1, nn, ci, code) ? 1 : -1;
        const int d_min = delta_dev(i, i - d, nn, ci, code);
        int64_t l_max = 2;
        while (delta_dev(i, i + l_max * d, nn, ci, code) > d_min) l_max *= 2;
        int64_t l = 0;
        for (int64_t t = l_max / 2; t >= 1; t /= 2)
            if (delta_dev(i, i + (l + t) * d, nn, ci, code) > d_min) l += t;
        const int64_t j = i + l * d;
        const int d_node = delta_dev(i, j, nn, ci, code);
        int64_t s = 0;
        int64_t t2 = (l + 1) / 2;
        while (t2 >= 1) {
            if (delta_dev(i, i + (s + t2) * d, nn, ci, code) > d_node) s += t2;
            t2 = t2 > 1 ? (t2 + 1) / 2 : 0;
        }
        const int64_t gamma = i + s * d + (d < 0 ? d : 0);
        const int64_t lo = i < j ? i : j;
        const int64_t hi = i < j ? j : i;
        const int lc = static_cast<int>(lo == gamma ? gamma : gamma + nn);
        const int rc = static_cast<int>(hi == gamma + 1 ? gamma + 1 : gamma + 1 + nn);
        const int self = static_cast<int>(i + nn);
        left[i] = lc;
        right[i] = rc;
        parent[lc] = self;
        parent[rc] = self;
        range[i] = make_int2(static_cast<int>(lo), static_cast<int>(hi));
        left_at_split[gamma] = lc;
        if (i == 0) parent[nn] = -1;
        if (hi - lo + 1 > kTreelet) {
            if (gamma - lo + 1 <= kTreelet) tl_of[lo].x = lc;
            if (hi - gamma <= kTreelet) tl_of[gamma + 1].x = rc;
        } else if (i == 0) {  // the root is itself a treelet
            tl_of[0].x = self;
        }
        atomicAdd(&tl_of[hi].y, 1);
    }
}

// Right-first DFS preorder slots (the order the reference's stack walk pops
// nodes in).  A node v with leaf range [lo, hi] is preceded by the nodes of
// the subtrees right of it, 2 (n-1-hi) - (their count), and by its
// ancestors; each ancestor where the path turns left roots one of those
// subtrees, so pre(v) = 2 (n-1-hi) + R(v), R the right turns on the
// root-to-v path.  The right children containing leaf lo are exactly the
// nodes of the path (and of v's leftmost descent, all left children) where
// it turns right, and each internal node a has one right child, with range
// [split(a) + 1, hi(a)].  Every leaf position p >= 1 is split(a) + 1 for
// exactly one a, so R(v) = lo - F(lo), F(x) the number of internal nodes
// whose range ends before x: an exclusive scan of the per-leaf counts that
// k_build leaves in tl_of[.].y.  For a leaf, pre(p) = 2 (n-1-p) + p - F(p);
// an internal node's subtree ends with its leftmost leaf, so
// pre(v) = pre(leaf lo) - 2 (hi - lo).

// --------------------------------------------------------- aggregation

// Node aggregate used when combining children: one 80-byte record.
struct Agg {
    double m;
    double mom[3];
    double mn[3];
    double mx[3];
};

// Walk records.  FP64 (reference parity): indexed by node id, with the
// right child and the escape (next node in right-first DFS order after the
// subtree).  FP32: stored in right-first DFS preorder, so opening a node
// moves to the next slot and accepting it skips its subtree: recA = (com,
// accept threshold), recB = (G*mass bits, skip slot); leaves have threshold
// -inf.  The 16-byte half is what every examined node needs; G*mass only an
// accepted one.
struct WalkOut {
    WalkRec64 *rec64;
    float4 *recA;
    int2 *recB;
    const int *pre;
    double G, inv_theta2;
    int theta_zero;
    const unsigned long long *stats;  // tree_aborted() is read from it
    unsigned long long *host_flag;    // pinned host word: set when the compute aborted
};

// pslot: the node's preorder slot (FP32 records; the caller has it at hand).
template <int MODE>
__device__ __forceinline__ void write_node(Agg *agg, const WalkOut &wo,
                                           int node, int n, double m, const double mom[3],
                                           const double mn[3], const double mx[3],
                                           const double c[3], int lo, int hi, int right_child,
                                           int escape, int pslot)
{
    Agg g;
    g.m = m;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        g.mom[a] = mom[a];
        g.mn[a] = mn[a];
        g.mx[a] = mx[a];
    }
    agg[node] = g;
    double size = mx[0] - mn[0];
    size = mx[1] - mn[1] > size ? mx[1] - mn[1] : size;
    size = mx[2] - mn[2] > size ? mx[2] - mn[2] : size;
    const bool leaf = node < n;
    if (MODE == 0) {
        WalkRec64 r;
        r.x = c[0]; r.y = c[1]; r.z = c[2]; r.m = m;
        r.size = leaf ? 0.0 : size;
        r.right = leaf ? -1 : right_child;
        r.escape = escape;
        wo.rec64[node] = r;
    } else {
        const int p = pslot;
        const float crit = leaf ? -INFINITY
                                : (wo.theta_zero ? INFINITY
                                                 : static_cast<float>(size * size * wo.inv_theta2));
        wo.recA[p] = make_float4(static_cast<float>(c[0]), static_cast<float>(c[1]),
                                 static_cast<float>(c[2]), crit);
        wo.recB[p] = make_int2(__float_as_int(static_cast<float>(wo.G * m)),
                               p + 2 * (hi - lo + 1) - 1);
    }
}

constexpr int kRankItems = 16;
constexpr int kRankTile = 256 * kRankItems;  // leaves per treelet-rank tile

// Rank of every treelet in leaf order, and the preorder slot of every leaf:
// two exclusive scans over the leaf positions done together -- of "first
// leaf of a treelet" (tl_of[i].x >= 0) and of the internal-node range-end
// counts (tl_of[i].y) -- one tile of kRankTile leaves per workgroup with
// decoupled look-back (the sort's scheme: one 64-bit word per tile, status +
// the two counts side by side, bounded waits).  trank[first leaf] = rank,
// tsorted[rank] = root, counts[3] = the number of treelets, and (pre set)
// pre[leaf] = its preorder slot.
__device__ void treelet_rank_tile(const int2 *__restrict__ tl_of, int n, int *__restrict__ trank,
                                  int *__restrict__ tsorted, int *__restrict__ pre,
                                  unsigned long long *__restrict__ lb, int *__restrict__ counts,
                                  int ntiles, unsigned long long *__restrict__ stats,
                                  int spin_limit)
{
    using W = unsigned long long;
    constexpr int kHi = 31;  // the range-end count's field
    constexpr W kLo = (1ull << kHi) - 1ull;
    __shared__ int s_tile;
    __shared__ W s_wsum[4];
    __shared__ W s_excl;
    const int t = threadIdx.x;
    const int w = t >> 6;
    const int lane = t & 63;
    if (t == 0) s_tile = atomicAdd(&counts[2], 1);
    __syncthreads();
    const int tile = s_tile;
    const int base = tile * kRankTile + t * kRankItems;
    unsigned bits = 0;
    int root[kRankItems], ends[kRankItems];
    unsigned esum = 0;
#pragma unroll
    for (int r = 0; r < kRankItems; ++r) {
        const int i = base + r;
        const int2 e = i < n ? tl_of[i] : make_int2(-1, 0);
        root[r] = e.x;
        ends[r] = e.y;
        if (e.x >= 0) bits |= 1u << r;
        esum += static_cast<unsigned>(e.y);
    }
    const W mine = (static_cast<W>(esum) << kHi) | __popc(bits);
    W incl = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const W v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) s_wsum[w] = incl;
    __syncthreads();
    W woff = 0, total = 0;
#pragma unroll
    for (int ww = 0; ww < 4; ++ww) {
        if (ww < w) woff += s_wsum[ww];
        total += s_wsum[ww];
    }
    if (t == 0) {
        W *slot = &lb[tile];
        __hip_atomic_store(slot, (tile == 0 ? LbWord<W>::inc : LbWord<W>::agg) | total,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        W excl = 0;
        if (tile > 0) {
            excl = look_back<kLookWin, W>(lb, tile, 1, stats, spin_limit);
            __hip_atomic_store(slot, LbWord<W>::inc | (excl + total), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        s_excl = excl;
        if (tile == ntiles - 1) counts[3] = static_cast<int>((excl + total) & kLo);
    }
    __syncthreads();
    const W start = s_excl + woff + incl - mine;
    unsigned r0 = static_cast<unsigned>(start & kLo);
    int f = static_cast<int>(start >> kHi);  // internal nodes ending before leaf base
#pragma unroll
    for (int r = 0; r < kRankItems; ++r) {
        const int i = base + r;
        if ((bits >> r) & 1u) {
            trank[i] = static_cast<int>(r0);
            tsorted[r0] = root[r];
            ++r0;
        }
        if (pre && i < n) pre[i] = 2 * (n - 1 - i) + i - f;
        f += ends[r];
    }
}

// One workgroup per ranking tile.
__global__ __launch_bounds__(256) void k_rank_preorder(
    const int2 *__restrict__ tl_of, int n, int *__restrict__ trank, int *__restrict__ tsorted,
    int *__restrict__ pre, unsigned long long *__restrict__ lb, int *__restrict__ counts,
    int ntiles, unsigned long long *__restrict__ stats, int spin_limit)
{
    if (tree_aborted(stats)) return;
    treelet_rank_tile(tl_of, n, trank, tsorted, pre, lb, counts, ntiles, stats, spin_limit);
}

// Pass 1: one workgroup per treelet (grid-stride over the list).  Every
// node of the treelet has both children inside it, so the whole treelet is
// summed in LDS.  Thread t owns leaf lo + t and internal node lo + t (the
// treelet's internal ids all lie in its leaf range; the one id there that
// belongs to an ancestor is skipped).  The internal nodes are summed level
// synchronously: in each pass a node whose two children were finished in
// earlier passes is combined, and one barrier ends the pass -- a node
// finished in pass k carries k in s_done, and readers in pass k accept
// only children with a smaller value, so no flag is read before the data
// it guards is visible.  The passes number the treelet's height (at most
// 17 at 120k).  Only the sums (mass, moment, box) are on that chain; the
// COM divisions and the node and walk-record stores follow in parallel,
// one node per thread.  Each node is the same expression of its children
// (left + right) as a bottom-up climb, so the doubles do not depend on the
// schedule.  A leaf keeps only its position and mass in LDS (its moment is
// m * x with the same rounding, its box the point).
template <int MODE>
__global__ __launch_bounds__(kTreelet) void k_agg_treelet(
    const int *__restrict__ order, const double *__restrict__ pos,
    const double *__restrict__ mass, int n, const int *__restrict__ left,
    const int *__restrict__ right, const int *__restrict__ parent,
    const int2 *__restrict__ range, const int *__restrict__ left_at_split,
    const int *__restrict__ tsorted, const int *__restrict__ counts, Agg *agg,
    WalkOut wo)
{
    if (tree_aborted(wo.stats)) return;
    constexpr int C = kTreelet;
    __shared__ double s_lx[3][C];   // leaves: position
    __shared__ double s_lm[C];      //         mass
    __shared__ double s_m[C];       // internal nodes: mass, moment, box
    __shared__ double s_mom[3][C];
    __shared__ double s_mn[3][C];
    __shared__ double s_mx[3][C];
    __shared__ int s_done[C];       // pass that finished internal node lo + t (0: not yet)
    __shared__ int s_pre[MODE != 0 ? C : 1];  // leaves' preorder slots
    const int tid = threadIdx.x;
    const int cnt = counts[3];  // treelets, in leaf order in tsorted
    for (int e = blockIdx.x; e < cnt; e += gridDim.x) {
        const int v = tsorted[e];
        const int lo = v < n ? v : range[v - n].x;
        const int size = v < n ? 1 : range[v - n].y - lo + 1;
        const int hi = lo + size - 1;
        // internal node lo + tid: topology in registers
        bool own = false;
        int L = 0, R = 0, npre = 0;
        int2 rg = make_int2(0, 0);
        const int ii = lo + tid;
        if (tid < size && ii < n - 1) {  // independent loads, issued together
            rg = range[ii];
            L = left[ii];
            R = right[ii];
            own = rg.x >= lo && rg.y <= hi;  // inside this treelet
        }
        s_done[tid] = 0;
        if (tid < size) {
            // barnes_hut.py:129-134 (leaf init).
            const int li = lo + tid;
            const int p = order[li];
            const double x[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
            const double m = mass[p];
            double mom[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                mom[a] = m * x[a];
                s_lx[a][tid] = x[a];
            }
            s_lm[tid] = m;
            const int lpre = MODE != 0 ? wo.pre[li] : 0;
            if (MODE != 0) s_pre[tid] = lpre;
            write_node<MODE>(agg, wo, li, n, m, mom, x, x, x, li, li, -1,
                             li == 0 ? -1 : left_at_split[li - 1], lpre);
        }
        __syncthreads();
        // an internal node's slot from its leftmost leaf's (its subtree's
        // last slot): pre(v) = pre(lo) - 2 (hi - lo)
        if (MODE != 0 && own) npre = s_pre[rg.x - lo] - 2 * (rg.y - rg.x);
        if (v >= n) {  // uniform: the treelet has internal nodes
            const int troot = v - n - lo;
            double mm = 0.0, mom[3] = {0.0, 0.0, 0.0}, mn[3] = {0.0, 0.0, 0.0},
                   mx[3] = {0.0, 0.0, 0.0};
            bool done = !own;
            // at most C passes (a treelet's height is below its leaf count)
            for (int pass = 1; pass <= C; ++pass) {
                if (!done) {
                    const int sL = L < n ? L - lo : L - n - lo;
                    const int sR = R < n ? R - lo : R - n - lo;
                    const int dL = L < n ? 0 : s_done[sL];
                    const int dR = R < n ? 0 : s_done[sR];
                    const bool okL = L < n || (dL > 0 && dL < pass);
                    const bool okR = R < n || (dR > 0 && dR < pass);
                    if (okL && okR) {
                        double mL, mR, momL[3], momR[3], mnL[3], mnR[3], mxL[3], mxR[3];
                        if (L < n) {
                            mL = s_lm[sL];
#pragma unroll
                            for (int a = 0; a < 3; ++a) {
                                momL[a] = mL * s_lx[a][sL];
                                mnL[a] = mxL[a] = s_lx[a][sL];
                            }
                        } else {
                            mL = s_m[sL];
#pragma unroll
                            for (int a = 0; a < 3; ++a) {
                                momL[a] = s_mom[a][sL];
                                mnL[a] = s_mn[a][sL];
                                mxL[a] = s_mx[a][sL];
                            }
                        }
                        if (R < n) {
                            mR = s_lm[sR];
#pragma unroll
                            for (int a = 0; a < 3; ++a) {
                                momR[a] = mR * s_lx[a][sR];
                                mnR[a] = mxR[a] = s_lx[a][sR];
                            }
                        } else {
                            mR = s_m[sR];
#pragma unroll
                            for (int a = 0; a < 3; ++a) {
                                momR[a] = s_mom[a][sR];
                                mnR[a] = s_mn[a][sR];
                                mxR[a] = s_mx[a][sR];
                            }
                        }
                        mm = mL + mR;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            mom[a] = momL[a] + momR[a];
                            mn[a] = fmin(mnL[a], mnR[a]);
                            mx[a] = fmax(mxL[a], mxR[a]);
                        }
                        s_m[tid] = mm;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            s_mom[a][tid] = mom[a];
                            s_mn[a][tid] = mn[a];
                            s_mx[a][tid] = mx[a];
                        }
                        s_done[tid] = pass;
                        done = true;
                    }
                }
                __syncthreads();
                // The root finished by this pass: every thread reads the same
                // value here, since a root finished in the next pass would
                // carry pass + 1.
                const int dr = s_done[troot];
                if (dr != 0 && dr <= pass) break;
            }
            if (own) {
                double c[3];
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    // barnes_hut.py:175-177: com = mom / mass where mass > 0.
                    c[a] = mm > 0.0 ? mom[a] / mm : 0.0;
                write_node<MODE>(agg, wo, ii + n, n, mm, mom, mn, mx, c, rg.x, rg.y, R,
                                 rg.x == 0 ? -1 : left_at_split[rg.x - 1], npre);
            }
        }
        __syncthreads();  // the LDS slots are reused by the next treelet
    }
}

// The nodes above the treelets: each is the sum of a run of consecutive
// treelets in leaf order, [trank[lo], trank[hi + 1]), so every upper node is
// summed on its own, with no ordering between nodes -- one workgroup per
// node.  They are the T - 1 nodes whose split falls between two consecutive
// treelets: the node split at leaf b - 1, b the first leaf of treelet j + 1,
// is parent[left_at_split[b - 1]].  Thread t adds treelets A + t, A + t + 256, ... in order; the 256
// partials are combined by a butterfly within each wave (both operands of
// every add are the same on the two lanes, so all lanes agree) and the four
// wave results in wave order: a fixed summation order.  Min/max are exact.
template <int MODE>
__global__ __launch_bounds__(256) void k_agg_upper(
    int n, const int *__restrict__ right, const int *__restrict__ parent,
    const int2 *__restrict__ range, const int *__restrict__ left_at_split,
    const int *__restrict__ counts, const int *__restrict__ trank,
    const int *__restrict__ tsorted, Agg *agg, WalkOut wo)
{
    if (tree_aborted(wo.stats)) {
        // Every kernel that can give up has finished: tell the host (the
        // next call on the handle reports it without synchronising).
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            __hip_atomic_store(wo.host_flag, wo.stats[3], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        }
        return;
    }
    __shared__ double s_r[4][10];
    const int t = threadIdx.x;
    const int w = t >> 6;
    const int lane = t & 63;
    const int T = counts[3];
    for (int e = blockIdx.x; e < T - 1; e += gridDim.x) {
        const int tr = tsorted[e + 1];
        const int b = tr < n ? tr : range[tr - n].x;  // first leaf of treelet e + 1
        const int v = parent[left_at_split[b - 1]];
        const int ii = v - n;
        const int2 rg = range[ii];
        const int A = trank[rg.x];
        const int B = rg.y == n - 1 ? T : trank[rg.y + 1];
        double r[10] = {0.0, 0.0, 0.0, 0.0, INFINITY, INFINITY, INFINITY,
                        -INFINITY, -INFINITY, -INFINITY};
        for (int j = A + t; j < B; j += 256) {
            const Agg a = agg[tsorted[j]];
            r[0] += a.m;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                r[1 + c] += a.mom[c];
                r[4 + c] = fmin(r[4 + c], a.mn[c]);
                r[7 + c] = fmax(r[7 + c], a.mx[c]);
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int c = 0; c < 4; ++c) r[c] += __shfl_xor(r[c], off, 64);
#pragma unroll
            for (int c = 4; c < 7; ++c) r[c] = fmin(r[c], __shfl_xor(r[c], off, 64));
#pragma unroll
            for (int c = 7; c < 10; ++c) r[c] = fmax(r[c], __shfl_xor(r[c], off, 64));
        }
        if (lane == 0)
#pragma unroll
            for (int c = 0; c < 10; ++c) s_r[w][c] = r[c];
        __syncthreads();
        if (t == 0) {
            double q[10];
#pragma unroll
            for (int c = 0; c < 10; ++c) q[c] = s_r[0][c];
            for (int ww = 1; ww < 4; ++ww) {
#pragma unroll
                for (int c = 0; c < 4; ++c) q[c] += s_r[ww][c];
#pragma unroll
                for (int c = 4; c < 7; ++c) q[c] = fmin(q[c], s_r[ww][c]);
#pragma unroll
                for (int c = 7; c < 10; ++c) q[c] = fmax(q[c], s_r[ww][c]);
            }
            const double mm = q[0];
            double mom[3], mn[3], mx[3], cm[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                mom[c] = q[1 + c];
                mn[c] = q[4 + c];
                mx[c] = q[7 + c];
                // barnes_hut.py:175-177: com = mom / mass where mass > 0.
                cm[c] = mm > 0.0 ? mom[c] / mm : 0.0;
            }
            write_node<MODE>(agg, wo, v, n, mm, mom, mn, mx, cm, rg.x, rg.y, right[ii],
                             rg.x == 0 ? -1 : left_at_split[rg.x - 1],
                             MODE == 0 ? 0 : wo.pre[rg.x] - 2 * (rg.y - rg.x));
        }
        __syncthreads();  // s_r is reused by the next node
    }
}

// Reference-layout node arrays (barnes_hut.py:232-236) from the aggregates,
// on demand: com is the leaf position for leaves (exactly its AABB) and
// mom / mass (0 when massless) for internal nodes -- the expression the
// walk records use.
__global__ void k_export_nodes(const Agg *__restrict__ agg, int n, double *__restrict__ com,
                               double *__restrict__ nmass, double *__restrict__ nmin,
                               double *__restrict__ nmax)
{
    const int m = 2 * n - 1;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m; v += gridDim.x * blockDim.x) {
        const Agg g = agg[v];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (com) com[3 * v + a] = v < n ? g.mn[a] : (g.m > 0.0 ? g.mom[a] / g.m : 0.0);
            if (nmin) nmin[3 * v + a] = g.mn[a];
            if (nmax) nmax[3 * v + a] = g.mx[a];
        }
        if (nmass) nmass[v] = g.m;
    }
}

// ---------------------------------------------------------------- walk
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    // Blocks are dispatched round-robin over the 8 XCDs.  Give each XCD a
    // contiguous run of Morton-ordered particles so the nodes its blocks
    // share stay in that XCD's L2.
    constexpr int kXcd = 8;
    const int q = nb / kXcd, r = nb % kXcd;
    const int x = b % kXcd, k = b / kXcd;
    return x * q + min(x, r) + k;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int wave_max_i32(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Where a particle's walk result goes: acc = a, or acc = acc + a when the
// caller folds the tree field into an existing acceleration (the
// "acc += acc_g" of living_galaxy_engine.py:116-120 and gas_disk_engine.py:
// 84-89); then, when vel is set, the closing half-kick of the KDK step,
// vel += acc * dt_half (leapfrog.py:14, engine.py:113-114).
struct AccOut {
    double *acc;
    double *vel;
    double dt_half;
    int add;
};

__device__ __forceinline__ void store_acc(const AccOut &o, int p, double ax, double ay, double az)
{
    double *__restrict__ acc = o.acc;
    if (o.add) {
        ax = acc[3 * p] + ax;
        ay = acc[3 * p + 1] + ay;
        az = acc[3 * p + 2] + az;
    }
    acc[3 * p] = ax;
    acc[3 * p + 1] = ay;
    acc[3 * p + 2] = az;
    if (o.vel) {
        double *__restrict__ v = o.vel;
        v[3 * p] = v[3 * p] + ax * o.dt_half;
        v[3 * p + 1] = v[3 * p + 1] + ay * o.dt_half;
        v[3 * p + 2] = v[3 * p + 2] + az * o.dt_half;
    }
}

// FP64: the reference's criterion and arithmetic (barnes_hut.py:196-216),
// nodes in the reference's pop order via right child / escape links.
__global__ void k_walk64(const WalkRec64 *__restrict__ rec, const int *__restrict__ order,
                         AccOut out, int n, double G, double theta2, double eps2,
                         unsigned long long *__restrict__ stats)
{
    if (tree_aborted(stats)) return;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int t = b * blockDim.x + threadIdx.x;
    const bool active = t < n;
    int visits = 0;
    if (active) {
        const WalkRec64 self = rec[t];
        const double px = self.x, py = self.y, pz = self.z;
        double ax = 0.0, ay = 0.0, az = 0.0;
        int node = n;  // root
        const int max_it = 2 * n;
        const int m = 2 * n - 1;
        while (node >= 0 && node < m && visits < max_it) {
            ++visits;
            const WalkRec64 r = rec[node];
            const double dx = r.x - px;
            const double dy = r.y - py;
            const double dz = r.z - pz;
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            // Leaves always interact (the self leaf adds an exact zero: d == 0).
            if (node < n || r.size * r.size < theta2 * r2) {
                const double inv = 1.0 / sqrt(r2);
                const double s = G * r.m * inv * inv * inv;
                ax += s * dx;
                ay += s * dy;
                az += s * dz;
                node = r.escape;
            } else {
                node = r.right;
            }
        }
        if (node >= 0) atomicAdd(&stats[4], 1ull);  // walk did not terminate in bounds
        const int p = order[t];
        store_acc(out, p, ax, ay, az);
    }
    const unsigned long long tot = wave_sum_u64(static_cast<unsigned long long>(visits));
    const int mx = wave_max_i32(visits);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], tot);
        atomicMax(&stats[1], static_cast<unsigned long long>(mx));
    }
}

// FP32: a forward scan of the preorder records with subtree skips.  Slot
// k+1 is the right child of an opened internal slot k; skip is the first
// slot after the subtree.  Every step moves forward, so a walk ends after at
// most 2n-1 steps.
__global__ void k_walk32(const float4 *__restrict__ recA, const int2 *__restrict__ recB,
                         const int *__restrict__ pre, const int *__restrict__ order,
                         AccOut out, int n, float eps2,
                         unsigned long long *__restrict__ stats)
{
    if (tree_aborted(stats)) return;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int t = b * blockDim.x + threadIdx.x;
    const bool active = t < n;
    int visits = 0;
    if (active) {
        const float4 self = recA[pre[t]];
        const float px = self.x, py = self.y, pz = self.z;
        float ax = 0.f, ay = 0.f, az = 0.f;
        const int m = 2 * n - 1;
        int k = 0;  // the root's slot
        while (k < m) {
            ++visits;
            const float4 c = recA[k];
            const int2 sb = recB[k];
            const float dx = c.x - px;
            const float dy = c.y - py;
            const float dz = c.z - pz;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
            if (c.w < r2) {
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float s = __int_as_float(sb.x) * (inv * inv) * inv;
                ax = fmaf(s, dx, ax);
                ay = fmaf(s, dy, ay);
                az = fmaf(s, dz, az);
                k = max(sb.y, k + 1);
            } else {
                k = k + 1;
            }
        }
        const int p = order[t];
        store_acc(out, p, static_cast<double>(ax), static_cast<double>(ay),
                  static_cast<double>(az));
    }
    const unsigned long long tot = wave_sum_u64(static_cast<unsigned long long>(visits));
    const int mx = wave_max_i32(visits);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], tot);
        atomicMax(&stats[1], static_cast<unsigned long long>(mx));
    }
}

// FP32, two lanes per particle: the particle's own leaf slot ps splits the
// preorder records into [0, ps) (lane 0) and [ps, 2n-1) (lane 1).  A lane
// interacts only with the nodes it owns; a node it does not own but whose
// subtree crosses into its range (an ancestor of the particle) is still
// tested, and accepting it means the owning lane adds its monopole, so the
// subtree is skipped.  The two partial sums are combined with a lane
// shuffle.  Every node of the single-lane walk is examined by exactly one
// owner, so the accepted set and visit counts are unchanged.
__global__ void k_walk32_pair(const float4 *__restrict__ recA, const int2 *__restrict__ recB,
                              const int *__restrict__ pre, const int *__restrict__ order,
                              AccOut out, int n, float eps2,
                              unsigned long long *__restrict__ stats)
{
    if (tree_aborted(stats)) return;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int g = b * blockDim.x + threadIdx.x;
    const int t = g >> 1;
    const int half = g & 1;
    const bool active = t < n;
    int visits = 0;
    float ax = 0.f, ay = 0.f, az = 0.f;
    if (active) {
        const int ps = pre[t];
        const float4 self = recA[ps];
        const float px = self.x, py = self.y, pz = self.z;
        const int m = 2 * n - 1;
        const int lo = half ? ps : 0;
        const int hi = half ? m : ps;
        int k = 0;
        while (k < hi) {
            const int2 sb = recB[k];
            const int skip = max(sb.y, k + 1);
            if (skip <= lo) {  // subtree entirely before this lane's range
                k = skip;
                continue;
            }
            const bool own = k >= lo;
            visits += own;
            const float4 c = recA[k];
            const float dx = c.x - px;
            const float dy = c.y - py;
            const float dz = c.z - pz;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
            if (c.w < r2) {
                if (own) {
                    const float inv = __builtin_amdgcn_rsqf(r2);
                    const float s = __int_as_float(sb.x) * (inv * inv) * inv;
                    ax = fmaf(s, dx, ax);
                    ay = fmaf(s, dy, ay);
                    az = fmaf(s, dz, az);
                }
                k = skip;
            } else {
                k = k + 1;
            }
        }
    }
    ax += __shfl_xor(ax, 1, 64);
    ay += __shfl_xor(ay, 1, 64);
    az += __shfl_xor(az, 1, 64);
    if (active && half == 0) {
        const int p = order[t];
        store_acc(out, p, static_cast<double>(ax), static_cast<double>(ay),
                  static_cast<double>(az));
    }
    const unsigned long long tot = wave_sum_u64(static_cast<unsigned long long>(visits));
    // per-particle visits = the pair's sum
    const int pv = visits + __shfl_xor(visits, 1, 64);
    const int mx = wave_max_i32(pv);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], tot);
        atomicMax(&stats[1], static_cast<unsigned long long>(mx));
    }
}

// FP32, one wave walks together: the wave steps through the preorder
// records in increasing slot order, visiting every slot that at least one
// of its 64 (Morton-adjacent) particles wants next; a lane whose next wanted
// slot is the current one tests it, the others wait.  The per-lane accepted
// set, visit count and summation order are exactly the single-lane walk's.
// Because the slots only increase, the wave keeps a window of 64
// consecutive records in registers (lane l holds slot k0 + l), refilled by
// one coalesced load when the walk leaves it; the current record is read
// with v_readlane into scalar registers, so most steps touch no memory.
__global__ __launch_bounds__(256) void k_walk32_wave(const float4 *__restrict__ recA,
                                                     const int2 *__restrict__ recB,
                                                     const int *__restrict__ pre,
                                                     const int *__restrict__ order,
                                                     AccOut out, int n, float eps2,
                                                     unsigned long long *__restrict__ stats)
{
    if (tree_aborted(stats)) return;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int t = b * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool active = t < n;
    const int m = 2 * n - 1;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (active) {
        const float4 self = recA[pre[t]];
        px = self.x;
        py = self.y;
        pz = self.z;
    }
    float ax = 0.f, ay = 0.f, az = 0.f;
    int visits = 0;
    int next = active ? 0 : m;
    int k = __builtin_amdgcn_readfirstlane(__ockl_wfred_min_i32(next));
    int k0 = -64;  // window start (no window yet)
    float4 wa = make_float4(0.f, 0.f, 0.f, 0.f);
    int2 wb = make_int2(0, 0);
    while (k < m) {
        if (k - k0 >= 64) {
            k0 = k;
            const int slot = k0 + lane;
            if (slot < m) {
                wa = recA[slot];
                wb = recB[slot];
            }
        }
        const int l = k - k0;
        const float cx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.x), l));
        const float cy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.y), l));
        const float cz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.z), l));
        const float cw = __int_as_float(__builtin_amdgcn_readlane(wb.x, l));
        const float crit = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.w), l));
        const int skip = __builtin_amdgcn_readlane(wb.y, l);
        if (next == k) {
            ++visits;
            const float dx = cx - px;
            const float dy = cy - py;
            const float dz = cz - pz;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
            if (crit < r2) {
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float s = cw * (inv * inv) * inv;
                ax = fmaf(s, dx, ax);
                ay = fmaf(s, dy, ay);
                az = fmaf(s, dz, az);
                next = max(skip, k + 1);
            } else {
                next = k + 1;
            }
        }
        // Every lane not at k waits at the skip slot of an ancestor of k,
        // which is >= skip(k); so the smallest next slot is k + 1 if a lane
        // opened node k, else skip(k), with no reduction over the lanes.
        k = __ballot(next == k + 1) ? k + 1 : max(skip, k + 1);
    }
    if (active) {
        const int p = order[t];
        store_acc(out, p, static_cast<double>(ax), static_cast<double>(ay),
                  static_cast<double>(az));
    }
    const unsigned long long tot = wave_sum_u64(static_cast<unsigned long long>(visits));
    const int mx = wave_max_i32(visits);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], tot);
        atomicMax(&stats[1], static_cast<unsigned long long>(mx));
    }
}

// The wave walk advancing two slots per iteration: slots k and k+1 (usually
// the next two the wave needs, since opening a node leads to k+1) are
// evaluated together -- their arithmetic is independent, so it overlaps --
// and applied in slot order, keeping each lane's participation and
// summation order exactly those of the one-slot walk.
__global__ __launch_bounds__(256) void k_walk32_wave2(const float4 *__restrict__ recA,
                                                      const int2 *__restrict__ recB,
                                                      const int *__restrict__ pre,
                                                      const int *__restrict__ order,
                                                      AccOut out, int n,
                                                      float eps2,
                                                      unsigned long long *__restrict__ stats)
{
    if (tree_aborted(stats)) return;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int t = b * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool active = t < n;
    const int m = 2 * n - 1;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (active) {
        const float4 self = recA[pre[t]];
        px = self.x;
        py = self.y;
        pz = self.z;
    }
    float ax = 0.f, ay = 0.f, az = 0.f;
    int visits = 0;
    int next = active ? 0 : m;
    int k = __builtin_amdgcn_readfirstlane(__ockl_wfred_min_i32(next));
    int k0 = -64;
    float4 wa = make_float4(0.f, 0.f, 0.f, 0.f);
    int2 wb = make_int2(0, 0);
    while (k < m) {
        if (k + 1 - k0 >= 64) {  // the window must hold slots k and k + 1
            k0 = k;
            const int slot = k0 + lane;
            if (slot < m) {
                wa = recA[slot];
                wb = recB[slot];
            }
        }
        const int l = k - k0;
        const bool has1 = k + 1 < m;  // uniform
        float c[2][4], crit[2];
        int skip[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            c[q][0] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.x), l + q));
            c[q][1] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.y), l + q));
            c[q][2] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.z), l + q));
            c[q][3] = __int_as_float(__builtin_amdgcn_readlane(wb.x, l + q));
            crit[q] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.w), l + q));
            skip[q] = __builtin_amdgcn_readlane(wb.y, l + q);
        }
        float dx[2], dy[2], dz[2], r2[2], s[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            dx[q] = c[q][0] - px;
            dy[q] = c[q][1] - py;
            dz[q] = c[q][2] - pz;
            r2[q] = fmaf(dz[q], dz[q], fmaf(dy[q], dy[q], fmaf(dx[q], dx[q], eps2)));
            const float inv = __builtin_amdgcn_rsqf(r2[q]);
            s[q] = c[q][3] * (inv * inv) * inv;
        }
        if (next == k) {
            ++visits;
            if (crit[0] < r2[0]) {
                ax = fmaf(s[0], dx[0], ax);
                ay = fmaf(s[0], dy[0], ay);
                az = fmaf(s[0], dz[0], az);
                next = max(skip[0], k + 1);
            } else {
                next = k + 1;
            }
        }
        if (has1 && next == k + 1) {
            ++visits;
            if (crit[1] < r2[1]) {
                ax = fmaf(s[1], dx[1], ax);
                ay = fmaf(s[1], dy[1], ay);
                az = fmaf(s[1], dz[1], az);
                next = max(skip[1], k + 2);
            } else {
                next = k + 2;
            }
        }
        // every lane now wants a slot >= kk; kk itself whenever a lane
        // opened (or stepped onto) slot k + 1
        const int kk = has1 ? k + 2 : k + 1;
        if (__ballot(next == kk))
            k = kk;
        else
            k = __builtin_amdgcn_readfirstlane(__ockl_wfred_min_i32(next));
    }
    if (active) {
        const int p = order[t];
        store_acc(out, p, static_cast<double>(ax), static_cast<double>(ay),
                  static_cast<double>(az));
    }
    const unsigned long long tot = wave_sum_u64(static_cast<unsigned long long>(visits));
    const int mx = wave_max_i32(visits);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], tot);
        atomicMax(&stats[1], static_cast<unsigned long long>(mx));
    }
}

// The wave walk with each group of 64 particles split over SPLIT waves of
// one workgroup (SPLIT = 2 or 4).  The boundaries are the preorder slots of
// the group's own leaves at lanes 64 (SPLIT - j) / SPLIT, j = 1 .. SPLIT-1:
// a right-first preorder visits leaves right to left, so these slots
// increase with j.  Wave j owns the slots [b_j, b_j+1) (b_0 = 0, b_SPLIT =
// 2n-1).  A node is tested, counted and (if accepted) added only by its
// owning wave.  A wave also tests the ancestors of its first slot (slot <
// b_j, subtree reaching past b_j): accepting one means an earlier wave adds
// its monopole, so this wave skips the subtree; subtrees ending at or
// before b_j are stepped over untested, since the walk resumes after a
// subtree at its skip slot whether the subtree was accepted or opened.  The
// visited set per particle is the one-wave walk's; the SPLIT FP32 partial
// sums are added through LDS in wave order.
template <int SPLIT>
__global__ __launch_bounds__(256) void k_walk32_wave_split(
    const float4 *__restrict__ recA, const int2 *__restrict__ recB, const int *__restrict__ pre,
    const int *__restrict__ order, AccOut out, int n, float eps2,
    unsigned long long *__restrict__ stats)
{
    if (tree_aborted(stats)) return;
    constexpr int G = 4 / SPLIT;  // groups of 64 particles per 256-thread block
    __shared__ float s_part[G][SPLIT - 1][64][3];  // waves 1.. partial sums
    __shared__ int s_vis[G][SPLIT - 1][64];
    const int wave = threadIdx.x >> 6;  // 0..3
    const int grp = wave / SPLIT;       // group within the block
    const int part = wave % SPLIT;
    const int lane = threadIdx.x & 63;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int t = (b * G + grp) * 64 + lane;
    const bool active = t < n;
    const int m = 2 * n - 1;
    int ps = 0;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (active) {
        ps = pre[t];
        const float4 self = recA[ps];
        px = self.x;
        py = self.y;
        pz = self.z;
    }
    // the boundary slots (the last active lane stands in for inactive ones)
    const int nact = max(0, min(64, n - (t - lane)));
    int lo = 0, hi = m;
    if (nact > 0) {
        if (part > 0)
            lo = __builtin_amdgcn_readlane(ps, min(64 - 64 * part / SPLIT, nact - 1));
        if (part < SPLIT - 1)
            hi = __builtin_amdgcn_readlane(ps, min(64 - 64 * (part + 1) / SPLIT, nact - 1));
    } else {
        lo = hi = m;
    }
    float ax = 0.f, ay = 0.f, az = 0.f;
    int visits = 0;
    int next = (active && lo < hi) ? 0 : m;
    int k = __builtin_amdgcn_readfirstlane(__ockl_wfred_min_i32(next));
    int k0 = -64;
    float4 wa = make_float4(0.f, 0.f, 0.f, 0.f);
    int2 wb = make_int2(0, 0);
    while (k < hi) {
        if (k - k0 >= 64) {
            k0 = k;
            const int slot = k0 + lane;
            if (slot < m) {
                wa = recA[slot];
                wb = recB[slot];
            }
        }
        const int l = k - k0;
        const float cx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.x), l));
        const float cy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.y), l));
        const float cz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.z), l));
        const float cw = __int_as_float(__builtin_amdgcn_readlane(wb.x, l));
        const float crit = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.w), l));
        const int skip = max(__builtin_amdgcn_readlane(wb.y, l), k + 1);
        if (next == k) {
            if (skip <= lo) {
                next = skip;  // subtree entirely before this wave's range (uniform per node)
            } else {
                const bool own = k >= lo;
                const float dx = cx - px;
                const float dy = cy - py;
                const float dz = cz - pz;
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
                if (crit < r2) {
                    if (own) {
                        ++visits;
                        const float inv = __builtin_amdgcn_rsqf(r2);
                        const float s = cw * (inv * inv) * inv;
                        ax = fmaf(s, dx, ax);
                        ay = fmaf(s, dy, ay);
                        az = fmaf(s, dz, az);
                    }
                    next = skip;
                } else {
                    visits += own ? 1 : 0;
                    next = k + 1;
                }
            }
        }
        // the smallest next slot: k + 1 if a lane opened k, else skip(k)
        // (every other lane waits at an ancestor's skip, >= skip(k))
        k = __ballot(next == k + 1) ? k + 1 : skip;
    }
    if (part > 0) {
        s_part[grp][part - 1][lane][0] = ax;
        s_part[grp][part - 1][lane][1] = ay;
        s_part[grp][part - 1][lane][2] = az;
        s_vis[grp][part - 1][lane] = visits;
    }
    __syncthreads();
    int pv = 0;
    if (part == 0) {
        pv = visits;
#pragma unroll
        for (int j = 0; j < SPLIT - 1; ++j) {
            ax += s_part[grp][j][lane][0];
            ay += s_part[grp][j][lane][1];
            az += s_part[grp][j][lane][2];
            pv += s_vis[grp][j][lane];
        }
        if (active) {
            const int p = order[t];
            store_acc(out, p, static_cast<double>(ax), static_cast<double>(ay),
                      static_cast<double>(az));
        }
    }
    const unsigned long long tot = wave_sum_u64(static_cast<unsigned long long>(visits));
    const int mx = wave_max_i32(pv);
    if (lane == 0) {
        atomicAdd(&stats[0], tot);
        if (part == 0) atomicMax(&stats[1], static_cast<unsigned long long>(mx));
    }
}

// Block b -> group, for launches of one block per group: the blocks go
// round-robin over the 8 XCDs, and each XCD gets runs of R consecutive
// (Morton-adjacent) groups, every 8th run, so the dense parts of a
// distribution spread over all XCDs while neighbouring groups still share
// an L2.  Blocks past the last full set of 8 runs map to themselves.
__device__ __forceinline__ int xcd_run_block(int b, int nb, int R)
{
    const int span = 8 * R;
    const int full = (nb / span) * span;
    if (b >= full) return b;
    const int x = b % 8, k = b / 8;
    return ((k / R) * 8 + x) * R + k % R;
}

constexpr int kChunkSide = 14;                     // chunk boundaries per side, at most
constexpr int kChunkMaxB = 2 * kChunkSide + 2;    // + 1 inside the group + 1 end

// Chunks of group g in k_walk32_chunked (its boundary count + 1).
inline int group_chunks(long long g, long long n, long long dmin, int shift)
{
    const long long t0 = 64 * g;
    const long long nact = std::min<long long>(64, n - t0);
    const long long last = t0 + nact - 1;
    int nr = 0, nl = 0;
    for (long long d = dmin; last + d <= n - 1 && nr < kChunkSide; d <<= shift) ++nr;
    for (long long d = dmin; t0 - d >= 0 && nl < kChunkSide; d <<= shift) ++nl;
    return nr + nl + (nact > 33 ? 1 : 0) + 1;
}

// The most chunks any group has: the side counts change only where a
// boundary distance d meets either end, so the maximum is taken at a group
// next to one of those places (or at either end).
inline int max_group_chunks(long long n, long long dmin, int shift)
{
    const long long ng = (n + 63) / 64;
    int best = 0;
    auto at = [&](long long g) {
        if (g >= 0 && g < ng) best = std::max(best, group_chunks(g, n, dmin, shift));
    };
    at(0);
    at(ng - 1);
    int i = 0;
    for (long long d = dmin; d <= n && i < kChunkSide; d <<= shift, ++i) {
        for (long long q = -1; q <= 1; ++q) {
            at(d / 64 + q);              // where t0 - d >= 0 starts
            at((n - 1 - d) / 64 + q);    // where last + d <= n - 1 ends
        }
    }
    return best;
}

// The wave walk of one group of 64 particles, cut into chunks of preorder
// slots that the block's waves take in turn.  Most of a group's walk is in
// its far field: the subtrees of leaves 2^10 .. 2^17 positions away in
// Morton order.  The boundaries are the slots of the leaves 64 * 4^i
// positions beyond each end of the group (on both sides) and of its own
// leaf at lane 32; a right-first preorder visits leaves right to left, so
// these slots increase from the far right through the group to the far
// left.  A chunk [lo, hi) is walked like one part of the split walk
// (ancestors of lo tested, earlier subtrees stepped over).  Waves take the
// chunks farthest-first, alternating ends, from an LDS counter; each
// chunk's FP32 partial sums go to LDS and are added in slot order at the
// end, so the result does not depend on which wave walked which chunk.
struct ChunkShape {
    int dmin;   // distance (in leaves) of the nearest boundary beyond each end
    int shift;  // boundary distances grow by 2^shift
    int run;    // XCD run length (groups)
};

// The union walk's next slot: l + 1 if an active lane opened slot l, else
// skip(l).  No other lane can need l + 1: a lane waiting at l + 1 waits at the
// skip of an ancestor of l whose subtree ends at l, so l is a leaf, and then
// skip(l) = l + 1 as well.  (s_andn2 sets SCC when its result is nonzero: no
// compare.)
__device__ __forceinline__ int walk_next_slot(unsigned long long mact, unsigned long long mlt,
                                              int l1, int skip)
{
    int kn;
    unsigned long long t_;
    asm volatile("s_andn2_b64 %1, %2, %3\n\ts_cselect_b32 %0, %4, %5"
                 : "=s"(kn), "=&s"(t_)
                 : "s"(mact), "s"(mlt), "s"(l1), "s"(skip)
                 : "scc");
    return kn;
}

// PF (with LW): the next record's LDS read is issued before the force, so its
// latency overlaps the force, for two more scalar instructions a step (the
// slot clamp into the window, and a copy of the slot the force still reads).
template <int NW, bool LW = false, bool PF = false>
__global__ __launch_bounds__(64 * NW) void k_walk32_chunked(
    const float4 *__restrict__ recA, const int2 *__restrict__ recB, const int *__restrict__ pre,
    const int *__restrict__ order, AccOut out, int n, float eps2,
    unsigned long long *__restrict__ stats, ChunkShape cs)
{
    if (tree_aborted(stats)) return;
    using v4f = float __attribute__((ext_vector_type(4)));
    using lds_v4f = __attribute__((address_space(3))) v4f;
    __shared__ int s_b[kChunkMaxB + 1];
    __shared__ int s_nc;
    __shared__ int s_ctr;
    __shared__ int s_done;
    // chunk partial sums [chunk][axis][lane], sized by the launch for the
    // most chunks a group can have at this n
    extern __shared__ float s_part_dyn[];
    float(*s_part)[3][64] = reinterpret_cast<float(*)[3][64]>(s_part_dyn);
    __shared__ int s_vis[64];
    // LW: each wave's window of records sits in LDS as (com, accept
    // threshold), 16 bytes a slot, and a step reads its record with one
    // broadcast load instead of four v_readlane.  G m and the skip slot stay
    // in registers (v_readlane): G m is needed only on accept, so an accept
    // costs no second LDS round trip, and the skip feeds the scalar choice of
    // the next slot.
    __shared__ v4f s_rec[LW ? NW : 1][64];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int g = xcd_run_block(blockIdx.x, gridDim.x, cs.run);
    const int t0 = g * 64;
    const int t = t0 + lane;
    const bool active = t < n;
    const int m = 2 * n - 1;
    const int nact = min(64, n - t0);
    const int last = t0 + nact - 1;
    // boundary leaves: right side at distances dmin * 2^(shift i) (farthest
    // first), then lane 32, then the left side (nearest first); at most
    // kChunkSide per side, the nearest ones (the farthest chunk takes the rest)
    int nr = 0, nl = 0;
    for (long long d = cs.dmin; last + d <= n - 1 && nr < kChunkSide; d <<= cs.shift) ++nr;
    for (long long d = cs.dmin; t0 - d >= 0 && nl < kChunkSide; d <<= cs.shift) ++nl;
    const int ni = nact > 33 ? 1 : 0;
    const int nbd = nr + ni + nl;  // interior boundaries (<= 2 kChunkSide + 1)
    if (wave == 0) {
        if (lane < nbd) {
            long long leaf;
            if (lane < nr)
                leaf = last + (static_cast<long long>(cs.dmin) << (cs.shift * (nr - 1 - lane)));
            else if (lane < nr + ni)
                leaf = t0 + 32;
            else
                leaf = t0 - (static_cast<long long>(cs.dmin) << (cs.shift * (lane - nr - ni)));
            s_b[lane + 1] = pre[leaf];
        }
        if (lane == 0) {
            s_b[0] = 0;
            s_b[nbd + 1] = m;
            s_nc = nbd + 1;
            s_ctr = 0;
            s_done = 0;
        }
        s_vis[lane] = 0;
    }
    float px = 0.f, py = 0.f, pz = 0.f;
    if (active) {
        const float4 self = recA[pre[t]];
        px = self.x;
        py = self.y;
        pz = self.z;
    }
    __syncthreads();
    const int nc = __builtin_amdgcn_readfirstlane(s_nc);
    // The window [k0, k0 + 64): lane l holds record k0 + l in wa / wb (and,
    // LW, its com and threshold in LDS; rb is the LDS byte address of slot 0
    // in the window's frame).  It carries over from chunk to chunk; a chunk's
    // walk starts again at the root, below k0, so every window check is
    // unsigned.
    const unsigned rec0 =
        LW ? static_cast<unsigned>((size_t)(const lds_v4f *)&s_rec[LW ? wave : 0][0]) : 0u;
    unsigned rb = 0;
    int k0 = -64;
    float4 wa = make_float4(0.f, 0.f, 0.f, 0.f);
    int2 wb = make_int2(0, 0);
    const int lm = m - 1;
    // Loads the window at slot KK (slots past the last record read the last
    // one, never used).
#define RG_WALK7_REFILL(KK)                                                   \
    do {                                                                      \
        k0 = (KK);                                                            \
        const int sl_ = min(k0 + lane, lm);                                   \
        wa = recA[sl_];                                                       \
        wb = recB[sl_];                                                       \
        /* the skip slot, relative to k0 and past its own slot (in a valid \
           tree it always is; this keeps every walk finite whatever the      \
           records hold) */                                                  \
        wb.y = max(wb.y - k0, lane + 1);                                      \
        if (LW) {                                                             \
            v4f w_;                                                           \
            w_.x = wa.x;                                                      \
            w_.y = wa.y;                                                      \
            w_.z = wa.z;                                                      \
            w_.w = wa.w;  /* the record's 16-byte half as loaded */          \
            s_rec[LW ? wave : 0][lane] = w_;                                  \
            rb = rec0 - 16u * static_cast<unsigned>(k0);                      \
            /* the window is read across lanes: reads stay after writes */    \
            __builtin_amdgcn_wave_barrier();                                  \
        }                                                                     \
    } while (0)
#define RG_WALK7_READ(KK) (*(const lds_v4f *)(size_t)(rb + 16u * static_cast<unsigned>(KK)))
    while (true) {
        int j = 0;
        if (lane == 0) j = atomicAdd(&s_ctr, 1);
        j = __builtin_amdgcn_readfirstlane(__shfl(j, 0, 64));
        if (j >= nc) break;
        const int c = (j & 1) ? nc - 1 - (j >> 1) : (j >> 1);
        const int lo = __builtin_amdgcn_readfirstlane(s_b[c]);
        const int hi = __builtin_amdgcn_readfirstlane(s_b[c + 1]);
        float ax = 0.f, ay = 0.f, az = 0.f;
        int accepts = 0;
        // A lane takes part at slot k when next <= k.  Accepting a node sets
        // next to its skip slot; opening it leaves next alone (it stays <=
        // k + 1), which saves a register write per step.
        int next = (active && lo < hi) ? 0 : m;
        // the smallest next: 0 when any lane walks this chunk, else m
        int k = __ballot(next == 0) ? 0 : m;
        // Ancestors of lo (slots before the chunk whose subtree reaches into
        // it): tested, not counted or added -- an earlier chunk owns them.
        // Subtrees ending at or before lo are stepped over untested.
        while (k < lo) {
            if (__builtin_expect(static_cast<unsigned>(k - k0) >= 64u, 0)) RG_WALK7_REFILL(k);
            const int l = k - k0;
            const int skip = __builtin_amdgcn_readlane(wb.y, l) + k0;
            if (skip <= lo) {
                if (next <= k) next = skip;
            } else {
                float cx, cy, cz, crit;
                if (LW) {
                    const v4f r = RG_WALK7_READ(k);
                    cx = r.x;
                    cy = r.y;
                    cz = r.z;
                    crit = r.w;
                } else {
                    cx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.x), l));
                    cy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.y), l));
                    cz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.z), l));
                    crit = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.w), l));
                }
                // a select, not a branch: lanes not at k keep their next
                const float dx = cx - px;
                const float dy = cy - py;
                const float dz = cz - pz;
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
                next = (next <= k && crit < r2) ? skip : next;  // opened: stays <= k + 1
            }
            // the smallest next slot: k + 1 if a lane opened k, else skip(k)
            // (every other lane waits at an ancestor's skip, >= skip(k))
            k = __ballot(next <= k + 1) ? k + 1 : skip;
        }
        // The chunk's own slots: every node a lane reaches is tested, counted
        // and, when accepted, added.  Slots are counted from the window's
        // first slot k0 here (l for the slot, nr for a lane's next, hr for
        // the chunk end), so a step needs no offset; a refill moves all three
        // by the shift.  The window check is at the end of a step, on the
        // next slot.
        if (k < hi) {
            if (static_cast<unsigned>(k - k0) >= 64u) RG_WALK7_REFILL(k);
            int l = k - k0;
            int nr = next - k0;
            int hr = hi - k0;
            // LW and PF: the record of the step's slot, read during the step
            // before
            v4f rn = {0.f, 0.f, 0.f, 0.f};
            if (LW && PF) rn = *(const lds_v4f *)(size_t)(rec0 + 16u * static_cast<unsigned>(l));
            do {
                float cx, cy, cz, crit;
                if (LW) {
                    const v4f r =
                        PF ? rn : *(const lds_v4f *)(size_t)(rec0 + 16u * static_cast<unsigned>(l));
                    cx = r.x;
                    cy = r.y;
                    cz = r.z;
                    crit = r.w;
                } else {
                    cx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.x), l));
                    cy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.y), l));
                    cz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.z), l));
                    crit = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wa.w), l));
                }
                const int skip = __builtin_amdgcn_readlane(wb.y, l);
                // The test is a select; only the force and the accepting
                // lanes' next slot are under a branch.
                const bool act = nr <= l;
                const unsigned long long mact = __ballot(act);
                const float dx = cx - px;
                const float dy = cy - py;
                const float dz = cz - pz;
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
                const bool lt = crit < r2;
                const unsigned long long mlt = __ballot(lt);
                // The next slot: with PF before the force, followed by the
                // next record's read (a slot past the window reads slot 63,
                // unused: the refill reads its slot 0 again); else after the
                // force, where l has had its last use, so the new slot can
                // take l's register.
                int kn = 0;
                if (LW && PF) {
                    kn = walk_next_slot(mact, mlt, l + 1, skip);
                    rn = *(const lds_v4f *)(size_t)(rec0 + 16u * min(static_cast<unsigned>(kn), 63u));
                }
                if (act && lt) {
                    nr = skip;
                    const float cw =
                        __int_as_float(__builtin_amdgcn_readlane(wb.x, l));
                    const float inv = __builtin_amdgcn_rsqf(r2);
                    const float s = cw * (inv * inv) * inv;
                    ax = fmaf(s, dx, ax);
                    ay = fmaf(s, dy, ay);
                    az = fmaf(s, dz, az);
                    ++accepts;
                }
                if (!(LW && PF)) kn = walk_next_slot(mact, mlt, l + 1, skip);
                l = kn;
                if (__builtin_expect(static_cast<unsigned>(l) >= 64u, 0)) {
                    if (l < hr) {
                        RG_WALK7_REFILL(k0 + l);
                        nr -= l;
                        hr -= l;
                        l = 0;
                        if (LW && PF) rn = *(const lds_v4f *)(size_t)rec0;
                    }
                }
            } while (l < hr);
        }
        s_part[c][0][lane] = ax;
        s_part[c][1][lane] = ay;
        s_part[c][2][lane] = az;
        if (accepts) atomicAdd(&s_vis[lane], accepts);
    }
    // No block barrier: a wave that runs out of chunks leaves at once (its
    // slot can take another block's wave), and the last wave to finish adds
    // the chunk partials.  The release orders this wave's partial and visit
    // writes before its count; the last wave's acquire sees every wave's.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    int done = 0;
    if (lane == 0) done = atomicAdd(&s_done, 1);
    done = __builtin_amdgcn_readfirstlane(__shfl(done, 0, 64));
    if (done == NW - 1) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        float ax = 0.f, ay = 0.f, az = 0.f;
        for (int c = 0; c < nc; ++c) {
            ax += s_part[c][0][lane];
            ay += s_part[c][1][lane];
            az += s_part[c][2][lane];
        }
        if (active) {
            const int p = order[t];
            store_acc(out, p, static_cast<double>(ax), static_cast<double>(ay),
                      static_cast<double>(az));
        }
        // A lane's walk examines a full binary tree: every node it examines
        // is accepted (leaves always are) or opened, and an opened node's two
        // children are examined, so 2 a - 1 nodes for a accepted ones (over
        // all chunks: the ancestor phases examine nothing of their own).
        const int vis = active ? 2 * s_vis[lane] - 1 : 0;
        const unsigned long long tot = wave_sum_u64(static_cast<unsigned long long>(vis));
        const int mx = wave_max_i32(vis);
        if (lane == 0) {
            atomicAdd(&stats[0], tot);
            atomicMax(&stats[1], static_cast<unsigned long long>(mx));
        }
    }
#undef RG_WALK7_READ
#undef RG_WALK7_REFILL
}


// The chunked wave walk (k_walk32_chunked, record window in LDS) with two
// chunk walks in flight per wave.  A step of the wave walk is one chain of
// dependent operations -- the slot's skip (v_readlane), its record (LDS),
// the criterion, the ballot that picks the next slot -- so a wave with one
// walk mostly waits.  Here each wave holds walks A and B on two chunks and
// advances both in one straight-line step: both records are read before
// either is tested, so one walk's LDS and ballot latency hides behind the
// other's arithmetic.  The ancestor and own phases of a chunk are the same
// code here (a scalar flag per step): before the chunk's first slot a node
// is only tested (subtrees ending at or before it are stepped over), from
// it on every node a lane reaches is counted and, when accepted, added.  Per
// lane and chunk the nodes, their order and the arithmetic are those of
// k_walk32_chunked, and the chunk partials are added in slot order at the
// end, so the accelerations are the same floats.
struct ChunkWalk {
    int c, lo, hi;  // chunk and its slot range (uniform)
    int k, k0;      // current slot and window base (uniform)
    int next;       // per lane: first slot this lane takes part in again
    int skipw;      // per lane: skip slot of window record k0 + lane
    float ax, ay, az;
    int visits;
};

template <int NW>
__global__ __launch_bounds__(64 * NW) void k_walk32_chunked2(
    const float4 *__restrict__ recA, const int2 *__restrict__ recB, const int *__restrict__ pre,
    const int *__restrict__ order, AccOut out, int n, float eps2,
    unsigned long long *__restrict__ stats, ChunkShape cs)
{
    if (tree_aborted(stats)) return;
    __shared__ int s_b[kChunkMaxB + 1];
    __shared__ int s_nc;
    __shared__ int s_ctr;
    __shared__ int s_done;
    extern __shared__ float s_part_dyn[];
    float(*s_part)[3][64] = reinterpret_cast<float(*)[3][64]>(s_part_dyn);
    __shared__ int s_vis[64];
    // per wave and walk: the window's rows (com x, y, z, G m, threshold)
    __shared__ float s_win[NW][2][5][64];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int g = xcd_run_block(blockIdx.x, gridDim.x, cs.run);
    const int t0 = g * 64;
    const int t = t0 + lane;
    const bool active = t < n;
    const int m = 2 * n - 1;
    const int nact = min(64, n - t0);
    const int last = t0 + nact - 1;
    int nr = 0, nl = 0;
    for (long long d = cs.dmin; last + d <= n - 1 && nr < kChunkSide; d <<= cs.shift) ++nr;
    for (long long d = cs.dmin; t0 - d >= 0 && nl < kChunkSide; d <<= cs.shift) ++nl;
    const int ni = nact > 33 ? 1 : 0;
    const int nbd = nr + ni + nl;
    if (wave == 0) {
        if (lane < nbd) {
            long long leaf;
            if (lane < nr)
                leaf = last + (static_cast<long long>(cs.dmin) << (cs.shift * (nr - 1 - lane)));
            else if (lane < nr + ni)
                leaf = t0 + 32;
            else
                leaf = t0 - (static_cast<long long>(cs.dmin) << (cs.shift * (lane - nr - ni)));
            s_b[lane + 1] = pre[leaf];
        }
        if (lane == 0) {
            s_b[0] = 0;
            s_b[nbd + 1] = m;
            s_nc = nbd + 1;
            s_ctr = 0;
            s_done = 0;
        }
        s_vis[lane] = 0;
    }
    float px = 0.f, py = 0.f, pz = 0.f;
    if (active) {
        const float4 self = recA[pre[t]];
        px = self.x;
        py = self.y;
        pz = self.z;
    }
    __syncthreads();
    const int nc = s_nc;
    int wvis = 0;

    // Store a finished chunk's partial sums and visits.
    auto finish = [&](ChunkWalk &w) {
        s_part[w.c][0][lane] = w.ax;
        s_part[w.c][1][lane] = w.ay;
        s_part[w.c][2][lane] = w.az;
        if (w.visits) atomicAdd(&s_vis[lane], w.visits);
        wvis += w.visits;
    };
    // Take the next non-empty chunk (farthest first, alternating ends);
    // false when the block has none left.
    auto fetch = [&](ChunkWalk &w) -> bool {
        while (true) {
            int j = 0;
            if (lane == 0) j = atomicAdd(&s_ctr, 1);
            j = __builtin_amdgcn_readfirstlane(__shfl(j, 0, 64));
            if (j >= nc) return false;
            w.c = (j & 1) ? nc - 1 - (j >> 1) : (j >> 1);
            w.lo = s_b[w.c];
            w.hi = s_b[w.c + 1];
            w.ax = w.ay = w.az = 0.f;
            w.visits = 0;
            w.next = (active && w.lo < w.hi) ? 0 : m;
            w.k = __builtin_amdgcn_readfirstlane(__ockl_wfred_min_i32(w.next));
            w.k0 = -64;
            if (w.k < w.hi) return true;
            finish(w);  // empty: its partials are zero
        }
    };
    auto refill = [&](ChunkWalk &w, int b) {
        if (w.k - w.k0 >= 64) {
            w.k0 = w.k;
            const int slot = w.k0 + lane;
            if (slot < m) {
                const float4 a = recA[slot];
                const int2 r = recB[slot];
                float *row = &s_win[wave][b][0][lane];
                row[0] = a.x;
                row[64] = a.y;
                row[128] = a.z;
                row[192] = __int_as_float(r.x);
                row[256] = a.w;
                w.skipw = r.y;
            }
            // the window is read across lanes: keep the reads after the writes
            __builtin_amdgcn_wave_barrier();
        }
    };
    // One step at slot w.k, in three parts so that two walks interleave:
    // load (skip and record), test and accumulate, advance.
    struct Rec {
        int skip;
        float cx, cy, cz, cw, crit;
    };
    auto load = [&](const ChunkWalk &w, int b) -> Rec {
        const int l = w.k - w.k0;
        Rec r;
        r.skip = max(__builtin_amdgcn_readlane(w.skipw, l), w.k + 1);
        const float *row = &s_win[wave][b][0][l];
        r.cx = row[0];
        r.cy = row[64];
        r.cz = row[128];
        r.cw = row[192];
        r.crit = row[256];
        return r;
    };
    auto test = [&](ChunkWalk &w, const Rec &r) {
        const bool own = w.k >= w.lo;             // uniform
        const bool over = !own && r.skip <= w.lo;  // uniform: subtree ends before the chunk
        const bool part = w.next <= w.k;
        const float dx = r.cx - px;
        const float dy = r.cy - py;
        const float dz = r.cz - pz;
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
        const bool acc = r.crit < r2;
        if (own && part) ++w.visits;
        if (part && (over || acc)) w.next = r.skip;
        if (own && part && acc) {
            const float inv = __builtin_amdgcn_rsqf(r2);
            const float sc = r.cw * (inv * inv) * inv;
            w.ax = fmaf(sc, dx, w.ax);
            w.ay = fmaf(sc, dy, w.ay);
            w.az = fmaf(sc, dz, w.az);
        }
    };
    auto advance = [&](ChunkWalk &w, const Rec &r) {
        // k + 1 if a lane opened k, else skip(k): every other lane waits at
        // an ancestor's skip, >= skip(k)
        w.k = __ballot(w.next <= w.k + 1) ? w.k + 1 : r.skip;
    };

    ChunkWalk A, B;
    bool aliveA = fetch(A);
    bool aliveB = aliveA && fetch(B);
    while (aliveA && aliveB) {
        refill(A, 0);
        refill(B, 1);
        const Rec ra = load(A, 0);
        const Rec rb = load(B, 1);
        test(A, ra);
        test(B, rb);
        advance(A, ra);
        advance(B, rb);
        if (A.k >= A.hi) {
            finish(A);
            aliveA = fetch(A);
        }
        if (B.k >= B.hi) {
            finish(B);
            aliveB = fetch(B);
        }
    }
    // the block's chunks are all taken: finish the walk still in flight
    if (aliveB) {
        A = B;
        aliveA = true;
        // its window sits in row set 1
        while (aliveA) {
            refill(A, 1);
            const Rec ra = load(A, 1);
            test(A, ra);
            advance(A, ra);
            if (A.k >= A.hi) {
                finish(A);
                aliveA = fetch(A);
            }
        }
    }
    while (aliveA) {
        refill(A, 0);
        const Rec ra = load(A, 0);
        test(A, ra);
        advance(A, ra);
        if (A.k >= A.hi) {
            finish(A);
            aliveA = fetch(A);
        }
    }
    {
        const unsigned long long tot = wave_sum_u64(static_cast<unsigned long long>(wvis));
        if (lane == 0) atomicAdd(&stats[0], tot);
    }
    // As k_walk32_chunked: no block barrier; the last wave to finish adds
    // the chunk partials in slot order.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    int done = 0;
    if (lane == 0) done = atomicAdd(&s_done, 1);
    done = __builtin_amdgcn_readfirstlane(__shfl(done, 0, 64));
    if (done == NW - 1) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        float ax = 0.f, ay = 0.f, az = 0.f;
        for (int c = 0; c < nc; ++c) {
            ax += s_part[c][0][lane];
            ay += s_part[c][1][lane];
            az += s_part[c][2][lane];
        }
        if (active) {
            const int p = order[t];
            store_acc(out, p, static_cast<double>(ax), static_cast<double>(ay),
                      static_cast<double>(az));
        }
        const int mx = wave_max_i32(s_vis[lane]);
        if (lane == 0) atomicMax(&stats[1], static_cast<unsigned long long>(mx));
    }
}

}  // namespace

// ================================================================ handle
struct rg_bh {
    int64_t n = 0;
    double theta = 0, eps = 0, G = 0;
    int bits = 21;
    int precision = RG_FP64;
    int device = 0;
    int nblk = 0;           // sort tiles
    int sort_items = 0;     // keys per thread in a sort tile
    int sort_bs = 256;      // threads per sort tile
    int passes = 0;         // radix passes
    int final_buf = 0;      // ping-pong buffer holding the sorted result
    hipStream_t last_stream = nullptr;
    bool computed = false;
    int walk_used = -1;   // walk variant of the last compute (-1 FP64)
    int walk_waves = 0;   // waves per group of 64 particles in that walk (0: per lane)
    int sort_passes = 0;  // global radix passes of the last compute's sort
    int sort_buckets = 0; // its sample-sort buckets (0: LSD sort)

    uint64_t *keys[2] = {nullptr, nullptr};
    int *idx[2] = {nullptr, nullptr};
    uint8_t *bkt = nullptr;       // sample sort: bucket id per particle
    unsigned *ghist = nullptr;    // kMaxPasses x kBuckets global digit counts
    unsigned *lookback = nullptr; // passes x nblk x kBuckets look-back words
    int *tickets = nullptr;
    int *left = nullptr, *right = nullptr, *parent = nullptr;
    int2 *range = nullptr;
    int *left_at_split = nullptr;
    int *pre = nullptr;  // FP32: the leaves' preorder slots
    int *counts = nullptr, *trank = nullptr, *tsorted = nullptr;
    int2 *tl_of = nullptr;  // per leaf: treelet starting there, internal nodes ending there
    unsigned long long *rank_lb = nullptr;  // treelet-rank look-back words, one per tile
    int rank_tiles = 0;
    Agg *agg = nullptr;
    WalkRec64 *rec64 = nullptr;  // FP64 walk records (node-id order)
    float4 *recA = nullptr;      // FP32 walk records (preorder)
    int2 *recB = nullptr;
    double *box = nullptr;    // bbox partials followed by the 4-double box
    unsigned long long *stats = nullptr;
    // Pinned host word the device sets when a compute aborts (a look-back
    // wait gave up); read, without synchronising, by the next call.
    unsigned long long *host_flag = nullptr;
    unsigned long long *host_flag_dev = nullptr;  // its device address
    std::vector<void *> allocs;

    // Optional per-phase timing: 5 events per compute (start, sorted, built,
    // aggregated, walked), up to kTimingCap computes between reads.
    static constexpr int kTimingCap = 4096;
    bool timing = false;
    int timed = 0;
    std::vector<hipEvent_t> ev;
};

namespace {

int bh_alloc(rg_bh *h, void **p, size_t bytes)
{
    RG_HIP(hipMalloc(p, bytes > 0 ? bytes : 16));
    h->allocs.push_back(*p);
    return RG_OK;
}

void bh_free(rg_bh *h)
{
    for (void *p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
    if (h->host_flag) (void)hipHostFree(h->host_flag);
    h->host_flag = nullptr;
    for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
    h->ev.clear();
}

int bh_mark(rg_bh *h, int phase, hipStream_t st)
{
    if (!h->timing || h->timed >= rg_bh::kTimingCap) return RG_OK;
    RG_HIP(hipEventRecord(h->ev[5 * h->timed + phase], st));
    return RG_OK;
}

}  // namespace

int rg::rg_bh_set_walk_mode_impl(int64_t v)
{
    RG_REQUIRE(v >= -1 && v <= 8, RG_ERR_ARG, "walk_mode must be -1 (auto) or 0..8");
    g_walk_mode = static_cast<int>(v);
    return RG_OK;
}

int rg::rg_bh_set_chunk_impl(const char *what, int64_t v)
{
    if (std::strcmp(what, "dmin") == 0) {
        RG_REQUIRE(v >= 1 && v <= (1 << 20), RG_ERR_ARG, "chunk_dmin must be in [1, 2^20]");
        g_chunk_dmin = static_cast<int>(v);
    } else if (std::strcmp(what, "shift") == 0) {
        RG_REQUIRE(v >= 0 && v <= 4, RG_ERR_ARG, "chunk_shift must be 0 (auto) or 1..4");
        g_chunk_shift = static_cast<int>(v);
    } else if (std::strcmp(what, "run") == 0) {
        RG_REQUIRE(v >= 1 && v <= 4096, RG_ERR_ARG, "chunk_run must be in [1, 4096]");
        g_chunk_run = static_cast<int>(v);
    } else if (std::strcmp(what, "read_ahead") == 0) {
        RG_REQUIRE(v >= -1 && v <= 1, RG_ERR_ARG,
                   "chunk_read_ahead must be -1 (auto), 0 (off) or 1 (on)");
        g_chunk_read_ahead = static_cast<int>(v);
    } else if (std::strcmp(what, "waves") == 0) {
        RG_REQUIRE(v == 0 || v == 2 || v == 4 || v == 8, RG_ERR_ARG,
                   "chunk_waves must be 0 (auto), 2 (walk 8 only), 4 or 8");
        g_chunk_waves = static_cast<int>(v);
    } else {
        rg::set_error("unknown parameter 'chunk_%s'", what);
        return RG_ERR_ARG;
    }
    return RG_OK;
}

int rg::rg_bh_set_walk_block_impl(int64_t v)
{
    RG_REQUIRE(v == 64 || v == 128 || v == 256 || v == 512, RG_ERR_ARG,
               "walk_block must be 64, 128, 256 or 512");
    g_walk_bs = static_cast<int>(v);
    return RG_OK;
}

extern "C" int rg_bh_create(int64_t n, double theta, double eps, double G, int morton_bits,
                            int precision, rg_bh **out)
{
    RG_REQUIRE(out, RG_ERR_ARG, "null out");
    *out = nullptr;
    RG_REQUIRE(n >= 2 && n <= (int64_t(1) << 29), RG_ERR_CAPACITY, "BH needs 2 <= n <= 2^29 (n=%lld)",
               static_cast<long long>(n));
    RG_REQUIRE(morton_bits >= 1 && morton_bits <= 21, RG_ERR_ARG, "morton_bits must be in [1, 21]");
    RG_REQUIRE(theta >= 0.0 && std::isfinite(theta), RG_ERR_ARG, "theta must be finite and >= 0");
    RG_REQUIRE(eps > 0.0 && std::isfinite(eps), RG_ERR_ARG, "softening must be > 0");
    RG_REQUIRE(precision == RG_FP64 || precision == RG_FP32, RG_ERR_ARG,
               "BH precision must be RG_FP64 or RG_FP32 (the walk has no compensated form), "
               "got %d", precision);
    rg_bh *h = new (std::nothrow) rg_bh();
    RG_REQUIRE(h, RG_ERR_CAPACITY, "out of host memory");
    h->n = n;
    h->theta = theta;
    h->eps = eps;
    h->G = G;
    h->bits = morton_bits;
    h->precision = precision;
    hipError_t e = hipGetDevice(&h->device);
    if (e != hipSuccess) {
        delete h;
        return rg::hip_fail(e, "hipGetDevice", __FILE__, __LINE__);
    }
    const size_t N = static_cast<size_t>(n);
    const size_t M = 2 * N - 1;
    h->sort_bs = sort_block_for(static_cast<int64_t>(N));
    h->sort_items = h->sort_bs == kSortBSWide ? kSortItemsSmall
                                              : sort_items_for(static_cast<int64_t>(N));
    h->nblk = sort_tiles(static_cast<int64_t>(N), h->sort_items, h->sort_bs);
    h->passes = (3 * morton_bits + kDigitBits - 1) / kDigitBits;
    int st = RG_OK;
#define ALLOC(ptr, bytes)                                                  \
    do {                                                                   \
        void *p_ = nullptr;                                                \
        st = bh_alloc(h, &p_, (bytes));                                    \
        if (st != RG_OK) {                                                 \
            bh_free(h);                                                    \
            delete h;                                                      \
            return st;                                                     \
        }                                                                  \
        (ptr) = reinterpret_cast<decltype(ptr)>(p_);                       \
    } while (0)
    ALLOC(h->keys[0], 8 * N);
    ALLOC(h->keys[1], 8 * N);
    ALLOC(h->idx[0], 4 * N);
    ALLOC(h->idx[1], 4 * N);
    ALLOC(h->bkt, N);
    ALLOC(h->ghist, 4 * kBuckets * kMaxPasses);
    ALLOC(h->tickets, 4 * kMaxPasses);
    ALLOC(h->lookback, 4 * kBuckets * static_cast<size_t>(h->nblk) * h->passes);
    ALLOC(h->left, 4 * (N - 1));
    ALLOC(h->right, 4 * (N - 1));
    ALLOC(h->range, 8 * (N - 1));
    ALLOC(h->left_at_split, 4 * (N - 1));
    ALLOC(h->parent, 4 * M);
    ALLOC(h->counts, 16);
    ALLOC(h->tl_of, 8 * N);
    ALLOC(h->trank, 4 * N);
    ALLOC(h->tsorted, 4 * N);
    h->rank_tiles = static_cast<int>((N + kRankTile - 1) / kRankTile);
    ALLOC(h->rank_lb, 8 * static_cast<size_t>(h->rank_tiles));
    ALLOC(h->agg, sizeof(Agg) * M);
    if (precision == RG_FP64) {
        ALLOC(h->rec64, sizeof(WalkRec64) * M);
    } else {
        ALLOC(h->pre, 4 * N);
        ALLOC(h->recA, 16 * M);
        ALLOC(h->recB, 8 * M);
    }
    ALLOC(h->box, 8 * bbox_part_doubles(static_cast<int>(n)));
    ALLOC(h->stats, 8 * kStatWords);
#undef ALLOC
    // Coherent (fine-grained) pinned memory: the device's store reaches the
    // host without a synchronisation.
    void *hf = nullptr, *hfd = nullptr;
    e = hipHostMalloc(&hf, 64, hipHostMallocCoherent);
    if (e == hipSuccess) {
        h->host_flag = static_cast<unsigned long long *>(hf);
        e = hipHostGetDevicePointer(&hfd, hf, 0);
    }
    if (e != hipSuccess) {
        bh_free(h);
        delete h;
        return rg::hip_fail(e, "hipHostMalloc", __FILE__, __LINE__);
    }
    h->host_flag_dev = static_cast<unsigned long long *>(hfd);
    __atomic_store_n(h->host_flag, 0ull, __ATOMIC_RELAXED);
    *out = h;
    return RG_OK;
}

namespace {

// A compute that aborted (a radix-sort or treelet-ranking look-back wait gave
// up) and has not been reported yet: reports it once and clears the flag.
int bh_take_abort(rg_bh *h)
{
    const unsigned long long f = __atomic_exchange_n(h->host_flag, 0ull, __ATOMIC_ACQ_REL);
    RG_REQUIRE(f == 0, RG_ERR_CAPACITY,
               "BH radix sort: %llu look-back waits gave up (sort_spin_limit) in an earlier "
               "compute on this handle; its tree and accelerations were not computed",
               f);
    return RG_OK;
}

// The sample sort needs the previous compute's order for its splitters and
// at most kMaxBuckets buckets of about kSampleTarget keys (LDS holds
// kBucketCap; buckets stay near their previous share while particles move
// little between computes).
constexpr int kSampleTarget = 1024;
constexpr int64_t kSampleMinN = 4096;
constexpr int64_t kSampleMaxN = 1572864;  // 256 buckets of 6,144 keys

int sample_buckets(int n)
{
    return static_cast<int>(std::min<int64_t>(kMaxBuckets,
                                              std::max<int64_t>(1, n / kSampleTarget)));
}

bool sample_sort_wanted(const rg_bh *h)
{
    if (rg::g_sort_sample == 0 || !h->computed || h->n > kSampleMaxN) return false;
    return rg::g_sort_sample == 1 || h->n >= kSampleMinN;
}

int bh_compute(rg_bh *h, const double *pos, const double *mass, const AccOut &out,
               void *stream, const KickDrift *kd = nullptr)
{
    RG_REQUIRE(h, RG_ERR_ARG, "null handle");
    {
        const int ab = bh_take_abort(h);
        if (ab != RG_OK) return ab;
    }
    RG_REQUIRE(pos && mass && out.acc, RG_ERR_ARG, "null pointer argument");
    const int n = static_cast<int>(h->n);
    const size_t b3 = 24 * static_cast<size_t>(n), b1 = 8 * static_cast<size_t>(n);
    RG_REQUIRE(!rg_overlap(out.acc, b3, pos, b3) && !rg_overlap(out.acc, b3, mass, b1),
               RG_ERR_ALIAS, "acc aliases pos or mass");
    if (out.vel)
        RG_REQUIRE(!rg_overlap(out.vel, b3, pos, b3) && !rg_overlap(out.vel, b3, mass, b1) &&
                       !rg_overlap(out.vel, b3, out.acc, b3),
                   RG_ERR_ALIAS, "vel aliases pos, mass or acc");
    int dev = 0;
    RG_HIP(hipGetDevice(&dev));
    RG_REQUIRE(dev == h->device, RG_ERR_ARG, "handle was created on device %d, current is %d",
               h->device, dev);
    hipStream_t st = rg::as_stream(stream);

    double *part = h->box + 8;
    double *box = h->box;
    int rc = bh_mark(h, 0, st);
    if (rc != RG_OK) return rc;
    int cur = 0;
    if (sample_sort_wanted(h)) {
        // Sample sort: buckets from the previous order's splitters, one
        // stable pass over the bucket ids, then one workgroup per bucket.
        // The splitters come from buffer f (the previous result); this
        // compute's keys go to the other one, and the result back to f.
        const int f = h->final_buf;
        const int nbk = sample_buckets(n);
        SampleSort ss{h->keys[f], h->idx[f], (n + nbk - 1) / nbk, nbk, h->bkt};
        const int lb_words = kBuckets * h->nblk;
        rc = launch_bbox_morton(pos, n, h->bits, part, box, h->stats, h->keys[f ^ 1], h->passes,
                                h->ghist, h->tickets, h->lookback, lb_words, h->counts,
                                h->tl_of, st, kd, &ss);
        if (rc != RG_OK) return rc;
        launch_onesweep(h->sort_items, h->sort_bs, h->nblk, st, h->keys[f ^ 1], nullptr,
                        h->keys[f], h->idx[f], n, 0, h->ghist, h->lookback, h->tickets,
                        h->stats, h->bkt);
        RG_CHECK_LAUNCH();
        k_bucket_sort<<<nbk, kBucketBS, 0, st>>>(h->keys[f], h->idx[f], h->keys[f ^ 1],
                                                 h->idx[f ^ 1], h->ghist, 3 * h->bits, h->stats);
        RG_CHECK_LAUNCH();
        cur = f;
        h->sort_passes = 1;
        h->sort_buckets = nbk;
    } else {
        const int lb_words = kBuckets * h->nblk * h->passes;
        rc = launch_bbox_morton(pos, n, h->bits, part, box, h->stats, h->keys[0], h->passes,
                                h->ghist, h->tickets, h->lookback, lb_words, h->counts,
                                h->tl_of, st, kd);
        if (rc != RG_OK) return rc;
        // Stable LSD radix sort, kDigitBits-bit digits, one kernel per pass.
        for (int p = 0; p < h->passes; ++p) {
            launch_onesweep(h->sort_items, h->sort_bs, h->nblk, st, h->keys[cur],
                            p == 0 ? nullptr : h->idx[cur], h->keys[cur ^ 1], h->idx[cur ^ 1], n,
                            kDigitBits * p, h->ghist + kBuckets * p,
                            h->lookback + static_cast<size_t>(kBuckets) * h->nblk * p,
                            h->tickets + p, h->stats);
            RG_CHECK_LAUNCH();
            cur ^= 1;
        }
        h->sort_passes = h->passes;
        h->sort_buckets = 0;
    }
    h->final_buf = cur;
    rc = bh_mark(h, 1, st);
    if (rc != RG_OK) return rc;
    const uint64_t *code = h->keys[cur];
    const int *order = h->idx[cur];

    const int g1 = min((n + 255) / 256, 16384);
    k_build<<<g1, 256, 0, st>>>(code, n, h->left, h->right, h->parent, h->range,
                                h->left_at_split, h->tl_of, h->rank_lb, h->rank_tiles, h->stats);
    RG_CHECK_LAUNCH();
    const bool fp64 = h->precision == RG_FP64;
    // FP32 walks also need the leaves' preorder slots
    k_rank_preorder<<<h->rank_tiles, 256, 0, st>>>(h->tl_of, n, h->trank, h->tsorted,
                                                   fp64 ? nullptr : h->pre, h->rank_lb, h->counts,
                                                   h->rank_tiles, h->stats, rg::g_spin_limit);
    RG_CHECK_LAUNCH();

    const double theta2 = h->theta * h->theta;
    WalkOut wo;
    wo.rec64 = h->rec64;
    wo.recA = h->recA;
    wo.recB = h->recB;
    wo.pre = h->pre;
    wo.G = h->G;
    wo.theta_zero = !(theta2 > 0.0);
    wo.inv_theta2 = wo.theta_zero ? 0.0 : 1.0 / theta2;
    wo.stats = h->stats;
    wo.host_flag = h->host_flag_dev;
    rc = bh_mark(h, 2, st);
    if (rc != RG_OK) return rc;
    const int gt = min(n / 64 + 64, 4096);
    const int gu = min(n / 128 + 8, 8192);
#define RG_AGG_LAUNCH(MODE)                                                                  \
    do {                                                                                     \
        k_agg_treelet<MODE><<<gt, kTreelet, 0, st>>>(order, pos, mass, n, h->left, h->right, \
                                                    h->parent, h->range, h->left_at_split,   \
                                                    h->tsorted, h->counts, h->agg,    \
                                                    wo);                                     \
        RG_CHECK_LAUNCH();                                                                   \
        k_agg_upper<MODE><<<gu, 256, 0, st>>>(n, h->right, h->parent, h->range,              \
                                              h->left_at_split, h->counts, h->trank,         \
                                              h->tsorted, h->agg, wo);                       \
    } while (0)
    if (fp64)
        RG_AGG_LAUNCH(0);
    else
        RG_AGG_LAUNCH(1);
#undef RG_AGG_LAUNCH
    RG_CHECK_LAUNCH();
    rc = bh_mark(h, 3, st);
    if (rc != RG_OK) return rc;

    const int wb = g_walk_bs;
    const int wg = (n + wb - 1) / wb;
    const double eps2 = h->eps * h->eps;
    h->walk_used = -1;
    h->walk_waves = 0;
    if (fp64) {
        k_walk64<<<wg, wb, 0, st>>>(h->rec64, order, out, n, h->G, theta2, eps2, h->stats);
    } else {
        // The chunked walk with its record window in LDS measured fastest
        // from 60k to 2M particles: its time is set by the longest wave,
        // chunking the far field shortens that wave the most, and a step
        // reads its record with one LDS broadcast load.
        const int mode = g_walk_mode >= 0 ? g_walk_mode : 7;
        h->walk_used = mode;
        if (mode >= 6) {  // 6, 7 and 8 (the chunked walks)
            // Measured on disk + bulge galaxies, 60k .. 2M particles: octave
            // chunks (growth 2) below ~350k, then growth 4 (fewer ancestor
            // tests once the walk is throughput-bound); 8 waves per group pay
            // below ~90k, where the groups are too few to fill the GPU, and
            // below ~300k with 10-bit keys, whose equal-key runs in a dense
            // centre make a few groups far heavier than the rest.
            const int shift = g_chunk_shift ? g_chunk_shift : (n < 350000 ? 1 : 2);
            int waves = g_chunk_waves ? g_chunk_waves
                                      : ((n < 90000 || (h->bits <= 10 && n < 300000)) ? 8 : 4);
            if (mode != 8 && waves == 2) waves = 4;  // two waves only for the two-walk kernel
            const ChunkShape cs{g_chunk_dmin, shift, g_chunk_run};
            const int ng = (n + 63) / 64;
            h->walk_waves = waves;
            // LDS for the chunk partial sums: the most chunks any group has
            // at this n (a block more per CU fits than with the two-sided
            // bound when no group can have full sides on both ends)
            const int ncmax = max_group_chunks(n, g_chunk_dmin, shift);
            const size_t smem = sizeof(float) * 3 * 64 * static_cast<size_t>(ncmax);
            const float e2 = static_cast<float>(eps2);
            if (mode == 8 && waves == 8)
                k_walk32_chunked2<8><<<ng, 512, smem, st>>>(h->recA, h->recB, h->pre, order, out,
                                                            n, e2, h->stats, cs);
            else if (mode == 8 && waves == 2)
                k_walk32_chunked2<2><<<ng, 128, smem, st>>>(h->recA, h->recB, h->pre, order, out,
                                                            n, e2, h->stats, cs);
            else if (mode == 8)
                k_walk32_chunked2<4><<<ng, 256, smem, st>>>(h->recA, h->recB, h->pre, order, out,
                                                            n, e2, h->stats, cs);
            else if (mode == 7) {
                // the read-ahead form unless switched off (measured no slower
                // from 120k to 10^6 particles, faster between)
                const bool pf = g_chunk_read_ahead != 0;
                if (waves == 8 && pf)
                    k_walk32_chunked<8, true, true><<<ng, 512, smem, st>>>(
                        h->recA, h->recB, h->pre, order, out, n, e2, h->stats, cs);
                else if (waves == 8)
                    k_walk32_chunked<8, true, false><<<ng, 512, smem, st>>>(
                        h->recA, h->recB, h->pre, order, out, n, e2, h->stats, cs);
                else if (pf)
                    k_walk32_chunked<4, true, true><<<ng, 256, smem, st>>>(
                        h->recA, h->recB, h->pre, order, out, n, e2, h->stats, cs);
                else
                    k_walk32_chunked<4, true, false><<<ng, 256, smem, st>>>(
                        h->recA, h->recB, h->pre, order, out, n, e2, h->stats, cs);
            }
            else if (waves == 8)
                k_walk32_chunked<8><<<ng, 512, smem, st>>>(h->recA, h->recB, h->pre, order, out,
                                                           n, e2, h->stats, cs);
            else
                k_walk32_chunked<4><<<ng, 256, smem, st>>>(h->recA, h->recB, h->pre, order, out,
                                                           n, e2, h->stats, cs);
        } else if (mode == 5) {
            h->walk_waves = 4;
            k_walk32_wave_split<4><<<(n + 63) / 64, 256, 0, st>>>(h->recA, h->recB, h->pre,
                                                                  order, out, n,
                                                                  static_cast<float>(eps2),
                                                                  h->stats);
        } else if (mode == 4) {
            h->walk_waves = 2;
            k_walk32_wave_split<2><<<(n + 127) / 128, 256, 0, st>>>(h->recA, h->recB, h->pre,
                                                                   order, out, n,
                                                                   static_cast<float>(eps2),
                                                                   h->stats);
        } else if (mode == 3) {
            h->walk_waves = 1;
            k_walk32_wave2<<<(n + 255) / 256, 256, 0, st>>>(h->recA, h->recB, h->pre, order,
                                                            out, n, static_cast<float>(eps2),
                                                            h->stats);
        } else if (mode == 2) {
            h->walk_waves = 1;
            k_walk32_wave<<<(n + 255) / 256, 256, 0, st>>>(h->recA, h->recB, h->pre, order,
                                                           out, n, static_cast<float>(eps2),
                                                           h->stats);
        } else if (mode == 1) {
            const int wg2 = (2 * n + wb - 1) / wb;
            k_walk32_pair<<<wg2, wb, 0, st>>>(h->recA, h->recB, h->pre, order, out, n,
                                              static_cast<float>(eps2), h->stats);
        } else {
            k_walk32<<<wg, wb, 0, st>>>(h->recA, h->recB, h->pre, order, out, n,
                                        static_cast<float>(eps2), h->stats);
        }
    }
    RG_CHECK_LAUNCH();
    rc = bh_mark(h, 4, st);
    if (rc != RG_OK) return rc;
    if (h->timing && h->timed < rg_bh::kTimingCap) ++h->timed;
    h->last_stream = st;
    h->computed = true;
    return RG_OK;
}

}  // namespace

extern "C" int rg_bh_compute(rg_bh *h, const double *pos, const double *mass, double *acc,
                             void *stream)
{
    return bh_compute(h, pos, mass, AccOut{acc, nullptr, 0.0, 0}, stream);
}

extern "C" int rg_bh_compute_add(rg_bh *h, const double *pos, const double *mass, double *acc,
                                 void *stream)
{
    return bh_compute(h, pos, mass, AccOut{acc, nullptr, 0.0, 1}, stream);
}

extern "C" int rg_bh_compute_kick(rg_bh *h, const double *pos, const double *mass, double *acc,
                                  double *vel, double dt_half, void *stream)
{
    RG_REQUIRE(vel, RG_ERR_ARG, "null vel");
    return bh_compute(h, pos, mass, AccOut{acc, vel, dt_half, 0}, stream);
}

extern "C" int rg_bh_step(rg_bh *h, double *pos, double *vel, const double *mass, double *acc,
                          double dt_half, double dt, void *stream)
{
    RG_REQUIRE(h, RG_ERR_ARG, "null handle");
    RG_REQUIRE(pos && vel && mass && acc, RG_ERR_ARG, "null pointer argument");
    const size_t b3 = 24 * static_cast<size_t>(h->n);
    RG_REQUIRE(!rg_overlap(pos, b3, vel, b3), RG_ERR_ALIAS, "vel aliases pos");
    const KickDrift kd{pos, vel, acc, dt_half, dt};
    return bh_compute(h, pos, mass, AccOut{acc, vel, dt_half, 0}, stream, &kd);
}

extern "C" int rg_bh_set_timing(rg_bh *h, int enable)
{
    RG_REQUIRE(h, RG_ERR_ARG, "null handle");
    if (enable && h->ev.empty()) {
        h->ev.resize(5 * rg_bh::kTimingCap, nullptr);
        for (auto &e : h->ev) RG_HIP(hipEventCreate(&e));
    }
    h->timing = enable != 0;
    h->timed = 0;
    return RG_OK;
}

extern "C" int rg_bh_timing(rg_bh *h, double *ms4, int64_t *count)
{
    RG_REQUIRE(h && ms4 && count, RG_ERR_ARG, "null argument");
    ms4[0] = ms4[1] = ms4[2] = ms4[3] = 0.0;
    *count = h->timed;
    for (int k = 0; k < h->timed; ++k) {
        hipEvent_t *e = &h->ev[5 * k];
        RG_HIP(hipEventSynchronize(e[4]));
        for (int ph = 0; ph < 4; ++ph) {
            float ms = 0.f;
            RG_HIP(hipEventElapsedTime(&ms, e[ph], e[ph + 1]));
            ms4[ph] += ms;
        }
    }
    h->timed = 0;
    return RG_OK;
}

extern "C" int rg_bh_export_tree(rg_bh *h, uint64_t *keys_sorted, int32_t *order, int32_t *left,
                                 int32_t *right, int32_t *parent, void *stream)
{
    RG_REQUIRE(h, RG_ERR_ARG, "null handle");
    RG_REQUIRE(h->computed, RG_ERR_ARG, "rg_bh_compute has not run on this handle");
    RG_REQUIRE(__atomic_load_n(h->host_flag, __ATOMIC_ACQUIRE) == 0, RG_ERR_CAPACITY,
               "the last compute on this handle aborted (radix-sort look-back gave up)");
    hipStream_t st = rg::as_stream(stream);
    const size_t N = static_cast<size_t>(h->n);
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    if (keys_sorted) RG_HIP(hipMemcpyAsync(keys_sorted, h->keys[h->final_buf], 8 * N, k, st));
    if (order) RG_HIP(hipMemcpyAsync(order, h->idx[h->final_buf], 4 * N, k, st));
    if (left) RG_HIP(hipMemcpyAsync(left, h->left, 4 * (N - 1), k, st));
    if (right) RG_HIP(hipMemcpyAsync(right, h->right, 4 * (N - 1), k, st));
    if (parent) RG_HIP(hipMemcpyAsync(parent, h->parent, 4 * (2 * N - 1), k, st));
    return RG_OK;
}

extern "C" int rg_bh_export_nodes(rg_bh *h, double *com, double *nmass, double *nmin,
                                  double *nmax, void *stream)
{
    RG_REQUIRE(h, RG_ERR_ARG, "null handle");
    RG_REQUIRE(h->computed, RG_ERR_ARG, "rg_bh_compute has not run on this handle");
    RG_REQUIRE(__atomic_load_n(h->host_flag, __ATOMIC_ACQUIRE) == 0, RG_ERR_CAPACITY,
               "the last compute on this handle aborted (radix-sort look-back gave up)");
    const int n = static_cast<int>(h->n);
    const int g = min((2 * n - 1 + 255) / 256, 16384);
    k_export_nodes<<<g, 256, 0, rg::as_stream(stream)>>>(h->agg, n, com, nmass, nmin, nmax);
    RG_CHECK_LAUNCH();
    return RG_OK;
}

extern "C" int rg_bh_stats(rg_bh *h, rg_bh_stats_t *out)
{
    RG_REQUIRE(h && out, RG_ERR_ARG, "null argument");
    RG_REQUIRE(h->computed, RG_ERR_ARG, "rg_bh_compute has not run on this handle");
    RG_HIP(hipStreamSynchronize(h->last_stream));
    unsigned long long s[7] = {0, 0, 0, 0, 0, 0, 0};
    RG_HIP(hipMemcpy(s, h->stats, sizeof(s), hipMemcpyDeviceToHost));
    out->n = h->n;
    out->visits_total = static_cast<int64_t>(s[0]);
    out->visits_max = static_cast<int64_t>(s[1]);
    out->nonfinite = static_cast<int64_t>(s[2]);
    out->walk_mode = h->walk_used;
    out->walk_waves = h->walk_waves;
    out->sort_passes = h->sort_passes;
    out->sort_buckets = h->sort_buckets;
    out->sort_spilled = static_cast<int64_t>(s[6]);
    if (s[3] != 0) {
        __atomic_store_n(h->host_flag, 0ull, __ATOMIC_RELEASE);  // reported here
        rg::set_error("BH radix sort: %llu look-back waits gave up (sort_spin_limit); the last "
                      "compute's tree and accelerations were not computed",
                      s[3]);
        return RG_ERR_CAPACITY;
    }
    RG_REQUIRE(s[4] == 0, RG_ERR_CAPACITY,
               "FP64 walk ran out of bounds for %llu particles (corrupt tree)", s[4]);
    return RG_OK;
}

extern "C" int rg_bh_destroy(rg_bh *h)
{
    if (!h) return RG_OK;
    int prev = 0;
    RG_HIP(hipGetDevice(&prev));
    if (prev != h->device) RG_HIP(hipSetDevice(h->device));
    bh_free(h);
    if (prev != h->device) RG_HIP(hipSetDevice(prev));
    delete h;
    return RG_OK;
}

extern "C" int rg_morton_codes(const double *pos, int64_t n, int bits, uint64_t *keys,
                               void *stream)
{
    RG_REQUIRE(n >= 0 && n <= (int64_t(1) << 29), RG_ERR_CAPACITY, "n out of range");
    if (n == 0) return RG_OK;
    RG_REQUIRE(pos && keys, RG_ERR_ARG, "null pointer argument");
    RG_REQUIRE(bits >= 1 && bits <= 21, RG_ERR_ARG, "bits must be in [1, 21]");
    void *scratch = nullptr;
    const int ni = static_cast<int>(n);
    const int st = rg::device_scratch(8 * bbox_part_doubles(ni), rg::as_stream(stream), &scratch);
    if (st != RG_OK) return st;
    double *box = static_cast<double *>(scratch);
    return launch_bbox_morton(pos, ni, bits, box + 8, box, nullptr, keys, 0, nullptr, nullptr,
                              nullptr, 0, nullptr, nullptr, rg::as_stream(stream));
}
