// Uniform-grid binning of the SPH gas particles on the device.
//
// Reference: core/solvers/neighbor_grid.py:49-136 (NeighborGrid.rebuild /
// rebuild_all; paths relative to the reference checkout), which bins on the
// host in numpy and uploads the result on every density pass.  Here the
// same outputs are computed on the device, bit-identically:
//
//   1. k_grid_partial / k_grid_setup  gas count, bounding box and h_max (exact
//      min/max, fixed-order reduction); then the cell size and the grid shape
//      by the reference's loop in FP64 (:64-74 and :107-117)
//   2. k_grid_keys  flat cell index (i * ny + j) * nz + k per gas particle;
//      other particles get the key max_cells, which sorts after every cell
//   3. k_onesweep   the stable LSD radix sort shared with the Barnes-Hut tree
//      (np.argsort(flat, kind="stable") over the gas list, :76)
//   4. k_grid_finish  gsort (gas indices in cell order, 0 past ngas) and
//      cell_start (first sorted position of each cell, 0 past n_cells), the
//      reference's zero-padded upload (:80-86)
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "rg_internal.h"
#include "rg_sort.h"

namespace {

constexpr int kGridBS = 256;

// Device-side geometry of the last rebuild.
struct GridDev {
    double lo[3];
    double cell;
    double inv_cell;
    long long dims[3];
    long long n_cells;
    long long ngas;       // -1: non-finite input
    long long nonfinite;
};

// Per block: gas count, non-finite count, lo[3], hi[3], h_max.
__global__ __launch_bounds__(kGridBS) void k_grid_partial(const double *__restrict__ pos,
                                                          const double *__restrict__ h,
                                                          const int *__restrict__ is_star, int n,
                                                          double *__restrict__ part)
{
    __shared__ double s[9][kGridBS];
    double v[9] = {0.0, 0.0, INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY,
                   -INFINITY};
    for (int i = blockIdx.x * kGridBS + threadIdx.x; i < n; i += gridDim.x * kGridBS) {
        if (is_star && is_star[i] != 0) continue;
        v[0] += 1.0;
        const double hi = h[i];
        bool bad = !isfinite(hi);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double x = pos[3 * i + a];
            bad = bad || !isfinite(x);
            v[2 + a] = fmin(v[2 + a], x);
            v[5 + a] = fmax(v[5 + a], x);
        }
        v[8] = fmax(v[8], hi);
        if (bad) v[1] += 1.0;
    }
    const int t = threadIdx.x;
#pragma unroll
    for (int c = 0; c < 9; ++c) s[c][t] = v[c];
    __syncthreads();
    for (int w = kGridBS / 2; w > 0; w >>= 1) {
        if (t < w) {
            s[0][t] += s[0][t + w];
            s[1][t] += s[1][t + w];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                s[2 + a][t] = fmin(s[2 + a][t], s[2 + a][t + w]);
                s[5 + a][t] = fmax(s[5 + a][t], s[5 + a][t + w]);
            }
            s[8][t] = fmax(s[8][t], s[8][t + w]);
        }
        __syncthreads();
    }
    if (t == 0)
#pragma unroll
        for (int c = 0; c < 9; ++c) part[c * gridDim.x + blockIdx.x] = s[c][0];
}

// The grid shape, neighbor_grid.py:62-74: cell = max(2 h_max, min_radius)
// (1.0 when not above 1e-9); dims = max(floor(span / cell) + 1, 1) per axis;
// while the cell count exceeds max_cells, cell *= 1.3.  Also clears the
// sort's histograms and tickets.
__device__ long long grid_dim(double span, double cell)
{
    const double f = floor(span / cell);
    const long long d = f < 4.0e18 ? static_cast<long long>(f) + 1 : (1ll << 62);
    return d > 1 ? d : 1;
}

__global__ __launch_bounds__(kGridBS) void k_grid_setup(const double *__restrict__ part, int nb,
                                                        double min_radius, long long max_cells,
                                                        GridDev *__restrict__ gd,
                                                        unsigned *__restrict__ ghist,
                                                        int *__restrict__ tickets,
                                                        unsigned long long *__restrict__ stats)
{
    for (int k = threadIdx.x; k < kMaxPasses * kBuckets; k += kGridBS) ghist[k] = 0u;
    if (threadIdx.x < kMaxPasses) tickets[threadIdx.x] = 0;
    if (threadIdx.x == 0) stats[3] = 0ull;  // this rebuild's look-back give-ups
    __shared__ double s[9][kGridBS];
    const int t = threadIdx.x;
    double v[9] = {0.0, 0.0, INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY,
                   -INFINITY};
    for (int b = t; b < nb; b += kGridBS) {
        v[0] += part[b];
        v[1] += part[nb + b];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            v[2 + a] = fmin(v[2 + a], part[(2 + a) * nb + b]);
            v[5 + a] = fmax(v[5 + a], part[(5 + a) * nb + b]);
        }
        v[8] = fmax(v[8], part[8 * nb + b]);
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) s[c][t] = v[c];
    __syncthreads();
    for (int w = kGridBS / 2; w > 0; w >>= 1) {
        if (t < w) {
            s[0][t] += s[0][t + w];
            s[1][t] += s[1][t + w];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                s[2 + a][t] = fmin(s[2 + a][t], s[2 + a][t + w]);
                s[5 + a][t] = fmax(s[5 + a][t], s[5 + a][t + w]);
            }
            s[8][t] = fmax(s[8][t], s[8][t + w]);
        }
        __syncthreads();
    }
    if (t != 0) return;
    GridDev g;
    g.ngas = static_cast<long long>(s[0][0]);
    g.nonfinite = static_cast<long long>(s[1][0]);
    double span[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = s[2 + a][0];
        span[a] = s[5 + a][0] - s[2 + a][0];
    }
    // Python's max(2.0 * h_max, min_radius) returns the second only when larger
    double cell = 2.0 * s[8][0];
    if (min_radius > cell) cell = min_radius;
    if (!(cell > 1e-9)) cell = 1.0;
    long long cells = 0;
    for (int it = 0; it < 4096; ++it) {
#pragma unroll
        for (int a = 0; a < 3; ++a) g.dims[a] = grid_dim(span[a], cell);
        // saturating product (numpy's int64 product, without its wrap-around)
        const bool big = g.dims[0] > max_cells || g.dims[1] > max_cells || g.dims[2] > max_cells ||
                         g.dims[0] * g.dims[1] > max_cells;
        cells = big ? max_cells + 1 : g.dims[0] * g.dims[1] * g.dims[2];
        if (cells <= max_cells) break;
        cell *= 1.3;
    }
    g.n_cells = cells;
    g.cell = cell;
    g.inv_cell = 1.0 / cell;
    if (g.nonfinite > 0) g.ngas = -1;
    *gd = g;
}

// Flat cell keys (neighbor_grid.py:75-76), sentinel max_cells for non-gas;
// the per-pass digit histograms for the sort; the look-back words cleared.
__global__ __launch_bounds__(256) void k_grid_keys(const double *__restrict__ pos,
                                                   const int *__restrict__ is_star, int n,
                                                   const GridDev *__restrict__ gd,
                                                   long long max_cells, uint64_t *__restrict__ keys,
                                                   int passes, unsigned *__restrict__ ghist,
                                                   unsigned *__restrict__ lookback, int lb_words)
{
    __shared__ unsigned s_h[kMaxPasses * kBuckets];
    for (int k = threadIdx.x; k < passes * kBuckets; k += blockDim.x) s_h[k] = 0u;
    __syncthreads();
    const GridDev g = *gd;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint64_t key = static_cast<uint64_t>(max_cells);
        if (g.ngas > 0 && !(is_star && is_star[i] != 0)) {
            long long ijk[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                // ((gpos - lo) / cell).astype(int64), clipped to [0, dims - 1]
                const double tq = (pos[3 * i + a] - g.lo[a]) / g.cell;
                long long q = tq < 4.0e18 ? static_cast<long long>(tq) : (1ll << 62);
                q = q < 0 ? 0 : q;
                ijk[a] = q > g.dims[a] - 1 ? g.dims[a] - 1 : q;
            }
            key = static_cast<uint64_t>((ijk[0] * g.dims[1] + ijk[1]) * g.dims[2] + ijk[2]);
        }
        keys[i] = key;
        for (int p = 0; p < passes; ++p)
            atomicAdd(&s_h[p * kBuckets + ((key >> (kDigitBits * p)) & kDigitMask)], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < passes * kBuckets; k += blockDim.x)
        if (s_h[k]) atomicAdd(&ghist[k], s_h[k]);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < lb_words; k += gridDim.x * blockDim.x)
        lookback[k] = 0u;
}

// gsort and cell_start (neighbor_grid.py:77-86); nothing is written when
// there is no gas (or the input was rejected), as the reference returns
// before touching its fields.
__global__ __launch_bounds__(256) void k_grid_finish(const uint64_t *__restrict__ skeys,
                                                     const int *__restrict__ sidx, int n,
                                                     const GridDev *__restrict__ gd,
                                                     long long max_cells, int *__restrict__ gsort,
                                                     int *__restrict__ cell_start)
{
    const long long ngas = gd->ngas;
    if (ngas <= 0) return;
    const long long n_cells = gd->n_cells;
    const long long total = (n > max_cells + 1 ? n : max_cells + 1);
    for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        if (i < n) gsort[i] = i < ngas ? sidx[i] : 0;
        if (i <= max_cells) {
            int v = 0;
            if (i <= n_cells) {
                // first sorted position whose cell is >= i
                long long a = 0, b = ngas;
                while (a < b) {
                    const long long mid = (a + b) >> 1;
                    if (skeys[mid] < static_cast<uint64_t>(i))
                        a = mid + 1;
                    else
                        b = mid;
                }
                v = static_cast<int>(a);
            }
            cell_start[i] = v;
        }
    }
}

}  // namespace

struct rg_grid {
    int64_t n = 0;
    double min_radius = 0.0;
    long long max_cells = 0;
    int device = 0;
    int passes = 0;
    int nblk = 0;  // sort tiles
    int sort_items = 0;  // keys per thread in a sort tile
    int sort_bs = 256;      // threads per sort tile
    int npart = 0; // reduction blocks
    uint64_t *keys[2] = {nullptr, nullptr};
    int *idx[2] = {nullptr, nullptr};
    unsigned *ghist = nullptr;
    int *tickets = nullptr;
    unsigned *lookback = nullptr;
    double *part = nullptr;
    GridDev *gd = nullptr;
    unsigned long long *stats = nullptr;
    std::vector<void *> allocs;
};

namespace {

void grid_free(rg_grid *g)
{
    for (void *p : g->allocs) (void)hipFree(p);
    g->allocs.clear();
}

int grid_alloc(rg_grid *g, void **p, size_t bytes)
{
    RG_HIP(hipMalloc(p, bytes > 0 ? bytes : 16));
    g->allocs.push_back(*p);
    return RG_OK;
}

}  // namespace

extern "C" int rg_grid_create(int64_t n, double min_radius, rg_grid **out)
{
    RG_REQUIRE(out, RG_ERR_ARG, "null out");
    *out = nullptr;
    RG_REQUIRE(n >= 0 && n <= (int64_t(1) << 28), RG_ERR_CAPACITY,
               "grid needs 0 <= n <= 2^28 (n=%lld)", static_cast<long long>(n));
    RG_REQUIRE(std::isfinite(min_radius), RG_ERR_ARG, "min_radius must be finite");
    rg_grid *g = new (std::nothrow) rg_grid;
    RG_REQUIRE(g, RG_ERR_CAPACITY, "out of host memory");
    g->n = n;
    g->min_radius = min_radius;
    // neighbor_grid.py:42: a cell budget of about 4 per particle
    g->max_cells = n * 4 > 64 ? n * 4 : 64;
    int bits = 0;
    while ((1ll << bits) <= g->max_cells) ++bits;  // keys are <= max_cells
    g->passes = (bits + kDigitBits - 1) / kDigitBits;
    const int64_t N = n > 0 ? n : 1;
    g->sort_bs = sort_block_for(N);
    g->sort_items = g->sort_bs == kSortBSWide ? kSortItemsSmall : sort_items_for(N);
    g->nblk = sort_tiles(N, g->sort_items, g->sort_bs);
    g->npart = static_cast<int>(std::min<int64_t>((N + kGridBS - 1) / kGridBS, 1024));
    int st = RG_OK;
    RG_HIP(hipGetDevice(&g->device));
#define GALLOC(ptr, bytes)                                                 \
    do {                                                                   \
        void *p_ = nullptr;                                                \
        st = grid_alloc(g, &p_, (bytes));                                  \
        if (st != RG_OK) {                                                 \
            grid_free(g);                                                  \
            delete g;                                                      \
            return st;                                                     \
        }                                                                  \
        (ptr) = reinterpret_cast<decltype(ptr)>(p_);                       \
    } while (0)
    GALLOC(g->keys[0], 8 * N);
    GALLOC(g->keys[1], 8 * N);
    GALLOC(g->idx[0], 4 * N);
    GALLOC(g->idx[1], 4 * N);
    GALLOC(g->ghist, 4 * kBuckets * kMaxPasses);
    GALLOC(g->tickets, 4 * kMaxPasses);
    GALLOC(g->lookback, 4 * kBuckets * static_cast<size_t>(g->nblk) * g->passes);
    GALLOC(g->part, 8 * 9 * static_cast<size_t>(g->npart));
    GALLOC(g->gd, sizeof(GridDev));
    GALLOC(g->stats, 8 * 4);
#undef GALLOC
    RG_HIP(hipMemset(g->stats, 0, 8 * 4));
    *out = g;
    return RG_OK;
}

extern "C" int rg_grid_rebuild(rg_grid *g, const double *pos, const double *h,
                               const int32_t *is_star, int32_t *gsort, int32_t *cell_start,
                               rg_grid_info_t *info, void *stream)
{
    RG_REQUIRE(g && info, RG_ERR_ARG, "null argument");
    info->ngas = 0;
    if (g->n == 0) return RG_OK;
    RG_REQUIRE(pos && h && gsort && cell_start, RG_ERR_ARG, "null pointer argument");
    int dev = 0;
    RG_HIP(hipGetDevice(&dev));
    RG_REQUIRE(dev == g->device, RG_ERR_ARG, "grid was created on device %d, current is %d",
               g->device, dev);
    const int n = static_cast<int>(g->n);
    const size_t b3 = 24 * static_cast<size_t>(n);
    const size_t bc = 4 * static_cast<size_t>(g->max_cells + 1);
    RG_REQUIRE(!rg_overlap(gsort, 4 * static_cast<size_t>(n), cell_start, bc) &&
                   !rg_overlap(gsort, 4 * static_cast<size_t>(n), pos, b3) &&
                   !rg_overlap(cell_start, bc, pos, b3),
               RG_ERR_ALIAS, "outputs overlap each other or pos");
    hipStream_t st = rg::as_stream(stream);
    k_grid_partial<<<g->npart, kGridBS, 0, st>>>(pos, h, is_star, n, g->part);
    RG_CHECK_LAUNCH();
    k_grid_setup<<<1, kGridBS, 0, st>>>(g->part, g->npart, g->min_radius, g->max_cells, g->gd,
                                        g->ghist, g->tickets, g->stats);
    RG_CHECK_LAUNCH();
    const int lb_words = kBuckets * g->nblk * g->passes;
    const int kg = std::min((n + 255) / 256, 256);
    k_grid_keys<<<kg, 256, 0, st>>>(pos, is_star, n, g->gd, g->max_cells, g->keys[0], g->passes,
                                    g->ghist, g->lookback, lb_words);
    RG_CHECK_LAUNCH();
    int cur = 0;
    for (int p = 0; p < g->passes; ++p) {
        launch_onesweep(g->sort_items, g->sort_bs, g->nblk, st, g->keys[cur], p == 0 ? nullptr : g->idx[cur],
                        g->keys[cur ^ 1], g->idx[cur ^ 1], n, kDigitBits * p,
                        g->ghist + kBuckets * p,
                        g->lookback + static_cast<size_t>(kBuckets) * g->nblk * p,
                        g->tickets + p, g->stats);
        RG_CHECK_LAUNCH();
        cur ^= 1;
    }
    const long long total = std::max<long long>(n, g->max_cells + 1);
    const int fg = static_cast<int>(std::min<long long>((total + 255) / 256, 8192));
    k_grid_finish<<<fg, 256, 0, st>>>(g->keys[cur], g->idx[cur], n, g->gd, g->max_cells, gsort,
                                      cell_start);
    RG_CHECK_LAUNCH();
    GridDev hd;
    RG_HIP(hipMemcpyAsync(&hd, g->gd, sizeof(GridDev), hipMemcpyDeviceToHost, st));
    unsigned long long flags[4];
    RG_HIP(hipMemcpyAsync(flags, g->stats, sizeof(flags), hipMemcpyDeviceToHost, st));
    RG_HIP(hipStreamSynchronize(st));
    RG_REQUIRE(flags[3] == 0, RG_ERR_CAPACITY,
               "grid radix sort: %llu look-back waits gave up (sort_spin_limit); gsort and "
               "cell_start are not valid",
               flags[3]);
    RG_REQUIRE(hd.ngas >= 0, RG_ERR_NONFINITE,
               "%lld gas particles have a non-finite position or smoothing length",
               hd.nonfinite);
    info->ngas = hd.ngas;
    if (hd.ngas == 0) return RG_OK;
    info->nx = hd.dims[0];
    info->ny = hd.dims[1];
    info->nz = hd.dims[2];
    info->n_cells = hd.n_cells;
    info->max_cells = g->max_cells;
    for (int a = 0; a < 3; ++a) info->lo[a] = hd.lo[a];
    info->cell = hd.cell;
    info->inv_cell = hd.inv_cell;
    return RG_OK;
}

extern "C" int rg_grid_max_cells(const rg_grid *g, int64_t *max_cells)
{
    RG_REQUIRE(g && max_cells, RG_ERR_ARG, "null argument");
    *max_cells = g->max_cells;
    return RG_OK;
}

extern "C" int rg_grid_destroy(rg_grid *g)
{
    if (!g) return RG_OK;
    int prev = 0;
    RG_HIP(hipGetDevice(&prev));
    if (prev != g->device) RG_HIP(hipSetDevice(g->device));
    grid_free(g);
    if (prev != g->device) RG_HIP(hipSetDevice(prev));
    delete g;
    return RG_OK;
}
