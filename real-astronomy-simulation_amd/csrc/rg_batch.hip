// Batching form: an ensemble of independent systems of uneven N advanced
// together (no reference counterpart: the reference runs one system per
// Engine, core/engine.py:32).
//
// Layout: all systems' particles are concatenated in the caller's FP64
// arrays (pos/vel/acc (Ntot,3), mass (Ntot)); system k owns the index range
// [off_k, off_k + n_k).  One step is two launches:
//   k_batch_kick_drift  kick(dt_k/2) + drift(dt_k) over every particle, and
//                       the reset of the work-queue counter;
//   k_batch_force       a persistent grid: each workgroup repeatedly takes
//                       the next work item (system, i-tile) from an atomic
//                       counter -- items are ordered largest system first
//                       (LPT) -- sums the tile's forces over its own system
//                       with LDS j-tiles, writes acc and applies the closing
//                       kick(dt_k/2).
// The pair arithmetic and the per-particle summation order are the direct
// kernels' (rg_direct.hip): for each i the j-loop runs over the system in
// order with the same 32-term compensation chunks, so a batched force is
// bitwise the per-system rg_direct_acc result for the same precision when
// that call runs unsplit (rg_set_param("direct_split", 1); FP64 always is).
// Steps can be captured into a hipGraph (no memset nodes: the counter is
// reset by a kernel).
#include <algorithm>
#include <new>
#include <numeric>
#include <vector>

#include "rg_internal.h"

namespace {

constexpr int kChunk = 32;

struct SysInfo {
    int off, n;
    double G, eps2, dt;
};

__device__ __forceinline__ void split_df(double x, float &hi, float &lo)
{
    hi = static_cast<float>(x);
    lo = static_cast<float>(x - static_cast<double>(hi));
}

__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e)
{
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__global__ void k_batch_kick_drift(double *__restrict__ pos, double *__restrict__ vel,
                                   const double *__restrict__ acc, const int *__restrict__ psys,
                                   const SysInfo *__restrict__ sys, int ntot,
                                   int *__restrict__ counter)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ntot; i += gridDim.x * blockDim.x) {
        const double dt = sys[psys[i]].dt;
        const double h = 0.5 * dt;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double v = vel[3 * i + a] + acc[3 * i + a] * h;
            vel[3 * i + a] = v;
            pos[3 * i + a] = pos[3 * i + a] + v * dt;
        }
    }
}

// MODE 0: FP64 reference arithmetic; MODE 2: FP32 on hi/lo positions with
// chunked two-sum compensation (the direct kernels' MODE numbering).  A work
// item is BS * IPT i-particles: thread t holds i0 + t + k * BS, k < IPT,
// and every j-tile staged in LDS feeds all of them.  Each i-particle's sum
// runs over its system's j in order with the 32-term chunks at the same j
// (multiples of 32), so the result does not depend on BS or IPT.
template <int MODE, int BS, int IPT>
__global__ __launch_bounds__(BS) void k_batch_force(
    const double *__restrict__ pos, const double *__restrict__ mass, double *__restrict__ acc,
    double *__restrict__ vel, const int2 *__restrict__ items, int n_items,
    const SysInfo *__restrict__ sys, int *__restrict__ counter, int kick)
{
    using T4 = typename std::conditional<MODE == 0, double4, float4>::type;
    __shared__ T4 s_hi[BS];
    __shared__ float4 s_lo[MODE == 2 ? BS : 1];
    __shared__ int s_item;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int item = s_item;
        __syncthreads();  // everyone has read s_item before it is overwritten
        if (item >= n_items) break;
        const int2 it = items[item];
        const SysInfo S = sys[it.x];
        double pd[IPT][3];
        double axd[IPT], ayd[IPT], azd[IPT];
        float xh[IPT], xl[IPT], yh[IPT], yl[IPT], zh[IPT], zl[IPT];
        float sx[IPT], sy[IPT], sz[IPT], cx[IPT], cy[IPT], cz[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const int li = it.y + tid + k * BS;
            const bool ok = li < S.n;
            const int i = S.off + li;
#pragma unroll
            for (int a = 0; a < 3; ++a) pd[k][a] = ok ? pos[3 * i + a] : 0.0;
            axd[k] = ayd[k] = azd[k] = 0.0;
            split_df(pd[k][0], xh[k], xl[k]);
            split_df(pd[k][1], yh[k], yl[k]);
            split_df(pd[k][2], zh[k], zl[k]);
            sx[k] = sy[k] = sz[k] = cx[k] = cy[k] = cz[k] = 0.f;
        }
        const float eps2f = static_cast<float>(S.eps2);
        for (int t0 = 0; t0 < S.n; t0 += BS) {
            const int j = t0 + tid;
            if (j < S.n) {
                const int jj = S.off + j;
                const double qx = pos[3 * jj], qy = pos[3 * jj + 1], qz = pos[3 * jj + 2];
                const double qm = mass[jj];
                if constexpr (MODE == 0) {
                    s_hi[tid] = make_double4(qx, qy, qz, qm);
                } else {
                    float hx, lx, hy, ly, hz, lz;
                    split_df(qx, hx, lx);
                    split_df(qy, hy, ly);
                    split_df(qz, hz, lz);
                    s_hi[tid] = make_float4(hx, hy, hz, static_cast<float>(S.G * qm));
                    s_lo[tid] = make_float4(lx, ly, lz, 0.f);
                }
            }
            __syncthreads();
            const int cnt = min(BS, S.n - t0);
            if constexpr (MODE == 0) {
                for (int jj = 0; jj < cnt; ++jj) {
                    const double4 q = s_hi[jj];
#pragma unroll
                    for (int k = 0; k < IPT; ++k) {
                        const double dx = q.x - pd[k][0];
                        const double dy = q.y - pd[k][1];
                        const double dz = q.z - pd[k][2];
                        const double r2 = dx * dx + dy * dy + dz * dz + S.eps2;
                        const double inv = 1.0 / sqrt(r2);
                        const double s = S.G * q.w * inv * inv * inv;
                        axd[k] += s * dx;
                        ayd[k] += s * dy;
                        azd[k] += s * dz;
                    }
                }
            } else {
                for (int c0 = 0; c0 < cnt; c0 += kChunk) {
                    const int c1 = min(c0 + kChunk, cnt);
                    float qxp[IPT], qyp[IPT], qzp[IPT];
#pragma unroll
                    for (int k = 0; k < IPT; ++k) qxp[k] = qyp[k] = qzp[k] = 0.f;
#pragma unroll 4
                    for (int jj = c0; jj < c1; ++jj) {
                        const float4 q = s_hi[jj];
                        const float4 l = s_lo[jj];
#pragma unroll
                        for (int k = 0; k < IPT; ++k) {
                            const float dx = (q.x - xh[k]) + (l.x - xl[k]);
                            const float dy = (q.y - yh[k]) + (l.y - yl[k]);
                            const float dz = (q.z - zh[k]) + (l.z - zl[k]);
                            const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2f)));
                            const float inv = __builtin_amdgcn_rsqf(r2);
                            const float s = q.w * (inv * inv) * inv;
                            qxp[k] = fmaf(s, dx, qxp[k]);
                            qyp[k] = fmaf(s, dy, qyp[k]);
                            qzp[k] = fmaf(s, dz, qzp[k]);
                        }
                    }
#pragma unroll
                    for (int k = 0; k < IPT; ++k) {
                        float e;
                        two_sum(sx[k], qxp[k], sx[k], e);
                        cx[k] += e;
                        two_sum(sy[k], qyp[k], sy[k], e);
                        cy[k] += e;
                        two_sum(sz[k], qzp[k], sz[k], e);
                        cz[k] += e;
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const int li = it.y + tid + k * BS;
            if (li >= S.n) continue;
            const int i = S.off + li;
            double ax, ay, az;
            if constexpr (MODE == 0) {
                ax = axd[k]; ay = ayd[k]; az = azd[k];
            } else {
                ax = static_cast<double>(sx[k]) + static_cast<double>(cx[k]);
                ay = static_cast<double>(sy[k]) + static_cast<double>(cy[k]);
                az = static_cast<double>(sz[k]) + static_cast<double>(cz[k]);
            }
            acc[3 * i] = ax;
            acc[3 * i + 1] = ay;
            acc[3 * i + 2] = az;
            if (kick) {
                const double h = 0.5 * S.dt;
                vel[3 * i] = vel[3 * i] + ax * h;
                vel[3 * i + 1] = vel[3 * i + 1] + ay * h;
                vel[3 * i + 2] = vel[3 * i + 2] + az * h;
            }
        }
    }
}

// Per-item FP64 energy partials (KE, PE with j != i, momentum), items in
// natural order; then a fixed-order reduction per system.
constexpr int kEB = 256;

__global__ __launch_bounds__(kEB) void k_batch_energy_items(
    const double *__restrict__ pos, const double *__restrict__ vel,
    const double *__restrict__ mass, const int2 *__restrict__ items,
    const SysInfo *__restrict__ sys, double *__restrict__ part)
{
    __shared__ double4 sj[kEB];
    __shared__ double red[5][kEB];
    const int2 it = items[blockIdx.x];
    const SysInfo S = sys[it.x];
    const int tid = threadIdx.x;
    const int li = it.y + tid;
    const bool ok = li < S.n;
    const int i = S.off + li;
    const double px = ok ? pos[3 * i] : 0.0;
    const double py = ok ? pos[3 * i + 1] : 0.0;
    const double pz = ok ? pos[3 * i + 2] : 0.0;
    double acc_i = 0.0;
    for (int t0 = 0; t0 < S.n; t0 += kEB) {
        const int j = t0 + tid;
        if (j < S.n) {
            const int jj = S.off + j;
            sj[tid] = make_double4(pos[3 * jj], pos[3 * jj + 1], pos[3 * jj + 2], mass[jj]);
        }
        __syncthreads();
        const int cnt = min(kEB, S.n - t0);
        for (int jj = 0; jj < cnt; ++jj) {
            if (t0 + jj == li) continue;
            const double4 q = sj[jj];
            const double dx = q.x - px, dy = q.y - py, dz = q.z - pz;
            acc_i += q.w / sqrt(dx * dx + dy * dy + dz * dz + S.eps2);
        }
        __syncthreads();
    }
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (ok) {
        const double mi = mass[i];
        const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        v[0] = 0.5 * mi * (vx * vx + vy * vy + vz * vz);
        v[1] = -0.5 * S.G * mi * acc_i;
        v[2] = mi * vx;
        v[3] = mi * vy;
        v[4] = mi * vz;
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) red[c][tid] = v[c];
    __syncthreads();
    for (int w = kEB / 2; w > 0; w >>= 1) {
        if (tid < w)
#pragma unroll
            for (int c = 0; c < 5; ++c) red[c][tid] += red[c][tid + w];
        __syncthreads();
    }
    if (tid == 0)
#pragma unroll
        for (int c = 0; c < 5; ++c) part[5 * blockIdx.x + c] = red[c][0];
}

__global__ void k_batch_energy_reduce(const double *__restrict__ part,
                                      const int *__restrict__ first_item, int nsys,
                                      double *__restrict__ out)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nsys; k += gridDim.x * blockDim.x) {
        double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int it = first_item[k]; it < first_item[k + 1]; ++it)
#pragma unroll
            for (int c = 0; c < 5; ++c) v[c] += part[5 * it + c];
#pragma unroll
        for (int c = 0; c < 5; ++c) out[5 * k + c] = v[c];
    }
}

template <int MODE>
int launch_force(int tile, int grid, const double *pos, const double *mass, double *acc,
                 double *vel, const int2 *items, int n_items, const SysInfo *sys, int *counter,
                 int kick, hipStream_t st)
{
    // tile = i-particles per work item: one per thread up to 256, then 256
    // threads with 2 or 4 each
#define RG_BATCH_LAUNCH(BS, IPT)                                                              \
    k_batch_force<MODE, BS, IPT><<<grid, BS, 0, st>>>(pos, mass, acc, vel, items, n_items, sys, \
                                                      counter, kick)
    switch (tile) {
    case 64: RG_BATCH_LAUNCH(64, 1); break;
    case 128: RG_BATCH_LAUNCH(128, 1); break;
    case 256: RG_BATCH_LAUNCH(256, 1); break;
    case 512: RG_BATCH_LAUNCH(256, 2); break;
    case 1024: RG_BATCH_LAUNCH(256, 4); break;
    default:
        rg::set_error("batch tile must be 64, 128, 256, 512 or 1024 (got %d)", tile);
        return RG_ERR_ARG;
    }
#undef RG_BATCH_LAUNCH
    RG_CHECK_LAUNCH();
    return RG_OK;
}

}  // namespace

struct rg_batch {
    int nsys = 0;
    int ntot = 0;
    int precision = RG_FP32C;
    int tile = 128;          // i-particles per work item (threads: min(tile, 256))
    int grid = 0;            // persistent workgroups (0 = 4 per CU)
    int device = 0;
    std::vector<int> off;    // host copy, nsys + 1
    SysInfo *d_sys = nullptr;
    int *d_psys = nullptr;
    int2 *d_items = nullptr;     // LPT order, for the force queue
    int2 *d_items_nat = nullptr; // natural order, tile 256, for energies
    int *d_first = nullptr;      // per-system first natural item (nsys + 1)
    int n_items = 0, n_items_nat = 0;
    int *d_counter = nullptr;
    double *d_epart = nullptr;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;
};

namespace {

void batch_release_graph(rg_batch *h)
{
    if (h->exec) (void)hipGraphExecDestroy(h->exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->exec = nullptr;
    h->graph = nullptr;
    h->graph_steps = 0;
}

void batch_free(rg_batch *h)
{
    batch_release_graph(h);
    if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
    void *ps[] = {h->d_sys, h->d_psys, h->d_items, h->d_items_nat, h->d_first, h->d_counter,
                  h->d_epart};
    for (void *p : ps)
        if (p) (void)hipFree(p);
}

int build_items(rg_batch *h)
{
    // Force queue: tiles of h->tile particles, largest system first (LPT);
    // within a system, tiles in index order.
    std::vector<int> ord(h->nsys);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        return h->off[a + 1] - h->off[a] > h->off[b + 1] - h->off[b];
    });
    std::vector<int2> items;
    for (int k : ord) {
        const int n = h->off[k + 1] - h->off[k];
        for (int i0 = 0; i0 < n; i0 += h->tile) items.push_back(make_int2(k, i0));
    }
    if (h->d_items) RG_HIP(hipFree(h->d_items));
    h->d_items = nullptr;
    h->n_items = static_cast<int>(items.size());
    RG_HIP(hipMalloc(&h->d_items, sizeof(int2) * std::max<size_t>(items.size(), 1)));
    RG_HIP(hipMemcpy(h->d_items, items.data(), sizeof(int2) * items.size(),
                     hipMemcpyHostToDevice));
    return RG_OK;
}

}  // namespace

extern "C" int rg_batch_create(int64_t nsys, const int64_t *offsets, const double *G,
                               const double *eps2, const double *dt, int precision,
                               rg_batch **out)
{
    RG_REQUIRE(out, RG_ERR_ARG, "null out");
    *out = nullptr;
    RG_REQUIRE(nsys >= 1 && nsys <= (1 << 24), RG_ERR_ARG, "nsys out of range");
    RG_REQUIRE(offsets && G && eps2 && dt, RG_ERR_ARG, "null host array");
    RG_REQUIRE(precision == RG_FP64 || precision == RG_FP32C, RG_ERR_ARG,
               "batch precision must be RG_FP64 or RG_FP32C");
    RG_REQUIRE(offsets[0] == 0, RG_ERR_ARG, "offsets[0] must be 0");
    for (int64_t k = 0; k < nsys; ++k) {
        RG_REQUIRE(offsets[k + 1] > offsets[k], RG_ERR_ARG, "system %lld is empty",
                   static_cast<long long>(k));
        RG_REQUIRE(eps2[k] > 0.0, RG_ERR_ARG, "eps2 of system %lld must be > 0",
                   static_cast<long long>(k));
    }
    RG_REQUIRE(offsets[nsys] <= (int64_t(1) << 29), RG_ERR_CAPACITY, "too many particles");
    rg_batch *h = new (std::nothrow) rg_batch();
    RG_REQUIRE(h, RG_ERR_CAPACITY, "out of host memory");
    h->nsys = static_cast<int>(nsys);
    h->ntot = static_cast<int>(offsets[nsys]);
    h->precision = precision;
    h->off.assign(offsets, offsets + nsys + 1);
    int st = RG_OK;
#define TRY(expr)                                                          \
    do {                                                                   \
        hipError_t e_ = (expr);                                            \
        if (e_ != hipSuccess) {                                            \
            st = rg::hip_fail(e_, #expr, __FILE__, __LINE__);              \
            batch_free(h);                                                 \
            delete h;                                                      \
            return st;                                                     \
        }                                                                  \
    } while (0)
    TRY(hipGetDevice(&h->device));
    std::vector<SysInfo> sys(h->nsys);
    std::vector<int> psys(h->ntot);
    std::vector<int2> nat;
    std::vector<int> first(h->nsys + 1);
    for (int k = 0; k < h->nsys; ++k) {
        sys[k].off = h->off[k];
        sys[k].n = h->off[k + 1] - h->off[k];
        sys[k].G = G[k];
        sys[k].eps2 = eps2[k];
        sys[k].dt = dt[k];
        for (int i = h->off[k]; i < h->off[k + 1]; ++i) psys[i] = k;
        first[k] = static_cast<int>(nat.size());
        for (int i0 = 0; i0 < sys[k].n; i0 += kEB) nat.push_back(make_int2(k, i0));
    }
    first[h->nsys] = static_cast<int>(nat.size());
    h->n_items_nat = static_cast<int>(nat.size());
    TRY(hipMalloc(&h->d_sys, sizeof(SysInfo) * sys.size()));
    TRY(hipMemcpy(h->d_sys, sys.data(), sizeof(SysInfo) * sys.size(), hipMemcpyHostToDevice));
    TRY(hipMalloc(&h->d_psys, sizeof(int) * psys.size()));
    TRY(hipMemcpy(h->d_psys, psys.data(), sizeof(int) * psys.size(), hipMemcpyHostToDevice));
    TRY(hipMalloc(&h->d_items_nat, sizeof(int2) * nat.size()));
    TRY(hipMemcpy(h->d_items_nat, nat.data(), sizeof(int2) * nat.size(), hipMemcpyHostToDevice));
    TRY(hipMalloc(&h->d_first, sizeof(int) * first.size()));
    TRY(hipMemcpy(h->d_first, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice));
    TRY(hipMalloc(&h->d_counter, 64));
    TRY(hipMalloc(&h->d_epart, sizeof(double) * 5 * nat.size()));
    TRY(hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
#undef TRY
    st = build_items(h);
    if (st != RG_OK) {
        batch_free(h);
        delete h;
        return st;
    }
    *out = h;
    return RG_OK;
}

extern "C" int rg_batch_set_shape(rg_batch *h, int tile, int grid)
{
    RG_REQUIRE(h, RG_ERR_ARG, "null handle");
    RG_REQUIRE(tile == 64 || tile == 128 || tile == 256 || tile == 512 || tile == 1024,
               RG_ERR_ARG, "tile must be 64, 128, 256, 512 or 1024");
    RG_REQUIRE(grid >= 0 && grid <= 65536, RG_ERR_ARG, "grid out of range");
    if (tile != h->tile) {
        h->tile = tile;
        const int st = build_items(h);
        if (st != RG_OK) return st;
    }
    h->grid = grid;
    batch_release_graph(h);  // a captured graph holds the old shape
    return RG_OK;
}

namespace {

int batch_grid(const rg_batch *h)
{
    if (h->grid > 0) return h->grid;
    int cus = 256;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, h->device) == hipSuccess) cus = p.multiProcessorCount;
    return std::max(1, std::min(4 * cus, h->n_items));
}

int launch_step(rg_batch *h, double *pos, double *vel, double *acc, const double *mass,
                hipStream_t st)
{
    const int g = std::min((h->ntot + 255) / 256, 4096);
    k_batch_kick_drift<<<g, 256, 0, st>>>(pos, vel, acc, h->d_psys, h->d_sys, h->ntot,
                                          h->d_counter);
    RG_CHECK_LAUNCH();
    if (h->precision == RG_FP64)
        return launch_force<0>(h->tile, batch_grid(h), pos, mass, acc, vel, h->d_items,
                               h->n_items, h->d_sys, h->d_counter, 1, st);
    return launch_force<2>(h->tile, batch_grid(h), pos, mass, acc, vel, h->d_items, h->n_items,
                           h->d_sys, h->d_counter, 1, st);
}

__global__ void k_reset_counter(int *c) { *c = 0; }

int check_arrays(rg_batch *h, const void *pos, const void *vel, const void *acc,
                 const void *mass)
{
    RG_REQUIRE(h, RG_ERR_ARG, "null handle");
    RG_REQUIRE(pos && mass && acc, RG_ERR_ARG, "null pointer argument");
    const size_t b3 = 24 * static_cast<size_t>(h->ntot), b1 = 8 * static_cast<size_t>(h->ntot);
    RG_REQUIRE(!rg_overlap(acc, b3, pos, b3) && !rg_overlap(acc, b3, mass, b1), RG_ERR_ALIAS,
               "acc aliases pos or mass");
    if (vel)
        RG_REQUIRE(!rg_overlap(vel, b3, pos, b3) && !rg_overlap(vel, b3, acc, b3) &&
                       !rg_overlap(vel, b3, mass, b1),
                   RG_ERR_ALIAS, "vel aliases another array");
    return RG_OK;
}

}  // namespace

extern "C" int rg_batch_acc(rg_batch *h, const double *pos, const double *mass, double *acc,
                            void *stream)
{
    int rc = check_arrays(h, pos, nullptr, acc, mass);
    if (rc != RG_OK) return rc;
    hipStream_t st = rg::as_stream(stream);
    k_reset_counter<<<1, 1, 0, st>>>(h->d_counter);
    RG_CHECK_LAUNCH();
    if (h->precision == RG_FP64)
        return launch_force<0>(h->tile, batch_grid(h), pos, mass, acc, nullptr, h->d_items,
                               h->n_items, h->d_sys, h->d_counter, 0, st);
    return launch_force<2>(h->tile, batch_grid(h), pos, mass, acc, nullptr, h->d_items,
                           h->n_items, h->d_sys, h->d_counter, 0, st);
}

extern "C" int rg_batch_step(rg_batch *h, double *pos, double *vel, double *acc,
                             const double *mass, int64_t nsteps, void *stream)
{
    int rc = check_arrays(h, pos, vel, acc, mass);
    if (rc != RG_OK) return rc;
    RG_REQUIRE(vel, RG_ERR_ARG, "null vel");
    RG_REQUIRE(nsteps >= 0, RG_ERR_ARG, "nsteps < 0");
    hipStream_t st = rg::as_stream(stream);
    for (int64_t s = 0; s < nsteps; ++s) {
        rc = launch_step(h, pos, vel, acc, mass, st);
        if (rc != RG_OK) return rc;
    }
    return RG_OK;
}

extern "C" int rg_batch_graph_capture(rg_batch *h, double *pos, double *vel, double *acc,
                                      const double *mass, int64_t nsteps)
{
    int rc = check_arrays(h, pos, vel, acc, mass);
    if (rc != RG_OK) return rc;
    RG_REQUIRE(vel, RG_ERR_ARG, "null vel");
    RG_REQUIRE(nsteps >= 1 && nsteps <= 100000, RG_ERR_ARG, "nsteps out of range");
    batch_release_graph(h);
    RG_HIP(hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
    for (int64_t s = 0; s < nsteps && rc == RG_OK; ++s)
        rc = launch_step(h, pos, vel, acc, mass, h->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->cap_stream, &g);
    if (rc != RG_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return rg::hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    h->graph = g;
    RG_HIP(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
    h->graph_steps = static_cast<int>(nsteps);
    return RG_OK;
}

extern "C" int rg_batch_graph_launch(rg_batch *h, void *stream)
{
    RG_REQUIRE(h, RG_ERR_ARG, "null handle");
    RG_REQUIRE(h->exec, RG_ERR_ARG, "no captured graph (call rg_batch_graph_capture)");
    RG_HIP(hipGraphLaunch(h->exec, rg::as_stream(stream)));
    return RG_OK;
}

extern "C" int rg_batch_energy(rg_batch *h, const double *pos, const double *vel,
                               const double *mass, double *out, void *stream)
{
    RG_REQUIRE(h && pos && vel && mass && out, RG_ERR_ARG, "null argument");
    hipStream_t st = rg::as_stream(stream);
    k_batch_energy_items<<<h->n_items_nat, kEB, 0, st>>>(pos, vel, mass, h->d_items_nat,
                                                         h->d_sys, h->d_epart);
    RG_CHECK_LAUNCH();
    k_batch_energy_reduce<<<(h->nsys + 255) / 256, 256, 0, st>>>(h->d_epart, h->d_first,
                                                                 h->nsys, out);
    RG_CHECK_LAUNCH();
    return RG_OK;
}

extern "C" int rg_batch_destroy(rg_batch *h)
{
    if (!h) return RG_OK;
    batch_free(h);
    delete h;
    return RG_OK;
}
