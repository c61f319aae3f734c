// Adjoint (vector-Jacobian product) of the direct N^2 acceleration.
//
// The reference has no gradients; this is the differentiation form built on
// the acceleration of core/solvers/gravity.py:13-26
//   a_i = G sum_j m_j d_ij / s_ij^3,   d_ij = x_j - x_i,  s^2 = |d|^2 + eps2.
// For a cotangent lam (n,3) on a, with the symmetric tidal tensor
//   T_kj = I / s^3 - 3 d d^T / s^5   (T_kj == T_jk)
// the VJP is
//   g_x[k] = G sum_j T_kj (m_k lam_j - m_j lam_k)
//   g_m[k] = -G sum_j lam_j . d_kj / s^3
// (the j == k terms vanish).  One fused pass per k over all j: the same LDS
// j-tiling as the forward kernel, with (x, m, lam) staged per j.
#include <algorithm>

#include "rg_internal.h"

namespace {

constexpr int kBS = 256;
constexpr int kChunk = 32;

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

// Particle i's sums over the block's j-slice (blockIdx.y): scaled into the
// outputs when there is one slice (part == nullptr), else stored as
// part[(y*4 + c)*n + i] for k_vjp_reduce.
__device__ __forceinline__ void finish_vjp(int i, double gx, double gy, double gz, double gm,
                                           double *__restrict__ gpos,
                                           double *__restrict__ gmass, int n, double G,
                                           double *__restrict__ part)
{
    if (part) {
        double *p = part + static_cast<size_t>(blockIdx.y) * 4 * n;
        p[i] = gx;
        p[n + i] = gy;
        p[2 * static_cast<size_t>(n) + i] = gz;
        p[3 * static_cast<size_t>(n) + i] = gm;
    } else {
        gpos[3 * i] = G * gx;
        gpos[3 * i + 1] = G * gy;
        gpos[3 * i + 2] = G * gz;
        gmass[i] = -G * gm;
    }
}

__global__ __launch_bounds__(kBS) void k_vjp64(const double *__restrict__ pos,
                                               const double *__restrict__ mass,
                                               const double *__restrict__ lam,
                                               double *__restrict__ gpos,
                                               double *__restrict__ gmass, int n, double G,
                                               double eps2, int jlen, double *__restrict__ part)
{
    __shared__ double4 sx[kBS];
    __shared__ double4 sl[kBS];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * kBS + tid;
    const bool ok = i < n;
    const double xi = ok ? pos[3 * i] : 0.0, yi = ok ? pos[3 * i + 1] : 0.0,
                 zi = ok ? pos[3 * i + 2] : 0.0;
    const double mi = ok ? mass[i] : 0.0;
    const double lxi = ok ? lam[3 * i] : 0.0, lyi = ok ? lam[3 * i + 1] : 0.0,
                 lzi = ok ? lam[3 * i + 2] : 0.0;
    double gx = 0.0, gy = 0.0, gz = 0.0, gm = 0.0;
    const int j0 = blockIdx.y * jlen;
    const int j1 = min(n, j0 + jlen);
    for (int t0 = j0; t0 < j1; t0 += kBS) {
        const int j = t0 + tid;
        if (j < j1) {
            sx[tid] = make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], mass[j]);
            sl[tid] = make_double4(lam[3 * j], lam[3 * j + 1], lam[3 * j + 2], 0.0);
        }
        __syncthreads();
        const int cnt = min(kBS, j1 - t0);
        for (int jj = 0; jj < cnt; ++jj) {
            const double4 q = sx[jj];
            const double4 l = sl[jj];
            const double dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
            const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, eps2)));
            const double inv = 1.0 / sqrt(r2);
            const double inv2 = inv * inv;
            const double inv3 = inv2 * inv;
            const double inv5 = inv3 * inv2;
            const double vx = mi * l.x - q.w * lxi;
            const double vy = mi * l.y - q.w * lyi;
            const double vz = mi * l.z - q.w * lzi;
            const double dv = fma(dz, vz, fma(dy, vy, dx * vx));
            const double c = 3.0 * dv * inv5;
            gx = fma(vx, inv3, fma(-c, dx, gx));
            gy = fma(vy, inv3, fma(-c, dy, gy));
            gz = fma(vz, inv3, fma(-c, dz, gz));
            gm = fma(fma(l.z, dz, fma(l.y, dy, l.x * dx)), inv3, gm);
        }
        __syncthreads();
    }
    if (ok) finish_vjp(i, gx, gy, gz, gm, gpos, gmass, n, G, part);
}

__global__ __launch_bounds__(kBS) void k_vjp32c(const double *__restrict__ pos,
                                                const double *__restrict__ mass,
                                                const double *__restrict__ lam,
                                                double *__restrict__ gpos,
                                                double *__restrict__ gmass, int n, double G,
                                                double eps2, int jlen, double *__restrict__ part)
{
    __shared__ float4 sh[kBS];
    __shared__ float4 slo[kBS];
    __shared__ float4 sl[kBS];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * kBS + tid;
    const bool ok = i < n;
    float xh, xl, yh, yl, zh, zl;
    split_df(ok ? pos[3 * i] : 0.0, xh, xl);
    split_df(ok ? pos[3 * i + 1] : 0.0, yh, yl);
    split_df(ok ? pos[3 * i + 2] : 0.0, zh, zl);
    const float mi = ok ? static_cast<float>(mass[i]) : 0.f;
    const float lxi = ok ? static_cast<float>(lam[3 * i]) : 0.f;
    const float lyi = ok ? static_cast<float>(lam[3 * i + 1]) : 0.f;
    const float lzi = ok ? static_cast<float>(lam[3 * i + 2]) : 0.f;
    const float eps2f = static_cast<float>(eps2);
    float s[4] = {0.f, 0.f, 0.f, 0.f}, c[4] = {0.f, 0.f, 0.f, 0.f};
    const int j0 = blockIdx.y * jlen;
    const int j1 = min(n, j0 + jlen);
    for (int t0 = j0; t0 < j1; t0 += kBS) {
        const int j = t0 + tid;
        if (j < j1) {
            float hx, lx, hy, ly, hz, lz;
            split_df(pos[3 * j], hx, lx);
            split_df(pos[3 * j + 1], hy, ly);
            split_df(pos[3 * j + 2], hz, lz);
            sh[tid] = make_float4(hx, hy, hz, static_cast<float>(mass[j]));
            slo[tid] = make_float4(lx, ly, lz, 0.f);
            sl[tid] = make_float4(static_cast<float>(lam[3 * j]), static_cast<float>(lam[3 * j + 1]),
                                  static_cast<float>(lam[3 * j + 2]), 0.f);
        }
        __syncthreads();
        const int cnt = min(kBS, j1 - t0);
        for (int c0 = 0; c0 < cnt; c0 += kChunk) {
            const int c1 = min(c0 + kChunk, cnt);
            float p[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
            for (int jj = c0; jj < c1; ++jj) {
                const float4 q = sh[jj];
                const float4 o = slo[jj];
                const float4 l = sl[jj];
                const float dx = (q.x - xh) + (o.x - xl);
                const float dy = (q.y - yh) + (o.y - yl);
                const float dz = (q.z - zh) + (o.z - zl);
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2f)));
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float inv2 = inv * inv;
                const float inv3 = inv2 * inv;
                const float inv5 = inv3 * inv2;
                const float vx = fmaf(mi, l.x, -q.w * lxi);
                const float vy = fmaf(mi, l.y, -q.w * lyi);
                const float vz = fmaf(mi, l.z, -q.w * lzi);
                const float dv = fmaf(dz, vz, fmaf(dy, vy, dx * vx));
                const float cc = 3.f * dv * inv5;
                p[0] = fmaf(vx, inv3, fmaf(-cc, dx, p[0]));
                p[1] = fmaf(vy, inv3, fmaf(-cc, dy, p[1]));
                p[2] = fmaf(vz, inv3, fmaf(-cc, dz, p[2]));
                p[3] = fmaf(fmaf(l.z, dz, fmaf(l.y, dy, l.x * dx)), inv3, p[3]);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float e;
                two_sum(s[k], p[k], s[k], e);
                c[k] += e;
            }
        }
        __syncthreads();
    }
    if (ok)
        finish_vjp(i, static_cast<double>(s[0]) + static_cast<double>(c[0]),
                   static_cast<double>(s[1]) + static_cast<double>(c[1]),
                   static_cast<double>(s[2]) + static_cast<double>(c[2]),
                   static_cast<double>(s[3]) + static_cast<double>(c[3]), gpos, gmass, n, G,
                   part);
}

// Adds the ns slice sums of each particle in slice order (FP64) and scales
// them as the single-slice kernels do.
__global__ __launch_bounds__(256) void k_vjp_reduce(const double *__restrict__ part, int ns,
                                                    int n, double *__restrict__ gpos,
                                                    double *__restrict__ gmass, double G)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = part[static_cast<size_t>(c) * n + i];
        for (int s = 1; s < ns; ++s) {
            const double *p = part + static_cast<size_t>(s) * 4 * n;
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = v[c] + p[static_cast<size_t>(c) * n + i];
        }
        gpos[3 * i] = G * v[0];
        gpos[3 * i + 1] = G * v[1];
        gpos[3 * i + 2] = G * v[2];
        gmass[i] = -G * v[3];
    }
}

}  // namespace

extern "C" int rg_direct_vjp(const double *pos, const double *mass, const double *lam,
                             double *g_pos, double *g_mass, int64_t n, double G, double eps2,
                             int precision, void *stream)
{
    RG_REQUIRE(n >= 0 && n <= (int64_t(1) << 29), RG_ERR_CAPACITY, "n out of range");
    if (n == 0) return RG_OK;
    RG_REQUIRE(pos && mass && lam && g_pos && g_mass, RG_ERR_ARG, "null pointer argument");
    RG_REQUIRE(eps2 > 0.0, RG_ERR_ARG, "eps2 must be > 0 (got %g)", eps2);
    const size_t b3 = 24 * static_cast<size_t>(n), b1 = 8 * static_cast<size_t>(n);
    const void *ins[3] = {pos, mass, lam};
    const size_t insz[3] = {b3, b1, b3};
    for (int k = 0; k < 3; ++k)
        RG_REQUIRE(!rg_overlap(g_pos, b3, ins[k], insz[k]) && !rg_overlap(g_mass, b1, ins[k], insz[k]),
                   RG_ERR_ALIAS, "outputs alias inputs");
    RG_REQUIRE(!rg_overlap(g_pos, b3, g_mass, b1), RG_ERR_ALIAS, "g_pos aliases g_mass");
    RG_REQUIRE(precision == RG_FP64 || precision == RG_FP32C || precision == RG_FP32, RG_ERR_ARG,
               "unknown precision %d", precision);
    const int ni = static_cast<int>(n);
    const int ib = (ni + kBS - 1) / kBS;
    // The VJP has no reference order to keep: every precision may split j.
    int jlen = ni;
    const int ns = rg::direct_slices(ni, 1, &jlen);
    hipStream_t st = rg::as_stream(stream);
    double *part = nullptr;
    if (ns > 1) {
        void *scratch = nullptr;
        const int rc = rg::device_scratch(sizeof(double) * 4 * static_cast<size_t>(ns) * ni,
                                          st, &scratch);
        if (rc != RG_OK) return rc;
        part = static_cast<double *>(scratch);
    }
    const dim3 grid(ib, ns);
    if (precision == RG_FP64)
        k_vjp64<<<grid, kBS, 0, st>>>(pos, mass, lam, g_pos, g_mass, ni, G, eps2, jlen, part);
    else
        k_vjp32c<<<grid, kBS, 0, st>>>(pos, mass, lam, g_pos, g_mass, ni, G, eps2, jlen, part);
    RG_CHECK_LAUNCH();
    if (part) {
        const int rg_grid = std::min((ni + 255) / 256, 4096);
        k_vjp_reduce<<<rg_grid, 256, 0, st>>>(part, ns, ni, g_pos, g_mass, G);
        RG_CHECK_LAUNCH();
    }
    return RG_OK;
}

// ---------------------------------------------------------------------
// Reverse of one leapfrog KDK step (core/engine.py:108-116), elementwise
// part.  Forward:  v_h = v + a h;  x' = x + v_h dt;  a' = A(x');  v' = v_h + a' h
// with h = dt/2.  Given the cotangents (xbar', vbar') of (x', v') and the
// VJP (g_x, g_m) of A at x' for lam' = abar' + vbar' h, the cotangents of
// the step's inputs are
//   xbar = xbar' + g_x,   vbar = vbar' + xbar dt,   mbar += g_m,
// and the cotangent of a (= A(x) from the previous step) is vbar h, so the
// previous step's VJP weight is lam = vbar h + vbar h_prev, with h_prev = h
// inside the trajectory and 0 for the initial acceleration.
namespace {

__global__ void k_kdk_adjoint_update(double *__restrict__ xbar, double *__restrict__ vbar,
                                     double *__restrict__ mbar, const double *__restrict__ gx,
                                     const double *__restrict__ gm, double *__restrict__ lam,
                                     int n, double dt, double lam_coef)
{
    const int m3 = 3 * n;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m3; k += gridDim.x * blockDim.x) {
        const double xb = xbar[k] + gx[k];
        const double vb = fma(xb, dt, vbar[k]);
        xbar[k] = xb;
        vbar[k] = vb;
        lam[k] = vb * lam_coef;
        if (k < n) mbar[k] += gm[k];
    }
}

}  // namespace

extern "C" int rg_kdk_adjoint_update(double *xbar, double *vbar, double *mbar, const double *g_pos,
                                     const double *g_mass, double *lam, int64_t n, double dt,
                                     double lam_coef, void *stream)
{
    RG_REQUIRE(n >= 0 && n <= (int64_t(1) << 29), RG_ERR_CAPACITY, "n out of range");
    if (n == 0) return RG_OK;
    RG_REQUIRE(xbar && vbar && mbar && g_pos && g_mass && lam, RG_ERR_ARG, "null pointer argument");
    const int m3 = static_cast<int>(3 * n);
    const int grid = min((m3 + 255) / 256, 4096);
    k_kdk_adjoint_update<<<grid, 256, 0, rg::as_stream(stream)>>>(
        xbar, vbar, mbar, g_pos, g_mass, lam, static_cast<int>(n), dt, lam_coef);
    RG_CHECK_LAUNCH();
    return RG_OK;
}
