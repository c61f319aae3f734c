// Direct N^2 gravity, leapfrog kick/drift and FP64 energy diagnostics.
//
// Reference: core/solvers/gravity.py:13-57, core/integrators/leapfrog.py:11-22
// (paths relative to the reference checkout).
//
// Layout: the caller's FP64 (n,3) positions and (n,) masses are read in
// place.  Each block stages a tile of j-particles into LDS, converting on
// the fly to the precision of the pair arithmetic, and every lane then
// streams the tile from LDS (all lanes read the same address: a broadcast)
// against the i-particles it holds in registers.  The pair sum is a
// vector-ALU workload (one rsqrt, about 16 VALU slots per pair); it is not a
// dense contraction, so there is no MFMA here.
//
// This file is compiled with -ffp-contract=off: every multiply-add that is
// fused is written as fma()/fmaf().  The FP64 kernels keep the reference's
// expression order with IEEE 1.0/sqrt, so each particle's FP64 result is the
// same double the sequential-j CPU evaluation gives.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <mutex>

#include "rg_internal.h"

// Scratch of rg_direct_acc_noalloc: the FP32/FP32C slice partials live in
// this static array instead of library-allocated memory.  Each GPU holds its
// own copy from the library's first launch on it until the process ends;
// host code uses the per-device address from hipGetSymbolAddress, never the
// array's name (which is a host-side stand-in).
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// Where a launch's scratch comes from: the library's grow-only buffer for
// the stream (ptr == nullptr), or a fixed region of cap bytes.
struct ScratchSrc {
    void *ptr;
    size_t cap;
};

constexpr int kBS = 256;      // threads per block = j-tile length
constexpr int kChunk = 32;    // FP32C: pair terms per partial before the two-sum fold
// j-split of the FP32/FP32C sums.  Slices per i-block by default (32 with
// one i-particle per lane, 64 with more), and the most slice-particles
// (slices x n) the FP64 slice sums may hold (2^24: 402 MB of 3 doubles).
// Measured on one MI355X, FP32C, one i per lane, against the unsplit sum:
// n = 16,384 0.120 ms against 0.894; 65,536 1.69 against 3.59; 262,144 26.1
// against 30.7 (more resident waves per SIMD hide the LDS and rsqrt
// latency).
constexpr int kSplitOne = 32;
constexpr int kSplitMany = 64;
constexpr int64_t kMaxSliceParticles = int64_t(1) << 24;
constexpr int kMaxSplit = 256;
// i-particles per lane by default.  FP64: 1 (108.3 ms with 2 at 262,144,
// 104.8 with 1).  FP32C: 4 from n = 65,536 (each LDS broadcast feeds four
// pairs; 262,144: 25.1 ms with 4 and 64 slices, 26.4 with 1 and 32;
// 65,536: 1.66 against 1.70), 1 below (16,384: 0.121 against 0.145).
// FP32: 2 from n = 131,072 (262,144: 18.63 ms against 18.91), 1 below.
constexpr int kIpt4From = 65536;
constexpr int kIpt2From = 131072;
// FP64 below kPipeBelow runs the pipelined kernel (k_direct64_pipe): one
// wave per 64 i-particles leaves most SIMDs idle there.  Measured on one
// MI355X (ms per evaluation, one wave / 8 / 16 waves per 64 i-particles):
// n = 4,096 0.512 / 0.192 / 0.155; 16,384 2.05 / 0.780 / 0.624; 20,000
// 2.50 / 1.44 / 1.50; 32,768 4.15 / 2.41 / 2.47.  Above, the one-wave
// kernel needs at least two waves on every SIMD, evenly spread, to beat the
// 8-wave pipeline (about 570 G interactions/s from n = 49,152); it reaches
// 620 with two waves per SIMD and 657 with four.  Measured (one wave / 8
// waves, ms): 49,152 6.16 / 4.60; 65,536 (one wave per SIMD) 8.20 / 7.89;
// 81,920 (1.25) 17.5 / 12.1; 98,304 (1.5) 21.0 / 17.1; 110,000 (1.68, fill
// 0.84) 23.4 / 22.1; 120,000 (fill 0.92) 25.5 / 27.8; 131,072 (2) 27.8 /
// 29.7; 163,840 (2.5, fill 0.83) 51.5 / 46.3; 196,608 (3) 60.4 / 67.1;
// 262,144 (4) 104.6 / 118.9.
constexpr int kPipe16Upto = 16384;
constexpr int kPipeBelow = 49152;
// one-wave waves on the busiest SIMD / the least even fill at which the
// one-wave kernel is used
constexpr int kOneWaveMinPerSimd = 2;
constexpr double kOneWaveMinFill = 0.9;

int g_direct_ipt = 0;         // 0 = heuristic; else i-particles per thread (1, 2, 4)
int g_direct_split = 0;       // 0 = by n; else j-slices per i-block (FP32/FP32C)
int g_fp64_pipe = 0;          // FP64 waves per 64 i-particles: 0 by n, 1 off, 4, 8, 16
int g_fp32c_d64 = 0;          // FP32C pair difference: 0 hi/lo FP32, 1 FP64 then rounded

__device__ __forceinline__ void split_df(double x, float &hi, float &lo)
{
    hi = static_cast<float>(x);
    lo = static_cast<float>(x - static_cast<double>(hi));
}

// The same split of a hi/lo position X = x + off (x FP64, off an FP32
// offset): hi = x rounded to FP32, lo = ((x - hi) + off) rounded to FP32;
// x - hi is exact, so hi + lo carries X to about 2^-48 of itself.
__device__ __forceinline__ void split_df_off(double x, float off, float &hi, float &lo)
{
    hi = static_cast<float>(x);
    lo = static_cast<float>((x - static_cast<double>(hi)) + static_cast<double>(off));
}

// Compensated drift of a hi/lo coordinate X = xh + xl (xh FP64, xl an FP32
// offset of at most half an ulp of xh): X += v dt.  The increment (v * dt,
// rounded once, as leapfrog.py:21 rounds it) is added to the offset first,
// then TwoSum(xh, t) renormalises: xh takes the rounded sum and xl its
// exact rounding error, rounded to FP32.  A step therefore keeps the
// increment to about 2^-53 of itself, where an FP64 drift rounds the new
// position to half an ulp of |X|; the FP32 rounding of the offset loses at
// most 2^-25 ulp(xh) per step.
__device__ __forceinline__ void drift_hilo(double &xh, float &xl, double v, double dt)
{
    const double t = static_cast<double>(xl) + v * dt;
    const double s = xh + t;
    const double bb = s - xh;
    const double e = (xh - (s - bb)) + (t - bb);
    xh = s;
    xl = static_cast<float>(e);
}

// What the kernel does with an i-particle's sum once the j loop is done.
enum EpiMode { kStore = 0, kAdd = 1, kAddExt = 2 };

struct Epilogue {
    int mode;                                   // EpiMode
    int is_hernquist;
    double M_d, a, b, M_h, a_h, smbh_mass;      // analytic potential (kAddExt)
};

Epilogue make_epilogue(int mode, const rg_potential_t *pot)
{
    Epilogue e{};
    e.mode = mode;
    if (pot) {
        e.M_d = pot->M_d; e.a = pot->a; e.b = pot->b;
        e.M_h = pot->M_h; e.a_h = pot->a_h; e.smbh_mass = pot->smbh_mass;
        e.is_hernquist = pot->is_hernquist;
    }
    return e;
}

// The analytic galaxy potential of core/gas_disk_engine.py:34-49 (the same
// arithmetic as core/living_galaxy_engine.py:134-149), FP64, in that order:
// a Miyamoto-Nagai disk, then a Hernquist (is_hernquist == 1) or Plummer
// halo, then a central point mass when smbh_mass > 0.  With init the disk
// vector starts the sum (living_galaxy_engine.py:138); otherwise it is
// added to the self-gravity sum already in av (gas_disk_engine.py:38).
// Taichi's `x ** 2` is x*x, `v.norm()` sqrt((x*x + y*y) + z*z), and
// `x ** (-1.5)` pow(x, -1.5).
__device__ void ext_terms(double px, double py, double pz, double G, const Epilogue &e,
                          double av[3], bool init)
{
    const double R2 = px * px + py * py;
    const double S = sqrt(pz * pz + e.b * e.b);
    const double aS = e.a + S;
    const double invD3 = pow(R2 + aS * aS, -1.5);
    const double gmd = -G * e.M_d;
    const double dx = gmd * px * invD3;
    const double dy = gmd * py * invD3;
    const double dz = gmd * pz * aS * invD3 / S;
    if (init) {
        av[0] = dx; av[1] = dy; av[2] = dz;
    } else {
        av[0] += dx; av[1] += dy; av[2] += dz;
    }
    double c;
    if (e.is_hernquist == 1) {
        const double r = sqrt(px * px + py * py + pz * pz);
        const double r_safe = fmax(r, 1e-6);
        const double t = r_safe + e.a_h;
        c = -(G * e.M_h / (t * t * r_safe));
    } else {
        const double rh = R2 + pz * pz + e.a_h * e.a_h;
        c = -(G * e.M_h * pow(rh, -1.5));
    }
    av[0] += c * px; av[1] += c * py; av[2] += c * pz;
    if (e.smbh_mass > 0.0) {
        const double r_safe = sqrt(px * px + py * py + pz * pz) + 1e-6;
        const double cb = -(G * e.smbh_mass / (r_safe * r_safe * r_safe));
        av[0] += cb * px; av[1] += cb * py; av[2] += cb * pz;
    }
}

// Knuth's TwoSum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e)
{
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

int elementwise_grid(int m);

// What happens to particle i's finished sum: stored, added to acc (with the
// analytic terms first for kAddExt), then the closing kick when vel is set.
__device__ __forceinline__ void finish_particle(int i, double ax, double ay, double az,
                                                const double *__restrict__ pos,
                                                double *__restrict__ acc,
                                                double *__restrict__ vel, double G,
                                                double dt_half, const Epilogue &epi)
{
    if (epi.mode != kStore) {  // uniform across the grid
        if (epi.mode == kAddExt) {
            double av[3] = {ax, ay, az};
            ext_terms(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], G, epi, av, false);
            ax = av[0]; ay = av[1]; az = av[2];
        }
        ax = acc[3 * i] + ax;
        ay = acc[3 * i + 1] + ay;
        az = acc[3 * i + 2] + az;
    }
    acc[3 * i] = ax;
    acc[3 * i + 1] = ay;
    acc[3 * i + 2] = az;
    if (vel) {  // leapfrog.py:14, v += a * dt_half
        vel[3 * i] = vel[3 * i] + ax * dt_half;
        vel[3 * i + 1] = vel[3 * i + 1] + ay * dt_half;
        vel[3 * i + 2] = vel[3 * i + 2] + az * dt_half;
    }
}

// MODE 0: FP64, reference order.  MODE 1: FP32.  MODE 2: FP32 on hi/lo
// positions with chunked two-sum compensation.  MODE 3: FP32C with the pair
// difference formed in FP64.
//
// blockIdx.y is the j-slice [y*jlen, min(n, (y+1)*jlen)).  With one slice
// (part == nullptr) the block finishes its particles itself; otherwise it
// writes its slice's sum as FP64 to part[(y*3 + axis)*n + i] and
// k_direct_reduce adds the slices in order.  FP64 always runs one slice, so
// its sum keeps the sequential j order of gravity.py:21.
//
// HILO (MODE 2 only): pos_lo holds the FP32 offsets of hi/lo positions, X =
// pos + pos_lo, split by split_df_off.  A template flag, so that the plain
// kernels compile as they do without it.
template <int MODE, int IPT, bool HILO = false>
__global__ __launch_bounds__(kBS) void k_direct(const double *__restrict__ pos,
                                                const float *__restrict__ pos_lo,
                                                const double *__restrict__ mass,
                                                double *__restrict__ acc,
                                                double *__restrict__ vel, int n, double G,
                                                double eps2, double dt_half, Epilogue epi,
                                                int jlen, double *__restrict__ part)
{
    using T4 = typename std::conditional<MODE == 0 || MODE == 3, double4, float4>::type;
    __shared__ T4 s_hi[kBS];
    __shared__ float4 s_lo[MODE == 2 ? kBS : 1];

    const int tid = threadIdx.x;
    const int base = blockIdx.x * (kBS * IPT) + tid;

    double xd[IPT], yd[IPT], zd[IPT];             // MODE 0 state
    double axd[IPT], ayd[IPT], azd[IPT];
    float xh[IPT], yh[IPT], zh[IPT];              // MODE 1/2 state
    float xl[IPT], yl[IPT], zl[IPT];
    float sx[IPT], sy[IPT], sz[IPT];              // running sums
    float cx[IPT], cy[IPT], cz[IPT];              // compensation terms

#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int i = base + k * kBS;
        const bool ok = i < n;
        const double px = ok ? pos[3 * i] : 0.0;
        const double py = ok ? pos[3 * i + 1] : 0.0;
        const double pz = ok ? pos[3 * i + 2] : 0.0;
        xd[k] = px; yd[k] = py; zd[k] = pz;
        axd[k] = ayd[k] = azd[k] = 0.0;
        if constexpr (HILO) {
            split_df_off(px, ok ? pos_lo[3 * i] : 0.f, xh[k], xl[k]);
            split_df_off(py, ok ? pos_lo[3 * i + 1] : 0.f, yh[k], yl[k]);
            split_df_off(pz, ok ? pos_lo[3 * i + 2] : 0.f, zh[k], zl[k]);
        } else {
            split_df(px, xh[k], xl[k]);
            split_df(py, yh[k], yl[k]);
            split_df(pz, zh[k], zl[k]);
        }
        sx[k] = sy[k] = sz[k] = 0.f;
        cx[k] = cy[k] = cz[k] = 0.f;
    }
    const float eps2f = static_cast<float>(eps2);
    const int j0 = blockIdx.y * jlen;
    const int j1 = min(n, j0 + jlen);

    for (int t0 = j0; t0 < j1; t0 += kBS) {
        const int j = t0 + tid;
        if (j < j1) {
            const double qx = pos[3 * j], qy = pos[3 * j + 1], qz = pos[3 * j + 2];
            const double qm = mass[j];
            if constexpr (MODE == 0) {
                s_hi[tid] = make_double4(qx, qy, qz, qm);
            } else if constexpr (MODE == 3) {
                s_hi[tid] = make_double4(qx, qy, qz, G * qm);
            } else if constexpr (MODE == 1) {
                s_hi[tid] = make_float4(static_cast<float>(qx), static_cast<float>(qy),
                                        static_cast<float>(qz), static_cast<float>(G * qm));
            } else {
                float hx, lx, hy, ly, hz, lz;
                if constexpr (HILO) {
                    split_df_off(qx, pos_lo[3 * j], hx, lx);
                    split_df_off(qy, pos_lo[3 * j + 1], hy, ly);
                    split_df_off(qz, pos_lo[3 * j + 2], hz, lz);
                } else {
                    split_df(qx, hx, lx);
                    split_df(qy, hy, ly);
                    split_df(qz, hz, lz);
                }
                s_hi[tid] = make_float4(hx, hy, hz, static_cast<float>(G * qm));
                s_lo[tid] = make_float4(lx, ly, lz, 0.f);
            }
        }
        __syncthreads();
        const int cnt = min(kBS, j1 - t0);
        if constexpr (MODE == 0) {
            for (int jj = 0; jj < cnt; ++jj) {
                const double4 q = s_hi[jj];
#pragma unroll
                for (int k = 0; k < IPT; ++k) {
                    // gravity.py:22-25, same association, IEEE sqrt and divide.
                    const double dx = q.x - xd[k];
                    const double dy = q.y - yd[k];
                    const double dz = q.z - zd[k];
                    const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                    const double inv = 1.0 / sqrt(r2);
                    const double s = G * q.w * inv * inv * inv;
                    axd[k] += s * dx;
                    ayd[k] += s * dy;
                    azd[k] += s * dz;
                }
            }
        } else if constexpr (MODE == 1) {
#pragma unroll 4
            for (int jj = 0; jj < cnt; ++jj) {
                const float4 q = s_hi[jj];
#pragma unroll
                for (int k = 0; k < IPT; ++k) {
                    const float dx = q.x - xh[k];
                    const float dy = q.y - yh[k];
                    const float dz = q.z - zh[k];
                    const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2f)));
                    const float inv = __builtin_amdgcn_rsqf(r2);
                    const float s = q.w * (inv * inv) * inv;
                    sx[k] = fmaf(s, dx, sx[k]);
                    sy[k] = fmaf(s, dy, sy[k]);
                    sz[k] = fmaf(s, dz, sz[k]);
                }
            }
        } else if constexpr (MODE == 3) {
            for (int c0 = 0; c0 < cnt; c0 += kChunk) {
                const int c1 = min(c0 + kChunk, cnt);
                float px[IPT], py[IPT], pz[IPT];
#pragma unroll
                for (int k = 0; k < IPT; ++k) px[k] = py[k] = pz[k] = 0.f;
#pragma unroll 4
                for (int jj = c0; jj < c1; ++jj) {
                    const double4 q = s_hi[jj];
                    const float gm = static_cast<float>(q.w);
#pragma unroll
                    for (int k = 0; k < IPT; ++k) {
                        // FP64 difference (full rate), rounded once to FP32
                        const float dx = static_cast<float>(q.x - xd[k]);
                        const float dy = static_cast<float>(q.y - yd[k]);
                        const float dz = static_cast<float>(q.z - zd[k]);
                        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2f)));
                        const float inv = __builtin_amdgcn_rsqf(r2);
                        const float s = gm * (inv * inv) * inv;
                        px[k] = fmaf(s, dx, px[k]);
                        py[k] = fmaf(s, dy, py[k]);
                        pz[k] = fmaf(s, dz, pz[k]);
                    }
                }
#pragma unroll
                for (int k = 0; k < IPT; ++k) {
                    float e;
                    two_sum(sx[k], px[k], sx[k], e);
                    cx[k] += e;
                    two_sum(sy[k], py[k], sy[k], e);
                    cy[k] += e;
                    two_sum(sz[k], pz[k], sz[k], e);
                    cz[k] += e;
                }
            }
        } else {
            for (int c0 = 0; c0 < cnt; c0 += kChunk) {
                const int c1 = min(c0 + kChunk, cnt);
                float px[IPT], py[IPT], pz[IPT];
#pragma unroll
                for (int k = 0; k < IPT; ++k) px[k] = py[k] = pz[k] = 0.f;
#pragma unroll 4
                for (int jj = c0; jj < c1; ++jj) {
                    const float4 q = s_hi[jj];
                    const float4 l = s_lo[jj];
#pragma unroll
                    for (int k = 0; k < IPT; ++k) {
                        // hi difference is exact for nearby pairs (Sterbenz);
                        // the lo difference restores the FP64 separation.
                        const float dx = (q.x - xh[k]) + (l.x - xl[k]);
                        const float dy = (q.y - yh[k]) + (l.y - yl[k]);
                        const float dz = (q.z - zh[k]) + (l.z - zl[k]);
                        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2f)));
                        const float inv = __builtin_amdgcn_rsqf(r2);
                        const float s = q.w * (inv * inv) * inv;
                        px[k] = fmaf(s, dx, px[k]);
                        py[k] = fmaf(s, dy, py[k]);
                        pz[k] = fmaf(s, dz, pz[k]);
                    }
                }
#pragma unroll
                for (int k = 0; k < IPT; ++k) {
                    float e;
                    two_sum(sx[k], px[k], sx[k], e);
                    cx[k] += e;
                    two_sum(sy[k], py[k], sy[k], e);
                    cy[k] += e;
                    two_sum(sz[k], pz[k], sz[k], e);
                    cz[k] += e;
                }
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int i = base + k * kBS;
        if (i >= n) continue;
        double ax, ay, az;
        if constexpr (MODE == 0) {
            ax = axd[k]; ay = ayd[k]; az = azd[k];
        } else {
            ax = static_cast<double>(sx[k]) + static_cast<double>(cx[k]);
            ay = static_cast<double>(sy[k]) + static_cast<double>(cy[k]);
            az = static_cast<double>(sz[k]) + static_cast<double>(cz[k]);
        }
        if (part) {  // uniform across the grid
            double *p = part + static_cast<size_t>(blockIdx.y) * 3 * n;
            p[i] = ax;
            p[n + i] = ay;
            p[2 * n + i] = az;
        } else {
            finish_particle(i, ax, ay, az, pos, acc, vel, G, dt_half, epi);
        }
    }
}

// FP64 in the reference's j order with W waves per 64 i-particles.  The
// sum a_i = (((0 + t_i0) + t_i1) + ...) must be formed in j order to be the
// reference's double, so it cannot be split over j; but the pair terms
// t_ij = (((G m_j) inv) inv) inv * d_ij (gravity.py:22-25) are independent.
// Each round, wave w computes the terms of S consecutive j (w*S .. w*S+S-1
// of the round's R = W*S) for the block's 64 particles (lane = i) and
// writes them to LDS; after the round's barrier, waves 0, 1 and 2 add the
// round's R terms of axis x, y and z, in j order, to their running sums
// while every wave computes the next round's terms into the other LDS
// buffer.  One barrier per round keeps the buffers apart: a buffer is
// rewritten two rounds later, after the barrier that the adding waves reach
// only when they are done reading it.  The j-particle data are uniform
// across the wave (scalar loads).  Same doubles as k_direct<0>.
template <int W, int S>
__global__ __launch_bounds__(64 * W) void k_direct64_pipe(const double *__restrict__ pos,
                                                         const double *__restrict__ mass,
                                                         double *__restrict__ acc,
                                                         double *__restrict__ vel, int n,
                                                         double G, double eps2, double dt_half,
                                                         Epilogue epi)
{
    constexpr int R = W * S;
    __shared__ double s_term[2][R][3][64];
    __shared__ double s_fin[3][64];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 64 + lane;
    const bool ok = i < n;
    const double xd = ok ? pos[3 * i] : 0.0;
    const double yd = ok ? pos[3 * i + 1] : 0.0;
    const double zd = ok ? pos[3 * i + 2] : 0.0;
    double sum = 0.0;  // waves 0..2: axis `wave` of particle i
    int r = 0;
    for (int r0 = 0; r0 < n; r0 += R, ++r) {
        const int buf = r & 1;
#pragma unroll
        for (int q = 0; q < S; ++q) {
            const int jl = wave * S + q;
            const int j = r0 + jl;
            if (j < n) {  // uniform
                const double qx = pos[3 * j], qy = pos[3 * j + 1], qz = pos[3 * j + 2];
                const double qm = mass[j];
                // gravity.py:22-25, same association, IEEE sqrt and divide.
                const double dx = qx - xd;
                const double dy = qy - yd;
                const double dz = qz - zd;
                const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                const double inv = 1.0 / sqrt(r2);
                const double sj = G * qm * inv * inv * inv;
                s_term[buf][jl][0][lane] = sj * dx;
                s_term[buf][jl][1][lane] = sj * dy;
                s_term[buf][jl][2][lane] = sj * dz;
            }
        }
        __syncthreads();
        if (wave < 3) {  // uniform per wave
            const int cnt = min(R, n - r0);
            const double *t = &s_term[buf][0][wave][lane];
            for (int q = 0; q < cnt; ++q) sum = sum + t[q * 3 * 64];
        }
    }
    if (wave < 3) s_fin[wave][lane] = sum;
    __syncthreads();
    if (wave == 0 && ok)
        finish_particle(i, s_fin[0][lane], s_fin[1][lane], s_fin[2][lane], pos, acc, vel, G,
                        dt_half, epi);
}

// Adds the ns slice sums of each particle in slice order (FP64), then
// finishes it as the single-slice kernel does.
__global__ __launch_bounds__(256) void k_direct_reduce(const double *__restrict__ part, int ns,
                                                       int n, const double *__restrict__ pos,
                                                       double *__restrict__ acc,
                                                       double *__restrict__ vel, double G,
                                                       double dt_half, Epilogue epi)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double ax = part[i], ay = part[n + i], az = part[2 * n + i];
        for (int s = 1; s < ns; ++s) {
            const double *p = part + static_cast<size_t>(s) * 3 * n;
            ax = ax + p[i];
            ay = ay + p[n + i];
            az = az + p[2 * n + i];
        }
        finish_particle(i, ax, ay, az, pos, acc, vel, G, dt_half, epi);
    }
}

// ---------------------------------------------- symmetric FP32/FP32C sums
// Each unordered pair is evaluated once (Newton's third law): i-block B
// (1024 particles, four per lane) against the j-tiles beyond it adds the
// pair term to i (i-side, in registers, compensated as in k_direct) and
// its opposite, with m_i for m_j, to j (j-side).  The j-side needs no
// cross-lane reduction: in step s of a group of 64 j, lane l takes j = (l +
// s) mod 64, so every j visits each lane once per group, and the lane adds
// its four-term j-side partial to the wave's FP64 accumulator of that j in
// LDS (a non-returning ds_add_f64; within one instruction the 64 lanes hold
// 64 different j, and one wave's instructions reach LDS in order, so there
// is no race and the order is fixed).  After a tile the four waves'
// accumulators are added in wave order and written to jpart[B][axis][j -
// 1024 (B + 1)].  A block's own four tiles (the diagonal) are summed
// one-sided.  k_direct_sym_reduce adds, per particle, its i-side partials in
// segment order and then the j-side partials of the earlier i-blocks in
// block order, all in FP64: the result is deterministic for a given n.  A
// pair costs about 23 VALU instructions (one rsqrt) against 19 in k_direct,
// for half the pairs.
constexpr int kSymIpt = 4;
constexpr int kSymIB = kBS * kSymIpt;       // i-particles per block
constexpr int kSymTilesIB = kSymIB / kBS;   // their tiles (4)
constexpr int kSymSeg = 16;                 // j-tiles per segment of a block row

// Segments of block row b: the diagonal, then ceil(beyond / kSymSeg).
__host__ __device__ __forceinline__ int sym_nseg(int b, int nt)
{
    const int beyond = nt - kSymTilesIB * (b + 1);
    return 1 + (beyond > 0 ? (beyond + kSymSeg - 1) / kSymSeg : 0);
}
// First j-side entry (doubles) of block row b: rows hold 3 x (n - 1024 (b + 1)).
__host__ __device__ __forceinline__ int64_t sym_jrow(int b, int n)
{
    return 3 * (static_cast<int64_t>(b) * n - static_cast<int64_t>(kSymIB) / 2 * b * (b + 1));
}

// Lane l gets lane (l + 1) mod 64's value (DPP wave_rol:1).
__device__ __forceinline__ float wave_rol1(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x134, 0xF, 0xF,
                                                      false));
}

// JROT: the j-side sums rotate through the lanes in registers (FP32, one
// DPP rotation per step fused into the add) instead of FP64 LDS adds.
template <int MODE, bool HILO, bool JROT = false>
__global__ __launch_bounds__(kBS, 4) void k_direct_sym(const double *__restrict__ pos,
                                                    const float *__restrict__ pos_lo,
                                                    const double *__restrict__ mass, int n,
                                                    double G, double eps2,
                                                    double *__restrict__ ipart,
                                                    double *__restrict__ jpart)
{
    static_assert(MODE == 1 || MODE == 2, "FP32 and FP32C only");
    constexpr int IPT = kSymIpt;
    constexpr int kW = kBS / 64;
    __shared__ float4 s_hi[kBS];
    __shared__ float4 s_lo[MODE == 2 ? kBS : 1];
    __shared__ double s_ja[kW][3][kBS];  // per wave: the j-side sums of the tile
    const int B = blockIdx.y;
    const int seg = blockIdx.x;
    const int nt = (n + kBS - 1) / kBS;
    const int nseg = sym_nseg(B, nt);
    if (seg >= nseg) return;  // uniform: rows beyond the diagonal are shorter
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const int base = B * kSymIB + tid;
    float xh[IPT], yh[IPT], zh[IPT], xl[IPT], yl[IPT], zl[IPT], gmi[IPT];
    float sx[IPT], sy[IPT], sz[IPT], cx[IPT], cy[IPT], cz[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int i = base + k * kBS;
        const bool ok = i < n;
        const double px = ok ? pos[3 * i] : 0.0;
        const double py = ok ? pos[3 * i + 1] : 0.0;
        const double pz = ok ? pos[3 * i + 2] : 0.0;
        if constexpr (HILO) {
            split_df_off(px, ok ? pos_lo[3 * i] : 0.f, xh[k], xl[k]);
            split_df_off(py, ok ? pos_lo[3 * i + 1] : 0.f, yh[k], yl[k]);
            split_df_off(pz, ok ? pos_lo[3 * i + 2] : 0.f, zh[k], zl[k]);
        } else {
            split_df(px, xh[k], xl[k]);
            split_df(py, yh[k], yl[k]);
            split_df(pz, zh[k], zl[k]);
        }
        if (MODE == 1) {  // plain FP32: the rounded position
            xl[k] = yl[k] = zl[k] = 0.f;
        }
        gmi[k] = ok ? static_cast<float>(G * mass[i]) : 0.f;  // no j-side term from padding
        sx[k] = sy[k] = sz[k] = 0.f;
        cx[k] = cy[k] = cz[k] = 0.f;
    }
    const float eps2f = static_cast<float>(eps2);
    const bool diag = seg == 0;
    const int tb = kSymTilesIB * (B + 1);  // first tile beyond the block
    const int tlo = diag ? kSymTilesIB * B : tb + (seg - 1) * kSymSeg;
    const int thi = diag ? min(tb, nt) : min(nt, tb + seg * kSymSeg);
    const int jrow0 = kSymIB * (B + 1);
    double *jp = jpart + sym_jrow(B, n);
    const int jlen = n - jrow0;

    for (int t = tlo; t < thi; ++t) {
        const int j = t * kBS + tid;
        if (j < n) {
            const double qx = pos[3 * j], qy = pos[3 * j + 1], qz = pos[3 * j + 2];
            const float gm = static_cast<float>(G * mass[j]);
            float hx, lx, hy, ly, hz, lz;
            if constexpr (HILO) {
                split_df_off(qx, pos_lo[3 * j], hx, lx);
                split_df_off(qy, pos_lo[3 * j + 1], hy, ly);
                split_df_off(qz, pos_lo[3 * j + 2], hz, lz);
            } else {
                split_df(qx, hx, lx);
                split_df(qy, hy, ly);
                split_df(qz, hz, lz);
            }
            s_hi[tid] = make_float4(hx, hy, hz, gm);
            if constexpr (MODE == 2) s_lo[tid] = make_float4(lx, ly, lz, 0.f);
        } else {
            // past n: a massless point, so the i-side terms are zero (its
            // j-side sums are never stored)
            s_hi[tid] = make_float4(0.f, 0.f, 0.f, 0.f);
            if constexpr (MODE == 2) s_lo[tid] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (!diag) {
#pragma unroll
            for (int w = 0; w < kW; ++w) {
                s_ja[w][0][tid] = 0.0;
                s_ja[w][1][tid] = 0.0;
                s_ja[w][2][tid] = 0.0;
            }
        }
        __syncthreads();
        const int cnt = min(kBS, n - t * kBS);
        if (diag) {  // uniform: one-sided, as k_direct
            for (int c0 = 0; c0 < cnt; c0 += kChunk) {
                const int c1 = min(c0 + kChunk, cnt);
                float px[IPT], py[IPT], pz[IPT];
#pragma unroll
                for (int k = 0; k < IPT; ++k) px[k] = py[k] = pz[k] = 0.f;
#pragma unroll 4
                for (int jj = c0; jj < c1; ++jj) {
                    const float4 q = s_hi[jj];
                    float4 l = make_float4(0.f, 0.f, 0.f, 0.f);
                    if constexpr (MODE == 2) l = s_lo[jj];
#pragma unroll
                    for (int k = 0; k < IPT; ++k) {
                        const float dx = MODE == 2 ? (q.x - xh[k]) + (l.x - xl[k]) : q.x - xh[k];
                        const float dy = MODE == 2 ? (q.y - yh[k]) + (l.y - yl[k]) : q.y - yh[k];
                        const float dz = MODE == 2 ? (q.z - zh[k]) + (l.z - zl[k]) : q.z - zh[k];
                        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2f)));
                        const float inv = __builtin_amdgcn_rsqf(r2);
                        const float sc = q.w * (inv * inv) * inv;
                        px[k] = fmaf(sc, dx, px[k]);
                        py[k] = fmaf(sc, dy, py[k]);
                        pz[k] = fmaf(sc, dz, pz[k]);
                    }
                }
#pragma unroll
                for (int k = 0; k < IPT; ++k) {
                    float e;
                    two_sum(sx[k], px[k], sx[k], e);
                    cx[k] += e;
                    two_sum(sy[k], py[k], sy[k], e);
                    cy[k] += e;
                    two_sum(sz[k], pz[k], sz[k], e);
                    cz[k] += e;
                }
            }
        } else {
            // groups of 64 j; chunks of 32 steps for the i-side compensation
            for (int g0 = 0; g0 < cnt; g0 += 64) {
                // JROT: R* hold, at step st, lane l's running sum for j = (l + st)
                // mod 64 of the group: R = rol(R) + this step's term, flushed to
                // FP64 at the end of each 32-step chunk
                float Rx = 0.f, Ry = 0.f, Rz = 0.f;
                for (int c0 = 0; c0 < 64; c0 += kChunk) {
                    float px[IPT], py[IPT], pz[IPT];
#pragma unroll
                    for (int k = 0; k < IPT; ++k) px[k] = py[k] = pz[k] = 0.f;
#pragma unroll 2
                    for (int st = c0; st < c0 + kChunk; ++st) {
                        const int jl = g0 + ((lane + st) & 63);
                        const float4 q = s_hi[jl];
                        float4 l = make_float4(0.f, 0.f, 0.f, 0.f);
                        if constexpr (MODE == 2) l = s_lo[jl];
                        float ajx = 0.f, ajy = 0.f, ajz = 0.f;
#pragma unroll
                        for (int k = 0; k < IPT; ++k) {
                            const float dx =
                                MODE == 2 ? (q.x - xh[k]) + (l.x - xl[k]) : q.x - xh[k];
                            const float dy =
                                MODE == 2 ? (q.y - yh[k]) + (l.y - yl[k]) : q.y - yh[k];
                            const float dz =
                                MODE == 2 ? (q.z - zh[k]) + (l.z - zl[k]) : q.z - zh[k];
                            const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2f)));
                            const float inv = __builtin_amdgcn_rsqf(r2);
                            const float inv3 = (inv * inv) * inv;
                            const float si = q.w * inv3;
                            const float sj = gmi[k] * inv3;
                            px[k] = fmaf(si, dx, px[k]);
                            py[k] = fmaf(si, dy, py[k]);
                            pz[k] = fmaf(si, dz, pz[k]);
                            ajx = fmaf(sj, dx, ajx);
                            ajy = fmaf(sj, dy, ajy);
                            ajz = fmaf(sj, dz, ajz);
                        }
                        if constexpr (JROT) {  // at st = 0 the rotated sums are zeros
                            Rx = wave_rol1(Rx) + ajx;
                            Ry = wave_rol1(Ry) + ajy;
                            Rz = wave_rol1(Rz) + ajz;
                        } else {
                            __hip_atomic_fetch_add(&s_ja[wave][0][jl], static_cast<double>(ajx),
                                                   __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WAVEFRONT);
                            __hip_atomic_fetch_add(&s_ja[wave][1][jl], static_cast<double>(ajy),
                                                   __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WAVEFRONT);
                            __hip_atomic_fetch_add(&s_ja[wave][2][jl], static_cast<double>(ajz),
                                                   __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WAVEFRONT);
                        }
                    }
#pragma unroll
                    for (int k = 0; k < IPT; ++k) {
                        float e;
                        two_sum(sx[k], px[k], sx[k], e);
                        cx[k] += e;
                        two_sum(sy[k], py[k], sy[k], e);
                        cy[k] += e;
                        two_sum(sz[k], pz[k], sz[k], e);
                        cz[k] += e;
                    }
                    if constexpr (JROT) {
                        // end of the chunk: lane l's sum (32 terms at most) belongs
                        // to j = (l + c0 + 31) mod 64; add it to that j's FP64
                        // accumulator and start again from zero
                        const int jf = g0 + ((lane + c0 + kChunk - 1) & 63);
                        __hip_atomic_fetch_add(&s_ja[wave][0][jf], static_cast<double>(Rx),
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                        __hip_atomic_fetch_add(&s_ja[wave][1][jf], static_cast<double>(Ry),
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                        __hip_atomic_fetch_add(&s_ja[wave][2][jf], static_cast<double>(Rz),
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                        Rx = Ry = Rz = 0.f;
                    }
                }
            }
        }
        __syncthreads();
        if (!diag && j < n) {  // j-side of j = t * 256 + tid: the waves' sums in wave order
            double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
            for (int w = 0; w < kW; ++w) {
                ax += s_ja[w][0][tid];
                ay += s_ja[w][1][tid];
                az += s_ja[w][2][tid];
            }
            const int jj = j - jrow0;
            jp[jj] = -ax;  // the force on j is minus the sum of sj d
            jp[jlen + jj] = -ay;
            jp[2 * static_cast<int64_t>(jlen) + jj] = -az;
        }
        __syncthreads();  // s_hi, s_lo and s_ja are rewritten by the next tile
    }
    // i-side partial of this segment
    int64_t segoff = 0;
    for (int b = 0; b < B; ++b) segoff += sym_nseg(b, nt);
    double *ip = ipart + (segoff + seg) * 3 * kSymIB;
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int li = tid + k * kBS;
        ip[li] = static_cast<double>(sx[k]) + static_cast<double>(cx[k]);
        ip[kSymIB + li] = static_cast<double>(sy[k]) + static_cast<double>(cy[k]);
        ip[2 * kSymIB + li] = static_cast<double>(sz[k]) + static_cast<double>(cz[k]);
    }
}

// Per particle: its block row's i-side partials in segment order, then the
// j-side partials of every earlier block row in row order (FP64), then the
// epilogue.
__global__ __launch_bounds__(256) void k_direct_sym_reduce(const double *__restrict__ ipart,
                                                           const double *__restrict__ jpart,
                                                           int n, const double *__restrict__ pos,
                                                           double *__restrict__ acc,
                                                           double *__restrict__ vel, double G,
                                                           double dt_half, Epilogue epi)
{
    const int nt = (n + kBS - 1) / kBS;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int B = i / kSymIB;
        const int li = i - B * kSymIB;
        int64_t segoff = 0;
        for (int b = 0; b < B; ++b) segoff += sym_nseg(b, nt);
        const int ns = sym_nseg(B, nt);
        const double *ip = ipart + segoff * 3 * kSymIB;
        double ax = ip[li], ay = ip[kSymIB + li], az = ip[2 * kSymIB + li];
        for (int s = 1; s < ns; ++s) {
            const double *q = ip + static_cast<int64_t>(s) * 3 * kSymIB;
            ax = ax + q[li];
            ay = ay + q[kSymIB + li];
            az = az + q[2 * kSymIB + li];
        }
        for (int b = 0; b < B; ++b) {
            const int64_t len = n - static_cast<int64_t>(kSymIB) * (b + 1);
            const double *q = jpart + sym_jrow(b, n) + (i - kSymIB * (b + 1));
            ax = ax + q[0];
            ay = ay + q[len];
            az = az + q[2 * len];
        }
        finish_particle(i, ax, ay, az, pos, acc, vel, G, dt_half, epi);
    }
}

// Scratch (doubles) of the symmetric sum: the i-side partials of every
// segment, then the j-side rows.
int64_t sym_scratch_doubles(int n)
{
    const int nt = (n + kBS - 1) / kBS;
    const int nb = (n + kSymIB - 1) / kSymIB;
    int64_t segs = 0;
    for (int b = 0; b < nb; ++b) segs += sym_nseg(b, nt);
    return segs * 3 * kSymIB + sym_jrow(nb > 0 ? nb - 1 : 0, n);
}

// The symmetric sum is used for FP32/FP32C from n = kSymFrom (four
// i-particles per lane there anyway) while its scratch fits kSymMaxBytes;
// "direct_sym" forces it on (1) or off (-1), and any explicit "direct_split"
// or "direct_ipt" selects the j-split kernel.
constexpr int kSymFrom = 131072;
constexpr int64_t kSymMaxBytes = int64_t(3) << 30;
int g_direct_sym = 0;
int g_sym_jrot = 0;  // j-side sums: 0 FP64 LDS adds, 1 rotating FP32 registers

bool sym_wanted(int mode, int n)
{
    if (mode != 1 && mode != 2) return false;
    if (g_direct_sym < 0) return false;
    if (g_direct_sym == 0 && (n < kSymFrom || g_direct_split || g_direct_ipt)) return false;
    return n >= 2 && 8 * sym_scratch_doubles(n) <= kSymMaxBytes;
}

template <int MODE>
int launch_direct_sym(const double *pos, const float *pos_lo, const double *mass, double *acc,
                      double *vel, int n, double G, double eps2, double dt_half,
                      const Epilogue &epi, hipStream_t st, const ScratchSrc &ss)
{
    const size_t need = 8 * static_cast<size_t>(sym_scratch_doubles(n));
    const int nt = (n + kBS - 1) / kBS;
    const int nb = (n + kSymIB - 1) / kSymIB;
    const int maxseg = sym_nseg(0, nt);
    const int64_t segs = sym_scratch_doubles(n) >= 0 ? [&] {
        int64_t c = 0;
        for (int b = 0; b < nb; ++b) c += sym_nseg(b, nt);
        return c;
    }() : 0;
    void *buf = nullptr;
    if (ss.ptr) {
        RG_REQUIRE(need <= ss.cap, RG_ERR_CAPACITY,
                   "n=%d: the symmetric sum needs %zu MiB of partials; the static scratch "
                   "holds %zu MiB", n, (need + (1u << 20) - 1) >> 20, ss.cap >> 20);
        buf = ss.ptr;
    } else {
        const int rc = rg::device_scratch(need, st, &buf);
        if (rc != RG_OK) return rc;
    }
    double *ipart = static_cast<double *>(buf);
    double *jpart = ipart + segs * 3 * kSymIB;
    const dim3 grid(maxseg, nb);
    if (g_sym_jrot) {
        if (MODE == 2 && pos_lo)
            k_direct_sym<MODE, true, true><<<grid, kBS, 0, st>>>(pos, pos_lo, mass, n, G, eps2,
                                                                 ipart, jpart);
        else
            k_direct_sym<MODE, false, true><<<grid, kBS, 0, st>>>(pos, nullptr, mass, n, G, eps2,
                                                                  ipart, jpart);
    } else if (MODE == 2 && pos_lo) {
        k_direct_sym<MODE, true><<<grid, kBS, 0, st>>>(pos, pos_lo, mass, n, G, eps2, ipart,
                                                       jpart);
    } else {
        k_direct_sym<MODE, false><<<grid, kBS, 0, st>>>(pos, nullptr, mass, n, G, eps2, ipart,
                                                        jpart);
    }
    RG_CHECK_LAUNCH();
    k_direct_sym_reduce<<<elementwise_grid(n), 256, 0, st>>>(ipart, jpart, n, pos, acc, vel, G,
                                                             dt_half, epi);
    RG_CHECK_LAUNCH();
    return RG_OK;
}

template <int MODE>
int launch_direct(const double *pos, const float *pos_lo, const double *mass, double *acc,
                  double *vel, int n, double G, double eps2, double dt_half,
                  const Epilogue &epi, hipStream_t st, const ScratchSrc &ss)
{
    if constexpr (MODE == 1 || MODE == 2) {
        if (sym_wanted(MODE, n))
            return launch_direct_sym<MODE>(pos, pos_lo, mass, acc, vel, n, G, eps2, dt_half, epi,
                                           st, ss);
    }
    int ipt = g_direct_ipt;
    if (ipt == 0) {
        ipt = 1;
        if ((MODE == 2 || MODE == 3) && n >= kIpt4From) ipt = 4;
        if (MODE == 1 && n >= kIpt2From) ipt = 2;
    }
    const int per_block = kBS * ipt;
    const int ib = (n + per_block - 1) / per_block;
    int jlen = n;
    const int ns = MODE == 0 ? 1 : rg::direct_slices(n, ipt, &jlen);
    double *part = nullptr;
    if (ns > 1) {
        const size_t need = sizeof(double) * 3 * static_cast<size_t>(ns) * n;
        void *buf = nullptr;
        if (ss.ptr) {
            RG_REQUIRE(need <= ss.cap, RG_ERR_CAPACITY,
                       "n=%d with %d j-slices needs %zu MiB of slice partials; the static "
                       "scratch holds %zu MiB (lower \"direct_split\" or use rg_direct_acc)",
                       n, ns, (need + (1u << 20) - 1) >> 20, ss.cap >> 20);
            buf = ss.ptr;
        } else {
            const int rc = rg::device_scratch(need, st, &buf);
            if (rc != RG_OK) return rc;
        }
        part = static_cast<double *>(buf);
    }
    const dim3 grid(ib, ns);
    constexpr bool kCanHilo = MODE == 2;
#define RG_DIRECT_LAUNCH(IPT)                                                                  \
    do {                                                                                       \
        if (kCanHilo && pos_lo)                                                                \
            k_direct<MODE, IPT, kCanHilo><<<grid, kBS, 0, st>>>(pos, pos_lo, mass, acc, vel, n, \
                                                                G, eps2, dt_half, epi, jlen,   \
                                                                part);                         \
        else                                                                                   \
            k_direct<MODE, IPT><<<grid, kBS, 0, st>>>(pos, nullptr, mass, acc, vel, n, G, eps2, \
                                                      dt_half, epi, jlen, part);               \
    } while (0)
    switch (ipt) {
    case 1:
        RG_DIRECT_LAUNCH(1);
        break;
    case 2:
        RG_DIRECT_LAUNCH(2);
        break;
    case 4:
        RG_DIRECT_LAUNCH(4);
        break;
    default:
        rg::set_error("direct_ipt must be 1, 2 or 4 (got %d)", ipt);
        return RG_ERR_ARG;
    }
#undef RG_DIRECT_LAUNCH
    RG_CHECK_LAUNCH();
    if (part) {
        k_direct_reduce<<<elementwise_grid(n), 256, 0, st>>>(part, ns, n, pos, acc, vel, G,
                                                             dt_half, epi);
        RG_CHECK_LAUNCH();
    }
    return RG_OK;
}

// SIMDs of the current device (4 per CU), cached per device; threads that
// race on the first query store the same value.
int device_simds()
{
    static std::atomic<int> simds[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
    int v = simds[dev].load(std::memory_order_relaxed);
    if (v == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess || cus <= 0)
            return 1024;
        v = 4 * cus;
        simds[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}

// Waves per 64 i-particles of the FP64 pipeline (1: the one-wave kernel).
int fp64_pipe_waves(int n)
{
    if (g_fp64_pipe) return g_fp64_pipe;
    if (g_direct_ipt) return 1;  // an explicit i-per-lane choice picks k_direct<0>
    if (n <= kPipe16Upto) return 16;
    if (n < kPipeBelow) return 8;
    // the one-wave kernel: waves on the busiest SIMD and how evenly the
    // others are filled
    const long long waves = (static_cast<long long>(n) + 63) / 64;
    const long long simds = device_simds();
    const long long k = (waves + simds - 1) / simds;
    const double fill = static_cast<double>(waves) / static_cast<double>(k * simds);
    if (k < kOneWaveMinPerSimd || (k < 4 && fill < kOneWaveMinFill)) return 8;
    return 1;
}

int launch_direct64_pipe(int w, const double *pos, const double *mass, double *acc, double *vel,
                         int n, double G, double eps2, double dt_half, const Epilogue &epi,
                         hipStream_t st)
{
    const int grid = (n + 63) / 64;
    switch (w) {
    case 4:
        k_direct64_pipe<4, 4><<<grid, 256, 0, st>>>(pos, mass, acc, vel, n, G, eps2, dt_half, epi);
        break;
    case 8:
        k_direct64_pipe<8, 2><<<grid, 512, 0, st>>>(pos, mass, acc, vel, n, G, eps2, dt_half, epi);
        break;
    case 16:
        k_direct64_pipe<16, 2><<<grid, 1024, 0, st>>>(pos, mass, acc, vel, n, G, eps2, dt_half,
                                                      epi);
        break;
    default:
        rg::set_error("fp64_pipe must be 0, 1, 4, 8 or 16 (got %d)", w);
        return RG_ERR_ARG;
    }
    RG_CHECK_LAUNCH();
    return RG_OK;
}

int direct_common(const double *pos, const double *mass, double *acc, double *vel,
                  int64_t n, double G, double eps2, int precision, double dt_half,
                  const Epilogue &epi, void *stream, const ScratchSrc &ss = {nullptr, 0},
                  const float *pos_lo = nullptr)
{
    RG_REQUIRE(n >= 0 && n <= (int64_t(1) << 29), RG_ERR_CAPACITY,
               "n=%lld outside [0, 2^29]", static_cast<long long>(n));
    if (n == 0) return RG_OK;
    RG_REQUIRE(pos && mass && acc, RG_ERR_ARG, "null pointer argument");
    RG_REQUIRE(eps2 > 0.0, RG_ERR_ARG, "eps2 must be > 0 (got %g)", eps2);
    const size_t b3 = sizeof(double) * 3 * static_cast<size_t>(n);
    const size_t b1 = sizeof(double) * static_cast<size_t>(n);
    RG_REQUIRE(!rg_overlap(acc, b3, pos, b3) && !rg_overlap(acc, b3, mass, b1), RG_ERR_ALIAS,
               "acc aliases pos or mass");
    if (vel)
        RG_REQUIRE(!rg_overlap(vel, b3, pos, b3) && !rg_overlap(vel, b3, mass, b1) &&
                       !rg_overlap(vel, b3, acc, b3),
                   RG_ERR_ALIAS, "vel aliases pos, mass or acc");
    if (pos_lo) {
        RG_REQUIRE(precision == RG_FP32C && !g_fp32c_d64, RG_ERR_ARG,
                   "hi/lo positions take the FP32C pair sum (with fp32c_d64 off)");
        const size_t bl = sizeof(float) * 3 * static_cast<size_t>(n);
        RG_REQUIRE(!rg_overlap(acc, b3, pos_lo, bl) && (!vel || !rg_overlap(vel, b3, pos_lo, bl)),
                   RG_ERR_ALIAS, "acc or vel aliases pos_lo");
    }
    hipStream_t st = rg::as_stream(stream);
    const int ni = static_cast<int>(n);
    switch (precision) {
    case RG_FP64: {
        const int w = fp64_pipe_waves(ni);
        if (w > 1) return launch_direct64_pipe(w, pos, mass, acc, vel, ni, G, eps2, dt_half, epi, st);
        return launch_direct<0>(pos, nullptr, mass, acc, vel, ni, G, eps2, dt_half, epi, st, ss);
    }
    case RG_FP32:
        return launch_direct<1>(pos, nullptr, mass, acc, vel, ni, G, eps2, dt_half, epi, st, ss);
    case RG_FP32C:
        if (g_fp32c_d64)
            return launch_direct<3>(pos, nullptr, mass, acc, vel, ni, G, eps2, dt_half, epi, st,
                                    ss);
        return launch_direct<2>(pos, pos_lo, mass, acc, vel, ni, G, eps2, dt_half, epi, st, ss);
    default: break;
    }
    rg::set_error("unknown precision %d", precision);
    return RG_ERR_ARG;
}

// ---------------------------------------------------------------- leapfrog
__global__ void k_kick(double *__restrict__ v, const double *__restrict__ a, int m, double h)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x)
        v[k] = v[k] + a[k] * h;
}

__global__ void k_drift(double *__restrict__ x, const double *__restrict__ v, int m, double dt)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x)
        x[k] = x[k] + v[k] * dt;
}

__global__ void k_kick_drift(double *__restrict__ x, double *__restrict__ v,
                             const double *__restrict__ a, int m, double h, double dt)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        const double vk = v[k] + a[k] * h;
        v[k] = vk;
        x[k] = x[k] + vk * dt;
    }
}

// leapfrog.drift (leapfrog.py:18-22) on hi/lo positions: X = x + xl += v dt
// per component, compensated (drift_hilo).
__global__ void k_drift_hilo(double *__restrict__ x, float *__restrict__ xl,
                             const double *__restrict__ v, int m, double dt)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        double xh = x[k];
        float lo = xl[k];
        drift_hilo(xh, lo, v[k], dt);
        x[k] = xh;
        xl[k] = lo;
    }
}

// kick (v += a h, leapfrog.py:14) then the hi/lo drift with the new v.
__global__ void k_kick_drift_hilo(double *__restrict__ x, float *__restrict__ xl,
                                  double *__restrict__ v, const double *__restrict__ a, int m,
                                  double h, double dt)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        const double vk = v[k] + a[k] * h;
        v[k] = vk;
        double xh = x[k];
        float lo = xl[k];
        drift_hilo(xh, lo, vk, dt);
        x[k] = xh;
        xl[k] = lo;
    }
}

// k_kick_drift_hilo two components per thread (16-byte FP64 and 8-byte FP32
// accesses): pairs = m / 2 pairs of components; an odd last component is
// done by k_kick_drift_hilo on its own.
__global__ void k_kick_drift_hilo2(double2 *__restrict__ x, float2 *__restrict__ xl,
                                   double2 *__restrict__ v, const double2 *__restrict__ a,
                                   int pairs, double h, double dt)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < pairs;
         k += gridDim.x * blockDim.x) {
        double2 vk = v[k];
        const double2 ak = a[k];
        double2 xk = x[k];
        float2 lk = xl[k];
        vk.x = vk.x + ak.x * h;
        vk.y = vk.y + ak.y * h;
        v[k] = vk;
        drift_hilo(xk.x, lk.x, vk.x, dt);
        drift_hilo(xk.y, lk.y, vk.y, dt);
        x[k] = xk;
        xl[k] = lk;
    }
}

// living_galaxy_engine.py:116-120 add_acc / gas_disk_engine.py:84-89 add_field.
__global__ void k_add_field(double *__restrict__ acc, const double *__restrict__ src, int m)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x)
        acc[k] = acc[k] + src[k];
}

// living_galaxy_engine.py:123-150 grav_analytic / gas_disk_engine.py:54-82
// add_analytic: acc[i] += ((disk + halo) + point mass).
__global__ void k_analytic_add(const double *__restrict__ pos, double *__restrict__ acc, int n,
                               double G, Epilogue epi)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double av[3];
        ext_terms(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], G, epi, av, true);
        acc[3 * i] = acc[3 * i] + av[0];
        acc[3 * i + 1] = acc[3 * i + 1] + av[1];
        acc[3 * i + 2] = acc[3 * i + 2] + av[2];
    }
}

int elementwise_grid(int m)
{
    const int g = (m + 255) / 256;
    return g < 4096 ? (g > 0 ? g : 1) : 4096;
}

// --------------------------------------------------------------- energies
constexpr int kEB = 256;

// Block-level fixed-order tree reduction of 5 doubles per thread.
__device__ void block_reduce5(double v[5], double *out, int stride)
{
    __shared__ double red[5][kEB];
    const int t = threadIdx.x;
#pragma unroll
    for (int c = 0; c < 5; ++c) red[c][t] = v[c];
    __syncthreads();
    for (int w = kEB / 2; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int c = 0; c < 5; ++c) red[c][t] = red[c][t] + red[c][t + w];
        __syncthreads();
    }
    if (t == 0)
#pragma unroll
        for (int c = 0; c < 5; ++c) out[c * stride] = red[c][0];
}

// Per-particle KE (gravity.py:35), PE (gravity.py:47-56, j != i) and
// momentum; one partial row per block.  blockIdx.y is the j-slice
// [y*jlen, min(n, (y+1)*jlen)) of the PE sum (the slices give the N^2 pass
// enough blocks to fill the GPU); KE and momentum come from slice 0 only.
// Row (y, x) of the partials is column y*gridDim.x + x.
__global__ __launch_bounds__(kEB) void k_energy_partial(const double *__restrict__ pos,
                                                         const double *__restrict__ vel,
                                                         const double *__restrict__ mass,
                                                         int n, double G, double eps2,
                                                         int jlen, double *__restrict__ part)
{
    __shared__ double4 sj[kEB];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * kEB + tid;
    const bool ok = i < n;
    const bool with_pe = pos != nullptr;  // uniform across the grid
    const double px = ok && with_pe ? pos[3 * i] : 0.0;
    const double py = ok && with_pe ? pos[3 * i + 1] : 0.0;
    const double pz = ok && with_pe ? pos[3 * i + 2] : 0.0;
    const int j0 = blockIdx.y * jlen;
    const int j1 = min(n, j0 + jlen);
    double acc_i = 0.0;
    for (int t0 = j0; with_pe && t0 < j1; t0 += kEB) {
        const int j = t0 + tid;
        if (j < j1) sj[tid] = make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], mass[j]);
        __syncthreads();
        const int cnt = min(kEB, j1 - t0);
        for (int jj = 0; jj < cnt; ++jj) {
            const double4 q = sj[jj];
            if (t0 + jj != i) {
                const double dx = q.x - px;
                const double dy = q.y - py;
                const double dz = q.z - pz;
                const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                acc_i += q.w / sqrt(r2);
            }
        }
        __syncthreads();
    }
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (ok) {
        const double mi = mass[i];
        v[1] = -0.5 * G * mi * acc_i;
        if (blockIdx.y == 0) {
            const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
            v[0] = 0.5 * mi * (vx * vx + vy * vy + vz * vz);
            v[2] = mi * vx;
            v[3] = mi * vy;
            v[4] = mi * vz;
        }
    }
    const int cols = gridDim.x * gridDim.y;
    block_reduce5(v, part + blockIdx.y * gridDim.x + blockIdx.x, cols);
}

// Single block: thread t sums partials t, t+kEB, ... in order, then a tree.
__global__ __launch_bounds__(kEB) void k_energy_final(const double *__restrict__ part, int nb,
                                                       double *__restrict__ out5)
{
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += kEB)
#pragma unroll
        for (int c = 0; c < 5; ++c) v[c] = v[c] + part[c * nb + b];
    block_reduce5(v, out5, 1);
}

}  // namespace

int rg::direct_slices(int n, int ipt, int *jlen)
{
    int ns = g_direct_split;
    if (ns == 0)
        ns = static_cast<int>(
            std::min<int64_t>(ipt == 1 ? kSplitOne : kSplitMany, kMaxSliceParticles / n));
    ns = std::max(1, std::min(ns, (n + kBS - 1) / kBS));  // each slice holds a tile
    const int per = (n + ns - 1) / ns;
    const int len = (per + kBS - 1) / kBS * kBS;
    *jlen = len;
    return (n + len - 1) / len;
}

int rg::rg_direct_set_fp64_pipe_impl(int64_t v)
{
    RG_REQUIRE(v == 0 || v == 1 || v == 4 || v == 8 || v == 16, RG_ERR_ARG,
               "fp64_pipe must be 0, 1, 4, 8 or 16");
    g_fp64_pipe = static_cast<int>(v);
    return RG_OK;
}

int rg::rg_direct_set_split_impl(int64_t v)
{
    RG_REQUIRE(v >= 0 && v <= kMaxSplit, RG_ERR_ARG, "direct_split must be in [0, %d]", kMaxSplit);
    g_direct_split = static_cast<int>(v);
    return RG_OK;
}

int rg::rg_direct_set_sym_impl(int64_t v)
{
    RG_REQUIRE(v >= -1 && v <= 3, RG_ERR_ARG, "direct_sym must be -1, 0, 1, 2 or 3");
    // 2 and 3: as 0 and 1 with the j-side sums rotating through registers
    g_direct_sym = v == 2 ? 0 : (v == 3 ? 1 : static_cast<int>(v));
    g_sym_jrot = v >= 2 ? 1 : 0;
    return RG_OK;
}

int rg::rg_direct_set_ipt_impl(int64_t v)
{
    RG_REQUIRE(v == 0 || v == 1 || v == 2 || v == 4, RG_ERR_ARG, "direct_ipt must be 0, 1, 2 or 4");
    g_direct_ipt = static_cast<int>(v);
    return RG_OK;
}

extern "C" int rg_direct_acc(const double *pos, const double *mass, double *acc, int64_t n,
                             double G, double eps2, int precision, void *stream)
{
    return direct_common(pos, mass, acc, nullptr, n, G, eps2, precision, 0.0,
                         make_epilogue(kStore, nullptr), stream);
}

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxDevices];  // device address of `scratch`, per device

// The current device's copy of `scratch`, looked up once per device with
// that device current.
int noalloc_scratch(void **out)
{
    int dev = 0;
    RG_HIP(hipGetDevice(&dev));
    RG_REQUIRE(dev >= 0 && dev < kMaxDevices, RG_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        RG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        RG_REQUIRE(p, RG_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch");
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return RG_OK;
}

}  // namespace

extern "C" int rg_direct_acc_noalloc(const double *pos, const double *mass, double *acc,
                                     int64_t n, double G, double eps2, int precision,
                                     void *stream)
{
    void *p = nullptr;
    const int rc = noalloc_scratch(&p);
    if (rc != RG_OK) return rc;
    return direct_common(pos, mass, acc, nullptr, n, G, eps2, precision, 0.0,
                         make_epilogue(kStore, nullptr), stream,
                         ScratchSrc{p, kNoAllocScratchBytes});
}

extern "C" int rg_direct_acc_add(const double *pos, const double *mass, double *acc, int64_t n,
                                 double G, double eps2, int precision, void *stream)
{
    return direct_common(pos, mass, acc, nullptr, n, G, eps2, precision, 0.0,
                         make_epilogue(kAdd, nullptr), stream);
}

extern "C" int rg_direct_ext_acc_add(const double *pos, const double *mass, double *acc,
                                     int64_t n, double G, double eps2,
                                     const rg_potential_t *pot, int precision, void *stream)
{
    RG_REQUIRE(pot, RG_ERR_ARG, "null potential");
    return direct_common(pos, mass, acc, nullptr, n, G, eps2, precision, 0.0,
                         make_epilogue(kAddExt, pot), stream);
}

extern "C" int rg_analytic_acc_add(const double *pos, double *acc, int64_t n, double G,
                                   const rg_potential_t *pot, void *stream)
{
    RG_REQUIRE(n >= 0 && n <= (int64_t(1) << 29), RG_ERR_CAPACITY, "n out of range");
    if (n == 0) return RG_OK;
    RG_REQUIRE(pos && acc && pot, RG_ERR_ARG, "null pointer argument");
    const size_t b3 = sizeof(double) * 3 * static_cast<size_t>(n);
    RG_REQUIRE(!rg_overlap(acc, b3, pos, b3), RG_ERR_ALIAS, "acc aliases pos");
    const int ni = static_cast<int>(n);
    k_analytic_add<<<elementwise_grid(ni), 256, 0, rg::as_stream(stream)>>>(
        pos, acc, ni, G, make_epilogue(kAddExt, pot));
    RG_CHECK_LAUNCH();
    return RG_OK;
}

extern "C" int rg_add_field(double *acc, const double *src, int64_t n, void *stream)
{
    RG_REQUIRE(n >= 0 && n <= (int64_t(1) << 29), RG_ERR_CAPACITY, "n out of range");
    if (n == 0) return RG_OK;
    RG_REQUIRE(acc && src, RG_ERR_ARG, "null pointer argument");
    const size_t b3 = sizeof(double) * 3 * static_cast<size_t>(n);
    RG_REQUIRE(!rg_overlap(acc, b3, src, b3), RG_ERR_ALIAS, "acc and src overlap");
    const int m = static_cast<int>(3 * n);
    k_add_field<<<elementwise_grid(m), 256, 0, rg::as_stream(stream)>>>(acc, src, m);
    RG_CHECK_LAUNCH();
    return RG_OK;
}

extern "C" int rg_direct_acc_kick(const double *pos, const double *mass, double *acc,
                                  double *vel, int64_t n, double G, double eps2,
                                  int precision, double dt_half, void *stream)
{
    RG_REQUIRE(vel || n == 0, RG_ERR_ARG, "null vel");
    return direct_common(pos, mass, acc, vel, n, G, eps2, precision, dt_half,
                         make_epilogue(kStore, nullptr), stream);
}

extern "C" int rg_direct_acc_hilo(const double *pos, const float *pos_lo, const double *mass,
                                  double *acc, int64_t n, double G, double eps2, void *stream)
{
    RG_REQUIRE(n == 0 || pos_lo, RG_ERR_ARG, "null pos_lo");
    return direct_common(pos, mass, acc, nullptr, n, G, eps2, RG_FP32C, 0.0,
                         make_epilogue(kStore, nullptr), stream, ScratchSrc{nullptr, 0}, pos_lo);
}

extern "C" int rg_direct_acc_kick_hilo(const double *pos, const float *pos_lo,
                                       const double *mass, double *acc, double *vel, int64_t n,
                                       double G, double eps2, double dt_half, void *stream)
{
    RG_REQUIRE(n == 0 || (vel && pos_lo), RG_ERR_ARG, "null vel or pos_lo");
    return direct_common(pos, mass, acc, vel, n, G, eps2, RG_FP32C, dt_half,
                         make_epilogue(kStore, nullptr), stream, ScratchSrc{nullptr, 0}, pos_lo);
}

namespace {

int hilo_check(const double *pos, const float *pos_lo, const void *other, int64_t n)
{
    RG_REQUIRE(n >= 0 && n <= (int64_t(1) << 29), RG_ERR_CAPACITY, "n out of range");
    if (n == 0) return RG_OK;
    RG_REQUIRE(pos && pos_lo && other, RG_ERR_ARG, "null pointer argument");
    const size_t b3 = 24 * static_cast<size_t>(n), bl = 12 * static_cast<size_t>(n);
    RG_REQUIRE(!rg_overlap(pos, b3, pos_lo, bl) && !rg_overlap(pos, b3, other, b3) &&
                   !rg_overlap(pos_lo, bl, other, b3),
               RG_ERR_ALIAS, "pos, pos_lo and vel must not overlap");
    return RG_OK;
}

}  // namespace

extern "C" int rg_drift_hilo(double *pos, float *pos_lo, const double *vel, int64_t n, double dt,
                             void *stream)
{
    const int rc = hilo_check(pos, pos_lo, vel, n);
    if (rc != RG_OK || n == 0) return rc;
    const int m = static_cast<int>(3 * n);
    k_drift_hilo<<<elementwise_grid(m), 256, 0, rg::as_stream(stream)>>>(pos, pos_lo, vel, m, dt);
    RG_CHECK_LAUNCH();
    return RG_OK;
}

extern "C" int rg_kick_drift_hilo(double *pos, float *pos_lo, double *vel, const double *acc,
                                  int64_t n, double dt_half, double dt, void *stream)
{
    int rc = hilo_check(pos, pos_lo, vel, n);
    if (rc != RG_OK || n == 0) return rc;
    RG_REQUIRE(acc, RG_ERR_ARG, "null acc");
    const size_t b3 = 24 * static_cast<size_t>(n), bl = 12 * static_cast<size_t>(n);
    RG_REQUIRE(!rg_overlap(acc, b3, vel, b3) && !rg_overlap(acc, b3, pos, b3) &&
                   !rg_overlap(acc, b3, pos_lo, bl),
               RG_ERR_ALIAS, "acc aliases vel, pos or pos_lo");
    const int m = static_cast<int>(3 * n);
    hipStream_t st = rg::as_stream(stream);
    const auto aligned = [](const void *p, uintptr_t b) {
        return (reinterpret_cast<uintptr_t>(p) & (b - 1)) == 0;
    };
    if (aligned(pos, 16) && aligned(vel, 16) && aligned(acc, 16) && aligned(pos_lo, 8)) {
        const int pairs = m / 2;
        k_kick_drift_hilo2<<<elementwise_grid(pairs), 256, 0, st>>>(
            reinterpret_cast<double2 *>(pos), reinterpret_cast<float2 *>(pos_lo),
            reinterpret_cast<double2 *>(vel), reinterpret_cast<const double2 *>(acc), pairs,
            dt_half, dt);
        RG_CHECK_LAUNCH();
        if (m & 1) {
            const int last = m - 1;
            k_kick_drift_hilo<<<1, 64, 0, st>>>(pos + last, pos_lo + last, vel + last,
                                                acc + last, 1, dt_half, dt);
            RG_CHECK_LAUNCH();
        }
        return RG_OK;
    }
    k_kick_drift_hilo<<<elementwise_grid(m), 256, 0, st>>>(pos, pos_lo, vel, acc, m, dt_half, dt);
    RG_CHECK_LAUNCH();
    return RG_OK;
}

extern "C" int rg_kick(double *vel, const double *acc, int64_t n, double dt_half, void *stream)
{
    RG_REQUIRE(n >= 0 && n <= (int64_t(1) << 29), RG_ERR_CAPACITY, "n out of range");
    if (n == 0) return RG_OK;
    RG_REQUIRE(vel && acc, RG_ERR_ARG, "null pointer argument");
    const int m = static_cast<int>(3 * n);
    k_kick<<<elementwise_grid(m), 256, 0, rg::as_stream(stream)>>>(vel, acc, m, dt_half);
    RG_CHECK_LAUNCH();
    return RG_OK;
}

extern "C" int rg_drift(double *pos, const double *vel, int64_t n, double dt, void *stream)
{
    RG_REQUIRE(n >= 0 && n <= (int64_t(1) << 29), RG_ERR_CAPACITY, "n out of range");
    if (n == 0) return RG_OK;
    RG_REQUIRE(pos && vel, RG_ERR_ARG, "null pointer argument");
    const int m = static_cast<int>(3 * n);
    k_drift<<<elementwise_grid(m), 256, 0, rg::as_stream(stream)>>>(pos, vel, m, dt);
    RG_CHECK_LAUNCH();
    return RG_OK;
}

extern "C" int rg_kick_drift(double *pos, double *vel, const double *acc, int64_t n,
                             double dt_half, double dt, void *stream)
{
    RG_REQUIRE(n >= 0 && n <= (int64_t(1) << 29), RG_ERR_CAPACITY, "n out of range");
    if (n == 0) return RG_OK;
    RG_REQUIRE(pos && vel && acc, RG_ERR_ARG, "null pointer argument");
    const size_t b3 = sizeof(double) * 3 * static_cast<size_t>(n);
    RG_REQUIRE(!rg_overlap(pos, b3, vel, b3) && !rg_overlap(pos, b3, acc, b3) &&
                   !rg_overlap(vel, b3, acc, b3),
               RG_ERR_ALIAS, "pos, vel and acc must not overlap");
    const int m = static_cast<int>(3 * n);
    k_kick_drift<<<elementwise_grid(m), 256, 0, rg::as_stream(stream)>>>(pos, vel, acc, m,
                                                                         dt_half, dt);
    RG_CHECK_LAUNCH();
    return RG_OK;
}

extern "C" int rg_energy(const double *pos, const double *vel, const double *mass, int64_t n,
                         double G, double eps2, double *out5, void *stream)
{
    RG_REQUIRE(n >= 1 && n <= (int64_t(1) << 29), RG_ERR_CAPACITY, "n=%lld outside [1, 2^29]",
               static_cast<long long>(n));
    RG_REQUIRE(vel && mass && out5, RG_ERR_ARG, "null pointer argument");
    RG_REQUIRE(!pos || eps2 > 0.0, RG_ERR_ARG, "eps2 must be > 0 (got %g)", eps2);
    const int ni = static_cast<int>(n);
    const int nb = (ni + kEB - 1) / kEB;
    // the PE pass is split over j like the FP32 sums (it has no j order to
    // keep: its partials are reduced in a fixed order, to within 1e-13 of
    // the sequential sum)
    int jlen = ni;
    const int ns = pos ? rg::direct_slices(ni, 1, &jlen) : 1;
    void *buf = nullptr;
    hipStream_t s = rg::as_stream(stream);
    const int st = rg::device_scratch(sizeof(double) * 5 * static_cast<size_t>(nb) * ns, s,
                                      &buf);
    if (st != RG_OK) return st;
    double *part = static_cast<double *>(buf);
    k_energy_partial<<<dim3(nb, ns), kEB, 0, s>>>(pos, vel, mass, ni, G, eps2, jlen, part);
    RG_CHECK_LAUNCH();
    k_energy_final<<<1, kEB, 0, s>>>(part, nb * ns, out5);
    RG_CHECK_LAUNCH();
    return RG_OK;
}

int rg::rg_direct_set_fp32c_d64_impl(int64_t v)
{
    RG_REQUIRE(v == 0 || v == 1, RG_ERR_ARG, "fp32c_d64 must be 0 or 1");
    g_fp32c_d64 = static_cast<int>(v);
    return RG_OK;
}
