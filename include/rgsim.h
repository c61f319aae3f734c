/*
 * rgsim.h -- C ABI of the MI355X gravity solver (librgsim.so).
 *
 * Conventions
 *  - Every entry point returns an int status (RG_OK = 0).  On failure,
 *    rg_last_error() returns a thread-local message.  Nothing throws across
 *    the ABI and nothing truncates silently.
 *  - Array arguments are DEVICE pointers (e.g. torch tensor data_ptr()) in
 *    the reference's layout: positions/velocities/accelerations are
 *    (n, 3) row-major FP64, masses (n,) FP64.  The caller owns them.
 *  - `stream` is a hipStream_t (NULL = the default stream).  Calls are
 *    stream-ordered and asynchronous unless stated otherwise.
 *  - Scratch memory: a BH handle owns its scratch.  rg_energy,
 *    rg_morton_codes, and the FP32/FP32C direct sums and rg_direct_vjp when
 *    they split j (from n = 257 up by default; at most 2^24 x 32 B) use a
 *    library buffer per (device, stream), held until the process ends; the
 *    symmetric FP32/FP32C sums ("direct_sym", from n = 131,072) take about
 *    12 n^2 / 1024 + 24 n^2 / 4096 B there (1.0 GB at n = 262,144, at most
 *    3 GiB).  Calls on different streams never share it; calls issued to
 *    one stream from several host threads at once race.
 *  - `precision`: RG_FP64 reproduces the reference arithmetic; RG_FP32 is
 *    plain FP32 pair arithmetic; RG_FP32C is FP32 pair arithmetic on
 *    double-float (hi+lo) positions with compensated accumulation.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *  core/solvers/gravity.py, core/integrators/leapfrog.py,
 *  core/solvers/barnes_hut.py, core/engine.py.
 */
#ifndef RGSIM_H
#define RGSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RG_ABI_VERSION 1

enum {
    RG_OK = 0,
    RG_ERR_ARG = 1,        /* invalid argument (null pointer, n, eps2 <= 0, ...) */
    RG_ERR_HIP = 2,        /* HIP runtime error */
    RG_ERR_NONFINITE = 3,  /* non-finite input detected */
    RG_ERR_CAPACITY = 4,   /* buffer or size limit exceeded */
    RG_ERR_ALIAS = 5       /* output aliases an input */
};

enum { RG_FP64 = 0, RG_FP32 = 1, RG_FP32C = 2 };

/* Library ABI version (RG_ABI_VERSION). */
int rg_version(void);

/* Message of the last failed call on this host thread ("" if none). */
const char *rg_last_error(void);

/* Launch-shape tunables (no reference counterpart; used by the autotuner):
 *   "direct_ipt"  i-particles per lane in the direct kernels: 0 (by precision
 *                 and n: FP64 1; FP32C 4 from n = 65,536; FP32 2 from
 *                 n = 131,072; else 1), 1, 2 or 4
 *   "fp64_pipe"   waves per 64 i-particles of the FP64 direct sums: 0 (by n:
 *                 16 up to n = 16,384, 8 below 49,152; above, 1 when that
 *                 puts at least two waves on every SIMD, filled to 90% of the
 *                 busiest below four per SIMD, else 8), 1 (one wave per 64
 *                 i-particles), 4, 8 or 16.  Every setting gives the
 *                 same doubles (the reference's j order); with more than one
 *                 wave, the waves compute pair terms in parallel and three
 *                 waves add them in j order
 *   "sort_ticket" radix-sort passes (BH keys, SPH cells): 0 (default) take
 *                 each tile's index from its workgroup when all tiles fit on
 *                 the GPU at once (<= 256), else from an atomic ticket; 1
 *                 always from the ticket.  Same results either way
 *   "sort_spin_limit" re-polls a decoupled look-back wait (radix sort,
 *                 treelet ranking) makes before it gives up and flags the
 *                 failure: -1 restores the default 2^20, 0 gives up at once
 *                 (to test the failure path; every kernel still exits)
 *   "sort_block"  threads per radix-sort tile, for BH and SPH-grid handles
 *                 created afterwards: 0 (default: 1024 from 200,000 keys,
 *                 else 256), 256, or 1024 (with 4 keys per thread, whatever
 *                 "sort_items" says).  Same results either way
 *   "sort_sample" the BH tree's sort: -1 (default) the sample sort from the
 *                 handle's second compute when 4,096 <= n <= 1,572,864, 0
 *                 always the 8-pass (4 with 30-bit keys) LSD radix sort, 1
 *                 the sample sort from the second compute at any n up to
 *                 1,572,864.  The sample sort takes splitters from the
 *                 previous compute's order, moves the keys into up to 256
 *                 buckets in one radix pass and sorts each bucket in one
 *                 workgroup (in LDS up to 8,192 keys, else in global
 *                 memory, counted in rg_bh_stats).  Same order either way
 *   "sort_items"  keys per thread in a 256-thread radix-sort tile, for BH and
 *                 SPH-grid handles created afterwards: 0 (default: 4 below
 *                 524,288 keys, 8 below 786,432, else 16), 4, 8 or 16.  Same
 *                 results either way
 *   "direct_split" j-slices per i-block of the FP32/FP32C direct sums and of
 *                 rg_direct_vjp: 0 (default: 32 with one i-particle per
 *                 lane, 64 with more, fewer when slices x n would pass 2^24)
 *                 or 1..256, and never more than n/256 (a slice holds at
 *                 least one 256-particle tile).  Slice sums are added in
 *                 slice order in FP64, so results are deterministic for a
 *                 given n and setting; 1 gives the unsplit sum.  FP64 direct
 *                 sums are never split (they keep the reference's j order).
 *   "direct_sym"  the symmetric FP32/FP32C direct sums (each pair evaluated
 *                 once, its term added to both particles; deterministic,
 *                 the j-side sums added in FP64): 0 (default: from n =
 *                 131,072 while "direct_split" and "direct_ipt" are 0 and the
 *                 scratch stays within 3 GiB), 1 always, -1 never; 2 and 3
 *                 as 0 and 1 with the j-side sums rotating through the lanes
 *                 in FP32 (flushed to FP64 every 32 steps)
 *   "walk_block"  threads per block of the BH walk: 64, 128, 256 or 512
 *   "walk_mode"   FP32 BH walk: -1 auto (default: 7), 0 one lane per
 *                 particle, 1 two lanes per particle, 2 one wave walking
 *                 together, 3 the same advancing two slots per step, 4 / 5
 *                 the wave walk split over two / four waves per particle
 *                 group, 6 the wave walk over far-field chunks of the
 *                 preorder that the group's waves take in turn, 7 walk 6
 *                 with its record window in LDS (same results as 6, bitwise),
 *                 8 walk 7 with two chunks in flight per wave (same results)
 *   "chunk_dmin", "chunk_shift", "chunk_run", "chunk_waves", "chunk_read_ahead"
 *                 shape of walks 6 and 7: leaf distance of the nearest chunk
 *                 boundary beyond each end of a group (default 256); growth
 *                 factor 2^shift of the next ones (0 = by n: 1 below 350k
 *                 particles, else 2); runs of groups per XCD (32); waves per
 *                 group, 4 or 8, or 2 for walk 8 (0 = by n: 8 below 90k
 *                 particles, and below 300k with 10-bit keys); walk 7 reading
 *                 the next record before the force: -1 (default) or 1 on,
 *                 0 off (same results)
 * Process-wide; takes effect at the next launch. */
int rg_set_param(const char *name, int64_t value);

/* ------------------------------------------------------------- direct N^2 */

/* Replaces gravity.compute_acc_direct(pos, mass, acc, n, grav, eps2)
 * (core/solvers/gravity.py:13-26).  acc[i] = G sum_j m_j d_ij / (|d_ij|^2 +
 * eps2)^{3/2} over ALL j (j == i adds zero).  Overwrites acc.  eps2 must be
 * > 0 (the reference yields NaN for eps2 == 0).  acc must not alias pos or
 * mass. */
int rg_direct_acc(const double *pos, const double *mass, double *acc, int64_t n,
                  double G, double eps2, int precision, void *stream);

/* rg_direct_acc with no allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels,
 * launch shapes and results (bit for bit) as rg_direct_acc; the FP32/FP32C
 * j-slice partials go to a static 128 MiB device array that each GPU holds
 * from the library's first launch on it until the process ends, instead of
 * the library's per-stream buffer.  n whose split needs more (3 doubles x
 * slices x n; with the default split, n up to 87,381 in every precision)
 * is rejected with RG_ERR_CAPACITY before any launch.  The array needs no
 * clearing (every slice overwrites its partials), but every call on one GPU
 * shares it: calls that overlap on one GPU (other streams, other host
 * threads) race. */
int rg_direct_acc_noalloc(const double *pos, const double *mass, double *acc, int64_t n,
                          double G, double eps2, int precision, void *stream);

/* rg_direct_acc followed by leapfrog.kick(vel, acc, n, dt_half)
 * (core/integrators/leapfrog.py:11-15) fused into the force kernel's
 * epilogue: the closing half-kick of Engine.step (core/engine.py:113-114). */
int rg_direct_acc_kick(const double *pos, const double *mass, double *acc,
                       double *vel, int64_t n, double G, double eps2,
                       int precision, double dt_half, void *stream);

/* ---------------------------------------- hi/lo positions (precision form)
 * No reference counterpart: the reference holds FP64 positions only.  A
 * position is X = pos + pos_lo, pos (n,3) FP64 and pos_lo (n,3) FP32 an
 * offset of at most half an ulp of pos (zero for positions that are plain
 * doubles), so pos is X rounded to FP64.  The drift below keeps the sub-ulp
 * part of every increment that an FP64 drift rounds away. */

/* rg_direct_acc with FP32C pair arithmetic on hi/lo positions: each
 * position is split into FP32 hi + lo parts from both arrays
 * (gravity.py:13-26 otherwise, in the FP32C form). */
int rg_direct_acc_hilo(const double *pos, const float *pos_lo, const double *mass,
                       double *acc, int64_t n, double G, double eps2, void *stream);

/* rg_direct_acc_hilo followed by leapfrog.kick(vel, acc, n, dt_half)
 * (core/integrators/leapfrog.py:11-15) in the force kernel's epilogue. */
int rg_direct_acc_kick_hilo(const double *pos, const float *pos_lo, const double *mass,
                            double *acc, double *vel, int64_t n, double G, double eps2,
                            double dt_half, void *stream);

/* leapfrog.drift (core/integrators/leapfrog.py:18-22) on hi/lo positions,
 * compensated: per component t = pos_lo + vel * dt, (pos, e) = TwoSum(pos,
 * t), pos_lo = e rounded to FP32. */
int rg_drift_hilo(double *pos, float *pos_lo, const double *vel, int64_t n, double dt,
                  void *stream);

/* leapfrog.kick(dt_half) then rg_drift_hilo(dt) in one pass (the opening
 * half of Engine.step, core/engine.py:111-112). */
int rg_kick_drift_hilo(double *pos, float *pos_lo, double *vel, const double *acc, int64_t n,
                       double dt_half, double dt, void *stream);

/* ------------------------------------- self-gravity of the other engines */

/* The analytic galaxy potential the gas-disk and living-galaxy engines add
 * (their `pot` dict, core/gas_disk_engine.py:220-224): a Miyamoto-Nagai disk
 * (M_d, a, b), a Plummer halo -- or a Hernquist halo when is_hernquist == 1
 * -- (M_h, a_h), and a central point mass when smbh_mass > 0.  Evaluated in
 * FP64 whatever the precision of the self-gravity sum. */
typedef struct {
    double M_d, a, b;
    double M_h, a_h;
    double smbh_mass;
    int32_t is_hernquist;
} rg_potential_t;

/* Replaces living_galaxy_engine.grav_self_direct(pos, mass, acc, n, eps2)
 * (core/living_galaxy_engine.py:101-113): acc[i] += the rg_direct_acc sum
 * (same j loop, same precision choices). */
int rg_direct_acc_add(const double *pos, const double *mass, double *acc, int64_t n,
                      double G, double eps2, int precision, void *stream);

/* Replaces gas_disk_engine.add_grav_ext(pos, mass, acc, n, eps2, M_d, a, b,
 * M_h, a_h, is_hernquist, smbh_mass) (core/gas_disk_engine.py:20-51):
 * acc[i] += (((self-gravity sum + disk) + halo) + point mass), the analytic
 * terms fused into the force kernel's epilogue. */
int rg_direct_ext_acc_add(const double *pos, const double *mass, double *acc, int64_t n,
                          double G, double eps2, const rg_potential_t *pot, int precision,
                          void *stream);

/* Replaces living_galaxy_engine.grav_analytic (core/living_galaxy_engine.py:
 * 123-150) and gas_disk_engine.add_analytic (core/gas_disk_engine.py:54-82),
 * the same arithmetic: acc[i] += ((disk + halo) + point mass).  acc must not
 * alias pos. */
int rg_analytic_acc_add(const double *pos, double *acc, int64_t n, double G,
                        const rg_potential_t *pot, void *stream);

/* Replaces living_galaxy_engine.add_acc(acc, acc_g, n) (core/
 * living_galaxy_engine.py:116-120) and gas_disk_engine.add_field(acc, src, n)
 * (core/gas_disk_engine.py:84-89): acc += src.  The two must not overlap. */
int rg_add_field(double *acc, const double *src, int64_t n, void *stream);

/* ------------------------------------------------------------ integrator */

/* leapfrog.kick (core/integrators/leapfrog.py:11-15): vel += acc * dt_half. */
int rg_kick(double *vel, const double *acc, int64_t n, double dt_half, void *stream);

/* leapfrog.drift (core/integrators/leapfrog.py:18-22): pos += vel * dt. */
int rg_drift(double *pos, const double *vel, int64_t n, double dt, void *stream);

/* kick(dt_half) then drift(dt) in one pass, bit-identical to the two calls
 * in sequence (core/engine.py:111-112). */
int rg_kick_drift(double *pos, double *vel, const double *acc, int64_t n,
                  double dt_half, double dt, void *stream);

/* ----------------------------------------------------------- diagnostics */

/* Engine.energies (core/engine.py:119-124) -> gravity.kinetic_energy and
 * gravity.potential_energy (core/solvers/gravity.py:29-57), plus total
 * momentum.  out5 (device, 5 doubles) = {KE, PE, Px, Py, Pz}, all FP64,
 * reduced in a fixed order (bitwise reproducible for a given n and
 * "direct_split").  PE excludes j == i; its j range is split like the FP32
 * direct sums.
 * pos == NULL skips the N^2 potential pass (PE = 0). */
int rg_energy(const double *pos, const double *vel, const double *mass, int64_t n,
              double G, double eps2, double *out5, void *stream);

/* ------------------------------------------------------------ Barnes-Hut */

/* barnes_hut.morton_codes(pos, bits) (core/solvers/barnes_hut.py:38-46):
 * keys over the bounding cube with bits per axis (21 -> 63-bit keys,
 * 10 -> 30-bit keys).  A NaN coordinate quantises to 0. */
int rg_morton_codes(const double *pos, int64_t n, int bits, uint64_t *keys,
                    void *stream);

typedef struct rg_bh rg_bh;

/* BarnesHut(n, theta, softening) (core/solvers/barnes_hut.py:223-240).
 * n >= 2.  G is explicit (the reference hard-codes core.units.G).
 * precision RG_FP64 walks FP64 node records with the reference's exact
 * criterion and operation order; RG_FP32 walks FP32 node records
 * (aggregation stays FP64).  RG_FP32C is rejected (RG_ERR_ARG): the walk
 * has no compensated form.  Allocates all scratch on the current device.
 *
 * Failure reporting of the compute calls below: the radix sort's and the
 * treelet ranking's look-back waits are bounded ("sort_spin_limit").  When
 * one gives up, every kernel of that compute after the sort returns without
 * touching acc (or vel), and the device sets a pinned host word of the
 * handle.  The compute calls are asynchronous, so the failure is returned
 * (RG_ERR_CAPACITY, a radix-sort message) by the next rg_bh_compute*,
 * rg_bh_step or rg_bh_export_* call on the handle, or by rg_bh_stats, which
 * synchronises first; it is reported once. */
int rg_bh_create(int64_t n, double theta, double eps, double G, int morton_bits,
                 int precision, rg_bh **out);

/* BarnesHut.compute(pos, mass, acc) (core/solvers/barnes_hut.py:242-259):
 * Morton keys, stable radix sort, Karras build, bottom-up aggregation and
 * the theta walk, all on the device.  Overwrites acc. */
int rg_bh_compute(rg_bh *h, const double *pos, const double *mass, double *acc,
                  void *stream);

/* rg_bh_compute followed by leapfrog.kick(vel, acc, n, dt_half)
 * (core/integrators/leapfrog.py:11-15) fused into the walk's store: the
 * closing half-kick of Engine.step (core/engine.py:113-114), bit-identical
 * to the two calls. */
int rg_bh_compute_kick(rg_bh *h, const double *pos, const double *mass, double *acc,
                       double *vel, double dt_half, void *stream);

/* Engine.step's tree path (core/engine.py:108-116) in one call: kick(dt_half)
 * and drift(dt) (core/integrators/leapfrog.py:11-22), the tree force into
 * acc, and the closing kick(dt_half).  Bit-identical to rg_kick_drift
 * followed by rg_bh_compute_kick; the opening kick + drift pass also boxes
 * the drifted positions for the tree, which saves the separate box pass.
 * acc holds the previous force on entry and the new one on return; pos, vel
 * and acc must not overlap each other or mass. */
int rg_bh_step(rg_bh *h, double *pos, double *vel, const double *mass, double *acc,
               double dt_half, double dt, void *stream);

/* BarnesHut.compute(pos, mass, acc_g) followed by add_acc(acc, acc_g, n), as
 * core/living_galaxy_engine.py:1017-1018 and core/gas_disk_engine.py:218-219
 * call them, with the add fused into the walk's store: acc[p] += a_p, the
 * same doubles as the two calls, and no acc_g buffer. */
int rg_bh_compute_add(rg_bh *h, const double *pos, const double *mass, double *acc,
                      void *stream);

/* Copies of the last compute's tree in the reference's encoding (device
 * pointers): keys_sorted (n) = code (barnes_hut.py:248), order (n),
 * left/right (n-1), parent (2n-1; internal ids offset by n, root -1).
 * Any pointer may be NULL. */
int rg_bh_export_tree(rg_bh *h, uint64_t *keys_sorted, int32_t *order, int32_t *left,
                      int32_t *right, int32_t *parent, void *stream);

/* FP64 node data of the last compute, indexed by node id (2n-1):
 * com (x3), nmass, nmin (x3), nmax (x3)  (barnes_hut.py:232-236).
 * Any pointer may be NULL. */
int rg_bh_export_nodes(rg_bh *h, double *com, double *nmass, double *nmin,
                       double *nmax, void *stream);

typedef struct {
    int64_t n;
    int64_t visits_total;   /* nodes examined by the walk, all particles */
    int64_t visits_max;     /* max nodes examined by one particle */
    int64_t nonfinite;      /* non-finite input coordinates seen */
    int64_t walk_mode;      /* FP32 walk variant used (0..8, see "walk_mode"); -1 FP64 */
    int64_t walk_waves;     /* waves per group of 64 particles in it (0: one lane each) */
    int64_t sort_passes;    /* global radix passes of its sort: 8 or 4 (LSD), 1 (sample sort) */
    int64_t sort_buckets;   /* sample sort: buckets (0 for the LSD sort) */
    int64_t sort_spilled;   /* sample sort: buckets too large for LDS, sorted in global memory */
} rg_bh_stats_t;

/* Statistics of the last compute.  Synchronises that compute's stream.
 * Returns RG_ERR_CAPACITY (with *out filled) if that compute aborted (a
 * look-back wait gave up) or its walk or preorder pass ran out of bounds (a
 * corrupt tree). */
int rg_bh_stats(rg_bh *h, rg_bh_stats_t *out);

/* Per-phase timing with HIP events recorded on the compute's stream (no
 * reference counterpart; the reference measures wall time only).  While
 * enabled, each rg_bh_compute records 5 events (up to 4096 computes).
 * rg_bh_timing waits for the recorded computes and returns the summed
 * milliseconds of {bbox+Morton+sort, build, aggregation, walk} and how many
 * computes they cover, then resets the count. */
int rg_bh_set_timing(rg_bh *h, int enable);
int rg_bh_timing(rg_bh *h, double *ms4, int64_t *count);

int rg_bh_destroy(rg_bh *h);

/* ------------------------------------------------------ SPH neighbour grid */

/* Replaces core.solvers.neighbor_grid.NeighborGrid (core/solvers/
 * neighbor_grid.py:36-136): a uniform grid over the gas particles whose cell
 * size is the SPH search radius.  The reference bins on the host in numpy
 * and uploads; here every step runs on the device, bit-identical. */
typedef struct rg_grid rg_grid;

typedef struct {
    int64_t ngas;            /* gas particles binned (0: no grid, outputs untouched) */
    int64_t nx, ny, nz;      /* grid shape */
    int64_t n_cells;         /* nx * ny * nz */
    int64_t max_cells;       /* cell budget, max(64, 4 n) (neighbor_grid.py:42) */
    double lo[3];            /* grid origin: min gas position per axis */
    double cell;             /* cell size */
    double inv_cell;         /* 1 / cell */
} rg_grid_info_t;

/* NeighborGrid(n, min_radius) (neighbor_grid.py:39-50).  n >= 0. */
int rg_grid_create(int64_t n, double min_radius, rg_grid **out);

/* NeighborGrid.rebuild(pos, h, is_star) (neighbor_grid.py:52-92), or
 * rebuild_all(pos, h) (:94-136) when is_star is NULL.  pos (n,3) and h (n)
 * FP64, is_star (n) int32 (0 = gas), all device pointers.  Writes gsort (n:
 * gas indices in cell order, stable, 0 past ngas) and cell_start
 * (max_cells + 1: first gsort position of each cell, 0 past n_cells), and
 * fills *info (host).  Synchronises the stream, as the reference's rebuild
 * returns the gas count to the host.  A gas particle with a non-finite
 * position or smoothing length is rejected (RG_ERR_NONFINITE); the
 * reference would bin garbage. */
int rg_grid_rebuild(rg_grid *g, const double *pos, const double *h, const int32_t *is_star,
                    int32_t *gsort, int32_t *cell_start, rg_grid_info_t *info, void *stream);

/* The cell budget, for sizing cell_start (max_cells + 1 entries). */
int rg_grid_max_cells(const rg_grid *g, int64_t *max_cells);

int rg_grid_destroy(rg_grid *g);

/* --------------------------------------------------------------- adjoint */

/* Vector-Jacobian product of the direct acceleration (no reference
 * counterpart; the reference has no gradients).  With d = x_j - x_k,
 * s2 = |d|^2 + eps2 and T = I/s^3 - 3 d d^T / s^5:
 *   g_pos[k]  = G sum_j T_kj (m_k lam_j - m_j lam_k)
 *   g_mass[k] = -G sum_j lam_j . d / s^3
 * Overwrites g_pos (n,3) and g_mass (n).  precision RG_FP64 or RG_FP32C.
 * The j range is split as in the FP32 direct sums ("direct_split"), in
 * either precision, and the slice sums are added in slice order in FP64. */
int rg_direct_vjp(const double *pos, const double *mass, const double *lam,
                  double *g_pos, double *g_mass, int64_t n, double G, double eps2,
                  int precision, void *stream);

/* Elementwise part of the reverse of one KDK step (no reference
 * counterpart).  With g_pos/g_mass the VJP of the acceleration at the
 * step's end positions:  xbar += g_pos;  vbar += xbar * dt;
 * mbar += g_mass;  lam = vbar * lam_coef  (the VJP weight for the previous
 * step: lam_coef = dt inside a trajectory, dt/2 for the initial
 * acceleration).  All arrays (n,3) except mbar/g_mass (n). */
int rg_kdk_adjoint_update(double *xbar, double *vbar, double *mbar, const double *g_pos,
                          const double *g_mass, double *lam, int64_t n, double dt,
                          double lam_coef, void *stream);

/* ------------------------------------------------------------ batching */

/* An ensemble of independent systems advanced together (no reference
 * counterpart; the reference runs one system per Engine, core/engine.py:32).
 * System k owns particles [offsets[k], offsets[k+1]) of the concatenated
 * (Ntot,3)/(Ntot) FP64 arrays; G, eps2 (> 0) and dt are per system
 * (HOST arrays of nsys, copied).  precision RG_FP64 or RG_FP32C; each
 * system's accelerations are bitwise those of rg_direct_acc on that system
 * alone, run unsplit ("direct_split" 1; FP64 sums always are).  Not
 * thread-safe per handle. */
typedef struct rg_batch rg_batch;
int rg_batch_create(int64_t nsys, const int64_t *offsets, const double *G,
                    const double *eps2, const double *dt, int precision, rg_batch **out);

/* Launch shape (the autotuner's knobs): tile = i-particles per work item:
 * 64, 128 or 256 (one per thread), 512 or 1024 (256 threads with 2 or 4
 * each); grid = persistent workgroups (0 = 4 per CU).  Every shape gives
 * the same doubles.  Drops a captured graph. */
int rg_batch_set_shape(rg_batch *h, int tile, int grid);

/* Accelerations of every system (overwrites acc). */
int rg_batch_acc(rg_batch *h, const double *pos, const double *mass, double *acc,
                 void *stream);

/* nsteps KDK steps of every system with its own dt: per step one
 * kick+drift launch and one persistent force(+closing kick) launch.  acc
 * must hold the accelerations of the current positions on entry (as
 * Engine.step assumes, core/engine.py:108-116). */
int rg_batch_step(rg_batch *h, double *pos, double *vel, double *acc, const double *mass,
                  int64_t nsteps, void *stream);

/* Capture nsteps steps on these buffers into a hipGraph (replacing any
 * previous one); rg_batch_graph_launch replays it on `stream`. */
int rg_batch_graph_capture(rg_batch *h, double *pos, double *vel, double *acc,
                           const double *mass, int64_t nsteps);
int rg_batch_graph_launch(rg_batch *h, void *stream);

/* Per-system FP64 {KE, PE, Px, Py, Pz} into out (device, nsys x 5), fixed
 * reduction order. */
int rg_batch_energy(rg_batch *h, const double *pos, const double *vel, const double *mass,
                    double *out, void *stream);

int rg_batch_destroy(rg_batch *h);

#ifdef __cplusplus
}
#endif

#endif /* RGSIM_H */
