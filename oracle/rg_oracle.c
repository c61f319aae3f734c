/*
 * rg_oracle.c -- CPU restatement of the reference gravity hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg may load it.  It is never
 * linked into, called by, or used as a fallback for the product library.
 *
 * Each function restates one reference kernel (paths relative to the
 * reference checkout of Khaserdene/real-astronomy-simulation):
 *
 *   orc_direct_acc        core/solvers/gravity.py:13-26
 *   orc_kinetic_energy    core/solvers/gravity.py:29-36
 *   orc_potential_energy  core/solvers/gravity.py:39-57
 *   orc_kick / orc_drift  core/integrators/leapfrog.py:11-22
 *   orc_morton            core/solvers/barnes_hut.py:27-46
 *   orc_stable_argsort    core/solvers/barnes_hut.py:247 (np.argsort kind="stable")
 *   orc_build_radix       core/solvers/barnes_hut.py:50-121
 *   orc_aggregate         core/solvers/barnes_hut.py:124-177
 *   orc_traverse          core/solvers/barnes_hut.py:180-217
 *   orc_grav_self_direct  core/living_galaxy_engine.py:101-113
 *   orc_analytic_add      core/living_galaxy_engine.py:123-150
 *                         (= core/gas_disk_engine.py:54-82)
 *   orc_grav_ext_add      core/gas_disk_engine.py:20-51
 *
 * Arithmetic is FP64 in the reference's operation order.  The file is built
 * with -ffp-contract=off so no multiply-add is fused, matching numpy and the
 * reference's explicit expression order.  Outer loops are OpenMP-parallel and
 * inner loops serial, the way Taichi's CPU backend parallelises the outermost
 * for of a kernel.
 *
 * Pinning (what checks each function independently of this text) is listed
 * in the header of oracle/np_oracle.py and exercised by tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_threads(int t)
{
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* gravity.py:13-26.  a_i = sum_j (G*m_j*inv*inv*inv) * d, d = pos_j - pos_i,
 * r2 = d.d + eps2, inv = 1/sqrt(r2); j runs over ALL particles (j == i adds
 * an exact zero).  acc is overwritten. */
void orc_direct_acc(const double *pos, const double *mass, double *acc,
                    int64_t n, double grav, double eps2)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            const double dx = pos[3 * j] - px;
            const double dy = pos[3 * j + 1] - py;
            const double dz = pos[3 * j + 2] - pz;
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            const double inv = 1.0 / sqrt(r2);
            const double s = grav * mass[j] * inv * inv * inv;
            ax += s * dx;
            ay += s * dy;
            az += s * dz;
        }
        acc[3 * i] = ax;
        acc[3 * i + 1] = ay;
        acc[3 * i + 2] = az;
    }
}

/* gravity.py:29-36.  KE = sum_i 0.5*m_i*(v_i.v_i), summed in index order. */
double orc_kinetic_energy(const double *vel, const double *mass, int64_t n)
{
    double ke = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        ke += 0.5 * mass[i] * (vx * vx + vy * vy + vz * vz);
    }
    return ke;
}

/* gravity.py:39-57.  Per-particle pe_i = -0.5*G*m_i*sum_{j!=i} m_j/sqrt(r2);
 * the per-particle values are written to pe_i (may be NULL) and their
 * index-order sum is returned. */
double orc_potential_energy(const double *pos, const double *mass, int64_t n,
                            double grav, double eps2, double *pe_i)
{
    double *tmp = pe_i ? pe_i : (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        double acc_i = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            if (j != i) {
                const double dx = pos[3 * j] - px;
                const double dy = pos[3 * j + 1] - py;
                const double dz = pos[3 * j + 2] - pz;
                const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                acc_i += mass[j] / sqrt(r2);
            }
        }
        tmp[i] = -0.5 * grav * mass[i] * acc_i;
    }
    double pe = 0.0;
    for (int64_t i = 0; i < n; ++i) pe += tmp[i];
    if (!pe_i) free(tmp);
    return pe;
}

/* leapfrog.py:11-15.  v += a*dt_half, elementwise. */
void orc_kick(double *vel, const double *acc, int64_t n, double dt_half)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < 3 * n; ++k) vel[k] += acc[k] * dt_half;
}

/* leapfrog.py:18-22.  x += v*dt, elementwise. */
void orc_drift(double *pos, const double *vel, int64_t n, double dt)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < 3 * n; ++k) pos[k] += vel[k] * dt;
}

/* barnes_hut.py:27-35.  Spread the low 21 bits of x to every third bit. */
static inline uint64_t part1by2(uint64_t x)
{
    x &= 0x1FFFFFull;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

uint64_t orc_part1by2(uint64_t x) { return part1by2(x); }

/* barnes_hut.py:38-46.
 *   lo    = pos.min(axis=0)
 *   span  = (pos.max(axis=0) - lo).max();  span = span if span > 0 else 1.0
 *   scale = (2**bits - 1) / (span * 1.0000001)
 *   q     = uint64(clip((pos - lo) * scale, 0, 2**bits - 1))
 *   key   = p(qx) | p(qy) << 1 | p(qz) << 2
 * numpy's min/max propagate NaN; casting NaN to uint64 is undefined, so the
 * oracle (like the product) maps a NaN quantity to 0.  Returns the number of
 * non-finite coordinates seen. */
int64_t orc_morton(const double *pos, int64_t n, int bits, uint64_t *keys)
{
    double lo[3] = {INFINITY, INFINITY, INFINITY};
    double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    int64_t bad = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            const double v = pos[3 * i + a];
            if (!isfinite(v)) ++bad;
            if (v < lo[a]) lo[a] = v;
            if (v > hi[a]) hi[a] = v;
        }
    double span = hi[0] - lo[0];
    for (int a = 1; a < 3; ++a) {
        const double s = hi[a] - lo[a];
        if (s > span) span = s;
    }
    if (!(span > 0)) span = 1.0;
    const double qmax = (double)((1ull << bits) - 1ull);
    const double scale = qmax / (span * 1.0000001);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        uint64_t key = 0;
        for (int a = 0; a < 3; ++a) {
            double t = (pos[3 * i + a] - lo[a]) * scale;
            if (!(t >= 0.0)) t = 0.0; /* also maps NaN to 0 */
            if (t > qmax) t = qmax;
            key |= part1by2((uint64_t)t) << a;
        }
        keys[i] = key;
    }
    return bad;
}

/* barnes_hut.py:247 -- np.argsort(codes, kind="stable").  Stable LSD radix
 * sort on 8-bit digits; stability is the whole contract of the reference's
 * call, so any stable sort gives the identical permutation. */
void orc_stable_argsort(const uint64_t *keys, int64_t n, int32_t *order)
{
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
    for (int pass = 0; pass < 8; ++pass) {
        const int sh = 8 * pass;
        int64_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < n; ++i) cnt[((keys[order[i]] >> sh) & 0xFF) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (int64_t i = 0; i < n; ++i) {
            const int d = (int)((keys[order[i]] >> sh) & 0xFF);
            tmp[cnt[d]++] = order[i];
        }
        memcpy(order, tmp, sizeof(int32_t) * (size_t)n);
    }
    free(tmp);
}

/* barnes_hut.py:50-58.  Leading zeros of a 63-bit value, counted from bit 62
 * with the reference's bit loop. */
static inline int clz64_ref(int64_t x)
{
    int c = 0;
    int64_t mask = (int64_t)1 << 62;
    while (mask > 0 && (x & mask) == 0) {
        ++c;
        mask >>= 1;
    }
    return c;
}

/* barnes_hut.py:61-68.  Leading zeros counted from bit 30. */
static inline int clz_i32_ref(int32_t x)
{
    int c = 0;
    int32_t mask = (int32_t)1 << 30;
    while (mask > 0 && (x & mask) == 0) {
        ++c;
        mask >>= 1;
    }
    return c;
}

/* barnes_hut.py:71-81. */
static inline int delta_ref(int64_t i, int64_t j, int64_t n, const int64_t *code)
{
    int res = -1;
    if (0 <= j && j < n) {
        const int64_t x = code[i] ^ code[j];
        if (x == 0)
            res = 63 + clz_i32_ref((int32_t)(i ^ j));
        else
            res = clz64_ref(x);
    }
    return res;
}

int orc_delta(int64_t i, int64_t j, int64_t n, const int64_t *code)
{
    return delta_ref(i, j, n, code);
}

int orc_clz64(int64_t x) { return clz64_ref(x); }
int orc_clz_i32(int32_t x) { return clz_i32_ref(x); }

/* barnes_hut.py:84-121 (Karras 2012).  Internal node i (0..n-2) has id i+n;
 * child ids are leaf (<n) or internal (+n); parent is indexed by node id and
 * holds internal ids (+n); the root (id n) keeps parent -1, as does every
 * slot the build does not write (barnes_hut.py:250 resets parent to -1). */
void orc_build_radix(const int64_t *code, int64_t n, int32_t *left,
                     int32_t *right, int32_t *parent)
{
    for (int64_t k = 0; k < 2 * n - 1; ++k) parent[k] = -1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n - 1; ++i) {
        const int64_t d = delta_ref(i, i + 1, n, code) > delta_ref(i, i - 1, n, code) ? 1 : -1;
        const int d_min = delta_ref(i, i - d, n, code);
        int64_t l_max = 2;
        while (delta_ref(i, i + l_max * d, n, code) > d_min) l_max *= 2;
        int64_t l = 0;
        int64_t t = l_max / 2;
        while (t >= 1) {
            if (delta_ref(i, i + (l + t) * d, n, code) > d_min) l += t;
            t /= 2;
        }
        const int64_t j = i + l * d;
        const int d_node = delta_ref(i, j, n, code);
        int64_t s = 0;
        int64_t t2 = (l + 1) / 2;
        while (t2 >= 1) {
            if (delta_ref(i, i + (s + t2) * d, n, code) > d_node) s += t2;
            t2 = t2 > 1 ? (t2 + 1) / 2 : 0;
        }
        const int64_t gamma = i + s * d + (d < 0 ? d : 0);
        const int64_t lo = i < j ? i : j;
        const int64_t hi = i < j ? j : i;
        const int64_t lc = (lo == gamma) ? gamma : gamma + n;
        const int64_t rc = (hi == gamma + 1) ? gamma + 1 : gamma + 1 + n;
        left[i] = (int32_t)lc;
        right[i] = (int32_t)rc;
        parent[lc] = (int32_t)(i + n);
        parent[rc] = (int32_t)(i + n);
    }
}

/* barnes_hut.py:124-177.  Leaf init, then every leaf adds its mass, mass
 * moment and AABB into every ancestor (the reference does this with atomics
 * in scheduler order; the oracle applies the leaves in index order, one of
 * the orders the reference can produce), then com = mom / mass where
 * mass > 0.  Internal nodes with zero mass keep com = 0 (a freshly allocated
 * reference field).  Arrays are (2n-1)x3 / (2n-1), indexed by node id. */
void orc_aggregate(const int32_t *order, const double *pos, const double *mass,
                   const int32_t *parent, int64_t n, double *com, double *nmass,
                   double *nmin, double *nmax)
{
    const int64_t m = 2 * n - 1;
    double *nmom = (double *)calloc((size_t)(3 * m), sizeof(double));
    for (int64_t li = 0; li < n; ++li) {
        const int64_t p = order[li];
        for (int a = 0; a < 3; ++a) {
            com[3 * li + a] = pos[3 * p + a];
            nmin[3 * li + a] = pos[3 * p + a];
            nmax[3 * li + a] = pos[3 * p + a];
        }
        nmass[li] = mass[p];
    }
    for (int64_t node = n; node < m; ++node) {
        nmass[node] = 0.0;
        for (int a = 0; a < 3; ++a) {
            com[3 * node + a] = 0.0;
            nmin[3 * node + a] = 1e30;
            nmax[3 * node + a] = -1e30;
        }
    }
    for (int64_t li = 0; li < n; ++li) {
        const double mm = nmass[li];
        const double *pp = &com[3 * li];
        int64_t node = parent[li];
        while (node >= 0) {
            nmass[node] += mm;
            for (int a = 0; a < 3; ++a) {
                nmom[3 * node + a] += mm * pp[a];
                if (pp[a] < nmin[3 * node + a]) nmin[3 * node + a] = pp[a];
                if (pp[a] > nmax[3 * node + a]) nmax[3 * node + a] = pp[a];
            }
            node = parent[node];
        }
    }
    for (int64_t node = n; node < m; ++node) {
        const double mm = nmass[node];
        if (mm > 0.0)
            for (int a = 0; a < 3; ++a) com[3 * node + a] = nmom[3 * node + a] / mm;
    }
    free(nmom);
}

/* barnes_hut.py:180-217.  Per particle p (original order): stack walk from
 * the root, push left then right (so right is popped first), drop both
 * children when the stack would exceed max_stack (counted in *overflow).
 * Leaves skip order[leaf] == p; internal nodes are accepted when
 * size*size < theta2*r2 with size = max-axis AABB extent.  visits[p] (may be
 * NULL) receives the number of nodes popped.  Returns total visits. */
int64_t orc_traverse(const double *pos, const int32_t *order, const double *com,
                     const double *nmass, const double *nmin, const double *nmax,
                     const int32_t *left, const int32_t *right, double *acc,
                     int64_t n, double grav, double theta2, double eps2,
                     int max_stack, int64_t *visits, int64_t *overflow)
{
    int64_t total = 0, ovf = 0;
    const int64_t root = n;
#pragma omp parallel reduction(+ : total, ovf)
    {
        int32_t *stack = (int32_t *)malloc(sizeof(int32_t) * (size_t)(max_stack > 0 ? max_stack : 1));
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; ++p) {
            const double px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
            double ax = 0.0, ay = 0.0, az = 0.0;
            int sp = 0;
            int64_t vis = 0;
            stack[sp++] = (int32_t)root;
            while (sp > 0) {
                const int64_t node = stack[--sp];
                ++vis;
                if (node < n) {
                    const int64_t q = order[node];
                    if (q != p) {
                        const double dx = com[3 * node] - px;
                        const double dy = com[3 * node + 1] - py;
                        const double dz = com[3 * node + 2] - pz;
                        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                        const double inv = 1.0 / sqrt(r2);
                        const double s = grav * nmass[node] * inv * inv * inv;
                        ax += s * dx;
                        ay += s * dy;
                        az += s * dz;
                    }
                } else {
                    const int64_t ii = node - n;
                    const double dx = com[3 * node] - px;
                    const double dy = com[3 * node + 1] - py;
                    const double dz = com[3 * node + 2] - pz;
                    const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                    double size = nmax[3 * node] - nmin[3 * node];
                    const double sy = nmax[3 * node + 1] - nmin[3 * node + 1];
                    const double sz = nmax[3 * node + 2] - nmin[3 * node + 2];
                    if (sy > size) size = sy;
                    if (sz > size) size = sz;
                    if (size * size < theta2 * r2) {
                        const double inv = 1.0 / sqrt(r2);
                        const double s = grav * nmass[node] * inv * inv * inv;
                        ax += s * dx;
                        ay += s * dy;
                        az += s * dz;
                    } else if (sp + 2 <= max_stack) {
                        stack[sp++] = left[ii];
                        stack[sp++] = right[ii];
                    } else {
                        ++ovf;
                    }
                }
            }
            acc[3 * p] = ax;
            acc[3 * p + 1] = ay;
            acc[3 * p + 2] = az;
            if (visits) visits[p] = vis;
            total += vis;
        }
        free(stack);
    }
    if (overflow) *overflow = ovf;
    return total;
}

/* barnes_hut.py:242-259, BarnesHut.compute: Morton keys from the current
 * positions, stable argsort, build, aggregate, traverse.  Scratch is
 * allocated per call.  Returns total visits; *overflow as orc_traverse. */
int64_t orc_bh_compute(const double *pos, const double *mass, double *acc,
                       int64_t n, double grav, double theta, double eps,
                       int bits, int max_stack, int64_t *overflow)
{
    const int64_t m = 2 * n - 1;
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    int64_t *code = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int32_t *left = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 1 ? n - 1 : 1));
    int32_t *right = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 1 ? n - 1 : 1));
    int32_t *parent = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
    double *com = (double *)malloc(sizeof(double) * (size_t)(3 * m));
    double *nmass = (double *)malloc(sizeof(double) * (size_t)m);
    double *nmin = (double *)malloc(sizeof(double) * (size_t)(3 * m));
    double *nmax = (double *)malloc(sizeof(double) * (size_t)(3 * m));
    orc_morton(pos, n, bits, keys);
    orc_stable_argsort(keys, n, order);
    for (int64_t i = 0; i < n; ++i) code[i] = (int64_t)keys[order[i]];
    orc_build_radix(code, n, left, right, parent);
    orc_aggregate(order, pos, mass, parent, n, com, nmass, nmin, nmax);
    const int64_t tot = orc_traverse(pos, order, com, nmass, nmin, nmax, left, right, acc,
                                     n, grav, theta * theta, eps * eps, max_stack, NULL,
                                     overflow);
    free(keys); free(code); free(order); free(left); free(right); free(parent);
    free(com); free(nmass); free(nmin); free(nmax);
    return tot;
}

/* ------------------------------------------------------------------------
 * Self-gravity and the analytic galaxy potential of the two other engines.
 * The potential is a Miyamoto-Nagai disk (M_d, a, b), a Plummer halo, or a
 * Hernquist halo when is_hernquist == 1 (M_h, a_h), and an optional central
 * point mass smbh_mass.  Taichi's `x ** 2` is x*x, `v.norm()` is
 * sqrt((x*x + y*y) + z*z), and `x ** (-1.5)` is pow(x, -1.5).
 */

/* The disk term, then halo and point-mass terms added to av in the order of
 * gas_disk_engine.py:34-49.  With init != 0 av starts as the disk vector
 * (living_galaxy_engine.py:134-140); otherwise the disk vector is added to
 * the self-gravity sum already in av (gas_disk_engine.py:34-39). */
static void ext_terms(double px, double py, double pz, double grav, double M_d, double a,
                      double b, double M_h, double a_h, int is_hernquist, double smbh_mass,
                      double av[3], int init)
{
    const double R2 = px * px + py * py;
    const double S = sqrt(pz * pz + b * b);
    const double aS = a + S;
    const double invD3 = pow(R2 + aS * aS, -1.5);
    const double gmd = -grav * M_d;
    const double dx = gmd * px * invD3;
    const double dy = gmd * py * invD3;
    const double dz = gmd * pz * aS * invD3 / S;
    if (init) {
        av[0] = dx; av[1] = dy; av[2] = dz;
    } else {
        av[0] += dx; av[1] += dy; av[2] += dz;
    }
    if (is_hernquist == 1) {
        const double r = sqrt(px * px + py * py + pz * pz);
        const double r_safe = r > 1e-6 ? r : 1e-6;          /* ti.max(r, 1e-6) */
        const double t = r_safe + a_h;
        const double c = -(grav * M_h / (t * t * r_safe));
        av[0] += c * px; av[1] += c * py; av[2] += c * pz;
    } else {
        const double rh = R2 + pz * pz + a_h * a_h;
        const double c = -(grav * M_h * pow(rh, -1.5));
        av[0] += c * px; av[1] += c * py; av[2] += c * pz;
    }
    if (smbh_mass > 0.0) {
        const double r_safe = sqrt(px * px + py * py + pz * pz) + 1e-6;
        const double c = -(grav * smbh_mass / (r_safe * r_safe * r_safe));
        av[0] += c * px; av[1] += c * py; av[2] += c * pz;
    }
}

/* Softened sum over all j, the loop of gravity.py:21-25. */
static void self_sum(const double *pos, const double *mass, int64_t n, int64_t i, double grav,
                     double eps2, double av[3])
{
    const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        const double dx = pos[3 * j] - px;
        const double dy = pos[3 * j + 1] - py;
        const double dz = pos[3 * j + 2] - pz;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        const double inv = 1.0 / sqrt(r2);
        const double s = grav * mass[j] * inv * inv * inv;
        ax += s * dx;
        ay += s * dy;
        az += s * dz;
    }
    av[0] = ax; av[1] = ay; av[2] = az;
}

/* living_galaxy_engine.py:101-113: acc[i] += sum_j (softened N^2). */
void orc_grav_self_direct(const double *pos, const double *mass, double *acc, int64_t n,
                          double grav, double eps2)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double av[3];
        self_sum(pos, mass, n, i, grav, eps2, av);
        acc[3 * i] = acc[3 * i] + av[0];
        acc[3 * i + 1] = acc[3 * i + 1] + av[1];
        acc[3 * i + 2] = acc[3 * i + 2] + av[2];
    }
}

/* living_galaxy_engine.py:123-150 (and gas_disk_engine.py:54-82, the same
 * arithmetic): acc[i] += ((disk + halo) + point mass). */
void orc_analytic_add(const double *pos, double *acc, int64_t n, double grav, double M_d,
                      double a, double b, double M_h, double a_h, int is_hernquist,
                      double smbh_mass)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double av[3];
        ext_terms(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], grav, M_d, a, b, M_h, a_h,
                  is_hernquist, smbh_mass, av, 1);
        acc[3 * i] = acc[3 * i] + av[0];
        acc[3 * i + 1] = acc[3 * i + 1] + av[1];
        acc[3 * i + 2] = acc[3 * i + 2] + av[2];
    }
}

/* gas_disk_engine.py:20-51: acc[i] += (((N^2 sum + disk) + halo) + point mass). */
void orc_grav_ext_add(const double *pos, const double *mass, double *acc, int64_t n,
                      double grav, double eps2, double M_d, double a, double b, double M_h,
                      double a_h, int is_hernquist, double smbh_mass)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double av[3];
        self_sum(pos, mass, n, i, grav, eps2, av);
        ext_terms(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], grav, M_d, a, b, M_h, a_h,
                  is_hernquist, smbh_mass, av, 0);
        acc[3 * i] = acc[3 * i] + av[0];
        acc[3 * i + 1] = acc[3 * i + 1] + av[1];
        acc[3 * i + 2] = acc[3 * i + 2] + av[2];
    }
}
