/*
 * tip_oracle.c -- CPU restatement of the reference's 3D Langevin RK2 lattice step and its
 * in-loop observables.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark call it; the product path never
 * does.  Every function cites the reference text it restates (paths relative to the
 * reference checkout).
 *
 * Precision semantics:
 *   *_f64       : float64 fields, float64 arithmetic, operations in the reference's textual
 *                 order (this file is compiled with -ffp-contract=off).
 *   *_f32mixed  : float32 fields as the reference ships them (field_dtype = float32,
 *                 simulation/latticeSimeRescale_numba.py:348,736).  Numba types the
 *                 six-neighbour sum float32+float32 -> float32 and everything that touches a
 *                 float64 literal or scalar (6.0*ph, inv_dx2, the f64 table) in float64; stores
 *                 round to float32.  This is restated literally and is also the CPU baseline.
 *
 * Pinning: every function here is checked in tests/test_oracle_*.py against closed forms,
 * invariants, brute force, an independent numpy restatement, or the reference's own numbers
 * (see the test docstrings).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------------------------
 * Counter-based hash RNG.
 * simulation/latticeSimeRescale_numba.py:2322-2336 (_hash_rng_pair), identical text in
 * simulation/latticeSimComplex_numba.py:2747-2761 and cosmolattice_ext/evolvers/stochasticrk.h:205-225.
 * x1 is the state after the first mixer (source of u1), x2 after the second (source of u2).
 * ------------------------------------------------------------------------------------------ */
static inline void hash_pair(uint64_t seed, uint64_t *x1, uint64_t *x2, double *u1, double *u2)
{
    uint64_t x = seed;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    x = x ^ (x >> 31);
    *x1 = x;
    double a = (double)(x >> 11) * (1.0 / 9007199254740992.0);
    x = (x * 0x2545F4914F6CDD1DULL) ^ 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = x ^ (x >> 31);
    *x2 = x;
    double b = (double)(x >> 11) * (1.0 / 9007199254740992.0);
    if (a < 1e-12) a = 1e-12;
    *u1 = a;
    *u2 = b;
}

/* Seed of one site: simulation/latticeSimeRescale_numba.py:2511-2516 (pass 2, K=73856093) and
 * :2636-2641 (pass 4, K=19349669): uint64(site * K) ^ uint64(noise_seed). */
static inline uint64_t site_seed(int64_t site, int64_t mult, uint64_t noise_seed)
{
    return ((uint64_t)(site * mult)) ^ noise_seed;
}

ORC_API void orc_hash_pairs(const uint64_t *seeds, int64_t n, uint64_t *x1, uint64_t *x2,
                            double *u1, double *u2)
{
    for (int64_t i = 0; i < n; ++i) hash_pair(seeds[i], &x1[i], &x2[i], &u1[i], &u2[i]);
}

/* Box-Muller of the fused kernels: z = sqrt(-2 ln u1) * cos(2 pi u2)
 * (latticeSimeRescale_numba.py:2518-2520); the complex pair uses sin for the second
 * component (latticeSimComplex_numba.py:2965-2974). */
static inline void box_muller(double u1, double u2, double *zc, double *zs)
{
    double r = sqrt(-2.0 * log(u1));
    double ang = 6.283185307179586 * u2;
    *zc = r * cos(ang);
    *zs = r * sin(ang);
}

/* Hash noise field: noise[site] = scale * z(site), site = i*Ny*Nz + j*Nz + k.  This is the
 * noise stream of the reference's fused kernels, written to an array so that the two-pass
 * rk2_step_table (which takes a pre-generated noise field, latticeSimeRescale_numba.py:1348-1361)
 * can consume it.  out2 (nullable) receives scale * r * sin (complex second component). */
ORC_API void orc_noise_field_f64(double *out1, double *out2, int64_t nx, int64_t ny, int64_t nz,
                                 uint64_t noise_seed, int64_t mult, double scale)
{
    int64_t n = nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < n; ++s) {
        uint64_t x1, x2;
        double u1, u2, zc, zs;
        hash_pair(site_seed(s, mult, noise_seed), &x1, &x2, &u1, &u2);
        box_muller(u1, u2, &zc, &zs);
        out1[s] = scale * zc;
        if (out2) out2[s] = scale * zs;
    }
}

/* float32 noise array (the reference allocates noise in field_dtype,
 * latticeSimeRescale_numba.py:3020): value computed in float64 then rounded. */
ORC_API void orc_noise_field_f32(float *out1, float *out2, int64_t nx, int64_t ny, int64_t nz,
                                 uint64_t noise_seed, int64_t mult, double scale)
{
    int64_t n = nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < n; ++s) {
        uint64_t x1, x2;
        double u1, u2, zc, zs;
        hash_pair(site_seed(s, mult, noise_seed), &x1, &x2, &u1, &u2);
        box_muller(u1, u2, &zc, &zs);
        out1[s] = (float)(scale * zc);
        if (out2) out2[s] = (float)(scale * zs);
    }
}

/* ------------------------------------------------------------------------------------------
 * V' table lookup, latticeSimeRescale_numba.py:1923-1930 (complex :2137-2144):
 *   fidx = (x - tmin) * dinv; idx = int(fidx) (truncation toward zero);
 *   idx clamped to [0, npts-2]; frac = fidx - idx left unclamped (linear extrapolation);
 *   table[idx] + frac * (table[idx+1] - table[idx]).
 * odd != 0 evaluates a table tabulated on x >= 0 at |x| and restores the sign (V' is odd in
 * phi for the real field: every term of Vprime_scalar, :805-837, is proportional to phi).
 * ------------------------------------------------------------------------------------------ */
static inline double lookup(double x, const double *t, double tmin, double dinv, int64_t npts, int odd)
{
    double sgn = 1.0;
    if (odd && x < 0.0) { x = -x; sgn = -1.0; }
    double fidx = (x - tmin) * dinv;
    int64_t idx = (int64_t)fidx;
    int64_t last = npts - 2;
    if (idx < 0) idx = 0;
    else if (idx > last) idx = last;
    double frac = fidx - (double)idx;
    double v = t[idx] + frac * (t[idx + 1] - t[idx]);
    return odd ? sgn * v : v;
}

ORC_API void orc_lookup(const double *x, int64_t n, double *out, const double *table, double tmin,
                        double dinv, int64_t npts, int odd)
{
    for (int64_t i = 0; i < n; ++i) out[i] = lookup(x[i], table, tmin, dinv, npts, odd);
}

#define IDX(i, j, k) (((int64_t)(i) * ny + (j)) * nz + (k))

/* ------------------------------------------------------------------------------------------
 * rk2_step_table, real field, float64.  latticeSimeRescale_numba.py:1875-1966.
 * Pass 1 writes (phi_tmp, pi_tmp); pass 2 reads phi_tmp's stencil and updates phi, pi in place.
 * ------------------------------------------------------------------------------------------ */
ORC_API void orc_rk2_step_table_f64(double *phi, double *pi, double dt, double dx, double eta_eff,
                                    double mu, const double *noise, double *phi_tmp,
                                    double *pi_tmp, double inv_a2, const double *table,
                                    double tmin, double dinv, int64_t npts, int odd, int64_t nx,
                                    int64_t ny, int64_t nz)
{
    const double inv_dx2 = 1.0 / (dx * dx);
    const double inv_mu2 = 1.0 / (mu * mu);
    const double half_dt = 0.5 * dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nx; ++i) {
        int64_t ip = i + 1 < nx ? i + 1 : 0, im = i - 1 >= 0 ? i - 1 : nx - 1;
        for (int64_t j = 0; j < ny; ++j) {
            int64_t jp = j + 1 < ny ? j + 1 : 0, jm = j - 1 >= 0 ? j - 1 : ny - 1;
            for (int64_t k = 0; k < nz; ++k) {
                int64_t kp = k + 1 < nz ? k + 1 : 0, km = k - 1 >= 0 ? k - 1 : nz - 1;
                double ph = phi[IDX(i, j, k)], pv = pi[IDX(i, j, k)];
                double lap = (phi[IDX(ip, j, k)] + phi[IDX(im, j, k)] + phi[IDX(i, jp, k)] +
                              phi[IDX(i, jm, k)] + phi[IDX(i, j, kp)] + phi[IDX(i, j, km)] -
                              6.0 * ph) * inv_dx2;
                double dV = lookup(ph, table, tmin, dinv, npts, odd);
                double kpi = inv_a2 * lap - eta_eff * pv - dV * inv_mu2;
                phi_tmp[IDX(i, j, k)] = ph + half_dt * pv;
                pi_tmp[IDX(i, j, k)] = pv + half_dt * kpi;
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nx; ++i) {
        int64_t ip = i + 1 < nx ? i + 1 : 0, im = i - 1 >= 0 ? i - 1 : nx - 1;
        for (int64_t j = 0; j < ny; ++j) {
            int64_t jp = j + 1 < ny ? j + 1 : 0, jm = j - 1 >= 0 ? j - 1 : ny - 1;
            for (int64_t k = 0; k < nz; ++k) {
                int64_t kp = k + 1 < nz ? k + 1 : 0, km = k - 1 >= 0 ? k - 1 : nz - 1;
                double ph = phi_tmp[IDX(i, j, k)], pv = pi_tmp[IDX(i, j, k)];
                double lap = (phi_tmp[IDX(ip, j, k)] + phi_tmp[IDX(im, j, k)] +
                              phi_tmp[IDX(i, jp, k)] + phi_tmp[IDX(i, jm, k)] +
                              phi_tmp[IDX(i, j, kp)] + phi_tmp[IDX(i, j, km)] - 6.0 * ph) * inv_dx2;
                double dV = lookup(ph, table, tmin, dinv, npts, odd);
                double kpi = inv_a2 * lap - eta_eff * pv - dV * inv_mu2;
                phi[IDX(i, j, k)] += dt * pv;
                pi[IDX(i, j, k)] += dt * kpi + (noise ? noise[IDX(i, j, k)] : 0.0);
            }
        }
    }
}

/* rk2_step_table, real field, float32 storage with the reference's mixed precision (see file
 * header).  Same text as above, latticeSimeRescale_numba.py:1875-1966.  Used as the CPU
 * baseline: two prange-style passes over x with the noise field generated beforehand. */
ORC_API void orc_rk2_step_table_f32mixed(float *phi, float *pi, double dt, double dx,
                                         double eta_eff, double mu, const float *noise,
                                         float *phi_tmp, float *pi_tmp, double inv_a2,
                                         const double *table, double tmin, double dinv,
                                         int64_t npts, int odd, int64_t nx, int64_t ny, int64_t nz)
{
    const double inv_dx2 = 1.0 / (dx * dx);
    const double inv_mu2 = 1.0 / (mu * mu);
    const double half_dt = 0.5 * dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nx; ++i) {
        int64_t ip = i + 1 < nx ? i + 1 : 0, im = i - 1 >= 0 ? i - 1 : nx - 1;
        for (int64_t j = 0; j < ny; ++j) {
            int64_t jp = j + 1 < ny ? j + 1 : 0, jm = j - 1 >= 0 ? j - 1 : ny - 1;
            for (int64_t k = 0; k < nz; ++k) {
                int64_t kp = k + 1 < nz ? k + 1 : 0, km = k - 1 >= 0 ? k - 1 : nz - 1;
                float ph = phi[IDX(i, j, k)], pv = pi[IDX(i, j, k)];
                float s6 = phi[IDX(ip, j, k)] + phi[IDX(im, j, k)];
                s6 = s6 + phi[IDX(i, jp, k)];
                s6 = s6 + phi[IDX(i, jm, k)];
                s6 = s6 + phi[IDX(i, j, kp)];
                s6 = s6 + phi[IDX(i, j, km)];
                double lap = ((double)s6 - 6.0 * (double)ph) * inv_dx2;
                double dV = lookup((double)ph, table, tmin, dinv, npts, odd);
                double kpi = inv_a2 * lap - eta_eff * (double)pv - dV * inv_mu2;
                phi_tmp[IDX(i, j, k)] = (float)((double)ph + half_dt * (double)pv);
                pi_tmp[IDX(i, j, k)] = (float)((double)pv + half_dt * kpi);
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nx; ++i) {
        int64_t ip = i + 1 < nx ? i + 1 : 0, im = i - 1 >= 0 ? i - 1 : nx - 1;
        for (int64_t j = 0; j < ny; ++j) {
            int64_t jp = j + 1 < ny ? j + 1 : 0, jm = j - 1 >= 0 ? j - 1 : ny - 1;
            for (int64_t k = 0; k < nz; ++k) {
                int64_t kp = k + 1 < nz ? k + 1 : 0, km = k - 1 >= 0 ? k - 1 : nz - 1;
                float ph = phi_tmp[IDX(i, j, k)], pv = pi_tmp[IDX(i, j, k)];
                float s6 = phi_tmp[IDX(ip, j, k)] + phi_tmp[IDX(im, j, k)];
                s6 = s6 + phi_tmp[IDX(i, jp, k)];
                s6 = s6 + phi_tmp[IDX(i, jm, k)];
                s6 = s6 + phi_tmp[IDX(i, j, kp)];
                s6 = s6 + phi_tmp[IDX(i, j, km)];
                double lap = ((double)s6 - 6.0 * (double)ph) * inv_dx2;
                double dV = lookup((double)ph, table, tmin, dinv, npts, odd);
                double kpi = inv_a2 * lap - eta_eff * (double)pv - dV * inv_mu2;
                phi[IDX(i, j, k)] = (float)((double)phi[IDX(i, j, k)] + dt * (double)pv);
                pi[IDX(i, j, k)] = (float)((double)pi[IDX(i, j, k)] +
                                           (dt * kpi + (noise ? (double)noise[IDX(i, j, k)] : 0.0)));
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * rk2_step_table, complex field (phi1 + i phi2), float64.
 * simulation/latticeSimComplex_numba.py:2067-2215; Z_N term :2147-2152 (zn_* are the
 * _ZN_PARAMS[0..2] globals, :227-231).
 * ------------------------------------------------------------------------------------------ */
static inline void complex_force(double p1, double p2, const double *table, double tmin,
                                 double dinv, int64_t npts, double zn_n, double zn_dv,
                                 double zn_active, double rho, double *dV1, double *dV2)
{
    double rho_safe = rho > 1e-30 ? rho : 1e-30;
    double dVdr = lookup(rho, table, tmin, dinv, npts, 0);
    double a = dVdr * p1 / rho_safe;
    double b = dVdr * p2 / rho_safe;
    if (zn_active > 0.5 && zn_n > 0.5 && rho > 1e-20) {
        double theta = atan2(p2, p1);
        double sn = sin(zn_n * theta);
        double inv_rho2 = (1.0 / rho_safe) * (1.0 / rho_safe);
        a += -zn_n * zn_dv * sn * p2 * inv_rho2;
        b += zn_n * zn_dv * sn * p1 * inv_rho2;
    }
    *dV1 = a;
    *dV2 = b;
}

/* The complex-field force (dV1, dV2) of complex_force at n points, exported so that its
 * Z_N term can be pinned on its own (finite differences of the potential, the reference's
 * C++ statement of the same force). */
ORC_API void orc_complex_force(const double *p1, const double *p2, int64_t n,
                               const double *table, double tmin, double dinv, int64_t npts,
                               double zn_n, double zn_dv, double zn_active, double *dV1,
                               double *dV2)
{
    for (int64_t i = 0; i < n; ++i) {
        double rho = sqrt(p1[i] * p1[i] + p2[i] * p2[i]);
        complex_force(p1[i], p2[i], table, tmin, dinv, npts, zn_n, zn_dv, zn_active, rho,
                      &dV1[i], &dV2[i]);
    }
}

ORC_API void orc_rk2_step_table_c64(double *phi1, double *phi2, double *pi1, double *pi2,
                                    double dt, double dx, double eta_eff, double mu,
                                    const double *noise1, const double *noise2, double *phi1_tmp,
                                    double *phi2_tmp, double *pi1_tmp, double *pi2_tmp,
                                    double inv_a2, const double *table, double tmin, double dinv,
                                    int64_t npts, double zn_n, double zn_dv, double zn_active,
                                    int64_t nx, int64_t ny, int64_t nz)
{
    const double inv_dx2 = 1.0 / (dx * dx);
    const double inv_mu2 = 1.0 / (mu * mu);
    const double half_dt = 0.5 * dt;
    for (int pass = 0; pass < 2; ++pass) {
        const double *q1 = pass == 0 ? phi1 : phi1_tmp;
        const double *q2 = pass == 0 ? phi2 : phi2_tmp;
        const double *r1 = pass == 0 ? pi1 : pi1_tmp;
        const double *r2 = pass == 0 ? pi2 : pi2_tmp;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t ip = i + 1 < nx ? i + 1 : 0, im = i - 1 >= 0 ? i - 1 : nx - 1;
            for (int64_t j = 0; j < ny; ++j) {
                int64_t jp = j + 1 < ny ? j + 1 : 0, jm = j - 1 >= 0 ? j - 1 : ny - 1;
                for (int64_t k = 0; k < nz; ++k) {
                    int64_t kp = k + 1 < nz ? k + 1 : 0, km = k - 1 >= 0 ? k - 1 : nz - 1;
                    int64_t s = IDX(i, j, k);
                    double p1 = q1[s], p2 = q2[s];
                    double rho = sqrt(p1 * p1 + p2 * p2);
                    double pv1 = r1[s], pv2 = r2[s];
                    double lap1 = (q1[IDX(ip, j, k)] + q1[IDX(im, j, k)] + q1[IDX(i, jp, k)] +
                                   q1[IDX(i, jm, k)] + q1[IDX(i, j, kp)] + q1[IDX(i, j, km)] -
                                   6.0 * p1) * inv_dx2;
                    double lap2 = (q2[IDX(ip, j, k)] + q2[IDX(im, j, k)] + q2[IDX(i, jp, k)] +
                                   q2[IDX(i, jm, k)] + q2[IDX(i, j, kp)] + q2[IDX(i, j, km)] -
                                   6.0 * p2) * inv_dx2;
                    double dV1, dV2;
                    complex_force(p1, p2, table, tmin, dinv, npts, zn_n, zn_dv, zn_active, rho,
                                  &dV1, &dV2);
                    double k1 = inv_a2 * lap1 - eta_eff * pv1 - dV1 * inv_mu2;
                    double k2 = inv_a2 * lap2 - eta_eff * pv2 - dV2 * inv_mu2;
                    if (pass == 0) {
                        phi1_tmp[s] = p1 + half_dt * pv1;
                        phi2_tmp[s] = p2 + half_dt * pv2;
                        pi1_tmp[s] = pv1 + half_dt * k1;
                        pi2_tmp[s] = pv2 + half_dt * k2;
                    } else {
                        phi1[s] += dt * pv1;
                        phi2[s] += dt * pv2;
                        pi1[s] += dt * k1 + (noise1 ? noise1[s] : 0.0);
                        pi2[s] += dt * k2 + (noise2 ? noise2[s] : 0.0);
                    }
                }
            }
        }
    }
}

/* Complex field, float32 storage with the reference's mixed precision.
 * latticeSimComplex_numba.py:2067-2215: with float32 p1, p2, Numba evaluates p1*p1 + p2*p2 and
 * math.sqrt in float32, the neighbour sums in float32, and the rest in float64. */
ORC_API void orc_rk2_step_table_c32mixed(float *phi1, float *phi2, float *pi1, float *pi2,
                                         double dt, double dx, double eta_eff, double mu,
                                         const float *noise1, const float *noise2,
                                         float *phi1_tmp, float *phi2_tmp, float *pi1_tmp,
                                         float *pi2_tmp, double inv_a2, const double *table,
                                         double tmin, double dinv, int64_t npts, double zn_n,
                                         double zn_dv, double zn_active, int64_t nx, int64_t ny,
                                         int64_t nz)
{
    const double inv_dx2 = 1.0 / (dx * dx);
    const double inv_mu2 = 1.0 / (mu * mu);
    const double half_dt = 0.5 * dt;
    for (int pass = 0; pass < 2; ++pass) {
        const float *q1 = pass == 0 ? phi1 : phi1_tmp;
        const float *q2 = pass == 0 ? phi2 : phi2_tmp;
        const float *r1 = pass == 0 ? pi1 : pi1_tmp;
        const float *r2 = pass == 0 ? pi2 : pi2_tmp;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t ip = i + 1 < nx ? i + 1 : 0, im = i - 1 >= 0 ? i - 1 : nx - 1;
            for (int64_t j = 0; j < ny; ++j) {
                int64_t jp = j + 1 < ny ? j + 1 : 0, jm = j - 1 >= 0 ? j - 1 : ny - 1;
                for (int64_t k = 0; k < nz; ++k) {
                    int64_t kp = k + 1 < nz ? k + 1 : 0, km = k - 1 >= 0 ? k - 1 : nz - 1;
                    int64_t s = IDX(i, j, k);
                    float p1 = q1[s], p2 = q2[s];
                    float rho = sqrtf(p1 * p1 + p2 * p2);
                    float pv1 = r1[s], pv2 = r2[s];
                    float s1 = q1[IDX(ip, j, k)] + q1[IDX(im, j, k)];
                    s1 = s1 + q1[IDX(i, jp, k)];
                    s1 = s1 + q1[IDX(i, jm, k)];
                    s1 = s1 + q1[IDX(i, j, kp)];
                    s1 = s1 + q1[IDX(i, j, km)];
                    float s2 = q2[IDX(ip, j, k)] + q2[IDX(im, j, k)];
                    s2 = s2 + q2[IDX(i, jp, k)];
                    s2 = s2 + q2[IDX(i, jm, k)];
                    s2 = s2 + q2[IDX(i, j, kp)];
                    s2 = s2 + q2[IDX(i, j, km)];
                    double lap1 = ((double)s1 - 6.0 * (double)p1) * inv_dx2;
                    double lap2 = ((double)s2 - 6.0 * (double)p2) * inv_dx2;
                    double dV1, dV2;
                    complex_force((double)p1, (double)p2, table, tmin, dinv, npts, zn_n, zn_dv,
                                  zn_active, (double)rho, &dV1, &dV2);
                    double k1 = inv_a2 * lap1 - eta_eff * (double)pv1 - dV1 * inv_mu2;
                    double k2 = inv_a2 * lap2 - eta_eff * (double)pv2 - dV2 * inv_mu2;
                    if (pass == 0) {
                        phi1_tmp[s] = (float)((double)p1 + half_dt * (double)pv1);
                        phi2_tmp[s] = (float)((double)p2 + half_dt * (double)pv2);
                        pi1_tmp[s] = (float)((double)pv1 + half_dt * k1);
                        pi2_tmp[s] = (float)((double)pv2 + half_dt * k2);
                    } else {
                        phi1[s] = (float)((double)phi1[s] + dt * (double)pv1);
                        phi2[s] = (float)((double)phi2[s] + dt * (double)pv2);
                        pi1[s] = (float)((double)pi1[s] + (dt * k1 + (noise1 ? (double)noise1[s] : 0.0)));
                        pi2[s] = (float)((double)pi2[s] + (dt * k2 + (noise2 ? (double)noise2[s] : 0.0)));
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * compute_winding_number, latticeSimComplex_numba.py:2683-2744 (restated in tools/winding.py:26-64).
 * Inputs are the float64 copies the driver passes (:4627-4629).  w_out receives the per-site
 * float32 sum of the XY, XZ and YZ plaquette windings as the reference writes it; plaq
 * (nullable) receives the three per-plaquette integers rint(sum/2pi) as int8, layout
 * [orientation][site] with orientation 0=XY, 1=XZ, 2=YZ.
 * ------------------------------------------------------------------------------------------ */
static inline double wrapd(double d) { return atan2(sin(d), cos(d)); }

ORC_API void orc_winding(const double *phi1, const double *phi2, float *w_out, int8_t *plaq,
                         int64_t nx, int64_t ny, int64_t nz)
{
    const double twopi = 2.0 * M_PI;
    int64_t n = nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nx; ++i) {
        int64_t ip = i + 1 < nx ? i + 1 : 0;
        for (int64_t j = 0; j < ny; ++j) {
            int64_t jp = j + 1 < ny ? j + 1 : 0;
            for (int64_t k = 0; k < nz; ++k) {
                int64_t kp = k + 1 < nz ? k + 1 : 0;
                double th00 = atan2(phi2[IDX(i, j, k)], phi1[IDX(i, j, k)]);
                double th10 = atan2(phi2[IDX(ip, j, k)], phi1[IDX(ip, j, k)]);
                double th11 = atan2(phi2[IDX(ip, jp, k)], phi1[IDX(ip, jp, k)]);
                double th01 = atan2(phi2[IDX(i, jp, k)], phi1[IDX(i, jp, k)]);
                double wxy = (wrapd(th10 - th00) + wrapd(th11 - th10) + wrapd(th01 - th11) +
                              wrapd(th00 - th01)) / twopi;
                double th1kp = atan2(phi2[IDX(ip, j, kp)], phi1[IDX(ip, j, kp)]);
                double th0kp = atan2(phi2[IDX(i, j, kp)], phi1[IDX(i, j, kp)]);
                double wxz = (wrapd(th10 - th00) + wrapd(th1kp - th10) + wrapd(th0kp - th1kp) +
                              wrapd(th00 - th0kp)) / twopi;
                double thj1kp = atan2(phi2[IDX(i, jp, kp)], phi1[IDX(i, jp, kp)]);
                double wyz = (wrapd(th01 - th00) + wrapd(thj1kp - th01) + wrapd(th0kp - thj1kp) +
                              wrapd(th00 - th0kp)) / twopi;
                double w_total = 0.0;
                w_total += wxy;
                w_total += wxz;
                w_total += wyz;
                int64_t s = IDX(i, j, k);
                if (w_out) w_out[s] = (float)w_total;
                if (plaq) {
                    plaq[s] = (int8_t)lrint(wxy);
                    plaq[n + s] = (int8_t)lrint(wxz);
                    plaq[2 * n + s] = (int8_t)lrint(wyz);
                }
            }
        }
    }
}

/* String count of the driver, latticeSimComplex_numba.py:4661: int(sum(|W| > 0.5)). */
ORC_API int64_t orc_string_count(const float *w, int64_t n)
{
    int64_t c = 0;
    for (int64_t s = 0; s < n; ++s) c += fabsf(w[s]) > 0.5f;
    return c;
}

/* ------------------------------------------------------------------------------------------
 * True-vacuum count: postprocess/compute_lattice_Tc1.py:10,24 counts |phi| <= escape as false
 * vacuum, so the true-vacuum count is #(|phi| > esc).  Complex amplitude form rho^2 > esc^2
 * (cosmolattice_ext/models/thermal_inflation.h:331-358).
 * ------------------------------------------------------------------------------------------ */
ORC_API int64_t orc_vacuum_count_f64(const double *phi, int64_t n, double esc)
{
    int64_t c = 0;
    for (int64_t s = 0; s < n; ++s) c += fabs(phi[s]) > esc;
    return c;
}

ORC_API int64_t orc_vacuum_count_f32(const float *phi, int64_t n, double esc)
{
    int64_t c = 0;
    for (int64_t s = 0; s < n; ++s) c += fabs((double)phi[s]) > esc;
    return c;
}

ORC_API int64_t orc_vacuum_count_c32(const float *phi1, const float *phi2, int64_t n, double esc)
{
    int64_t c = 0;
    double e2 = esc * esc;
    for (int64_t s = 0; s < n; ++s) {
        double a = phi1[s], b = phi2[s];
        c += a * a + b * b > e2;
    }
    return c;
}

ORC_API int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------------------------
 * Overdamped Euler-Maruyama and BAOAB, float64, race-free (Jacobi) form.
 *   overdamped_euler_step_table: real  simulation/latticeSimeRescale_numba.py:2115-2192,
 *                                complex simulation/latticeSimComplex_numba.py:2422-2517
 *   baoab_step_table:            real  simulation/latticeSimeRescale_numba.py:1974-2112,
 *                                complex simulation/latticeSimComplex_numba.py:2223-2419
 * The reference writes phi in place inside its parallel loop, so some neighbours in a site's
 * Laplacian are already advanced (schedule dependent).  Here every force is evaluated on one
 * field: the stencil reads phi, the result goes to phi_out.  Per-site arithmetic in the
 * reference's textual order:
 *   force = inv_a2 * lap - dV * inv_mu2
 *   overdamped: phi_out = ph + dt_over_eta * force + noise_scale * z          (:2192)
 *   BAOAB stage OPEN (1):  pi_v = pi + half_dt * force; ph_a = ph + half_dt * pi_v;  (:2058-2060)
 *                          pi_b = c1 * pi_v + c2 * z;                               (:2073)
 *                          phi_out = ph_a + half_dt * pi_b; pi_out = pi_b           (:2094)
 *   stage CLOSE (3):       pi_out = pi + half_dt * force_prev                       (:2112)
 *   stage FUSED (2):       CLOSE with force_prev, then OPEN on the result
 * force_prev uses inv_a2_prev (the previous step's scale factor).  c1 = exp(-gamma dt),
 * c2 = sqrt(1 - c1 c1) noise_scale (:2008-2009).  z: the hash pair of the site
 * (mult 73856093); the real field takes the cos member, the complex field (cos, sin).
 * ------------------------------------------------------------------------------------------ */
static inline void langevin_site(int op, double ph, double p, double F, double Fp, double z,
                                 double half_dt, double dt_over_eta, double c1, double c2,
                                 double ns, double *ph_out, double *p_out)
{
    if (op == 0) {
        *ph_out = ph + dt_over_eta * F + ns * z;
        return;
    }
    if (op == 2 || op == 3) p = p + half_dt * Fp;
    if (op == 3) {
        *p_out = p;
        *ph_out = ph;
        return;
    }
    double pv = p + half_dt * F;
    double pa = ph + half_dt * pv;
    double pb = c1 * pv + c2 * z;
    *ph_out = pa + half_dt * pb;
    *p_out = pb;
}

ORC_API void orc_langevin_f64(int op, const double *phi, const double *pi, double *phi_out,
                              double *pi_out, double dt, double dx, double eta_or_gamma,
                              double mu, double inv_a2, double inv_a2_prev, const double *table,
                              double tmin, double dinv, int64_t npts, int odd,
                              uint64_t noise_seed, double noise_scale, int64_t nx, int64_t ny,
                              int64_t nz)
{
    const double inv_dx2 = 1.0 / (dx * dx);
    const double inv_mu2 = 1.0 / (mu * mu);
    const double half_dt = 0.5 * dt;
    const double dt_over_eta = dt / eta_or_gamma;
    const double c1 = exp(-eta_or_gamma * dt);
    const double c2 = sqrt(1.0 - c1 * c1) * noise_scale;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nx; ++i) {
        int64_t ip = i + 1 < nx ? i + 1 : 0, im = i - 1 >= 0 ? i - 1 : nx - 1;
        for (int64_t j = 0; j < ny; ++j) {
            int64_t jp = j + 1 < ny ? j + 1 : 0, jm = j - 1 >= 0 ? j - 1 : ny - 1;
            for (int64_t k = 0; k < nz; ++k) {
                int64_t kp = k + 1 < nz ? k + 1 : 0, km = k - 1 >= 0 ? k - 1 : nz - 1;
                int64_t s = IDX(i, j, k);
                double ph = phi[s];
                double lap = (phi[IDX(ip, j, k)] + phi[IDX(im, j, k)] + phi[IDX(i, jp, k)] +
                              phi[IDX(i, jm, k)] + phi[IDX(i, j, kp)] + phi[IDX(i, j, km)] -
                              6.0 * ph) * inv_dx2;
                double dV = lookup(ph, table, tmin, dinv, npts, odd);
                double F = inv_a2 * lap - dV * inv_mu2;
                double Fp = inv_a2_prev * lap - dV * inv_mu2;
                double z = 0.0;
                if (op != 3 && noise_scale != 0.0) {
                    uint64_t x1, x2;
                    double u1, u2, zs;
                    hash_pair(site_seed(s, 73856093, noise_seed), &x1, &x2, &u1, &u2);
                    box_muller(u1, u2, &z, &zs);
                }
                double a, p;
                langevin_site(op, ph, op == 0 ? 0.0 : pi[s], F, Fp, z, half_dt, dt_over_eta,
                              c1, c2, noise_scale, &a, &p);
                if (op != 3) phi_out[s] = a;
                if (op != 0) pi_out[s] = p;
            }
        }
    }
}

ORC_API void orc_langevin_c64(int op, const double *phi1, const double *phi2, const double *pi1,
                              const double *pi2, double *phi1_out, double *phi2_out,
                              double *pi1_out, double *pi2_out, double dt, double dx,
                              double eta_or_gamma, double mu, double inv_a2, double inv_a2_prev,
                              const double *table, double tmin, double dinv, int64_t npts,
                              double zn_n, double zn_dv, double zn_active, uint64_t noise_seed,
                              double noise_scale, int64_t nx, int64_t ny, int64_t nz)
{
    const double inv_dx2 = 1.0 / (dx * dx);
    const double inv_mu2 = 1.0 / (mu * mu);
    const double half_dt = 0.5 * dt;
    const double dt_over_eta = dt / eta_or_gamma;
    const double c1 = exp(-eta_or_gamma * dt);
    const double c2 = sqrt(1.0 - c1 * c1) * noise_scale;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nx; ++i) {
        int64_t ip = i + 1 < nx ? i + 1 : 0, im = i - 1 >= 0 ? i - 1 : nx - 1;
        for (int64_t j = 0; j < ny; ++j) {
            int64_t jp = j + 1 < ny ? j + 1 : 0, jm = j - 1 >= 0 ? j - 1 : ny - 1;
            for (int64_t k = 0; k < nz; ++k) {
                int64_t kp = k + 1 < nz ? k + 1 : 0, km = k - 1 >= 0 ? k - 1 : nz - 1;
                int64_t s = IDX(i, j, k);
                double p1 = phi1[s], p2 = phi2[s];
                double rho = sqrt(p1 * p1 + p2 * p2);
                double lap1 = (phi1[IDX(ip, j, k)] + phi1[IDX(im, j, k)] + phi1[IDX(i, jp, k)] +
                               phi1[IDX(i, jm, k)] + phi1[IDX(i, j, kp)] + phi1[IDX(i, j, km)] -
                               6.0 * p1) * inv_dx2;
                double lap2 = (phi2[IDX(ip, j, k)] + phi2[IDX(im, j, k)] + phi2[IDX(i, jp, k)] +
                               phi2[IDX(i, jm, k)] + phi2[IDX(i, j, kp)] + phi2[IDX(i, j, km)] -
                               6.0 * p2) * inv_dx2;
                double dV1, dV2;
                complex_force(p1, p2, table, tmin, dinv, npts, zn_n, zn_dv, zn_active, rho, &dV1,
                              &dV2);
                double F1 = inv_a2 * lap1 - dV1 * inv_mu2, F2 = inv_a2 * lap2 - dV2 * inv_mu2;
                double F1p = inv_a2_prev * lap1 - dV1 * inv_mu2;
                double F2p = inv_a2_prev * lap2 - dV2 * inv_mu2;
                double z1 = 0.0, z2 = 0.0;
                if (op != 3 && noise_scale != 0.0) {
                    uint64_t x1, x2;
                    double u1, u2;
                    hash_pair(site_seed(s, 73856093, noise_seed), &x1, &x2, &u1, &u2);
                    box_muller(u1, u2, &z1, &z2);
                }
                double a, b, u, v;
                langevin_site(op, p1, op == 0 ? 0.0 : pi1[s], F1, F1p, z1, half_dt, dt_over_eta,
                              c1, c2, noise_scale, &a, &u);
                langevin_site(op, p2, op == 0 ? 0.0 : pi2[s], F2, F2p, z2, half_dt, dt_over_eta,
                              c1, c2, noise_scale, &b, &v);
                if (op != 3) {
                    phi1_out[s] = a;
                    phi2_out[s] = b;
                }
                if (op != 0) {
                    pi1_out[s] = u;
                    pi2_out[s] = v;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * CosmoLattice port's numba-compatible initial conditions,
 * cosmolattice_ext/models/thermal_inflation.h:174-222 (globalSiteIndex, hashMix64,
 * hashGaussian, applyNumbaInitialConditions): component c of site s is
 * amp * hashGaussian(s ^ kIcSeed) (c = 0) or amp * hashGaussian(s ^ golden ^ kIcSeed) (c = 1).
 * ------------------------------------------------------------------------------------------ */
static inline uint64_t cl_mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}

static inline double cl_hash_gaussian(uint64_t seed)
{
    uint64_t h1 = cl_mix64(seed);
    uint64_t h2 = cl_mix64(seed ^ 0x9E3779B97F4A7C15ULL);
    double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
    double r = sqrt(-2.0 * log(u1 > 1e-300 ? u1 : 1e-300));
    return r * cos(6.28318530717958647692 * u2);
}

ORC_API void orc_cl_hash_ics(double *f0, double *f1, int64_t n, double amp)
{
    const uint64_t ic = 0x4E554D42412D4943ULL;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < n; ++s) {
        f0[s] = amp * cl_hash_gaussian((uint64_t)s ^ ic);
        if (f1) f1[s] = amp * cl_hash_gaussian((uint64_t)s ^ 0x9E3779B97F4A7C15ULL ^ ic);
    }
}
