/*
 * tip.h -- C-ABI of libtip, the MI355X (gfx950) implementation of the 3D stochastic Langevin
 * RK2 lattice step of thermal-inflation-phase-transition and its in-loop observables.
 *
 * Conventions
 *  - Every entry point returns TIP_OK (0) or a negative TIP_ERR_* code; tip_last_error()
 *    returns the message of the last failure on the calling thread.
 *  - Field pointers are DEVICE pointers (allocate with tip_malloc or any HIP allocator).
 *    Fields are C-contiguous (nx, ny, nz) arrays, z fastest, exactly the reference's numpy
 *    layout.  `stream` is a hipStream_t (NULL = the default stream); calls are asynchronous
 *    with respect to the host unless stated otherwise.
 *  - Library scratch (the float32 drop-ins' prepared segment table, reductions, labelling)
 *    is one buffer per device and purpose.  A call on another stream than the buffer's last
 *    user waits on the device for that use to end, so calls on several streams are safe;
 *    calls from several host threads at once must be serialised by the caller.
 *  - Scalars follow the reference's meaning and units (rescaled lattice units: dx = mu*dx_phys,
 *    dt = mu*dt_phys, eta_eff, mu, inv_a2 = 1/a^2).
 *  - Noise: a non-NULL `noise` array is added as the reference adds its pre-generated field
 *    (pi += dt*k2 + noise).  With noise == NULL the library draws the reference's hash noise
 *    inline: noise = noise_scale * z, z = sqrt(-2 ln u1) cos(2 pi u2), (u1, u2) from
 *    _hash_rng_pair(uint64(site*73856093) ^ noise_seed).  The reference passes
 *    noise_seed = uint64(step) (simulation/latticeSimeRescale_numba.py:3738); tip_noise_seed()
 *    folds an ensemble seed into it so that seed 0 reproduces the reference stream.
 *  - V' tables are float64 arrays of npts values on x_i = tmin + i/dinv, looked up exactly as
 *    the reference does (truncating index clamped to [0, npts-2], unclamped linear fraction).
 *    odd != 0 marks a table tabulated on x >= 0 that is read at |phi| with the sign restored.
 */
#ifndef TIP_H
#define TIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The library is compiled with hidden visibility; only the declarations below are exported. */
#if defined(TIP_BUILDING) && defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

#define TIP_OK 0
#define TIP_ERR_ARG -1      /* invalid argument (shape, NULL pointer, range) */
#define TIP_ERR_HIP -2      /* a HIP runtime call failed */
#define TIP_ERR_NOMEM -3    /* device allocation failed */
#define TIP_ERR_STATE -4    /* object used in the wrong state */

/* ---- library / memory plumbing ------------------------------------------------------- */
int tip_version(int *major, int *minor);
int tip_last_error(char *buf, size_t len);
int tip_device_count(int *count);
int tip_set_device(int device);
int tip_malloc(void **ptr, size_t bytes);
int tip_free(void *ptr);
int tip_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int tip_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int tip_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int tip_memset(void *ptr, int value, size_t bytes, void *stream);
int tip_stream_create(void **stream);
int tip_stream_destroy(void *stream);
int tip_stream_synchronize(void *stream);
int tip_device_synchronize(void);

/* Per-step noise seed: step ^ (seed * 0x9E3779B97F4A7C15).  seed = 0 gives uint64(step), the
 * reference's noise_seed (simulation/latticeSimeRescale_numba.py:3738). */
uint64_t tip_noise_seed(uint64_t step, uint64_t seed);

/* ---- hash RNG (bit-exact with _hash_rng_pair) ----------------------------------------
 * Replaces: _hash_rng_pair, simulation/latticeSimeRescale_numba.py:2322-2336
 * (same text latticeSimComplex_numba.py:2747-2761).
 * For n device seeds writes the mixer states x1, x2 and u1, u2 (u1 clamped at 1e-12), and
 * z = sqrt(-2 ln u1) cos(2 pi u2), zs = ... sin(...) in float64.  Any output may be NULL. */
int tip_hash_pairs(const uint64_t *seeds, int64_t n, uint64_t *x1, uint64_t *x2, double *u1,
                   double *u2, double *z, double *zs, void *stream);

/* Hash noise field of the fused kernels written to an array:
 * noise1[site] = scale * r cos(2 pi u2), noise2 (nullable) = scale * r sin(2 pi u2),
 * seed uint64(site*mult) ^ noise_seed.  mult = 73856093 (pass 2) or 19349669 (pass 4) in
 * the reference (latticeSimeRescale_numba.py:2511-2522,2636-2647).
 * Replaces: generate_noise_field(noise, scale, seed), latticeSimeRescale_numba.py:1348-1361,
 * whose numpy per-thread stream is not reproducible; this is the deterministic stream. */
int tip_noise_field_f64(double *noise1, double *noise2, int64_t nx, int64_t ny, int64_t nz,
                        uint64_t noise_seed, int64_t mult, double scale, void *stream);
int tip_noise_field_f32(float *noise1, float *noise2, int64_t nx, int64_t ny, int64_t nz,
                        uint64_t noise_seed, int64_t mult, double scale, void *stream);

/* ---- fused RK2 Langevin step (drop-in) -----------------------------------------------
 * Replaces: rk2_step_table(phi, pi, dt, dx, eta_eff, mu, noise, phi_tmp, pi_tmp, inv_a2,
 *           vp_table, vp_tmin, vp_dinv, vp_npts)
 *           simulation/latticeSimeRescale_numba.py:1875-1966 (caller :3752-3768).
 * phi and pi are updated in place; phi_tmp and pi_tmp are caller scratch (clobbered) as in
 * the reference.  One fused kernel computes both RK2 stages from the 7-point stencil of
 * (phi, pi), writes the result to (phi_tmp, pi_tmp), and the result is copied back.
 * _f64: float64 fields and arithmetic in the reference's operation order.
 * _f32: float32 fields, float32 arithmetic (float32 hash-noise transcendentals; u1, u2 exact).
 * vac_count (nullable, device int64): += #sites with |phi_new| > vac_esc. */
int tip_rk2_step_table_f64(double *phi, double *pi, double dt, double dx, double eta_eff,
                           double mu, const double *noise, double *phi_tmp, double *pi_tmp,
                           double inv_a2, const double *vp_table, double vp_tmin,
                           double vp_dinv, int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny,
                           int64_t nz, uint64_t noise_seed, double noise_scale,
                           unsigned long long *vac_count, double vac_esc, void *stream);
int tip_rk2_step_table_f32(float *phi, float *pi, double dt, double dx, double eta_eff,
                           double mu, const float *noise, float *phi_tmp, float *pi_tmp,
                           double inv_a2, const double *vp_table, double vp_tmin,
                           double vp_dinv, int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny,
                           int64_t nz, uint64_t noise_seed, double noise_scale,
                           unsigned long long *vac_count, double vac_esc, void *stream);

/* Allocation-free variant of tip_rk2_step_table_f32 (same arguments, same results).
 * No counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The float32 segment table goes to a static device array of at most 65537 table points
 * (512 KiB per device, held until the process ends) instead of library scratch; vp_npts
 * above that returns TIP_ERR_ARG.  No hipMalloc/hipFree is called.  All calls on one device
 * share that array: calls that overlap on one device race. */
int tip_rk2_step_table_f32_noalloc(float *phi, float *pi, double dt, double dx,
                                   double eta_eff, double mu, const float *noise,
                                   float *phi_tmp, float *pi_tmp, double inv_a2,
                                   const double *vp_table, double vp_tmin, double vp_dinv,
                                   int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny,
                                   int64_t nz, uint64_t noise_seed, double noise_scale,
                                   unsigned long long *vac_count, double vac_esc, void *stream);

/* Out-of-place fused step (the performance form: ping-pong buffers, no copy back).
 * (phi, pi) -> (phi_out, pi_out); the input and output buffers must not overlap.
 * table_ab is a float32 segment table from tip_table_prepare_f32 (npts-1 float2 entries). */
int tip_rk2_fused_f64(const double *phi, const double *pi, double *phi_out, double *pi_out,
                      double dt, double dx, double eta_eff, double mu, double inv_a2,
                      const double *noise, const double *vp_table, double vp_tmin,
                      double vp_dinv, int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny,
                      int64_t nz, uint64_t noise_seed, double noise_scale,
                      unsigned long long *vac_count, double vac_esc, void *stream);
int tip_rk2_fused_f32(const float *phi, const float *pi, float *phi_out, float *pi_out,
                      double dt, double dx, double eta_eff, double mu, double inv_a2,
                      const float *noise, const void *table_ab, double vp_tmin,
                      double vp_dinv, int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny,
                      int64_t nz, uint64_t noise_seed, double noise_scale,
                      unsigned long long *vac_count, double vac_esc, void *stream);

/* float64 table (device, npts values) -> float32 segment table (device, (npts-1) float2:
 * V'(x) = A_i + B_i x on segment i).  Buffer size: 8*(npts-1) bytes. */
int tip_table_prepare_f32(const double *vp_table, int64_t npts, double vp_tmin, double vp_dinv,
                          void *table_ab, void *stream);

/* ---- complex field (phi1 + i phi2) ---------------------------------------------------
 * Replaces: rk2_step_table(phi1, phi2, pi1, pi2, dt, dx, eta_eff, mu, noise1, noise2,
 *           phi1_tmp, phi2_tmp, pi1_tmp, pi2_tmp, inv_a2, vp_table, vp_tmin, vp_dinv, vp_npts)
 *           simulation/latticeSimComplex_numba.py:2067-2215 (caller :4377-4398); the Z_N
 *           parameters are the _ZN_PARAMS global (:227-231) passed explicitly.
 * Inline hash noise (noise1 == NULL): one hash per site gives (cos, sin) components, as the
 * reference's fused complex kernels (latticeSimComplex_numba.py:2961-2974).
 * vac_count counts rho^2 > vac_esc^2. */
int tip_rk2_step_table_c64(double *phi1, double *phi2, double *pi1, double *pi2, double dt,
                           double dx, double eta_eff, double mu, const double *noise1,
                           const double *noise2, double *phi1_tmp, double *phi2_tmp,
                           double *pi1_tmp, double *pi2_tmp, double inv_a2,
                           const double *vp_table, double vp_tmin, double vp_dinv,
                           int64_t vp_npts, double zn_n, double zn_dv, double zn_active,
                           int64_t nx, int64_t ny, int64_t nz, uint64_t noise_seed,
                           double noise_scale, unsigned long long *vac_count, double vac_esc,
                           void *stream);
int tip_rk2_step_table_c32(float *phi1, float *phi2, float *pi1, float *pi2, double dt,
                           double dx, double eta_eff, double mu, const float *noise1,
                           const float *noise2, float *phi1_tmp, float *phi2_tmp,
                           float *pi1_tmp, float *pi2_tmp, double inv_a2,
                           const double *vp_table, double vp_tmin, double vp_dinv,
                           int64_t vp_npts, double zn_n, double zn_dv, double zn_active,
                           int64_t nx, int64_t ny, int64_t nz, uint64_t noise_seed,
                           double noise_scale, unsigned long long *vac_count, double vac_esc,
                           void *stream);
int tip_rk2_fused_c32(const float *phi1, const float *phi2, const float *pi1, const float *pi2,
                      float *phi1_out, float *phi2_out, float *pi1_out, float *pi2_out,
                      double dt, double dx, double eta_eff, double mu, double inv_a2,
                      const float *noise1, const float *noise2, const void *table_ab,
                      double vp_tmin, double vp_dinv, int64_t vp_npts, double zn_n,
                      double zn_dv, double zn_active, int64_t nx, int64_t ny, int64_t nz,
                      uint64_t noise_seed, double noise_scale, unsigned long long *vac_count,
                      double vac_esc, void *stream);

/* ---- the reference's other two table integrators ---------------------------------------
 * Overdamped Euler-Maruyama.  Replaces: overdamped_euler_step_table(phi, dt, dx, eta_eff, mu,
 * inv_a2, vp_table, vp_tmin, vp_dinv, vp_npts, noise_scale, seed),
 * simulation/latticeSimeRescale_numba.py:2115-2192 (complex (phi1, phi2, ...):
 * simulation/latticeSimComplex_numba.py:2422-2517):
 *   phi_out = phi + (dt / eta_eff) (inv_a2 lap(phi) - V'(phi) / mu^2) + noise_scale z.
 * BAOAB.  Replaces: baoab_step_table(phi, pi, dt, dx, gamma, mu, inv_a2, vp_table, vp_tmin,
 * vp_dinv, vp_npts, noise_scale, seed), simulation/latticeSimeRescale_numba.py:1974-2112
 * (complex: simulation/latticeSimComplex_numba.py:2223-2419), as stages:
 *   TIP_BAOAB_OPEN   B A O A from the true momentum: phi_out, pi_out = post-O momentum
 *   TIP_BAOAB_FUSED  the closing half-kick of the previous step (with inv_a2_prev) then OPEN
 *   TIP_BAOAB_CLOSE  the closing half-kick only: pi_out = pi + dt/2 F(phi) (phi_out unused)
 * One step is OPEN then CLOSE; k steps are OPEN, FUSED x (k-1), CLOSE (k + 1 lattice passes).
 * O: pi = c1 pi + c2 z, c1 = exp(-gamma dt), c2 = sqrt(1 - c1^2) noise_scale.
 * Both evaluate every force on the input field (the reference updates phi in place inside its
 * parallel loop, so part of each Laplacian sees advanced neighbours); phi_out must not alias
 * phi, pi_out may alias pi.  Noise: the hash draw of site s = uint64(s * 73856093) ^
 * noise_seed (z = the cos member; complex (cos, sin)); noise_scale = 0 switches it off.
 * The reference's seed argument is the step index, so noise_seed = tip_noise_seed(step, 0).
 * vp_table: device float64 table (the float32 entry points convert it to segments in a
 * per-device scratch); vp_odd: real tables tabulated on phi >= 0 and read at |phi|. */
#define TIP_BAOAB_OPEN 1
#define TIP_BAOAB_FUSED 2
#define TIP_BAOAB_CLOSE 3
int tip_overdamped_f32(const float *phi, float *phi_out, double dt, double dx, double eta_eff,
                       double mu, double inv_a2, const double *vp_table, double vp_tmin,
                       double vp_dinv, int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny,
                       int64_t nz, uint64_t noise_seed, double noise_scale, void *stream);
int tip_overdamped_f64(const double *phi, double *phi_out, double dt, double dx, double eta_eff,
                       double mu, double inv_a2, const double *vp_table, double vp_tmin,
                       double vp_dinv, int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny,
                       int64_t nz, uint64_t noise_seed, double noise_scale, void *stream);
int tip_baoab_f32(int stage, const float *phi, const float *pi, float *phi_out, float *pi_out,
                  double dt, double dx, double gamma, double mu, double inv_a2,
                  double inv_a2_prev, const double *vp_table, double vp_tmin, double vp_dinv,
                  int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny, int64_t nz,
                  uint64_t noise_seed, double noise_scale, void *stream);
int tip_baoab_f64(int stage, const double *phi, const double *pi, double *phi_out,
                  double *pi_out, double dt, double dx, double gamma, double mu, double inv_a2,
                  double inv_a2_prev, const double *vp_table, double vp_tmin, double vp_dinv,
                  int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny, int64_t nz,
                  uint64_t noise_seed, double noise_scale, void *stream);
int tip_overdamped_c32(const float *phi1, const float *phi2, float *phi1_out, float *phi2_out,
                       double dt, double dx, double eta_eff, double mu, double inv_a2,
                       const double *vp_table, double vp_tmin, double vp_dinv, int64_t vp_npts,
                       double zn_n, double zn_dv, double zn_active, int64_t nx, int64_t ny,
                       int64_t nz, uint64_t noise_seed, double noise_scale, void *stream);
int tip_overdamped_c64(const double *phi1, const double *phi2, double *phi1_out,
                       double *phi2_out, double dt, double dx, double eta_eff, double mu,
                       double inv_a2, const double *vp_table, double vp_tmin, double vp_dinv,
                       int64_t vp_npts, double zn_n, double zn_dv, double zn_active, int64_t nx,
                       int64_t ny, int64_t nz, uint64_t noise_seed, double noise_scale,
                       void *stream);
int tip_baoab_c32(int stage, const float *phi1, const float *phi2, const float *pi1,
                  const float *pi2, float *phi1_out, float *phi2_out, float *pi1_out,
                  float *pi2_out, double dt, double dx, double gamma, double mu, double inv_a2,
                  double inv_a2_prev, const double *vp_table, double vp_tmin, double vp_dinv,
                  int64_t vp_npts, double zn_n, double zn_dv, double zn_active, int64_t nx,
                  int64_t ny, int64_t nz, uint64_t noise_seed, double noise_scale,
                  void *stream);
int tip_baoab_c64(int stage, const double *phi1, const double *phi2, const double *pi1,
                  const double *pi2, double *phi1_out, double *phi2_out, double *pi1_out,
                  double *pi2_out, double dt, double dx, double gamma, double mu, double inv_a2,
                  double inv_a2_prev, const double *vp_table, double vp_tmin, double vp_dinv,
                  int64_t vp_npts, double zn_n, double zn_dv, double zn_active, int64_t nx,
                  int64_t ny, int64_t nz, uint64_t noise_seed, double noise_scale,
                  void *stream);

/* ---- observables ---------------------------------------------------------------------
 * Winding.  Replaces: compute_winding_number(phi1, phi2, winding_out),
 * simulation/latticeSimComplex_numba.py:2683-2744 (numpy restatement tools/winding.py:26-64),
 * and the string count int(sum(|W| > 0.5)) of :4661.
 * Per site and plaquette orientation (0 = XY, 1 = XZ, 2 = YZ, corner order as the reference)
 * the integer winding sum(wrap(dtheta))/2pi.  Outputs (all nullable, device):
 *   plaq   int8 [3][nx*ny*nz]  per-plaquette windings
 *   wsum   float [nx*ny*nz]    per-site sum (the reference's winding_out)
 *   counts uint64 [2]          += {#sites with |wsum| > 0.5, #plaquettes with winding != 0}
 * The integers are computed exactly from the signs of the float32 field (the phase
 * difference crosses the branch cut iff the quadrant test and the sign of the exact float64
 * cross product say so); edges whose cross product is exactly zero (zero vectors with signed
 * zeros, the two sides of the negative real axis, antiparallel pairs) take the reference's
 * float64 atan2 expression.  They agree with the reference's arithmetic except where a phase
 * difference of a non-degenerate edge lies within rounding of +-pi. */
int tip_winding_f32(const float *phi1, const float *phi2, int8_t *plaq, float *wsum,
                    unsigned long long *counts, int64_t nx, int64_t ny, int64_t nz,
                    void *stream);
int tip_winding_f64(const double *phi1, const double *phi2, int8_t *plaq, float *wsum,
                    unsigned long long *counts, int64_t nx, int64_t ny, int64_t nz,
                    void *stream);

/* Connected-component labelling of {|v| > threshold} (the comparison in v's precision), as
 * scipy.ndimage.label with its default 6-connectivity structure numbers them for the bubble
 * (|phi| > esc) and string (|W| > 0.5) analysis of postprocess/revisualize_snapshots.py:
 * 171-213 (_identify_bubbles) and :360-392 (_identify_strings).  labels (device int32 [N]):
 * 0 = background, components 1..n in raster order of their first site (scipy's numbering);
 * *n_components (host) = n.  periodic != 0 also joins across the lattice faces (the reference's
 * analysis does not).  Synchronous on the stream; uses a per-device scratch of 4 N bytes. */
int tip_label_f32(const float *v, double threshold, int64_t nx, int64_t ny, int64_t nz,
                  int periodic, int32_t *labels, int64_t *n_components, void *stream);
int tip_label_f64(const double *v, double threshold, int64_t nx, int64_t ny, int64_t nz,
                  int periodic, int32_t *labels, int64_t *n_components, void *stream);
/* Per-component sums over the labelled sites (device outputs, n_components entries each,
 * sum_xyz 3 per component): volume, sum of x, y, z (site coordinates), sum of x^2 + y^2 + z^2
 * (exact integers), sum of v (float64) and max |v| (float64 bit patterns).  From these the
 * host forms the reference's centroid, r_eff = (3V/4pi)^(1/3), r_rms, sign of the mean and
 * peak (revisualize_snapshots.py:189-205, :372-388). */
int tip_label_stats_f32(const float *v, const int32_t *labels, int64_t n_components, int64_t nx,
                        int64_t ny, int64_t nz, unsigned long long *volume,
                        unsigned long long *sum_xyz, unsigned long long *sum_sq, double *sum_v,
                        unsigned long long *peak_bits, void *stream);
int tip_label_stats_f64(const double *v, const int32_t *labels, int64_t n_components,
                        int64_t nx, int64_t ny, int64_t nz, unsigned long long *volume,
                        unsigned long long *sum_xyz, unsigned long long *sum_sq, double *sum_v,
                        unsigned long long *peak_bits, void *stream);

/* The CosmoLattice port's numba-compatible initial field (applyNumbaInitialConditions,
 * cosmolattice_ext/models/thermal_inflation.h:174-222): fld0[s] = amp * z(s ^ kIcSeed),
 * fld1[s] (nullable) = amp * z(s ^ 0x9E3779B97F4A7C15 ^ kIcSeed), kIcSeed = 0x4E554D42412D4943,
 * z = the port's hashGaussian of the splitmix64 finaliser, s = (ix ny + iy) nz + iz; computed
 * in float64 and rounded to the output type. */
int tip_cl_hash_ics_f32(float *fld0, float *fld1, int64_t nx, int64_t ny, int64_t nz,
                        double amp, void *stream);
int tip_cl_hash_ics_f64(double *fld0, double *fld1, int64_t nx, int64_t ny, int64_t nz,
                        double amp, void *stream);

/* True-vacuum count: += #sites with |phi| > esc (real) or phi1^2 + phi2^2 > esc^2 (complex).
 * Replaces the off-line mean(|phi| > esc) of postprocess/compute_lattice_Tc1.py:24 and the
 * falseVacuumFraction reduction of cosmolattice_ext/models/thermal_inflation.h:331-358. */
int tip_vacuum_count_f32(const float *phi, int64_t n, double esc, unsigned long long *count,
                         void *stream);
int tip_vacuum_count_f64(const double *phi, int64_t n, double esc, unsigned long long *count,
                         void *stream);
int tip_vacuum_count_c32(const float *phi1, const float *phi2, int64_t n, double esc,
                         unsigned long long *count, void *stream);

/* min / max of a field (double[2] device output {min, max}); replaces the per-step host scans
 * phi.min()/phi.max() of the table-range check, latticeSimeRescale_numba.py:3655-3656, and
 * max(rho) of latticeSimComplex_numba.py:4268 (tip_rho_max_c32). */
int tip_minmax_f32(const float *phi, int64_t n, double *out2, void *stream);
int tip_minmax_f64(const double *phi, int64_t n, double *out2, void *stream);
/* Snapshot quantities of the complex driver (latticeSimComplex_numba.py:4625-4626, 4638-4646):
 * rho = sqrt(phi1^2 + phi2^2) and theta = atan2(phi2, phi1), both formed in float64 from the
 * float32 field and rounded to float32 (rho, theta: device float[n], nullable), and
 * {min, max} of the float64 rho (rho_minmax: device double[2], nullable). */
int tip_rho_theta_f32(const float *phi1, const float *phi2, int64_t n, float *rho, float *theta,
                      double *rho_minmax, void *stream);
int tip_rho_max_c32(const float *phi1, const float *phi2, int64_t n, double *out1,
                    void *stream);

/* ---- device-resident lattice (the performance path) -----------------------------------
 * Owns ping-pong field buffers in HBM, the V' table (float64 + float32 segments), a stream and
 * HIP events, so state stays on the device between save points and each step is one kernel
 * launch with no copy.  Replaces the reference's main loop body around rk2_step_table
 * (simulation/latticeSimeRescale_numba.py:3619-3768; complex latticeSimComplex_numba.py:
 * 4237-4398) for a fixed parameter set over nsteps; the host calls it once per schedule
 * update (T / Hubble, latticeSimeRescale_numba.py:3622-3647).
 * field_kind: TIP_REAL (phi, pi) or TIP_COMPLEX (phi1, phi2, pi1, pi2); precision 4 or 8
 * (complex: 4 only).  Field index `which`: real 0 = phi, 1 = pi; complex 0 = phi1, 1 = pi1,
 * 2 = phi2, 3 = pi2.  Step n of a call uses noise_seed = tip_noise_seed(step0 + n, seed). */
#define TIP_REAL 0
#define TIP_COMPLEX 1

/* ---- V' table producer ----------------------------------------------------------------
 * Replaces: build_vprime_table(T, phi_min, phi_max, n) -> (table, tmin, dinv) with
 *           Vprime_scalar(phi, T), simulation/latticeSimeRescale_numba.py:805-849 (complex
 *           driver simulation/latticeSimComplex_numba.py:818-862).
 * table (device, n float64) [i] = V'(phi_i, T), phi_i = numpy.linspace(phi_min, phi_max, n)[i]:
 * V' = lam phi^3 - mphi^2 phi + T^4/(2 pi^2) (coeff_b dJ_b(u_b) du_b/dphi
 *      + coeff_f dJ_f(u_f) du_f/dphi), the dJ splines evaluated with the reference's clamp and
 * Horner rule.  cb, cf (device): the splines' piecewise-cubic coefficients, scipy
 * CubicSpline.c layout (4, nseg) row-major, rows = powers 3, 2, 1, 0, on knots x_min + i h
 * (:432-455; the lattice counterterm, :758-795, is folded into them by the caller).
 * coeff_b, coeff_f are _POT_COEFFS (convention x multiplicity, :231-241).  float64 arithmetic
 * in the reference's operation order.  tmin, dinv (host, nullable) receive the reference's
 * return values phi_min and 1 / ((phi_max - phi_min) / (n - 1)). */
typedef struct tip_vprime_model {
    double lam, mphi, mb2, yb, gb, yf, gf, coeff_b, coeff_f;
    double x_min, h; /* spline knots x_min + i h, i = 0..nseg */
    int64_t nseg;
} tip_vprime_model;

int tip_build_vprime_table(double *table, int64_t n, double T, double phi_min, double phi_max,
                           const tip_vprime_model *model, const double *cb, const double *cf,
                           double *tmin, double *dinv, void *stream);

typedef struct tip_step_params {
    double dt, dx, eta_eff, mu, inv_a2;
    double noise_scale; /* 0 disables the inline hash noise */
} tip_step_params;

typedef struct tip_lattice tip_lattice;

int tip_lattice_create(tip_lattice **out, int field_kind, int precision, int64_t nx, int64_t ny,
                       int64_t nz);
int tip_lattice_destroy(tip_lattice *lat);
int tip_lattice_stream(tip_lattice *lat, void **stream);
int tip_lattice_field_ptr(tip_lattice *lat, int which, void **ptr);
int tip_lattice_upload(tip_lattice *lat, int which, const void *host);
int tip_lattice_download(tip_lattice *lat, int which, void *host);
int tip_lattice_set_table(tip_lattice *lat, const double *host_table, int64_t npts,
                          double tmin, double dinv, int odd);
/* Build the lattice's V' table on the device (tip_build_vprime_table into the lattice's own
 * table buffer, then the float32 segments): no host evaluation and no upload.  Replaces the
 * host build + set_table of the driver's rebuild rule (latticeSimeRescale_numba.py:3650-3672);
 * the lattice records tmin = phi_min, dinv = (npts - 1) / (phi_max - phi_min). */
int tip_lattice_build_table(tip_lattice *lat, int64_t npts, double T, double phi_min,
                            double phi_max, const tip_vprime_model *model, const double *cb,
                            const double *cf, int odd);
int tip_lattice_set_zn(tip_lattice *lat, double zn_n, double zn_dv, double zn_active);
/* Launch geometry of the float32 step kernels (the autotuner's knobs).  kernel:
 * TIP_KERNEL_AUTO (row tiles where they apply, else flat tiles), TIP_KERNEL_ROWS (whole z
 * rows staged in LDS, `rows` = y rows per thread, 1, 2 or 4, dividing ny; needs nz <= 1024) or
 * TIP_KERNEL_FLAT (`rows` = rows of 64 z-sites per block, 1, 2, 4 or 8).  float32 lattices;
 * the float64 drop-in has one geometry.  rows = 0 picks the default, planes_per_block = x
 * planes each block marches through (0 = derived), xcd_remap = XCD-aware block order (0/1). */
#define TIP_KERNEL_AUTO 0
#define TIP_KERNEL_ROWS 1
#define TIP_KERNEL_FLAT 2
int tip_lattice_set_geometry(tip_lattice *lat, int kernel, int rows, int planes_per_block,
                             int xcd_remap);
/* nsteps RK2 steps; elapsed_ms (nullable) receives the HIP-event time of the launches on the
 * lattice's stream and makes the call synchronous. */
int tip_lattice_step(tip_lattice *lat, const tip_step_params *p, int64_t nsteps, uint64_t step0,
                     uint64_t seed, float *elapsed_ms);
/* The overdamped and BAOAB integrators on the lattice (tip_overdamped_*, tip_baoab_*), with
 * eta_eff as eta (overdamped) or gamma (BAOAB) and noise_scale as the integrator's own
 * amplitude: overdamped sqrt(2 T dt / (eta_eff mu^2 dx^3)), BAOAB sqrt(T / (mu^2 dx^3))
 * (latticeSimComplex_numba.py:4293, 4313).  A BAOAB call of nsteps runs OPEN (or FUSED,
 * completing the previous call's last step) and FUSED x (nsteps - 1): one lattice pass per
 * step.  The closing half-kick of its last step is deferred: the next BAOAB call fuses it
 * (when dt, dx and mu agree; its inv_a2 stays that of the step it belongs to), and any
 * operation that needs the true pi or changes the force runs it first with the old
 * parameters -- tip_lattice_download / _field_ptr of a momentum, _upload, _set_table,
 * _set_zn, the other steppers and tip_lattice_finish.  Field reads of phi (vacuum count,
 * winding, min/max, rho/theta) do not.  The overdamped step leaves pi untouched. */
int tip_lattice_step_overdamped(tip_lattice *lat, const tip_step_params *p, int64_t nsteps,
                                uint64_t step0, uint64_t seed, float *elapsed_ms);
int tip_lattice_step_baoab(tip_lattice *lat, const tip_step_params *p, int64_t nsteps,
                           uint64_t step0, uint64_t seed, float *elapsed_ms);
/* Complete deferred work (the BAOAB closing half-kick) and synchronise the lattice stream. */
int tip_lattice_finish(tip_lattice *lat);
/* Synchronous observables on the current state. */
int tip_lattice_vacuum_count(tip_lattice *lat, double esc, unsigned long long *count);
/* plaq_dev (int8 [3][N]) and wsum_dev (float [N], the per-site winding density) are nullable
 * device outputs; counts2 (host) receives {string sites, nonzero plaquettes}. */
int tip_lattice_winding(tip_lattice *lat, int8_t *plaq_dev, float *wsum_dev,
                        unsigned long long *counts2);
/* rho, theta (device float[N], nullable) and {min, max} of rho (host double[2], nullable) of
 * the current complex field, as tip_rho_theta_f32. */
int tip_lattice_rho_theta(tip_lattice *lat, float *rho_dev, float *theta_dev,
                          double *rho_minmax_host);
int tip_lattice_minmax(tip_lattice *lat, double *out2);

/* ---- batched ensemble ---------------------------------------------------------------
 * n independent real float32 realisations of uneven sizes (dims = n x {nx, ny, nz}), each with
 * its own seed, V' table and step parameters, advanced together.  No counterpart in the
 * reference, which runs one realisation per process; each instance follows rk2_step_table
 * (simulation/latticeSimeRescale_numba.py:1875-1966) with hash noise
 * noise_seed = tip_noise_seed(step, seed_i).
 * Modes: TIP_ENS_SEQUENTIAL launches the single-lattice kernel once per instance per step
 * (the plain loop); TIP_ENS_BATCHED launches one kernel per step over a work list of every
 * instance's tiles; TIP_ENS_GRAPH captures the nsteps batched launches into a hipGraph once per
 * (nsteps, buffer parity) and replays it.  All modes give identical fields. */
#define TIP_ENS_SEQUENTIAL 0
#define TIP_ENS_BATCHED 1
#define TIP_ENS_GRAPH 2

typedef struct tip_ensemble tip_ensemble;

int tip_ensemble_create(tip_ensemble **out, int n, const int64_t *dims, const uint64_t *seeds);
int tip_ensemble_destroy(tip_ensemble *ens);
/* which: 0 = phi, 1 = pi; host float32 arrays of the instance's size */
int tip_ensemble_upload(tip_ensemble *ens, int inst, int which, const float *host);
int tip_ensemble_download(tip_ensemble *ens, int inst, int which, float *host);
int tip_ensemble_set_table(tip_ensemble *ens, int inst, const double *host_table, int64_t npts,
                           double tmin, double dinv, int odd);
int tip_ensemble_set_params(tip_ensemble *ens, int inst, const tip_step_params *p);
/* kernel TIP_KERNEL_ROWS (LDS row tiles, `rows` = y rows per thread, 1, 2 or 4 dividing every
 * instance's ny, 0 = default; needs every nz <= 256), TIP_KERNEL_FLAT (`rows` = rows of 64
 * z-sites per block, 1, 2, 4 or 8) or TIP_KERNEL_AUTO (row tiles where they apply); x planes
 * per work item, XCD-aware block remap, and persistent_blocks > 0 to launch that many blocks
 * each walking a contiguous item range (0 = one block per item). */
int tip_ensemble_set_geometry(tip_ensemble *ens, int kernel, int rows, int planes_per_block,
                              int xcd_remap, int persistent_blocks);
/* Step schedule of the batched modes.  The row kernel launches instances whose z extent
 * divides 256, those whose z extent divides 192 (rows packed densely in 3-wave blocks) and the
 * rest separately.  With group_bytes > 0 each such class of B bytes of field data (phi and
 * pi, both buffers: 16 B per site) is cut into ceil(B / group_bytes) groups of about equal
 * bytes (contiguous runs of its instances, a group holding one instance at least), and each
 * group runs all nsteps of a call before the next group starts, so a group's fields stay in
 * the GPU's last-level cache (Infinity Cache) from one step to the next.  0 (default) = one
 * launch per step per class.  Results are identical either way. */
int tip_ensemble_set_schedule(tip_ensemble *ens, int64_t group_bytes);
/* Groups stepping at once in the batched and graph modes: 1 (default) or 2 (groups alternate
 * between two streams, so two independent groups run concurrently and one group's launch
 * tails share the GPU with the other's blocks).  Results are identical either way. */
int tip_ensemble_set_concurrency(tip_ensemble *ens, int groups_at_once);
int tip_ensemble_step(tip_ensemble *ens, int64_t nsteps, int mode, float *elapsed_ms);
int tip_ensemble_vacuum_counts(tip_ensemble *ens, double esc, unsigned long long *host_counts);
int tip_ensemble_step_count(tip_ensemble *ens, uint64_t *step);
int tip_ensemble_set_step(tip_ensemble *ens, uint64_t step);

/* ---- learned V' surrogate -----------------------------------------------------------
 * No counterpart in the reference (which evaluates V' by host tabulation, Vprime_scalar /
 * build_vprime_table, simulation/latticeSimeRescale_numba.py:805-849, rebuilt whenever T moves
 * by 1e-4, :3650-3672).  A 2 -> 32 -> 32 -> 1 tanh MLP, float32, parameter vector of
 * TIP_MLP_NPARAMS floats laid out W1[32][2], b1[32], W2[32][32], b2[32], w3[32], b3.
 * Inputs: features x = (f1, f2), f1 = asinh(rho/rho_scale)/asinh(rho_max/rho_scale),
 * f2 = (T - T_mid)/T_half.
 * V'(rho, T) = rho * (lam rho^2 - m2 + out_scale * e(rho) * (g0(f2) + f1 * y)),
 * e = 1/(1 + (rho/rho_cut)^2), g0(f2) = g0[0] + g0[1] f2 + g0[2] f2^2 + g0[3] f2^3: the tree
 * term is exact, g0 is the thermal part of V'/rho at rho -> 0 (the effective mass, fitted to
 * the float64 potential; f1 = 0 there), and the network carries the rest of the thermal +
 * Coleman-Weinberg part of V'/rho over a fixed envelope. */
#define TIP_MLP_HIDDEN 32
#define TIP_MLP_NPARAMS 1185

typedef struct tip_mlp_features {
    double rho_scale, rho_max, T_mid, T_half, out_scale, lam, m2, rho_cut;
    double g0[4];
} tip_mlp_features;

/* y[i] = MLP(x[2i], x[2i+1]) */
int tip_mlp_forward(const float *params, const float *x, float *y, int64_t n, void *stream);
/* gradient of L = (1/2n) sum w (y - y_target)^2 over the n samples (device float[1185]) and
 * L (device float, nullable); weights w nullable (= 1) */
int tip_mlp_grad(const float *params, const float *x, const float *y_target,
                 const float *weights, int64_t n, float *grad, float *loss, void *stream);
/* one Adam step on a minibatch (t = 1, 2, ... for the bias corrections) */
int tip_mlp_adam_step(float *params, float *m, float *v, const float *x, const float *y_target,
                      const float *weights, int64_t n, double lr, double beta1, double beta2,
                      double eps, int64_t t, float *loss, void *stream);
/* V'(rho[i], T[i]) */
int tip_mlp_vprime(const float *params, const tip_mlp_features *fm, const float *rho,
                   const float *T, float *vp, int64_t n, void *stream);
/* On-device re-tabulation: table[i] = V'(x_min + i*(x_max-x_min)/(npts-1), T) (float64 out,
 * odd in x), the table the reference would rebuild on the host. */
int tip_mlp_tabulate(const float *params, const tip_mlp_features *fm, double T, double x_min,
                     double x_max, int64_t npts, double *table, void *stream);
/* Complex fused RK2 step with V'(rho, T) evaluated by the MLP at every site and stage instead
 * of the table (the table entry point is tip_rk2_fused_c32): row tiles with the 32 x 32 layer
 * on the matrix cores (half-precision head/tail products accumulated in float32); nz <= 512. */
int tip_rk2_fused_c32_mlp(const float *phi1, const float *phi2, const float *pi1,
                          const float *pi2, float *phi1_out, float *phi2_out, float *pi1_out,
                          float *pi2_out, double dt, double dx, double eta_eff, double mu,
                          double inv_a2, const float *mlp_params, const tip_mlp_features *fm,
                          double T, double zn_n, double zn_dv, double zn_active, int64_t nx,
                          int64_t ny, int64_t nz, uint64_t noise_seed, double noise_scale,
                          unsigned long long *vac_count, double vac_esc, void *stream);

/* nsteps complex-field steps of a device-resident lattice with V'(rho, T) from the surrogate,
 * T = T_start + dT_per_step * n at step n (continuous cooling, no re-tabulation from the
 * potential).  The noise amplitude p->noise_scale is the one at T_start; step n uses
 * noise_scale * sqrt(T_n / T_start), the T dependence of sqrt(2 eta T dt / (mu^2 dx^3)).
 * T must stay positive.  grid_points = 0: the network is evaluated inline at every site and
 * stage; grid_points >= 2: each step first evaluates the network at T_n on that many points of
 * its compressed rho coordinate (a device buffer the lattice keeps) and the step interpolates
 * them (same network and T(t); about 10x faster; nz <= 512). */
int tip_lattice_step_mlp(tip_lattice *lat, const tip_step_params *p, const float *mlp_params,
                         const tip_mlp_features *fm, double T_start, double dT_per_step,
                         int64_t nsteps, uint64_t step0, uint64_t seed, int64_t grid_points,
                         float *elapsed_ms);

#if defined(TIP_BUILDING) && defined(__GNUC__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif
#endif /* TIP_H */
