// langevin_kernels.h -- the reference's two other table integrators on the lattice layout of
// the RK2 step, written once for float64 and float32 (instantiated by tip_langevin_f64.hip,
// compiled with -ffp-contract=off, and tip_langevin_f32.hip):
//
//   overdamped Euler-Maruyama  overdamped_euler_step_table
//       real    simulation/latticeSimeRescale_numba.py:2115-2192
//       complex simulation/latticeSimComplex_numba.py:2422-2517
//     phi' = phi + (dt / eta) F(phi) + ns z,       F = inv_a2 lap(phi) - V'(phi) / mu^2
//
//   BAOAB                      baoab_step_table
//       real    simulation/latticeSimeRescale_numba.py:1974-2112
//       complex simulation/latticeSimComplex_numba.py:2223-2419
//     B: pi += dt/2 F(phi)   A: phi += dt/2 pi   O: pi = c1 pi + c2 z   A   B
//     c1 = exp(-gamma dt), c2 = sqrt(1 - c1^2) ns
//
// The reference updates phi in place inside its parallel loops, so a site's Laplacian reads
// neighbours of which some were already advanced (which ones depends on the thread schedule;
// its docstring calls this the in-place neighbour race).  These kernels evaluate every force
// on one consistent field (Jacobi form): the result is deterministic and independent of the
// launch geometry.  Noise: the hash draw of the RK2 path per site and step (the reference
// seeds np.random per call, a per-thread stream no other program reproduces); z = the cos
// member of the Box-Muller pair, the complex field takes (cos, sin) as the reference does.
//
// BAOAB as three stencil stages on the same kernel.  The last half-kick of step n and the
// first of step n+1 both evaluate F at phi_{n+1}; a run of k steps is
//   OPEN (B A O A from the true pi) , FUSED x (k-1) (B_prev B A O A) , CLOSE (B_prev)
// with one Laplacian and one table lookup per site and stage: k + 1 lattice passes against
// the reference's 2k, and between OPEN and CLOSE the stored pi is the post-O momentum (the
// value before the final half-kick).  B_prev uses the previous step's inv_a2.
#pragma once
#include "step_kernels.h"

namespace tip {

enum LangevinOp { LV_OVERDAMPED = 0, LV_BAOAB_OPEN = 1, LV_BAOAB_FUSED = 2, LV_BAOAB_CLOSE = 3 };

template <typename R>
struct LangevinScalars {
    R half_dt;      // BAOAB kicks and drifts
    R dt_over_eta;  // overdamped
    R inv_dx2, inv_mu2, inv_a2, inv_a2_prev;
    R c1, c2;  // BAOAB O step
    R ns;      // overdamped noise amplitude
    uint32_t word_mask;
};

template <typename R>
inline LangevinScalars<R> make_langevin(double dt, double dx, double eta_or_gamma, double mu,
                                        double inv_a2, double inv_a2_prev, double noise_scale)
{
    LangevinScalars<R> c;
    c.half_dt = (R)(0.5 * dt);
    c.dt_over_eta = (R)(dt / eta_or_gamma);
    c.inv_dx2 = (R)(1.0 / (dx * dx));
    c.inv_mu2 = (R)(1.0 / (mu * mu));
    c.inv_a2 = (R)inv_a2;
    c.inv_a2_prev = (R)inv_a2_prev;
    const double c1 = exp(-eta_or_gamma * dt);
    c.c1 = (R)c1;
    c.c2 = (R)(sqrt(1.0 - c1 * c1) * noise_scale);
    c.ns = (R)noise_scale;
    c.word_mask = 0xFFFFFFFFu;
    return c;
}

// Per-site update shared by the real and complex kernels (one component).  F and F_prev are
// the forces with this step's and the previous step's inv_a2; z the normal draw.
template <typename R, int OP>
__device__ __forceinline__ void langevin_site(const LangevinScalars<R>& c, R ph, R p, R F,
                                              R F_prev, R z, R& ph_out, R& p_out)
{
    if (OP == LV_OVERDAMPED) {
        ph_out = ph + c.dt_over_eta * F + c.ns * z;
        return;
    }
    if (OP == LV_BAOAB_FUSED || OP == LV_BAOAB_CLOSE) p = p + c.half_dt * F_prev;
    if (OP == LV_BAOAB_CLOSE) {
        p_out = p;
        return;
    }
    const R pv = p + c.half_dt * F;      // B
    const R pa = ph + c.half_dt * pv;    // A
    const R pb = c.c1 * pv + c.c2 * z;   // O
    ph_out = pa + c.half_dt * pb;        // A
    p_out = pb;
}

// One lane per (y, z) site of a 64*BY tile of the flattened plane, marching over the x planes
// [x0, x1) of its chunk with the x-1, x, x+1 values of its own column in registers.  pi is read
// and written at the lane's own site only, so pi_out may equal pi (the closing half-kick runs
// in place).
template <typename R, typename Tab, int OP, int BY>
__global__ __launch_bounds__(64 * BY) void k_langevin_real(const R* __restrict__ phi,
                                                           const R* pi, R* __restrict__ phi_out,
                                                           R* pi_out, LangevinScalars<R> c,
                                                           Tab tab, int nx, int ny, int nz,
                                                           TileGrid g, uint64_t noise_seed,
                                                           int hash_noise)
{
    int tp, ch;
    g.decode((int)blockIdx.x, (int)gridDim.x, tp, ch);
    const int x0 = ch * g.cx;
    const int x1 = min(x0 + g.cx, nx);
    const int nyz = ny * nz;
    const int pl = tp * (64 * BY) + (int)threadIdx.y * 64 + (int)threadIdx.x;
    if (pl >= nyz) return;  // no barrier or wave operation below
    const int jj = pl / nz, kk = pl - jj * nz;
    const int kp = kk + 1 < nz ? kk + 1 : 0, km = kk - 1 >= 0 ? kk - 1 : nz - 1;
    const int jp = jj + 1 < ny ? jj + 1 : 0, jm = jj - 1 >= 0 ? jj - 1 : ny - 1;
    const int o_c = pl, o_kp = jj * nz + kp, o_km = jj * nz + km;
    const int o_jp = jp * nz + kk, o_jm = jm * nz + kk;
    const int xm = x0 - 1 >= 0 ? x0 - 1 : nx - 1;
    R am = phi[xm * nyz + o_c];
    R ac = phi[x0 * nyz + o_c];
    for (int x = x0; x < x1; ++x) {
        const int xp = x + 1 < nx ? x + 1 : 0;
        const int base = x * nyz;
        const R ap = phi[xp * nyz + o_c];
        R s1 = ap + am;
        s1 = s1 + phi[base + o_jp];
        s1 = s1 + phi[base + o_jm];
        s1 = s1 + phi[base + o_kp];
        s1 = s1 + phi[base + o_km];
        const R lap = (s1 - (R)6 * ac) * c.inv_dx2;
        const R dV = tab(ac);
        const R F = c.inv_a2 * lap - dV * c.inv_mu2;
        const R Fp = c.inv_a2_prev * lap - dV * c.inv_mu2;
        const int s = base + o_c;
        R z = 0;
        if (OP != LV_BAOAB_CLOSE && hash_noise)
            z = gauss_of((R)0, site_seed((int64_t)s, noise_seed), c.word_mask);
        R ph_new = 0, p_new = 0;
        langevin_site<R, OP>(c, ac, OP == LV_OVERDAMPED ? (R)0 : pi[s], F, Fp, z, ph_new, p_new);
        if (OP != LV_BAOAB_CLOSE) phi_out[s] = ph_new;
        if (OP != LV_OVERDAMPED) pi_out[s] = p_new;
        am = ac;
        ac = ap;
    }
}

// Complex force of the two integrators (latticeSimComplex_numba.py:2286-2327, 2369-2412,
// 2475-2516):
// float64 in the reference's order (complex_force); float32 with the hardware square root and
// reciprocal of the RK2 row kernel (complex_force_f32).
template <typename Tab>
__device__ __forceinline__ void lv_complex_force(const Tab& tab, const ZnParams<double>& zn,
                                                 double a, double b, double& d1, double& d2)
{
    complex_force(tab, zn, a, b, d1, d2);
}
__device__ __forceinline__ void lv_complex_force(const TableF32<false>& tab,
                                                 const ZnParams<float>& zn, float a, float b,
                                                 float& d1, float& d2)
{
    TableRadial rad;
    rad.t = tab;
    complex_force_f32(rad, zn, a, b, d1, d2);
}

// Complex field phi1 + i phi2: V'(rho) from the radial table projected on phi / rho plus the
// Z_N term (latticeSimComplex_numba.py:2130-2152 and the same lines of the two integrators);
// noise (cos, sin) of one hash draw per site.
template <typename R, typename Tab, int OP, int BY>
__global__ __launch_bounds__(64 * BY) void k_langevin_complex(
    const R* __restrict__ phi1, const R* __restrict__ phi2, const R* pi1, const R* pi2,
    R* __restrict__ phi1_out, R* __restrict__ phi2_out, R* pi1_out, R* pi2_out,
    LangevinScalars<R> c, Tab tab, ZnParams<R> zn, int nx, int ny, int nz, TileGrid g,
    uint64_t noise_seed, int hash_noise)
{
    int tp, ch;
    g.decode((int)blockIdx.x, (int)gridDim.x, tp, ch);
    const int x0 = ch * g.cx;
    const int x1 = min(x0 + g.cx, nx);
    const int nyz = ny * nz;
    const int pl = tp * (64 * BY) + (int)threadIdx.y * 64 + (int)threadIdx.x;
    if (pl >= nyz) return;  // no barrier or wave operation below
    const int jj = pl / nz, kk = pl - jj * nz;
    const int kp = kk + 1 < nz ? kk + 1 : 0, km = kk - 1 >= 0 ? kk - 1 : nz - 1;
    const int jp = jj + 1 < ny ? jj + 1 : 0, jm = jj - 1 >= 0 ? jj - 1 : ny - 1;
    const int o_c = pl, o_kp = jj * nz + kp, o_km = jj * nz + km;
    const int o_jp = jp * nz + kk, o_jm = jm * nz + kk;
    const int xm = x0 - 1 >= 0 ? x0 - 1 : nx - 1;
    R am = phi1[xm * nyz + o_c], bm = phi2[xm * nyz + o_c];
    R ac = phi1[x0 * nyz + o_c], bc = phi2[x0 * nyz + o_c];
    for (int x = x0; x < x1; ++x) {
        const int xp = x + 1 < nx ? x + 1 : 0;
        const int base = x * nyz;
        const R ap = phi1[xp * nyz + o_c], bp = phi2[xp * nyz + o_c];
        R s1 = ap + am;
        s1 = s1 + phi1[base + o_jp];
        s1 = s1 + phi1[base + o_jm];
        s1 = s1 + phi1[base + o_kp];
        s1 = s1 + phi1[base + o_km];
        R s2 = bp + bm;
        s2 = s2 + phi2[base + o_jp];
        s2 = s2 + phi2[base + o_jm];
        s2 = s2 + phi2[base + o_kp];
        s2 = s2 + phi2[base + o_km];
        const R lap1 = (s1 - (R)6 * ac) * c.inv_dx2;
        const R lap2 = (s2 - (R)6 * bc) * c.inv_dx2;
        R dV1, dV2;
        lv_complex_force(tab, zn, ac, bc, dV1, dV2);
        const R F1 = c.inv_a2 * lap1 - dV1 * c.inv_mu2;
        const R F2 = c.inv_a2 * lap2 - dV2 * c.inv_mu2;
        const R F1p = c.inv_a2_prev * lap1 - dV1 * c.inv_mu2;
        const R F2p = c.inv_a2_prev * lap2 - dV2 * c.inv_mu2;
        const int s = base + o_c;
        R z1 = 0, z2 = 0;
        if (OP != LV_BAOAB_CLOSE && hash_noise)
            gauss_pair_of((R)0, site_seed((int64_t)s, noise_seed), c.word_mask, z1, z2);
        R u = 0, v = 0;
        if (OP != LV_OVERDAMPED) {
            u = pi1[s];
            v = pi2[s];
        }
        R a_new = 0, b_new = 0, u_new = 0, v_new = 0;
        langevin_site<R, OP>(c, ac, u, F1, F1p, z1, a_new, u_new);
        langevin_site<R, OP>(c, bc, v, F2, F2p, z2, b_new, v_new);
        if (OP != LV_BAOAB_CLOSE) {
            phi1_out[s] = a_new;
            phi2_out[s] = b_new;
        }
        if (OP != LV_OVERDAMPED) {
            pi1_out[s] = u_new;
            pi2_out[s] = v_new;
        }
        am = ac;
        bm = bc;
        ac = ap;
        bc = bp;
    }
}

// Host launchers (defined once per precision by the two translation units).
template <typename R, typename Tab>
int launch_langevin_real(int op, const R* phi, const R* pi, R* phi_out, R* pi_out,
                         const LangevinScalars<R>& c, const Tab& tab, int64_t nx, int64_t ny,
                         int64_t nz, uint64_t noise_seed, int hash, hipStream_t st)
{
    StepGeometry g;
    TileGrid tg;
    int nb;
    int rc = resolve_geometry(nullptr, nx, ny, nz, 0, 0, g, tg, nb);
    if (rc) return rc;
    const dim3 blk(64, 4);
#define TIP_LV_REAL(OPV)                                                                       \
    hipLaunchKernelGGL((k_langevin_real<R, Tab, OPV, 4>), dim3(nb), blk, 0, st, phi, pi,       \
                       phi_out, pi_out, c, tab, (int)nx, (int)ny, (int)nz, tg, noise_seed, hash)
    switch (op) {
        case LV_OVERDAMPED: TIP_LV_REAL(LV_OVERDAMPED); break;
        case LV_BAOAB_OPEN: TIP_LV_REAL(LV_BAOAB_OPEN); break;
        case LV_BAOAB_FUSED: TIP_LV_REAL(LV_BAOAB_FUSED); break;
        case LV_BAOAB_CLOSE: TIP_LV_REAL(LV_BAOAB_CLOSE); break;
        default: TIP_REQUIRE(false, "langevin: unknown stage %d", op);
    }
#undef TIP_LV_REAL
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

template <typename R, typename Tab>
int launch_langevin_complex(int op, const R* phi1, const R* phi2, const R* pi1, const R* pi2,
                            R* phi1_out, R* phi2_out, R* pi1_out, R* pi2_out,
                            const LangevinScalars<R>& c, const Tab& tab, const ZnParams<R>& zn,
                            int64_t nx, int64_t ny, int64_t nz, uint64_t noise_seed, int hash,
                            hipStream_t st)
{
    StepGeometry g;
    TileGrid tg;
    int nb;
    int rc = resolve_geometry(nullptr, nx, ny, nz, 1, 0, g, tg, nb);
    if (rc) return rc;
    const dim3 blk(64, 4);
#define TIP_LV_CPLX(OPV)                                                                       \
    hipLaunchKernelGGL((k_langevin_complex<R, Tab, OPV, 4>), dim3(nb), blk, 0, st, phi1, phi2, \
                       pi1, pi2, phi1_out, phi2_out, pi1_out, pi2_out, c, tab, zn, (int)nx,     \
                       (int)ny, (int)nz, tg, noise_seed, hash)
    switch (op) {
        case LV_OVERDAMPED: TIP_LV_CPLX(LV_OVERDAMPED); break;
        case LV_BAOAB_OPEN: TIP_LV_CPLX(LV_BAOAB_OPEN); break;
        case LV_BAOAB_FUSED: TIP_LV_CPLX(LV_BAOAB_FUSED); break;
        case LV_BAOAB_CLOSE: TIP_LV_CPLX(LV_BAOAB_CLOSE); break;
        default: TIP_REQUIRE(false, "langevin: unknown stage %d", op);
    }
#undef TIP_LV_CPLX
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

}  // namespace tip

namespace tip {

// Precision-specific launchers used by the C-ABI entry points and the lattice
// (tip_langevin_f32.hip, tip_langevin_f64.hip).
int langevin_real_f32(int op, const float* phi, const float* pi, float* phi_out, float* pi_out,
                      const LangevinScalars<float>& c, const SegTable& tab, int64_t nx,
                      int64_t ny, int64_t nz, uint64_t noise_seed, int hash, hipStream_t st);
int langevin_complex_f32(int op, const float* phi1, const float* phi2, const float* pi1,
                         const float* pi2, float* phi1_out, float* phi2_out, float* pi1_out,
                         float* pi2_out, const LangevinScalars<float>& c, const SegTable& tab,
                         const ZnParams<float>& zn, int64_t nx, int64_t ny, int64_t nz,
                         uint64_t noise_seed, int hash, hipStream_t st);
int langevin_real_f64(int op, const double* phi, const double* pi, double* phi_out,
                      double* pi_out, const LangevinScalars<double>& c, const double* table,
                      double tmin, double dinv, int64_t npts, int odd, int64_t nx, int64_t ny,
                      int64_t nz, uint64_t noise_seed, int hash, hipStream_t st);

}  // namespace tip
