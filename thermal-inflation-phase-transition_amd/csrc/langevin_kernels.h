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

// ---------------------------------------------------------------------------------------
// float32 row tiles for the two integrators (the default where the shape allows): the tile
// and LDS scheme of the RK2 row kernel, with only phi staged -- the force needs the stencil of
// phi and pi only at the site.  NF = 1 (real) or 2 (complex) stencil fields; a group of
// W = ceil(nz/64) waves covers whole z rows, each thread TY consecutive y rows; per plane the
// group writes its rows of phi to LDS (K = 2 NF TY + 1 words per z column: two buffers, odd
// stride) behind one barrier; z neighbours are LDS reads, y neighbours registers except the
// two halo rows, x neighbours the register window (x-1, x, x+1, and x+2 in flight).  Loads
// per site: phi at x+2, pi at x+1 (BAOAB) and 2/TY halo rows per field.
// ---------------------------------------------------------------------------------------
// ZN = false: the Z_N term compiled out (complex launches that switch it off)
template <int NF, int TY, int OP, bool ODD, bool HS = false, bool ZN = true>
__device__ __forceinline__ void lv_rows_tile(
    const float* const (&phi)[NF], const float* const (&pi)[NF], float* const (&phi_out)[NF],
    float* const (&pi_out)[NF], const LangevinScalars<float>& c, const TableF32<ODD>& tab,
    const ZnParams<float>& zn, int nx, int ny, int nz, int W, int tile, int x0, int x1,
    uint64_t noise_seed, int hash, float* __restrict__ lds)
{
    constexpr int K = 2 * NF * TY + 1;
    constexpr bool MOM = OP != LV_OVERDAMPED;  // reads and writes pi
    const int G = (int)blockDim.y / W;
    const int w = (int)threadIdx.y, grp = w / W, zw = w - grp * W;
    const int z = zw * 64 + (int)threadIdx.x;
    const bool zact = z < nz;
    const int zc = zact ? z : 0;
    const int zp = zc + 1 < nz ? zc + 1 : 0, zm = zc > 0 ? zc - 1 : nz - 1;
    const int y0 = (tile * G + grp) * TY;
    const bool act = zact && y0 < ny;
    const int yb = y0 < ny ? y0 : 0;
    float* Lg = lds + (size_t)grp * (64 * W) * K;
    float* Lz = Lg + z * K;
    const float* Lzp = Lg + zp * K;
    const float* Lzm = Lg + zm * K;
    // HS (two row groups per block holding consecutive y tiles): each group loads one halo row
    // of phi (group 0 the block's top one, group 1 the bottom one) and reads the other from the
    // neighbour group's staged rows
    const bool g0 = grp == 0;
    const float* Lh = Lz + (g0 ? 64 * W * K : (TY - 1) - 64 * W * K);

    const int nyz = ny * nz;
    const size_t plane = (size_t)nyz * 4u;
    const uint32_t pb = (uint32_t)plane;
    auto at = [&](const float* b, int x) {
        return make_rsrc(reinterpret_cast<const char*>(b) + (size_t)x * plane, pb);
    };
    const __amdgpu_buffer_rsrc_t T = make_rsrc(tab.ab, (uint32_t)(tab.last + 1) * 8u);
    TableRadial rad;  // complex tables are even (ODD = false); unused for the real field
    rad.t.ab = tab.ab;
    rad.t.tmin = tab.tmin;
    rad.t.dinv = tab.dinv;
    rad.t.off = tab.off;
    rad.t.last = tab.last;
    uint32_t off[TY], offs[TY];
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        off[r] = (uint32_t)((yb + r) * nz + zc) * 4u;
        offs[r] = act ? off[r] : 0xFFFFFFF0u;  // idle lanes store past the end (dropped)
    }
    const uint32_t offm = (uint32_t)((yb > 0 ? yb - 1 : ny - 1) * nz + zc) * 4u;
    const uint32_t offp = (uint32_t)((yb + TY < ny ? yb + TY : 0) * nz + zc) * 4u;
    const uint32_t offh = g0 ? offm : offp;  // HS: the group's one loaded halo row
    const int64_t col0 = (int64_t)yb * nz + zc;

    // phi window (x-1, x, x+1, x+2 in the four role slots), pi at x (and x+1 in flight), halo
    // rows y0-1 and y0+TY of phi for the current / next plane.
    float w0[NF][TY], w1[NF][TY], w2[NF][TY], w3[NF][TY];
    float m0[NF][TY], m1[NF][TY];
    float hA[NF][2], hB[NF][2];
    {
        const int xm = x0 - 1 >= 0 ? x0 - 1 : nx - 1;
        const int xq = x0 + 1 < nx ? x0 + 1 : 0;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                w0[f][r] = buf_load<float>(at(phi[f], xm), off[r]);
                w1[f][r] = buf_load<float>(at(phi[f], x0), off[r]);
                w2[f][r] = buf_load<float>(at(phi[f], xq), off[r]);
                w3[f][r] = 0.0f;
                m0[f][r] = MOM ? buf_load<float>(at(pi[f], x0), off[r]) : 0.0f;
                m1[f][r] = 0.0f;
            }
            if constexpr (HS) {
                hA[f][0] = buf_load<float>(at(phi[f], x0), offh);
                hA[f][1] = 0.0f;
            } else {
                hA[f][0] = buf_load<float>(at(phi[f], x0), offm);
                hA[f][1] = buf_load<float>(at(phi[f], x0), offp);
            }
            hB[f][0] = hB[f][1] = 0.0f;
        }
    }

    auto plane_step = [&](auto bufc, int x, const float (&pm)[NF][TY], const float (&pc)[NF][TY],
                          const float (&pp)[NF][TY], float (&pn)[NF][TY],
                          const float (&qc)[NF][TY], float (&qn)[NF][TY],
                          const float (&hc)[NF][2], float (&hn)[NF][2]) {
        constexpr int b = decltype(bufc)::value;
        // Complex field: the radii and table gathers of every row depend only on this plane's
        // phi; they are issued before the next plane's loads so that waiting for a gather
        // (vector memory operations complete in issue order) does not wait for the next plane.
        float rho[TY];
        TableRadial::Pre seg[TY];
        if constexpr (NF == 2) {
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                rho[r] = complex_rho(pc[0][r], pc[1][r]);
                seg[r] = rad.fetch(rho[r]);
            }
        }
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int r = 0; r < TY; ++r) Lz[(b * NF + f) * TY + r] = pc[f][r];
        if (x + 1 < x1) {  // the chunk's last plane has no next plane
            const int xa = x + 1 < nx ? x + 1 : 0;
            const int xb = xa + 1 < nx ? xa + 1 : 0;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
#pragma unroll
                for (int r = 0; r < TY; ++r) {
                    pn[f][r] = buf_load<float>(at(phi[f], xb), off[r]);
                    if (MOM) qn[f][r] = buf_load<float>(at(pi[f], xa), off[r]);
                }
                if constexpr (HS) {
                    hn[f][0] = buf_load<float>(at(phi[f], xa), offh);
                } else {
                    hn[f][0] = buf_load<float>(at(phi[f], xa), offm);
                    hn[f][1] = buf_load<float>(at(phi[f], xa), offp);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            float F[NF], Fp[NF];
            float lap[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const int o = (b * NF + f) * TY + r;
                float yp, ym;
                if constexpr (HS) {
                    const float hl = (r == 0 || r + 1 == TY) ? Lh[(b * NF + f) * TY] : 0.0f;
                    yp = r + 1 < TY ? pc[f][r + 1 < TY ? r + 1 : r] : (g0 ? hl : hc[f][0]);
                    ym = r > 0 ? pc[f][r > 0 ? r - 1 : 0] : (g0 ? hc[f][0] : hl);
                } else {
                    yp = r + 1 < TY ? pc[f][r + 1 < TY ? r + 1 : r] : hc[f][1];
                    ym = r > 0 ? pc[f][r > 0 ? r - 1 : 0] : hc[f][0];
                }
                float s1 = pp[f][r] + pm[f][r];
                s1 = s1 + yp;
                s1 = s1 + ym;
                s1 = s1 + Lzp[o];
                s1 = s1 + Lzm[o];
                lap[f] = (s1 - 6.0f * pc[f][r]) * c.inv_dx2;
            }
            float dV[NF];
            if constexpr (NF == 1) {
                dV[0] = seg_eval<ODD>(seg_load<ODD>(T, tab, pc[0][r]), pc[0][r]);
            } else {
                complex_force_finish<TableRadial, ZN>(rad, zn, pc[0][r], pc[NF - 1][r], rho[r],
                                                      seg[r], dV[0],
                                     dV[NF - 1]);
            }
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                F[f] = c.inv_a2 * lap[f] - dV[f] * c.inv_mu2;
                Fp[f] = c.inv_a2_prev * lap[f] - dV[f] * c.inv_mu2;
            }
            float zv[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) zv[f] = 0.0f;
            if (OP != LV_BAOAB_CLOSE && hash) {
                const uint64_t seed = site_seed((int64_t)x * nyz + col0 + (int64_t)r * nz,
                                                noise_seed);
                if constexpr (NF == 1) zv[0] = gauss_of(0.0f, seed, c.word_mask);
                else gauss_pair_of(0.0f, seed, c.word_mask, zv[0], zv[NF - 1]);
            }
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                float a_new = 0.0f, p_new = 0.0f;
                langevin_site<float, OP>(c, pc[f][r], MOM ? qc[f][r] : 0.0f, F[f], Fp[f], zv[f],
                                         a_new, p_new);
                if (OP != LV_BAOAB_CLOSE) buf_store(a_new, at(phi_out[f], x), offs[r]);
                if (MOM) buf_store(p_new, at(pi_out[f], x), offs[r]);
            }
        }
    };
    using B0 = std::integral_constant<int, 0>;
    using B1 = std::integral_constant<int, 1>;
    int x = x0;
    for (; x + 4 <= x1; x += 4) {
        plane_step(B0{}, x + 0, w0, w1, w2, w3, m0, m1, hA, hB);
        plane_step(B1{}, x + 1, w1, w2, w3, w0, m1, m0, hB, hA);
        plane_step(B0{}, x + 2, w2, w3, w0, w1, m0, m1, hA, hB);
        plane_step(B1{}, x + 3, w3, w0, w1, w2, m1, m0, hB, hA);
    }
    // remainder: canonical roles with an explicit rotation; the buffers keep alternating
    auto rotate = [&]() {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                w0[f][r] = w1[f][r];
                w1[f][r] = w2[f][r];
                w2[f][r] = w3[f][r];
                m0[f][r] = m1[f][r];
            }
            hA[f][0] = hB[f][0];
            hA[f][1] = hB[f][1];
        }
    };
    if (x < x1) {
        plane_step(B0{}, x, w0, w1, w2, w3, m0, m1, hA, hB);
        rotate();
        ++x;
    }
    if (x < x1) {
        plane_step(B1{}, x, w0, w1, w2, w3, m0, m1, hA, hB);
        rotate();
        ++x;
    }
    if (x < x1) plane_step(B0{}, x, w0, w1, w2, w3, m0, m1, hA, hB);
}

template <int NF, int TY, int OP, bool ODD, bool HS = false, bool ZN = true>
__global__ __launch_bounds__(NF == 1 ? 1024 : 512) void k_langevin_rows(
    const float* __restrict__ f0, const float* __restrict__ f1, const float* q0,
    const float* q1, float* __restrict__ o0, float* __restrict__ o1, float* r0, float* r1,
    LangevinScalars<float> c, TableF32<ODD> tab, ZnParams<float> zn, int nx, int ny, int nz,
    int W, TileGrid g, uint64_t noise_seed, int hash)
{
    extern __shared__ float lds_lv[];
    int tp, ch;
    g.decode((int)blockIdx.x, (int)gridDim.x, tp, ch);
    const int x0 = ch * g.cx;
    const int x1 = min(x0 + g.cx, nx);
    if constexpr (NF == 1) {
        const float* const phi[1] = {f0};
        const float* const pi[1] = {q0};
        float* const po[1] = {o0};
        float* const qo[1] = {r0};
        lv_rows_tile<1, TY, OP, ODD, HS>(phi, pi, po, qo, c, tab, zn, nx, ny, nz, W, tp, x0,
                                          x1, noise_seed, hash, lds_lv);
    } else {
        const float* const phi[2] = {f0, f1};
        const float* const pi[2] = {q0, q1};
        float* const po[2] = {o0, o1};
        float* const qo[2] = {r0, r1};
        lv_rows_tile<2, TY, OP, ODD, HS, ZN>(phi, pi, po, qo, c, tab, zn, nx, ny, nz, W, tp, x0,
                                              x1, noise_seed, hash, lds_lv);
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
// geom (nullable): TIP_KERNEL_FLAT selects the flat tile, otherwise the row tile where the
// shape allows (rows per thread and planes per block from geom when set).
int langevin_real_f32(int op, const float* phi, const float* pi, float* phi_out, float* pi_out,
                      const LangevinScalars<float>& c, const SegTable& tab, int64_t nx,
                      int64_t ny, int64_t nz, uint64_t noise_seed, int hash, hipStream_t st,
                      const StepGeometry* geom);
int langevin_complex_f32(int op, const float* phi1, const float* phi2, const float* pi1,
                         const float* pi2, float* phi1_out, float* phi2_out, float* pi1_out,
                         float* pi2_out, const LangevinScalars<float>& c, const SegTable& tab,
                         const ZnParams<float>& zn, int64_t nx, int64_t ny, int64_t nz,
                         uint64_t noise_seed, int hash, hipStream_t st,
                         const StepGeometry* geom);
int langevin_real_f64(int op, const double* phi, const double* pi, double* phi_out,
                      double* pi_out, const LangevinScalars<double>& c, const double* table,
                      double tmin, double dinv, int64_t npts, int odd, int64_t nx, int64_t ny,
                      int64_t nz, uint64_t noise_seed, int hash, hipStream_t st);

}  // namespace tip
