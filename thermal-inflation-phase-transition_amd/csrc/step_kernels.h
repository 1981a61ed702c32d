// step_kernels.h -- the fused RK2 Langevin step (kernel (a)), written once for float64 and
// float32 and instantiated by tip_step_f64.hip (compiled with -ffp-contract=off so the float64
// path keeps the reference's rounding sequence) and tip_step_f32.hip.
//
// Algorithm: rk2_step_table, simulation/latticeSimeRescale_numba.py:1875-1966 (real) and
// simulation/latticeSimComplex_numba.py:2067-2215 (complex).  The reference makes two passes
// over the lattice (pass 1 writes phi_tmp, pi_tmp; pass 2 reads the stencil of phi_tmp).
// Because phi_tmp = phi + dt/2 * pi is local to a site, the second-stage Laplacian at a site
// only needs (phi, pi) of its six neighbours: one kernel reads the 7-point stencil of
// (phi, pi) once and writes (phi_new, pi_new) once -- 16 B/site in float32 (read 2 x 4 B,
// write 2 x 4 B) against the reference's 5 arrays x 2 passes.
//
// Work decomposition: a 64 x BY thread block covers a (z, y) tile, one lane per z, and marches
// along x over `cx` planes keeping the x-1, x, x+1 values of its own column in registers.
// Each plane costs two new HBM loads per site (phi, pi at x+1); the four y/z neighbours are
// row segments that the same or adjacent waves just loaded, served from L1/L2.
#pragma once
#include "tip_internal.h"

namespace tip {

template <typename R>
struct StepScalars {
    R dt, half_dt, inv_dx2, inv_mu2, eta, inv_a2, ns;
    uint32_t word_mask;  // all ones (see unit_f32)
};

template <typename R>
inline StepScalars<R> make_scalars(double dt, double dx, double eta_eff, double mu,
                                   double inv_a2, double noise_scale)
{
    StepScalars<R> c;
    c.dt = (R)dt;
    c.half_dt = (R)(0.5 * dt);
    c.inv_dx2 = (R)(1.0 / (dx * dx));
    c.inv_mu2 = (R)(1.0 / (mu * mu));
    c.eta = (R)eta_eff;
    c.inv_a2 = (R)inv_a2;
    c.ns = (R)noise_scale;
    c.word_mask = 0xFFFFFFFFu;
    return c;
}

__device__ __forceinline__ double gauss_of(double, uint64_t s, uint32_t) { return gauss_f64(s); }
__device__ __forceinline__ float gauss_of(float, uint64_t s, uint32_t m) { return gauss_f32(s, m); }
__device__ __forceinline__ void gauss_pair_of(double, uint64_t s, uint32_t, double& a, double& b)
{
    gauss_pair_f64(s, a, b);
}
__device__ __forceinline__ void gauss_pair_of(float, uint64_t s, uint32_t m, float& a, float& b)
{
    gauss_pair_f32(s, m, a, b);
}

__device__ __forceinline__ double r_abs(double x) { return fabs(x); }
__device__ __forceinline__ float r_abs(float x) { return fabsf(x); }

// Wave-level population count of a predicate; all 64 lanes of the wave must execute it.
__device__ __forceinline__ unsigned wave_count(bool p)
{
    return (unsigned)__popcll(__ballot(p));
}

// ---------------------------------------------------------------------------------------
// Real field.
// ---------------------------------------------------------------------------------------
// One tile of 64*BY consecutive sites of the flattened (y, z) plane, marching over x planes
// [x0, x1).  Flattening keeps every lane busy whatever nz is (a 96-wide row would leave a
// quarter of a 64 x k tile idle).  Lanes past the plane compute on site 0 and store nothing,
// so loads need no predication.  Returns the wave's count of sites with |phi_new| > esc when
// `count` is set (all lanes of every wave call it with the same x0, x1 and count).
template <typename R, typename Tab, int BY>
__device__ __forceinline__ unsigned rk2_real_tile(
    const R* __restrict__ phi, const R* __restrict__ pi, R* __restrict__ phi_out,
    R* __restrict__ pi_out, const R* __restrict__ noise, const StepScalars<R>& c,
    const Tab& tab, int nx, int ny, int nz, int tile_p, int x0, int x1,
    uint64_t noise_seed, int hash_noise, bool count, R esc)
{
    const int nyz = ny * nz;
    const int pl = tile_p * (64 * BY) + (int)threadIdx.y * 64 + (int)threadIdx.x;
    const bool act = pl < nyz;
    const int p0 = act ? pl : 0;
    const int jj = p0 / nz, kk = p0 - jj * nz;
    const int kp = kk + 1 < nz ? kk + 1 : 0, km = kk - 1 >= 0 ? kk - 1 : nz - 1;
    const int jp = jj + 1 < ny ? jj + 1 : 0, jm = jj - 1 >= 0 ? jj - 1 : ny - 1;
    // Byte offsets inside one x plane (< 2^32 by the dimension check).  Each load is a uniform
    // plane base (scalar registers) plus a loop-invariant 32-bit lane offset, so the loop needs
    // no per-load 64-bit address arithmetic.
    const uint32_t E = (uint32_t)sizeof(R);
    const uint32_t o_c = (uint32_t)(jj * nz + kk) * E;
    const uint32_t o_kp = (uint32_t)(jj * nz + kp) * E, o_km = (uint32_t)(jj * nz + km) * E;
    const uint32_t o_jp = (uint32_t)(jp * nz + kk) * E, o_jm = (uint32_t)(jm * nz + kk) * E;
    const size_t plane = (size_t)nyz * E;
    const uint32_t pb = (uint32_t)plane;  // plane bytes < 2^32 by the dimension check
    const char* P = reinterpret_cast<const char*>(phi);
    const char* Q = reinterpret_cast<const char*>(pi);
    const char* PO = reinterpret_cast<const char*>(phi_out);
    const char* QO = reinterpret_cast<const char*>(pi_out);
    const char* NZ = reinterpret_cast<const char*>(noise);
    auto at = [&](const char* b, int x) { return make_rsrc(b + (size_t)x * plane, pb); };
    // Site seed uint64(site*K) ^ noise_seed with site = x*nyz + column: the product splits into
    // a per-lane part and a per-plane part (exact mod 2^64).
    const uint64_t colK = (uint64_t)(uint32_t)(jj * nz + kk) * (uint64_t)kNoiseMult;
    const uint64_t planeK = (uint64_t)(uint32_t)nyz * (uint64_t)kNoiseMult;
    uint64_t xK = (uint64_t)(uint32_t)x0 * planeK;

    unsigned cnt = 0;
    const int xm = x0 - 1 >= 0 ? x0 - 1 : nx - 1;
    R pm = buf_load<R>(at(P, xm), o_c);
    R qm = buf_load<R>(at(Q, xm), o_c);
    R pc = buf_load<R>(at(P, x0), o_c);
    R qc = buf_load<R>(at(Q, x0), o_c);
    for (int x = x0; x < x1; ++x) {
        const int xp = x + 1 < nx ? x + 1 : 0;
        const __amdgpu_buffer_rsrc_t Pb = at(P, x), Qb = at(Q, x);
        const R pp = buf_load<R>(at(P, xp), o_c);
        const R qp = buf_load<R>(at(Q, xp), o_c);
        const R pyp = buf_load<R>(Pb, o_jp), pym = buf_load<R>(Pb, o_jm);
        const R pzp = buf_load<R>(Pb, o_kp), pzm = buf_load<R>(Pb, o_km);
        const R qyp = buf_load<R>(Qb, o_jp), qym = buf_load<R>(Qb, o_jm);
        const R qzp = buf_load<R>(Qb, o_kp), qzm = buf_load<R>(Qb, o_km);

        // Stage 1 (reference pass 1): k1 at (phi, pi); midpoint (phi_t, pi_t).
        R s6 = pp + pm;
        s6 = s6 + pyp;
        s6 = s6 + pym;
        s6 = s6 + pzp;
        s6 = s6 + pzm;
        R lap = (s6 - (R)6 * pc) * c.inv_dx2;
        R dV = tab(pc);
        R k1 = c.inv_a2 * lap - c.eta * qc - dV * c.inv_mu2;
        R ftc = pc + c.half_dt * qc;
        R ptc = qc + c.half_dt * k1;

        // Stage 2 (reference pass 2): phi_t of the six neighbours is local to each of them.
        R t_ip = pp + c.half_dt * qp;
        R t_im = pm + c.half_dt * qm;
        R t_jp = pyp + c.half_dt * qyp;
        R t_jm = pym + c.half_dt * qym;
        R t_kp = pzp + c.half_dt * qzp;
        R t_km = pzm + c.half_dt * qzm;
        R s6b = t_ip + t_im;
        s6b = s6b + t_jp;
        s6b = s6b + t_jm;
        s6b = s6b + t_kp;
        s6b = s6b + t_km;
        R lap2 = (s6b - (R)6 * ftc) * c.inv_dx2;
        R dV2 = tab(ftc);
        R k2 = c.inv_a2 * lap2 - c.eta * ptc - dV2 * c.inv_mu2;

        R nz_ = 0;
        if (noise) nz_ = buf_load<R>(at(NZ, x), o_c);
        else if (hash_noise) nz_ = c.ns * gauss_of((R)0, (colK + xK) ^ noise_seed, c.word_mask);
        R pn = pc + c.dt * ptc;
        R qn = qc + (c.dt * k2 + nz_);
        if (act) {
            buf_store(pn, at(PO, x), o_c);
            buf_store(qn, at(QO, x), o_c);
        }
        if (count) cnt += wave_count(act && r_abs(pn) > esc);
        pm = pc; qm = qc; pc = pp; qc = qp;
        xK += planeK;
    }
    return cnt;
}

// Block -> tile decode.  Blocks are numbered with the z tile fastest, then the y tile, then
// the x chunk.  With `xcd_remap` the hardware's round-robin dispatch of consecutive block ids
// over the 8 XCDs is undone: block b runs tile (b % 8) * (nb / 8) + b / 8, so each XCD works
// through one contiguous range of x chunks and the neighbour planes a chunk re-reads stay in
// that XCD's L2.  Requires nb % 8 == 0 (the launcher clears the flag otherwise).
struct TileGrid {
    int tiles, chunks, cx, xcd_remap;  // tiles: plane tiles of 64*BY sites
    __device__ __forceinline__ void decode(int b, int nb, int& tp, int& ch) const
    {
        if (xcd_remap) b = (b & 7) * (nb >> 3) + (b >> 3);
        tp = b % tiles;
        ch = b / tiles;
    }
};

template <typename R, typename Tab, int BY>
__global__ __launch_bounds__(64 * BY) void k_rk2_real(
    const R* __restrict__ phi, const R* __restrict__ pi, R* __restrict__ phi_out,
    R* __restrict__ pi_out, const R* __restrict__ noise, StepScalars<R> c, Tab tab, int nx,
    int ny, int nz, TileGrid g, uint64_t noise_seed, int hash_noise,
    unsigned long long* __restrict__ vac, R esc)
{
    int tp, ch;
    g.decode((int)blockIdx.x, (int)gridDim.x, tp, ch);
    const int x0 = ch * g.cx;
    const int x1 = min(x0 + g.cx, nx);
    unsigned cnt = rk2_real_tile<R, Tab, BY>(phi, pi, phi_out, pi_out, noise, c, tab, nx, ny, nz,
                                             tp, x0, x1, noise_seed, hash_noise, vac != nullptr,
                                             esc);
    if (vac && cnt && (threadIdx.x & 63) == 0) atomicAdd(vac, (unsigned long long)cnt);
}

// ---------------------------------------------------------------------------------------
// Complex field (phi1 + i phi2), latticeSimComplex_numba.py:2067-2215.
// ---------------------------------------------------------------------------------------
template <typename R>
struct ZnParams {
    R n, dv;
    int active;  // zn_active > 0.5 && zn_n > 0.5
    int ipow;    // n as a small integer (1..32) for the float32 power form, else 0
};

// Z_N force term, latticeSimComplex_numba.py:2147-2152:
//   dV1 += -N dV sin(N theta) phi2 / rho^2,  dV2 += +N dV sin(N theta) phi1 / rho^2.
__device__ __forceinline__ double zn_sin(const ZnParams<double>& z, double p1, double p2, double)
{
    return sin(z.n * atan2(p2, p1));
}
// float32: sin(N theta) = Im((phi1 + i phi2)^N) / rho^N for integer N, no atan2/sin needed.
__device__ __forceinline__ float zn_sin(const ZnParams<float>& z, float p1, float p2, float rho)
{
    if (z.ipow > 0) {
        float inv = 1.0f / rho;
        float a = p1 * inv, b = p2 * inv, re = 1.0f, im = 0.0f;
        for (int e = 0; e < z.ipow; ++e) {
            float t = re * a - im * b;
            im = re * b + im * a;
            re = t;
        }
        return im;
    }
    return sinf(z.n * atan2f(p2, p1));
}

__device__ __forceinline__ double r_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float r_sqrt(float x) { return __builtin_sqrtf(x); }

template <typename R, typename Tab>
__device__ __forceinline__ void complex_force(const Tab& tab, const ZnParams<R>& zn, R p1, R p2,
                                              R& dV1, R& dV2)
{
    R rho = r_sqrt(p1 * p1 + p2 * p2);
    R rho_safe = rho > (R)1e-30 ? rho : (R)1e-30;
    R dvr = tab(rho);
    dV1 = dvr * p1 / rho_safe;
    dV2 = dvr * p2 / rho_safe;
    if (zn.active && rho > (R)1e-20) {
        R sn = zn_sin(zn, p1, p2, rho);
        R inv_rho2 = ((R)1 / rho_safe) * ((R)1 / rho_safe);
        dV1 += -zn.n * zn.dv * sn * p2 * inv_rho2;
        dV2 += zn.n * zn.dv * sn * p1 * inv_rho2;
    }
}

template <typename R, typename Tab, int BY>
__global__ __launch_bounds__(64 * BY) void k_rk2_complex(
    const R* __restrict__ phi1, const R* __restrict__ phi2, const R* __restrict__ pi1,
    const R* __restrict__ pi2, R* __restrict__ phi1_out, R* __restrict__ phi2_out,
    R* __restrict__ pi1_out, R* __restrict__ pi2_out, const R* __restrict__ noise1,
    const R* __restrict__ noise2, StepScalars<R> c, Tab tab, ZnParams<R> zn, int nx, int ny,
    int nz, TileGrid g, uint64_t noise_seed, int hash_noise,
    unsigned long long* __restrict__ vac, R esc2)
{
    int tp, ch;
    g.decode((int)blockIdx.x, (int)gridDim.x, tp, ch);
    const int x0 = ch * g.cx;
    const int x1 = min(x0 + g.cx, nx);
    const int nyz = ny * nz;
    const int pl = tp * (64 * BY) + (int)threadIdx.y * 64 + (int)threadIdx.x;
    const bool act = pl < nyz;
    const int p0 = act ? pl : 0;
    const int jj = p0 / nz, kk = p0 - jj * nz;
    const int kp = kk + 1 < nz ? kk + 1 : 0, km = kk - 1 >= 0 ? kk - 1 : nz - 1;
    const int jp = jj + 1 < ny ? jj + 1 : 0, jm = jj - 1 >= 0 ? jj - 1 : ny - 1;
    const int row = jj * nz;
    const int o_kp = row + kp, o_km = row + km, o_c = row + kk;
    const int o_jp = jp * nz + kk, o_jm = jm * nz + kk;

    unsigned cnt = 0;
    R am = 0, bm = 0, um = 0, vm = 0, ac = 0, bc = 0, uc = 0, vc = 0;  // a=phi1 b=phi2 u=pi1 v=pi2
    if (act) {
        const int xm = x0 - 1 >= 0 ? x0 - 1 : nx - 1;
        am = phi1[xm * nyz + o_c];
        bm = phi2[xm * nyz + o_c];
        um = pi1[xm * nyz + o_c];
        vm = pi2[xm * nyz + o_c];
        ac = phi1[x0 * nyz + o_c];
        bc = phi2[x0 * nyz + o_c];
        uc = pi1[x0 * nyz + o_c];
        vc = pi2[x0 * nyz + o_c];
    }
    for (int x = x0; x < x1; ++x) {
        const int xp = x + 1 < nx ? x + 1 : 0;
        const int base = x * nyz;
        R ap = 0, bp = 0, up = 0, vp = 0;
        bool tv = false;
        if (act) {
            ap = phi1[xp * nyz + o_c];
            bp = phi2[xp * nyz + o_c];
            up = pi1[xp * nyz + o_c];
            vp = pi2[xp * nyz + o_c];
            R a_jp = phi1[base + o_jp], a_jm = phi1[base + o_jm];
            R a_kp = phi1[base + o_kp], a_km = phi1[base + o_km];
            R b_jp = phi2[base + o_jp], b_jm = phi2[base + o_jm];
            R b_kp = phi2[base + o_kp], b_km = phi2[base + o_km];
            R u_jp = pi1[base + o_jp], u_jm = pi1[base + o_jm];
            R u_kp = pi1[base + o_kp], u_km = pi1[base + o_km];
            R v_jp = pi2[base + o_jp], v_jm = pi2[base + o_jm];
            R v_kp = pi2[base + o_kp], v_km = pi2[base + o_km];

            // Stage 1.
            R s1 = ap + am;
            s1 = s1 + a_jp;
            s1 = s1 + a_jm;
            s1 = s1 + a_kp;
            s1 = s1 + a_km;
            R s2 = bp + bm;
            s2 = s2 + b_jp;
            s2 = s2 + b_jm;
            s2 = s2 + b_kp;
            s2 = s2 + b_km;
            R lap1 = (s1 - (R)6 * ac) * c.inv_dx2;
            R lap2 = (s2 - (R)6 * bc) * c.inv_dx2;
            R dV1, dV2;
            complex_force(tab, zn, ac, bc, dV1, dV2);
            R k1a = c.inv_a2 * lap1 - c.eta * uc - dV1 * c.inv_mu2;
            R k1b = c.inv_a2 * lap2 - c.eta * vc - dV2 * c.inv_mu2;
            R fa = ac + c.half_dt * uc, fb = bc + c.half_dt * vc;
            R ga = uc + c.half_dt * k1a, gb = vc + c.half_dt * k1b;

            // Stage 2.
            R t1 = (ap + c.half_dt * up) + (am + c.half_dt * um);
            t1 = t1 + (a_jp + c.half_dt * u_jp);
            t1 = t1 + (a_jm + c.half_dt * u_jm);
            t1 = t1 + (a_kp + c.half_dt * u_kp);
            t1 = t1 + (a_km + c.half_dt * u_km);
            R t2 = (bp + c.half_dt * vp) + (bm + c.half_dt * vm);
            t2 = t2 + (b_jp + c.half_dt * v_jp);
            t2 = t2 + (b_jm + c.half_dt * v_jm);
            t2 = t2 + (b_kp + c.half_dt * v_kp);
            t2 = t2 + (b_km + c.half_dt * v_km);
            R l1 = (t1 - (R)6 * fa) * c.inv_dx2;
            R l2 = (t2 - (R)6 * fb) * c.inv_dx2;
            R e1, e2;
            complex_force(tab, zn, fa, fb, e1, e2);
            R k2a = c.inv_a2 * l1 - c.eta * ga - e1 * c.inv_mu2;
            R k2b = c.inv_a2 * l2 - c.eta * gb - e2 * c.inv_mu2;

            const int s = base + o_c;
            R n1 = 0, n2 = 0;
            if (noise1) {
                n1 = noise1[s];
                n2 = noise2 ? noise2[s] : (R)0;
            } else if (hash_noise) {
                R z1, z2;
                gauss_pair_of((R)0, site_seed((int64_t)s, noise_seed), c.word_mask, z1, z2);
                n1 = c.ns * z1;
                n2 = c.ns * z2;
            }
            R na = ac + c.dt * ga, nb = bc + c.dt * gb;
            phi1_out[s] = na;
            phi2_out[s] = nb;
            pi1_out[s] = uc + (c.dt * k2a + n1);
            pi2_out[s] = vc + (c.dt * k2b + n2);
            tv = na * na + nb * nb > esc2;
        }
        if (vac) cnt += wave_count(tv);
        am = ac; bm = bc; um = uc; vm = vc;
        ac = ap; bc = bp; uc = up; vc = vp;
    }
    if (vac && cnt && (threadIdx.x & 63) == 0) atomicAdd(vac, (unsigned long long)cnt);
}

// Launch geometry shared by the real and complex launchers.
struct StepGeometry {
    int by;   // rows per block (1, 2, 4 or 8)
    int cx;   // x planes per block (0 = choose from the lattice size)
    int xcd;  // 1 = XCD-aware block remap
};
StepGeometry step_geometry(int64_t nx, int64_t ny, int64_t nz, int complex_field, int f64);
// Resolve a requested geometry (nullable = default) into the tile grid and the block count.
int resolve_geometry(const StepGeometry* req, int64_t nx, int64_t ny, int64_t nz,
                     int complex_field, int f64, StepGeometry& g, TileGrid& tg, int& nblocks);

// Launchers shared by the C-ABI entry points, the lattice and the ensemble.
int launch_real_f32(const float* phi, const float* pi, float* phi_out, float* pi_out,
                    const StepScalars<float>& c, const float* noise, const SegTable& tab,
                    int64_t nx, int64_t ny, int64_t nz, uint64_t noise_seed, int hash,
                    unsigned long long* vac, float esc, const StepGeometry* geom,
                    hipStream_t st);
int launch_complex_f32(const float* phi1, const float* phi2, const float* pi1, const float* pi2,
                       float* o1, float* o2, float* o3, float* o4, const StepScalars<float>& c,
                       const float* noise1, const float* noise2, const SegTable& tab,
                       const ZnParams<float>& zn, int64_t nx, int64_t ny, int64_t nz,
                       uint64_t noise_seed, int hash, unsigned long long* vac, float esc2,
                       const StepGeometry* geom, hipStream_t st);
SegTable make_table32(const void* ab, double tmin, double dinv, int64_t npts, int odd);
// Complex fused step with the V' surrogate evaluated inline (tip_mlp.hip).
int launch_complex_mlp(const float* phi1, const float* phi2, const float* pi1, const float* pi2,
                       float* o1, float* o2, float* o3, float* o4, const StepScalars<float>& c,
                       const float* mlp_params, const tip_mlp_features* fm, double T,
                       const ZnParams<float>& zn, int64_t nx, int64_t ny, int64_t nz,
                       uint64_t noise_seed, int hash, unsigned long long* vac, float esc2,
                       hipStream_t st);
ZnParams<float> make_zn32(double zn_n, double zn_dv, double zn_active);

}  // namespace tip
