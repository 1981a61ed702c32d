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
#include <type_traits>

#include "tip_internal.h"

namespace tip {

template <typename R>
struct StepScalars {
    R dt, half_dt, inv_dx2, inv_mu2, eta, inv_a2, ns;
    R c_lap;  // inv_a2 * inv_dx2 (float32 path)
    R ns_l2;  // ns * sqrt(2 ln 2): amplitude applied to sqrt(-log2 u1) cos(2 pi u2)
    uint32_t word_mask;  // all ones (see unit_f32)
};

// Noise source of the fused step, a compile-time choice so that no branch joins the three
// paths inside the plane loop.
enum NoiseMode { NOISE_NONE = 0, NOISE_ARRAY = 1, NOISE_HASH = 2 };

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
    c.c_lap = (R)(inv_a2 / (dx * dx));
    c.ns_l2 = (R)(noise_scale * 1.1774100225154747);  // sqrt(2 ln 2)
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
// CNT is set (all lanes of every wave call it with the same x0 and x1).
// One site of the RK2 step in the reference's operation order (pass 1 :1912-1933, pass 2
// :1945-1966): the six neighbours' phi_t = phi + dt/2 pi are formed from (phi, pi), so both
// passes read only the step's input.  dV = V'(pc), dV2 = V'(ftc), ftc = pc + dt/2 qc (computed
// by the caller for its table lookup with the same expression).
template <typename R>
__device__ __forceinline__ void rk2_site(const StepScalars<R>& c, R pc, R qc, R pp, R qp, R pm,
                                         R qm, R pyp, R qyp, R pym, R qym, R pzp, R qzp, R pzm,
                                         R qzm, R ftc, R dV, R dV2, R nz_, R& pn, R& qn)
{
    // Stage 1 (reference pass 1): k1 at (phi, pi); midpoint (phi_t, pi_t).
    R s6 = pp + pm;
    s6 = s6 + pyp;
    s6 = s6 + pym;
    s6 = s6 + pzp;
    s6 = s6 + pzm;
    R lap = (s6 - (R)6 * pc) * c.inv_dx2;
    R k1 = c.inv_a2 * lap - c.eta * qc - dV * c.inv_mu2;
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
    R k2 = c.inv_a2 * lap2 - c.eta * ptc - dV2 * c.inv_mu2;
    pn = pc + c.dt * ptc;
    qn = qc + (c.dt * k2 + nz_);
}

template <typename R, typename Tab, int BY, bool CNT>
__device__ __forceinline__ unsigned rk2_real_tile(
    const R* __restrict__ phi, const R* __restrict__ pi, R* __restrict__ phi_out,
    R* __restrict__ pi_out, const R* __restrict__ noise, const StepScalars<R>& c,
    const Tab& tab, int nx, int ny, int nz, int tile_p, int x0, int x1,
    uint64_t noise_seed, int hash_noise, R esc)
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

        const R ftc = pc + c.half_dt * qc;
        const R dV = tab(pc), dV2 = tab(ftc);
        R nz_ = 0;
        if (noise) nz_ = buf_load<R>(at(NZ, x), o_c);
        else if (hash_noise) nz_ = c.ns * gauss_of((R)0, (colK + xK) ^ noise_seed, c.word_mask);
        R pn, qn;
        rk2_site<R>(c, pc, qc, pp, qp, pm, qm, pyp, qyp, pym, qym, pzp, qzp, pzm, qzm, ftc, dV, dV2,
                    nz_, pn, qn);
        if (act) {
            buf_store(pn, at(PO, x), o_c);
            buf_store(qn, at(QO, x), o_c);
        }
        if constexpr (CNT) cnt += wave_count(act && r_abs(pn) > esc);
        pm = pc; qm = qc; pc = pp; qc = qp;
        xK += planeK;
    }
    return cnt;
}

// ---------------------------------------------------------------------------------------
// float32 tile: the same step, rearranged for the float32 path's instruction count and
// software-pipelined over the x march.
//
//  * lap(phi_t) = lap(phi) + dt/2 lap(pi): the stage-2 numerator is t1 + dt/2 (sum6(pi) - 6 pi)
//    (algebraically the reference's sum of the six neighbours' phi_t minus 6 phi_t);
//    inv_a2 * inv_dx2 is one constant.
//  * Both table gathers depend only on (phi, pi) of the current plane, so they are issued
//    first; the loads of the next plane (two centre values, eight neighbours) follow, and the
//    hash noise is computed while all of them are in flight.  Loads complete in order, so the
//    gathers are waited for without waiting for the next plane.
//  * noise: ns * sqrt(-2 ln u1) cos(2 pi u2) = ns sqrt(2 ln 2) * sqrt(-log2 u1) * cos(u2 rev),
//    u2 from the high word of the second state (its low word moves u2 by < 2^-32).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float gauss_f32_l2(uint64_t seed, uint32_t word_mask)
{
    uint64_t x1, x2;
    hash_states(seed, x1, x2);
    const float u1 = fmaxf(unit_f32(x1, word_mask), 1e-12f);
    const float u2 = (float)((uint32_t)(x2 >> 32) & word_mask) * 0x1p-32f;
    return __builtin_amdgcn_sqrtf(-__builtin_amdgcn_logf(u1)) * __builtin_amdgcn_cosf(u2);
}

template <bool ODD>
__device__ __forceinline__ float2 seg_load(__amdgpu_buffer_rsrc_t r, const TableF32<ODD>& t,
                                           float x)
{
    const float ax = ODD ? fabsf(x) : x;
    const float fidx = fmaf(ax, t.dinv, t.off);
    const uint32_t idx = (uint32_t)(int)__builtin_amdgcn_fmed3f(fidx, 0.0f, (float)t.last);
    using f2 = float __attribute__((ext_vector_type(2)));
    const f2 v = __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(idx * 8u), 0, 0));
    return make_float2(v.x, v.y);
}

template <bool ODD>
__device__ __forceinline__ float seg_eval(float2 s, float x)
{
    if (!ODD) return fmaf(s.y, x, s.x);
    const float w = fmaf(s.y, fabsf(x), s.x);
    return __uint_as_float(__float_as_uint(w) ^ (__float_as_uint(x) & 0x80000000u));
}

template <bool ODD, int BY, int NM>
__device__ __forceinline__ unsigned rk2_real_tile_f32(
    const float* __restrict__ phi, const float* __restrict__ pi, float* __restrict__ phi_out,
    float* __restrict__ pi_out, const float* __restrict__ noise, const StepScalars<float>& c,
    const TableF32<ODD>& tab, int nx, int ny, int nz, int tile_p, int x0, int x1,
    uint64_t noise_seed, bool count, float esc)
{
    const int nyz = ny * nz;
    const int pl = tile_p * (64 * BY) + (int)threadIdx.y * 64 + (int)threadIdx.x;
    const bool act = pl < nyz;
    const int p0 = act ? pl : 0;
    const int jj = p0 / nz, kk = p0 - jj * nz;
    const int kp = kk + 1 < nz ? kk + 1 : 0, km = kk - 1 >= 0 ? kk - 1 : nz - 1;
    const int jp = jj + 1 < ny ? jj + 1 : 0, jm = jj - 1 >= 0 ? jj - 1 : ny - 1;
    const uint32_t o_c = (uint32_t)(jj * nz + kk) * 4u;
    const uint32_t o_kp = (uint32_t)(jj * nz + kp) * 4u, o_km = (uint32_t)(jj * nz + km) * 4u;
    const uint32_t o_jp = (uint32_t)(jp * nz + kk) * 4u, o_jm = (uint32_t)(jm * nz + kk) * 4u;
    const size_t plane = (size_t)nyz * 4u;
    const uint32_t pb = (uint32_t)plane;
    const char* P = reinterpret_cast<const char*>(phi);
    const char* Q = reinterpret_cast<const char*>(pi);
    const char* PO = reinterpret_cast<const char*>(phi_out);
    const char* QO = reinterpret_cast<const char*>(pi_out);
    const char* NZ = reinterpret_cast<const char*>(noise);
    auto at = [&](const char* b, int x) { return make_rsrc(b + (size_t)x * plane, pb); };
    const __amdgpu_buffer_rsrc_t T = make_rsrc(tab.ab, (uint32_t)(tab.last + 1) * 8u);
    const uint64_t colK = (uint64_t)(uint32_t)(jj * nz + kk) * (uint64_t)kNoiseMult;
    const uint64_t planeK = (uint64_t)(uint32_t)nyz * (uint64_t)kNoiseMult;
    uint64_t xK = (uint64_t)(uint32_t)x0 * planeK;

    unsigned cnt = 0;
    const int xm = x0 - 1 >= 0 ? x0 - 1 : nx - 1;
    const int xq = x0 + 1 < nx ? x0 + 1 : 0;
    // Window of centre values at x-1, x, x+1 and the prefetch slot for x+2 (w*), and two sets
    // of the eight in-plane neighbours (current plane / next plane).
    float wp0 = buf_load<float>(at(P, xm), o_c), wq0 = buf_load<float>(at(Q, xm), o_c);
    float wp1 = buf_load<float>(at(P, x0), o_c), wq1 = buf_load<float>(at(Q, x0), o_c);
    float wp2 = buf_load<float>(at(P, xq), o_c), wq2 = buf_load<float>(at(Q, xq), o_c);
    float wp3, wq3;
    float A[8], B[8];  // [phi y+, y-, z+, z-, pi y+, y-, z+, z-]
    {
        const __amdgpu_buffer_rsrc_t Pb = at(P, x0), Qb = at(Q, x0);
        A[0] = buf_load<float>(Pb, o_jp); A[1] = buf_load<float>(Pb, o_jm);
        A[2] = buf_load<float>(Pb, o_kp); A[3] = buf_load<float>(Pb, o_km);
        A[4] = buf_load<float>(Qb, o_jp); A[5] = buf_load<float>(Qb, o_jm);
        A[6] = buf_load<float>(Qb, o_kp); A[7] = buf_load<float>(Qb, o_km);
    }

    // One plane.  The next plane's loads are issued unconditionally (periodic indices keep
    // them in bounds); on the chunk's last plane they are simply not used.
    auto plane_step = [&](int x, float pm, float qm, float pc, float qc, float pp, float qp,
                          float& pn2, float& qn2, const float (&cur)[8], float (&nxt)[8]) {
        const float ftc = fmaf(c.half_dt, qc, pc);
        const float2 s1 = seg_load<ODD>(T, tab, pc);
        const float2 s2 = seg_load<ODD>(T, tab, ftc);
        float nzv = 0.0f;
        if (NM == NOISE_ARRAY) nzv = buf_load<float>(at(NZ, x), o_c);
        {
            const int xa = x + 1 < nx ? x + 1 : 0;
            const int xb = xa + 1 < nx ? xa + 1 : 0;
            pn2 = buf_load<float>(at(P, xb), o_c);
            qn2 = buf_load<float>(at(Q, xb), o_c);
            const __amdgpu_buffer_rsrc_t Pa = at(P, xa), Qa = at(Q, xa);
            nxt[0] = buf_load<float>(Pa, o_jp); nxt[1] = buf_load<float>(Pa, o_jm);
            nxt[2] = buf_load<float>(Pa, o_kp); nxt[3] = buf_load<float>(Pa, o_km);
            nxt[4] = buf_load<float>(Qa, o_jp); nxt[5] = buf_load<float>(Qa, o_jm);
            nxt[6] = buf_load<float>(Qa, o_kp); nxt[7] = buf_load<float>(Qa, o_km);
        }
        float g = 0.0f;
        if (NM == NOISE_HASH) g = gauss_f32_l2((colK + xK) ^ noise_seed, c.word_mask);

        const float s6 = ((((pp + pm) + cur[0]) + cur[1]) + cur[2]) + cur[3];
        const float t1 = fmaf(-6.0f, pc, s6);
        const float q6 = ((((qp + qm) + cur[4]) + cur[5]) + cur[6]) + cur[7];
        const float tq = fmaf(-6.0f, qc, q6);
        const float t2 = fmaf(c.half_dt, tq, t1);
        const float k1 = fmaf(-c.inv_mu2, seg_eval<ODD>(s1, pc), fmaf(c.c_lap, t1, -c.eta * qc));
        const float ptc = fmaf(c.half_dt, k1, qc);
        const float k2 =
            fmaf(-c.inv_mu2, seg_eval<ODD>(s2, ftc), fmaf(c.c_lap, t2, -c.eta * ptc));
        const float pn = fmaf(c.dt, ptc, pc);
        float qn;
        if (NM == NOISE_HASH) qn = fmaf(c.dt, k2, fmaf(c.ns_l2, g, qc));
        else if (NM == NOISE_ARRAY) qn = qc + fmaf(c.dt, k2, nzv);
        else qn = fmaf(c.dt, k2, qc);
        if (act) {
            buf_store(pn, at(PO, x), o_c);
            buf_store(qn, at(QO, x), o_c);
        }
        if (count) cnt += wave_count(act && fabsf(pn) > esc);
        xK += planeK;
    };

    // Four planes per trip: the window and neighbour sets change roles instead of moving.
    int x = x0;
    for (; x + 4 <= x1; x += 4) {
        plane_step(x + 0, wp0, wq0, wp1, wq1, wp2, wq2, wp3, wq3, A, B);
        plane_step(x + 1, wp1, wq1, wp2, wq2, wp3, wq3, wp0, wq0, B, A);
        plane_step(x + 2, wp2, wq2, wp3, wq3, wp0, wq0, wp1, wq1, A, B);
        plane_step(x + 3, wp3, wq3, wp0, wq0, wp1, wq1, wp2, wq2, B, A);
    }
    for (; x < x1; ++x) {
        plane_step(x, wp0, wq0, wp1, wq1, wp2, wq2, wp3, wq3, A, B);
        wp0 = wp1; wq0 = wq1; wp1 = wp2; wq1 = wq2; wp2 = wp3; wq2 = wq3;
#pragma unroll
        for (int k = 0; k < 8; ++k) A[k] = B[k];
    }
    return cnt;
}

// ---------------------------------------------------------------------------------------
// float32 row-tile kernel with LDS staging (the default float32 real-field step).
//
// A group of W = ceil(nz/64) waves covers whole z rows (lane z of wave zw is site
// z = 64 zw + lane), and each thread carries TY consecutive y rows of its z column through
// the x march.  Per plane, the group stages its TY rows of (phi, pi) in LDS once, behind one
// barrier (double-buffered, so the next plane's staging needs no second barrier); z
// neighbours, periodic inside the row, are LDS reads; y neighbours are registers of the same
// thread except for the two halo rows, which are loaded; x neighbours are the register window.
// That is 2 + 4/TY global loads per site against 10 for the flat tile above, which is what
// bounds the flat tile (its y/z neighbours are L1/L2 requests).  A block holds G groups of
// consecutive row blocks (G = 4/W for narrow rows, so a block has 4 waves).  TY must divide
// ny; groups past the last row are idle.  The next plane's centres and halo rows are
// loaded one plane ahead (not on the chunk's last plane).  With DENSE (the ensemble's small lattices) a group is nz
// consecutive threads rather than whole waves, so rows whose length is not a multiple of 64
// leave no lanes idle when the block size is a multiple of nz.
// ---------------------------------------------------------------------------------------
template <bool ODD, int TY, int NM, bool CNT, bool DENSE = false>
__device__ __forceinline__ unsigned rk2_real_rows_tile(
    const float* __restrict__ phi, const float* __restrict__ pi, float* __restrict__ phi_out,
    float* __restrict__ pi_out, const float* __restrict__ noise, const StepScalars<float>& c,
    const TableF32<ODD>& tab, int nx, int ny, int nz, int W, int tile, int x0, int x1,
    uint64_t noise_seed, float esc, float* __restrict__ lds)
{
    // LDS words per z column: 2 buffers x 2 fields x TY rows, plus one pad word so that
    // consecutive lanes (stride K words) hit distinct banks.  Every slot of a column is a
    // compile-time offset from one of three per-thread base addresses (z, z+1, z-1).
    constexpr int K = 4 * TY + 1;
    // Thread -> (row group, z).  Default: a group is W whole waves (z = 64 zw + lane).  DENSE:
    // groups of nz threads packed back to back (W unused); the block's thread count must be
    // a multiple of nz, so a row of 96 sites takes 96 lanes instead of two waves.
    int G, grp, z, pitch;
    if constexpr (DENSE) {
        const int t = (int)threadIdx.y * 64 + (int)threadIdx.x;
        G = ((int)blockDim.y * 64) / nz;
        grp = t / nz;
        z = t - grp * nz;
        pitch = nz;
    } else {
        G = (int)blockDim.y / W;
        const int w = (int)threadIdx.y;
        grp = w / W;
        z = (w - grp * W) * 64 + (int)threadIdx.x;
        pitch = 64 * W;
    }
    const bool zact = z < nz;
    const int zc = zact ? z : 0;
    const int zp = zc + 1 < nz ? zc + 1 : 0, zm = zc > 0 ? zc - 1 : nz - 1;
    const int y0 = (tile * G + grp) * TY;
    const bool act = zact && y0 < ny;
    const int yb = y0 < ny ? y0 : 0;
    float* Lg = lds + (size_t)grp * pitch * K;
    float* Lz = Lg + z * K;
    const float* Lzp = Lg + zp * K;
    const float* Lzm = Lg + zm * K;

    const int nyz = ny * nz;
    const size_t plane = (size_t)nyz * 4u;
    const uint32_t pb = (uint32_t)plane;
    const char* P = reinterpret_cast<const char*>(phi);
    const char* Q = reinterpret_cast<const char*>(pi);
    const char* PO = reinterpret_cast<const char*>(phi_out);
    const char* QO = reinterpret_cast<const char*>(pi_out);
    const char* NZ = reinterpret_cast<const char*>(noise);
    auto at = [&](const char* b, int x) { return make_rsrc(b + (size_t)x * plane, pb); };
    const __amdgpu_buffer_rsrc_t T = make_rsrc(tab.ab, (uint32_t)(tab.last + 1) * 8u);
    uint32_t off[TY], offs[TY];
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        off[r] = (uint32_t)((yb + r) * nz + zc) * 4u;
        // idle lanes store past the buffer end, where buffer stores are dropped: no branch
        offs[r] = act ? off[r] : 0xFFFFFFF0u;
    }
    const uint32_t offm = (uint32_t)((yb > 0 ? yb - 1 : ny - 1) * nz + zc) * 4u;
    const uint32_t offp = (uint32_t)((yb + TY < ny ? yb + TY : 0) * nz + zc) * 4u;
    const uint64_t planeK = (uint64_t)(uint32_t)nyz * (uint64_t)kNoiseMult;
    const uint64_t col0K = (uint64_t)(uint32_t)(yb * nz + zc) * (uint64_t)kNoiseMult;
    const uint64_t rowK = (uint64_t)(uint32_t)nz * (uint64_t)kNoiseMult;
    uint64_t xK = (uint64_t)(uint32_t)x0 * planeK;

    // Centre window x-1, x, x+1 and the slot for x+2 (w0..w3 change roles each plane), and
    // halo rows (y0-1, y0+TY) of the current / next plane.
    float p0[TY], q0[TY], p1[TY], q1[TY], p2[TY], q2[TY];
    float p3[TY] = {}, q3[TY] = {};  // zero where a chunk's last plane loads nothing
    {
        const int xm = x0 - 1 >= 0 ? x0 - 1 : nx - 1;
        const int xq = x0 + 1 < nx ? x0 + 1 : 0;
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            p0[r] = buf_load<float>(at(P, xm), off[r]); q0[r] = buf_load<float>(at(Q, xm), off[r]);
            p1[r] = buf_load<float>(at(P, x0), off[r]); q1[r] = buf_load<float>(at(Q, x0), off[r]);
            p2[r] = buf_load<float>(at(P, xq), off[r]); q2[r] = buf_load<float>(at(Q, xq), off[r]);
        }
    }
    float hA[4], hB[4] = {};  // phi y-1, pi y-1, phi y+TY, pi y+TY
    hA[0] = buf_load<float>(at(P, x0), offm); hA[1] = buf_load<float>(at(Q, x0), offm);
    hA[2] = buf_load<float>(at(P, x0), offp); hA[3] = buf_load<float>(at(Q, x0), offp);
    unsigned cnt = 0;

    auto plane_step = [&](auto bufc, int x, const float (&pm)[TY], const float (&qm)[TY],
                          const float (&pc)[TY], const float (&qc)[TY], const float (&pp)[TY],
                          const float (&qp)[TY], float (&pn)[TY], float (&qn)[TY],
                          const float (&hc)[4], float (&hn)[4]) {
        constexpr int b = decltype(bufc)::value;
        constexpr int oP = (b * 2 + 0) * TY, oQ = (b * 2 + 1) * TY;
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            Lz[oP + r] = pc[r];
            Lz[oQ + r] = qc[r];
        }
        // the next plane's centres (x + 2) and halo rows (x + 1); the chunk's last plane has
        // no next plane (a wave-uniform branch): those planes belong to the next chunk
        if (x + 1 < x1) {
            const int xa = x + 1 < nx ? x + 1 : 0;
            const int xb = xa + 1 < nx ? xa + 1 : 0;
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                pn[r] = buf_load<float>(at(P, xb), off[r]);
                qn[r] = buf_load<float>(at(Q, xb), off[r]);
            }
            hn[0] = buf_load<float>(at(P, xa), offm); hn[1] = buf_load<float>(at(Q, xa), offm);
            hn[2] = buf_load<float>(at(P, xa), offp); hn[3] = buf_load<float>(at(Q, xa), offp);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            const float pyp = r + 1 < TY ? pc[r + 1 < TY ? r + 1 : r] : hc[2];
            const float pym = r > 0 ? pc[r > 0 ? r - 1 : 0] : hc[0];
            const float qyp = r + 1 < TY ? qc[r + 1 < TY ? r + 1 : r] : hc[3];
            const float qym = r > 0 ? qc[r > 0 ? r - 1 : 0] : hc[1];
            const float pzp = Lzp[oP + r], pzm = Lzm[oP + r];
            const float qzp = Lzp[oQ + r], qzm = Lzm[oQ + r];
            const float ftc = fmaf(c.half_dt, qc[r], pc[r]);
            const float2 s1 = seg_load<ODD>(T, tab, pc[r]);
            const float2 s2 = seg_load<ODD>(T, tab, ftc);
            float nzv = 0.0f, g = 0.0f;
            if (NM == NOISE_ARRAY) nzv = buf_load<float>(at(NZ, x), off[r]);
            if (NM == NOISE_HASH)
                g = gauss_f32_l2((col0K + (uint64_t)r * rowK + xK) ^ noise_seed, c.word_mask);
            const float s6 = ((((pp[r] + pm[r]) + pyp) + pym) + pzp) + pzm;
            const float t1 = fmaf(-6.0f, pc[r], s6);
            const float q6 = ((((qp[r] + qm[r]) + qyp) + qym) + qzp) + qzm;
            const float tq = fmaf(-6.0f, qc[r], q6);
            const float t2 = fmaf(c.half_dt, tq, t1);
            const float k1 =
                fmaf(-c.inv_mu2, seg_eval<ODD>(s1, pc[r]), fmaf(c.c_lap, t1, -c.eta * qc[r]));
            const float ptc = fmaf(c.half_dt, k1, qc[r]);
            const float k2 =
                fmaf(-c.inv_mu2, seg_eval<ODD>(s2, ftc), fmaf(c.c_lap, t2, -c.eta * ptc));
            const float pnew = fmaf(c.dt, ptc, pc[r]);
            float qnew;
            if (NM == NOISE_HASH) qnew = fmaf(c.dt, k2, fmaf(c.ns_l2, g, qc[r]));
            else if (NM == NOISE_ARRAY) qnew = qc[r] + fmaf(c.dt, k2, nzv);
            else qnew = fmaf(c.dt, k2, qc[r]);
            buf_store(pnew, at(PO, x), offs[r]);
            buf_store(qnew, at(QO, x), offs[r]);
            if constexpr (CNT) cnt += wave_count(act && fabsf(pnew) > esc);
        }
        xK += planeK;
    };
    using B0 = std::integral_constant<int, 0>;
    using B1 = std::integral_constant<int, 1>;
    int x = x0;
    for (; x + 4 <= x1; x += 4) {
        plane_step(B0{}, x + 0, p0, q0, p1, q1, p2, q2, p3, q3, hA, hB);
        plane_step(B1{}, x + 1, p1, q1, p2, q2, p3, q3, p0, q0, hB, hA);
        plane_step(B0{}, x + 2, p2, q2, p3, q3, p0, q0, p1, q1, hA, hB);
        plane_step(B1{}, x + 3, p3, q3, p0, q0, p1, q1, p2, q2, hB, hA);
    }
    // remainder: canonical roles, explicit rotation; buffers keep alternating
    auto rotate = [&]() {
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            p0[r] = p1[r]; q0[r] = q1[r]; p1[r] = p2[r]; q1[r] = q2[r];
            p2[r] = p3[r]; q2[r] = q3[r];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) hA[k] = hB[k];
    };
    if (x < x1) {
        plane_step(B0{}, x, p0, q0, p1, q1, p2, q2, p3, q3, hA, hB);
        rotate();
        ++x;
    }
    if (x < x1) {
        plane_step(B1{}, x, p0, q0, p1, q1, p2, q2, p3, q3, hA, hB);
        rotate();
        ++x;
    }
    if (x < x1) plane_step(B0{}, x, p0, q0, p1, q1, p2, q2, p3, q3, hA, hB);
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

template <typename R, typename Tab, int BY, bool CNT>
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
    unsigned cnt = rk2_real_tile<R, Tab, BY, CNT>(phi, pi, phi_out, pi_out, noise, c, tab, nx,
                                                  ny, nz, tp, x0, x1, noise_seed, hash_noise,
                                                  esc);
    if (CNT && cnt && (threadIdx.x & 63) == 0) atomicAdd(vac, (unsigned long long)cnt);
}

template <bool ODD, int BY, int NM>
__global__ __launch_bounds__(64 * BY) void k_rk2_real_f32(
    const float* __restrict__ phi, const float* __restrict__ pi, float* __restrict__ phi_out,
    float* __restrict__ pi_out, const float* __restrict__ noise, StepScalars<float> c,
    TableF32<ODD> tab, int nx, int ny, int nz, TileGrid g, uint64_t noise_seed,
    unsigned long long* __restrict__ vac, float esc)
{
    int tp, ch;
    g.decode((int)blockIdx.x, (int)gridDim.x, tp, ch);
    const int x0 = ch * g.cx;
    const int x1 = min(x0 + g.cx, nx);
    unsigned cnt = rk2_real_tile_f32<ODD, BY, NM>(phi, pi, phi_out, pi_out, noise, c, tab, nx,
                                                  ny, nz, tp, x0, x1, noise_seed, vac != nullptr,
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

// CNT: also count the sites with |phi_new| > esc into *vac (a separate instantiation, so the
// plain step carries no per-row test of the flag).
template <bool ODD, int TY, int NM, bool CNT>
__global__ __launch_bounds__(1024) void k_rk2_real_rows(
    const float* __restrict__ phi, const float* __restrict__ pi, float* __restrict__ phi_out,
    float* __restrict__ pi_out, const float* __restrict__ noise, StepScalars<float> c,
    TableF32<ODD> tab, int nx, int ny, int nz, int W, TileGrid g, uint64_t noise_seed,
    unsigned long long* __restrict__ vac, float esc)
{
    extern __shared__ float lds_rows[];
    int tp, ch;
    g.decode((int)blockIdx.x, (int)gridDim.x, tp, ch);
    const int x0 = ch * g.cx;
    const int x1 = min(x0 + g.cx, nx);
    unsigned cnt = rk2_real_rows_tile<ODD, TY, NM, CNT>(phi, pi, phi_out, pi_out, noise, c, tab,
                                                        nx, ny, nz, W, tp, x0, x1, noise_seed,
                                                        esc, lds_rows);
    if (CNT && cnt && (threadIdx.x & 63) == 0) atomicAdd(vac, (unsigned long long)cnt);
}

// ---------------------------------------------------------------------------------------
// Complex float32 row-tile kernel (LDS staging; default for the complex field).  Same tile,
// window and LDS scheme as the real row kernel with four fields (phi1, phi2, pi1, pi2): 4 + 8/TY
// global loads per site.  Force: rho = |phi|, V'(rho) from the (even) segment table,
// dV_i = V'(rho) phi_i / max(rho, 1e-30) (latticeSimComplex_numba.py:2130-2146), Z_N term
// (:2147-2152) for integer orders as Im((phi1 + i phi2)^N) / rho^N, otherwise sin(N atan2).
// Noise: (cos, sin) of one hash draw per site, as the reference's complex kernels.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void gauss_pair_f32_l2(uint64_t seed, uint32_t word_mask, float& zc,
                                                  float& zs)
{
    uint64_t x1, x2;
    hash_states(seed, x1, x2);
    const float u1 = fmaxf(unit_f32(x1, word_mask), 1e-12f);
    const float u2 = (float)((uint32_t)(x2 >> 32) & word_mask) * 0x1p-32f;
    const float r = __builtin_amdgcn_sqrtf(-__builtin_amdgcn_logf(u1));
    zc = r * __builtin_amdgcn_cosf(u2);
    zs = r * __builtin_amdgcn_sinf(u2);
}

// V'(rho) of the complex field's radial potential: the float32 segment table ...
// fetch() issues the gather, eval() finishes V' from what it returned, so a caller can issue
// the gathers of several sites before it needs any of them.  A Radial with kBatch instead
// evaluates a whole plane's radii (both stages, every row of the wave) in one wave-collective
// batch() call after the next plane's loads are issued; kLdsFloats of LDS past the staging
// area are its own, filled once per block by block_init().
struct TableRadial {
    TableF32<false> t;
    using Pre = float2;
    static constexpr bool kEarly = true;  // issue the gathers of a plane before its loads
    static constexpr bool kBatch = false;
    static constexpr int kLdsFloats = 0;
    __device__ __forceinline__ void block_init(float*) {}
    __device__ __forceinline__ Pre fetch(float rho) const
    {
        const __amdgpu_buffer_rsrc_t T = make_rsrc(t.ab, (uint32_t)(t.last + 1) * 8u);
        return seg_load<false>(T, t, rho);
    }
    __device__ __forceinline__ float eval(const Pre& s, float rho) const
    {
        return seg_eval<false>(s, rho);
    }
    __device__ __forceinline__ float operator()(float rho) const { return eval(fetch(rho), rho); }
};

// ... or the learned network at this step's temperature, read from its per-step grid in the
// compressed coordinate g = asinh(rho / rho_s) * gscale (segments (y_i, y_{i+1} - y_i), linear
// in g), with the exact tree term and the fixed envelope of the learned part:
// V' = rho (lam rho^2 - m2 + out_scale y / (1 + (rho / rho_cut)^2)), formed in float64.
struct MlpGridRadial {
    const float2* seg;
    int last;  // number of segments - 1
    float inv_rho_s, gscale, out_scale, inv_rho_cut;
    double lam, m2;
    struct Pre {
        float a, b, frac;  // segment (y_i, y_{i+1} - y_i) and the position inside it
    };
    // the gathers are issued where they are used: issuing them before the next plane's loads
    // (three registers per pending gather) measured slower at 512^3, 1.58-1.63 against
    // 1.52 ms per step
    static constexpr bool kEarly = false;
    static constexpr bool kBatch = false;
    static constexpr int kLdsFloats = 0;
    __device__ __forceinline__ void block_init(float*) {}
    __device__ __forceinline__ Pre fetch(float rho) const
    {
        const __amdgpu_buffer_rsrc_t Y = make_rsrc(seg, (uint32_t)(last + 1) * 8u);
        const float s = rho * inv_rho_s;
        const float g = __logf(s + __builtin_amdgcn_sqrtf(fmaf(s, s, 1.0f))) * gscale;
        const int idx = (int)__builtin_amdgcn_fmed3f(g, 0.0f, (float)last);
        using f2 = float __attribute__((ext_vector_type(2)));
        const f2 ab =
            __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(Y, idx * 8, 0, 0));
        return Pre{ab.x, ab.y, g - (float)idx};
    }
    __device__ __forceinline__ float eval(const Pre& p, float rho) const
    {
        const float y = fmaf(p.b, p.frac, p.a);
        const float q = rho * inv_rho_cut;
        const float th = out_scale * y * __builtin_amdgcn_rcpf(fmaf(q, q, 1.0f));
        const double r = (double)rho;
        return (float)(r * (fma(lam * r, r, -m2) + (double)th));
    }
    __device__ __forceinline__ float operator()(float rho) const { return eval(fetch(rho), rho); }
};

// dV_i = V'(rho) phi_i / max(rho, 1e-30) plus the Z_N term.  complex_rho + Radial::fetch
// start it (the gather), complex_force_finish completes it.
__device__ __forceinline__ float complex_rho(float a, float b)
{
    return __builtin_amdgcn_sqrtf(fmaf(a, a, b * b));
}

// ZN = false compiles the Z_N term out (launches with it switched off): a loop body a third
// of the size.
template <typename Radial, bool ZN = true>
__device__ __forceinline__ void complex_force_finish(const Radial& rad, const ZnParams<float>& zn,
                                                     float a, float b, float rho,
                                                     const typename Radial::Pre& pre, float& d1,
                                                     float& d2)
{
    const float inv = __builtin_amdgcn_rcpf(fmaxf(rho, 1e-30f));
    const float w = rad.eval(pre, rho) * inv;
    d1 = w * a;
    d2 = w * b;
    if (ZN && zn.active && rho > 1e-20f) {
        float sn;
        if (zn.ipow > 0) {
            const float ca = a * inv, sb = b * inv;
            float re = 1.0f, im = 0.0f;
            for (int e = 0; e < zn.ipow; ++e) {
                const float t = fmaf(re, ca, -im * sb);
                im = fmaf(re, sb, im * ca);
                re = t;
            }
            sn = im;
        } else {
            sn = sinf(zn.n * atan2f(b, a));
        }
        const float cz = zn.n * zn.dv * sn * inv * inv;
        d1 = fmaf(-cz, b, d1);
        d2 = fmaf(cz, a, d2);
    }
}

template <typename Radial, bool ZN = true>
__device__ __forceinline__ void complex_force_f32(const Radial& rad, const ZnParams<float>& zn,
                                                  float a, float b, float& d1, float& d2)
{
    const float rho = complex_rho(a, b);
    complex_force_finish<Radial, ZN>(rad, zn, a, b, rho, rad.fetch(rho), d1, d2);
}

// The segment table read through the vector cache or from a per-block LDS copy, chosen per
// block at the start of each launch.  A field spread over many segments within a wave (walls
// between the false and the true vacuum during the roll, string cores) costs the vector L1 one
// tag lookup per distinct cache line of a gather, which LDS reads do not pay; a narrow field is
// served from L1 without the copy.  The choice looks at the block's first plane: if the table
// indices of any wave there span more than kSpread segments, the block copies the table (at
// most kMaxSeg segments, 16-byte aligned) into LDS beside the staging area and reads it there.
// Both paths return the same entry.  With one 8-wave block per CU (four rows per thread, shared
// halo rows) the 64 KiB costs no waves.
struct TableRadialAuto {
    TableF32<false> t;
    const float2* lt;     // the block's copy (LDS)
    mutable int use_lds;  // block-uniform
    using Pre = float2;
    static constexpr bool kEarly = true;
    static constexpr bool kBatch = false;
    static constexpr int kMaxSeg = 8192;
    static constexpr int kLdsFloats = 2 * kMaxSeg;
    static constexpr int kSpread = 128;  // segments: 8 cache lines of 16 entries
    __device__ __forceinline__ void block_init(float* lds)
    {
        lt = reinterpret_cast<const float2*>(lds);
        use_lds = 0;
    }
    __device__ __forceinline__ uint32_t index(float rho) const
    {
        return (uint32_t)(int)__builtin_amdgcn_fmed3f(fmaf(rho, t.dinv, t.off), 0.0f,
                                                      (float)t.last);
    }
    // all threads of the block, with the block's first plane (phi1, phi2 rows) loaded
    template <int TY>
    __device__ __forceinline__ void stage(const float (&a)[TY], const float (&b)[TY]) const
    {
        int lo = t.last, hi = 0;
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            const int i = (int)index(complex_rho(a[r], b[r]));
            lo = min(lo, i);
            hi = max(hi, i);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lo = min(lo, __shfl_xor(lo, o));
            hi = max(hi, __shfl_xor(hi, o));
        }
        use_lds = __builtin_amdgcn_readfirstlane(__syncthreads_or(hi - lo > kSpread));
        if (use_lds) {
            const int n2 = t.last + 1, n4 = n2 >> 1;
            const int nt = (int)(blockDim.x * blockDim.y);
            const int tid = (int)(threadIdx.y * blockDim.x + threadIdx.x);
            const float4* src = reinterpret_cast<const float4*>(t.ab);
            float4* dst = reinterpret_cast<float4*>(const_cast<float2*>(lt));
#pragma unroll 8
            for (int i = tid; i < n4; i += nt) dst[i] = src[i];
            if ((n2 & 1) && tid == 0) const_cast<float2*>(lt)[n2 - 1] = t.ab[n2 - 1];
            __syncthreads();
        }
    }
    // the tile runs one of two compiled plane loops per block, so the test is not per gather
    template <bool LDS>
    __device__ __forceinline__ Pre fetch_from(float rho) const
    {
        const uint32_t i = index(rho);
        if constexpr (LDS) {
            return lt[i];
        } else {
            const __amdgpu_buffer_rsrc_t T = make_rsrc(t.ab, (uint32_t)(t.last + 1) * 8u);
            using f2 = float __attribute__((ext_vector_type(2)));
            const f2 v = __builtin_bit_cast(
                f2, __builtin_amdgcn_raw_buffer_load_b64(T, (int)(i * 8u), 0, 0));
            return make_float2(v.x, v.y);
        }
    }
    __device__ __forceinline__ Pre fetch(float rho) const
    {
        return use_lds ? fetch_from<true>(rho) : fetch_from<false>(rho);
    }
    __device__ __forceinline__ float eval(const Pre& s, float rho) const
    {
        return seg_eval<false>(s, rho);
    }
};

// radial forms that prepare per-block data inside the tile, after its first loads
template <typename Radial>
struct radial_stages : std::false_type {};
template <>
struct radial_stages<TableRadialAuto> : std::true_type {};

// the radial's fetch inside the tile's plane loop compiled for one table source (L)
template <typename Radial, typename L>
__device__ __forceinline__ typename Radial::Pre rad_fetch(const Radial& rad, float rho, L)
{
    if constexpr (radial_stages<Radial>::value)
        return rad.template fetch_from<L::value>(rho);
    else
        return rad.fetch(rho);
}

template <int TY, int NM, typename Radial, bool CNT, bool HS = false, bool ZN = true>
__device__ __forceinline__ unsigned rk2_complex_rows_tile(
    const float* __restrict__ phi1, const float* __restrict__ phi2, const float* __restrict__ pi1,
    const float* __restrict__ pi2, float* __restrict__ o1, float* __restrict__ o2,
    float* __restrict__ o3, float* __restrict__ o4, const float* __restrict__ noise1,
    const float* __restrict__ noise2, const StepScalars<float>& c, const Radial& rad,
    const ZnParams<float>& zn, int nx, int ny, int nz, int W, int tile, int x0, int x1,
    uint64_t noise_seed, float esc2, float* __restrict__ lds)
{
    constexpr int NF = 4;                 // phi1, phi2, pi1, pi2
    constexpr int K = 2 * NF * TY + 1;    // LDS words per z column
    const int G = (int)blockDim.y / W;
    // blockDim.x is one wave: threadIdx.y and the row offsets derived from it are wave-uniform
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y), grp = w / W, zw = w - grp * W;
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
    // HS (two row groups per block, consecutive y tiles): group 0 loads the block's top halo
    // row and reads its bottom halo from group 1's first staged row; group 1 loads the bottom
    // halo and reads its top halo from group 0's last staged row
    constexpr int NH = HS ? NF : 2 * NF;
    const bool g0 = grp == 0;
    const float* Lh = Lz + (g0 ? 64 * W * K : (TY - 1) - 64 * W * K);

    const int nyz = ny * nz;
    const size_t plane = (size_t)nyz * 4u;
    const uint32_t pb = (uint32_t)plane;
    const char* F[NF] = {reinterpret_cast<const char*>(phi1), reinterpret_cast<const char*>(phi2),
                         reinterpret_cast<const char*>(pi1), reinterpret_cast<const char*>(pi2)};
    char* O[NF] = {reinterpret_cast<char*>(o1), reinterpret_cast<char*>(o2),
                   reinterpret_cast<char*>(o3), reinterpret_cast<char*>(o4)};
    auto at = [&](const char* bp, int x) { return make_rsrc(bp + (size_t)x * plane, pb); };
    // z column offset per lane (vector register), row offsets wave-uniform (scalar registers);
    // idle lanes store at the plane's end (< 2^31 bytes, no wrap), where stores are dropped
    const uint32_t vo = (uint32_t)zc * 4u;
    const uint32_t vs = act ? vo : pb;
    uint32_t so[TY];
#pragma unroll
    for (int r = 0; r < TY; ++r) so[r] = (uint32_t)((yb + r) * nz) * 4u;
    const uint32_t som = (uint32_t)((yb > 0 ? yb - 1 : ny - 1) * nz) * 4u;
    const uint32_t sop = (uint32_t)((yb + TY < ny ? yb + TY : 0) * nz) * 4u;
    const uint64_t planeK = (uint64_t)(uint32_t)nyz * (uint64_t)kNoiseMult;
    const uint64_t col0K = (uint64_t)(uint32_t)(yb * nz + zc) * (uint64_t)kNoiseMult;
    const uint64_t rowK = (uint64_t)(uint32_t)nz * (uint64_t)kNoiseMult;
    uint64_t xK = (uint64_t)(uint32_t)x0 * planeK;

    float v0[NF][TY], v1[NF][TY], v2[NF][TY], v3[NF][TY] = {};
    {
        const int xm = x0 - 1 >= 0 ? x0 - 1 : nx - 1;
        const int xq = x0 + 1 < nx ? x0 + 1 : 0;
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                v0[f][r] = buf_load_s(at(F[f], xm), vo, so[r]);
                v1[f][r] = buf_load_s(at(F[f], x0), vo, so[r]);
                v2[f][r] = buf_load_s(at(F[f], xq), vo, so[r]);
            }
    }
    // [f] row y0-1, [NF + f] row y0+TY; HS: [f] the group's one loaded halo row
    float hA[NH], hB[NH] = {};
    const uint32_t soh = g0 ? som : sop;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        if constexpr (HS) {
            hA[f] = buf_load_s(at(F[f], x0), vo, soh);
        } else {
            hA[f] = buf_load_s(at(F[f], x0), vo, som);
            hA[NF + f] = buf_load_s(at(F[f], x0), vo, sop);
        }
    }
    if constexpr (radial_stages<Radial>::value) rad.stage(v1[0], v1[1]);
    unsigned cnt = 0;

    auto plane_step = [&](auto ldsc, auto bufc, int x, const float (&vm)[NF][TY],
                          const float (&vc)[NF][TY],
                          const float (&vp)[NF][TY], float (&vn)[NF][TY],
                          const float (&hc)[NH], float (&hn)[NH]) {
        constexpr int b = decltype(bufc)::value;
        // Both stages' radii and table gathers of every row depend only on this plane's
        // fields: they are issued before the next plane's loads, so waiting for a gather
        // (vector memory operations complete in issue order) never waits for the next plane.
        float rho1[TY], rho2[TY];
        typename Radial::Pre P1[TY], P2[TY];
        if constexpr (Radial::kEarly) {
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                rho1[r] = complex_rho(vc[0][r], vc[1][r]);
                P1[r] = rad_fetch(rad, rho1[r], ldsc);
                rho2[r] = complex_rho(fmaf(c.half_dt, vc[2][r], vc[0][r]),
                                      fmaf(c.half_dt, vc[3][r], vc[1][r]));
                P2[r] = rad_fetch(rad, rho2[r], ldsc);
            }
        }
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int r = 0; r < TY; ++r) Lz[(b * NF + f) * TY + r] = vc[f][r];
        if (x + 1 < x1) {  // the chunk's last plane has no next plane
            const int xa = x + 1 < nx ? x + 1 : 0;
            const int xb = xa + 1 < nx ? xa + 1 : 0;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
#pragma unroll
                for (int r = 0; r < TY; ++r) vn[f][r] = buf_load_s(at(F[f], xb), vo, so[r]);
                if constexpr (HS) {
                    hn[f] = buf_load_s(at(F[f], xa), vo, soh);
                } else {
                    hn[f] = buf_load_s(at(F[f], xa), vo, som);
                    hn[NF + f] = buf_load_s(at(F[f], xa), vo, sop);
                }
            }
        }
        if constexpr (Radial::kBatch) {
            // every lane of the wave takes part (idle lanes carry radii of row offset 0)
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                rho1[r] = complex_rho(vc[0][r], vc[1][r]);
                rho2[r] = complex_rho(fmaf(c.half_dt, vc[2][r], vc[0][r]),
                                      fmaf(c.half_dt, vc[3][r], vc[1][r]));
            }
            rad.template batch<TY>(rho1, rho2, P1, P2);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            float sum[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                float yp, ym;
                if constexpr (HS) {
                    const float hl = (r == 0 || r + 1 == TY) ? Lh[(b * NF + f) * TY] : 0.0f;
                    yp = r + 1 < TY ? vc[f][r + 1 < TY ? r + 1 : r] : (g0 ? hl : hc[f]);
                    ym = r > 0 ? vc[f][r > 0 ? r - 1 : 0] : (g0 ? hc[f] : hl);
                } else {
                    yp = r + 1 < TY ? vc[f][r + 1 < TY ? r + 1 : r] : hc[NF + f];
                    ym = r > 0 ? vc[f][r > 0 ? r - 1 : 0] : hc[f];
                }
                const float zpv = Lzp[(b * NF + f) * TY + r], zmv = Lzm[(b * NF + f) * TY + r];
                sum[f] = ((((vp[f][r] + vm[f][r]) + yp) + ym) + zpv) + zmv;
            }
            const float a = vc[0][r], bb = vc[1][r], u = vc[2][r], v = vc[3][r];
            const float t1a = fmaf(-6.0f, a, sum[0]), t1b = fmaf(-6.0f, bb, sum[1]);
            const float fa = fmaf(c.half_dt, u, a), fb = fmaf(c.half_dt, v, bb);
            float n1 = 0.0f, n2 = 0.0f;
            if (NM == NOISE_ARRAY) {
                n1 = buf_load_s(at(reinterpret_cast<const char*>(noise1), x), vo, so[r]);
                if (noise2)
                    n2 = buf_load_s(at(reinterpret_cast<const char*>(noise2), x), vo, so[r]);
            }
            float g1 = 0.0f, g2 = 0.0f;
            if (NM == NOISE_HASH)
                gauss_pair_f32_l2((col0K + (uint64_t)r * rowK + xK) ^ noise_seed, c.word_mask, g1,
                                  g2);
            float d1, d2;
            if constexpr (Radial::kEarly || Radial::kBatch)
                complex_force_finish<Radial, ZN>(rad, zn, a, bb, rho1[r], P1[r], d1, d2);
            else
                complex_force_f32<Radial, ZN>(rad, zn, a, bb, d1, d2);
            const float k1a = fmaf(-c.inv_mu2, d1, fmaf(c.c_lap, t1a, -c.eta * u));
            const float k1b = fmaf(-c.inv_mu2, d2, fmaf(c.c_lap, t1b, -c.eta * v));
            const float ga = fmaf(c.half_dt, k1a, u), gb = fmaf(c.half_dt, k1b, v);
            const float t2a = fmaf(c.half_dt, fmaf(-6.0f, u, sum[2]), t1a);
            const float t2b = fmaf(c.half_dt, fmaf(-6.0f, v, sum[3]), t1b);
            float e1, e2;
            if constexpr (Radial::kEarly || Radial::kBatch)
                complex_force_finish<Radial, ZN>(rad, zn, fa, fb, rho2[r], P2[r], e1, e2);
            else
                complex_force_f32<Radial, ZN>(rad, zn, fa, fb, e1, e2);
            const float k2a = fmaf(-c.inv_mu2, e1, fmaf(c.c_lap, t2a, -c.eta * ga));
            const float k2b = fmaf(-c.inv_mu2, e2, fmaf(c.c_lap, t2b, -c.eta * gb));
            const float na = fmaf(c.dt, ga, a), nb = fmaf(c.dt, gb, bb);
            float pa, pbv;
            if (NM == NOISE_HASH) {
                pa = fmaf(c.dt, k2a, fmaf(c.ns_l2, g1, u));
                pbv = fmaf(c.dt, k2b, fmaf(c.ns_l2, g2, v));
            } else if (NM == NOISE_ARRAY) {
                pa = u + fmaf(c.dt, k2a, n1);
                pbv = v + fmaf(c.dt, k2b, n2);
            } else {
                pa = fmaf(c.dt, k2a, u);
                pbv = fmaf(c.dt, k2b, v);
            }
            buf_store_s(na, at(O[0], x), vs, so[r]);
            buf_store_s(nb, at(O[1], x), vs, so[r]);
            buf_store_s(pa, at(O[2], x), vs, so[r]);
            buf_store_s(pbv, at(O[3], x), vs, so[r]);
            if constexpr (CNT) cnt += wave_count(act && fmaf(na, na, nb * nb) > esc2);
        }
        xK += planeK;
    };
    using B0 = std::integral_constant<int, 0>;
    using B1 = std::integral_constant<int, 1>;
    auto rotate = [&]() {
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                v0[f][r] = v1[f][r]; v1[f][r] = v2[f][r]; v2[f][r] = v3[f][r];
            }
#pragma unroll
        for (int k = 0; k < NH; ++k) hA[k] = hB[k];
    };
    auto march = [&](auto ldsc) {
        int x = x0;
        for (; x + 4 <= x1; x += 4) {
            plane_step(ldsc, B0{}, x + 0, v0, v1, v2, v3, hA, hB);
            plane_step(ldsc, B1{}, x + 1, v1, v2, v3, v0, hB, hA);
            plane_step(ldsc, B0{}, x + 2, v2, v3, v0, v1, hA, hB);
            plane_step(ldsc, B1{}, x + 3, v3, v0, v1, v2, hB, hA);
        }
        if (x < x1) { plane_step(ldsc, B0{}, x, v0, v1, v2, v3, hA, hB); rotate(); ++x; }
        if (x < x1) { plane_step(ldsc, B1{}, x, v0, v1, v2, v3, hA, hB); rotate(); ++x; }
        if (x < x1) plane_step(ldsc, B0{}, x, v0, v1, v2, v3, hA, hB);
    };
    if constexpr (radial_stages<Radial>::value) {
        if (rad.use_lds)  // block-uniform
            march(std::true_type{});
        else
            march(std::false_type{});
    } else {
        march(std::false_type{});
    }
    return cnt;
}

// CNT: also count the sites with |phi_new|^2 > esc2 into *vac (separate instantiation).
// HS: two row groups per block that share their inner y-halo rows through LDS.
template <int TY, int NM, typename Radial, bool CNT, bool HS = false, bool ZN = true>
__global__ __launch_bounds__(512) void k_rk2_complex_rows(
    const float* __restrict__ phi1, const float* __restrict__ phi2, const float* __restrict__ pi1,
    const float* __restrict__ pi2, float* __restrict__ o1, float* __restrict__ o2,
    float* __restrict__ o3, float* __restrict__ o4, const float* __restrict__ noise1,
    const float* __restrict__ noise2, StepScalars<float> c, Radial rad, ZnParams<float> zn,
    int nx, int ny, int nz, int W, TileGrid g, uint64_t noise_seed,
    unsigned long long* __restrict__ vac, float esc2)
{
    extern __shared__ float lds_crows[];
    Radial r = rad;
    if constexpr (Radial::kLdsFloats > 0)
        r.block_init(lds_crows + (size_t)blockDim.y * 64 * (8 * TY + 1));
    int tp, ch;
    g.decode((int)blockIdx.x, (int)gridDim.x, tp, ch);
    const int x0 = ch * g.cx;
    const int x1 = min(x0 + g.cx, nx);
    unsigned cnt = rk2_complex_rows_tile<TY, NM, Radial, CNT, HS, ZN>(
        phi1, phi2, pi1, pi2, o1, o2, o3, o4, noise1, noise2, c, r, zn, nx, ny, nz, W, tp, x0,
        x1, noise_seed, esc2, lds_crows);
    if (CNT && cnt && (threadIdx.x & 63) == 0) atomicAdd(vac, (unsigned long long)cnt);
}

// Launch geometry shared by the real and complex launchers.
//   kind TIP_KERNEL_ROWS (real float32 default where it applies): by = y rows per thread
//   (1, 2 or 4; must divide ny), cx = x planes per block.
//   kind TIP_KERNEL_FLAT: by = rows of 64 z-sites per block (1, 2, 4 or 8), cx as above.
struct StepGeometry {
    int by;   // see above (0 = default)
    int cx;   // x planes per block (0 = choose from the lattice size)
    int xcd;  // 1 = XCD-aware block remap
    int kind; // TIP_KERNEL_AUTO / TIP_KERNEL_ROWS / TIP_KERNEL_FLAT
};
// Row-tile shape of a real float32 lattice: waves per row W, groups per block G, and whether
// the row kernel applies at all (nz <= 1024, ny divisible by TY).
struct RowShape {
    int W, G, TY;
    bool ok;
};
RowShape row_shape(int64_t ny, int64_t nz, int ty, int max_w = 16);
int default_rows_per_thread(int64_t nx, int64_t ny, int64_t nz);
// Rows per thread and target block count of the real-field row kernels (RK2, BAOAB, overdamped).
int real_rows_per_thread(int64_t nx, int64_t ny, int64_t nz);
int64_t real_row_blocks(int64_t nx, int64_t ny, int64_t nz);
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
// The same step with V'(rho, T) from a per-step grid of the network (tip_mlp.hip): one small
// kernel tabulates the network at T on npts points of its compressed coordinate into `grid`
// ((npts - 1) float2), then the row-tile step reads it.
int launch_complex_mlp_grid(const float* phi1, const float* phi2, const float* pi1,
                            const float* pi2, float* o1, float* o2, float* o3, float* o4,
                            const StepScalars<float>& c, const float* mlp_params,
                            const tip_mlp_features* fm, double T, const ZnParams<float>& zn,
                            int64_t nx, int64_t ny, int64_t nz, uint64_t noise_seed, int hash,
                            unsigned long long* vac, float esc2, float2* grid, int64_t npts,
                            hipStream_t st);
ZnParams<float> make_zn32(double zn_n, double zn_dv, double zn_active);

}  // namespace tip
