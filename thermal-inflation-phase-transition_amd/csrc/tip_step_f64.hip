// tip_step_f64.hip -- float64 fused RK2 step (the drop-in parity path).
// Compiled with -ffp-contract=off: every product and sum rounds separately, in the order the
// reference writes them (simulation/latticeSimeRescale_numba.py:1912-1933,1945-1966), so the
// only differences from a float64 CPU run are the libm transcendentals of the hash noise.
#include "step_kernels.h"

using namespace tip;

namespace {

// ---------------------------------------------------------------------------------------
// float64 row tiles: the float32 row kernel's scheme (a group of W waves covers whole z rows,
// each thread TY consecutive y rows of its z column, marching along x with a register window,
// the plane's (phi, pi) staged in LDS for the z neighbours, two halo rows loaded) with the
// per-site arithmetic of rk2_site in the reference's order -- the same values in the same
// operations as the flat tile, so every geometry gives identical fields.  Both table lookups of
// a row depend only on the current plane, so their loads are issued before the LDS store and
// the next plane's loads (vector memory completes in issue order).  LDS: 2 buffers x 2 fields
// x TY rows of doubles per z column, K = 4 TY + 1 doubles (an odd stride of 2K dwords: the
// 16 lanes of an LDS pass hit distinct bank pairs).
// ---------------------------------------------------------------------------------------
template <int TY, int NM, bool CNT>
__device__ __forceinline__ unsigned rk2_real_rows_f64_tile(
    const double* __restrict__ phi, const double* __restrict__ pi, double* __restrict__ phi_out,
    double* __restrict__ pi_out, const double* __restrict__ noise, const StepScalars<double>& c,
    const TableF64& tab, int nx, int ny, int nz, int W, int tile, int x0, int x1,
    uint64_t noise_seed, double esc, double* __restrict__ lds)
{
    constexpr int K = 4 * TY + 1;
    const int G = (int)blockDim.y / W;
    const int w = (int)threadIdx.y, grp = w / W;
    const int z = (w - grp * W) * 64 + (int)threadIdx.x;
    const bool zact = z < nz;
    const int zc = zact ? z : 0;
    const int zp = zc + 1 < nz ? zc + 1 : 0, zm = zc > 0 ? zc - 1 : nz - 1;
    const int y0 = (tile * G + grp) * TY;
    const bool act = zact && y0 < ny;
    const int yb = y0 < ny ? y0 : 0;
    double* Lg = lds + (size_t)grp * (64 * W) * K;
    double* Lz = Lg + z * K;
    const double* Lzp = Lg + zp * K;
    const double* Lzm = Lg + zm * K;

    const int nyz = ny * nz;
    const size_t plane = (size_t)nyz * 8u;
    const uint32_t pb = (uint32_t)plane;  // < 2^32 by the dimension check
    const char* P = reinterpret_cast<const char*>(phi);
    const char* Q = reinterpret_cast<const char*>(pi);
    const char* PO = reinterpret_cast<const char*>(phi_out);
    const char* QO = reinterpret_cast<const char*>(pi_out);
    const char* NZ = reinterpret_cast<const char*>(noise);
    auto at = [&](const char* b, int x) { return make_rsrc(b + (size_t)x * plane, pb); };
    uint32_t off[TY], offs[TY];
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        off[r] = (uint32_t)((yb + r) * nz + zc) * 8u;
        // idle lanes store past the buffer end, where buffer stores are dropped
        offs[r] = act ? off[r] : 0xFFFFFFF0u;
    }
    const uint32_t offm = (uint32_t)((yb > 0 ? yb - 1 : ny - 1) * nz + zc) * 8u;
    const uint32_t offp = (uint32_t)((yb + TY < ny ? yb + TY : 0) * nz + zc) * 8u;
    const uint64_t planeK = (uint64_t)(uint32_t)nyz * (uint64_t)kNoiseMult;
    const uint64_t col0K = (uint64_t)(uint32_t)(yb * nz + zc) * (uint64_t)kNoiseMult;
    const uint64_t rowK = (uint64_t)(uint32_t)nz * (uint64_t)kNoiseMult;
    uint64_t xK = (uint64_t)(uint32_t)x0 * planeK;

    double p0[TY], q0[TY], p1[TY], q1[TY], p2[TY], q2[TY];
    double p3[TY] = {}, q3[TY] = {};  // zero where a chunk's last plane loads nothing
    {
        const int xm = x0 - 1 >= 0 ? x0 - 1 : nx - 1;
        const int xq = x0 + 1 < nx ? x0 + 1 : 0;
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            p0[r] = buf_load<double>(at(P, xm), off[r]);
            q0[r] = buf_load<double>(at(Q, xm), off[r]);
            p1[r] = buf_load<double>(at(P, x0), off[r]);
            q1[r] = buf_load<double>(at(Q, x0), off[r]);
            p2[r] = buf_load<double>(at(P, xq), off[r]);
            q2[r] = buf_load<double>(at(Q, xq), off[r]);
        }
    }
    double hA[4], hB[4] = {};  // phi y-1, pi y-1, phi y+TY, pi y+TY
    hA[0] = buf_load<double>(at(P, x0), offm);
    hA[1] = buf_load<double>(at(Q, x0), offm);
    hA[2] = buf_load<double>(at(P, x0), offp);
    hA[3] = buf_load<double>(at(Q, x0), offp);
    unsigned cnt = 0;

    auto plane_step = [&](auto bufc, int x, const double (&pm)[TY], const double (&qm)[TY],
                          const double (&pc)[TY], const double (&qc)[TY],
                          const double (&pp)[TY], const double (&qp)[TY], double (&pn)[TY],
                          double (&qn)[TY], const double (&hc)[4], double (&hn)[4]) {
        constexpr int b = decltype(bufc)::value;
        constexpr int oP = (b * 2 + 0) * TY, oQ = (b * 2 + 1) * TY;
        double ftc[TY];
        double2 g1[TY], g2[TY];
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            ftc[r] = pc[r] + c.half_dt * qc[r];
            g1[r] = tab.fetch(pc[r]);
            g2[r] = tab.fetch(ftc[r]);
        }
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            Lz[oP + r] = pc[r];
            Lz[oQ + r] = qc[r];
        }
        if (x + 1 < x1) {  // the chunk's last plane has no next plane
            const int xa = x + 1 < nx ? x + 1 : 0;
            const int xb = xa + 1 < nx ? xa + 1 : 0;
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                pn[r] = buf_load<double>(at(P, xb), off[r]);
                qn[r] = buf_load<double>(at(Q, xb), off[r]);
            }
            hn[0] = buf_load<double>(at(P, xa), offm);
            hn[1] = buf_load<double>(at(Q, xa), offm);
            hn[2] = buf_load<double>(at(P, xa), offp);
            hn[3] = buf_load<double>(at(Q, xa), offp);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            const double pyp = r + 1 < TY ? pc[r + 1 < TY ? r + 1 : r] : hc[2];
            const double pym = r > 0 ? pc[r > 0 ? r - 1 : 0] : hc[0];
            const double qyp = r + 1 < TY ? qc[r + 1 < TY ? r + 1 : r] : hc[3];
            const double qym = r > 0 ? qc[r > 0 ? r - 1 : 0] : hc[1];
            const double pzp = Lzp[oP + r], pzm = Lzm[oP + r];
            const double qzp = Lzp[oQ + r], qzm = Lzm[oQ + r];
            double nz_ = 0.0;
            if (NM == NOISE_ARRAY) nz_ = buf_load<double>(at(NZ, x), off[r]);
            if (NM == NOISE_HASH)
                nz_ = c.ns * gauss_f64((col0K + (uint64_t)r * rowK + xK) ^ noise_seed);
            double pnew, qnew;
            rk2_site<double>(c, pc[r], qc[r], pp[r], qp[r], pm[r], qm[r], pyp, qyp, pym, qym,
                             pzp, qzp, pzm, qzm, ftc[r], tab.eval(g1[r], pc[r]),
                             tab.eval(g2[r], ftc[r]), nz_, pnew, qnew);
            buf_store(pnew, at(PO, x), offs[r]);
            buf_store(qnew, at(QO, x), offs[r]);
            if constexpr (CNT) cnt += wave_count(act && fabs(pnew) > esc);
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

template <int TY, int NM, bool CNT>
__global__ __launch_bounds__(512) void k_rk2_real_rows_f64(
    const double* __restrict__ phi, const double* __restrict__ pi, double* __restrict__ phi_out,
    double* __restrict__ pi_out, const double* __restrict__ noise, StepScalars<double> c,
    TableF64 tab, int nx, int ny, int nz, int W, TileGrid g, uint64_t noise_seed,
    unsigned long long* __restrict__ vac, double esc)
{
    extern __shared__ double lds_f64rows[];
    int tp, ch;
    g.decode((int)blockIdx.x, (int)gridDim.x, tp, ch);
    const int x0 = ch * g.cx;
    const int x1 = min(x0 + g.cx, nx);
    unsigned cnt = rk2_real_rows_f64_tile<TY, NM, CNT>(phi, pi, phi_out, pi_out, noise, c, tab,
                                                      nx, ny, nz, W, tp, x0, x1, noise_seed, esc,
                                                      lds_f64rows);
    if (CNT && cnt && (threadIdx.x & 63) == 0) atomicAdd(vac, (unsigned long long)cnt);
}

TableF64 make_table(const double* t, double tmin, double dinv, int64_t npts, int odd)
{
    TableF64 tab;
    tab.t = t;
    tab.tmin = tmin;
    tab.dinv = dinv;
    tab.last = (int)(npts - 2);
    tab.odd = odd ? 1 : 0;
    return tab;
}

int launch_real_f64(const double* phi, const double* pi, double* phi_out, double* pi_out,
                    double dt, double dx, double eta_eff, double mu, double inv_a2,
                    const double* noise, const double* vp_table, double vp_tmin, double vp_dinv,
                    int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny, int64_t nz,
                    uint64_t noise_seed, double noise_scale, unsigned long long* vac,
                    double vac_esc, hipStream_t st)
{
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(phi && pi && phi_out && pi_out && vp_table, "rk2 f64: NULL field or table");
    TIP_REQUIRE(vp_npts >= 2 && vp_npts < (int64_t(1) << 30), "rk2 f64: vp_npts=%lld out of range",
                (long long)vp_npts);
    TIP_REQUIRE(phi != phi_out && pi != pi_out, "rk2 f64: output must not alias input");
    StepScalars<double> c = make_scalars<double>(dt, dx, eta_eff, mu, inv_a2, noise_scale);
    int hash = (noise == nullptr && noise_scale != 0.0) ? 1 : 0;
    const TableF64 tab = make_table(vp_table, vp_tmin, vp_dinv, vp_npts, vp_odd);
    const int nm = noise ? NOISE_ARRAY : (hash ? NOISE_HASH : NOISE_NONE);
    // row tiles (z rows of up to 8 waves): two y rows per thread on large lattices
    const int ty = (nx * ny * nz >= (int64_t(1) << 23) && ny % 2 == 0) ? 2 : 1;
    const RowShape rs = row_shape(ny, nz, ty, 8);
    if (rs.ok) {
        TileGrid tg;
        tg.tiles = (int)((ny + (int64_t)rs.G * rs.TY - 1) / ((int64_t)rs.G * rs.TY));
        int64_t cx = (nx * tg.tiles) / 2048;
        cx = cx < 1 ? 1 : (cx > 32 ? 32 : cx);
        tg.cx = (int)cx;
        tg.chunks = (int)((nx + cx - 1) / cx);
        const int64_t nb64 = (int64_t)tg.tiles * tg.chunks;
        TIP_REQUIRE(nb64 < (int64_t(1) << 31), "rk2 f64: too many blocks");
        tg.xcd_remap = nb64 % 8 == 0 ? 1 : 0;
        const size_t lds = (size_t)rs.G * (64 * rs.W) * (4 * rs.TY + 1) * sizeof(double);
        const dim3 blk(64, rs.W * rs.G);
#define TIP_LAUNCH_F64R(TYV, NMV, CNTV)                                                          \
    hipLaunchKernelGGL((k_rk2_real_rows_f64<TYV, NMV, CNTV>), dim3((int)nb64), blk, lds, st, phi, \
                       pi, phi_out, pi_out, noise, c, tab, (int)nx, (int)ny, (int)nz, rs.W, tg,   \
                       noise_seed, vac, vac_esc)
#define TIP_LAUNCH_F64R_C(TYV, NMV)            \
    do {                                       \
        if (vac)                               \
            TIP_LAUNCH_F64R(TYV, NMV, true);   \
        else                                   \
            TIP_LAUNCH_F64R(TYV, NMV, false);  \
    } while (0)
#define TIP_LAUNCH_F64R_NM(TYV)                                         \
    switch (nm) {                                                       \
        case NOISE_ARRAY: TIP_LAUNCH_F64R_C(TYV, NOISE_ARRAY); break;   \
        case NOISE_HASH: TIP_LAUNCH_F64R_C(TYV, NOISE_HASH); break;     \
        default: TIP_LAUNCH_F64R_C(TYV, NOISE_NONE); break;             \
    }
        if (rs.TY == 2) TIP_LAUNCH_F64R_NM(2)
        else TIP_LAUNCH_F64R_NM(1)
#undef TIP_LAUNCH_F64R_NM
#undef TIP_LAUNCH_F64R_C
#undef TIP_LAUNCH_F64R
        TIP_HIP(hipGetLastError());
        return TIP_OK;
    }
    // z rows wider than 8 waves: the flat tile
    StepGeometry g;
    TileGrid tg;
    int nb;
    rc = resolve_geometry(nullptr, nx, ny, nz, 0, 1, g, tg, nb);
    if (rc) return rc;
    if (vac)
        hipLaunchKernelGGL((k_rk2_real<double, TableF64, 4, true>), dim3(nb), dim3(64, 4), 0, st,
                           phi, pi, phi_out, pi_out, noise, c, tab, (int)nx, (int)ny, (int)nz, tg,
                           noise_seed, hash, vac, vac_esc);
    else
        hipLaunchKernelGGL((k_rk2_real<double, TableF64, 4, false>), dim3(nb), dim3(64, 4), 0, st,
                           phi, pi, phi_out, pi_out, noise, c, tab, (int)nx, (int)ny, (int)nz, tg,
                           noise_seed, hash, vac, vac_esc);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

}  // namespace

extern "C" int tip_rk2_fused_f64(const double* phi, const double* pi, double* phi_out,
                                 double* pi_out, double dt, double dx, double eta_eff, double mu,
                                 double inv_a2, const double* noise, const double* vp_table,
                                 double vp_tmin, double vp_dinv, int64_t vp_npts, int vp_odd,
                                 int64_t nx, int64_t ny, int64_t nz, uint64_t noise_seed,
                                 double noise_scale, unsigned long long* vac_count,
                                 double vac_esc, void* stream)
{
    return launch_real_f64(phi, pi, phi_out, pi_out, dt, dx, eta_eff, mu, inv_a2, noise, vp_table,
                           vp_tmin, vp_dinv, vp_npts, vp_odd, nx, ny, nz, noise_seed, noise_scale,
                           vac_count, vac_esc, as_stream(stream));
}

extern "C" int tip_rk2_step_table_f64(double* phi, double* pi, double dt, double dx,
                                      double eta_eff, double mu, const double* noise,
                                      double* phi_tmp, double* pi_tmp, double inv_a2,
                                      const double* vp_table, double vp_tmin, double vp_dinv,
                                      int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny,
                                      int64_t nz, uint64_t noise_seed, double noise_scale,
                                      unsigned long long* vac_count, double vac_esc,
                                      void* stream)
{
    hipStream_t st = as_stream(stream);
    TIP_REQUIRE(phi_tmp && pi_tmp, "rk2_step_table_f64: phi_tmp/pi_tmp are required");
    int rc = launch_real_f64(phi, pi, phi_tmp, pi_tmp, dt, dx, eta_eff, mu, inv_a2, noise,
                             vp_table, vp_tmin, vp_dinv, vp_npts, vp_odd, nx, ny, nz, noise_seed,
                             noise_scale, vac_count, vac_esc, st);
    if (rc) return rc;
    size_t bytes = (size_t)(nx * ny * nz) * sizeof(double);
    TIP_HIP(hipMemcpyAsync(phi, phi_tmp, bytes, hipMemcpyDeviceToDevice, st));
    TIP_HIP(hipMemcpyAsync(pi, pi_tmp, bytes, hipMemcpyDeviceToDevice, st));
    return TIP_OK;
}

extern "C" int tip_rk2_step_table_c64(
    double* phi1, double* phi2, double* pi1, double* pi2, double dt, double dx, double eta_eff,
    double mu, const double* noise1, const double* noise2, double* phi1_tmp, double* phi2_tmp,
    double* pi1_tmp, double* pi2_tmp, double inv_a2, const double* vp_table, double vp_tmin,
    double vp_dinv, int64_t vp_npts, double zn_n, double zn_dv, double zn_active, int64_t nx,
    int64_t ny, int64_t nz, uint64_t noise_seed, double noise_scale,
    unsigned long long* vac_count, double vac_esc, void* stream)
{
    hipStream_t st = as_stream(stream);
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(phi1 && phi2 && pi1 && pi2 && vp_table, "rk2 c64: NULL field or table");
    TIP_REQUIRE(phi1_tmp && phi2_tmp && pi1_tmp && pi2_tmp, "rk2 c64: tmp arrays are required");
    TIP_REQUIRE(vp_npts >= 2 && vp_npts < (int64_t(1) << 30), "rk2 c64: vp_npts out of range");
    TIP_REQUIRE(noise1 != nullptr || noise2 == nullptr, "rk2 c64: noise2 without noise1");
    StepScalars<double> c = make_scalars<double>(dt, dx, eta_eff, mu, inv_a2, noise_scale);
    ZnParams<double> zn;
    zn.n = zn_n;
    zn.dv = zn_dv;
    zn.active = (zn_active > 0.5 && zn_n > 0.5) ? 1 : 0;
    zn.ipow = 0;
    StepGeometry g;
    TileGrid tg;
    int nb;
    rc = resolve_geometry(nullptr, nx, ny, nz, 1, 1, g, tg, nb);
    if (rc) return rc;
    int hash = (noise1 == nullptr && noise_scale != 0.0) ? 1 : 0;
    hipLaunchKernelGGL((k_rk2_complex<double, TableF64, 4>), dim3(nb), dim3(64, 4), 0, st, phi1,
                       phi2, pi1, pi2, phi1_tmp, phi2_tmp, pi1_tmp, pi2_tmp, noise1, noise2, c,
                       make_table(vp_table, vp_tmin, vp_dinv, vp_npts, 0), zn, (int)nx, (int)ny,
                       (int)nz, tg, noise_seed, hash, vac_count, vac_esc * vac_esc);
    TIP_HIP(hipGetLastError());
    size_t bytes = (size_t)(nx * ny * nz) * sizeof(double);
    TIP_HIP(hipMemcpyAsync(phi1, phi1_tmp, bytes, hipMemcpyDeviceToDevice, st));
    TIP_HIP(hipMemcpyAsync(phi2, phi2_tmp, bytes, hipMemcpyDeviceToDevice, st));
    TIP_HIP(hipMemcpyAsync(pi1, pi1_tmp, bytes, hipMemcpyDeviceToDevice, st));
    TIP_HIP(hipMemcpyAsync(pi2, pi2_tmp, bytes, hipMemcpyDeviceToDevice, st));
    return TIP_OK;
}
