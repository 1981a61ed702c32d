// tip_langevin_f32.hip -- float32 overdamped Euler-Maruyama and BAOAB steps, real and complex,
// reading the float32 segment form of the V' table.
#include "langevin_kernels.h"

using namespace tip;

namespace {

// Row-tile launch (see lv_rows_tile); `launched` is false when the shape does not fit the row
// tile (nz beyond the block, rows per thread not dividing ny) and the caller falls back to the
// flat tile.
template <int NF, bool ODD>
int launch_lv_rows(int op, const float* f0, const float* f1, const float* q0, const float* q1,
                   float* o0, float* o1, float* r0, float* r1, const LangevinScalars<float>& c,
                   const TableF32<ODD>& tab, const ZnParams<float>& zn, int64_t nx, int64_t ny,
                   int64_t nz, uint64_t noise_seed, int hash, const StepGeometry* geom,
                   hipStream_t st, bool& launched)
{
    launched = false;
    int ty = (geom && geom->by) ? geom->by
                                : (NF == 1 ? real_rows_per_thread(nx, ny, nz)
                                           : default_rows_per_thread(nx, ny, nz));
    if (NF == 2 && !(geom && geom->by) && ty > 2) ty = 2;
    RowShape rs = row_shape(ny, nz, ty, NF == 1 ? 16 : 8);
    if (!rs.ok) return TIP_OK;
    // complex rows of two or four waves: two row groups per block share their inner y-halo
    // rows through LDS when ny divides into pairs of tiles
    const bool hs = NF == 2 && (rs.W == 2 || rs.W == 4) && ny % (2 * rs.TY) == 0;
    if (hs) rs.G = 2;
    TileGrid tg;
    tg.tiles = (int)((ny + (int64_t)rs.G * rs.TY - 1) / ((int64_t)rs.G * rs.TY));
    const int64_t target = NF == 1 ? real_row_blocks(nx, ny, nz) : 1024;
    int64_t cx = (geom && geom->cx) ? geom->cx : (nx * tg.tiles) / target;
    if (!(geom && geom->cx)) cx = cx < 1 ? 1 : (cx > 32 ? 32 : cx);
    TIP_REQUIRE(cx >= 1 && cx <= 4096, "geometry: planes per block out of range");
    tg.cx = (int)cx;
    tg.chunks = (int)((nx + cx - 1) / cx);
    const int64_t nb64 = (int64_t)tg.tiles * tg.chunks;
    TIP_REQUIRE(nb64 < (int64_t(1) << 31), "geometry: too many blocks");
    const int nb = (int)nb64;
    tg.xcd_remap = ((geom ? geom->xcd : 1) && nb % 8 == 0) ? 1 : 0;
    const size_t lds = (size_t)rs.G * (64 * rs.W) * (2 * NF * rs.TY + 1) * sizeof(float);
    const dim3 blk(64, rs.W * rs.G);
#define TIP_LV_ROWS_Z(TYV, OPV, HSV, ZNV)                                                     \
    hipLaunchKernelGGL((k_langevin_rows<NF, TYV, OPV, ODD, NF == 2 && HSV, NF == 1 || ZNV>),     \
                       dim3(nb), blk, lds, st, f0, f1, q0, q1, o0, o1, r0, r1, c, tab, zn,        \
                       (int)nx, (int)ny, (int)nz, rs.W, tg, noise_seed, hash)
// the Z_N term compiled out of complex launches that switch it off
#define TIP_LV_ROWS_H(TYV, OPV, HSV)                 \
    do {                                             \
        if (NF == 2 && !zn.active)                   \
            TIP_LV_ROWS_Z(TYV, OPV, HSV, false);     \
        else                                         \
            TIP_LV_ROWS_Z(TYV, OPV, HSV, true);      \
    } while (0)
#define TIP_LV_ROWS(TYV, OPV)                 \
    do {                                      \
        if (hs)                               \
            TIP_LV_ROWS_H(TYV, OPV, true);    \
        else                                  \
            TIP_LV_ROWS_H(TYV, OPV, false);   \
    } while (0)
#define TIP_LV_ROWS_OP(TYV)                                              \
    switch (op) {                                                        \
        case LV_OVERDAMPED: TIP_LV_ROWS(TYV, LV_OVERDAMPED); break;      \
        case LV_BAOAB_OPEN: TIP_LV_ROWS(TYV, LV_BAOAB_OPEN); break;      \
        case LV_BAOAB_FUSED: TIP_LV_ROWS(TYV, LV_BAOAB_FUSED); break;    \
        case LV_BAOAB_CLOSE: TIP_LV_ROWS(TYV, LV_BAOAB_CLOSE); break;    \
        default: TIP_REQUIRE(false, "langevin: unknown stage %d", op);   \
    }
    switch (rs.TY) {
        case 1: TIP_LV_ROWS_OP(1); break;
        case 4: TIP_LV_ROWS_OP(4); break;
        default: TIP_LV_ROWS_OP(2); break;
    }
#undef TIP_LV_ROWS_OP
#undef TIP_LV_ROWS
#undef TIP_LV_ROWS_H
#undef TIP_LV_ROWS_Z
    TIP_HIP(hipGetLastError());
    launched = true;
    return TIP_OK;
}

bool want_rows(const StepGeometry* geom)
{
    return !(geom && geom->kind == TIP_KERNEL_FLAT);
}

}  // namespace

namespace tip {

int langevin_real_f32(int op, const float* phi, const float* pi, float* phi_out, float* pi_out,
                      const LangevinScalars<float>& c, const SegTable& tab, int64_t nx,
                      int64_t ny, int64_t nz, uint64_t noise_seed, int hash, hipStream_t st,
                      const StepGeometry* geom)
{
    if (want_rows(geom)) {
        bool done = false;
        const ZnParams<float> zn{};
        int rc = tab.odd
            ? launch_lv_rows<1, true>(op, phi, nullptr, pi, nullptr, phi_out, nullptr, pi_out,
                                      nullptr, c, tab.as<true>(), zn, nx, ny, nz, noise_seed,
                                      hash, geom, st, done)
            : launch_lv_rows<1, false>(op, phi, nullptr, pi, nullptr, phi_out, nullptr, pi_out,
                                       nullptr, c, tab.as<false>(), zn, nx, ny, nz, noise_seed,
                                       hash, geom, st, done);
        if (rc || done) return rc;
        TIP_REQUIRE(!(geom && geom->kind == TIP_KERNEL_ROWS),
                    "langevin f32: the row kernel needs nz <= 1024 and rows per thread dividing "
                    "ny");
    }
    if (tab.odd)
        return launch_langevin_real<float, TableF32<true>>(op, phi, pi, phi_out, pi_out, c,
                                                           tab.as<true>(), nx, ny, nz,
                                                           noise_seed, hash, st);
    return launch_langevin_real<float, TableF32<false>>(op, phi, pi, phi_out, pi_out, c,
                                                        tab.as<false>(), nx, ny, nz, noise_seed,
                                                        hash, st);
}

int langevin_complex_f32(int op, const float* phi1, const float* phi2, const float* pi1,
                         const float* pi2, float* phi1_out, float* phi2_out, float* pi1_out,
                         float* pi2_out, const LangevinScalars<float>& c, const SegTable& tab,
                         const ZnParams<float>& zn, int64_t nx, int64_t ny, int64_t nz,
                         uint64_t noise_seed, int hash, hipStream_t st,
                         const StepGeometry* geom)
{
    if (want_rows(geom)) {
        bool done = false;
        int rc = launch_lv_rows<2, false>(op, phi1, phi2, pi1, pi2, phi1_out, phi2_out, pi1_out,
                                          pi2_out, c, tab.as<false>(), zn, nx, ny, nz,
                                          noise_seed, hash, geom, st, done);
        if (rc || done) return rc;
        TIP_REQUIRE(!(geom && geom->kind == TIP_KERNEL_ROWS),
                    "langevin c32: the row kernel needs nz <= 512 and rows per thread dividing "
                    "ny");
    }
    return launch_langevin_complex<float, TableF32<false>>(op, phi1, phi2, pi1, pi2, phi1_out,
                                                           phi2_out, pi1_out, pi2_out, c,
                                                           tab.as<false>(), zn, nx, ny, nz,
                                                           noise_seed, hash, st);
}

}  // namespace tip

namespace {

// The float32 kernels read the (A, B) segment form of the float64 device table, prepared into
// the per-device scratch (slot 0, as the float32 RK2 drop-in).  The caller holds a
// ScratchLease on slot 0 across its launches.
int prepared_table(const char* who, int64_t nx, int64_t ny, int64_t nz, const double* vp_table,
                   int64_t npts, double tmin, double dinv, int odd, double eta, hipStream_t st,
                   SegTable& out)
{
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(vp_table, "%s: NULL table", who);
    TIP_REQUIRE(npts >= 2 && npts < (int64_t(1) << 30), "%s: vp_npts out of range", who);
    TIP_REQUIRE(eta > 0.0, "%s: eta / gamma must be positive", who);
    float2* ab =
        reinterpret_cast<float2*>(scratch_get((size_t)(npts - 1) * sizeof(float2), 0, st));
    TIP_REQUIRE(ab != nullptr, "%s: scratch allocation failed", who);
    rc = prepare_table_f32(vp_table, npts, tmin, dinv, ab, st);
    if (rc) return rc;
    out = make_table32(ab, tmin, dinv, npts, odd);
    return TIP_OK;
}

int check_stage(const char* who, int stage, const void* phi, const void* pi, const void* phi_out,
                const void* pi_out)
{
    TIP_REQUIRE(stage == LV_BAOAB_OPEN || stage == LV_BAOAB_FUSED || stage == LV_BAOAB_CLOSE,
                "%s: stage must be TIP_BAOAB_OPEN, _FUSED or _CLOSE (got %d)", who, stage);
    TIP_REQUIRE(phi && pi && pi_out, "%s: NULL field", who);
    if (stage != LV_BAOAB_CLOSE) {
        TIP_REQUIRE(phi_out, "%s: NULL phi_out", who);
        TIP_REQUIRE(phi_out != phi, "%s: phi_out must not alias phi", who);
    }
    return TIP_OK;
}

}  // namespace

extern "C" int tip_overdamped_f32(const float* phi, float* phi_out, double dt, double dx,
                                  double eta_eff, double mu, double inv_a2,
                                  const double* vp_table, double vp_tmin, double vp_dinv,
                                  int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny,
                                  int64_t nz, uint64_t noise_seed, double noise_scale,
                                  void* stream)
{
    hipStream_t st = as_stream(stream);
    TIP_REQUIRE(phi && phi_out && phi != phi_out, "overdamped f32: NULL or aliased field");
    SegTable tab;
    ScratchLease lease{0, st};
    int rc = prepared_table("overdamped f32", nx, ny, nz, vp_table, vp_npts, vp_tmin, vp_dinv,
                            vp_odd, eta_eff, st, tab);
    if (rc) return rc;
    return langevin_real_f32(LV_OVERDAMPED, phi, nullptr, phi_out, nullptr,
                             make_langevin<float>(dt, dx, eta_eff, mu, inv_a2, inv_a2,
                                                  noise_scale),
                             tab, nx, ny, nz, noise_seed, noise_scale != 0.0, st, nullptr);
}

extern "C" int tip_baoab_f32(int stage, const float* phi, const float* pi, float* phi_out,
                             float* pi_out, double dt, double dx, double gamma, double mu,
                             double inv_a2, double inv_a2_prev, const double* vp_table,
                             double vp_tmin, double vp_dinv, int64_t vp_npts, int vp_odd,
                             int64_t nx, int64_t ny, int64_t nz, uint64_t noise_seed,
                             double noise_scale, void* stream)
{
    hipStream_t st = as_stream(stream);
    int rc = check_stage("baoab f32", stage, phi, pi, phi_out, pi_out);
    if (rc) return rc;
    SegTable tab;
    ScratchLease lease{0, st};
    rc = prepared_table("baoab f32", nx, ny, nz, vp_table, vp_npts, vp_tmin, vp_dinv, vp_odd,
                        gamma, st, tab);
    if (rc) return rc;
    return langevin_real_f32(stage, phi, pi, phi_out, pi_out,
                             make_langevin<float>(dt, dx, gamma, mu, inv_a2, inv_a2_prev,
                                                  noise_scale),
                             tab, nx, ny, nz, noise_seed, noise_scale != 0.0, st, nullptr);
}

extern "C" int tip_overdamped_c32(const float* phi1, const float* phi2, float* phi1_out,
                                  float* phi2_out, double dt, double dx, double eta_eff,
                                  double mu, double inv_a2, const double* vp_table,
                                  double vp_tmin, double vp_dinv, int64_t vp_npts, double zn_n,
                                  double zn_dv, double zn_active, int64_t nx, int64_t ny,
                                  int64_t nz, uint64_t noise_seed, double noise_scale,
                                  void* stream)
{
    hipStream_t st = as_stream(stream);
    TIP_REQUIRE(phi1 && phi2 && phi1_out && phi2_out && phi1 != phi1_out && phi2 != phi2_out,
                "overdamped c32: NULL or aliased field");
    SegTable tab;
    ScratchLease lease{0, st};
    int rc = prepared_table("overdamped c32", nx, ny, nz, vp_table, vp_npts, vp_tmin, vp_dinv,
                            0, eta_eff, st, tab);
    if (rc) return rc;
    return langevin_complex_f32(LV_OVERDAMPED, phi1, phi2, nullptr, nullptr, phi1_out, phi2_out,
                                nullptr, nullptr,
                                make_langevin<float>(dt, dx, eta_eff, mu, inv_a2, inv_a2,
                                                     noise_scale),
                                tab, make_zn32(zn_n, zn_dv, zn_active), nx, ny, nz, noise_seed,
                                noise_scale != 0.0, st, nullptr);
}

extern "C" int tip_baoab_c32(int stage, const float* phi1, const float* phi2, const float* pi1,
                             const float* pi2, float* phi1_out, float* phi2_out, float* pi1_out,
                             float* pi2_out, double dt, double dx, double gamma, double mu,
                             double inv_a2, double inv_a2_prev, const double* vp_table,
                             double vp_tmin, double vp_dinv, int64_t vp_npts, double zn_n,
                             double zn_dv, double zn_active, int64_t nx, int64_t ny, int64_t nz,
                             uint64_t noise_seed, double noise_scale, void* stream)
{
    hipStream_t st = as_stream(stream);
    int rc = check_stage("baoab c32", stage, phi1, pi1, phi1_out, pi1_out);
    if (rc) return rc;
    rc = check_stage("baoab c32", stage, phi2, pi2, phi2_out, pi2_out);
    if (rc) return rc;
    SegTable tab;
    ScratchLease lease{0, st};
    rc = prepared_table("baoab c32", nx, ny, nz, vp_table, vp_npts, vp_tmin, vp_dinv, 0, gamma,
                        st, tab);
    if (rc) return rc;
    return langevin_complex_f32(stage, phi1, phi2, pi1, pi2, phi1_out, phi2_out, pi1_out, pi2_out,
                                make_langevin<float>(dt, dx, gamma, mu, inv_a2, inv_a2_prev,
                                                     noise_scale),
                                tab, make_zn32(zn_n, zn_dv, zn_active), nx, ny, nz, noise_seed,
                                noise_scale != 0.0, st, nullptr);
}
