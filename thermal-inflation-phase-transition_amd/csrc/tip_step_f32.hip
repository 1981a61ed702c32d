// tip_step_f32.hip -- float32 fused RK2 step (the performance path), its launch geometry and
// the float32 segment table it reads.
#include <mutex>

#include "step_kernels.h"

using namespace tip;

namespace tip {

// Segment i of the reference's piecewise-linear lookup (latticeSimeRescale_numba.py:1923-1930)
// as V'(x) = A_i + B_i x: B_i = (t[i+1]-t[i]) * dinv, A_i = t[i] - B_i * (tmin + i/dinv),
// computed in float64 and rounded once.
__global__ void k_prepare_table_f32(const double* __restrict__ t, int nseg, double tmin,
                                    double dinv, float2* __restrict__ ab)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg) return;
    double b = (t[i + 1] - t[i]) * dinv;
    double xi = tmin + (double)i / dinv;
    double a = t[i] - b * xi;
    ab[i] = make_float2((float)a, (float)b);
}

int prepare_table_f32(const double* table_dev, int64_t npts, double tmin, double dinv,
                      float2* ab_dev, hipStream_t s)
{
    TIP_REQUIRE(table_dev && ab_dev, "table_prepare_f32: NULL pointer");
    TIP_REQUIRE(npts >= 2 && npts < (int64_t(1) << 30), "table_prepare_f32: npts out of range");
    TIP_REQUIRE(dinv > 0.0 && dinv < 1e300, "table_prepare_f32: dinv must be positive");
    int nseg = (int)(npts - 1);
    hipLaunchKernelGGL(k_prepare_table_f32, dim3((nseg + 255) / 256), dim3(256), 0, s, table_dev,
                       nseg, tmin, dinv, ab_dev);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

StepGeometry step_geometry(int64_t nx, int64_t ny, int64_t nz, int complex_field, int f64)
{
    (void)complex_field;
    (void)f64;
    StepGeometry g;
    g.by = 4;
    g.xcd = 1;
    int64_t per_plane = (ny * nz + 255) / 256;
    // Aim for about 4096 blocks (16 per CU) while marching at least 1 and at most 32 planes.
    int64_t cx = (nx * per_plane) / 4096;
    if (cx < 1) cx = 1;
    if (cx > 32) cx = 32;
    g.cx = (int)cx;
    return g;
}

int resolve_geometry(const StepGeometry* req, int64_t nx, int64_t ny, int64_t nz,
                     int complex_field, int f64, StepGeometry& g, TileGrid& tg, int& nblocks)
{
    g = step_geometry(nx, ny, nz, complex_field, f64);
    if (req) {
        TIP_REQUIRE(req->by == 0 || req->by == 1 || req->by == 2 || req->by == 4 || req->by == 8,
                    "geometry: rows per block must be 1, 2, 4 or 8 (got %d)", req->by);
        TIP_REQUIRE(req->cx >= 0 && req->cx <= 4096, "geometry: planes per block out of range");
        if (req->by) g.by = req->by;
        if (req->by && !req->cx) {  // re-derive the plane count for the new row count
            int64_t per_plane = (ny * nz + 64 * g.by - 1) / (64 * g.by);
            int64_t cx = (nx * per_plane) / 4096;
            g.cx = (int)(cx < 1 ? 1 : (cx > 32 ? 32 : cx));
        }
        if (req->cx) g.cx = req->cx;
        g.xcd = req->xcd ? 1 : 0;
    }
    tg.tiles = (int)((ny * nz + 64 * g.by - 1) / (64 * g.by));
    tg.chunks = (int)((nx + g.cx - 1) / g.cx);
    tg.cx = g.cx;
    int64_t nb = (int64_t)tg.tiles * tg.chunks;
    TIP_REQUIRE(nb < (int64_t(1) << 31), "geometry: too many blocks");
    nblocks = (int)nb;
    tg.xcd_remap = (g.xcd && nb % 8 == 0) ? 1 : 0;
    return TIP_OK;
}

// Rows per thread of the row kernel: 4 on large lattices (fewer halo loads per site, measured
// fastest at 256^3), 1 on small ones (more blocks to fill the machine); 4 and 2 need ny to be
// divisible by them.
int default_rows_per_thread(int64_t nx, int64_t ny, int64_t nz)
{
    if (nx * ny * nz >= (int64_t(1) << 23)) {
        if (ny % 4 == 0) return 4;
        if (ny % 2 == 0) return 2;
    }
    return 1;
}

// The real-field row kernels' default on large lattices: two rows per thread (79 registers,
// six waves per SIMD) and about 4096 blocks of 8 planes at 256^3 -- measured 46.8 us against
// 50.0 us for four rows (124 registers, four waves per SIMD) in 1024 blocks, and 47.7-48.7 us
// for 4, 6, 12 or 16 planes (BAOAB 45.2 against 49.3 us, overdamped 40.7 against 43.3 us).
static bool real_large(int64_t nx, int64_t ny, int64_t nz)
{
    return nx * ny * nz >= (int64_t(1) << 23);
}
int real_rows_per_thread(int64_t nx, int64_t ny, int64_t nz)
{
    return real_large(nx, ny, nz) && ny % 2 == 0 ? 2 : 1;
}
int64_t real_row_blocks(int64_t nx, int64_t ny, int64_t nz)
{
    return real_large(nx, ny, nz) ? 4096 : 1024;
}

RowShape row_shape(int64_t ny, int64_t nz, int ty, int max_w)
{
    RowShape r;
    r.W = (int)((nz + 63) / 64);
    r.G = r.W >= 4 ? 1 : 4 / r.W;
    r.TY = ty;
    r.ok = r.W <= max_w && (ty == 1 || ty == 2 || ty == 4) && ny % ty == 0;
    return r;
}

SegTable make_table32(const void* ab, double tmin, double dinv, int64_t npts, int odd)
{
    SegTable tab;
    tab.ab = reinterpret_cast<const float2*>(ab);
    tab.tmin = (float)tmin;
    tab.dinv = (float)dinv;
    tab.last = (int)(npts - 2);
    tab.odd = odd ? 1 : 0;
    tab.off = (float)(-tmin * dinv);
    return tab;
}

ZnParams<float> make_zn32(double zn_n, double zn_dv, double zn_active)
{
    ZnParams<float> zn;
    zn.n = (float)zn_n;
    zn.dv = (float)zn_dv;
    zn.active = (zn_active > 0.5 && zn_n > 0.5) ? 1 : 0;
    double r = zn_n - (double)(int64_t)zn_n;
    zn.ipow = (r == 0.0 && zn_n >= 1.0 && zn_n <= 32.0) ? (int)zn_n : 0;
    return zn;
}

int launch_real_f32(const float* phi, const float* pi, float* phi_out, float* pi_out,
                    const StepScalars<float>& c, const float* noise, const SegTable& tab,
                    int64_t nx, int64_t ny, int64_t nz, uint64_t noise_seed, int hash,
                    unsigned long long* vac, float esc, const StepGeometry* geom,
                    hipStream_t st)
{
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(phi && pi && phi_out && pi_out && tab.ab, "rk2 f32: NULL field or table");
    TIP_REQUIRE(phi != phi_out && pi != pi_out, "rk2 f32: output must not alias input");
    const int nm = noise ? NOISE_ARRAY : (hash ? NOISE_HASH : NOISE_NONE);
    const int kind = geom ? geom->kind : TIP_KERNEL_AUTO;
    TIP_REQUIRE(kind == TIP_KERNEL_AUTO || kind == TIP_KERNEL_ROWS || kind == TIP_KERNEL_FLAT,
                "rk2 f32: unknown kernel kind %d", kind);
    if (kind != TIP_KERNEL_FLAT) {
        const int ty_req = (geom && geom->by) ? geom->by : real_rows_per_thread(nx, ny, nz);
        const RowShape rs = row_shape(ny, nz, ty_req);
        TIP_REQUIRE(rs.ok || kind == TIP_KERNEL_AUTO,
                    "rk2 f32: row kernel needs nz <= 1024 and rows per thread (1, 2, 4) dividing "
                    "ny (got %d rows, ny = %lld)", ty_req, (long long)ny);
        if (rs.ok) {
            TileGrid tg;
            tg.tiles = (int)((ny + (int64_t)rs.G * rs.TY - 1) / ((int64_t)rs.G * rs.TY));
            // about 4096 blocks on large lattices (1024 on small ones), each marching enough
            // planes to amortise its window fill and halo rows
            int64_t cx = (geom && geom->cx) ? geom->cx
                                            : (nx * tg.tiles) / real_row_blocks(nx, ny, nz);
            if (!(geom && geom->cx)) cx = cx < 1 ? 1 : (cx > 32 ? 32 : cx);
            TIP_REQUIRE(cx >= 1 && cx <= 4096, "geometry: planes per block out of range");
            tg.cx = (int)cx;
            tg.chunks = (int)((nx + cx - 1) / cx);
            const int64_t nb64 = (int64_t)tg.tiles * tg.chunks;
            TIP_REQUIRE(nb64 < (int64_t(1) << 31), "geometry: too many blocks");
            const int nbr = (int)nb64;
            tg.xcd_remap = ((geom ? geom->xcd : 1) && nbr % 8 == 0) ? 1 : 0;
            const size_t lds = (size_t)rs.G * (64 * rs.W) * (4 * rs.TY + 1) * sizeof(float);
            const dim3 blk(64, rs.W * rs.G);
#define TIP_LAUNCH_ROWS_C(TYV, ODD, NM, CNT)                                                    \
    hipLaunchKernelGGL((k_rk2_real_rows<ODD, TYV, NM, CNT>), dim3(nbr), blk, lds, st, phi, pi, \
                       phi_out, pi_out, noise, c, tab.as<ODD>(), (int)nx, (int)ny, (int)nz,    \
                       rs.W, tg, noise_seed, vac, esc)
#define TIP_LAUNCH_ROWS(TYV, ODD, NM)                 \
    do {                                              \
        if (vac)                                      \
            TIP_LAUNCH_ROWS_C(TYV, ODD, NM, true);    \
        else                                          \
            TIP_LAUNCH_ROWS_C(TYV, ODD, NM, false);   \
    } while (0)
#define TIP_LAUNCH_ROWS_NM(TYV, ODD)                                      \
    switch (nm) {                                                         \
        case NOISE_ARRAY: TIP_LAUNCH_ROWS(TYV, ODD, NOISE_ARRAY); break;  \
        case NOISE_HASH: TIP_LAUNCH_ROWS(TYV, ODD, NOISE_HASH); break;    \
        default: TIP_LAUNCH_ROWS(TYV, ODD, NOISE_NONE); break;            \
    }
#define TIP_LAUNCH_ROWS_TY(ODD)                      \
    switch (rs.TY) {                                 \
        case 1: TIP_LAUNCH_ROWS_NM(1, ODD); break;   \
        case 4: TIP_LAUNCH_ROWS_NM(4, ODD); break;   \
        default: TIP_LAUNCH_ROWS_NM(2, ODD); break;  \
    }
            if (tab.odd) {
                TIP_LAUNCH_ROWS_TY(true)
            } else {
                TIP_LAUNCH_ROWS_TY(false)
            }
#undef TIP_LAUNCH_ROWS_TY
#undef TIP_LAUNCH_ROWS_NM
#undef TIP_LAUNCH_ROWS
#undef TIP_LAUNCH_ROWS_C
            TIP_HIP(hipGetLastError());
            return TIP_OK;
        }
    }
    StepGeometry g;
    TileGrid tg;
    int nb;
    rc = resolve_geometry(geom, nx, ny, nz, 0, 0, g, tg, nb);
    if (rc) return rc;
#define TIP_LAUNCH_REAL(BYV, ODD, NM)                                                         \
    hipLaunchKernelGGL((k_rk2_real_f32<ODD, BYV, NM>), dim3(nb), dim3(64, BYV), 0, st, phi,  \
                       pi, phi_out, pi_out, noise, c, tab.as<ODD>(), (int)nx, (int)ny,        \
                       (int)nz, tg, noise_seed, vac, esc)
#define TIP_LAUNCH_REAL_NM(BYV, ODD)                                  \
    switch (nm) {                                                     \
        case NOISE_ARRAY: TIP_LAUNCH_REAL(BYV, ODD, NOISE_ARRAY); break; \
        case NOISE_HASH: TIP_LAUNCH_REAL(BYV, ODD, NOISE_HASH); break;   \
        default: TIP_LAUNCH_REAL(BYV, ODD, NOISE_NONE); break;           \
    }
#define TIP_LAUNCH_REAL_BY(ODD)                          \
    switch (g.by) {                                      \
        case 1: TIP_LAUNCH_REAL_NM(1, ODD); break;       \
        case 2: TIP_LAUNCH_REAL_NM(2, ODD); break;       \
        case 8: TIP_LAUNCH_REAL_NM(8, ODD); break;       \
        default: TIP_LAUNCH_REAL_NM(4, ODD); break;      \
    }
    if (tab.odd) {
        TIP_LAUNCH_REAL_BY(true)
    } else {
        TIP_LAUNCH_REAL_BY(false)
    }
#undef TIP_LAUNCH_REAL_BY
#undef TIP_LAUNCH_REAL_NM
#undef TIP_LAUNCH_REAL
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

int launch_complex_f32(const float* phi1, const float* phi2, const float* pi1, const float* pi2,
                       float* o1, float* o2, float* o3, float* o4, const StepScalars<float>& c,
                       const float* noise1, const float* noise2, const SegTable& tab,
                       const ZnParams<float>& zn, int64_t nx, int64_t ny, int64_t nz,
                       uint64_t noise_seed, int hash, unsigned long long* vac, float esc2,
                       const StepGeometry* geom, hipStream_t st)
{
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(phi1 && phi2 && pi1 && pi2 && o1 && o2 && o3 && o4 && tab.ab,
                "rk2 c32: NULL field or table");
    TIP_REQUIRE(noise1 != nullptr || noise2 == nullptr, "rk2 c32: noise2 without noise1");
    const int nm = noise1 ? NOISE_ARRAY : (hash ? NOISE_HASH : NOISE_NONE);
    const int kind = geom ? geom->kind : TIP_KERNEL_AUTO;
    TIP_REQUIRE(kind == TIP_KERNEL_AUTO || kind == TIP_KERNEL_ROWS || kind == TIP_KERNEL_FLAT,
                "rk2 c32: unknown kernel kind %d", kind);
    if (kind != TIP_KERNEL_FLAT) {
        // Rows of two or four waves (nz in (64, 256]) run two row groups per block that share
        // their inner y-halo rows through LDS (HS) when ny divides into the pairs of tiles;
        // then four y rows per thread by default on large lattices (one 8-wave block per CU:
        // 134 -> 128 us at 256^3 against 140 -> 134 us for 2 rows and no sharing, measured
        // early / after the roll); otherwise at most two rows per thread (register budget).
        const int w_row = (int)((nz + 63) / 64);
        const int ty_def = default_rows_per_thread(nx, ny, nz);
        int ty_req = (geom && geom->by) ? geom->by : ty_def;
        const bool hs_shape = w_row == 2 || w_row == 4;
        if (!(geom && geom->by) && ty_req > 2 && !(hs_shape && ny % (2 * ty_req) == 0))
            ty_req = 2;
        RowShape rs = row_shape(ny, nz, ty_req, 8);  // blocks of at most 512 threads
        const bool hs = rs.ok && hs_shape && ny % (2 * rs.TY) == 0;
        if (hs) rs.G = 2;
        TIP_REQUIRE(rs.ok || kind == TIP_KERNEL_AUTO,
                    "rk2 c32: row kernel needs nz <= 512 and rows per thread (1, 2, 4) dividing "
                    "ny (got %d rows, ny = %lld)", ty_req, (long long)ny);
        if (rs.ok) {
            TileGrid tg;
            tg.tiles = (int)((ny + (int64_t)rs.G * rs.TY - 1) / ((int64_t)rs.G * rs.TY));
            // about 2048 blocks (16 planes per block at 256^3 with TY = 2), or one block per
            // CU for the shared-halo form with four rows per thread (32 planes at 256^3)
            const int64_t target = (hs && rs.TY == 4) ? 256 : 2048;
            int64_t cx = (geom && geom->cx) ? geom->cx : (nx * tg.tiles) / target;
            if (!(geom && geom->cx)) cx = cx < 1 ? 1 : (cx > 32 ? 32 : cx);
            TIP_REQUIRE(cx >= 1 && cx <= 4096, "geometry: planes per block out of range");
            tg.cx = (int)cx;
            tg.chunks = (int)((nx + cx - 1) / cx);
            const int64_t nb64 = (int64_t)tg.tiles * tg.chunks;
            TIP_REQUIRE(nb64 < (int64_t(1) << 31), "geometry: too many blocks");
            const int nbr = (int)nb64;
            tg.xcd_remap = ((geom ? geom->xcd : 1) && nbr % 8 == 0) ? 1 : 0;
            const size_t lds = (size_t)rs.G * (64 * rs.W) * (8 * rs.TY + 1) * sizeof(float);
            const dim3 blk(64, rs.W * rs.G);
            TableRadial t32;
            t32.t = tab.as<false>();
            // the shared-halo form with four rows per thread (one block per CU) chooses per
            // block between the vector-cache table and an LDS copy, when the table fits
            TableRadialAuto t32a;
            t32a.t = t32.t;
            t32a.lt = nullptr;
            t32a.use_lds = 0;
            const bool auto_tab = hs && rs.TY == 4 && tab.last + 1 <= TableRadialAuto::kMaxSeg &&
                                  ((uintptr_t)tab.ab & 15) == 0;
            const size_t lds_a = lds + (size_t)TableRadialAuto::kLdsFloats * sizeof(float);
#define TIP_LAUNCH_CROWS_HZ(TYV, NM, CNT, HSV, ZNV)                                             \
    hipLaunchKernelGGL((k_rk2_complex_rows<TYV, NM, TableRadial, CNT, HSV, ZNV>), dim3(nbr), blk, \
                       lds, st, phi1, phi2, pi1, pi2, o1, o2, o3, o4, noise1, noise2, c, t32, zn, \
                       (int)nx, (int)ny, (int)nz, rs.W, tg, noise_seed, vac, esc2)
#define TIP_LAUNCH_CROWS_H(TYV, NM, CNT, HSV)            \
    do {                                                 \
        if (zn.active)                                   \
            TIP_LAUNCH_CROWS_HZ(TYV, NM, CNT, HSV, true);  \
        else                                             \
            TIP_LAUNCH_CROWS_HZ(TYV, NM, CNT, HSV, false); \
    } while (0)
#define TIP_LAUNCH_CROWS_AZ(NM, CNT, ZNV)                                                      \
    hipLaunchKernelGGL((k_rk2_complex_rows<4, NM, TableRadialAuto, CNT, true, ZNV>), dim3(nbr),   \
                       blk, lds_a, st, phi1, phi2, pi1, pi2, o1, o2, o3, o4, noise1, noise2, c,   \
                       t32a, zn, (int)nx, (int)ny, (int)nz, rs.W, tg, noise_seed, vac, esc2)
// the Z_N term compiled out of launches that switch it off (a third of the loop code)
#define TIP_LAUNCH_CROWS_A(NM, CNT)                 \
    do {                                            \
        if (zn.active)                              \
            TIP_LAUNCH_CROWS_AZ(NM, CNT, true);     \
        else                                        \
            TIP_LAUNCH_CROWS_AZ(NM, CNT, false);    \
    } while (0)
#define TIP_LAUNCH_CROWS_C(TYV, NM, CNT)            \
    do {                                            \
        if (TYV == 4 && auto_tab)                   \
            TIP_LAUNCH_CROWS_A(NM, CNT);            \
        else if (hs)                                \
            TIP_LAUNCH_CROWS_H(TYV, NM, CNT, true); \
        else                                        \
            TIP_LAUNCH_CROWS_H(TYV, NM, CNT, false);\
    } while (0)
#define TIP_LAUNCH_CROWS(TYV, NM)              \
    do {                                       \
        if (vac)                               \
            TIP_LAUNCH_CROWS_C(TYV, NM, true); \
        else                                   \
            TIP_LAUNCH_CROWS_C(TYV, NM, false);\
    } while (0)
#define TIP_LAUNCH_CROWS_NM(TYV)                                      \
    switch (nm) {                                                     \
        case NOISE_ARRAY: TIP_LAUNCH_CROWS(TYV, NOISE_ARRAY); break;  \
        case NOISE_HASH: TIP_LAUNCH_CROWS(TYV, NOISE_HASH); break;    \
        default: TIP_LAUNCH_CROWS(TYV, NOISE_NONE); break;            \
    }
            switch (rs.TY) {
                case 1: TIP_LAUNCH_CROWS_NM(1); break;
                case 4: TIP_LAUNCH_CROWS_NM(4); break;
                default: TIP_LAUNCH_CROWS_NM(2); break;
            }
#undef TIP_LAUNCH_CROWS_NM
#undef TIP_LAUNCH_CROWS
#undef TIP_LAUNCH_CROWS_C
#undef TIP_LAUNCH_CROWS_H
#undef TIP_LAUNCH_CROWS_HZ
#undef TIP_LAUNCH_CROWS_A
#undef TIP_LAUNCH_CROWS_AZ
            TIP_HIP(hipGetLastError());
            return TIP_OK;
        }
    }
    StepGeometry g;
    TileGrid tg;
    int nb;
    rc = resolve_geometry(geom, nx, ny, nz, 1, 0, g, tg, nb);
    if (rc) return rc;
#define TIP_LAUNCH_CPLX(BYV)                                                                    \
    hipLaunchKernelGGL((k_rk2_complex<float, TableF32<false>, BYV>), dim3(nb), dim3(64, BYV), 0, \
                       st, phi1, phi2, pi1, pi2, o1, o2, o3, o4, noise1, noise2, c,              \
                       tab.as<false>(), zn, (int)nx,                                              \
                       (int)ny, (int)nz, tg, noise_seed, hash, vac, esc2)
    switch (g.by) {
        case 1: TIP_LAUNCH_CPLX(1); break;
        case 2: TIP_LAUNCH_CPLX(2); break;
        case 8: TIP_LAUNCH_CPLX(8); break;
        default: TIP_LAUNCH_CPLX(4); break;
    }
#undef TIP_LAUNCH_CPLX
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

}  // namespace tip

namespace {

int check_npts(int64_t vp_npts)
{
    TIP_REQUIRE(vp_npts >= 2 && vp_npts < (int64_t(1) << 30), "rk2 f32: vp_npts out of range");
    return TIP_OK;
}

}  // namespace

extern "C" int tip_table_prepare_f32(const double* vp_table, int64_t npts, double vp_tmin,
                                     double vp_dinv, void* table_ab, void* stream)
{
    return prepare_table_f32(vp_table, npts, vp_tmin, vp_dinv,
                             reinterpret_cast<float2*>(table_ab), as_stream(stream));
}

extern "C" int tip_rk2_fused_f32(const float* phi, const float* pi, float* phi_out, float* pi_out,
                                 double dt, double dx, double eta_eff, double mu, double inv_a2,
                                 const float* noise, const void* table_ab, double vp_tmin,
                                 double vp_dinv, int64_t vp_npts, int vp_odd, int64_t nx,
                                 int64_t ny, int64_t nz, uint64_t noise_seed, double noise_scale,
                                 unsigned long long* vac_count, double vac_esc, void* stream)
{
    int rc = check_npts(vp_npts);
    if (rc) return rc;
    return launch_real_f32(phi, pi, phi_out, pi_out,
                           make_scalars<float>(dt, dx, eta_eff, mu, inv_a2, noise_scale), noise,
                           make_table32(table_ab, vp_tmin, vp_dinv, vp_npts, vp_odd), nx, ny, nz,
                           noise_seed, (noise == nullptr && noise_scale != 0.0) ? 1 : 0,
                           vac_count, (float)vac_esc, nullptr, as_stream(stream));
}

extern "C" int tip_rk2_step_table_f32(float* phi, float* pi, double dt, double dx,
                                      double eta_eff, double mu, const float* noise,
                                      float* phi_tmp, float* pi_tmp, double inv_a2,
                                      const double* vp_table, double vp_tmin, double vp_dinv,
                                      int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny,
                                      int64_t nz, uint64_t noise_seed, double noise_scale,
                                      unsigned long long* vac_count, double vac_esc,
                                      void* stream)
{
    hipStream_t st = as_stream(stream);
    TIP_REQUIRE(phi_tmp && pi_tmp, "rk2_step_table_f32: phi_tmp/pi_tmp are required");
    TIP_REQUIRE(vp_table, "rk2_step_table_f32: NULL table");
    int rc = check_npts(vp_npts);
    if (rc) return rc;
    ScratchLease lease{0, st};
    float2* ab =
        reinterpret_cast<float2*>(scratch_get((size_t)(vp_npts - 1) * sizeof(float2), 0, st));
    TIP_REQUIRE(ab != nullptr, "rk2_step_table_f32: scratch allocation failed");
    rc = prepare_table_f32(vp_table, vp_npts, vp_tmin, vp_dinv, ab, st);
    if (rc) return rc;
    rc = launch_real_f32(phi, pi, phi_tmp, pi_tmp,
                         make_scalars<float>(dt, dx, eta_eff, mu, inv_a2, noise_scale), noise,
                         make_table32(ab, vp_tmin, vp_dinv, vp_npts, vp_odd), nx, ny, nz,
                         noise_seed, (noise == nullptr && noise_scale != 0.0) ? 1 : 0, vac_count,
                         (float)vac_esc, nullptr, st);
    if (rc) return rc;
    size_t bytes = (size_t)(nx * ny * nz) * sizeof(float);
    TIP_HIP(hipMemcpyAsync(phi, phi_tmp, bytes, hipMemcpyDeviceToDevice, st));
    TIP_HIP(hipMemcpyAsync(pi, pi_tmp, bytes, hipMemcpyDeviceToDevice, st));
    return TIP_OK;
}

extern "C" int tip_rk2_fused_c32(const float* phi1, const float* phi2, const float* pi1,
                                 const float* pi2, float* phi1_out, float* phi2_out,
                                 float* pi1_out, float* pi2_out, double dt, double dx,
                                 double eta_eff, double mu, double inv_a2, const float* noise1,
                                 const float* noise2, const void* table_ab, double vp_tmin,
                                 double vp_dinv, int64_t vp_npts, double zn_n, double zn_dv,
                                 double zn_active, int64_t nx, int64_t ny, int64_t nz,
                                 uint64_t noise_seed, double noise_scale,
                                 unsigned long long* vac_count, double vac_esc, void* stream)
{
    TIP_REQUIRE(phi1 != phi1_out && phi2 != phi2_out && pi1 != pi1_out && pi2 != pi2_out,
                "rk2 c32: output must not alias input");
    int rc = check_npts(vp_npts);
    if (rc) return rc;
    return launch_complex_f32(phi1, phi2, pi1, pi2, phi1_out, phi2_out, pi1_out, pi2_out,
                              make_scalars<float>(dt, dx, eta_eff, mu, inv_a2, noise_scale),
                              noise1, noise2, make_table32(table_ab, vp_tmin, vp_dinv, vp_npts, 0),
                              make_zn32(zn_n, zn_dv, zn_active), nx, ny, nz, noise_seed,
                              (noise1 == nullptr && noise_scale != 0.0) ? 1 : 0, vac_count,
                              (float)(vac_esc * vac_esc), nullptr, as_stream(stream));
}

extern "C" int tip_rk2_step_table_c32(
    float* phi1, float* phi2, float* pi1, float* pi2, double dt, double dx, double eta_eff,
    double mu, const float* noise1, const float* noise2, float* phi1_tmp, float* phi2_tmp,
    float* pi1_tmp, float* pi2_tmp, double inv_a2, const double* vp_table, double vp_tmin,
    double vp_dinv, int64_t vp_npts, double zn_n, double zn_dv, double zn_active, int64_t nx,
    int64_t ny, int64_t nz, uint64_t noise_seed, double noise_scale,
    unsigned long long* vac_count, double vac_esc, void* stream)
{
    hipStream_t st = as_stream(stream);
    TIP_REQUIRE(phi1_tmp && phi2_tmp && pi1_tmp && pi2_tmp, "rk2 c32: tmp arrays are required");
    TIP_REQUIRE(vp_table, "rk2 c32: NULL table");
    int rc = check_npts(vp_npts);
    if (rc) return rc;
    ScratchLease lease{0, st};
    float2* ab =
        reinterpret_cast<float2*>(scratch_get((size_t)(vp_npts - 1) * sizeof(float2), 0, st));
    TIP_REQUIRE(ab != nullptr, "rk2 c32: scratch allocation failed");
    rc = prepare_table_f32(vp_table, vp_npts, vp_tmin, vp_dinv, ab, st);
    if (rc) return rc;
    rc = launch_complex_f32(phi1, phi2, pi1, pi2, phi1_tmp, phi2_tmp, pi1_tmp, pi2_tmp,
                            make_scalars<float>(dt, dx, eta_eff, mu, inv_a2, noise_scale), noise1,
                            noise2, make_table32(ab, vp_tmin, vp_dinv, vp_npts, 0),
                            make_zn32(zn_n, zn_dv, zn_active), nx, ny, nz, noise_seed,
                            (noise1 == nullptr && noise_scale != 0.0) ? 1 : 0, vac_count,
                            (float)(vac_esc * vac_esc), nullptr, st);
    if (rc) return rc;
    size_t bytes = (size_t)(nx * ny * nz) * sizeof(float);
    TIP_HIP(hipMemcpyAsync(phi1, phi1_tmp, bytes, hipMemcpyDeviceToDevice, st));
    TIP_HIP(hipMemcpyAsync(phi2, phi2_tmp, bytes, hipMemcpyDeviceToDevice, st));
    TIP_HIP(hipMemcpyAsync(pi1, pi1_tmp, bytes, hipMemcpyDeviceToDevice, st));
    TIP_HIP(hipMemcpyAsync(pi2, pi2_tmp, bytes, hipMemcpyDeviceToDevice, st));
    return TIP_OK;
}

// ------------------------------------------------------------------------------------------
// Allocation-free form of tip_rk2_step_table_f32.  The float32 segment table lives in a
// file-scope device array instead of library scratch grown with hipMalloc, so the call
// allocates nothing.  Each device holds its own copy of the array at its own address; the
// address is resolved once per device with hipGetSymbolAddress and only that address is
// passed to kernels and copies.  Every call on one device shares the array: calls that
// overlap on one device (several host threads or streams) race.
// ------------------------------------------------------------------------------------------
namespace {

constexpr int64_t kNoAllocMaxPoints = 65537;  // table points: 65536 segments, 512 KiB
static __device__ __align__(256) float2 scratch[kNoAllocMaxPoints - 1];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
float2* g_noalloc_ptr[kNoAllocMaxDevices] = {};

int noalloc_scratch(float2** out)
{
    int dev = 0;
    TIP_HIP(hipGetDevice(&dev));
    TIP_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "no-alloc step: device %d out of range",
                dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        TIP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        TIP_REQUIRE(p != nullptr, "no-alloc step: scratch symbol has no device address");
        g_noalloc_ptr[dev] = reinterpret_cast<float2*>(p);
    }
    *out = g_noalloc_ptr[dev];
    return TIP_OK;
}

}  // namespace

extern "C" int tip_rk2_step_table_f32_noalloc(float* phi, float* pi, double dt, double dx,
                                              double eta_eff, double mu, const float* noise,
                                              float* phi_tmp, float* pi_tmp, double inv_a2,
                                              const double* vp_table, double vp_tmin,
                                              double vp_dinv, int64_t vp_npts, int vp_odd,
                                              int64_t nx, int64_t ny, int64_t nz,
                                              uint64_t noise_seed, double noise_scale,
                                              unsigned long long* vac_count, double vac_esc,
                                              void* stream)
{
    hipStream_t st = as_stream(stream);
    TIP_REQUIRE(phi_tmp && pi_tmp, "rk2_step_table_f32_noalloc: phi_tmp/pi_tmp are required");
    TIP_REQUIRE(vp_table, "rk2_step_table_f32_noalloc: NULL table");
    int rc = check_npts(vp_npts);
    if (rc) return rc;
    TIP_REQUIRE(vp_npts <= kNoAllocMaxPoints,
                "rk2_step_table_f32_noalloc: vp_npts = %lld exceeds the %lld points of its "
                "static table scratch", (long long)vp_npts, (long long)kNoAllocMaxPoints);
    float2* ab = nullptr;
    rc = noalloc_scratch(&ab);
    if (rc) return rc;
    rc = prepare_table_f32(vp_table, vp_npts, vp_tmin, vp_dinv, ab, st);
    if (rc) return rc;
    rc = launch_real_f32(phi, pi, phi_tmp, pi_tmp,
                         make_scalars<float>(dt, dx, eta_eff, mu, inv_a2, noise_scale), noise,
                         make_table32(ab, vp_tmin, vp_dinv, vp_npts, vp_odd), nx, ny, nz,
                         noise_seed, (noise == nullptr && noise_scale != 0.0) ? 1 : 0, vac_count,
                         (float)vac_esc, nullptr, st);
    if (rc) return rc;
    size_t bytes = (size_t)(nx * ny * nz) * sizeof(float);
    TIP_HIP(hipMemcpyAsync(phi, phi_tmp, bytes, hipMemcpyDeviceToDevice, st));
    TIP_HIP(hipMemcpyAsync(pi, pi_tmp, bytes, hipMemcpyDeviceToDevice, st));
    return TIP_OK;
}
