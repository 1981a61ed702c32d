// tip_lattice.hip -- device-resident lattice: ping-pong field buffers, the V' table in both
// precisions, a stream and timing events.  One RK2 step = one fused kernel launch; the
// buffers swap roles after every step, so no copy ever runs.
#include "langevin_kernels.h"

struct tip_lattice {
    int kind = TIP_REAL;
    int prec = 4;
    int64_t nx = 0, ny = 0, nz = 0;
    int nfields = 2;
    void* buf[2][4] = {};  // [parity][field]
    int cur = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double* table = nullptr;
    float2* table_ab = nullptr;
    int64_t npts = 0, cap = 0;
    double tmin = 0.0, dinv = 1.0;
    int odd = 0;
    double zn_n = 0.0, zn_dv = 0.0, zn_active = 0.0;
    unsigned long long* counter = nullptr;  // 2 entries
    double* mm = nullptr;                   // 2 entries
    tip::StepGeometry geom = {0, 0, 1, TIP_KERNEL_AUTO};  // defaults for the lattice size
    float2* mlp_grid = nullptr;  // per-step grid of the surrogate ((npts - 1) float2)
    int64_t mlp_grid_cap = 0;
    // BAOAB: the closing half-kick of the last step is deferred (pi holds the post-O
    // momentum) until the next BAOAB call fuses it or an operation needs the true pi.
    int baoab_pending = 0;
    tip_step_params baoab_prev = {};
};

using namespace tip;

namespace {

int close_baoab(tip_lattice* L);

size_t field_bytes(const tip_lattice* L) { return (size_t)(L->nx * L->ny * L->nz) * L->prec; }

int release(tip_lattice* L)
{
    if (!L) return TIP_OK;
    if (L->stream) (void)hipStreamSynchronize(L->stream);
    for (auto& p : L->buf)
        for (void*& q : p)
            if (q) { (void)hipFree(q); q = nullptr; }
    if (L->table) (void)hipFree(L->table);
    if (L->table_ab) (void)hipFree(L->table_ab);
    if (L->counter) (void)hipFree(L->counter);
    if (L->mm) (void)hipFree(L->mm);
    if (L->mlp_grid) (void)hipFree(L->mlp_grid);
    if (L->ev0) (void)hipEventDestroy(L->ev0);
    if (L->ev1) (void)hipEventDestroy(L->ev1);
    if (L->stream) (void)hipStreamDestroy(L->stream);
    delete L;
    return TIP_OK;
}

// Map the public field index to the internal slot: real {phi, pi}; complex
// {phi1, pi1, phi2, pi2} -> slots {0, 2, 1, 3} so that the kernel order is phi1, phi2, pi1, pi2.
int slot_of(const tip_lattice* L, int which)
{
    if (which < 0 || which >= L->nfields) return -1;
    if (L->kind == TIP_REAL) return which;
    static const int m[4] = {0, 2, 1, 3};
    return m[which];
}

}  // namespace

extern "C" int tip_lattice_create(tip_lattice** out, int field_kind, int precision, int64_t nx,
                                  int64_t ny, int64_t nz)
{
    TIP_REQUIRE(out, "tip_lattice_create: NULL out");
    *out = nullptr;
    TIP_REQUIRE(field_kind == TIP_REAL || field_kind == TIP_COMPLEX, "unknown field_kind %d",
                field_kind);
    TIP_REQUIRE(precision == 4 || precision == 8, "precision must be 4 or 8");
    TIP_REQUIRE(!(field_kind == TIP_COMPLEX && precision == 8),
                "complex lattices are float32 (use tip_rk2_step_table_c64 for float64)");
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    tip_lattice* L = new tip_lattice();
    L->kind = field_kind;
    L->prec = precision;
    L->nx = nx;
    L->ny = ny;
    L->nz = nz;
    L->nfields = field_kind == TIP_REAL ? 2 : 4;
    hipError_t e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&L->ev0);
    if (e == hipSuccess) e = hipEventCreate(&L->ev1);
    for (int p = 0; p < 2 && e == hipSuccess; ++p)
        for (int f = 0; f < L->nfields && e == hipSuccess; ++f) {
            e = hipMalloc(&L->buf[p][f], field_bytes(L));
            if (e == hipSuccess) e = hipMemsetAsync(L->buf[p][f], 0, field_bytes(L), L->stream);
        }
    if (e == hipSuccess) e = hipMalloc(&L->counter, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&L->mm, 2 * sizeof(double));
    if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
    if (e != hipSuccess) {
        int r = fail_hip(e, "tip_lattice_create");
        release(L);
        return r;
    }
    *out = L;
    return TIP_OK;
}

extern "C" int tip_lattice_destroy(tip_lattice* lat) { return release(lat); }

extern "C" int tip_lattice_stream(tip_lattice* lat, void** stream)
{
    TIP_REQUIRE(lat && stream, "tip_lattice_stream: NULL");
    *stream = lat->stream;
    return TIP_OK;
}

extern "C" int tip_lattice_field_ptr(tip_lattice* lat, int which, void** ptr)
{
    TIP_REQUIRE(lat && ptr, "tip_lattice_field_ptr: NULL");
    int s = slot_of(lat, which);
    TIP_REQUIRE(s >= 0, "tip_lattice_field_ptr: bad field index %d", which);
    int rc = close_baoab(lat);  // the caller may read or write any field through the pointer
    if (rc) return rc;
    *ptr = lat->buf[lat->cur][s];
    return TIP_OK;
}

extern "C" int tip_lattice_upload(tip_lattice* lat, int which, const void* host)
{
    TIP_REQUIRE(lat && host, "tip_lattice_upload: NULL");
    int s = slot_of(lat, which);
    TIP_REQUIRE(s >= 0, "tip_lattice_upload: bad field index %d", which);
    int rc = close_baoab(lat);
    if (rc) return rc;
    TIP_HIP(hipMemcpyAsync(lat->buf[lat->cur][s], host, field_bytes(lat), hipMemcpyHostToDevice,
                           lat->stream));
    TIP_HIP(hipStreamSynchronize(lat->stream));
    return TIP_OK;
}

extern "C" int tip_lattice_download(tip_lattice* lat, int which, void* host)
{
    TIP_REQUIRE(lat && host, "tip_lattice_download: NULL");
    int s = slot_of(lat, which);
    TIP_REQUIRE(s >= 0, "tip_lattice_download: bad field index %d", which);
    const bool momentum = lat->kind == TIP_REAL ? which == 1 : (which == 1 || which == 3);
    if (momentum) {
        int rc = close_baoab(lat);
        if (rc) return rc;
    }
    TIP_HIP(hipMemcpyAsync(host, lat->buf[lat->cur][s], field_bytes(lat), hipMemcpyDeviceToHost,
                           lat->stream));
    TIP_HIP(hipStreamSynchronize(lat->stream));
    return TIP_OK;
}

extern "C" int tip_lattice_set_table(tip_lattice* lat, const double* host_table, int64_t npts,
                                     double tmin, double dinv, int odd)
{
    TIP_REQUIRE(lat && host_table, "tip_lattice_set_table: NULL");
    TIP_REQUIRE(npts >= 2 && npts < (int64_t(1) << 30), "tip_lattice_set_table: npts out of range");
    TIP_REQUIRE(dinv > 0.0, "tip_lattice_set_table: dinv must be positive");
    TIP_REQUIRE(!(lat->kind == TIP_COMPLEX && odd), "complex tables are tabulated on rho >= 0");
    int rc0 = close_baoab(lat);  // the deferred half-kick belongs to the old table
    if (rc0) return rc0;
    if (npts > lat->cap) {
        TIP_HIP(hipStreamSynchronize(lat->stream));
        if (lat->table) TIP_HIP(hipFree(lat->table));
        if (lat->table_ab) TIP_HIP(hipFree(lat->table_ab));
        lat->table = nullptr;
        lat->table_ab = nullptr;
        TIP_HIP(hipMalloc(&lat->table, npts * sizeof(double)));
        TIP_HIP(hipMalloc(&lat->table_ab, (npts - 1) * sizeof(float2)));
        lat->cap = npts;
    }
    TIP_HIP(hipMemcpyAsync(lat->table, host_table, npts * sizeof(double), hipMemcpyHostToDevice,
                           lat->stream));
    int rc = prepare_table_f32(lat->table, npts, tmin, dinv, lat->table_ab, lat->stream);
    if (rc) return rc;
    TIP_HIP(hipStreamSynchronize(lat->stream));
    lat->npts = npts;
    lat->tmin = tmin;
    lat->dinv = dinv;
    lat->odd = odd ? 1 : 0;
    return TIP_OK;
}

extern "C" int tip_lattice_build_table(tip_lattice* lat, int64_t npts, double T, double phi_min,
                                       double phi_max, const tip_vprime_model* model,
                                       const double* cb, const double* cf, int odd)
{
    TIP_REQUIRE(lat && model && cb && cf, "tip_lattice_build_table: NULL");
    TIP_REQUIRE(npts >= 2 && npts < (int64_t(1) << 30),
                "tip_lattice_build_table: npts out of range");
    TIP_REQUIRE(!(lat->kind == TIP_COMPLEX && odd), "complex tables are tabulated on rho >= 0");
    double tmin, dinv;
    table_grid(npts, phi_min, phi_max, &tmin, &dinv);
    TIP_REQUIRE(dinv > 0.0 && dinv < 1e300, "tip_lattice_build_table: empty table range");
    int rc = close_baoab(lat);  // the deferred half-kick belongs to the old table
    if (rc) return rc;
    if (npts > lat->cap) {
        TIP_HIP(hipStreamSynchronize(lat->stream));
        if (lat->table) TIP_HIP(hipFree(lat->table));
        if (lat->table_ab) TIP_HIP(hipFree(lat->table_ab));
        lat->table = nullptr;
        lat->table_ab = nullptr;
        TIP_HIP(hipMalloc(&lat->table, npts * sizeof(double)));
        TIP_HIP(hipMalloc(&lat->table_ab, (npts - 1) * sizeof(float2)));
        lat->cap = npts;
    }
    // stream-ordered after every step already issued on the lattice: no host sync needed
    rc = build_vprime_table_dev(lat->table, npts, T, phi_min, phi_max, *model, cb, cf,
                                lat->stream);
    if (rc) return rc;
    rc = prepare_table_f32(lat->table, npts, tmin, dinv, lat->table_ab, lat->stream);
    if (rc) return rc;
    lat->npts = npts;
    lat->tmin = tmin;
    lat->dinv = dinv;
    lat->odd = odd ? 1 : 0;
    return TIP_OK;
}

extern "C" int tip_lattice_set_zn(tip_lattice* lat, double zn_n, double zn_dv, double zn_active)
{
    TIP_REQUIRE(lat, "tip_lattice_set_zn: NULL");
    int rc = close_baoab(lat);
    if (rc) return rc;
    lat->zn_n = zn_n;
    lat->zn_dv = zn_dv;
    lat->zn_active = zn_active;
    return TIP_OK;
}

extern "C" int tip_lattice_set_geometry(tip_lattice* lat, int kernel, int rows,
                                        int planes_per_block, int xcd_remap)
{
    TIP_REQUIRE(lat, "tip_lattice_set_geometry: NULL");
    TIP_REQUIRE(kernel == TIP_KERNEL_AUTO || kernel == TIP_KERNEL_ROWS ||
                    kernel == TIP_KERNEL_FLAT,
                "tip_lattice_set_geometry: unknown kernel %d", kernel);
    StepGeometry req = {rows, planes_per_block, xcd_remap, kernel};
    if (kernel == TIP_KERNEL_ROWS) {
        TIP_REQUIRE(lat->prec == 4,
                    "tip_lattice_set_geometry: the row kernel is the float32 step");
        const int ty_def = lat->kind == TIP_COMPLEX
                               ? default_rows_per_thread(lat->nx, lat->ny, lat->nz)
                               : real_rows_per_thread(lat->nx, lat->ny, lat->nz);
        const RowShape rs =
            row_shape(lat->ny, lat->nz, rows ? rows : ty_def, lat->kind == TIP_COMPLEX ? 8 : 16);
        TIP_REQUIRE(rs.ok, "tip_lattice_set_geometry: row kernel needs nz <= 1024 (complex: 512) "
                           "and rows per thread (1, 2, 4) dividing ny");
        TIP_REQUIRE(planes_per_block >= 0 && planes_per_block <= 4096,
                    "geometry: planes per block out of range");
    } else {
        StepGeometry g;
        TileGrid tg;
        int nb;
        StepGeometry flat = req;
        // under AUTO a rows value the flat kernel does not take is a row-kernel value
        if (kernel == TIP_KERNEL_AUTO && rows != 0 && rows != 1 && rows != 2 && rows != 4 &&
            rows != 8)
            TIP_REQUIRE(false, "geometry: rows must be 0, 1, 2, 4 or 8 (got %d)", rows);
        int rc = resolve_geometry(&flat, lat->nx, lat->ny, lat->nz, lat->kind == TIP_COMPLEX, 0,
                                  g, tg, nb);
        if (rc) return rc;
    }
    lat->geom = req;
    return TIP_OK;
}

extern "C" int tip_lattice_step(tip_lattice* L, const tip_step_params* p, int64_t nsteps,
                                uint64_t step0, uint64_t seed, float* elapsed_ms)
{
    TIP_REQUIRE(L && p, "tip_lattice_step: NULL");
    TIP_REQUIRE(L->npts >= 2, "tip_lattice_step: no V' table set");
    TIP_REQUIRE(nsteps >= 0, "tip_lattice_step: nsteps < 0");
    int rc0 = close_baoab(L);
    if (rc0) return rc0;
    const StepScalars<float> c32 =
        make_scalars<float>(p->dt, p->dx, p->eta_eff, p->mu, p->inv_a2, p->noise_scale);
    const SegTable tab32 = make_table32(L->table_ab, L->tmin, L->dinv, L->npts, L->odd);
    const ZnParams<float> zn32 = make_zn32(L->zn_n, L->zn_dv, L->zn_active);
    const int hash = p->noise_scale != 0.0 ? 1 : 0;
    if (elapsed_ms) TIP_HIP(hipEventRecord(L->ev0, L->stream));
    for (int64_t n = 0; n < nsteps; ++n) {
        const uint64_t ns = tip_noise_seed(step0 + (uint64_t)n, seed);
        void** in = L->buf[L->cur];
        void** out = L->buf[L->cur ^ 1];
        int rc;
        if (L->kind == TIP_REAL && L->prec == 4) {
            rc = launch_real_f32((const float*)in[0], (const float*)in[1], (float*)out[0],
                                 (float*)out[1], c32, nullptr, tab32, L->nx, L->ny, L->nz, ns,
                                 hash, nullptr, 0.0f, &L->geom, L->stream);
        } else if (L->kind == TIP_REAL) {
            rc = tip_rk2_fused_f64((const double*)in[0], (const double*)in[1], (double*)out[0],
                                   (double*)out[1], p->dt, p->dx, p->eta_eff, p->mu, p->inv_a2,
                                   nullptr, L->table, L->tmin, L->dinv, L->npts, L->odd, L->nx,
                                   L->ny, L->nz, ns, p->noise_scale, nullptr, 0.0, L->stream);
        } else {
            rc = launch_complex_f32((const float*)in[0], (const float*)in[1], (const float*)in[2],
                                    (const float*)in[3], (float*)out[0], (float*)out[1],
                                    (float*)out[2], (float*)out[3], c32, nullptr, nullptr, tab32,
                                    zn32, L->nx, L->ny, L->nz, ns, hash, nullptr, 0.0f, &L->geom,
                                    L->stream);
        }
        if (rc) return rc;
        L->cur ^= 1;
    }
    if (elapsed_ms) {
        TIP_HIP(hipEventRecord(L->ev1, L->stream));
        TIP_HIP(hipEventSynchronize(L->ev1));
        TIP_HIP(hipEventElapsedTime(elapsed_ms, L->ev0, L->ev1));
    }
    return TIP_OK;
}

namespace {

// One overdamped / BAOAB stage on the lattice's current state.  Stages that write phi write it
// to the other parity's phi slots; the caller swaps those (phi_swap) or the whole parity.
int lattice_langevin(tip_lattice* L, int op, const tip_step_params* p, double inv_a2_prev,
                     uint64_t noise_seed)
{
    void** in = L->buf[L->cur];
    void** out = L->buf[L->cur ^ 1];
    const int hash = p->noise_scale != 0.0 ? 1 : 0;
    // BAOAB writes pi to the other parity (open / fused) or in place (close); the overdamped
    // step leaves pi alone.
    const bool pi_in_place = op == LV_BAOAB_CLOSE;
    if (L->kind == TIP_REAL) {
        void* pi_out = pi_in_place ? in[1] : out[1];
        if (L->prec == 4) {
            return langevin_real_f32(
                op, (const float*)in[0], (const float*)in[1], (float*)out[0], (float*)pi_out,
                make_langevin<float>(p->dt, p->dx, p->eta_eff, p->mu, p->inv_a2, inv_a2_prev,
                                     p->noise_scale),
                make_table32(L->table_ab, L->tmin, L->dinv, L->npts, L->odd), L->nx, L->ny,
                L->nz, noise_seed, hash, L->stream, &L->geom);
        }
        return langevin_real_f64(
            op, (const double*)in[0], (const double*)in[1], (double*)out[0], (double*)pi_out,
            make_langevin<double>(p->dt, p->dx, p->eta_eff, p->mu, p->inv_a2, inv_a2_prev,
                                  p->noise_scale),
            L->table, L->tmin, L->dinv, L->npts, L->odd, L->nx, L->ny, L->nz, noise_seed, hash,
            L->stream);
    }
    void* q1 = pi_in_place ? in[2] : out[2];
    void* q2 = pi_in_place ? in[3] : out[3];
    return langevin_complex_f32(
        op, (const float*)in[0], (const float*)in[1], (const float*)in[2], (const float*)in[3],
        (float*)out[0], (float*)out[1], (float*)q1, (float*)q2,
        make_langevin<float>(p->dt, p->dx, p->eta_eff, p->mu, p->inv_a2, inv_a2_prev,
                             p->noise_scale),
        make_table32(L->table_ab, L->tmin, L->dinv, L->npts, 0),
        make_zn32(L->zn_n, L->zn_dv, L->zn_active), L->nx, L->ny, L->nz, noise_seed, hash,
        L->stream, &L->geom);
}

// Exchange the phi slots of the two parities (the overdamped step keeps pi where it is).
void phi_swap(tip_lattice* L)
{
    const int nphi = L->kind == TIP_REAL ? 1 : 2;
    for (int f = 0; f < nphi; ++f) {
        void* t = L->buf[L->cur][f];
        L->buf[L->cur][f] = L->buf[L->cur ^ 1][f];
        L->buf[L->cur ^ 1][f] = t;
    }
}

int langevin_checks(tip_lattice* L, const tip_step_params* p, int64_t nsteps, const char* who)
{
    TIP_REQUIRE(L && p, "%s: NULL", who);
    TIP_REQUIRE(L->npts >= 2, "%s: no V' table set", who);
    TIP_REQUIRE(nsteps >= 0, "%s: nsteps < 0", who);
    TIP_REQUIRE(p->eta_eff > 0.0, "%s: eta_eff (gamma) must be positive", who);
    return TIP_OK;
}

// Run the deferred closing half-kick of the last BAOAB step (with that step's parameters and
// the table and Z_N parameters still in place).
int close_baoab(tip_lattice* L)
{
    if (!L->baoab_pending) return TIP_OK;
    L->baoab_pending = 0;
    return lattice_langevin(L, LV_BAOAB_CLOSE, &L->baoab_prev, L->baoab_prev.inv_a2, 0);
}

}  // namespace

extern "C" int tip_lattice_finish(tip_lattice* L)
{
    TIP_REQUIRE(L, "tip_lattice_finish: NULL");
    int rc = close_baoab(L);
    if (rc) return rc;
    TIP_HIP(hipStreamSynchronize(L->stream));
    return TIP_OK;
}

extern "C" int tip_lattice_step_overdamped(tip_lattice* L, const tip_step_params* p,
                                           int64_t nsteps, uint64_t step0, uint64_t seed,
                                           float* elapsed_ms)
{
    int rc = langevin_checks(L, p, nsteps, "tip_lattice_step_overdamped");
    if (rc) return rc;
    rc = close_baoab(L);
    if (rc) return rc;
    if (elapsed_ms) TIP_HIP(hipEventRecord(L->ev0, L->stream));
    for (int64_t n = 0; n < nsteps; ++n) {
        rc = lattice_langevin(L, LV_OVERDAMPED, p, p->inv_a2,
                              tip_noise_seed(step0 + (uint64_t)n, seed));
        if (rc) return rc;
        phi_swap(L);
    }
    if (elapsed_ms) {
        TIP_HIP(hipEventRecord(L->ev1, L->stream));
        TIP_HIP(hipEventSynchronize(L->ev1));
        TIP_HIP(hipEventElapsedTime(elapsed_ms, L->ev0, L->ev1));
    }
    return TIP_OK;
}

extern "C" int tip_lattice_step_baoab(tip_lattice* L, const tip_step_params* p, int64_t nsteps,
                                      uint64_t step0, uint64_t seed, float* elapsed_ms)
{
    int rc = langevin_checks(L, p, nsteps, "tip_lattice_step_baoab");
    if (rc) return rc;
    if (nsteps == 0) return TIP_OK;
    // A deferred half-kick fuses into this call's first pass when the kick and drift
    // constants agree (dt, dx, mu; inv_a2 may differ: the fused pass takes both).
    const tip_step_params& q = L->baoab_prev;
    const bool fuse_first =
        L->baoab_pending && q.dt == p->dt && q.dx == p->dx && q.mu == p->mu;
    if (L->baoab_pending && !fuse_first) {
        rc = close_baoab(L);
        if (rc) return rc;
    }
    const double inv_a2_prev = fuse_first ? q.inv_a2 : p->inv_a2;
    if (elapsed_ms) TIP_HIP(hipEventRecord(L->ev0, L->stream));
    for (int64_t n = 0; n < nsteps; ++n) {
        const bool fused = n > 0 || fuse_first;
        rc = lattice_langevin(L, fused ? LV_BAOAB_FUSED : LV_BAOAB_OPEN, p,
                              n == 0 ? inv_a2_prev : p->inv_a2,
                              tip_noise_seed(step0 + (uint64_t)n, seed));
        if (rc) return rc;
        L->cur ^= 1;
    }
    L->baoab_pending = 1;
    L->baoab_prev = *p;
    if (elapsed_ms) {
        TIP_HIP(hipEventRecord(L->ev1, L->stream));
        TIP_HIP(hipEventSynchronize(L->ev1));
        TIP_HIP(hipEventElapsedTime(elapsed_ms, L->ev0, L->ev1));
    }
    return TIP_OK;
}

extern "C" int tip_lattice_step_mlp(tip_lattice* L, const tip_step_params* p,
                                    const float* mlp_params, const tip_mlp_features* fm, double T,
                                    double dT_per_step, int64_t nsteps, uint64_t step0,
                                    uint64_t seed, int64_t grid_points, float* elapsed_ms)
{
    TIP_REQUIRE(L && p && mlp_params && fm, "tip_lattice_step_mlp: NULL");
    int rc0 = close_baoab(L);
    if (rc0) return rc0;
    TIP_REQUIRE(L->kind == TIP_COMPLEX, "tip_lattice_step_mlp: needs a complex lattice");
    TIP_REQUIRE(nsteps >= 0, "tip_lattice_step_mlp: nsteps < 0");
    TIP_REQUIRE(grid_points == 0 || (grid_points >= 2 && grid_points <= (int64_t(1) << 24)),
                "tip_lattice_step_mlp: grid_points must be 0 or in [2, 2^24]");
    if (grid_points > 0 && grid_points - 1 > L->mlp_grid_cap) {
        TIP_HIP(hipStreamSynchronize(L->stream));
        if (L->mlp_grid) TIP_HIP(hipFree(L->mlp_grid));
        L->mlp_grid = nullptr;
        L->mlp_grid_cap = 0;
        TIP_HIP(hipMalloc(&L->mlp_grid, (size_t)(grid_points - 1) * sizeof(float2)));
        L->mlp_grid_cap = grid_points - 1;
    }
    TIP_REQUIRE(T > 0.0 && T + dT_per_step * (double)nsteps > 0.0,
                "tip_lattice_step_mlp: temperature must stay positive");
    const ZnParams<float> zn32 = make_zn32(L->zn_n, L->zn_dv, L->zn_active);
    const int hash = p->noise_scale != 0.0 ? 1 : 0;
    if (elapsed_ms) TIP_HIP(hipEventRecord(L->ev0, L->stream));
    for (int64_t n = 0; n < nsteps; ++n) {
        void** in = L->buf[L->cur];
        void** out = L->buf[L->cur ^ 1];
        // The Langevin amplitude sqrt(2 eta T dt / (mu^2 dx^3)) follows the cooling T.
        const double Tn = T + dT_per_step * (double)n;
        const StepScalars<float> c32 = make_scalars<float>(
            p->dt, p->dx, p->eta_eff, p->mu, p->inv_a2, p->noise_scale * sqrt(Tn / T));
        int rc = grid_points > 0
            ? launch_complex_mlp_grid(
                  (const float*)in[0], (const float*)in[1], (const float*)in[2],
                  (const float*)in[3], (float*)out[0], (float*)out[1], (float*)out[2],
                  (float*)out[3], c32, mlp_params, fm, Tn, zn32, L->nx, L->ny, L->nz,
                  tip_noise_seed(step0 + (uint64_t)n, seed), hash, nullptr, 0.0f, L->mlp_grid,
                  grid_points, L->stream)
            : launch_complex_mlp(
                  (const float*)in[0], (const float*)in[1], (const float*)in[2],
                  (const float*)in[3], (float*)out[0], (float*)out[1], (float*)out[2],
                  (float*)out[3], c32, mlp_params, fm, Tn, zn32, L->nx, L->ny, L->nz,
                  tip_noise_seed(step0 + (uint64_t)n, seed), hash, nullptr, 0.0f, L->stream);
        if (rc) return rc;
        L->cur ^= 1;
    }
    if (elapsed_ms) {
        TIP_HIP(hipEventRecord(L->ev1, L->stream));
        TIP_HIP(hipEventSynchronize(L->ev1));
        TIP_HIP(hipEventElapsedTime(elapsed_ms, L->ev0, L->ev1));
    }
    return TIP_OK;
}

extern "C" int tip_lattice_vacuum_count(tip_lattice* L, double esc, unsigned long long* count)
{
    TIP_REQUIRE(L && count, "tip_lattice_vacuum_count: NULL");
    TIP_HIP(hipMemsetAsync(L->counter, 0, sizeof(unsigned long long), L->stream));
    int64_t n = L->nx * L->ny * L->nz;
    void** b = L->buf[L->cur];
    int rc;
    if (L->kind == TIP_COMPLEX)
        rc = tip_vacuum_count_c32((const float*)b[0], (const float*)b[1], n, esc, L->counter,
                                  L->stream);
    else if (L->prec == 4)
        rc = tip_vacuum_count_f32((const float*)b[0], n, esc, L->counter, L->stream);
    else
        rc = tip_vacuum_count_f64((const double*)b[0], n, esc, L->counter, L->stream);
    if (rc) return rc;
    TIP_HIP(hipMemcpyAsync(count, L->counter, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           L->stream));
    TIP_HIP(hipStreamSynchronize(L->stream));
    return TIP_OK;
}

extern "C" int tip_lattice_winding(tip_lattice* L, int8_t* plaq_dev, float* wsum_dev,
                                   unsigned long long* counts2)
{
    TIP_REQUIRE(L && counts2, "tip_lattice_winding: NULL");
    TIP_REQUIRE(L->kind == TIP_COMPLEX, "tip_lattice_winding: needs a complex lattice");
    TIP_HIP(hipMemsetAsync(L->counter, 0, 2 * sizeof(unsigned long long), L->stream));
    void** b = L->buf[L->cur];
    int rc = tip_winding_f32((const float*)b[0], (const float*)b[1], plaq_dev, wsum_dev,
                             L->counter, L->nx, L->ny, L->nz, L->stream);
    if (rc) return rc;
    TIP_HIP(hipMemcpyAsync(counts2, L->counter, 2 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, L->stream));
    TIP_HIP(hipStreamSynchronize(L->stream));
    return TIP_OK;
}

extern "C" int tip_lattice_rho_theta(tip_lattice* L, float* rho_dev, float* theta_dev,
                                     double* rho_minmax_host)
{
    TIP_REQUIRE(L, "tip_lattice_rho_theta: NULL");
    TIP_REQUIRE(L->kind == TIP_COMPLEX && L->prec == 4,
                "tip_lattice_rho_theta: needs a float32 complex lattice");
    void** b = L->buf[L->cur];
    int rc = tip_rho_theta_f32((const float*)b[0], (const float*)b[1], L->nx * L->ny * L->nz,
                               rho_dev, theta_dev, rho_minmax_host ? L->mm : nullptr, L->stream);
    if (rc) return rc;
    if (rho_minmax_host)
        TIP_HIP(hipMemcpyAsync(rho_minmax_host, L->mm, 2 * sizeof(double),
                               hipMemcpyDeviceToHost, L->stream));
    TIP_HIP(hipStreamSynchronize(L->stream));
    return TIP_OK;
}

extern "C" int tip_lattice_minmax(tip_lattice* L, double* out2)
{
    TIP_REQUIRE(L && out2, "tip_lattice_minmax: NULL");
    int64_t n = L->nx * L->ny * L->nz;
    void** b = L->buf[L->cur];
    int rc;
    if (L->kind == TIP_COMPLEX) {
        TIP_HIP(hipMemsetAsync(L->mm, 0, 2 * sizeof(double), L->stream));
        rc = tip_rho_max_c32((const float*)b[0], (const float*)b[1], n, L->mm + 1, L->stream);
    } else if (L->prec == 4)
        rc = tip_minmax_f32((const float*)b[0], n, L->mm, L->stream);
    else
        rc = tip_minmax_f64((const double*)b[0], n, L->mm, L->stream);
    if (rc) return rc;
    TIP_HIP(hipMemcpyAsync(out2, L->mm, 2 * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    TIP_HIP(hipStreamSynchronize(L->stream));
    return TIP_OK;
}
