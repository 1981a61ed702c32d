// tip_langevin_f64.hip -- float64 overdamped Euler-Maruyama and BAOAB steps, real and complex.
// Compiled with -ffp-contract=off: every product and sum rounds separately in the order the
// reference writes them (latticeSimeRescale_numba.py:2049-2060, 2092-2112, 2170-2192), so the
// only differences from a float64 CPU run of the same form are the libm transcendentals of the
// hash noise and of the complex Z_N term.
#include "langevin_kernels.h"

using namespace tip;

namespace {

TableF64 table64(const double* t, double tmin, double dinv, int64_t npts, int odd)
{
    TableF64 tab;
    tab.t = t;
    tab.tmin = tmin;
    tab.dinv = dinv;
    tab.last = (int)(npts - 2);
    tab.odd = odd ? 1 : 0;
    return tab;
}

int check_args(const char* who, int64_t nx, int64_t ny, int64_t nz, const void* table,
               int64_t npts, double eta)
{
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(table, "%s: NULL table", who);
    TIP_REQUIRE(npts >= 2 && npts < (int64_t(1) << 30), "%s: vp_npts out of range", who);
    TIP_REQUIRE(eta > 0.0, "%s: eta / gamma must be positive", who);
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

namespace tip {

int langevin_real_f64(int op, const double* phi, const double* pi, double* phi_out,
                      double* pi_out, const LangevinScalars<double>& c, const double* table,
                      double tmin, double dinv, int64_t npts, int odd, int64_t nx, int64_t ny,
                      int64_t nz, uint64_t noise_seed, int hash, hipStream_t st)
{
    return launch_langevin_real<double, TableF64>(op, phi, pi, phi_out, pi_out, c,
                                                  table64(table, tmin, dinv, npts, odd), nx, ny,
                                                  nz, noise_seed, hash, st);
}

}  // namespace tip

extern "C" int tip_overdamped_f64(const double* phi, double* phi_out, double dt, double dx,
                                  double eta_eff, double mu, double inv_a2,
                                  const double* vp_table, double vp_tmin, double vp_dinv,
                                  int64_t vp_npts, int vp_odd, int64_t nx, int64_t ny,
                                  int64_t nz, uint64_t noise_seed, double noise_scale,
                                  void* stream)
{
    int rc = check_args("overdamped f64", nx, ny, nz, vp_table, vp_npts, eta_eff);
    if (rc) return rc;
    TIP_REQUIRE(phi && phi_out && phi != phi_out, "overdamped f64: NULL or aliased field");
    return langevin_real_f64(LV_OVERDAMPED, phi, nullptr, phi_out, nullptr,
                             make_langevin<double>(dt, dx, eta_eff, mu, inv_a2, inv_a2,
                                                   noise_scale),
                             vp_table, vp_tmin, vp_dinv, vp_npts, vp_odd, nx, ny, nz, noise_seed,
                             noise_scale != 0.0, as_stream(stream));
}

extern "C" int tip_baoab_f64(int stage, const double* phi, const double* pi, double* phi_out,
                             double* pi_out, double dt, double dx, double gamma, double mu,
                             double inv_a2, double inv_a2_prev, const double* vp_table,
                             double vp_tmin, double vp_dinv, int64_t vp_npts, int vp_odd,
                             int64_t nx, int64_t ny, int64_t nz, uint64_t noise_seed,
                             double noise_scale, void* stream)
{
    int rc = check_args("baoab f64", nx, ny, nz, vp_table, vp_npts, gamma);
    if (rc) return rc;
    rc = check_stage("baoab f64", stage, phi, pi, phi_out, pi_out);
    if (rc) return rc;
    return langevin_real_f64(stage, phi, pi, phi_out, pi_out,
                             make_langevin<double>(dt, dx, gamma, mu, inv_a2, inv_a2_prev,
                                                   noise_scale),
                             vp_table, vp_tmin, vp_dinv, vp_npts, vp_odd, nx, ny, nz, noise_seed,
                             noise_scale != 0.0, as_stream(stream));
}

namespace {

ZnParams<double> zn64(double zn_n, double zn_dv, double zn_active)
{
    ZnParams<double> zn;
    zn.n = zn_n;
    zn.dv = zn_dv;
    zn.active = (zn_active > 0.5 && zn_n > 0.5) ? 1 : 0;
    zn.ipow = 0;
    return zn;
}

}  // namespace

extern "C" int tip_overdamped_c64(const double* phi1, const double* phi2, double* phi1_out,
                                  double* phi2_out, double dt, double dx, double eta_eff,
                                  double mu, double inv_a2, const double* vp_table,
                                  double vp_tmin, double vp_dinv, int64_t vp_npts, double zn_n,
                                  double zn_dv, double zn_active, int64_t nx, int64_t ny,
                                  int64_t nz, uint64_t noise_seed, double noise_scale,
                                  void* stream)
{
    int rc = check_args("overdamped c64", nx, ny, nz, vp_table, vp_npts, eta_eff);
    if (rc) return rc;
    TIP_REQUIRE(phi1 && phi2 && phi1_out && phi2_out && phi1 != phi1_out && phi2 != phi2_out,
                "overdamped c64: NULL or aliased field");
    return launch_langevin_complex<double, TableF64>(
        LV_OVERDAMPED, phi1, phi2, nullptr, nullptr, phi1_out, phi2_out, nullptr, nullptr,
        make_langevin<double>(dt, dx, eta_eff, mu, inv_a2, inv_a2, noise_scale),
        table64(vp_table, vp_tmin, vp_dinv, vp_npts, 0), zn64(zn_n, zn_dv, zn_active), nx, ny,
        nz, noise_seed, noise_scale != 0.0, as_stream(stream));
}

extern "C" int tip_baoab_c64(int stage, const double* phi1, const double* phi2,
                             const double* pi1, const double* pi2, double* phi1_out,
                             double* phi2_out, double* pi1_out, double* pi2_out, double dt,
                             double dx, double gamma, double mu, double inv_a2,
                             double inv_a2_prev, const double* vp_table, double vp_tmin,
                             double vp_dinv, int64_t vp_npts, double zn_n, double zn_dv,
                             double zn_active, int64_t nx, int64_t ny, int64_t nz,
                             uint64_t noise_seed, double noise_scale, void* stream)
{
    int rc = check_args("baoab c64", nx, ny, nz, vp_table, vp_npts, gamma);
    if (rc) return rc;
    rc = check_stage("baoab c64", stage, phi1, pi1, phi1_out, pi1_out);
    if (rc) return rc;
    rc = check_stage("baoab c64", stage, phi2, pi2, phi2_out, pi2_out);
    if (rc) return rc;
    return launch_langevin_complex<double, TableF64>(
        stage, phi1, phi2, pi1, pi2, phi1_out, phi2_out, pi1_out, pi2_out,
        make_langevin<double>(dt, dx, gamma, mu, inv_a2, inv_a2_prev, noise_scale),
        table64(vp_table, vp_tmin, vp_dinv, vp_npts, 0), zn64(zn_n, zn_dv, zn_active), nx, ny,
        nz, noise_seed, noise_scale != 0.0, as_stream(stream));
}
