// tip_step_f64.hip -- float64 fused RK2 step (the drop-in parity path).
// Compiled with -ffp-contract=off: every product and sum rounds separately, in the order the
// reference writes them (simulation/latticeSimeRescale_numba.py:1912-1933,1945-1966), so the
// only differences from a float64 CPU run are the libm transcendentals of the hash noise.
#include "step_kernels.h"

using namespace tip;

namespace {

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
    StepGeometry g;
    TileGrid tg;
    int nb;
    rc = resolve_geometry(nullptr, nx, ny, nz, 0, 1, g, tg, nb);
    if (rc) return rc;
    int hash = (noise == nullptr && noise_scale != 0.0) ? 1 : 0;
    const TableF64 tab = make_table(vp_table, vp_tmin, vp_dinv, vp_npts, vp_odd);
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
