// tip_potential.hip -- the V' table producer on the device (float64).
//
// Reference: Vprime_scalar + build_vprime_table, simulation/latticeSimeRescale_numba.py:805-849
// (complex driver :818-862): table[i] = V'(phi_i, T) on the grid
// phi_i = numpy.linspace(phi_min, phi_max, n)[i], V' = tree + thermal boson + thermal fermion
// through the not-a-knot cubic splines of dJ_b and dJ_f (:432-455) with the reference's clamp
// and Horner rule.  With Hubble expansion the drivers rebuild the table every few steps
// (:3650-3672); building it here, into the lattice's own table buffer, removes the host
// evaluation and the upload from the loop.
//
// Compiled with -ffp-contract=off: every product and sum rounds as in the reference's
// float64 text (no fused multiply-adds).
#include "tip_internal.h"

namespace tip {

namespace {

// The reference's spline evaluation (:812-820): clamp u to [x_min, x_min + h nseg - 1e-12],
// truncating segment index clamped to [0, nseg - 1], Horner in the local coordinate.  c holds
// scipy's CubicSpline.c layout, rows = powers (3, 2, 1, 0), row-major (4, nseg).
__device__ __forceinline__ double spline_eval(const double* __restrict__ c, double x_min,
                                              double h, int nseg, double u)
{
    double uc = fmin(fmax(u, x_min), x_min + h * (double)nseg - 1e-12);
    int idx = (int)((uc - x_min) / h);
    idx = idx < 0 ? 0 : (idx > nseg - 1 ? nseg - 1 : idx);
    const double ds = uc - (x_min + (double)idx * h);
    return ((c[idx] * ds + c[nseg + idx]) * ds + c[2 * nseg + idx]) * ds + c[3 * nseg + idx];
}

// V'(phi, T) in the operation order of Vprime_scalar (:805-837).
__device__ double vprime_point(double phi, double T, const tip_vprime_model& m,
                               const double* __restrict__ cb, const double* __restrict__ cf)
{
    double dV = m.lam * (phi * phi * phi) - m.mphi * m.mphi * phi;
    const double T2 = T * T;
    const double pref = (T2 * T2) / (2.0 * M_PI * M_PI);
    const double gb2 = m.yb * m.yb;
    const double gg2 = m.gb * m.gb;
    const double coef_b = 0.25 * gb2 + (2.0 / 3.0) * gg2;
    const double xb_sq = m.mb2 + 0.5 * gb2 * (phi * phi) + coef_b * T2;
    if (xb_sq > 0.0) {
        const double xb = sqrt(xb_sq) / T;
        const double dJb = spline_eval(cb, m.x_min, m.h, (int)m.nseg, xb);
        const double dxb = 0.5 * gb2 * phi / (T2 * fmax(xb, 1e-20));
        dV += pref * m.coeff_b * dJb * dxb;
    }
    const double gf2 = m.yf * m.yf;
    const double gfg2 = m.gf * m.gf;
    const double xf_sq = 0.5 * gf2 * (phi * phi) + (1.0 / 6.0) * gfg2 * T2;
    if (xf_sq > 0.0) {
        const double xf = sqrt(xf_sq) / T;
        const double dJf = spline_eval(cf, m.x_min, m.h, (int)m.nseg, xf);
        const double dxf = 0.5 * gf2 * phi / (T2 * fmax(xf, 1e-20));
        dV += pref * m.coeff_f * dJf * dxf;
    }
    return dV;
}

// One point per thread.  The grid is numpy.linspace's: phi_i = i * step + phi_min with
// step = (phi_max - phi_min) / (n - 1), and the last point exactly phi_max.
__global__ __launch_bounds__(256) void k_build_vprime_table(
    double* __restrict__ table, int64_t n, double T, double phi_min, double phi_max, double step,
    tip_vprime_model m, const double* __restrict__ cb, const double* __restrict__ cf)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double phi = (i == n - 1) ? phi_max : (double)i * step + phi_min;
    table[i] = vprime_point(phi, T, m, cb, cf);
}

}  // namespace

int build_vprime_table_dev(double* table, int64_t n, double T, double phi_min, double phi_max,
                           const tip_vprime_model& m, const double* cb, const double* cf,
                           hipStream_t s)
{
    TIP_REQUIRE(table && cb && cf, "build_vprime_table: NULL pointer");
    TIP_REQUIRE(n >= 2 && n < (int64_t(1) << 30), "build_vprime_table: n out of range");
    TIP_REQUIRE(T > 0.0, "build_vprime_table: T must be positive");
    TIP_REQUIRE(phi_max > phi_min, "build_vprime_table: phi_max must exceed phi_min");
    TIP_REQUIRE(m.nseg >= 1 && m.nseg < (int64_t(1) << 24) && m.h > 0.0,
                "build_vprime_table: bad spline grid (nseg %lld, h %g)", (long long)m.nseg, m.h);
    const double step = (phi_max - phi_min) / (double)(n - 1);
    hipLaunchKernelGGL(k_build_vprime_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       table, n, T, phi_min, phi_max, step, m, cb, cf);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

void table_grid(int64_t n, double phi_min, double phi_max, double* tmin, double* dinv)
{
    // the reference's return value (:849): (float(phi_min), 1 / ((phi_max - phi_min) / (n - 1)))
    *tmin = phi_min;
    *dinv = 1.0 / ((phi_max - phi_min) / (double)(n - 1));
}

}  // namespace tip

using namespace tip;

extern "C" int tip_build_vprime_table(double* table, int64_t n, double T, double phi_min,
                                      double phi_max, const tip_vprime_model* model,
                                      const double* cb, const double* cf, double* tmin,
                                      double* dinv, void* stream)
{
    TIP_REQUIRE(model, "tip_build_vprime_table: NULL model");
    int rc = build_vprime_table_dev(table, n, T, phi_min, phi_max, *model, cb, cf,
                                    as_stream(stream));
    if (rc) return rc;
    double a, b;
    table_grid(n, phi_min, phi_max, &a, &b);
    if (tmin) *tmin = a;
    if (dinv) *dinv = b;
    return TIP_OK;
}
