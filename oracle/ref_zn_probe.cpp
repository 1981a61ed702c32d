// ref_zn_probe.cpp -- TEST INFRASTRUCTURE ONLY (the checker's checker).
//
// Evaluates the reference's own C++ force of the complex field,
// ThermalInflation::ThermalTables::vPrimeComponents (cosmolattice_ext/models/
// thermal_tables.hpp:161-178: radial V'(rho) by the chain rule plus the Z_N angular term),
// compiled from the header where it lies in the checkout.
//
// argv[1]: a J table in the header's binary layout.  stdin lines: phi1 phi2 T N dV.
// stdout per line: dV1 dV2 with the Z_N term active, then dV1 dV2 with it off (%.17g).
#include <cstdio>

#include "thermal_tables.hpp"

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    ThermalInflation::ThermalTables tab(argv[1]);
    double p1, p2, T, n, dv;
    while (std::scanf("%lf %lf %lf %lf %lf", &p1, &p2, &T, &n, &dv) == 5) {
        double a1, a2, b1, b2;
        tab.vPrimeComponents(p1, p2, T, n, dv, true, a1, a2);
        tab.vPrimeComponents(p1, p2, T, n, dv, false, b1, b2);
        std::printf("%.17g %.17g %.17g %.17g\n", a1, a2, b1, b2);
    }
    return 0;
}
