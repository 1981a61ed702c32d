// ref_hash_probe.cpp -- TEST INFRASTRUCTURE ONLY (the checker's checker).
//
// Runs the reference's own C++ text of the counter-based hash RNG, hashMix64 / hashRngPair of
// cosmolattice_ext/evolvers/stochasticrk.h:205-225.  That header needs TempLat (absent), so
// oracle/build_ref.py copies the two static member functions, unchanged, from the checkout
// into oracle/_ref/stochasticrk_hash.inc at build time, and this program includes them as
// members of a class template with T = double, exactly as the header declares them.
//
// stdin: one uint64 seed (decimal) per line.  stdout: u1 and u2 per seed as the hex bits of
// the float64 values; u = (x >> 11) * 2^-53 is exact, so the bits carry x >> 11.
#include <cinttypes>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

template <class T>
struct RefHash {
#include "stochasticrk_hash.inc"
};

int main()
{
    unsigned long long s = 0;
    while (std::scanf("%llu", &s) == 1) {
        double u1 = 0.0, u2 = 0.0;
        RefHash<double>::hashRngPair((uint64_t)s, u1, u2);
        uint64_t b1, b2;
        std::memcpy(&b1, &u1, 8);
        std::memcpy(&b2, &u2, 8);
        std::printf("%016" PRIx64 " %016" PRIx64 "\n", b1, b2);
    }
    return 0;
}
