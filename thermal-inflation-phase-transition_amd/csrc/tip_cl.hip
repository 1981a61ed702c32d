// tip_cl.hip -- the initial conditions of the CosmoLattice port's numba-compatible start
// (cosmolattice_ext/models/thermal_inflation.h:174-222, applyNumbaInitialConditions): white
// noise from a stateless hash of the global site index, so a CosmoLattice run and a run here
// start from the same field.
//   site = (ix N + iy) N + iz (thermal_inflation.h:176-186; (ix ny + iy) nz + iz here)
//   h(x)  = splitmix64 finaliser (:188-195)
//   z(s)  = sqrt(-2 ln u1) cos(2 pi u2), u1 = ((h(s) >> 11) + 1) 2^-53,
//           u2 = (h(s ^ 0x9E3779B97F4A7C15) >> 11) 2^-53  (:197-204)
//   component 0: amp z(site ^ kIcSeed); component 1: amp z(site ^ 0x9E3779B97F4A7C15 ^ kIcSeed)
//   with kIcSeed = 0x4E554D42412D4943 ("NUMBA-IC") (:214-227).
#include <hip/hip_runtime.h>

#include "tip_internal.h"

using namespace tip;

namespace {

constexpr uint64_t kIcSeed = 0x4E554D42412D4943ULL;
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}

__device__ __forceinline__ double hash_gaussian(uint64_t seed)
{
    const uint64_t h1 = mix64(seed);
    const uint64_t h2 = mix64(seed ^ kGolden);
    const double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * log(u1 > 1e-300 ? u1 : 1e-300));
    return r * cos(6.283185307179586 * u2);
}

template <typename R>
__global__ __launch_bounds__(256) void k_cl_ics(R* __restrict__ f0, R* __restrict__ f1,
                                                int64_t n, double amp)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t site = (uint64_t)s;
        f0[s] = (R)(amp * hash_gaussian(site ^ kIcSeed));
        if (f1) f1[s] = (R)(amp * hash_gaussian(site ^ kGolden ^ kIcSeed));
    }
}

template <typename R>
int cl_ics(R* f0, R* f1, int64_t nx, int64_t ny, int64_t nz, double amp, void* stream)
{
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(f0, "tip_cl_hash_ics: NULL field");
    const int64_t n = nx * ny * nz;
    int64_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    hipLaunchKernelGGL((k_cl_ics<R>), dim3((unsigned)b), dim3(256), 0, as_stream(stream), f0, f1,
                       n, amp);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

}  // namespace

extern "C" int tip_cl_hash_ics_f32(float* fld0, float* fld1, int64_t nx, int64_t ny, int64_t nz,
                                   double amp, void* stream)
{
    return cl_ics<float>(fld0, fld1, nx, ny, nz, amp, stream);
}

extern "C" int tip_cl_hash_ics_f64(double* fld0, double* fld1, int64_t nx, int64_t ny,
                                   int64_t nz, double amp, void* stream)
{
    return cl_ics<double>(fld0, fld1, nx, ny, nz, amp, stream);
}
