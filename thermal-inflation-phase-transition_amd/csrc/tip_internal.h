// tip_internal.h -- shared host/device helpers of libtip (MI355X / gfx950).
//
// Lattice layout in HBM: every field is one C-contiguous (Nx, Ny, Nz) array with z fastest,
// site = (i*Ny + j)*Nz + k, as the reference's numpy arrays
// (simulation/latticeSimeRescale_numba.py:1259-1260).  A wave of 64 lanes runs along z, so
// every centre and y/x-neighbour load is one coalesced 256-byte (float) row segment.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <string>

#include "../../include/tip.h"
#include "bm_math.h"

namespace tip {

// ---------------------------------------------------------------------------------------
// Error handling: every C-ABI entry point returns TIP_OK or a negative code and leaves a
// message for tip_last_error().
// ---------------------------------------------------------------------------------------
void set_error(const char* fmt, ...);
int fail_hip(hipError_t e, const char* what);

#define TIP_HIP(call)                                              \
    do {                                                           \
        hipError_t _e = (call);                                    \
        if (_e != hipSuccess) return ::tip::fail_hip(_e, #call);   \
    } while (0)

#define TIP_REQUIRE(cond, ...)                                     \
    do {                                                           \
        if (!(cond)) {                                             \
            ::tip::set_error(__VA_ARGS__);                         \
            return TIP_ERR_ARG;                                    \
        }                                                          \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Lattice dimensions are bounded so that a site index fits a signed 32-bit int inside the
// kernels (1024^3 = 2^30 sites); the hash seed still uses the 64-bit site index.
constexpr int64_t kMaxSites = int64_t(1) << 31;

int check_dims(int64_t nx, int64_t ny, int64_t nz);

// Library-owned device scratch, grown on demand (used by drop-in entry points that need a
// converted table or a counter).  One buffer per (device, slot).  scratch_get makes stream st
// wait for the slot's last use on another stream; scratch_release records the end of this
// use on st (ScratchLease does it when the entry point returns, after its launches).
// Calls from several host threads at once that use one slot must be serialised by the
// caller: the slot is one buffer.
void* scratch_get(size_t bytes, int slot, hipStream_t st);
void scratch_release(int slot, hipStream_t st);

struct ScratchLease {
    int slot;
    hipStream_t st;
    ~ScratchLease() { scratch_release(slot, st); }
};

// ---------------------------------------------------------------------------------------
// Counter-based hash RNG of the reference (simulation/latticeSimeRescale_numba.py:2322-2336).
// x1 = state after the first mixer, x2 after the second.  u = (x >> 11) * 2^-53 is exact in
// float64.  u1 is clamped below at 1e-12 as in the reference.
// ---------------------------------------------------------------------------------------
__host__ __device__ __forceinline__ void hash_states(uint64_t seed, uint64_t& x1, uint64_t& x2)
{
    uint64_t x = seed;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    x = x ^ (x >> 31);
    x1 = x;
    x = (x * 0x2545F4914F6CDD1DULL) ^ 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = x ^ (x >> 31);
    x2 = x;
}

__host__ __device__ __forceinline__ double hash_unit(uint64_t x)
{
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// Site seed of the fused reference kernels: uint64(site*K) ^ uint64(noise_seed), K = 73856093
// for the full-step noise (simulation/latticeSimeRescale_numba.py:2514-2516).
constexpr int64_t kNoiseMult = 73856093;

__host__ __device__ __forceinline__ uint64_t site_seed(int64_t site, uint64_t noise_seed)
{
    return (uint64_t)(site * kNoiseMult) ^ noise_seed;
}

// float64 Box-Muller as written in the reference (cos for the first component, sin for the
// second of the complex pair, latticeSimComplex_numba.py:2965-2974), with the log, sqrt and
// cos / sin of bm_math.h on the draw's bounded domains (within a few ulp of libm).  The real
// field's single draw takes cos(2 pi u2) from one polynomial (bm_cos2pi); the pair takes both
// from bm_sincos.
__device__ __forceinline__ double gauss_f64(uint64_t seed)
{
    uint64_t x1, x2;
    hash_states(seed, x1, x2);
    const double u1 = fmax(hash_unit(x1), 1e-12), u2 = hash_unit(x2);
    return bm_sqrt(bm_neg2log(u1)) * bm_cos2pi(u2);
}

__device__ __forceinline__ void gauss_pair_f64(uint64_t seed, double& zc, double& zs)
{
    uint64_t x1, x2;
    hash_states(seed, x1, x2);
    const double u1 = fmax(hash_unit(x1), 1e-12), u2 = hash_unit(x2);
    double r = bm_sqrt(bm_neg2log(u1));
    double s, c;
    bm_sincos(6.283185307179586 * u2, &s, &c);
    zc = r * c;
    zs = r * s;
}

// float32 Box-Muller from the same hash states: u's in float32, then the hardware
// transcendentals.  v_cos/v_sin take their argument in revolutions, so cos(2*pi*u2) is one
// instruction; -2 ln u1 = -2 ln2 * log2 u1.
// (x >> 11) * 2^-53 in float32 without a float64 detour: the high word converts exactly up to
// 24 bits and the (truncated) low word adds the remaining fraction; the result is within one
// float32 ulp of the rounded float64 value.  `word_mask` is a runtime all-ones word: ANDing
// the high word with it stops the compiler from re-widening the 32-bit conversion into a
// multi-instruction 64-bit one (it otherwise proves the word came from a 64-bit shift).
__device__ __forceinline__ float unit_f32(uint64_t x, uint32_t word_mask)
{
    const uint32_t hi = (uint32_t)(x >> 32) & word_mask;
    const uint32_t lo = (uint32_t)x & 0xFFFFF800u;
    return fmaf((float)lo, 0x1p-64f, (float)hi * 0x1p-32f);
}

__device__ __forceinline__ float gauss_f32(uint64_t seed, uint32_t word_mask)
{
    uint64_t x1, x2;
    hash_states(seed, x1, x2);
    float u1 = fmaxf(unit_f32(x1, word_mask), 1e-12f);
    float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
    return r * __builtin_amdgcn_cosf(unit_f32(x2, word_mask));
}

__device__ __forceinline__ void gauss_pair_f32(uint64_t seed, uint32_t word_mask, float& zc,
                                               float& zs)
{
    uint64_t x1, x2;
    hash_states(seed, x1, x2);
    float u1 = fmaxf(unit_f32(x1, word_mask), 1e-12f);
    float u2 = unit_f32(x2, word_mask);
    float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
    zc = r * __builtin_amdgcn_cosf(u2);
    zs = r * __builtin_amdgcn_sinf(u2);
}

// ---------------------------------------------------------------------------------------
// Raw buffer access.  A buffer resource (base, byte range) lives in scalar registers; a load
// is one instruction with a 32-bit per-lane byte offset and no vector address arithmetic.
// 0x00020000 is the GFX9-family (gfx950) raw-buffer word 3 (32-bit data format).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             0x00020000);
}

template <typename R>
__device__ __forceinline__ R buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

// The same with a wave-uniform byte offset in a scalar register (soffset), so row offsets that
// are equal across the wave cost no vector registers.  Callers keep voffset + soffset inside
// the resource's range for every lane that must not be dropped, and out-of-range lanes at or
// past the range without 32-bit wrap-around.
__device__ __forceinline__ float buf_load_s(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float,
                              __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void buf_store_s(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                            uint32_t soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff,
                                          (int)soff, 0);
}

__device__ __forceinline__ void buf_store(float v, __amdgpu_buffer_rsrc_t r, uint32_t off)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 0);
}
__device__ __forceinline__ void buf_store(double v, __amdgpu_buffer_rsrc_t r, uint32_t off)
{
    using u2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)off, 0, 0);
}

// ---------------------------------------------------------------------------------------
// V' tables.
//
// TableF64: the reference lookup verbatim (latticeSimeRescale_numba.py:1923-1930): truncating
// int(), index clamped to [0, npts-2], frac unclamped (linear extrapolation).  odd != 0 reads
// a table tabulated on x >= 0 at |x| and gives the result the sign of x (bit operations on the
// sign bit: a zero takes the sign of x's zero).
//
// TableF32: the same piecewise-linear function stored per segment as (A_i, B_i) with
// V'(x) = A_i + B_i x, prepared in float64 from the float64 table.  Evaluating in this form
// keeps float32 accurate where x is small next to the table origin (x - tmin would cancel).
// ---------------------------------------------------------------------------------------
struct TableF64 {
    const double* t;
    double tmin, dinv;
    int last;  // npts - 2
    int odd;
    // fetch() issues the two table loads of x's segment, eval() finishes the interpolation
    // (recomputing the index), so a caller can issue several sites' loads before it needs them.
    __device__ __forceinline__ uint64_t sign_mask() const
    {
        return odd ? 0x8000000000000000ull : 0ull;
    }
    __device__ __forceinline__ void index(double x, double& fidx, int& idx) const
    {
        const double ax =
            __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) & ~sign_mask());
        fidx = (ax - tmin) * dinv;
        // int() truncates toward zero, so clamping fidx to [0, last] first gives the
        // reference's clamp of the truncated index (and 0 for NaN)
        idx = (int)fmin(fmax(fidx, 0.0), (double)last);
    }
    __device__ __forceinline__ double2 fetch(double x) const
    {
        double fidx;
        int idx;
        index(x, fidx, idx);
        return make_double2(t[idx], t[idx + 1]);
    }
    __device__ __forceinline__ double eval(double2 s, double x) const
    {
        double fidx;
        int idx;
        index(x, fidx, idx);
        const double frac = fidx - (double)idx;
        const double v = s.x + frac * (s.y - s.x);
        return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, v) ^
                                              (__builtin_bit_cast(uint64_t, x) & sign_mask()));
    }
    __device__ __forceinline__ double operator()(double x) const { return eval(fetch(x), x); }
};

// ODD is a compile-time flag so the even table pays nothing for the odd extension.  The
// segment index is the truncation of fidx clamped to [0, last] in float: for the segment
// choice this equals the reference's clamp of the truncated index (frac is not needed in
// the A + B x form).
template <bool ODD>
struct TableF32 {
    const float2* ab;
    float tmin, dinv;
    int last;
    float off;  // -tmin * dinv: fidx = x * dinv + off in one fma
    __device__ __forceinline__ float operator()(float x) const
    {
        const float ax = ODD ? fabsf(x) : x;
        const float fidx = (ax - tmin) * dinv;
        const int idx = (int)__builtin_amdgcn_fmed3f(fidx, 0.0f, (float)last);
        const float2 s = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(ab) +
                                                          (uint32_t)idx * 8u);
        const float v = fmaf(s.y, ax, s.x);
        return (ODD && x < 0.0f) ? -v : v;
    }
};

// Host-side description of a float32 segment table (the ODD flag becomes a template argument
// at launch).
struct SegTable {
    const float2* ab;
    float tmin, dinv;
    int last;
    int odd;
    float off;
    template <bool ODD>
    TableF32<ODD> as() const
    {
        TableF32<ODD> t;
        t.ab = ab;
        t.tmin = tmin;
        t.dinv = dinv;
        t.last = last;
        t.off = off;
        return t;
    }
};

// Convert a float64 reference table into float32 (A, B) segments on the device.
int prepare_table_f32(const double* table_dev, int64_t npts, double tmin, double dinv,
                      float2* ab_dev, hipStream_t s);

// The V' table producer on the device (tip_potential.hip): table[i] = V'(phi_i, T) on
// numpy.linspace(phi_min, phi_max, n); table_grid gives the reference's (tmin, dinv).
int build_vprime_table_dev(double* table, int64_t n, double T, double phi_min, double phi_max,
                           const tip_vprime_model& m, const double* cb, const double* cf,
                           hipStream_t s);
void table_grid(int64_t n, double phi_min, double phi_max, double* tmin, double* dinv);

// Lookup on the host of one segment table built from a host table (used for argument checks).
int check_table(const void* table, int64_t npts, double dinv);

}  // namespace tip
