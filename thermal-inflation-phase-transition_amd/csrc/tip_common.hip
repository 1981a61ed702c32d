// tip_common.hip -- error state, memory plumbing, hash RNG entry points and the in-loop
// observables (kernel (b) plaquette winding, kernel (c) true-vacuum count, min/max).
#include <stdarg.h>
#include <stdio.h>

#include <mutex>
#include <vector>

#include "tip_internal.h"

namespace tip {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int fail_hip(hipError_t e, const char* what)
{
    set_error("%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
    return e == hipErrorOutOfMemory ? TIP_ERR_NOMEM : TIP_ERR_HIP;
}

int check_dims(int64_t nx, int64_t ny, int64_t nz)
{
    TIP_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "lattice dims must be >= 1 (got %lld,%lld,%lld)",
                (long long)nx, (long long)ny, (long long)nz);
    TIP_REQUIRE(nx <= 65535 && ny <= 65535 && nz <= 65535 && nx * ny * nz < kMaxSites,
                "lattice %lldx%lldx%lld exceeds 2^31 sites", (long long)nx, (long long)ny,
                (long long)nz);
    // byte offsets inside one x plane are 32-bit in the step kernels (float64: 8 B/site)
    TIP_REQUIRE(ny * nz < (int64_t(1) << 29), "lattice plane %lldx%lld exceeds 2^29 sites",
                (long long)ny, (long long)nz);
    return TIP_OK;
}

// Scratch buffers per (device, slot).  Growing a buffer synchronises the device first so
// no kernel still reads the old one.  Each slot also keeps an event recorded on the stream of
// its last use (scratch_release); a call on another stream waits for that event before it
// writes the slot, so entry points on different streams never overwrite a buffer a kernel is
// still reading.  Streams under graph capture neither wait nor record.
namespace {
std::mutex g_scratch_mu;
struct Scratch {
    void* p = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    hipStream_t last = nullptr;
    bool recorded = false;
};
std::vector<std::vector<Scratch>> g_scratch;

bool capturing(hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

Scratch* scratch_slot(int slot)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
    auto& v = g_scratch[dev];
    if ((int)v.size() <= slot) v.resize(slot + 1);
    return &v[slot];
}
}  // namespace

void* scratch_get(size_t bytes, int slot, hipStream_t st)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    Scratch* sp = scratch_slot(slot);
    if (!sp) return nullptr;
    Scratch& s = *sp;
    if (s.recorded && s.last != st && !capturing(st) &&
        hipStreamWaitEvent(st, s.ev, 0) != hipSuccess)
        return nullptr;
    if (s.bytes >= bytes && s.p) return s.p;
    if (s.p) {
        if (hipDeviceSynchronize() != hipSuccess) return nullptr;
        (void)hipFree(s.p);
        s.p = nullptr;
        s.bytes = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    if (hipMalloc(&s.p, want) != hipSuccess) {
        s.p = nullptr;
        return nullptr;
    }
    s.bytes = want;
    return s.p;
}

void scratch_release(int slot, hipStream_t st)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    Scratch* sp = scratch_slot(slot);
    if (!sp || capturing(st)) return;
    Scratch& s = *sp;
    if (!s.ev && hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess) {
        s.ev = nullptr;
        return;
    }
    if (hipEventRecord(s.ev, st) == hipSuccess) {
        s.last = st;
        s.recorded = true;
    }
}

}  // namespace tip

using namespace tip;

// ------------------------------------------------------------------------------------------
// Plumbing
// ------------------------------------------------------------------------------------------
extern "C" int tip_version(int* major, int* minor)
{
    if (major) *major = 0;
    if (minor) *minor = 1;
    return TIP_OK;
}

extern "C" int tip_last_error(char* buf, size_t len)
{
    if (!buf || len == 0) return TIP_ERR_ARG;
    snprintf(buf, len, "%s", g_err);
    return TIP_OK;
}

extern "C" int tip_device_count(int* count)
{
    TIP_REQUIRE(count, "tip_device_count: NULL");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return fail_hip(e, "hipGetDeviceCount");
    }
    return TIP_OK;
}

extern "C" int tip_set_device(int device)
{
    TIP_HIP(hipSetDevice(device));
    return TIP_OK;
}

extern "C" int tip_malloc(void** ptr, size_t bytes)
{
    TIP_REQUIRE(ptr, "tip_malloc: NULL out pointer");
    *ptr = nullptr;
    TIP_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return TIP_OK;
}

extern "C" int tip_free(void* ptr)
{
    if (ptr) TIP_HIP(hipFree(ptr));
    return TIP_OK;
}

extern "C" int tip_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream)
{
    TIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    TIP_HIP(hipStreamSynchronize(as_stream(stream)));
    return TIP_OK;
}

extern "C" int tip_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream)
{
    TIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    TIP_HIP(hipStreamSynchronize(as_stream(stream)));
    return TIP_OK;
}

extern "C" int tip_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream)
{
    TIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return TIP_OK;
}

extern "C" int tip_memset(void* ptr, int value, size_t bytes, void* stream)
{
    TIP_HIP(hipMemsetAsync(ptr, value, bytes, as_stream(stream)));
    return TIP_OK;
}

extern "C" int tip_stream_create(void** stream)
{
    TIP_REQUIRE(stream, "tip_stream_create: NULL");
    hipStream_t s;
    TIP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return TIP_OK;
}

extern "C" int tip_stream_destroy(void* stream)
{
    if (stream) TIP_HIP(hipStreamDestroy(as_stream(stream)));
    return TIP_OK;
}

extern "C" int tip_stream_synchronize(void* stream)
{
    TIP_HIP(hipStreamSynchronize(as_stream(stream)));
    return TIP_OK;
}

extern "C" int tip_device_synchronize(void)
{
    TIP_HIP(hipDeviceSynchronize());
    return TIP_OK;
}

extern "C" uint64_t tip_noise_seed(uint64_t step, uint64_t seed)
{
    return step ^ (seed * 0x9E3779B97F4A7C15ULL);
}

// ------------------------------------------------------------------------------------------
// Hash RNG
// ------------------------------------------------------------------------------------------
__global__ void k_hash_pairs(const uint64_t* __restrict__ seeds, int64_t n, uint64_t* x1o,
                             uint64_t* x2o, double* u1o, double* u2o, double* zo, double* zso)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t x1, x2;
    hash_states(seeds[i], x1, x2);
    double u1 = hash_unit(x1), u2 = hash_unit(x2);
    if (u1 < 1e-12) u1 = 1e-12;
    if (x1o) x1o[i] = x1;
    if (x2o) x2o[i] = x2;
    if (u1o) u1o[i] = u1;
    if (u2o) u2o[i] = u2;
    if (zo || zso) {
        // z: the real field's draw (cos from bm_cos2pi); zs: the pair's second member
        const double r = bm_sqrt(bm_neg2log(u1));
        if (zo) zo[i] = r * bm_cos2pi(u2);
        if (zso) {
            double sn, cs;
            bm_sincos(6.283185307179586 * u2, &sn, &cs);
            zso[i] = r * sn;
        }
    }
}

extern "C" int tip_hash_pairs(const uint64_t* seeds, int64_t n, uint64_t* x1, uint64_t* x2,
                              double* u1, double* u2, double* z, double* zs, void* stream)
{
    TIP_REQUIRE(n >= 0, "tip_hash_pairs: n < 0");
    if (n == 0) return TIP_OK;
    TIP_REQUIRE(seeds, "tip_hash_pairs: NULL seeds");
    int64_t blocks = (n + 255) / 256;
    TIP_REQUIRE(blocks < (int64_t(1) << 31), "tip_hash_pairs: n too large");
    hipLaunchKernelGGL(k_hash_pairs, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream),
                       seeds, n, x1, x2, u1, u2, z, zs);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

template <typename R>
__global__ void k_noise_field(R* __restrict__ o1, R* __restrict__ o2, int64_t n,
                              uint64_t noise_seed, int64_t mult, double scale)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint64_t x1, x2;
    hash_states((uint64_t)(s * mult) ^ noise_seed, x1, x2);
    double u1 = hash_unit(x1), u2 = hash_unit(x2);
    if (u1 < 1e-12) u1 = 1e-12;
    // the fused steps' draws: one (bm_cos2pi) for the real field, the (cos, sin) pair for two
    const double r = bm_sqrt(bm_neg2log(u1));
    if (o2) {
        double sn, cs;
        bm_sincos(6.283185307179586 * u2, &sn, &cs);
        o1[s] = (R)(scale * (r * cs));
        o2[s] = (R)(scale * (r * sn));
    } else {
        o1[s] = (R)(scale * (r * bm_cos2pi(u2)));
    }
}

template <typename R>
static int noise_field(R* o1, R* o2, int64_t nx, int64_t ny, int64_t nz, uint64_t noise_seed,
                       int64_t mult, double scale, void* stream)
{
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(o1, "tip_noise_field: NULL output");
    int64_t n = nx * ny * nz;
    hipLaunchKernelGGL((k_noise_field<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), o1, o2, n, noise_seed, mult, scale);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

extern "C" int tip_noise_field_f64(double* noise1, double* noise2, int64_t nx, int64_t ny,
                                   int64_t nz, uint64_t noise_seed, int64_t mult, double scale,
                                   void* stream)
{
    return noise_field<double>(noise1, noise2, nx, ny, nz, noise_seed, mult, scale, stream);
}

extern "C" int tip_noise_field_f32(float* noise1, float* noise2, int64_t nx, int64_t ny,
                                   int64_t nz, uint64_t noise_seed, int64_t mult, double scale,
                                   void* stream)
{
    return noise_field<float>(noise1, noise2, nx, ny, nz, noise_seed, mult, scale, stream);
}

// ------------------------------------------------------------------------------------------
// Kernel (b): plaquette winding.
//
// Reference: compute_winding_number, latticeSimComplex_numba.py:2683-2744.  For an edge a->b
// the reference adds wrap(theta_b - theta_a) = d - 2 pi m with m = +1 if d > pi, -1 if
// d < -pi; around a plaquette the raw d's telescope to zero, so the plaquette winding is
// -(sum of m).  With theta in (-pi, pi]:
//   m = +1  iff theta_a < 0 < theta_b and the counter-clockwise arc a->b exceeds pi,
//           i.e. cross(a, b) = a1 b2 - a2 b1 < 0;
//   m = -1  iff theta_b < 0 < theta_a and cross(a, b) > 0.
// For float32 inputs the float64 products are exact and the difference is correctly rounded,
// so the sign of cross is exact and decides every edge except the cross == 0 cases below,
// which take the reference's transcendental expression.
// ------------------------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ int half_plane(R x, R y)
{
    // sign of atan2(y, x) as the reference's atan2 returns it, including signed zeros:
    // +-1 for y != 0, and for y = +-0 the sign of y when x is on the negative side (x < 0 or
    // x = -0), else 0; NaN gives 0.  Selects only, no branches.
    const bool left = signbit(x) && !(x != x);
    const bool on = (y != (R)0 || left) && !(y != y);
    return on ? (signbit(y) ? -1 : 1) : 0;
}

// cross(a, b) == 0 with a and b in opposite open half planes: a zero vector whose signed zeros
// give it the angle +-pi, the two sides of the negative real axis (theta = pi and -pi), or an
// exactly antiparallel pair (|d| = pi up to rounding).  The sign of cross cannot decide these;
// evaluate the reference's own expression, m = (d - atan2(sin d, cos d)) / 2 pi.
__device__ __noinline__ int edge_wrap_degenerate(double ax, double ay, double bx, double by)
{
    const double d = atan2(by, bx) - atan2(ay, ax);
    const double w = atan2(sin(d), cos(d));
    return (int)rint((d - w) * 0.15915494309189535);  // 1 / (2 pi)
}

// A plaquette corner: the field in float64 (exact for float32 input) and its half plane h.
struct WCorner {
    double x, y;
    int h;
};

template <typename R>
__device__ __forceinline__ WCorner wcorner(R x, R y)
{
    WCorner c;
    c.x = (double)x;
    c.y = (double)y;
    c.h = half_plane(x, y);
    return c;
}

// Wrap count m of the edge a -> b: +1 for h_a = -1, h_b = +1 and cross < 0; -1 for h_a = +1,
// h_b = -1 and cross > 0; else 0.  With t = h_b - h_a and s = -sign(cross) this is
// m = (t == 2 s) ? s : 0 (t = +-2 exactly when the corners are in opposite half planes).
// `deg` is raised where they are and cross is exactly zero (edge_m_full decides those).
// m(b -> a) = -m(a -> b) in both the sign method and the degenerate expression (the
// difference d changes sign exactly, sin and atan2 are odd, rint is odd), so each edge is
// evaluated once and enters its plaquettes with the orientation's sign.
__device__ __forceinline__ int edge_m(const WCorner& a, const WCorner& b, bool& deg)
{
    const double cr = a.x * b.y - a.y * b.x;
    const int t = b.h - a.h;
    const int s = (int)(cr < 0.0) - (int)(cr > 0.0);
    deg |= (t * t == 4) & (cr == 0.0);
    return t == 2 * s ? s : 0;
}

__device__ __forceinline__ int edge_m_full(const WCorner& a, const WCorner& b)
{
    bool deg = false;
    const int m = edge_m(a, b, deg);
    return deg ? edge_wrap_degenerate(a.x, a.y, b.x, b.y) : m;
}

// One lane per (y, z) column, marching over a chunk of kWindCX x planes.  Plane x's corners
// (j, k), (j+1, k), (j, k+1), (j+1, k+1) are the previous plane's x+1 corners, and its y and
// z edges are the previous plane's x+1 edges, so each plane evaluates seven new edges: three
// along x, two along y and two along z at x+1.  The loads of plane x+2 are issued before
// plane x is evaluated.  The string and plaquette counts are reduced over the block and added
// with one pair of atomics per block.
constexpr int kWindCX = 16;

template <typename R>
__global__ __launch_bounds__(256) void k_winding(const R* __restrict__ p1,
                                                 const R* __restrict__ p2, int8_t* plaq,
                                                 float* wsum, unsigned long long* counts, int nx,
                                                 int ny, int nz)
{
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int i0 = blockIdx.z * kWindCX;
    const int i1 = min(i0 + kWindCX, nx);
    const bool act = k < nz && j < ny;
    unsigned nstr = 0, nplq = 0;
    if (act) {
        const int jp = j + 1 < ny ? j + 1 : 0;
        const int kp = k + 1 < nz ? k + 1 : 0;
        const int64_t n = (int64_t)nx * ny * nz;
        const int64_t nyz = (int64_t)ny * nz;
        const int64_t o00 = (int64_t)j * nz + k, o10 = (int64_t)jp * nz + k;
        const int64_t o01 = (int64_t)j * nz + kp, o11 = (int64_t)jp * nz + kp;
        // corners of plane x: a = (j, k), b = (j+1, k), c = (j, k+1), d = (j+1, k+1)
        int64_t base = (int64_t)i0 * nyz;
        WCorner a = wcorner(p1[base + o00], p2[base + o00]);
        WCorner b = wcorner(p1[base + o10], p2[base + o10]);
        WCorner c = wcorner(p1[base + o01], p2[base + o01]);
        WCorner d = wcorner(p1[base + o11], p2[base + o11]);
        int ey = edge_m_full(a, b), ez = edge_m_full(a, c), ezb = edge_m_full(b, d),
            eyc = edge_m_full(c, d);
        // raw corners of plane x+1
        int64_t bn = (int64_t)(i0 + 1 < nx ? i0 + 1 : 0) * nyz;
        R q[8] = {p1[bn + o00], p2[bn + o00], p1[bn + o10], p2[bn + o10],
                  p1[bn + o01], p2[bn + o01], p1[bn + o11], p2[bn + o11]};
        for (int i = i0; i < i1; ++i) {
            const WCorner A = wcorner(q[0], q[1]), B = wcorner(q[2], q[3]);
            const WCorner Cc = wcorner(q[4], q[5]), D = wcorner(q[6], q[7]);
            if (i + 1 < i1) {
                int i2 = i + 2;
                if (i2 >= nx) i2 -= nx;
                if (i2 >= nx) i2 -= nx;
                const int64_t b2 = (int64_t)i2 * nyz;
                q[0] = p1[b2 + o00];
                q[1] = p2[b2 + o00];
                q[2] = p1[b2 + o10];
                q[3] = p2[b2 + o10];
                q[4] = p1[b2 + o01];
                q[5] = p2[b2 + o01];
                q[6] = p1[b2 + o11];
                q[7] = p2[b2 + o11];
            }
            bool deg = false;
            int exa = edge_m(a, A, deg), exb = edge_m(b, B, deg), exc = edge_m(c, Cc, deg);
            int eyA = edge_m(A, B, deg), ezA = edge_m(A, Cc, deg), ezB = edge_m(B, D, deg);
            int eyC = edge_m(Cc, D, deg);
            if (deg) {  // rare: some edge of this plane needs the transcendental expression
                exa = edge_m_full(a, A);
                exb = edge_m_full(b, B);
                exc = edge_m_full(c, Cc);
                eyA = edge_m_full(A, B);
                ezA = edge_m_full(A, Cc);
                ezB = edge_m_full(B, D);
                eyC = edge_m_full(Cc, D);
            }
            // XY: (i,j,k)->(ip,j,k)->(ip,jp,k)->(i,jp,k); the winding is -(sum of m)
            const int wxy = ey + exb - exa - eyA;
            // XZ: (i,j,k)->(ip,j,k)->(ip,j,kp)->(i,j,kp)
            const int wxz = ez + exc - exa - ezA;
            // YZ: (i,j,k)->(i,jp,k)->(i,jp,kp)->(i,j,kp)
            const int wyz = ez + eyc - ey - ezb;
            const int64_t s = base + o00;
            if (plaq) {
                plaq[s] = (int8_t)wxy;
                plaq[n + s] = (int8_t)wxz;
                plaq[2 * n + s] = (int8_t)wyz;
            }
            const int w = wxy + wxz + wyz;
            if (wsum) wsum[s] = (float)w;
            nstr += w != 0;
            nplq += (wxy != 0) + (wxz != 0) + (wyz != 0);
            a = A;
            b = B;
            c = Cc;
            d = D;
            ey = eyA;
            ez = ezA;
            ezb = ezB;
            eyc = eyC;
            base = (int64_t)(i + 1 < nx ? i + 1 : 0) * nyz;
        }
    }
    if (counts) {
        // wave sums, then the four waves of the block through LDS; every thread gets here
        for (int off = 32; off > 0; off >>= 1) {
            nstr += __shfl_xor(nstr, off);
            nplq += __shfl_xor(nplq, off);
        }
        __shared__ unsigned red[2][4];
        if (threadIdx.x == 0) {
            red[0][threadIdx.y] = nstr;
            red[1][threadIdx.y] = nplq;
        }
        __syncthreads();
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            const unsigned long long a = (unsigned long long)red[0][0] + red[0][1] + red[0][2] +
                                         red[0][3];
            const unsigned long long t = (unsigned long long)red[1][0] + red[1][1] + red[1][2] +
                                         red[1][3];
            if (a) atomicAdd(&counts[0], a);
            if (t) atomicAdd(&counts[1], t);
        }
    }
}

template <typename R>
static int winding(const R* phi1, const R* phi2, int8_t* plaq, float* wsum,
                   unsigned long long* counts, int64_t nx, int64_t ny, int64_t nz, void* stream)
{
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(phi1 && phi2, "tip_winding: NULL field");
    dim3 grid((unsigned)((nz + 63) / 64), (unsigned)((ny + 3) / 4),
              (unsigned)((nx + kWindCX - 1) / kWindCX));
    hipLaunchKernelGGL((k_winding<R>), grid, dim3(64, 4), 0, as_stream(stream), phi1, phi2, plaq,
                       wsum, counts, (int)nx, (int)ny, (int)nz);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

extern "C" int tip_winding_f32(const float* phi1, const float* phi2, int8_t* plaq, float* wsum,
                               unsigned long long* counts, int64_t nx, int64_t ny, int64_t nz,
                               void* stream)
{
    return winding<float>(phi1, phi2, plaq, wsum, counts, nx, ny, nz, stream);
}

extern "C" int tip_winding_f64(const double* phi1, const double* phi2, int8_t* plaq, float* wsum,
                               unsigned long long* counts, int64_t nx, int64_t ny, int64_t nz,
                               void* stream)
{
    return winding<double>(phi1, phi2, plaq, wsum, counts, nx, ny, nz, stream);
}

// ------------------------------------------------------------------------------------------
// Kernel (c): true-vacuum count, and min/max reductions.
//
// The count reads one or two fields once and does almost no arithmetic, so it is bound by the
// bytes in flight: each thread issues kVacU independent 16-byte loads per field before testing
// any of them, and the block's count is added with one atomic (at most kVacMaxBlocks atomics
// per launch on one address).  Fields that are not 16-byte aligned take the same loop with
// element loads.
// ------------------------------------------------------------------------------------------
constexpr int kVacThreads = 1024;
constexpr int kVacU = 4;
constexpr int kVacMaxBlocks = 512;

template <typename R, int L>
struct Pack {
    R e[L];
};

template <typename R, int L>
__device__ __forceinline__ Pack<R, L> load_pack(const R* __restrict__ p, int64_t i)
{
    Pack<R, L> r;
    if constexpr (L == 1) {
        r.e[0] = p[i];
    } else {
        typedef R V __attribute__((ext_vector_type(L)));
        const V v = reinterpret_cast<const V*>(p)[i];
#pragma unroll
        for (int l = 0; l < L; ++l) r.e[l] = v[l];
    }
    return r;
}

template <typename R, bool CPLX>
__device__ __forceinline__ unsigned vac_hit(R u, R v, double esc, double e2)
{
    if (CPLX) {
        const double a = (double)u, b = (double)v;
        return (a * a + b * b > e2) ? 1u : 0u;
    }
    return (fabs((double)u) > esc) ? 1u : 0u;
}

template <typename R, bool CPLX, int L>
__global__ __launch_bounds__(kVacThreads) void k_vacuum(const R* __restrict__ a,
                                                        const R* __restrict__ b, int64_t n,
                                                        double esc,
                                                        unsigned long long* count)
{
    const double e2 = esc * esc;
    const int64_t nv = n / L;
    const int64_t span = (int64_t)gridDim.x * kVacThreads * kVacU;
    unsigned c = 0;  // per thread at most n / (kVacThreads * blocks) + kVacU * L + 1
    for (int64_t s0 = (int64_t)blockIdx.x * kVacThreads * kVacU + threadIdx.x; s0 < nv;
         s0 += span) {
        Pack<R, L> x[kVacU], y[kVacU];
#pragma unroll
        for (int u = 0; u < kVacU; ++u) {
            const int64_t s = s0 + (int64_t)u * kVacThreads;
            x[u] = Pack<R, L>{};
            y[u] = Pack<R, L>{};
            if (s < nv) {
                x[u] = load_pack<R, L>(a, s);
                if (CPLX) y[u] = load_pack<R, L>(b, s);
            }
        }
#pragma unroll
        for (int u = 0; u < kVacU; ++u) {
            if (s0 + (int64_t)u * kVacThreads < nv) {
#pragma unroll
                for (int l = 0; l < L; ++l) c += vac_hit<R, CPLX>(x[u].e[l], y[u].e[l], esc, e2);
            }
        }
    }
    // the n mod L elements past the last full vector
    if (blockIdx.x == 0 && (int64_t)threadIdx.x < n - nv * L) {
        const int64_t s = nv * L + threadIdx.x;
        c += vac_hit<R, CPLX>(a[s], CPLX ? b[s] : (R)0, esc, e2);
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    __shared__ unsigned red[kVacThreads / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kVacThreads / 64; ++w) t += red[w];
        if (t) atomicAdd(count, t);
    }
}

template <typename R, bool CPLX>
static int vacuum_count(const R* a, const R* b, int64_t n, double esc, unsigned long long* count,
                        void* stream)
{
    if (n == 0) return TIP_OK;
    constexpr int L = 16 / sizeof(R);
    const bool vec = ((uintptr_t)a % 16 == 0) && (!CPLX || (uintptr_t)b % 16 == 0);
    const int64_t nv = vec ? n / L : n;
    int64_t nb = (nv + (int64_t)kVacThreads * kVacU - 1) / ((int64_t)kVacThreads * kVacU);
    nb = nb < 1 ? 1 : (nb > kVacMaxBlocks ? kVacMaxBlocks : nb);
    hipStream_t st = as_stream(stream);
    if (vec)
        hipLaunchKernelGGL((k_vacuum<R, CPLX, L>), dim3((unsigned)nb), dim3(kVacThreads), 0, st,
                           a, b, n, esc, count);
    else
        hipLaunchKernelGGL((k_vacuum<R, CPLX, 1>), dim3((unsigned)nb), dim3(kVacThreads), 0, st,
                           a, b, n, esc, count);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

static unsigned reduce_blocks(int64_t n)
{
    int64_t b = (n + 255) / 256;
    if (b > 2048) b = 2048;
    if (b < 1) b = 1;
    return (unsigned)b;
}

extern "C" int tip_vacuum_count_f32(const float* phi, int64_t n, double esc,
                                    unsigned long long* count, void* stream)
{
    TIP_REQUIRE(phi && count && n >= 0, "tip_vacuum_count_f32: bad arguments");
    return vacuum_count<float, false>(phi, nullptr, n, esc, count, stream);
}

extern "C" int tip_vacuum_count_f64(const double* phi, int64_t n, double esc,
                                    unsigned long long* count, void* stream)
{
    TIP_REQUIRE(phi && count && n >= 0, "tip_vacuum_count_f64: bad arguments");
    return vacuum_count<double, false>(phi, nullptr, n, esc, count, stream);
}

extern "C" int tip_vacuum_count_c32(const float* phi1, const float* phi2, int64_t n, double esc,
                                    unsigned long long* count, void* stream)
{
    TIP_REQUIRE(phi1 && phi2 && count && n >= 0, "tip_vacuum_count_c32: bad arguments");
    return vacuum_count<float, true>(phi1, phi2, n, esc, count, stream);
}

// Two-stage min/max: per-block partials into scratch, then one block folds them.
// mode 0: min/max of a; mode 1: max of sqrt(a^2 + b^2) (written to out[0]); mode 2: min/max of
// sqrt(a^2 + b^2).  The partial pass reads like the vacuum count (kVacU independent 16-byte
// loads per field per thread before any is used; element loads for unaligned fields).
template <typename R, int MODE, int L>
__global__ __launch_bounds__(256) void k_minmax_part(const R* __restrict__ a,
                                                     const R* __restrict__ b, int64_t n,
                                                     double* part)
{
    __shared__ double smin[256], smax[256];
    double lo = INFINITY, hi = -INFINITY;
    // NaN propagates to both ends so the caller's finiteness check sees it
    auto take = [&](double v) {
        lo = (v < lo || v != v) ? v : lo;
        hi = (v > hi || v != v) ? v : hi;
    };
    auto value = [&](R u, R w) -> double {
        if (MODE == 0) return (double)u;
        const double x = (double)u, y = (double)w;  // modes 1, 2: rho
        return sqrt(x * x + y * y);
    };
    const int64_t nv = n / L;
    const int64_t span = (int64_t)gridDim.x * blockDim.x * kVacU;
    for (int64_t s0 = (int64_t)blockIdx.x * blockDim.x * kVacU + threadIdx.x; s0 < nv;
         s0 += span) {
        Pack<R, L> x[kVacU], y[kVacU];
#pragma unroll
        for (int u = 0; u < kVacU; ++u) {
            const int64_t s = s0 + (int64_t)u * blockDim.x;
            x[u] = Pack<R, L>{};
            y[u] = Pack<R, L>{};
            if (s < nv) {
                x[u] = load_pack<R, L>(a, s);
                if (MODE != 0) y[u] = load_pack<R, L>(b, s);
            }
        }
#pragma unroll
        for (int u = 0; u < kVacU; ++u) {
            if (s0 + (int64_t)u * blockDim.x < nv) {
#pragma unroll
                for (int l = 0; l < L; ++l) take(value(x[u].e[l], y[u].e[l]));
            }
        }
    }
    if (blockIdx.x == 0 && (int64_t)threadIdx.x < n - nv * L) {
        const int64_t s = nv * L + threadIdx.x;
        take(value(a[s], MODE != 0 ? b[s] : (R)0));
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double l2 = smin[threadIdx.x + w], h2 = smax[threadIdx.x + w];
            double l1 = smin[threadIdx.x], h1 = smax[threadIdx.x];
            smin[threadIdx.x] = (l2 < l1 || l2 != l2) ? l2 : l1;
            smax[threadIdx.x] = (h2 > h1 || h2 != h2) ? h2 : h1;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smin[0];
        part[2 * blockIdx.x + 1] = smax[0];
    }
}

__global__ __launch_bounds__(256) void k_minmax_fold(const double* __restrict__ part, int nparts,
                                                     double* out, int mode)
{
    __shared__ double smin[256], smax[256];
    double lo = INFINITY, hi = -INFINITY;
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) {
        double l = part[2 * p], h = part[2 * p + 1];
        lo = (l < lo || l != l) ? l : lo;
        hi = (h > hi || h != h) ? h : hi;
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double l2 = smin[threadIdx.x + w], h2 = smax[threadIdx.x + w];
            double l1 = smin[threadIdx.x], h1 = smax[threadIdx.x];
            smin[threadIdx.x] = (l2 < l1 || l2 != l2) ? l2 : l1;
            smax[threadIdx.x] = (h2 > h1 || h2 != h2) ? h2 : h1;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (mode != 1) {
            out[0] = smin[0];
            out[1] = smax[0];
        } else {
            out[0] = smax[0];
        }
    }
}

template <typename R, int MODE>
static int minmax(const R* a, const R* b, int64_t n, double* out, void* stream)
{
    TIP_REQUIRE(a && out && n > 0, "tip_minmax: bad arguments");
    constexpr int L = 16 / sizeof(R);
    const bool vec = ((uintptr_t)a % 16 == 0) && (MODE == 0 || (uintptr_t)b % 16 == 0);
    const int64_t nv = vec ? n / L : n;
    int64_t nb64 = (nv + 256 * kVacU - 1) / (256 * kVacU);
    nb64 = nb64 < 1 ? 1 : (nb64 > 1024 ? 1024 : nb64);
    const unsigned nb = (unsigned)nb64;
    hipStream_t st = as_stream(stream);
    ScratchLease lease{1, st};
    double* part = reinterpret_cast<double*>(scratch_get(sizeof(double) * 2 * nb, 1, st));
    TIP_REQUIRE(part, "tip_minmax: scratch allocation failed");
    if (vec)
        hipLaunchKernelGGL((k_minmax_part<R, MODE, L>), dim3(nb), dim3(256), 0, st, a, b, n,
                           part);
    else
        hipLaunchKernelGGL((k_minmax_part<R, MODE, 1>), dim3(nb), dim3(256), 0, st, a, b, n,
                           part);
    TIP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_minmax_fold, dim3(1), dim3(256), 0, st, part, (int)nb, out, MODE);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

extern "C" int tip_minmax_f32(const float* phi, int64_t n, double* out2, void* stream)
{
    return minmax<float, 0>(phi, nullptr, n, out2, stream);
}

extern "C" int tip_minmax_f64(const double* phi, int64_t n, double* out2, void* stream)
{
    return minmax<double, 0>(phi, nullptr, n, out2, stream);
}

// Snapshot fields of the complex driver (latticeSimComplex_numba.py:4625-4626): rho and theta
// formed in float64 from the float32 field and rounded to float32.
__global__ __launch_bounds__(256) void k_rho_theta(const float* __restrict__ a,
                                                   const float* __restrict__ b, int64_t n,
                                                   float* __restrict__ rho,
                                                   float* __restrict__ theta)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x) {
        const double u = (double)a[s], w = (double)b[s];
        if (rho) rho[s] = (float)sqrt(u * u + w * w);  // u*u, w*w exact: one rounding, then sqrt
        if (theta) theta[s] = (float)atan2(w, u);
    }
}

extern "C" int tip_rho_theta_f32(const float* phi1, const float* phi2, int64_t n, float* rho,
                                 float* theta, double* rho_minmax, void* stream)
{
    TIP_REQUIRE(phi1 && phi2 && n > 0, "tip_rho_theta_f32: bad arguments");
    hipStream_t st = as_stream(stream);
    if (rho || theta) {
        hipLaunchKernelGGL(k_rho_theta, dim3(reduce_blocks(n)), dim3(256), 0, st, phi1, phi2, n,
                           rho, theta);
        TIP_HIP(hipGetLastError());
    }
    if (rho_minmax) return minmax<float, 2>(phi1, phi2, n, rho_minmax, stream);
    return TIP_OK;
}

extern "C" int tip_rho_max_c32(const float* phi1, const float* phi2, int64_t n, double* out1,
                               void* stream)
{
    TIP_REQUIRE(phi2, "tip_rho_max_c32: NULL phi2");
    return minmax<float, 1>(phi1, phi2, n, out1, stream);
}
