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
// no kernel still reads the old one.
namespace {
std::mutex g_scratch_mu;
struct Scratch {
    void* p = nullptr;
    size_t bytes = 0;
};
std::vector<std::vector<Scratch>> g_scratch;
}  // namespace

void* scratch_get(size_t bytes, int slot)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
    auto& v = g_scratch[dev];
    if ((int)v.size() <= slot) v.resize(slot + 1);
    Scratch& s = v[slot];
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
        double r = sqrt(-2.0 * log(u1));
        double a = 6.283185307179586 * u2;
        if (zo) zo[i] = r * cos(a);
        if (zso) zso[i] = r * sin(a);
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
    double r = sqrt(-2.0 * log(u1));
    double a = 6.283185307179586 * u2;
    o1[s] = (R)(scale * (r * cos(a)));
    if (o2) o2[s] = (R)(scale * (r * sin(a)));
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
    // sign of atan2(y, x) as the reference's atan2 returns it, including signed zeros.
    if (y > (R)0) return 1;
    if (y < (R)0) return -1;
    if (y == (R)0) {
        bool left = (x < (R)0) || (x == (R)0 && signbit(x));
        if (!left) return 0;
        return signbit(y) ? -1 : 1;
    }
    return 0;  // NaN
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

template <typename R>
__device__ __forceinline__ int edge_wrap(R ax, R ay, int ha, R bx, R by, int hb)
{
    if (ha == hb || ha == 0 || hb == 0) return 0;
    double cr = (double)ax * (double)by - (double)ay * (double)bx;
    if (cr == 0.0) return edge_wrap_degenerate((double)ax, (double)ay, (double)bx, (double)by);
    if (ha < 0) return cr < 0.0 ? 1 : 0;
    return cr > 0.0 ? -1 : 0;
}

template <typename R>
__device__ __forceinline__ int plaquette(R x0, R y0, int h0, R x1, R y1, int h1, R x2, R y2, int h2,
                                         R x3, R y3, int h3)
{
    int m = edge_wrap(x0, y0, h0, x1, y1, h1) + edge_wrap(x1, y1, h1, x2, y2, h2) +
            edge_wrap(x2, y2, h2, x3, y3, h3) + edge_wrap(x3, y3, h3, x0, y0, h0);
    return -m;
}

// One lane per (y, z) column, marching over a chunk of kWindCX x planes: the four corners
// (j, k), (j+1, k), (j, k+1), (j+1, k+1) of plane x+1 loaded for one plane are plane x's
// corners for the next, so each plane costs four new (phi1, phi2) pairs.  The string and
// plaquette counts are reduced over the block and added with one pair of atomics per block.
constexpr int kWindCX = 16;

template <typename R>
struct Corner {
    R x, y;
    int h;
};

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
    unsigned long long nstr = 0, nplq = 0;
    if (act) {
        const int jp = j + 1 < ny ? j + 1 : 0;
        const int kp = k + 1 < nz ? k + 1 : 0;
        const int64_t n = (int64_t)nx * ny * nz;
        const int64_t nyz = (int64_t)ny * nz;
        const int64_t o00 = (int64_t)j * nz + k, o10 = (int64_t)jp * nz + k;
        const int64_t o01 = (int64_t)j * nz + kp, o11 = (int64_t)jp * nz + kp;
        auto corner = [&](int64_t s) {
            Corner<R> c;
            c.x = p1[s];
            c.y = p2[s];
            c.h = half_plane(c.x, c.y);
            return c;
        };
        // plane i corners: a = (j, k), b = (j+1, k), c = (j, k+1), d = (j+1, k+1)
        int64_t base = (int64_t)i0 * nyz;
        Corner<R> a = corner(base + o00), b = corner(base + o10), c = corner(base + o01),
                  d = corner(base + o11);
        for (int i = i0; i < i1; ++i) {
            const int ip = i + 1 < nx ? i + 1 : 0;
            const int64_t bp = (int64_t)ip * nyz;
            const Corner<R> A = corner(bp + o00), B = corner(bp + o10), Cc = corner(bp + o01),
                            D = corner(bp + o11);
            // XY: (i,j,k)->(ip,j,k)->(ip,jp,k)->(i,jp,k)
            const int wxy = plaquette(a.x, a.y, a.h, A.x, A.y, A.h, B.x, B.y, B.h, b.x, b.y, b.h);
            // XZ: (i,j,k)->(ip,j,k)->(ip,j,kp)->(i,j,kp)
            const int wxz =
                plaquette(a.x, a.y, a.h, A.x, A.y, A.h, Cc.x, Cc.y, Cc.h, c.x, c.y, c.h);
            // YZ: (i,j,k)->(i,jp,k)->(i,jp,kp)->(i,j,kp)
            const int wyz = plaquette(a.x, a.y, a.h, b.x, b.y, b.h, d.x, d.y, d.h, c.x, c.y, c.h);
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
            base = bp;
        }
    }
    if (counts) {
        // wave sums, then the four waves of the block through LDS; every thread gets here
        for (int off = 32; off > 0; off >>= 1) {
            nstr += __shfl_xor(nstr, off);
            nplq += __shfl_xor(nplq, off);
        }
        __shared__ unsigned long long red[2][4];
        if (threadIdx.x == 0) {
            red[0][threadIdx.y] = nstr;
            red[1][threadIdx.y] = nplq;
        }
        __syncthreads();
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            const unsigned long long a = red[0][0] + red[0][1] + red[0][2] + red[0][3];
            const unsigned long long t = red[1][0] + red[1][1] + red[1][2] + red[1][3];
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
// ------------------------------------------------------------------------------------------
template <typename R, bool CPLX>
__global__ __launch_bounds__(256) void k_vacuum(const R* __restrict__ a, const R* __restrict__ b,
                                                int64_t n, double esc,
                                                unsigned long long* count)
{
    unsigned long long c = 0;
    const double e2 = esc * esc;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x) {
        if (CPLX) {
            double u = (double)a[s], v = (double)b[s];
            c += (u * u + v * v > e2) ? 1 : 0;
        } else {
            c += (fabs((double)a[s]) > esc) ? 1 : 0;
        }
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
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
    if (n == 0) return TIP_OK;
    hipLaunchKernelGGL((k_vacuum<float, false>), dim3(reduce_blocks(n)), dim3(256), 0,
                       as_stream(stream), phi, nullptr, n, esc, count);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

extern "C" int tip_vacuum_count_f64(const double* phi, int64_t n, double esc,
                                    unsigned long long* count, void* stream)
{
    TIP_REQUIRE(phi && count && n >= 0, "tip_vacuum_count_f64: bad arguments");
    if (n == 0) return TIP_OK;
    hipLaunchKernelGGL((k_vacuum<double, false>), dim3(reduce_blocks(n)), dim3(256), 0,
                       as_stream(stream), phi, nullptr, n, esc, count);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

extern "C" int tip_vacuum_count_c32(const float* phi1, const float* phi2, int64_t n, double esc,
                                    unsigned long long* count, void* stream)
{
    TIP_REQUIRE(phi1 && phi2 && count && n >= 0, "tip_vacuum_count_c32: bad arguments");
    if (n == 0) return TIP_OK;
    hipLaunchKernelGGL((k_vacuum<float, true>), dim3(reduce_blocks(n)), dim3(256), 0,
                       as_stream(stream), phi1, phi2, n, esc, count);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

// Two-stage min/max: per-block partials into scratch, then one block folds them.
// mode 0: min/max of a; mode 1: max of sqrt(a^2 + b^2) (written to out[0]); mode 2: min/max of
// sqrt(a^2 + b^2).
template <typename R, int MODE>
__global__ __launch_bounds__(256) void k_minmax_part(const R* __restrict__ a,
                                                     const R* __restrict__ b, int64_t n,
                                                     double* part)
{
    __shared__ double smin[256], smax[256];
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x) {
        double v;
        if (MODE == 0) v = (double)a[s];
        else {  // modes 1, 2: rho
            double u = (double)a[s], w = (double)b[s];
            v = sqrt(u * u + w * w);
        }
        // NaN propagates to both ends so the caller's finiteness check sees it
        lo = (v < lo || v != v) ? v : lo;
        hi = (v > hi || v != v) ? v : hi;
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
    unsigned nb = reduce_blocks(n);
    if (nb > 1024) nb = 1024;
    double* part = reinterpret_cast<double*>(scratch_get(sizeof(double) * 2 * nb, 1));
    TIP_REQUIRE(part, "tip_minmax: scratch allocation failed");
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL((k_minmax_part<R, MODE>), dim3(nb), dim3(256), 0, st, a, b, n, part);
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
