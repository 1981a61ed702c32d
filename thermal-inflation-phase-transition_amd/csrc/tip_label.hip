// tip_label.hip -- connected-component labelling of a thresholded field and per-component
// statistics: the bubble (|phi| > esc) and string (|W| > 0.5) analysis of the post-processing
// tools (postprocess/revisualize_snapshots.py:171-213 _identify_bubbles, :360-392
// _identify_strings), which call scipy.ndimage.label with its default structuring element
// (6-connectivity, no wrap-around) and number components 1..n in raster order of their first
// site.
//
// Labelling is union-find in HBM (after Playne and Hawick's GPU union-find): every foreground
// site starts as its own parent; one pass merges each site with its -x, -y and -z foreground
// neighbours by atomicMin on the roots, so a component's root is its smallest site index (its
// first site in raster order); a compression pass points every site at its root; a scan of
// the root flags numbers the roots 1..n in raster order, which is scipy's numbering.
//
// Statistics: one thread per (x, y) row walks z and accumulates runs of equal label in
// registers; each run is added to a per-block table in LDS, and the table is added to the
// global sums once per block (a percolating cluster costs one set of global atomics per
// block, not one per run).  Coordinates are summed as exact 64-bit integers (sum x, sum
// x^2 over the three axes); the value sum is float64.
#include <hip/hip_runtime.h>

#include "tip_internal.h"

using namespace tip;

namespace {

// Parent reads bypass the non-coherent vector L1: a stale parent is still an ancestor, but a
// coherent read shortens the chains other waves are rewriting.
__device__ __forceinline__ int32_t parent(const int32_t* L, int32_t x)
{
    return __hip_atomic_load(L + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int32_t find_root(const int32_t* L, int32_t x)
{
    int32_t p = parent(L, x);
    while (p != x) {
        x = p;
        p = parent(L, x);
    }
    return x;
}

// Union of the trees holding a and b: the larger root is pointed at the smaller one with an
// atomicMin; a failed attempt (the root was re-parented meanwhile) retries from the value the
// atomic returned.  Indices only decrease, so the loop ends.
__device__ void unite(int32_t* L, int32_t a, int32_t b)
{
    for (;;) {
        a = find_root(L, a);
        b = find_root(L, b);
        if (a == b) return;
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        const int32_t old = atomicMin(L + b, a);
        if (old == b) return;
        b = old;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_ccl_init(const R* __restrict__ v, double thr, int64_t n,
                                                  int32_t* __restrict__ L)
{
    // the comparison runs in the field's precision, as numpy compares a float32 array with
    // a Python float
    const R t = (R)thr;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x) {
        const R a = v[s] < (R)0 ? -v[s] : v[s];
        L[s] = a > t ? (int32_t)s : -1;
    }
}

__global__ __launch_bounds__(256) void k_ccl_merge(int32_t* L, int nx, int ny, int nz,
                                                   int periodic)
{
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x) {
        if (L[s] < 0) continue;
        const int k = (int)(s % nz);
        const int64_t r = s / nz;
        const int j = (int)(r % ny);
        const int i = (int)(r / ny);
        const int64_t plane = (int64_t)ny * nz;
        if (k > 0) {
            if (L[s - 1] >= 0) unite(L, (int32_t)s, (int32_t)(s - 1));
        } else if (periodic && nz > 1 && L[s + nz - 1] >= 0) {
            unite(L, (int32_t)s, (int32_t)(s + nz - 1));
        }
        if (j > 0) {
            if (L[s - nz] >= 0) unite(L, (int32_t)s, (int32_t)(s - nz));
        } else if (periodic && ny > 1 && L[s + (int64_t)(ny - 1) * nz] >= 0) {
            unite(L, (int32_t)s, (int32_t)(s + (int64_t)(ny - 1) * nz));
        }
        if (i > 0) {
            if (L[s - plane] >= 0) unite(L, (int32_t)s, (int32_t)(s - plane));
        } else if (periodic && nx > 1 && L[s + (int64_t)(nx - 1) * plane] >= 0) {
            unite(L, (int32_t)s, (int32_t)(s + (int64_t)(nx - 1) * plane));
        }
    }
}

__global__ __launch_bounds__(256) void k_ccl_compress(int32_t* L, int64_t n)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x) {
        if (L[s] >= 0) L[s] = find_root(L, (int32_t)s);
    }
}

// Roots per tile of kTile sites (a root is a site that is its own parent).
constexpr int kTile = 1024;

__global__ __launch_bounds__(256) void k_ccl_count_roots(const int32_t* __restrict__ L, int64_t n,
                                                         int32_t* __restrict__ tile_count)
{
    __shared__ int32_t red[256];
    const int64_t base = (int64_t)blockIdx.x * kTile;
    int c = 0;
    for (int e = 0; e < kTile / 256; ++e) {
        const int64_t s = base + e * 256 + threadIdx.x;
        if (s < n && L[s] == (int32_t)s) ++c;
    }
    red[threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_count[blockIdx.x] = red[0];
}

// Exclusive scan of the tile counts in place (one block walks them in chunks of 1024) and the
// total into *total.
__global__ __launch_bounds__(1024) void k_ccl_scan_tiles(int32_t* tile_count, int ntiles,
                                                         int64_t* total)
{
    __shared__ int32_t buf[1024];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < ntiles; base += 1024) {
        const int t = base + (int)threadIdx.x;
        const int32_t x = t < ntiles ? tile_count[t] : 0;
        buf[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
            const int32_t y = (int)threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += y;
            __syncthreads();
        }
        if (t < ntiles) tile_count[t] = carry + buf[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// ids[root] = 1 + rank of the root in raster order.
__global__ __launch_bounds__(256) void k_ccl_rank_roots(const int32_t* __restrict__ L, int64_t n,
                                                        const int32_t* __restrict__ tile_off,
                                                        int32_t* __restrict__ ids)
{
    __shared__ int32_t sc[256];
    const int64_t s0 = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * (kTile / 256);
    bool root[kTile / 256];
    int c = 0;
    for (int e = 0; e < kTile / 256; ++e) {
        const int64_t s = s0 + e;
        root[e] = s < n && L[s] == (int32_t)s;
        c += root[e];
    }
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int32_t y = (int)threadIdx.x >= off ? sc[threadIdx.x - off] : 0;
        __syncthreads();
        sc[threadIdx.x] += y;
        __syncthreads();
    }
    int32_t rank = tile_off[blockIdx.x] + sc[threadIdx.x] - c;
    for (int e = 0; e < kTile / 256; ++e) {
        if (root[e]) ids[s0 + e] = ++rank;
    }
}

__global__ __launch_bounds__(256) void k_ccl_finalize(int32_t* L, const int32_t* __restrict__ ids,
                                                      int64_t n)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int32_t r = L[s];
        L[s] = r >= 0 ? ids[r] : 0;
    }
}

struct CclStatsOut {
    unsigned long long* volume;  // [n]
    unsigned long long* sx;      // [3n]: sum of x, y, z
    unsigned long long* sxx;     // [n]: sum of x^2 + y^2 + z^2
    double* sum_v;               // [n]
    unsigned long long* peak;    // [n]: max |v| as float64 bits (non-negative doubles order as
                                 // their bit patterns)
};

__device__ __forceinline__ void flush_global(const CclStatsOut& o, int64_t c,
                                             unsigned long long cnt, unsigned long long sx,
                                             unsigned long long sy, unsigned long long sz,
                                             unsigned long long sq, double sv,
                                             unsigned long long pk)
{
    atomicAdd(o.volume + c, cnt);
    atomicAdd(o.sx + 3 * c, sx);
    atomicAdd(o.sx + 3 * c + 1, sy);
    atomicAdd(o.sx + 3 * c + 2, sz);
    atomicAdd(o.sxx + c, sq);
    atomicAdd(o.sum_v + c, sv);
    atomicMax(o.peak + c, pk);
}

// Per-block accumulators for the labels met by the block's rows: an open-addressing table in
// LDS (key 0 = empty), so a large component costs one set of global atomics per block instead
// of one per run.  A run whose label finds no slot within kProbe probes goes to global memory
// directly.
constexpr int kSlots = 512;
constexpr int kProbe = 32;

struct StatsTable {
    int32_t key[kSlots];
    unsigned long long cnt[kSlots], sx[kSlots], sy[kSlots], sz[kSlots], sq[kSlots], pk[kSlots];
    double sv[kSlots];
};

template <typename R>
__global__ __launch_bounds__(256) void k_ccl_stats(const R* __restrict__ v,
                                                   const int32_t* __restrict__ L, int nx, int ny,
                                                   int nz, int64_t nlab, CclStatsOut o)
{
    __shared__ StatsTable tb;
    for (int e = threadIdx.x; e < kSlots; e += blockDim.x) {
        tb.key[e] = 0;
        tb.cnt[e] = tb.sx[e] = tb.sy[e] = tb.sz[e] = tb.sq[e] = tb.pk[e] = 0;
        tb.sv[e] = 0.0;
    }
    __syncthreads();
    auto flush = [&](int32_t lab, unsigned long long cnt, unsigned long long sx,
                     unsigned long long sy, unsigned long long sz, unsigned long long sq,
                     double sv, unsigned long long pk) {
        if (lab <= 0 || lab > nlab) return;
        unsigned h = ((unsigned)lab * 2654435761u) & (kSlots - 1);
        for (int p = 0; p < kProbe; ++p) {
            const int32_t old = atomicCAS(&tb.key[h], 0, lab);
            if (old == 0 || old == lab) {
                atomicAdd(&tb.cnt[h], cnt);
                atomicAdd(&tb.sx[h], sx);
                atomicAdd(&tb.sy[h], sy);
                atomicAdd(&tb.sz[h], sz);
                atomicAdd(&tb.sq[h], sq);
                atomicAdd(&tb.sv[h], sv);
                atomicMax(&tb.pk[h], pk);
                return;
            }
            h = (h + 1) & (kSlots - 1);
        }
        flush_global(o, lab - 1, cnt, sx, sy, sz, sq, sv, pk);
    };
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row < (int64_t)nx * ny) {
        const unsigned long long x = (unsigned long long)(row / ny);
        const unsigned long long y = (unsigned long long)(row % ny);
        const int64_t base = row * nz;
        int32_t cur = 0;
        unsigned long long cnt = 0, sz = 0, sq = 0;
        double sv = 0.0, pk = 0.0;
        for (int k = 0; k < nz; ++k) {
            const int32_t lab = L[base + k];
            if (lab != cur) {
                if (cnt)
                    flush(cur, cnt, cnt * x, cnt * y, sz, sq, sv,
                          (unsigned long long)__double_as_longlong(pk));
                cur = lab;
                cnt = sz = sq = 0;
                sv = 0.0;
                pk = 0.0;
            }
            if (lab > 0) {
                const unsigned long long z = (unsigned long long)k;
                const double val = (double)v[base + k];
                ++cnt;
                sz += z;
                sq += x * x + y * y + z * z;
                sv += val;
                pk = fmax(pk, fabs(val));
            }
        }
        if (cnt)
            flush(cur, cnt, cnt * x, cnt * y, sz, sq, sv,
                  (unsigned long long)__double_as_longlong(pk));
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kSlots; e += blockDim.x) {
        const int32_t lab = tb.key[e];
        if (lab > 0)
            flush_global(o, lab - 1, tb.cnt[e], tb.sx[e], tb.sy[e], tb.sz[e], tb.sq[e], tb.sv[e],
                         tb.pk[e]);
    }
}

unsigned grid_for(int64_t n)
{
    int64_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b < 1 ? 1 : b) : 8192);
}

template <typename R>
int label(const R* v, double thr, int64_t nx, int64_t ny, int64_t nz, int periodic,
          int32_t* labels, int64_t* n_components, void* stream)
{
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(v && labels && n_components, "tip_label: NULL argument");
    TIP_REQUIRE(thr >= 0.0, "tip_label: threshold must be >= 0");
    hipStream_t st = as_stream(stream);
    const int64_t n = nx * ny * nz;
    const int64_t ntiles = (n + kTile - 1) / kTile;
    int32_t* ids = reinterpret_cast<int32_t*>(scratch_get((size_t)n * sizeof(int32_t), 3));
    char* aux = reinterpret_cast<char*>(scratch_get((size_t)ntiles * sizeof(int32_t) + 64, 4));
    TIP_REQUIRE(ids && aux, "tip_label: scratch allocation failed");
    int64_t* total = reinterpret_cast<int64_t*>(aux);
    int32_t* tiles = reinterpret_cast<int32_t*>(aux + 64);
    const unsigned g = grid_for(n);
    hipLaunchKernelGGL((k_ccl_init<R>), dim3(g), dim3(256), 0, st, v, thr, n, labels);
    hipLaunchKernelGGL(k_ccl_merge, dim3(g), dim3(256), 0, st, labels, (int)nx, (int)ny, (int)nz,
                       periodic ? 1 : 0);
    hipLaunchKernelGGL(k_ccl_compress, dim3(g), dim3(256), 0, st, labels, n);
    hipLaunchKernelGGL(k_ccl_count_roots, dim3((unsigned)ntiles), dim3(256), 0, st, labels, n,
                       tiles);
    hipLaunchKernelGGL(k_ccl_scan_tiles, dim3(1), dim3(1024), 0, st, tiles, (int)ntiles, total);
    hipLaunchKernelGGL(k_ccl_rank_roots, dim3((unsigned)ntiles), dim3(256), 0, st, labels, n,
                       tiles, ids);
    hipLaunchKernelGGL(k_ccl_finalize, dim3(g), dim3(256), 0, st, labels, ids, n);
    TIP_HIP(hipGetLastError());
    TIP_HIP(hipMemcpyAsync(n_components, total, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    TIP_HIP(hipStreamSynchronize(st));
    return TIP_OK;
}

template <typename R>
int label_stats(const R* v, const int32_t* labels, int64_t n_components, int64_t nx, int64_t ny,
                int64_t nz, unsigned long long* volume, unsigned long long* sum_xyz,
                unsigned long long* sum_sq, double* sum_v, unsigned long long* peak_bits,
                void* stream)
{
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(v && labels, "tip_label_stats: NULL field or labels");
    TIP_REQUIRE(n_components >= 0, "tip_label_stats: n_components < 0");
    if (n_components == 0) return TIP_OK;
    TIP_REQUIRE(volume && sum_xyz && sum_sq && sum_v && peak_bits,
                "tip_label_stats: NULL output");
    hipStream_t st = as_stream(stream);
    const size_t nc = (size_t)n_components;
    TIP_HIP(hipMemsetAsync(volume, 0, nc * sizeof(unsigned long long), st));
    TIP_HIP(hipMemsetAsync(sum_xyz, 0, 3 * nc * sizeof(unsigned long long), st));
    TIP_HIP(hipMemsetAsync(sum_sq, 0, nc * sizeof(unsigned long long), st));
    TIP_HIP(hipMemsetAsync(sum_v, 0, nc * sizeof(double), st));
    TIP_HIP(hipMemsetAsync(peak_bits, 0, nc * sizeof(unsigned long long), st));
    CclStatsOut o{volume, sum_xyz, sum_sq, sum_v, peak_bits};
    const int64_t rows = nx * ny;
    hipLaunchKernelGGL((k_ccl_stats<R>), dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st,
                       v, labels, (int)nx, (int)ny, (int)nz, n_components, o);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

}  // namespace

extern "C" int tip_label_f32(const float* v, double threshold, int64_t nx, int64_t ny, int64_t nz,
                             int periodic, int32_t* labels, int64_t* n_components, void* stream)
{
    return label<float>(v, threshold, nx, ny, nz, periodic, labels, n_components, stream);
}

extern "C" int tip_label_f64(const double* v, double threshold, int64_t nx, int64_t ny,
                             int64_t nz, int periodic, int32_t* labels, int64_t* n_components,
                             void* stream)
{
    return label<double>(v, threshold, nx, ny, nz, periodic, labels, n_components, stream);
}

extern "C" int tip_label_stats_f32(const float* v, const int32_t* labels, int64_t n_components,
                                   int64_t nx, int64_t ny, int64_t nz, unsigned long long* volume,
                                   unsigned long long* sum_xyz, unsigned long long* sum_sq,
                                   double* sum_v, unsigned long long* peak_bits, void* stream)
{
    return label_stats<float>(v, labels, n_components, nx, ny, nz, volume, sum_xyz, sum_sq,
                              sum_v, peak_bits, stream);
}

extern "C" int tip_label_stats_f64(const double* v, const int32_t* labels, int64_t n_components,
                                   int64_t nx, int64_t ny, int64_t nz, unsigned long long* volume,
                                   unsigned long long* sum_xyz, unsigned long long* sum_sq,
                                   double* sum_v, unsigned long long* peak_bits, void* stream)
{
    return label_stats<double>(v, labels, n_components, nx, ny, nz, volume, sum_xyz, sum_sq,
                               sum_v, peak_bits, stream);
}
