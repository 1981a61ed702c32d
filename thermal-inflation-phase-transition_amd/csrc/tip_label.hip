// tip_label.hip -- connected-component labelling of a thresholded field and per-component
// statistics: the bubble (|phi| > esc) and string (|W| > 0.5) analysis of the post-processing
// tools (postprocess/revisualize_snapshots.py:171-213 _identify_bubbles, :360-392
// _identify_strings), which call scipy.ndimage.label with its default structuring element
// (6-connectivity, no wrap-around) and number components 1..n in raster order of their first
// site.
//
// Labelling is union-find (after Playne and Hawick's GPU union-find): each block first joins
// the sites of one 8 x 8 x 16 brick in LDS, then a global pass joins the pieces across brick
// faces by atomicMin on the roots in HBM, so a component's root is its smallest site index (its
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

// Local phase: a block labels one brick of kBX x kBY x kBZ sites with union-find in LDS (the
// brick's own -x, -y, -z edges), then writes every foreground site's brick-local root as a
// global index.  The local order (x, y, z within the brick) is monotone in the global site
// index, so a local root -- the smallest local index of its piece -- is also the piece's
// smallest global index; the global merge then only has to join pieces across brick faces.
constexpr int kBX = 8, kBY = 8, kBZ = 16, kBrick = kBX * kBY * kBZ;

__device__ __forceinline__ int lds_parent(const int* lab, int x)
{
    return __hip_atomic_load(lab + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int lds_find(const int* lab, int x)
{
    int p = lds_parent(lab, x);
    while (p != x) {
        x = p;
        p = lds_parent(lab, x);
    }
    return x;
}

__device__ void lds_unite(int* lab, int a, int b)
{
    for (;;) {
        a = lds_find(lab, a);
        b = lds_find(lab, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicMin(lab + b, a);
        if (old == b) return;
        b = old;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_ccl_local(const R* __restrict__ v, double thr,
                                                   int32_t* __restrict__ L, int nx, int ny,
                                                   int nz)
{
    __shared__ int lab[kBrick];
    const int nbz = (nz + kBZ - 1) / kBZ, nby = (ny + kBY - 1) / kBY;
    const int b = (int)blockIdx.x;
    const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
    const int x0 = bx * kBX, y0 = by * kBY, z0 = bz * kBZ;
    const R t = (R)thr;
    auto gidx = [&](int li) {
        const int lz = li % kBZ, ly = (li / kBZ) % kBY, lx = li / (kBZ * kBY);
        return ((int64_t)(x0 + lx) * ny + (y0 + ly)) * nz + (z0 + lz);
    };
    auto inside = [&](int li) {
        const int lz = li % kBZ, ly = (li / kBZ) % kBY, lx = li / (kBZ * kBY);
        return x0 + lx < nx && y0 + ly < ny && z0 + lz < nz;
    };
    for (int li = threadIdx.x; li < kBrick; li += blockDim.x) {
        int val = -1;
        if (inside(li)) {
            const R u = v[gidx(li)];
            const R a = u < (R)0 ? -u : u;
            if (a > t) val = li;
        }
        lab[li] = val;
    }
    __syncthreads();
    for (int li = threadIdx.x; li < kBrick; li += blockDim.x) {
        if (lab[li] < 0) continue;
        const int lz = li % kBZ, ly = (li / kBZ) % kBY, lx = li / (kBZ * kBY);
        if (lz > 0 && lab[li - 1] >= 0) lds_unite(lab, li, li - 1);
        if (ly > 0 && lab[li - kBZ] >= 0) lds_unite(lab, li, li - kBZ);
        if (lx > 0 && lab[li - kBZ * kBY] >= 0) lds_unite(lab, li, li - kBZ * kBY);
    }
    __syncthreads();
    for (int li = threadIdx.x; li < kBrick; li += blockDim.x) {
        if (!inside(li)) continue;
        const int r = lab[li] >= 0 ? lds_find(lab, li) : -1;
        L[gidx(li)] = r >= 0 ? (int32_t)gidx(r) : -1;
    }
}

// Global phase: join the bricks' pieces across brick faces (and across the lattice faces when
// periodic); edges inside a brick were joined in LDS.
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
            if (k % kBZ == 0 && L[s - 1] >= 0) unite(L, (int32_t)s, (int32_t)(s - 1));
        } else if (periodic && nz > 1 && L[s + nz - 1] >= 0) {
            unite(L, (int32_t)s, (int32_t)(s + nz - 1));
        }
        if (j > 0) {
            if (j % kBY == 0 && L[s - nz] >= 0) unite(L, (int32_t)s, (int32_t)(s - nz));
        } else if (periodic && ny > 1 && L[s + (int64_t)(ny - 1) * nz] >= 0) {
            unite(L, (int32_t)s, (int32_t)(s + (int64_t)(ny - 1) * nz));
        }
        if (i > 0) {
            if (i % kBX == 0 && L[s - plane] >= 0) unite(L, (int32_t)s, (int32_t)(s - plane));
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
    ScratchLease lease_ids{3, st}, lease_aux{4, st};
    int32_t* ids = reinterpret_cast<int32_t*>(scratch_get((size_t)n * sizeof(int32_t), 3, st));
    char* aux =
        reinterpret_cast<char*>(scratch_get((size_t)ntiles * sizeof(int32_t) + 64, 4, st));
    TIP_REQUIRE(ids && aux, "tip_label: scratch allocation failed");
    int64_t* total = reinterpret_cast<int64_t*>(aux);
    int32_t* tiles = reinterpret_cast<int32_t*>(aux + 64);
    const unsigned g = grid_for(n);
    const int64_t nbricks = ((nx + kBX - 1) / kBX) * ((ny + kBY - 1) / kBY) *
                            ((nz + kBZ - 1) / kBZ);
    TIP_REQUIRE(nbricks < (int64_t(1) << 31), "tip_label: too many bricks");
    hipLaunchKernelGGL((k_ccl_local<R>), dim3((unsigned)nbricks), dim3(256), 0, st, v, thr,
                       labels, (int)nx, (int)ny, (int)nz);
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
