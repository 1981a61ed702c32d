// tip_ensemble.hip -- many independent real-field realisations (uneven lattice sizes, seeds,
// temperatures and V' tables) advanced together.
//
// The reference runs one realisation per process (it has no ensemble layer); the
// nucleation-rate and string statistics need ensembles.  Here every instance's float32 fields
// live in one HBM pool, and one kernel launch per RK2 step covers all instances: the host
// builds a work list of (instance, plane tile, x chunk) items and each block runs the same
// tile body as the single-lattice kernel on its item.  Instance parameters, table pointers and
// buffers sit in a device descriptor array, and the noise step counter is a device word, so a
// K-step loop (K launches + one counter bump) is captured once into a hipGraph and replayed;
// parameter updates rewrite the descriptors and leave the graph valid.
#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

#include "step_kernels.h"

using namespace tip;

namespace {

struct EnsInst {
    float* f[2][2];  // [parity][phi, pi]
    const float2* ab;
    StepScalars<float> c;
    float tmin, dinv, off;
    int last, odd;
    int nx, ny, nz;
    int W;  // waves per z row in the row-tile kernel (1, 2 or 4)
    uint64_t seed_mix;
    int hash;
};

// Work item: instance, plane tile (64*BY consecutive (y, z) sites), first x plane; `cx`
// planes per item.
struct EnsItem {
    int inst, tp, x0, pad;
};

template <int BY, bool ODD>
__global__ __launch_bounds__(64 * BY) void k_rk2_ensemble(
    const EnsInst* __restrict__ insts, const EnsItem* __restrict__ items, int nitems, int cx,
    int parity, const uint64_t* __restrict__ step_base, int step_off, int xcd_remap,
    unsigned long long* __restrict__ vac, float esc)
{
    int b = (int)blockIdx.x;
    const int nb = (int)gridDim.x;
    if (xcd_remap) b = (b & 7) * (nb >> 3) + (b >> 3);
    // persistent blocks: block b owns the contiguous item range [b*q, min((b+1)*q, nitems))
    const int q = (nitems + nb - 1) / nb;
    const int i0 = b * q;
    const int i1 = min(i0 + q, nitems);
    const uint64_t step = step_base[0] + (uint64_t)step_off;
    for (int i = i0; i < i1; ++i) {
        const EnsItem it = items[i];
        const EnsInst& I = insts[it.inst];
        TableF32<ODD> tab;
        tab.ab = I.ab;
        tab.tmin = I.tmin;
        tab.dinv = I.dinv;
        tab.last = I.last;
        tab.off = I.off;
        const int x1 = min(it.x0 + cx, I.nx);
        // Instances without noise have ns = 0: fma(0, g, pi) = pi, so the hash path gives the
        // noise-free result bit for bit.
        unsigned cnt = rk2_real_tile_f32<ODD, BY, NOISE_HASH>(
            I.f[parity][0], I.f[parity][1], I.f[parity ^ 1][0], I.f[parity ^ 1][1], nullptr, I.c,
            tab, I.nx, I.ny, I.nz, it.tp, it.x0, x1, step ^ I.seed_mix, vac != nullptr, esc);
        if (vac && cnt && (threadIdx.x & 63) == 0) atomicAdd(&vac[it.inst], (unsigned long long)cnt);
    }
}

// Row-tile variant (LDS staging, y rows per thread TY).  Default mapping: a block is 4 waves
// and an instance with W waves per z row gets 4/W row groups per block.  DENSE: row groups of
// nz threads packed back to back; the launch's block size (256 or 192 threads) is a multiple
// of the nz of every instance it covers.
template <int TY, bool ODD, bool CNT, bool DENSE>
__global__ __launch_bounds__(256) void k_rk2_ensemble_rows(
    const EnsInst* __restrict__ insts, const EnsItem* __restrict__ items, int nitems, int cx,
    int parity, const uint64_t* __restrict__ step_base, int step_off, int xcd_remap,
    unsigned long long* __restrict__ vac, float esc)
{
    extern __shared__ float lds_ens[];
    int b = (int)blockIdx.x;
    const int nb = (int)gridDim.x;
    if (xcd_remap) b = (b & 7) * (nb >> 3) + (b >> 3);
    const int q = (nitems + nb - 1) / nb;
    const int i0 = b * q;
    const int i1 = min(i0 + q, nitems);
    const uint64_t step = step_base[0] + (uint64_t)step_off;
    for (int i = i0; i < i1; ++i) {
        const EnsItem it = items[i];
        const EnsInst& I = insts[it.inst];
        TableF32<ODD> tab;
        tab.ab = I.ab;
        tab.tmin = I.tmin;
        tab.dinv = I.dinv;
        tab.last = I.last;
        tab.off = I.off;
        const int x1 = min(it.x0 + cx, I.nx);
        __syncthreads();  // the previous item's last plane may still be read from LDS
        unsigned cnt = rk2_real_rows_tile<ODD, TY, NOISE_HASH, CNT, DENSE>(
            I.f[parity][0], I.f[parity][1], I.f[parity ^ 1][0], I.f[parity ^ 1][1], nullptr, I.c,
            tab, I.nx, I.ny, I.nz, I.W, it.tp, it.x0, x1, step ^ I.seed_mix, esc, lds_ens);
        if (CNT && cnt && (threadIdx.x & 63) == 0) atomicAdd(&vac[it.inst], (unsigned long long)cnt);
    }
}

__global__ void k_advance(uint64_t* step_base, uint64_t* step_base2, int n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        step_base[0] += (uint64_t)n;
        if (step_base2) step_base2[0] += (uint64_t)n;
    }
}

// A run of work items launched together: items [first, last), threads per block, and whether
// the dense row mapping applies.
struct EnsGroup {
    int first, last, threads, dense;
};

}  // namespace

struct tip_ensemble {
    int n = 0;
    std::vector<int64_t> dims;   // 3n
    std::vector<uint64_t> seeds;
    std::vector<size_t> off;     // element offset of instance i in each pool quarter
    size_t total = 0;            // sites over all instances
    float* pool = nullptr;       // 4 quarters: phi0, pi0, phi1, pi1 (each `total` floats)
    std::vector<EnsInst> host_inst;
    EnsInst* dev_inst = nullptr;
    std::vector<double*> table64;
    std::vector<float2*> table_ab;
    std::vector<int64_t> table_cap;
    std::vector<int> has_table, has_params;
    uint64_t* dev_step = nullptr;
    // the second stream's copy of the step counter: graph replay advances each stream's
    // chain of groups with its own counter, so the two chains never wait on each other
    // between replays; every call leaves both equal to `step`
    uint64_t* dev_step2 = nullptr;
    uint64_t step = 0;
    int parity = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;            // second stream of the concurrent schedule
    hipEvent_t fork = nullptr, join = nullptr;  // ordering between the two streams
    int concurrency = 1;                        // groups stepping at once (1 or 2)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    unsigned long long* dev_vac = nullptr;
    // geometry: kind (TIP_KERNEL_*), by = rows (flat: 64-site rows per block; row tiles: y rows
    // per thread, 0 = default), cx = x planes per work item
    int kind = TIP_KERNEL_AUTO, by = 0, cx = 8, xcd = 1, persistent = 0;
    int use_rows = 0, ty = 1, flat_by = 4;  // resolved by build_items
    // schedule: instances are stepped in groups of about group_bytes of field data (both
    // buffers of phi and pi), split evenly inside each block-size class; each group runs all
    // steps of a call before the next group
    int64_t group_bytes = 0;
    std::vector<EnsGroup> groups;
    EnsItem* dev_items = nullptr;
    int nitems = 0;
    bool items_dirty = true;
    bool desc_dirty = true;
    // (nsteps, parity, chain) -> exec; chain 1 is the second stream's groups
    std::map<std::tuple<int64_t, int, int>, hipGraphExec_t> graphs;
};

namespace {

void drop_graphs(tip_ensemble* E)
{
    for (auto& kv : E->graphs) (void)hipGraphExecDestroy(kv.second);
    E->graphs.clear();
}

int release(tip_ensemble* E)
{
    if (!E) return TIP_OK;
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    drop_graphs(E);
    if (E->pool) (void)hipFree(E->pool);
    if (E->dev_inst) (void)hipFree(E->dev_inst);
    if (E->dev_items) (void)hipFree(E->dev_items);
    if (E->dev_step) (void)hipFree(E->dev_step);
    if (E->dev_step2) (void)hipFree(E->dev_step2);
    if (E->dev_vac) (void)hipFree(E->dev_vac);
    for (auto p : E->table64)
        if (p) (void)hipFree(p);
    for (auto p : E->table_ab)
        if (p) (void)hipFree(p);
    if (E->stream2) (void)hipStreamSynchronize(E->stream2);
    if (E->ev0) (void)hipEventDestroy(E->ev0);
    if (E->ev1) (void)hipEventDestroy(E->ev1);
    if (E->fork) (void)hipEventDestroy(E->fork);
    if (E->join) (void)hipEventDestroy(E->join);
    if (E->stream2) (void)hipStreamDestroy(E->stream2);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
    return TIP_OK;
}

size_t inst_sites(const tip_ensemble* E, int i)
{
    return (size_t)(E->dims[3 * i] * E->dims[3 * i + 1] * E->dims[3 * i + 2]);
}

float* field(tip_ensemble* E, int i, int parity, int which)
{
    // quarters: [parity*2 + which]
    return E->pool + (size_t)(parity * 2 + which) * E->total + E->off[i];
}

void fill_pointers(tip_ensemble* E)
{
    for (int i = 0; i < E->n; ++i) {
        EnsInst& I = E->host_inst[i];
        for (int p = 0; p < 2; ++p)
            for (int w = 0; w < 2; ++w) I.f[p][w] = field(E, i, p, w);
        I.nx = (int)E->dims[3 * i];
        I.ny = (int)E->dims[3 * i + 1];
        I.nz = (int)E->dims[3 * i + 2];
        I.seed_mix = E->seeds[i] * 0x9E3779B97F4A7C15ULL;
    }
}

int sync_descriptors(tip_ensemble* E)
{
    if (!E->desc_dirty) return TIP_OK;
    TIP_HIP(hipMemcpyAsync(E->dev_inst, E->host_inst.data(), sizeof(EnsInst) * E->n,
                           hipMemcpyHostToDevice, E->stream));
    TIP_HIP(hipStreamSynchronize(E->stream));
    E->desc_dirty = false;
    return TIP_OK;
}

// Whether the row-tile kernel runs this ensemble and with how many rows per thread.
int resolve_kernel(tip_ensemble* E)
{
    bool rows_ok = E->kind != TIP_KERNEL_FLAT;
    int ty = E->by;
    if (E->kind == TIP_KERNEL_FLAT || (E->kind == TIP_KERNEL_AUTO && E->by == 8)) rows_ok = false;
    if (rows_ok && ty == 0) {  // default: the largest of 4, 2, 1 dividing every ny
        ty = 4;
        for (int i = 0; i < E->n; ++i)
            while (E->dims[3 * i + 1] % ty) ty >>= 1;
    }
    for (int i = 0; rows_ok && i < E->n; ++i)
        if (E->dims[3 * i + 2] > 256 || E->dims[3 * i + 1] % ty) rows_ok = false;
    TIP_REQUIRE(rows_ok || E->kind != TIP_KERNEL_ROWS,
                "ensemble: the row kernel needs nz <= 256 and rows per thread dividing every ny");
    E->use_rows = rows_ok ? 1 : 0;
    E->ty = rows_ok ? ty : 1;
    E->flat_by = (!rows_ok && E->by) ? E->by : 4;
    return TIP_OK;
}

// Threads per block of the dense row mapping for a z extent: 256 or 192 when that is a
// multiple of nz (rows of 96 sites: two per 3-wave block, where whole waves per row would
// leave a quarter of the lanes idle), else 0 (whole waves per z row, 256 threads).
int dense_threads(int64_t nz)
{
    if (256 % nz == 0) return 256;
    if (192 % nz == 0) return 192;
    return 0;
}

int build_items(tip_ensemble* E)
{
    if (!E->items_dirty) return TIP_OK;
    int rc = resolve_kernel(E);
    if (rc) return rc;
    // Instances ordered by block-size class (stable), so that each launch has one block size;
    // the stepping order of independent instances does not change any result.
    std::vector<int> order(E->n), cls(E->n, 0);
    for (int i = 0; i < E->n; ++i) {
        order[i] = i;
        cls[i] = E->use_rows ? dense_threads(E->dims[3 * i + 2]) : 0;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cls[a] < cls[b]; });
    std::vector<EnsItem> items;
    E->groups.clear();
    for (size_t k = 0; k < order.size();) {
        const int cl = cls[order[k]];
        size_t k1 = k;
        double cbytes = 0.0;
        for (; k1 < order.size() && cls[order[k1]] == cl; ++k1)
            cbytes += 16.0 * (double)inst_sites(E, order[k1]);
        const int threads = cl ? cl : 256;
        // a class of B bytes becomes ceil(B / group_bytes) groups of about equal bytes
        // (contiguous runs of its instances); no schedule: one group per class
        int64_t ng = 1;
        if (E->group_bytes > 0)
            ng = std::min<int64_t>((int64_t)(k1 - k),
                                   (int64_t)std::ceil(cbytes / (double)E->group_bytes));
        ng = std::max<int64_t>(ng, 1);
        double cum = 0.0;
        int64_t gcur = 0;
        size_t gfirst = items.size();
        for (size_t j = k; j < k1; ++j) {
            const int i = order[j];
            const int nx = (int)E->dims[3 * i], ny = (int)E->dims[3 * i + 1];
            const int nz = (int)E->dims[3 * i + 2];
            const double ibytes = 16.0 * (double)inst_sites(E, i);
            const int64_t gi =
                std::min<int64_t>(ng - 1, (int64_t)((cum + 0.5 * ibytes) / cbytes * (double)ng));
            if (gi != gcur && items.size() > gfirst) {
                E->groups.push_back({(int)gfirst, (int)items.size(), threads, cl ? 1 : 0});
                gfirst = items.size();
            }
            gcur = gi;
            cum += ibytes;
            int tiles;
            if (E->use_rows) {
                int W = (nz + 63) / 64;
                if (W == 3) W = 4;
                E->host_inst[i].W = W;
                const int G = cl ? threads / nz : 4 / W;
                tiles = (ny + G * E->ty - 1) / (G * E->ty);
            } else {
                tiles = (ny * nz + 64 * E->flat_by - 1) / (64 * E->flat_by);
            }
            for (int x0 = 0; x0 < nx; x0 += E->cx)
                for (int t = 0; t < tiles; ++t) items.push_back({i, t, x0, 0});
        }
        E->groups.push_back({(int)gfirst, (int)items.size(), threads, cl ? 1 : 0});
        k = k1;
    }
    E->desc_dirty = true;
    rc = sync_descriptors(E);
    if (rc) return rc;
    TIP_REQUIRE(!items.empty() && items.size() < (size_t(1) << 30), "ensemble: bad work list");
    TIP_HIP(hipStreamSynchronize(E->stream));
    // cached graphs hold the old work list's address in their kernel nodes: destroy them
    // before that list is freed
    drop_graphs(E);
    if (E->dev_items) TIP_HIP(hipFree(E->dev_items));
    E->dev_items = nullptr;
    TIP_HIP(hipMalloc(&E->dev_items, sizeof(EnsItem) * items.size()));
    TIP_HIP(hipMemcpy(E->dev_items, items.data(), sizeof(EnsItem) * items.size(),
                      hipMemcpyHostToDevice));
    E->nitems = (int)items.size();
    E->items_dirty = false;
    return TIP_OK;
}

// One step of the items of group g on stream st, its noise seeds from the step counter ctr.
int launch_batched(tip_ensemble* E, int parity, int step_off, const EnsGroup& g,
                   hipStream_t st, const uint64_t* ctr)
{
    const EnsItem* items = E->dev_items + g.first;
    const int nit = g.last - g.first;
    unsigned long long* const vac = nullptr;
    const float esc = 0.0f;
    int nb = nit;
    if (E->persistent > 0 && E->persistent < nb) nb = E->persistent;
    const int remap = (E->xcd && nb % 8 == 0) ? 1 : 0;
    if (E->use_rows) {
        const size_t lds = (size_t)g.threads * (4 * E->ty + 1) * sizeof(float);
        TIP_REQUIRE(g.threads == 256 || (g.dense && g.threads == 192),
                    "ensemble: bad block size %d", g.threads);
#define TIP_LAUNCH_ENSR_C(TYV, ODD, DENSE)                                                     \
    hipLaunchKernelGGL((k_rk2_ensemble_rows<TYV, ODD, false, DENSE>), dim3(nb),              \
                       dim3(64, g.threads / 64), lds, st, E->dev_inst, items, nit, E->cx, \
                       parity, ctr, step_off, remap, vac, esc)
#define TIP_LAUNCH_ENSR(TYV, ODD)                 \
    do {                                          \
        if (g.dense)                              \
            TIP_LAUNCH_ENSR_C(TYV, ODD, true);    \
        else                                      \
            TIP_LAUNCH_ENSR_C(TYV, ODD, false);   \
    } while (0)
#define TIP_LAUNCH_ENSR_TY(ODD)                        \
    switch (E->ty) {                                   \
        case 4: TIP_LAUNCH_ENSR(4, ODD); break;        \
        case 2: TIP_LAUNCH_ENSR(2, ODD); break;        \
        default: TIP_LAUNCH_ENSR(1, ODD); break;       \
    }
        if (E->host_inst[0].odd) {
            TIP_LAUNCH_ENSR_TY(true)
        } else {
            TIP_LAUNCH_ENSR_TY(false)
        }
#undef TIP_LAUNCH_ENSR_TY
#undef TIP_LAUNCH_ENSR
#undef TIP_LAUNCH_ENSR_C
        TIP_HIP(hipGetLastError());
        return TIP_OK;
    }
#define TIP_LAUNCH_ENS(BYV, ODD)                                                                \
    hipLaunchKernelGGL((k_rk2_ensemble<BYV, ODD>), dim3(nb), dim3(64, BYV), 0, st,            \
                       E->dev_inst, items, nit, E->cx, parity, ctr,                            \
                       step_off, remap, vac, esc)
#define TIP_LAUNCH_ENS_BY(ODD)                         \
    switch (E->flat_by) {                              \
        case 1: TIP_LAUNCH_ENS(1, ODD); break;         \
        case 2: TIP_LAUNCH_ENS(2, ODD); break;         \
        case 8: TIP_LAUNCH_ENS(8, ODD); break;         \
        default: TIP_LAUNCH_ENS(4, ODD); break;        \
    }
    if (E->host_inst[0].odd) {
        TIP_LAUNCH_ENS_BY(true)
    } else {
        TIP_LAUNCH_ENS_BY(false)
    }
#undef TIP_LAUNCH_ENS_BY
#undef TIP_LAUNCH_ENS
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

// All launches of nsteps steps starting at buffer parity par0 (no counter bump).  With
// concurrency 2 the groups alternate between the ensemble's stream and a second stream (forked
// from the first and joined back at the end), so two independent groups step at once and
// the last blocks of one group's launch share the GPU with the other group's blocks; each
// group's steps stay in order on its own stream.
int issue_steps(tip_ensemble* E, int64_t nsteps, int par0)
{
    const bool two = E->concurrency == 2 && E->groups.size() > 1;
    if (two) {
        TIP_HIP(hipEventRecord(E->fork, E->stream));
        TIP_HIP(hipStreamWaitEvent(E->stream2, E->fork, 0));
    }
    for (size_t gi = 0; gi < E->groups.size(); ++gi) {
        const hipStream_t st = (two && (gi & 1)) ? E->stream2 : E->stream;
        for (int64_t s = 0; s < nsteps; ++s) {
            int rc = launch_batched(E, par0 ^ (int)(s & 1), (int)s, E->groups[gi], st,
                                    E->dev_step);
            if (rc) return rc;
        }
    }
    if (two) {
        TIP_HIP(hipEventRecord(E->join, E->stream2));
        TIP_HIP(hipStreamWaitEvent(E->stream, E->join, 0));
    }
    return TIP_OK;
}

int ready(tip_ensemble* E)
{
    for (int i = 0; i < E->n; ++i) {
        TIP_REQUIRE(E->has_table[i], "ensemble: instance %d has no V' table", i);
        TIP_REQUIRE(E->host_inst[i].odd == E->host_inst[0].odd,
                    "ensemble: all instances must use odd tables or all even ones");
        TIP_REQUIRE(E->has_params[i], "ensemble: instance %d has no step parameters", i);
    }
    int rc = sync_descriptors(E);
    if (rc) return rc;
    return build_items(E);
}

// Steps per captured graph: bounds a graph's node count (groups x steps + 1) whatever the
// length of the call; an even count leaves the buffer parity unchanged, so one graph per
// parity serves every full chunk.  64 steps (a few hundred kernel nodes per chain at the
// nucleation ensemble's group sizes) make replay as fast as direct launches; 16 steps cost
// ~5 % in graph launches.
constexpr int64_t kGraphChunk = 64;

// The graph of k steps of one chain of groups starting at buffer parity `par`, captured on
// first use: with `chains` = 2 chain c holds groups c, c + 2, ... (the groups the batched
// schedule puts on stream c) and ends by advancing that stream's step counter; with one
// chain it holds every group.
int graph_for(tip_ensemble* E, int64_t k, int par, int chain, int chains, hipGraphExec_t* out)
{
    auto key = std::make_tuple(k, par, chain);
    auto itg = E->graphs.find(key);
    if (itg == E->graphs.end()) {
        hipGraph_t g = nullptr;
        const hipStream_t st = chain ? E->stream2 : E->stream;
        uint64_t* const ctr = chain ? E->dev_step2 : E->dev_step;
        TIP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = TIP_OK;
        for (size_t gi = (size_t)chain; gi < E->groups.size() && !rc; gi += (size_t)chains)
            for (int64_t s = 0; s < k && !rc; ++s)
                rc = launch_batched(E, par ^ (int)(s & 1), (int)s, E->groups[gi], st, ctr);
        if (rc) {
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, ctr, nullptr, (int)k);
        TIP_HIP(hipStreamEndCapture(st, &g));
        hipGraphExec_t ex;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail_hip(e, "hipGraphInstantiate");
        itg = E->graphs.emplace(key, ex).first;
    }
    *out = itg->second;
    return TIP_OK;
}

}  // namespace

extern "C" int tip_ensemble_create(tip_ensemble** out, int n, const int64_t* dims,
                                   const uint64_t* seeds)
{
    TIP_REQUIRE(out && dims, "tip_ensemble_create: NULL");
    *out = nullptr;
    TIP_REQUIRE(n >= 1 && n <= 65536, "tip_ensemble_create: n=%d out of range", n);
    tip_ensemble* E = new tip_ensemble();
    E->n = n;
    E->dims.assign(dims, dims + 3 * (size_t)n);
    E->seeds.resize(n);
    E->off.resize(n);
    for (int i = 0; i < n; ++i) {
        int rc = check_dims(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]);
        if (rc) {
            delete E;
            return rc;
        }
        E->seeds[i] = seeds ? seeds[i] : (uint64_t)i;
        E->off[i] = E->total;
        // keep every instance 256-byte aligned inside the pool
        E->total += (inst_sites(E, i) + 63) & ~size_t(63);
    }
    E->host_inst.assign(n, EnsInst{});
    E->table64.assign(n, nullptr);
    E->table_ab.assign(n, nullptr);
    E->table_cap.assign(n, 0);
    E->has_table.assign(n, 0);
    E->has_params.assign(n, 0);
    hipError_t e = hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&E->stream2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&E->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&E->join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&E->ev0);
    if (e == hipSuccess) e = hipEventCreate(&E->ev1);
    if (e == hipSuccess) e = hipMalloc(&E->pool, 4 * E->total * sizeof(float));
    if (e == hipSuccess) e = hipMemsetAsync(E->pool, 0, 4 * E->total * sizeof(float), E->stream);
    if (e == hipSuccess) e = hipMalloc(&E->dev_inst, sizeof(EnsInst) * n);
    if (e == hipSuccess) e = hipMalloc(&E->dev_step, sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemsetAsync(E->dev_step, 0, sizeof(uint64_t), E->stream);
    if (e == hipSuccess) e = hipMalloc(&E->dev_step2, sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemsetAsync(E->dev_step2, 0, sizeof(uint64_t), E->stream);
    if (e == hipSuccess) e = hipMalloc(&E->dev_vac, sizeof(unsigned long long) * n);
    if (e == hipSuccess) e = hipStreamSynchronize(E->stream);
    if (e != hipSuccess) {
        int r = fail_hip(e, "tip_ensemble_create");
        release(E);
        return r;
    }
    fill_pointers(E);
    *out = E;
    return TIP_OK;
}

extern "C" int tip_ensemble_destroy(tip_ensemble* ens) { return release(ens); }

extern "C" int tip_ensemble_upload(tip_ensemble* E, int i, int which, const float* host)
{
    TIP_REQUIRE(E && host, "tip_ensemble_upload: NULL");
    TIP_REQUIRE(i >= 0 && i < E->n && (which == 0 || which == 1), "tip_ensemble_upload: bad index");
    TIP_HIP(hipMemcpyAsync(field(E, i, E->parity, which), host, inst_sites(E, i) * sizeof(float),
                           hipMemcpyHostToDevice, E->stream));
    TIP_HIP(hipStreamSynchronize(E->stream));
    return TIP_OK;
}

extern "C" int tip_ensemble_download(tip_ensemble* E, int i, int which, float* host)
{
    TIP_REQUIRE(E && host, "tip_ensemble_download: NULL");
    TIP_REQUIRE(i >= 0 && i < E->n && (which == 0 || which == 1),
                "tip_ensemble_download: bad index");
    TIP_HIP(hipMemcpyAsync(host, field(E, i, E->parity, which), inst_sites(E, i) * sizeof(float),
                           hipMemcpyDeviceToHost, E->stream));
    TIP_HIP(hipStreamSynchronize(E->stream));
    return TIP_OK;
}

extern "C" int tip_ensemble_set_table(tip_ensemble* E, int i, const double* host_table,
                                      int64_t npts, double tmin, double dinv, int odd)
{
    TIP_REQUIRE(E && host_table, "tip_ensemble_set_table: NULL");
    TIP_REQUIRE(i >= 0 && i < E->n, "tip_ensemble_set_table: bad instance %d", i);
    TIP_REQUIRE(npts >= 2 && npts < (int64_t(1) << 30), "tip_ensemble_set_table: npts out of range");
    TIP_REQUIRE(dinv > 0.0, "tip_ensemble_set_table: dinv must be positive");
    if (npts > E->table_cap[i]) {
        TIP_HIP(hipStreamSynchronize(E->stream));
        if (E->table64[i]) TIP_HIP(hipFree(E->table64[i]));
        if (E->table_ab[i]) TIP_HIP(hipFree(E->table_ab[i]));
        E->table64[i] = nullptr;
        E->table_ab[i] = nullptr;
        TIP_HIP(hipMalloc(&E->table64[i], npts * sizeof(double)));
        TIP_HIP(hipMalloc(&E->table_ab[i], (npts - 1) * sizeof(float2)));
        E->table_cap[i] = npts;
    }
    TIP_HIP(hipMemcpyAsync(E->table64[i], host_table, npts * sizeof(double),
                           hipMemcpyHostToDevice, E->stream));
    int rc = prepare_table_f32(E->table64[i], npts, tmin, dinv, E->table_ab[i], E->stream);
    if (rc) return rc;
    TIP_HIP(hipStreamSynchronize(E->stream));
    EnsInst& I = E->host_inst[i];
    // A captured graph fixes the kernel's odd/even table variant (and the stream is idle
    // here): a change of that flag invalidates every cached graph.
    if (E->has_table[i] && I.odd != (odd ? 1 : 0)) drop_graphs(E);
    I.ab = E->table_ab[i];
    I.tmin = (float)tmin;
    I.dinv = (float)dinv;
    I.off = (float)(-tmin * dinv);
    I.last = (int)(npts - 2);
    I.odd = odd ? 1 : 0;
    E->has_table[i] = 1;
    E->desc_dirty = true;
    return TIP_OK;
}

extern "C" int tip_ensemble_set_params(tip_ensemble* E, int i, const tip_step_params* p)
{
    TIP_REQUIRE(E && p, "tip_ensemble_set_params: NULL");
    TIP_REQUIRE(i >= 0 && i < E->n, "tip_ensemble_set_params: bad instance %d", i);
    EnsInst& I = E->host_inst[i];
    I.c = make_scalars<float>(p->dt, p->dx, p->eta_eff, p->mu, p->inv_a2, p->noise_scale);
    I.hash = p->noise_scale != 0.0 ? 1 : 0;
    E->has_params[i] = 1;
    E->desc_dirty = true;
    return TIP_OK;
}

extern "C" int tip_ensemble_set_geometry(tip_ensemble* E, int kernel, int rows,
                                         int planes_per_block, int xcd_remap,
                                         int persistent_blocks)
{
    TIP_REQUIRE(E, "tip_ensemble_set_geometry: NULL");
    TIP_REQUIRE(kernel == TIP_KERNEL_AUTO || kernel == TIP_KERNEL_ROWS ||
                    kernel == TIP_KERNEL_FLAT,
                "ensemble geometry: unknown kernel %d", kernel);
    TIP_REQUIRE(kernel == TIP_KERNEL_ROWS ? (rows == 0 || rows == 1 || rows == 2 || rows == 4)
                : kernel == TIP_KERNEL_FLAT
                    ? (rows == 1 || rows == 2 || rows == 4 || rows == 8)
                    : (rows == 0 || rows == 1 || rows == 2 || rows == 4 || rows == 8),
                "ensemble geometry: rows %d not valid for kernel %d", rows, kernel);
    TIP_REQUIRE(planes_per_block >= 1 && planes_per_block <= 4096,
                "ensemble geometry: planes per block out of range");
    TIP_REQUIRE(persistent_blocks >= 0, "ensemble geometry: persistent_blocks < 0");
    const int old_kind = E->kind, old_by = E->by;
    E->kind = kernel;
    E->by = rows;
    int rc = resolve_kernel(E);
    if (rc) {
        E->kind = old_kind;
        E->by = old_by;
        return rc;
    }
    E->cx = planes_per_block;
    E->xcd = xcd_remap ? 1 : 0;
    E->persistent = persistent_blocks;
    E->items_dirty = true;
    return TIP_OK;
}

extern "C" int tip_ensemble_set_schedule(tip_ensemble* E, int64_t group_bytes)
{
    TIP_REQUIRE(E, "tip_ensemble_set_schedule: NULL");
    TIP_REQUIRE(group_bytes >= 0, "tip_ensemble_set_schedule: group_bytes < 0");
    E->group_bytes = group_bytes;
    E->items_dirty = true;
    return TIP_OK;
}

extern "C" int tip_ensemble_set_concurrency(tip_ensemble* E, int groups_at_once)
{
    TIP_REQUIRE(E, "tip_ensemble_set_concurrency: NULL");
    TIP_REQUIRE(groups_at_once == 1 || groups_at_once == 2,
                "tip_ensemble_set_concurrency: 1 or 2 groups at once (got %d)", groups_at_once);
    if (groups_at_once != E->concurrency) {
        TIP_HIP(hipStreamSynchronize(E->stream));
        drop_graphs(E);  // captured graphs hold the previous stream layout
        E->concurrency = groups_at_once;
    }
    return TIP_OK;
}

extern "C" int tip_ensemble_step(tip_ensemble* E, int64_t nsteps, int mode, float* elapsed_ms)
{
    TIP_REQUIRE(E, "tip_ensemble_step: NULL");
    TIP_REQUIRE(nsteps >= 0, "tip_ensemble_step: nsteps < 0");
    TIP_REQUIRE(mode == TIP_ENS_SEQUENTIAL || mode == TIP_ENS_BATCHED || mode == TIP_ENS_GRAPH,
                "tip_ensemble_step: unknown mode %d", mode);
    int rc = ready(E);
    if (rc) return rc;
    if (elapsed_ms) TIP_HIP(hipEventRecord(E->ev0, E->stream));
    if (mode == TIP_ENS_SEQUENTIAL) {
        // the plain loop: one single-lattice kernel launch per instance per step
        for (int64_t s = 0; s < nsteps; ++s) {
            const int par = E->parity;
            for (int i = 0; i < E->n; ++i) {
                const EnsInst& I = E->host_inst[i];
                SegTable tab;
                tab.ab = I.ab;
                tab.tmin = I.tmin;
                tab.dinv = I.dinv;
                tab.last = I.last;
                tab.odd = I.odd;
                tab.off = I.off;
                rc = launch_real_f32(I.f[par][0], I.f[par][1], I.f[par ^ 1][0], I.f[par ^ 1][1],
                                     I.c, nullptr, tab, I.nx, I.ny, I.nz,
                                     (E->step + (uint64_t)s) ^ I.seed_mix, I.hash, nullptr, 0.0f,
                                     nullptr, E->stream);
                if (rc) return rc;
            }
            E->parity ^= 1;
        }
        hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, E->stream, E->dev_step, E->dev_step2,
                           (int)nsteps);
        TIP_HIP(hipGetLastError());
    } else if (mode == TIP_ENS_BATCHED || nsteps == 0) {
        rc = issue_steps(E, nsteps, E->parity);
        if (rc) return rc;
        if (nsteps & 1) E->parity ^= 1;
        hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, E->stream, E->dev_step, E->dev_step2,
                           (int)nsteps);
        TIP_HIP(hipGetLastError());
    } else {
        // Each chain of groups (one per stream) is a graph of a chunk of at most kGraphChunk
        // steps (its groups x chunk kernel nodes plus its counter bump), replayed on its own
        // stream; a call of n steps is n / chunk replays of the full chunk and one of the
        // remainder.  The chains are forked from the ensemble's stream once per call and
        // joined at its end, so successive replays of one chain follow each other without
        // waiting for the other chain.  Each replay advances its chain's step counter by
        // its chunk, so the noise seeds equal those of batched launches.
        const bool two = E->concurrency == 2 && E->groups.size() > 1;
        const int chains = two ? 2 : 1;
        if (two) {
            TIP_HIP(hipEventRecord(E->fork, E->stream));
            TIP_HIP(hipStreamWaitEvent(E->stream2, E->fork, 0));
        }
        int64_t left = nsteps;
        while (left > 0) {
            const int64_t k = left < kGraphChunk ? left : kGraphChunk;
            for (int c = 0; c < chains; ++c) {
                hipGraphExec_t ex = nullptr;
                rc = graph_for(E, k, E->parity, c, chains, &ex);
                if (rc) return rc;
                TIP_HIP(hipGraphLaunch(ex, c ? E->stream2 : E->stream));
            }
            if (k & 1) E->parity ^= 1;
            left -= k;
        }
        if (two) {
            TIP_HIP(hipEventRecord(E->join, E->stream2));
            TIP_HIP(hipStreamWaitEvent(E->stream, E->join, 0));
        } else {
            // the second counter follows the first
            hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, E->stream, E->dev_step2,
                               nullptr, (int)nsteps);
            TIP_HIP(hipGetLastError());
        }
    }
    E->step += (uint64_t)nsteps;
    if (elapsed_ms) {
        TIP_HIP(hipEventRecord(E->ev1, E->stream));
        TIP_HIP(hipEventSynchronize(E->ev1));
        TIP_HIP(hipEventElapsedTime(elapsed_ms, E->ev0, E->ev1));
    }
    return TIP_OK;
}

extern "C" int tip_ensemble_vacuum_counts(tip_ensemble* E, double esc,
                                          unsigned long long* host_counts)
{
    TIP_REQUIRE(E && host_counts, "tip_ensemble_vacuum_counts: NULL");
    TIP_HIP(hipMemsetAsync(E->dev_vac, 0, sizeof(unsigned long long) * E->n, E->stream));
    for (int i = 0; i < E->n; ++i) {
        int rc = tip_vacuum_count_f32(field(E, i, E->parity, 0), (int64_t)inst_sites(E, i), esc,
                                      E->dev_vac + i, E->stream);
        if (rc) return rc;
    }
    TIP_HIP(hipMemcpyAsync(host_counts, E->dev_vac, sizeof(unsigned long long) * E->n,
                           hipMemcpyDeviceToHost, E->stream));
    TIP_HIP(hipStreamSynchronize(E->stream));
    return TIP_OK;
}

extern "C" int tip_ensemble_step_count(tip_ensemble* E, uint64_t* step)
{
    TIP_REQUIRE(E && step, "tip_ensemble_step_count: NULL");
    *step = E->step;
    return TIP_OK;
}

extern "C" int tip_ensemble_set_step(tip_ensemble* E, uint64_t step)
{
    TIP_REQUIRE(E, "tip_ensemble_set_step: NULL");
    TIP_HIP(hipStreamSynchronize(E->stream));
    TIP_HIP(hipMemcpy(E->dev_step, &step, sizeof(uint64_t), hipMemcpyHostToDevice));
    TIP_HIP(hipMemcpy(E->dev_step2, &step, sizeof(uint64_t), hipMemcpyHostToDevice));
    E->step = step;
    return TIP_OK;
}
