// tip_mlp.hip -- learned V'(rho, T) surrogate: a 2 -> 32 -> 32 -> 1 tanh MLP in float32,
// trained on the GPU by hand-written forward/backward/Adam kernels and evaluated inline in the
// complex-field step kernel so a continuously cooling T(t) needs no re-tabulation.
//
// Parameter vector (1185 float32): W1[32][2], b1[32], W2[32][32], b2[32], w3[32], b3.
// Features: f1 = asinh(rho / rho_scale) / asinh(rho_max / rho_scale)  (resolves the thermal
// barrier, a few T wide, and the vacuum at ~1e11 GeV in one input), f2 = (T - T_mid) / T_half.
// Output: V'(rho, T) = rho * (lam rho^2 - m2 + out_scale * (g0(f2) + f1 y) * e(rho)),
// e = 1 / (1 + (rho/rho_cut)^2): the tree term is exact; g0(f2), a cubic in f2, is the thermal part of V'/rho at rho -> 0
// (the effective mass, fitted on the host to the float64 potential), so the effective mass
// does not depend on the network (f1 = 0 at rho = 0); the network carries the rest of the
// thermal and Coleman-Weinberg part of V'/rho (smooth, Boltzmann-suppressed at large rho)
// divided by a fixed envelope, so its residual in the flat large-rho region is suppressed.
//
// Training is deterministic: each 64-sample block writes its partial gradient and one kernel
// sums the partials in block order and applies Adam.
#include "step_kernels.h"

using namespace tip;

namespace {

constexpr int H = TIP_MLP_HIDDEN;
constexpr int NP = TIP_MLP_NPARAMS;
constexpr int OW1 = 0, OB1 = 2 * H, OW2 = 3 * H, OB2 = 3 * H + H * H, OW3 = 4 * H + H * H,
              OB3 = 5 * H + H * H;
static_assert(OB3 + 1 == NP, "parameter layout");
constexpr int SB = 64;  // samples per training block (one wave)

__device__ __forceinline__ float tanh_f(float x)
{
    // tanh(x) = 1 - 2 / (exp(2x) + 1); exact limits for |x| large
    float e = __expf(2.0f * x);
    return 1.0f - 2.0f / (e + 1.0f);
}

struct FeatureMap {
    float inv_rho_s, inv_norm, T_mid, inv_T_half, out_scale, inv_rho_cut;
    float g0c[4];
    double lam, m2;
    // The tree term lam rho^3 - m2 rho cancels near the vacuum; it is formed in float64 (full
    // rate on MI355X) so V' is rounded to float32 once, relative to its own size.  b is the
    // bracket g0(f2) + f1 y.
    __device__ __forceinline__ float vprime(float rho, float b) const
    {
        const float q = rho * inv_rho_cut;
        const float th = out_scale * b * __builtin_amdgcn_rcpf(fmaf(q, q, 1.0f));
        const double r = (double)rho;
        return (float)(r * (fma(lam * r, r, -m2) + (double)th));
    }
    __host__ __device__ __forceinline__ float g0(float f2v) const
    {
        return fmaf(fmaf(fmaf(g0c[3], f2v, g0c[2]), f2v, g0c[1]), f2v, g0c[0]);
    }
    __device__ __forceinline__ float f1(float rho) const
    {
        float a = rho * inv_rho_s;
        return __logf(a + __builtin_amdgcn_sqrtf(fmaf(a, a, 1.0f))) * inv_norm;
    }
    __device__ __forceinline__ float f2(float T) const { return (T - T_mid) * inv_T_half; }
};

FeatureMap make_fm(const tip_mlp_features* fm)
{
    FeatureMap m;
    m.inv_rho_s = (float)(1.0 / fm->rho_scale);
    m.inv_norm = (float)(1.0 / asinh(fm->rho_max / fm->rho_scale));
    m.T_mid = (float)fm->T_mid;
    m.inv_T_half = (float)(1.0 / fm->T_half);
    m.out_scale = (float)fm->out_scale;
    m.lam = fm->lam;
    m.m2 = fm->m2;
    m.inv_rho_cut = (float)(1.0 / fm->rho_cut);
    for (int k = 0; k < 4; ++k) m.g0c[k] = (float)fm->g0[k];
    return m;
}

int check_fm(const tip_mlp_features* fm)
{
    TIP_REQUIRE(fm, "mlp: NULL feature map");
    TIP_REQUIRE(fm->rho_scale > 0 && fm->rho_max > 0 && fm->T_half > 0 && fm->rho_cut > 0,
                "mlp: feature map scales must be positive");
    return TIP_OK;
}

// y = MLP(f1, f2) with the weights read through uniform (scalar) loads.
__device__ __forceinline__ float mlp_eval(const float* __restrict__ P, float x0, float x1)
{
    float h1[H];
#pragma unroll
    for (int j = 0; j < H; ++j) h1[j] = tanh_f(fmaf(P[OW1 + 2 * j], x0, fmaf(P[OW1 + 2 * j + 1], x1, P[OB1 + j])));
    float y = P[OB3];
#pragma unroll 4
    for (int j = 0; j < H; ++j) {
        float z = P[OB2 + j];
#pragma unroll
        for (int i = 0; i < H; ++i) z = fmaf(P[OW2 + H * j + i], h1[i], z);
        y = fmaf(P[OW3 + j], tanh_f(z), y);
    }
    return y;
}

// V'(rho) at a fixed temperature inside the complex row-tile step, the network's 32 x 32 layer
// on the matrix cores.  A wave evaluates its sites 16 at a time: for a block of 16 sites
// (lanes 16 nb .. 16 nb + 15 hold their radii) lane l = 16 g + p computes the first layer for
// site p and neurons 8 g + j (j = 0..7) on the VALU; those eight tanh values are its B operand
// of v_mfma_f32_16x16x32_f16 (B[k = 8 g + j][n = p]), the A operand holds the weights
// (A[m = p][k = 8 g + j] = W2[16 m2 + p][8 g + j], two 16-row halves m2), the accumulators
// start at b2, and lane l receives z2 of neurons 16 m2 + 4 g + i (i = 0..3) of site p.  Their
// tanh, weighted by w3, summed over the lane's eight neurons and across the four lane groups
// (two butterfly exchanges), give y of site p in every group; the lane of group nb keeps it.
// Float32 accuracy from half-precision matrix cores: every operand is split into a
// half-precision head and tail (x = hi + lo, lo = half(x - hi)), and hi*hi + hi*lo + lo*hi is
// accumulated in float32 (three MFMAs per product; the dropped lo*lo is below 2^-22 of it).
// The per-lane weight halves (32 halves per lane) and the per-group first-layer, bias and
// output weights (32 floats per group, first layer with this launch's f2 folded in) sit in
// LDS, filled once per block, pre-multiplied by 2 log2(e) where they feed a tanh.  tanh comes
// from the hardware exp2 and reciprocal.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// 2 log2(e): the constants of both layers are stored multiplied by it, so the pre-activations
// arrive as x' = 2 x log2(e) and tanh(x) = 1 - 2 / (2^x' + 1) (+-1 exactly for |x| large).
constexpr float kTwoLog2e = 2.8853900817779268f;
__device__ __forceinline__ float tanh_scaled(float xs)
{
    return fmaf(-2.0f, __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(xs) + 1.0f), 1.0f);
}

struct MlpMfmaRadial {
    const float* P;
    FeatureMap fm;
    float f2, g0v;
    float* cst;  // this block's constants in LDS (block_init)
    static constexpr bool kEarly = false;
    static constexpr bool kBatch = true;
    static constexpr int kLaneFloats = 64 * 16;              // 32 weight halves per lane
    static constexpr int kLdsFloats = kLaneFloats + 4 * 32;  // + per-group constants
    struct Pre {
        float f1, y;
    };

    __device__ __forceinline__ void block_init(float* l)
    {
        cst = l;
        const int nt = (int)(blockDim.x * blockDim.y);
        const int tid = (int)(threadIdx.y * blockDim.x + threadIdx.x);
        _Float16* wh = reinterpret_cast<_Float16*>(l);
        for (int e = tid; e < 2 * kLaneFloats; e += nt) {
            // lane's halves: [m2][head, tail][j]
            const int lane = e >> 5, q = e & 31;
            const int g = lane >> 4, p = lane & 15, m2 = q >> 4, part = (q >> 3) & 1, j = q & 7;
            const float w = P[OW2 + H * (16 * m2 + p) + 8 * g + j] * kTwoLog2e;
            const _Float16 hi = (_Float16)w;
            wh[e] = part ? (_Float16)(w - (float)hi) : hi;
        }
        for (int e = tid; e < 4 * 32; e += nt) {
            const int g = e >> 5, c = e & 31, q = c & 7;
            const int n1 = 8 * g + q;                          // first-layer neuron
            const int n2 = 16 * (q >> 2) + 4 * g + (q & 3);    // second-layer neuron
            float v;
            if (c < 8) v = P[OW1 + 2 * n1] * kTwoLog2e;
            else if (c < 16) v = fmaf(P[OW1 + 2 * n1 + 1], f2, P[OB1 + n1]) * kTwoLog2e;
            else if (c < 24) v = P[OB2 + n2] * kTwoLog2e;
            else v = P[OW3 + n2];
            l[kLaneFloats + e] = v;
        }
        __syncthreads();
    }

    template <int TY>
    __device__ __forceinline__ void batch(const float (&rho1)[TY], const float (&rho2)[TY],
                                          Pre (&P1)[TY], Pre (&P2)[TY]) const
    {
        const int l = (int)threadIdx.x, g = l >> 4, p = l & 15;
        const f16x8* wl = reinterpret_cast<const f16x8*>(cst) + 4 * l;
        const f32x4* gc = reinterpret_cast<const f32x4*>(cst + kLaneFloats + g * 32);
        const f16x8 ah0 = wl[0], al0 = wl[1], ah1 = wl[2], al1 = wl[3];
        const f32x4 a10 = gc[0], a11 = gc[1], c10 = gc[2], c11 = gc[3];
        const f32x4 b2a = gc[4], b2b = gc[5], w3a = gc[6], w3b = gc[7];
        const float b3 = P[OB3];
        float x1[2 * TY], y[2 * TY];
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            x1[r] = fm.f1(rho1[r]);
            x1[TY + r] = fm.f1(rho2[r]);
        }
#pragma unroll
        for (int e = 0; e < 2 * TY; ++e) {
            float ye = 0.0f;
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const float xf = __shfl(x1[e], nb * 16 + p);
                f16x8 bh, bl;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const float h = tanh_scaled(j < 4 ? fmaf(a10[j & 3], xf, c10[j & 3])
                                                      : fmaf(a11[j & 3], xf, c11[j & 3]));
                    const _Float16 hh = (_Float16)h;
                    bh[j] = hh;
                    bl[j] = (_Float16)(h - (float)hh);
                }
                f32x4 z0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah0, bh, b2a, 0, 0, 0);
                f32x4 z1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah1, bh, b2b, 0, 0, 0);
                z0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah0, bl, z0, 0, 0, 0);
                z1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah1, bl, z1, 0, 0, 0);
                z0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(al0, bh, z0, 0, 0, 0);
                z1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(al1, bh, z1, 0, 0, 0);
                float part = 0.0f;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    part = fmaf(w3a[i], tanh_scaled(z0[i]), part);
                    part = fmaf(w3b[i], tanh_scaled(z1[i]), part);
                }
                part += __shfl_xor(part, 16);
                part += __shfl_xor(part, 32);
                ye = g == nb ? part : ye;
            }
            y[e] = ye + b3;
        }
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            P1[r] = Pre{x1[r], y[r]};
            P2[r] = Pre{x1[TY + r], y[TY + r]};
        }
    }

    __device__ __forceinline__ float eval(const Pre& q, float rho) const
    {
        return fm.vprime(rho, fmaf(q.f1, q.y, g0v));
    }
};

// ---- training --------------------------------------------------------------------------
// One block = 64 samples = one wave.  Phase 1: forward/backward per sample, activations and
// deltas staged in LDS.  Phase 2: every lane accumulates a slice of the 1185 parameter
// gradients over the block's samples (fixed order).
__global__ __launch_bounds__(SB) void k_mlp_fwd_bwd(const float* __restrict__ P,
                                                    const float2* __restrict__ X,
                                                    const float* __restrict__ Y,
                                                    const float* __restrict__ Wt, int n,
                                                    float inv_n, float* __restrict__ part,
                                                    float* __restrict__ loss_part)
{
    __shared__ float sh1[SB][H + 1], sh2[SB][H + 1], sd2[SB][H + 1], sd1[SB][H + 1];
    __shared__ float sx[SB][2], sdy[SB], sloss[SB];
    const int t = threadIdx.x;
    const int s = blockIdx.x * SB + t;
    const bool act = s < n;
    float2 x = act ? X[s] : make_float2(0.f, 0.f);
    float h1[H], h2[H];
#pragma unroll
    for (int j = 0; j < H; ++j)
        h1[j] = tanh_f(fmaf(P[OW1 + 2 * j], x.x, fmaf(P[OW1 + 2 * j + 1], x.y, P[OB1 + j])));
    float y = P[OB3];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        float z = P[OB2 + j];
#pragma unroll
        for (int i = 0; i < H; ++i) z = fmaf(P[OW2 + H * j + i], h1[i], z);
        h2[j] = tanh_f(z);
        y = fmaf(P[OW3 + j], h2[j], y);
    }
    const float r = act ? (y - Y[s]) : 0.0f;
    const float w = (act && Wt) ? Wt[s] : 1.0f;
    const float dy = w * r * inv_n;  // d(0.5 mean w r^2)/dy
    sloss[t] = 0.5f * w * r * r;
    sdy[t] = dy;
    sx[t][0] = act ? x.x : 0.f;
    sx[t][1] = act ? x.y : 0.f;
    float d2[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        d2[j] = dy * P[OW3 + j] * (1.0f - h2[j] * h2[j]);
        sh1[t][j] = act ? h1[j] : 0.f;
        sh2[t][j] = act ? h2[j] : 0.f;
        sd2[t][j] = act ? d2[j] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
        float g = 0.f;
#pragma unroll
        for (int j = 0; j < H; ++j) g = fmaf(P[OW2 + H * j + i], d2[j], g);
        sd1[t][i] = act ? g * (1.0f - h1[i] * h1[i]) : 0.f;
    }
    __syncthreads();
    float* out = part + (size_t)blockIdx.x * NP;
    for (int p = t; p < NP; p += SB) {
        float g = 0.f;
        if (p < OB1) {  // W1[j][k]
            const int j = p >> 1, k = p & 1;
            for (int q = 0; q < SB; ++q) g = fmaf(sd1[q][j], sx[q][k], g);
        } else if (p < OW2) {  // b1[j]
            const int j = p - OB1;
            for (int q = 0; q < SB; ++q) g += sd1[q][j];
        } else if (p < OB2) {  // W2[j][i]
            const int j = (p - OW2) / H, i = (p - OW2) % H;
            for (int q = 0; q < SB; ++q) g = fmaf(sd2[q][j], sh1[q][i], g);
        } else if (p < OW3) {  // b2[j]
            const int j = p - OB2;
            for (int q = 0; q < SB; ++q) g += sd2[q][j];
        } else if (p < OB3) {  // w3[j]
            const int j = p - OW3;
            for (int q = 0; q < SB; ++q) g = fmaf(sdy[q], sh2[q][j], g);
        } else {  // b3
            for (int q = 0; q < SB; ++q) g += sdy[q];
        }
        out[p] = g;
    }
    if (t == 0) {
        float l = 0.f;
        for (int q = 0; q < SB; ++q) l += sloss[q];
        loss_part[blockIdx.x] = l;
    }
}

// Sum the per-block partials (block order) into the gradient; optionally apply Adam.
__global__ void k_mlp_reduce(const float* __restrict__ part, const float* __restrict__ loss_part,
                             int nblocks, float* __restrict__ grad, float* __restrict__ loss,
                             float inv_n, float* __restrict__ P, float* __restrict__ m,
                             float* __restrict__ v, float lr, float b1, float b2, float eps,
                             float bc1, float bc2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && loss) {
        double l = 0.0;
        for (int b = 0; b < nblocks; ++b) l += loss_part[b];
        loss[0] = (float)(l * inv_n);
    }
    if (p >= NP) return;
    float g = 0.f;
    for (int b = 0; b < nblocks; ++b) g += part[(size_t)b * NP + p];
    if (grad) grad[p] = g;
    if (P) {
        const float mm = b1 * m[p] + (1.0f - b1) * g;
        const float vv = b2 * v[p] + (1.0f - b2) * g * g;
        m[p] = mm;
        v[p] = vv;
        P[p] -= lr * (mm / bc1) / (__builtin_sqrtf(vv / bc2) + eps);
    }
}

__global__ void k_mlp_forward(const float* __restrict__ P, const float2* __restrict__ X,
                              float* __restrict__ Y, int n)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float2 x = X[s];
    Y[s] = mlp_eval(P, x.x, x.y);
}

__global__ void k_mlp_vprime(const float* __restrict__ P, FeatureMap fm,
                             const float* __restrict__ rho, const float* __restrict__ T,
                             float* __restrict__ vp, int n)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float r = rho[s];
    const float x1 = fm.f1(r), x2 = fm.f2(T[s]);
    vp[s] = fm.vprime(r, fmaf(x1, mlp_eval(P, x1, x2), fm.g0(x2)));
}

__global__ void k_mlp_tabulate(const float* __restrict__ P, FeatureMap fm, float f2, double x0,
                               double h, int n, double* __restrict__ table)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float r = (float)(x0 + h * (double)i);
    const float a = fabsf(r);
    const float x1 = fm.f1(a);
    const float v = fm.vprime(a, fmaf(x1, mlp_eval(P, x1, f2), fm.g0(f2)));
    table[i] = (double)(r < 0.0f ? -v : v);
}

int train_pass(const float* P, const float* X, const float* Y, const float* Wt, int64_t n,
               float* grad,
               float* loss, float* Pw, float* m, float* v, double lr, double b1, double b2,
               double eps, int64_t t, hipStream_t st)
{
    TIP_REQUIRE(P && X && Y, "mlp: NULL parameters or data");
    TIP_REQUIRE(n >= 1 && n < (int64_t(1) << 30), "mlp: batch size out of range");
    const int nb = (int)((n + SB - 1) / SB);
    ScratchLease lease{2, st};
    float* part =
        reinterpret_cast<float*>(scratch_get(((size_t)nb * NP + nb) * sizeof(float), 2, st));
    TIP_REQUIRE(part, "mlp: scratch allocation failed");
    float* lpart = part + (size_t)nb * NP;
    const float inv_n = (float)(1.0 / (double)n);
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nb), dim3(SB), 0, st, P,
                       reinterpret_cast<const float2*>(X), Y, Wt, (int)n, inv_n, part, lpart);
    TIP_HIP(hipGetLastError());
    float bc1 = 1.f, bc2 = 1.f;
    if (Pw) {
        TIP_REQUIRE(m && v && t >= 1, "mlp adam: moments and step t >= 1 are required");
        bc1 = (float)(1.0 - pow(b1, (double)t));
        bc2 = (float)(1.0 - pow(b2, (double)t));
    }
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NP + 255) / 256), dim3(256), 0, st, part, lpart, nb,
                       grad, loss, inv_n, Pw, m, v, (float)lr, (float)b1,
                       (float)b2, (float)eps, bc1, bc2);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

}  // namespace

extern "C" int tip_mlp_forward(const float* params, const float* x, float* y, int64_t n,
                               void* stream)
{
    TIP_REQUIRE(params && x && y && n >= 0 && n < (int64_t(1) << 31), "tip_mlp_forward: bad args");
    if (n == 0) return TIP_OK;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), params, reinterpret_cast<const float2*>(x), y, (int)n);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

extern "C" int tip_mlp_grad(const float* params, const float* x, const float* y_target,
                            const float* weights, int64_t n, float* grad, float* loss,
                            void* stream)
{
    TIP_REQUIRE(grad, "tip_mlp_grad: NULL grad");
    return train_pass(params, x, y_target, weights, n, grad, loss, nullptr, nullptr, nullptr, 0,
                      0, 0, 0, 0, as_stream(stream));
}

extern "C" int tip_mlp_adam_step(float* params, float* m, float* v, const float* x,
                                 const float* y_target, const float* weights, int64_t n,
                                 double lr, double beta1, double beta2, double eps, int64_t t,
                                 float* loss, void* stream)
{
    return train_pass(params, x, y_target, weights, n, nullptr, loss, params, m, v, lr, beta1,
                      beta2, eps, t, as_stream(stream));
}

extern "C" int tip_mlp_vprime(const float* params, const tip_mlp_features* fm, const float* rho,
                              const float* T, float* vp, int64_t n, void* stream)
{
    int rc = check_fm(fm);
    if (rc) return rc;
    TIP_REQUIRE(params && rho && T && vp && n >= 0 && n < (int64_t(1) << 31),
                "tip_mlp_vprime: bad args");
    if (n == 0) return TIP_OK;
    hipLaunchKernelGGL(k_mlp_vprime, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), params, make_fm(fm), rho, T, vp, (int)n);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

extern "C" int tip_mlp_tabulate(const float* params, const tip_mlp_features* fm, double T,
                                double x_min, double x_max, int64_t npts, double* table,
                                void* stream)
{
    int rc = check_fm(fm);
    if (rc) return rc;
    TIP_REQUIRE(params && table && npts >= 2 && npts < (int64_t(1) << 30) && x_max > x_min,
                "tip_mlp_tabulate: bad args");
    FeatureMap m = make_fm(fm);
    hipLaunchKernelGGL(k_mlp_tabulate, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0,
                       as_stream(stream), params, m, (float)((T - fm->T_mid) / fm->T_half), x_min,
                       (x_max - x_min) / (double)(npts - 1), (int)npts, table);
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

namespace {

// Grid of the network at one temperature: segment i of the bracket b(g) = g0 + f1 y on g in
// [i, i+1], g the grid coordinate of f1 = i * f1_step.
__global__ void k_mlp_grid(const float* __restrict__ P, float f2, float g0v, int nseg,
                           float f1_step, float2* __restrict__ seg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg) return;
    const float a0 = (float)i * f1_step, a1 = (float)(i + 1) * f1_step;
    const float y0 = fmaf(a0, mlp_eval(P, a0, f2), g0v);
    const float y1 = fmaf(a1, mlp_eval(P, a1, f2), g0v);
    seg[i] = make_float2(y0, y1 - y0);
}

}  // namespace

namespace tip {

// Largest feature value of the per-step grid: rho up to rho_max^1.05 in the compressed
// coordinate; beyond it the last segment extrapolates (the learned part is negligible there,
// its envelope is (rho_cut / rho)^2 < 1e-12).
constexpr float kGridF1Max = 1.05f;

int launch_complex_mlp_grid(const float* phi1, const float* phi2, const float* pi1,
                            const float* pi2, float* o1, float* o2, float* o3, float* o4,
                            const StepScalars<float>& c, const float* mlp_params,
                            const tip_mlp_features* fm, double T, const ZnParams<float>& zn,
                            int64_t nx, int64_t ny, int64_t nz, uint64_t noise_seed, int hash,
                            unsigned long long* vac, float esc2, float2* grid, int64_t npts,
                            hipStream_t st)
{
    int rc = check_fm(fm);
    if (rc) return rc;
    rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(phi1 && phi2 && pi1 && pi2 && o1 && o2 && o3 && o4 && mlp_params && grid,
                "rk2 c32 mlp grid: NULL pointer");
    TIP_REQUIRE(npts >= 2 && npts <= (int64_t(1) << 24), "rk2 c32 mlp grid: npts out of range");
    const int ty = default_rows_per_thread(nx, ny, nz) > 1 && ny % 2 == 0 ? 2 : 1;
    const RowShape rs = row_shape(ny, nz, ty, 8);
    TIP_REQUIRE(rs.ok, "rk2 c32 mlp grid: needs nz <= 512");
    const FeatureMap f = make_fm(fm);
    const int nseg = (int)(npts - 1);
    const float f1_step = kGridF1Max / (float)nseg;
    const float f2 = (float)((T - fm->T_mid) / fm->T_half);
    const float g0v = (float)(fm->g0[0] + f2 * (fm->g0[1] + f2 * (fm->g0[2] + f2 * fm->g0[3])));
    hipLaunchKernelGGL(k_mlp_grid, dim3((nseg + 255) / 256), dim3(256), 0, st, mlp_params, f2,
                       g0v, nseg, f1_step, grid);
    TIP_HIP(hipGetLastError());
    MlpGridRadial rad;
    rad.seg = grid;
    rad.last = nseg - 1;
    rad.inv_rho_s = f.inv_rho_s;
    rad.gscale = (float)((double)f.inv_norm * (double)nseg / (double)kGridF1Max);
    rad.out_scale = f.out_scale;
    rad.inv_rho_cut = f.inv_rho_cut;
    rad.lam = f.lam;
    rad.m2 = f.m2;
    TileGrid tg;
    tg.tiles = (int)((ny + (int64_t)rs.G * rs.TY - 1) / ((int64_t)rs.G * rs.TY));
    int64_t cx = (nx * tg.tiles) / 1024;
    cx = cx < 1 ? 1 : (cx > 32 ? 32 : cx);
    tg.cx = (int)cx;
    tg.chunks = (int)((nx + cx - 1) / cx);
    const int64_t nb64 = (int64_t)tg.tiles * tg.chunks;
    TIP_REQUIRE(nb64 < (int64_t(1) << 31), "rk2 c32 mlp grid: too many blocks");
    tg.xcd_remap = nb64 % 8 == 0 ? 1 : 0;
    const size_t lds = (size_t)rs.G * (64 * rs.W) * (8 * rs.TY + 1) * sizeof(float);
    const dim3 blk(64, rs.W * rs.G);
#define TIP_LAUNCH_MG_Z(TYV, NM, CNT, ZNV)                                                     \
    hipLaunchKernelGGL((k_rk2_complex_rows<TYV, NM, MlpGridRadial, CNT, false, ZNV>),             \
                       dim3((int)nb64), blk, lds, st, phi1, phi2, pi1, pi2, o1, o2, o3, o4,       \
                       nullptr, nullptr, c, rad, zn, (int)nx, (int)ny, (int)nz, rs.W, tg,         \
                       noise_seed, vac, esc2)
// the Z_N term compiled out of launches that switch it off (-14 % at 512^3)
#define TIP_LAUNCH_MG_C(TYV, NM, CNT)                \
    do {                                             \
        if (zn.active)                               \
            TIP_LAUNCH_MG_Z(TYV, NM, CNT, true);     \
        else                                         \
            TIP_LAUNCH_MG_Z(TYV, NM, CNT, false);    \
    } while (0)
#define TIP_LAUNCH_MG(TYV, NM)                \
    do {                                      \
        if (vac)                              \
            TIP_LAUNCH_MG_C(TYV, NM, true);   \
        else                                  \
            TIP_LAUNCH_MG_C(TYV, NM, false);  \
    } while (0)
    if (rs.TY == 2) {
        if (hash) TIP_LAUNCH_MG(2, NOISE_HASH);
        else TIP_LAUNCH_MG(2, NOISE_NONE);
    } else {
        if (hash) TIP_LAUNCH_MG(1, NOISE_HASH);
        else TIP_LAUNCH_MG(1, NOISE_NONE);
    }
#undef TIP_LAUNCH_MG
#undef TIP_LAUNCH_MG_C
#undef TIP_LAUNCH_MG_Z
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

int launch_complex_mlp(const float* phi1, const float* phi2, const float* pi1, const float* pi2,
                       float* o1, float* o2, float* o3, float* o4, const StepScalars<float>& c,
                       const float* mlp_params, const tip_mlp_features* fm, double T,
                       const ZnParams<float>& zn, int64_t nx, int64_t ny, int64_t nz,
                       uint64_t noise_seed, int hash, unsigned long long* vac, float esc2,
                       hipStream_t st)
{
    int rc = check_fm(fm);
    if (rc) return rc;
    rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    TIP_REQUIRE(phi1 && phi2 && pi1 && pi2 && o1 && o2 && o3 && o4 && mlp_params,
                "rk2 c32 mlp: NULL pointer");
    TIP_REQUIRE(phi1 != o1 && phi2 != o2 && pi1 != o3 && pi2 != o4,
                "rk2 c32 mlp: output must not alias input");
    // one y row per thread: the network's constants and the matrix-core operands need the
    // registers a second row would take
    const RowShape rs = row_shape(ny, nz, 1, 8);
    TIP_REQUIRE(rs.ok, "rk2 c32 mlp: the inline-network step needs nz <= 512");
    MlpMfmaRadial rad;
    rad.P = mlp_params;
    rad.fm = make_fm(fm);
    rad.f2 = (float)((T - fm->T_mid) / fm->T_half);
    rad.g0v = rad.fm.g0(rad.f2);
    rad.cst = nullptr;
    TileGrid tg;
    tg.tiles = (int)((ny + (int64_t)rs.G - 1) / (int64_t)rs.G);
    int64_t cx = (nx * tg.tiles) / 2048;
    cx = cx < 1 ? 1 : (cx > 32 ? 32 : cx);
    tg.cx = (int)cx;
    tg.chunks = (int)((nx + cx - 1) / cx);
    const int64_t nb64 = (int64_t)tg.tiles * tg.chunks;
    TIP_REQUIRE(nb64 < (int64_t(1) << 31), "rk2 c32 mlp: too many blocks");
    tg.xcd_remap = nb64 % 8 == 0 ? 1 : 0;
    const size_t lds =
        ((size_t)rs.G * (64 * rs.W) * 9 + MlpMfmaRadial::kLdsFloats) * sizeof(float);
    const dim3 blk(64, rs.W * rs.G);
#define TIP_LAUNCH_MI(NM, CNT)                                                                 \
    hipLaunchKernelGGL((k_rk2_complex_rows<1, NM, MlpMfmaRadial, CNT>), dim3((int)nb64), blk,   \
                       lds, st, phi1, phi2, pi1, pi2, o1, o2, o3, o4, nullptr, nullptr, c, rad, \
                       zn, (int)nx, (int)ny, (int)nz, rs.W, tg, noise_seed, vac, esc2)
    if (hash) {
        if (vac) TIP_LAUNCH_MI(NOISE_HASH, true);
        else TIP_LAUNCH_MI(NOISE_HASH, false);
    } else {
        if (vac) TIP_LAUNCH_MI(NOISE_NONE, true);
        else TIP_LAUNCH_MI(NOISE_NONE, false);
    }
#undef TIP_LAUNCH_MI
    TIP_HIP(hipGetLastError());
    return TIP_OK;
}

}  // namespace tip

extern "C" int tip_rk2_fused_c32_mlp(const float* phi1, const float* phi2, const float* pi1,
                                     const float* pi2, float* phi1_out, float* phi2_out,
                                     float* pi1_out, float* pi2_out, double dt, double dx,
                                     double eta_eff, double mu, double inv_a2,
                                     const float* mlp_params, const tip_mlp_features* fm,
                                     double T, double zn_n, double zn_dv, double zn_active,
                                     int64_t nx, int64_t ny, int64_t nz, uint64_t noise_seed,
                                     double noise_scale, unsigned long long* vac_count,
                                     double vac_esc, void* stream)
{
    return launch_complex_mlp(phi1, phi2, pi1, pi2, phi1_out, phi2_out, pi1_out, pi2_out,
                              make_scalars<float>(dt, dx, eta_eff, mu, inv_a2, noise_scale),
                              mlp_params, fm, T, make_zn32(zn_n, zn_dv, zn_active), nx, ny, nz,
                              noise_seed, noise_scale != 0.0 ? 1 : 0, vac_count,
                              (float)(vac_esc * vac_esc), as_stream(stream));
}
