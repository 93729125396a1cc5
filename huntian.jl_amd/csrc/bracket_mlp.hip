// bracket_mlp.hip -- the learned shell-depth predictor: inner-shell features of each peak,
// and an MLP 18 -> 128 -> 128 -> 128 -> 3 (SiLU) run in BF16 on the matrix cores with FP32
// accumulation, with its backward pass and an Adam step for training on the GPU.
//
// The network has no counterpart in the reference.  It predicts, from the first 12 shells
// of a peak, the crossing shell and the gather extent (two shell indices, as log(1 + s))
// and a collapse logit.  hnt_shell_analyse may use the predicted extent to order peaks
// into wavefronts of equal depth; every lane still stops by the exact early-exit rule, so
// the prediction changes the schedule and never a result.
//
// Matrix-core tile: v_mfma_f32_16x16x32_bf16 (gfx950).  Lane l holds A[l % 16][8 (l / 16)
// + e] and B[8 (l / 16) + e][l % 16], e = 0..7, and receives C[4 (l / 16) + r][l % 16],
// r = 0..3.
#include "hnt_features.h"
#include "hnt_internal.h"

namespace {

using namespace hnt;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int NI = HNT_MLP_INPUTS, KI = 32, H = HNT_MLP_HIDDEN, NO = HNT_MLP_OUTPUTS, KO = 32;
constexpr int NFEAT_SHELLS = 12;

// parameter offsets in the flat FP32 vector (row-major W[k][o])
constexpr int64_t OFF_W1 = 0, OFF_B1 = OFF_W1 + NI * H, OFF_W2 = OFF_B1 + H,
                  OFF_B2 = OFF_W2 + H * H, OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H * H,
                  OFF_W4 = OFF_B3 + H, OFF_B4 = OFF_W4 + H * NO, NPARAM = OFF_B4 + NO;
static_assert(NPARAM == HNT_MLP_NPARAM, "parameter count");

__device__ __forceinline__ __bf16 to_bf16(float f) {
    // round to nearest even; NaN stays NaN
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return __builtin_bit_cast(__bf16, (uint16_t)0x7fc0);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __builtin_bit_cast(__bf16, (uint16_t)(u >> 16));
}
__device__ __forceinline__ float from_bf16(__bf16 h) {
    return __builtin_bit_cast(float, (uint32_t)__builtin_bit_cast(uint16_t, h) << 16);
}
__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// Epilogues of the GEMM
enum { EPI_STORE = 0, EPI_BIAS_SILU = 1, EPI_BIAS = 2, EPI_SILU_GRAD = 3 };

struct GemmArgs {
    const __bf16* A;  // A(m, k) = A[m * am + k * ak]
    const __bf16* B;  // B(k, n) = B[k * bk + n * bn]
    int64_t am, ak, bk, bn;
    int64_t M, N, K;  // M, N multiples of 16; K multiple of 32
    int epi;
    const float* bias;  // [N] (EPI_BIAS*)
    const float* zin;   // [M x N] pre-activation (EPI_SILU_GRAD), row stride N
    float* cf;          // fp32 out [M x N] (row stride N), may be null
    __bf16* cb;         // bf16 out [M x N] (row stride N), may be null
};

// One wavefront per 16 x 16 tile of C; 4 wavefronts per workgroup.
__global__ __launch_bounds__(256) void gemm_bf16(GemmArgs g) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t ntn = g.N / 16;
    const int64_t tile = (int64_t)blockIdx.x * 4 + wave;
    if (tile >= (g.M / 16) * ntn) return;
    const int64_t m0 = (tile / ntn) * 16, n0 = (tile % ntn) * 16;
    const int64_t ra = m0 + (l & 15), cb = n0 + (l & 15);
    const int kq = 8 * (l >> 4);
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int64_t k0 = 0; k0 < g.K; k0 += 32) {
        bf16x8 av, bv;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            av[e] = g.A[ra * g.am + (k0 + kq + e) * g.ak];
            bv[e] = g.B[(k0 + kq + e) * g.bk + cb * g.bn];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + 4 * (l >> 4) + r, n = cb;
        float v = acc[r];
        const int64_t o = m * g.N + n;
        if (g.epi == EPI_BIAS || g.epi == EPI_BIAS_SILU) v += g.bias[n];
        if (g.epi == EPI_BIAS_SILU) {
            if (g.cf) g.cf[o] = v;  // keep the pre-activation for the backward pass
            if (g.cb) g.cb[o] = to_bf16(silu(v));
        } else if (g.epi == EPI_SILU_GRAD) {
            const float d = v * silu_grad(g.zin[o]);
            if (g.cf) g.cf[o] = d;
            if (g.cb) g.cb[o] = to_bf16(d);
        } else {
            if (g.cf) g.cf[o] = v;
            if (g.cb) g.cb[o] = to_bf16(v);
        }
    }
}

// BF16 weight copies: W1t [H x KI], W2t/W3t [H x H] (transposed: [o][k]), W4t [KO x H]
// (rows >= NO zero), and the untransposed copies W2/W3/W4p [k][o] for the backward pass.
__global__ void pack_weights(const float* __restrict__ P, __bf16* __restrict__ w1t,
                             __bf16* __restrict__ w2t, __bf16* __restrict__ w3t,
                             __bf16* __restrict__ w4t, __bf16* __restrict__ w1,
                             __bf16* __restrict__ w2, __bf16* __restrict__ w3,
                             __bf16* __restrict__ w4) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < H * KI) {
        const int o = q / KI, k = q % KI;
        const float v = k < NI ? P[OFF_W1 + k * H + o] : 0.0f;
        w1t[q] = to_bf16(v);
        w1[k * H + o] = to_bf16(v);
    }
    if (q < H * H) {
        const int o = q / H, k = q % H;
        w2t[q] = to_bf16(P[OFF_W2 + k * H + o]);
        w3t[q] = to_bf16(P[OFF_W3 + k * H + o]);
        w2[q] = to_bf16(P[OFF_W2 + q]);
        w3[q] = to_bf16(P[OFF_W3 + q]);
    }
    if (q < KO * H) {
        const int o = q / H, k = q % H;
        const float v = o < NO ? P[OFF_W4 + k * NO + o] : 0.0f;
        w4t[q] = to_bf16(v);
        w4[k * KO + o] = to_bf16(v);
    }
}

// Features normalised with (x - mean) / scale into bf16 rows of KI (zero padded), rows
// beyond B zero.
__global__ void pack_inputs(const float* __restrict__ x, int64_t B, int64_t Bp,
                            const float* __restrict__ norm, __bf16* __restrict__ x0) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Bp * KI) return;
    const int64_t b = q / KI;
    const int k = (int)(q % KI);
    float v = 0.0f;
    if (b < B && k < NI) v = (x[b * NI + k] - norm[k]) / norm[NI + k];
    x0[q] = to_bf16(v);
}

__global__ void bias_out(const float* __restrict__ y, int64_t B, const float* __restrict__ b4,
                         float* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= B * NO) return;
    const int64_t b = q / NO;
    const int o = (int)(q % NO);
    out[q] = y[b * KO + o] + b4[o];
}

// Loss per sample: (y0 - t0)^2 + (y1 - t1)^2 + BCE(y2, t2); mean over B.  dY = dL/dy
// (bf16, [Bp x KO], zero outside), block sums of the loss into part[blockIdx].
__global__ __launch_bounds__(256) void loss_grad(const float* __restrict__ y,
                                                 const float* __restrict__ b4,
                                                 const float* __restrict__ t, int64_t B,
                                                 int64_t Bp, __bf16* __restrict__ dy,
                                                 float* __restrict__ dyf,
                                                 double* __restrict__ part) {
    __shared__ double red[256];
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double loss = 0.0;
    if (b < Bp) {
        float g[KO];
#pragma unroll
        for (int o = 0; o < KO; ++o) g[o] = 0.0f;
        if (b < B) {
            const float y0 = y[b * KO + 0] + b4[0], y1 = y[b * KO + 1] + b4[1],
                        y2 = y[b * KO + 2] + b4[2];
            const float t0 = t[b * NO + 0], t1 = t[b * NO + 1], t2 = t[b * NO + 2];
            const float inv = 1.0f / (float)B;
            g[0] = 2.0f * (y0 - t0) * inv;
            g[1] = 2.0f * (y1 - t1) * inv;
            const float s = 1.0f / (1.0f + expf(-y2));
            g[2] = (s - t2) * inv;
            // BCE with logits, stable form
            const float bce = fmaxf(y2, 0.0f) - y2 * t2 + log1pf(expf(-fabsf(y2)));
            loss = (double)((y0 - t0) * (y0 - t0) + (y1 - t1) * (y1 - t1) + bce);
        }
#pragma unroll
        for (int o = 0; o < KO; ++o) {
            dy[b * KO + o] = to_bf16(g[o]);
            dyf[b * KO + o] = g[o];
        }
    }
    red[threadIdx.x] = loss;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Column sums of an fp32 [Bp x N] matrix (bias gradients): one thread per column.
__global__ void col_sums(const float* __restrict__ d, int64_t Bp, int N, float* __restrict__ out,
                         int nout) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nout) return;
    float s = 0.0f;
    for (int64_t b = 0; b < Bp; ++b) s += d[b * N + n];
    out[n] = s;
}

// Scatter the weight gradients (fp32 GEMM outputs) into the flat gradient vector.
__global__ void gather_grads(const float* __restrict__ gw1, const float* __restrict__ gw2,
                             const float* __restrict__ gw3, const float* __restrict__ gw4,
                             float* __restrict__ grad) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < NI * H) grad[OFF_W1 + q] = gw1[q];           // gw1 is [KI x H], rows < NI
    if (q < H * H) {
        grad[OFF_W2 + q] = gw2[q];
        grad[OFF_W3 + q] = gw3[q];
    }
    if (q < H * NO) {
        const int k = q / NO, o = q % NO;
        grad[OFF_W4 + q] = gw4[k * KO + o];              // gw4 is [H x KO]
    }
}

__global__ void adam_step(float* __restrict__ p, const float* __restrict__ g,
                          float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                          float b1, float b2, float eps, float c1, float c2) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.0f - b1) * gq;
    const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

// ------------------------------------------------------------------ fused inference
// All four layers in one kernel: the BF16 weights (80 KB) are staged in LDS once per
// workgroup; each wavefront then pushes groups of MT = 4 tiles of 16 samples through the
// network, keeping their 64 x 128 activations in its own LDS tile.  Every B operand read
// from LDS serves the MT tiles' matrix products.  Inputs are normalised and rounded to BF16
// on the fly (the A operand of layer 1 comes straight from registers).
constexpr int FW = 4;  // wavefronts per workgroup
constexpr int MT = 4;  // 16-sample tiles per wavefront pass

// Inference activations: SiLU with the hardware exp2 and reciprocal (a few FP32 ulps from
// the correctly rounded form, far below the BF16 rounding that follows) and the hardware
// round-to-nearest-even BF16 conversion.  The pass is bound by these per-element VALU
// operations, not by the matrix products.
__device__ __forceinline__ float silu_fast(float z) {
    return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ __bf16 act_bf16(float f) { return (__bf16)f; }

// the wave's LDS writes are complete and visible to its other lanes before later reads
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(64 * FW) void mlp_forward_fused(
    const float* __restrict__ x, int64_t B, const float* __restrict__ norm,
    const __bf16* __restrict__ w1t, const __bf16* __restrict__ w2t,
    const __bf16* __restrict__ w3t, const __bf16* __restrict__ w4t,
    const float* __restrict__ P, float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) __bf16 sw1[H * KI], sw2[H * H], sw3[H * H],
        sw4[KO * H];
    __shared__ __attribute__((aligned(16))) __bf16 act[FW][MT * 16 * H];
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    // stage the weights (16-byte chunks)
    for (int q = tid; q < H * KI / 8; q += 64 * FW)
        ((uint4*)sw1)[q] = ((const uint4*)w1t)[q];
    for (int q = tid; q < H * H / 8; q += 64 * FW) {
        ((uint4*)sw2)[q] = ((const uint4*)w2t)[q];
        ((uint4*)sw3)[q] = ((const uint4*)w3t)[q];
    }
    for (int q = tid; q < KO * H / 8; q += 64 * FW)
        ((uint4*)sw4)[q] = ((const uint4*)w4t)[q];
    __syncthreads();
    const float* b1 = P + OFF_B1;
    const float* b2 = P + OFF_B2;
    const float* b3 = P + OFF_B3;
    const float* b4 = P + OFF_B4;
    __bf16* A = act[wave];  // tile mt's 16 x H activations at A + mt * 16 * H
    const int rl = l & 15, kq = 8 * (l >> 4);
    const int64_t ngroups = (B + 16 * MT - 1) / (16 * MT);
    for (int64_t gi = (int64_t)blockIdx.x * FW + wave; gi < ngroups;
         gi += (int64_t)gridDim.x * FW) {
        const int64_t row0 = gi * 16 * MT;
        // layer 1: A operands from x (k = kq..kq+7 of this lane's row of each tile)
        bf16x8 a0[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int64_t row = row0 + mt * 16 + rl;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = kq + e;
                float v = 0.0f;
                if (row < B && k < NI) v = (x[row * NI + k] - norm[k]) / norm[NI + k];
                a0[mt][e] = (__bf16)v;
            }
        }
#pragma unroll
        for (int nt = 0; nt < H / 16; ++nt) {
            const bf16x8 bv = *(const bf16x8*)&sw1[(nt * 16 + rl) * KI + kq];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0[mt], bv, c, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = 4 * (l >> 4) + r, n = nt * 16 + rl;
                    A[mt * 16 * H + m * H + n] = act_bf16(silu_fast(c[r] + b1[n]));
                }
            }
        }
        wave_lds_fence();
        // layers 2 and 3: 128 -> 128; all A operands in registers before the tile is
        // overwritten in place
#pragma unroll
        for (int layer = 0; layer < 2; ++layer) {
            const __bf16* W = layer == 0 ? sw2 : sw3;
            const float* bb = layer == 0 ? b2 : b3;
            bf16x8 av[MT][H / 32];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int ks = 0; ks < H / 32; ++ks)
                    av[mt][ks] = *(const bf16x8*)&A[mt * 16 * H + rl * H + ks * 32 + kq];
            wave_lds_fence();
#pragma unroll
            for (int nt = 0; nt < H / 16; ++nt) {
                f32x4 c[MT];
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) c[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int ks = 0; ks < H / 32; ++ks) {
                    const bf16x8 bv = *(const bf16x8*)&W[(nt * 16 + rl) * H + ks * 32 + kq];
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
                        c[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[mt][ks], bv, c[mt],
                                                                        0, 0, 0);
                }
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int m = 4 * (l >> 4) + r, n = nt * 16 + rl;
                        A[mt * 16 * H + m * H + n] = act_bf16(silu_fast(c[mt][r] + bb[n]));
                    }
            }
            wave_lds_fence();
        }
        // layer 4: 128 -> 16 (3 used)
        f32x4 c[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) c[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < H / 32; ++ks) {
            const bf16x8 bv = *(const bf16x8*)&sw4[rl * H + ks * 32 + kq];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const bf16x8 av = *(const bf16x8*)&A[mt * 16 * H + rl * H + ks * 32 + kq];
                c[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c[mt], 0, 0, 0);
            }
        }
        if (rl < NO) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t m = row0 + mt * 16 + 4 * (l >> 4) + r;
                    if (m < B) out[m * NO + rl] = c[mt][r] + b4[rl];
                }
        }
        wave_lds_fence();
    }
}

// ------------------------------------------------------------------ features
// Per peak, from delta over shells 1..12:
//  f[0..11]  mean delta of shell s (in-bounds cells; 0 if none)
//  f[12..15] e, p of the delta quadrupole sum_cells delta d_L d_K over shells 2..6, then
//            over shells 2..12 (e = (l3 - l1) / (2 F), p = (l3 + l1 - 2 l2) / (2 F),
//            F = l1 + l2 + l3; both 0 if F <= 0)
//  f[16]     mean delta over shells 1..12
//  f[17]     Rf / alatt
struct FeatArgs {
    const float* delta;
    int n;
    int64_t s1, s2;
    const int32_t* pi;
    const int32_t* pj;
    const int32_t* pk;
    const double* Rf;
    int64_t npeaks;
    const int4* cells;
    const int32_t* shell_start;
    const int32_t* shell_count;
    double alatt;
    float* out;
};



__global__ __launch_bounds__(256) void features_kernel(FeatArgs a) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.npeaks) return;
    const int ci = a.pi[q], cj = a.pj[q], ck = a.pk[q];
    const int64_t c0 = (int64_t)(ci - 1) + (cj - 1) * a.s1 + (ck - 1) * a.s2;
    auto load = [&](int64_t off) -> float { return a.delta[off]; };
    peak_features(a.cells, a.shell_start, a.shell_count, ci, cj, ck, c0, a.n, a.s1, a.s2,
                  a.Rf[q] / a.alatt, load, a.out + q * NI);
}

}  // namespace

// ------------------------------------------------------------------ host side

namespace {

hnt_status gemm(hipStream_t st, const GemmArgs& g) {
    HNT_REQUIRE(g.M % 16 == 0 && g.N % 16 == 0 && g.K % 32 == 0, "mlp gemm: bad shape");
    const int64_t tiles = (g.M / 16) * (g.N / 16);
    hipLaunchKernelGGL(gemm_bf16, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, st, g);
    HNT_HIP(hipGetLastError());
    return HNT_OK;
}

struct Views {
    __bf16 *w1t, *w2t, *w3t, *w4t, *w1, *w2, *w3, *w4;
    __bf16 *x0, *h1, *h2, *h3, *dy, *dz[2];
    float *z1, *z2, *z3, *y, *dyf, *dzf[2], *gw1, *gw2, *gw3, *gw4;
    double* part;
};

hnt_status prepare(hnt_ctx* ctx, int64_t B, Views& v, bool training) {
    const int64_t Bp = ((B + 31) / 32) * 32;
    HNT_REQUIRE(B >= 1 && Bp <= (1LL << 30), "mlp: batch size %lld", (long long)B);
    MlpScratch& S = ctx->mlp;
    const size_t wcount = (size_t)H * KI * 2 + (size_t)H * H * 4 + (size_t)KO * H * 2;
    HNT_TRY(reserve(ctx, S.wt, wcount * 2));
    HNT_TRY(reserve(ctx, S.x0, (size_t)Bp * KI * 2));
    if (training) HNT_TRY(reserve(ctx, S.z, (size_t)Bp * H * 4 * 3));
    HNT_TRY(reserve(ctx, S.h, (size_t)Bp * H * 2 * 3));
    HNT_TRY(reserve(ctx, S.y, (size_t)Bp * KO * 4));
    if (training) {
        HNT_TRY(reserve(ctx, S.dy, (size_t)Bp * KO * 2));
        HNT_TRY(reserve(ctx, S.dyf, (size_t)Bp * KO * 4));
        HNT_TRY(reserve(ctx, S.dz, (size_t)Bp * H * 12));
    }
    HNT_TRY(reserve(ctx, S.gw, ((size_t)KI * H + 2 * (size_t)H * H + (size_t)H * KO) * 4));
    HNT_TRY(reserve(ctx, S.part, (size_t)((Bp + 255) / 256) * 8));
    S.Bp = Bp;
    __bf16* w = (__bf16*)S.wt.p;
    v.w1t = w;
    v.w2t = v.w1t + H * KI;
    v.w3t = v.w2t + H * H;
    v.w4t = v.w3t + H * H;
    v.w1 = v.w4t + KO * H;
    v.w2 = v.w1 + KI * H;
    v.w3 = v.w2 + H * H;
    v.w4 = v.w3 + H * H;
    v.x0 = (__bf16*)S.x0.p;
    v.z1 = (float*)S.z.p;
    v.z2 = v.z1 + Bp * H;
    v.z3 = v.z2 + Bp * H;
    v.h1 = (__bf16*)S.h.p;
    v.h2 = v.h1 + Bp * H;
    v.h3 = v.h2 + Bp * H;
    v.y = (float*)S.y.p;
    v.dy = (__bf16*)S.dy.p;
    v.dyf = (float*)S.dyf.p;
    v.dz[0] = (__bf16*)S.dz.p;
    v.dz[1] = v.dz[0] + Bp * H;
    v.dzf[0] = (float*)(v.dz[1] + Bp * H);
    v.dzf[1] = v.dzf[0] + Bp * H;
    v.gw1 = (float*)S.gw.p;
    v.gw2 = v.gw1 + KI * H;
    v.gw3 = v.gw2 + H * H;
    v.gw4 = v.gw3 + H * H;
    v.part = (double*)S.part.p;
    return HNT_OK;
}

// forward through the four layers; y = H3 W4 (bias of the last layer added by the caller).
// keep_z: store the pre-activations for a backward pass.
hnt_status forward(hnt_ctx* ctx, const hnt_mlp* net, const float* x, int64_t B, Views& v,
                   bool keep_z) {
    hipStream_t st = ctx->stream;
    const int64_t Bp = ctx->mlp.Bp;
    hipLaunchKernelGGL(pack_weights, dim3((H * H + 255) / 256), dim3(256), 0, st, net->params,
                       v.w1t, v.w2t, v.w3t, v.w4t, v.w1, v.w2, v.w3, v.w4);
    hipLaunchKernelGGL(pack_inputs, dim3((unsigned)((Bp * KI + 255) / 256)), dim3(256), 0, st, x,
                       B, Bp, net->norm, v.x0);
    const float* P = net->params;
    // layer 1: [Bp x KI] x W1 (B(k, n) = w1t[n * KI + k])
    GemmArgs g{};
    g = GemmArgs{v.x0, v.w1t, KI, 1, 1, KI, Bp, H, KI, EPI_BIAS_SILU, P + OFF_B1, nullptr,
                 keep_z ? v.z1 : nullptr, v.h1};
    HNT_TRY(gemm(st, g));
    g = GemmArgs{v.h1, v.w2t, H, 1, 1, H, Bp, H, H, EPI_BIAS_SILU, P + OFF_B2, nullptr,
                 keep_z ? v.z2 : nullptr, v.h2};
    HNT_TRY(gemm(st, g));
    g = GemmArgs{v.h2, v.w3t, H, 1, 1, H, Bp, H, H, EPI_BIAS_SILU, P + OFF_B3, nullptr,
                 keep_z ? v.z3 : nullptr, v.h3};
    HNT_TRY(gemm(st, g));
    g = GemmArgs{v.h3, v.w4t, H, 1, 1, H, Bp, KO, H, EPI_STORE, nullptr, nullptr, v.y, nullptr};
    HNT_TRY(gemm(st, g));
    return HNT_OK;
}

hnt_status check_net(const hnt_mlp* net) {
    HNT_REQUIRE(net && net->params && net->norm, "mlp: NULL network");
    return HNT_OK;
}

}  // namespace

using namespace hnt;

namespace hnt {

// Network outputs for B rows of features, in chunks of 1M rows (bf16 activations only).
hnt_status mlp_predict(hnt_ctx* ctx, const hnt_mlp* net, const float* x, int64_t B, float* out) {
    if (B <= 0) return HNT_OK;
    MlpScratch& S = ctx->mlp;
    const size_t wcount = (size_t)H * KI * 2 + (size_t)H * H * 4 + (size_t)KO * H * 2;
    HNT_TRY(reserve(ctx, S.wt, wcount * 2));
    __bf16* w = (__bf16*)S.wt.p;
    __bf16 *w1t = w, *w2t = w1t + H * KI, *w3t = w2t + H * H, *w4t = w3t + H * H;
    __bf16 *w1 = w4t + KO * H, *w2 = w1 + KI * H, *w3 = w2 + H * H, *w4 = w3 + H * H;
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(pack_weights, dim3((H * H + 255) / 256), dim3(256), 0, st, net->params,
                       w1t, w2t, w3t, w4t, w1, w2, w3, w4);
    // one workgroup per CU (80 KB of weights and 64 KB of activations in LDS)
    const int64_t ngroups = (B + 16 * MT - 1) / (16 * MT);
    int64_t nblocks = (ngroups + FW - 1) / FW;
    const int64_t cap = (int64_t)ctx->num_cus;
    if (nblocks > cap) nblocks = cap;
    hipLaunchKernelGGL(mlp_forward_fused, dim3((unsigned)nblocks), dim3(64 * FW), 0, st, x, B,
                       net->norm, w1t, w2t, w3t, w4t, net->params, out);
    HNT_HIP(hipGetLastError());
    return HNT_OK;
}

// Features of the peaks of batch (tables already uploaded).
hnt_status bracket_features(hnt_ctx* ctx, const hnt_fields* fields, const hnt_batch* batch,
                            double alatt, float* out) {
    if (batch->npeaks == 0) return HNT_OK;
    FeatArgs a{};
    a.delta = fields->delta;
    a.n = fields->n1;
    a.s1 = fields->n1;
    a.s2 = (int64_t)fields->n1 * fields->n1;
    a.pi = batch->pi;
    a.pj = batch->pj;
    a.pk = batch->pk;
    a.Rf = batch->Rf;
    a.npeaks = batch->npeaks;
    a.cells = (const int4*)ctx->tables.cells.p;
    a.shell_start = (const int32_t*)ctx->tables.start.p;
    a.shell_count = (const int32_t*)ctx->tables.count.p;
    a.alatt = alatt;
    a.out = out;
    hipLaunchKernelGGL(features_kernel, dim3((unsigned)((batch->npeaks + 255) / 256)), dim3(256),
                       0, ctx->stream, a);
    HNT_HIP(hipGetLastError());
    return HNT_OK;
}

}  // namespace hnt

extern "C" hnt_status hnt_mlp_forward(hnt_ctx* ctx, const hnt_mlp* net, const float* x, int64_t B,
                                      float* out) {
    HNT_TRY(begin_call(ctx));
    HNT_TRY(check_net(net));
    HNT_REQUIRE(x && out && B >= 1, "hnt_mlp_forward: NULL array or empty batch");
    HNT_TRY(mlp_predict(ctx, net, x, B, out));
    return finish_call(ctx, "hnt_mlp_forward");
}

// Gradient of the mean loss with respect to the flat parameters (grad, device [NPARAM]);
// *loss gets the mean loss (host).
static hnt_status mlp_grad(hnt_ctx* ctx, const hnt_mlp* net, const float* x, const float* t,
                           int64_t B, float* grad, double* loss) {
    Views v;
    HNT_TRY(prepare(ctx, B, v, true));
    HNT_TRY(forward(ctx, net, x, B, v, true));
    hipStream_t st = ctx->stream;
    const int64_t Bp = ctx->mlp.Bp;
    const float* P = net->params;
    const unsigned nb = (unsigned)((Bp + 255) / 256);
    hipLaunchKernelGGL(loss_grad, dim3(nb), dim3(256), 0, st, v.y, P + OFF_B4, t, B, Bp, v.dy,
                       v.dyf, v.part);
    GemmArgs g{};
    // layer 4: dW4 = H3^T dY [H x KO]; dZ3 = (dY W4^T) * silu'(Z3)
    g = GemmArgs{v.h3, v.dy, 1, H, KO, 1, H, KO, Bp, EPI_STORE, nullptr, nullptr, v.gw4, nullptr};
    HNT_TRY(gemm(st, g));
    hipLaunchKernelGGL(col_sums, dim3(1), dim3(64), 0, st, v.dyf, Bp, KO, grad + OFF_B4, NO);
    g = GemmArgs{v.dy, v.w4, KO, 1, 1, KO, Bp, H, KO, EPI_SILU_GRAD, nullptr, v.z3, v.dzf[0],
                 v.dz[0]};
    HNT_TRY(gemm(st, g));
    // layer 3: dW3 = H2^T dZ3; dZ2 = (dZ3 W3^T) * silu'(Z2)
    g = GemmArgs{v.h2, v.dz[0], 1, H, H, 1, H, H, Bp, EPI_STORE, nullptr, nullptr, v.gw3,
                 nullptr};
    HNT_TRY(gemm(st, g));
    hipLaunchKernelGGL(col_sums, dim3(1), dim3(128), 0, st, v.dzf[0], Bp, H, grad + OFF_B3, H);
    g = GemmArgs{v.dz[0], v.w3, H, 1, 1, H, Bp, H, H, EPI_SILU_GRAD, nullptr, v.z2, v.dzf[1],
                 v.dz[1]};
    HNT_TRY(gemm(st, g));
    // layer 2: dW2 = H1^T dZ2; dZ1 = (dZ2 W2^T) * silu'(Z1)
    g = GemmArgs{v.h1, v.dz[1], 1, H, H, 1, H, H, Bp, EPI_STORE, nullptr, nullptr, v.gw2,
                 nullptr};
    HNT_TRY(gemm(st, g));
    hipLaunchKernelGGL(col_sums, dim3(1), dim3(128), 0, st, v.dzf[1], Bp, H, grad + OFF_B2, H);
    g = GemmArgs{v.dz[1], v.w2, H, 1, 1, H, Bp, H, H, EPI_SILU_GRAD, nullptr, v.z1, v.dzf[0],
                 v.dz[0]};
    HNT_TRY(gemm(st, g));
    // layer 1: dW1 = X0^T dZ1
    g = GemmArgs{v.x0, v.dz[0], 1, KI, H, 1, KI, H, Bp, EPI_STORE, nullptr, nullptr, v.gw1,
                 nullptr};
    HNT_TRY(gemm(st, g));
    hipLaunchKernelGGL(col_sums, dim3(1), dim3(128), 0, st, v.dzf[0], Bp, H, grad + OFF_B1, H);
    hipLaunchKernelGGL(gather_grads, dim3((H * H + 255) / 256), dim3(256), 0, st, v.gw1, v.gw2,
                       v.gw3, v.gw4, grad);
    if (loss) {
        std::vector<double> part(nb);
        HNT_HIP(hipMemcpyAsync(part.data(), v.part, nb * sizeof(double), hipMemcpyDeviceToHost,
                               st));
        HNT_HIP(hipStreamSynchronize(st));
        double s = 0.0;
        for (double q : part) s += q;
        *loss = s / (double)B;
    }
    return HNT_OK;
}

extern "C" hnt_status hnt_mlp_gradient(hnt_ctx* ctx, const hnt_mlp* net, const float* x,
                                       const float* target, int64_t B, float* grad,
                                       double* loss) {
    HNT_TRY(begin_call(ctx));
    HNT_TRY(check_net(net));
    HNT_REQUIRE(x && target && grad, "hnt_mlp_gradient: NULL array");
    HNT_TRY(mlp_grad(ctx, net, x, target, B, grad, loss));
    return finish_call(ctx, "hnt_mlp_gradient");
}

extern "C" hnt_status hnt_mlp_train_step(hnt_ctx* ctx, hnt_mlp* net, const float* x,
                                         const float* target, int64_t B, float lr,
                                         double* loss) {
    HNT_TRY(begin_call(ctx));
    HNT_TRY(check_net(net));
    HNT_REQUIRE(x && target && net->adam_m && net->adam_v, "hnt_mlp_train_step: NULL array");
    HNT_REQUIRE(lr > 0.0f, "hnt_mlp_train_step: lr must be > 0");
    HNT_TRY(reserve(ctx, ctx->mlp.gradv, (size_t)NPARAM * 4));
    float* grad = (float*)ctx->mlp.gradv.p;
    HNT_TRY(mlp_grad(ctx, net, x, target, B, grad, loss));
    net->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)net->step), c2 = 1.0f - powf(b2, (float)net->step);
    hipLaunchKernelGGL(adam_step, dim3((unsigned)((NPARAM + 255) / 256)), dim3(256), 0,
                       ctx->stream, net->params, grad, net->adam_m, net->adam_v, NPARAM, lr, b1,
                       b2, 1e-8f, c1, c2);
    return finish_call(ctx, "hnt_mlp_train_step");
}

extern "C" hnt_status hnt_bracket_features(hnt_ctx* ctx, const hnt_fields* fields,
                                           const hnt_batch* batch,
                                           const hnt_shell_tables* tables, double alatt,
                                           float* out) {
    HNT_TRY(begin_call(ctx));
    HNT_REQUIRE(fields && batch && tables && out, "hnt_bracket_features: NULL argument");
    HNT_REQUIRE(tables->nshells > NFEAT_SHELLS, "hnt_bracket_features: needs > %d shells",
                NFEAT_SHELLS);
    HNT_REQUIRE(fields->n1 == fields->n2 && fields->n2 == fields->n3 && fields->delta &&
                    fields->etax && fields->etay && fields->etaz,
                "hnt_bracket_features: needs a cubic grid with delta and eta");
    HNT_TRY(upload_tables(ctx, tables));
    HNT_TRY(bracket_features(ctx, fields, batch, alatt, out));
    return finish_call(ctx, "hnt_bracket_features");
}
