// Training and evaluation of the far-field model of mlp.h: forward, backward and Adam,
// hand-written on BF16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32 accumulation.
//
// Loss: L = (1/N) sum_s sum_o (y_so - t_so)^2 over the N samples of a batch.
// k_mlp_grad  one wave per 64 samples: forward (pre-activations Z and activations H kept in
//             LDS), backward through the layers (dH = W^T dZ on MFMA with the transposed
//             weights as A), weight gradients dW = dZ^T H on MFMA with the samples as the K
//             dimension, bias gradients by lane reductions; writes its partial gradient
// k_mlp_reduce  sums the partial gradients of the workgroups in a fixed order (deterministic)
// k_adam      Adam step on the FP32 master parameters
// k_mlp_eval  inference, four waves per workgroup sharing the BF16 weights
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "bl_internal.h"
#include "mlp.h"

namespace bl {
namespace {

constexpr int NP1 = MLP_NPARAM + 1;  // parameters + the loss sum

struct GradLds {
    MlpLds L;
    short w2t[MLP_H * MLP_H];  // W2^T, W3^T: rows = input units
    short w3t[MLP_H * MLP_H];
    short w4t[MLP_H * 16];     // W4^T [64][16], columns 2..15 zero
    short xs[64 * 16];         // inputs [sample][16]
    short hs[3][64 * MLP_H];   // H1, H2, H3 [sample][unit]
    float zs[3][64 * MLP_H];   // Z1, Z2, Z3 [sample][unit]
    short dzs[64 * MLP_H];     // dZ of the layer being differentiated [sample][unit]
    short dz4s[64 * 16];       // dZ4 [sample][16]
};

// element (s, o) of a tile held in the MFMA result layout
__device__ __forceinline__ int tile_s(int sb, int lane) { return sb * 16 + (lane & 15); }
__device__ __forceinline__ int tile_o(int ob, int lane, int i) { return ob * 16 + 4 * (lane >> 4) + i; }

// out[ob'][sb] = A (rows from ld-strided row-major M) * B[kb][sb], no bias
template <int KB>
__device__ __forceinline__ void mm_tiles(const short *M, int ld, const bf16x4 (&B)[4][4], f32x4 (&out)[4][4],
                                         int lane) {
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
        bf16x4 a[KB];
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) a[kb] = frag_a(M, ld, ob * 16, kb * 16, lane);
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) acc = mfma16(a[kb], B[kb][sb], acc);
            out[ob][sb] = acc;
        }
    }
}

// G[o][in] = sum over the wave's 64 samples of dZ[s][o] * H[s][in]; rows o < orows and
// columns in < icols are written to out (row stride ldo).  A = dZ^T and B = H are read from
// LDS with the sample as the K index.
__device__ void wgrad(const short *dZ, int ldz, int OB, const short *H, int ldh, int IB, float *out, int ldo,
                      int orows, int icols, int lane) {
    for (int ob = 0; ob < OB; ++ob) {
        for (int ib = 0; ib < IB; ++ib) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int sb = 0; sb < 4; ++sb) {
                bf16x4 a, b;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int s = sb * 16 + 4 * (lane >> 4) + i;
                    a[i] = dZ[s * ldz + ob * 16 + (lane & 15)];
                    b[i] = H[s * ldh + ib * 16 + (lane & 15)];
                }
                acc = mfma16(a, b, acc);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int o = tile_o(ob, lane, i), in = ib * 16 + (lane & 15);
                if (o < orows && in < icols) out[o * ldo + in] = acc[i];
            }
        }
    }
}

// sum over the 16 lanes of each lane group (same lane/16)
__device__ __forceinline__ float group16_sum(float v) {
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// db[o] = sum_s dZ[s][o] from the result-layout tiles
__device__ __forceinline__ void bgrad(const f32x4 (&dZ)[4][4], float *out, int lane) {
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float v = 0.f;
#pragma unroll
            for (int sb = 0; sb < 4; ++sb) v += dZ[ob][sb][i];
            v = group16_sum(v);
            if ((lane & 15) == 0) out[tile_o(ob, lane, i)] = v;
        }
    }
}

// dZ = dH * gelu'(Z); stores dZ (BF16) in dzs and returns its B fragments
__device__ __forceinline__ void act_back(const f32x4 (&dH)[4][4], const float *zs, short *dzs, f32x4 (&dZ)[4][4],
                                         bf16x4 (&dZb)[4][4], int lane) {
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
            f32x4 d;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int s = tile_s(sb, lane), o = tile_o(ob, lane, i);
                d[i] = dH[ob][sb][i] * gelu_grad(zs[s * MLP_H + o]);
            }
            dZ[ob][sb] = d;
            dZb[ob][sb] = pack4(d);
#pragma unroll
            for (int i = 0; i < 4; ++i) dzs[tile_s(sb, lane) * MLP_H + tile_o(ob, lane, i)] = dZb[ob][sb][i];
        }
    }
}

__device__ __forceinline__ void store_layer(const f32x4 (&Z)[4][4], const bf16x4 (&H)[4][4], float *zs, short *hs,
                                            int lane) {
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
        for (int sb = 0; sb < 4; ++sb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int at = tile_s(sb, lane) * MLP_H + tile_o(ob, lane, i);
                zs[at] = Z[ob][sb][i];
                hs[at] = H[ob][sb][i];
            }
}

// One wave: partial gradient of the samples idx[base + wg*64 + 0..63] (those below count).
__global__ __launch_bounds__(64) void k_mlp_grad(const float *__restrict__ P, const float *__restrict__ X,
                                                 const float *__restrict__ T, const int *__restrict__ idx, int base,
                                                 int count, float inv_n, float *__restrict__ part) {
    extern __shared__ __align__(16) unsigned char smem[];
    GradLds &S = *reinterpret_cast<GradLds *>(smem);
    const int lane = threadIdx.x;
    mlp_load(S.L, P, lane, 64);
    for (int k = lane; k < MLP_H * MLP_H; k += 64) {
        const int r = k >> 6, c = k & 63;  // transposed: [in r][out c]
        S.w2t[k] = f2bf(P[MLP_W2 + c * MLP_H + r]);
        S.w3t[k] = f2bf(P[MLP_W3 + c * MLP_H + r]);
    }
    for (int k = lane; k < MLP_H * 16; k += 64) {
        const int r = k >> 4, c = k & 15;
        S.w4t[k] = c < MLP_OUT ? f2bf(P[MLP_W4 + c * MLP_H + r]) : (short)0;
    }
    const int s0 = blockIdx.x * 64;
    auto sample = [&](int s) { return s0 + s < count ? (idx ? idx[base + s0 + s] : base + s0 + s) : -1; };
    for (int k = lane; k < 64 * 16; k += 64) {
        const int s = k >> 4, f = k & 15;
        const int g = sample(s);
        S.xs[k] = (g >= 0 && f < MLP_IN) ? f2bf(X[(size_t)g * MLP_IN + f]) : (short)0;
    }
    __syncthreads();

    // forward
    bf16x4 Xb[4][4], Ha[4][4], Hb[4][4];
    f32x4 Z[4][4];
#pragma unroll
    for (int sb = 0; sb < 4; ++sb)
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            bf16x4 v = {0, 0, 0, 0};
            if (kb == 0) v = *reinterpret_cast<const bf16x4 *>(S.xs + tile_s(sb, lane) * 16 + 4 * (lane >> 4));
            Xb[kb][sb] = v;
        }
    mlp_hidden<1>(S.L.w1, 16, S.L.b1, Xb, Z, Ha, lane);
    store_layer(Z, Ha, S.zs[0], S.hs[0], lane);
    mlp_hidden<4>(S.L.w2, MLP_H, S.L.b2, Ha, Z, Hb, lane);
    store_layer(Z, Hb, S.zs[1], S.hs[1], lane);
    mlp_hidden<4>(S.L.w3, MLP_H, S.L.b3, Hb, Z, Ha, lane);
    store_layer(Z, Ha, S.zs[2], S.hs[2], lane);
    f32x4 Y[4];
    mlp_out(S.L, Ha, Y, lane);

    // loss and dZ4 = 2 (y - t) / N (lanes 0..15 hold outputs 0..3 of sample sb*16 + lane)
    float lsum = 0.f;
    f32x4 dZ4[4];
    bf16x4 dZ4b[4][4];
#pragma unroll
    for (int sb = 0; sb < 4; ++sb) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        const int g = sample(tile_s(sb, lane));
        if (lane < 16 && g >= 0) {
            const float e0 = Y[sb][0] - T[2 * (size_t)g], e1 = Y[sb][1] - T[2 * (size_t)g + 1];
            lsum += e0 * e0 + e1 * e1;
            d[0] = 2.f * e0 * inv_n;
            d[1] = 2.f * e1 * inv_n;
        }
        dZ4[sb] = d;
        dZ4b[0][sb] = pack4(d);
        dZ4b[1][sb] = dZ4b[2][sb] = dZ4b[3][sb] = bf16x4{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 4; ++i) S.dz4s[tile_s(sb, lane) * 16 + 4 * (lane >> 4) + i] = dZ4b[0][sb][i];
    }
    __syncthreads();
    float *G = part + (size_t)blockIdx.x * NP1;
    // layer 4
    wgrad(S.dz4s, 16, 1, S.hs[2], MLP_H, 4, G + MLP_W4, MLP_H, MLP_OUT, MLP_H, lane);
#pragma unroll
    for (int o = 0; o < MLP_OUT; ++o) {
        float v = 0.f;
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) v += dZ4[sb][o];
        v = group16_sum(v);  // lanes 0..15 carry the outputs
        if (lane == 0) G[MLP_B4 + o] = v;
    }
    f32x4 dH[4][4], dZ[4][4];
    bf16x4 dZb[4][4];
    mm_tiles<1>(S.w4t, 16, dZ4b, dH, lane);
    // layer 3
    act_back(dH, S.zs[2], S.dzs, dZ, dZb, lane);
    __syncthreads();
    wgrad(S.dzs, MLP_H, 4, S.hs[1], MLP_H, 4, G + MLP_W3, MLP_H, MLP_H, MLP_H, lane);
    bgrad(dZ, G + MLP_B3, lane);
    mm_tiles<4>(S.w3t, MLP_H, dZb, dH, lane);
    __syncthreads();
    // layer 2
    act_back(dH, S.zs[1], S.dzs, dZ, dZb, lane);
    __syncthreads();
    wgrad(S.dzs, MLP_H, 4, S.hs[0], MLP_H, 4, G + MLP_W2, MLP_H, MLP_H, MLP_H, lane);
    bgrad(dZ, G + MLP_B2, lane);
    mm_tiles<4>(S.w2t, MLP_H, dZb, dH, lane);
    __syncthreads();
    // layer 1
    act_back(dH, S.zs[0], S.dzs, dZ, dZb, lane);
    __syncthreads();
    wgrad(S.dzs, MLP_H, 4, S.xs, 16, 1, G + MLP_W1, MLP_IN, MLP_H, MLP_IN, lane);
    bgrad(dZ, G + MLP_B1, lane);
    lsum = group16_sum(lsum);
    if (lane == 0) G[MLP_NPARAM] = lsum;
}

// grad[p] = sum over workgroups of part[w][p], in workgroup order; acc (nullable) += the loss
__global__ __launch_bounds__(256) void k_mlp_reduce(const float *__restrict__ part, int nwg, float *__restrict__ grad,
                                                    float *__restrict__ acc) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NP1) return;
    float s = 0.f;
    for (int w = 0; w < nwg; ++w) s += part[(size_t)w * NP1 + p];
    grad[p] = s;
    if (acc && p == MLP_NPARAM) *acc += s;
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ P, const float *__restrict__ g, float *__restrict__ m,
                                              float *__restrict__ v, float lr, float b1, float b2, float eps,
                                              float bc1, float bc2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.f - b1) * gp;
    const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    P[p] -= lr * (mp / bc1) / (sqrtf(vp / bc2) + eps);
}

__global__ __launch_bounds__(256) void k_mlp_eval(const float *__restrict__ P, const float *__restrict__ X, int n,
                                                  float *__restrict__ Y) {
    __shared__ MlpLds L;
    mlp_load(L, P, threadIdx.x, 256);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int s0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
    auto feat = [&](int s, int f) { return s0 + s < n ? X[(size_t)(s0 + s) * MLP_IN + f] : 0.0f; };
    f32x4 out[4];
    mlp_forward(L, feat, out, lane);
    if (lane < 16) {
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
            const int s = s0 + sb * 16 + lane;
            if (s < n) {
                Y[2 * (size_t)s] = out[sb][0];
                Y[2 * (size_t)s + 1] = out[sb][1];
            }
        }
    }
}

struct Mem {
    std::vector<void *> p;
    hipStream_t s = nullptr;
    ~Mem() {
        if (s) (void)hipStreamSynchronize(s);
        for (void *q : p) (void)hipFree(q);
        if (s) (void)hipStreamDestroy(s);
    }
    template <class T> int alloc(T **out, size_t count) {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) return BL_E_NOMEM;
        p.push_back(q);
        *out = static_cast<T *>(q);
        return BL_OK;
    }
};

int grad_launch(const float *P, const float *X, const float *T, const int *idx, int base, int count, float *part,
                float *grad, float *acc, hipStream_t s) {
    const int nwg = (count + 63) / 64;
    static bool attr = false;
    if (!attr) {
        BL_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_mlp_grad),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(GradLds)));
        attr = true;
    }
    hipLaunchKernelGGL(k_mlp_grad, dim3(nwg), dim3(64), sizeof(GradLds), s, P, X, T, idx, base, count,
                       1.0f / (float)count, part);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NP1 + 255) / 256), dim3(256), 0, s, part, nwg, grad, acc);
    BL_HIP(hipGetLastError());
    return BL_OK;
}

}  // namespace

int far_model_grad_host(const float *params, int32_t n, const float *feat, const float *targ, float *loss,
                        float *grad) {
    Mem M;
    BL_HIP(hipStreamCreateWithFlags(&M.s, hipStreamNonBlocking));
    float *dP, *dX, *dT, *dpart, *dg;
    const int nwg = (n + 63) / 64;
    if (M.alloc(&dP, MLP_NPARAM) || M.alloc(&dX, (size_t)n * MLP_IN) || M.alloc(&dT, (size_t)n * 2) ||
        M.alloc(&dpart, (size_t)nwg * NP1) || M.alloc(&dg, NP1))
        return BL_E_NOMEM;
    BL_HIP(hipMemcpyAsync(dP, params, 4 * MLP_NPARAM, hipMemcpyHostToDevice, M.s));
    BL_HIP(hipMemcpyAsync(dX, feat, 4 * (size_t)n * MLP_IN, hipMemcpyHostToDevice, M.s));
    BL_HIP(hipMemcpyAsync(dT, targ, 8 * (size_t)n, hipMemcpyHostToDevice, M.s));
    int rc = grad_launch(dP, dX, dT, nullptr, 0, n, dpart, dg, nullptr, M.s);
    if (rc) return rc;
    std::vector<float> g(NP1);
    BL_HIP(hipMemcpyAsync(g.data(), dg, 4 * NP1, hipMemcpyDeviceToHost, M.s));
    BL_HIP(hipStreamSynchronize(M.s));
    if (grad) memcpy(grad, g.data(), 4 * MLP_NPARAM);
    if (loss) *loss = g[MLP_NPARAM] / (float)n;
    return BL_OK;
}

int far_model_eval_host(const float *params, int32_t n, const float *feat, float *out) {
    Mem M;
    BL_HIP(hipStreamCreateWithFlags(&M.s, hipStreamNonBlocking));
    float *dP, *dX, *dY;
    if (M.alloc(&dP, MLP_NPARAM) || M.alloc(&dX, (size_t)n * MLP_IN) || M.alloc(&dY, (size_t)n * 2))
        return BL_E_NOMEM;
    BL_HIP(hipMemcpyAsync(dP, params, 4 * MLP_NPARAM, hipMemcpyHostToDevice, M.s));
    BL_HIP(hipMemcpyAsync(dX, feat, 4 * (size_t)n * MLP_IN, hipMemcpyHostToDevice, M.s));
    hipLaunchKernelGGL(k_mlp_eval, dim3((n + 255) / 256), dim3(256), 0, M.s, dP, dX, n, dY);
    BL_HIP(hipGetLastError());
    BL_HIP(hipMemcpyAsync(out, dY, 8 * (size_t)n, hipMemcpyDeviceToHost, M.s));
    BL_HIP(hipStreamSynchronize(M.s));
    return BL_OK;
}

// Deterministic Adam training: epochs over a permutation of the samples drawn from `seed`
// (xorshift64*, Fisher-Yates), batches of `batch` samples, learning rate lr (beta1 0.9,
// beta2 0.999, eps 1e-8).  losses[e] = mean loss over the batches of epoch e.
int far_model_train_host(float *params, int32_t n, const float *feat, const float *targ, int32_t epochs,
                         int32_t batch, float lr, uint64_t seed, float *losses) {
    Mem M;
    BL_HIP(hipStreamCreateWithFlags(&M.s, hipStreamNonBlocking));
    float *dP, *dX, *dT, *dpart, *dg, *dm, *dv, *dacc;
    int *didx;
    const int nwg = (batch + 63) / 64;
    if (M.alloc(&dP, MLP_NPARAM) || M.alloc(&dX, (size_t)n * MLP_IN) || M.alloc(&dT, (size_t)n * 2) ||
        M.alloc(&dpart, (size_t)nwg * NP1) || M.alloc(&dg, NP1) || M.alloc(&dm, MLP_NPARAM) ||
        M.alloc(&dv, MLP_NPARAM) || M.alloc(&dacc, std::max(epochs, 1)) || M.alloc(&didx, n))
        return BL_E_NOMEM;
    BL_HIP(hipMemcpyAsync(dP, params, 4 * MLP_NPARAM, hipMemcpyHostToDevice, M.s));
    BL_HIP(hipMemcpyAsync(dX, feat, 4 * (size_t)n * MLP_IN, hipMemcpyHostToDevice, M.s));
    BL_HIP(hipMemcpyAsync(dT, targ, 8 * (size_t)n, hipMemcpyHostToDevice, M.s));
    BL_HIP(hipMemsetAsync(dm, 0, 4 * MLP_NPARAM, M.s));
    BL_HIP(hipMemsetAsync(dv, 0, 4 * MLP_NPARAM, M.s));
    BL_HIP(hipMemsetAsync(dacc, 0, 4 * (size_t)std::max(epochs, 1), M.s));
    std::vector<int> perm(n);
    uint64_t st = seed ? seed : 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        st ^= st >> 12;
        st ^= st << 25;
        st ^= st >> 27;
        return st * 2685821657736338717ull;
    };
    const float b1 = 0.9f, b2 = 0.999f;
    double p1 = 1.0, p2 = 1.0;
    std::vector<int> nbatches(std::max(epochs, 1), 0);
    for (int e = 0; e < epochs; ++e) {
        for (int i = 0; i < n; ++i) perm[i] = i;
        for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[(int)(rnd() % (uint64_t)(i + 1))]);
        // the upload is ordered on the stream before this epoch's kernels; wait before the
        // next epoch rewrites perm
        BL_HIP(hipMemcpyAsync(didx, perm.data(), 4 * (size_t)n, hipMemcpyHostToDevice, M.s));
        for (int off = 0; off < n; off += batch) {
            const int cnt = std::min(batch, n - off);
            int rc = grad_launch(dP, dX, dT, didx, off, cnt, dpart, dg, dacc + e, M.s);
            if (rc) return rc;
            p1 *= b1;
            p2 *= b2;
            hipLaunchKernelGGL(k_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, M.s, dP, dg, dm, dv, lr, b1, b2,
                               1e-8f, (float)(1.0 - p1), (float)(1.0 - p2));
            ++nbatches[e];
        }
        BL_HIP(hipGetLastError());
        BL_HIP(hipStreamSynchronize(M.s));
    }
    std::vector<float> acc(std::max(epochs, 1));
    BL_HIP(hipMemcpyAsync(acc.data(), dacc, 4 * acc.size(), hipMemcpyDeviceToHost, M.s));
    BL_HIP(hipMemcpyAsync(params, dP, 4 * MLP_NPARAM, hipMemcpyDeviceToHost, M.s));
    BL_HIP(hipStreamSynchronize(M.s));
    // acc[e] is the sum over batches of the summed squared errors: mean per sample
    if (losses)
        for (int e = 0; e < epochs; ++e) losses[e] = acc[e] / (float)n;
    return BL_OK;
}

}  // namespace bl
