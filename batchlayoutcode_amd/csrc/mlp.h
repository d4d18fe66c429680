// Far-field model: a 3-hidden-layer width-64 GELU network, 8 -> 64 -> 64 -> 64 -> 2, on BF16
// MFMA with FP32 accumulation (v_mfma_f32_16x16x16_bf16).  Shared by the training kernels
// (mlp.hip) and the learned traversal (farfield.hip).
//
// Orientation: a layer computes Z^T = W H^T with the weights as the A operand (rows = output
// units) and the activations as the B operand (columns = samples).  The MFMA result layout
// (lane l holds rows 4(l/16)..4(l/16)+3 of column l%16) is then exactly the B-operand layout
// of the next layer (lane l supplies k = 4(l/16)..+3 of column l%16), so activations stay in
// registers from layer to layer.  A wave evaluates 64 samples: 4 blocks of 16 columns.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bl {

constexpr int MLP_IN = 8, MLP_H = 64, MLP_OUT = 2;
// parameter vector (FP32): W1[64][8] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[2][64] b4[2]
constexpr int MLP_W1 = 0, MLP_B1 = 512, MLP_W2 = 576, MLP_B2 = 4672, MLP_W3 = 4736, MLP_B3 = 8832,
              MLP_W4 = 8896, MLP_B4 = 9024, MLP_NPARAM = 9026;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (gfx950's v_cvt_pk_bf16_f32; inputs are finite)
__device__ __forceinline__ short f2bf(float x) { return __builtin_bit_cast(short, (__bf16)x); }
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float(((uint32_t)(uint16_t)h) << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// K = 32 step (gfx950's v_mfma_f32_16x16x32_bf16, twice the work of the K = 16 one per
// issue) over the k-blocks 2h and 2h+1 of 16: a lane's 8 operands are its 4 of block 2h
// followed by its 4 of block 2h+1, in A and B alike.  The contraction does not depend on
// the order of k, so this gives the same sum as two K = 16 steps without moving activations
// between lanes.
typedef short bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ bf16x8 cat8(bf16x4 lo, bf16x4 hi) {
    return bf16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 pack4(f32x4 v) {  // two v_cvt_pk_bf16_f32
    typedef __bf16 hb4 __attribute__((ext_vector_type(4)));
    hb4 h;
    h[0] = (__bf16)v[0];
    h[1] = (__bf16)v[1];
    h[2] = (__bf16)v[2];
    h[3] = (__bf16)v[3];
    return __builtin_bit_cast(bf16x4, h);
}

// A fragment of a row-major BF16 matrix M (row stride ld): rows r0..r0+15, k0..k0+15;
// lane l reads M[r0 + l%16][k0 + 4(l/16) .. +3] (one 8-byte LDS read).
__device__ __forceinline__ bf16x4 frag_a(const short *M, int ld, int r0, int k0, int lane) {
    return *reinterpret_cast<const bf16x4 *>(M + (r0 + (lane & 15)) * ld + k0 + 4 * (lane >> 4));
}

// GELU in tanh form, 0.5 z (1 + tanh(u)) with u = sqrt(2/pi) (z + 0.044715 z^3), evaluated as
// z - z / (1 + e) with e = exp(2u) = 2^(z (GA + GB z^2)): 0.5 (1 + tanh u) = 1 - 1/(1 + e).
// Six VALU operations and two transcendentals; z -> -inf gives e -> 0 and exactly 0, z -> inf
// gives e -> inf and z.
constexpr float GA = 2.3022082f;   // 2 sqrt(2/pi) / ln 2
constexpr float GB = 0.10294324f;  // GA * 0.044715
__device__ __forceinline__ float gelu_r(float z) {  // 1 / (1 + exp(2u))
    const float e = __builtin_amdgcn_exp2f(z * fmaf(GB, z * z, GA));
    return __builtin_amdgcn_rcpf(e + 1.0f);
}
__device__ __forceinline__ float gelu(float z) { return fmaf(-z, gelu_r(z), z); }
// d/dz: s + z * 2 s r u'(z) with s = 1 - r = 0.5 (1 + tanh u), 2 s r = 0.5 (1 - tanh^2 u) and
// u'(z) = sqrt(2/pi) (1 + 3 * 0.044715 z^2)
__device__ __forceinline__ float gelu_grad(float z) {
    const float r = gelu_r(z), sg = 1.0f - r;
    const float du = fmaf(0.10703222f, z * z, 0.7978845608f);  // sqrt(2/pi) (1 + 0.134145 z^2)
    return fmaf(z * (2.0f * sg * r), du, sg);
}

// Weights of the model in LDS as BF16, row-major, padded to MFMA tiles.
struct MlpLds {
    short w1[MLP_H * 16];      // [64][16], inputs 8..15 zero
    short w2[MLP_H * MLP_H];   // [64][64]
    short w3[MLP_H * MLP_H];
    short w4[16 * MLP_H];      // [16][64], outputs 2..15 zero
    float b1[MLP_H], b2[MLP_H], b3[MLP_H], b4[4];
};

// Fills L from the FP32 parameter vector P; every thread of the block takes part (the caller
// synchronises afterwards).
__device__ __forceinline__ void mlp_load(MlpLds &L, const float *__restrict__ P, int tid, int nthreads) {
    for (int k = tid; k < MLP_H * 16; k += nthreads) {
        const int o = k >> 4, i = k & 15;
        L.w1[k] = i < MLP_IN ? f2bf(P[MLP_W1 + o * MLP_IN + i]) : (short)0;
    }
    for (int k = tid; k < MLP_H * MLP_H; k += nthreads) {
        L.w2[k] = f2bf(P[MLP_W2 + k]);
        L.w3[k] = f2bf(P[MLP_W3 + k]);
    }
    for (int k = tid; k < 16 * MLP_H; k += nthreads) {
        const int o = k / MLP_H;
        L.w4[k] = o < MLP_OUT ? f2bf(P[MLP_W4 + k]) : (short)0;
    }
    for (int k = tid; k < MLP_H; k += nthreads) {
        L.b1[k] = P[MLP_B1 + k];
        L.b2[k] = P[MLP_B2 + k];
        L.b3[k] = P[MLP_B3 + k];
    }
    if (tid < 4) L.b4[tid] = tid < MLP_OUT ? P[MLP_B4 + tid] : 0.0f;
}

// One hidden layer for the wave's 64 samples: Z[ob][sb] = W (64 x 16*KB) H[kb][sb] + b, and
// the activations gelu(Z) as the next layer's B fragments.  Z is returned (FP32) when the
// caller needs it for the backward pass.  K32 selects the K = 32 MFMA for even KB: fewer
// issues, but more registers (the learned traversal, whose occupancy they set, uses K = 16).
template <int KB, bool K32 = true>
__device__ __forceinline__ void mlp_hidden(const short *W, int ld, const float *b, const bf16x4 (&Hin)[4][4],
                                           f32x4 (&Z)[4][4], bf16x4 (&Hout)[4][4], int lane) {
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
        bf16x4 a[KB];
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) a[kb] = frag_a(W, ld, ob * 16, kb * 16, lane);
        f32x4 bias;
#pragma unroll
        for (int i = 0; i < 4; ++i) bias[i] = b[ob * 16 + 4 * (lane >> 4) + i];
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
            f32x4 acc = bias;
            if constexpr (K32 && KB % 2 == 0) {
#pragma unroll
                for (int h = 0; h < KB / 2; ++h)
                    acc = mfma32(cat8(a[2 * h], a[2 * h + 1]), cat8(Hin[2 * h][sb], Hin[2 * h + 1][sb]), acc);
            } else {
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) acc = mfma16(a[kb], Hin[kb][sb], acc);
            }
            Z[ob][sb] = acc;
            f32x4 h;
#pragma unroll
            for (int i = 0; i < 4; ++i) h[i] = gelu(acc[i]);
            Hout[ob][sb] = pack4(h);
        }
    }
}

// Output layer: Y[sb] = W4 H3[.][sb] + b4; lanes 0..15 hold (y0, y1) of sample sb*16 + lane
// in elements 0 and 1.
template <bool K32 = true>
__device__ __forceinline__ void mlp_out(const MlpLds &L, const bf16x4 (&H)[4][4], f32x4 (&Y)[4], int lane) {
    bf16x4 a[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) a[kb] = frag_a(L.w4, MLP_H, 0, kb * 16, lane);
    f32x4 bias;
#pragma unroll
    for (int i = 0; i < 4; ++i) bias[i] = L.b4[(4 * (lane >> 4) + i) & 3] * ((lane >> 4) == 0 ? 1.0f : 0.0f);
#pragma unroll
    for (int sb = 0; sb < 4; ++sb) {
        f32x4 acc = bias;
        if constexpr (K32) {
#pragma unroll
            for (int h = 0; h < 2; ++h)
                acc = mfma32(cat8(a[2 * h], a[2 * h + 1]), cat8(H[2 * h][sb], H[2 * h + 1][sb]), acc);
        } else {
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) acc = mfma16(a[kb], H[kb][sb], acc);
        }
        Y[sb] = acc;
    }
}

// B fragments of the input features: lane l supplies features 4(l/16)..+3 (zero beyond 8)
// of sample sb*16 + l%16; feat(s, i) gives feature i of sample s.
template <class F>
__device__ __forceinline__ void mlp_input(F feat, bf16x4 (&X)[4][4], int lane) {
#pragma unroll
    for (int sb = 0; sb < 4; ++sb) {
        const int s = sb * 16 + (lane & 15);
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            bf16x4 v;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int f = kb * 16 + 4 * (lane >> 4) + i;
                v[i] = (kb == 0 && f < MLP_IN) ? f2bf(feat(s, f)) : (short)0;
            }
            X[kb][sb] = v;
        }
    }
}

// Full forward pass of the wave's 64 samples (inference).
template <bool K32 = true, class F>
__device__ __forceinline__ void mlp_forward(const MlpLds &L, F feat, f32x4 (&Y)[4], int lane) {
    bf16x4 X[4][4], H1[4][4], H2[4][4];
    f32x4 Z[4][4];
    mlp_input(feat, X, lane);
    mlp_hidden<1, K32>(L.w1, 16, L.b1, X, Z, H1, lane);
    mlp_hidden<4, K32>(L.w2, MLP_H, L.b2, H1, Z, H2, lane);
    mlp_hidden<4, K32>(L.w3, MLP_H, L.b3, H2, Z, H1, lane);
    mlp_out<K32>(L, H1, Y, lane);
}

}  // namespace bl
