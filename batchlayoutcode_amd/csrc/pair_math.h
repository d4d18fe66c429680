// Device arithmetic shared by the minibatch kernels (single graph and batched).
//
// Reference semantics (sample/algorithms.cpp:758-768): a neighbour j of i contributes the
// attraction A(d) with d = x_j - x_i; any other j with |d|^2 > 0 contributes -d/|d|^2.
// The dense loops add -d/|d|^2 for every j (the self pair gives 0); correct_pair() then
// removes, for each neighbour, exactly the product the dense loop added and adds A(d).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "bl_internal.h"

namespace bl {

// |d|^2 floor: the self pair and coincident vertices give 0 * (1/kTiny) = 0, as the
// reference's `dist > 0` test skips them.
constexpr float kTiny = 1e-30f;

// v + the value of the lane CTRL names (a DPP move; lanes without a source lane, or outside
// ROW_MASK, add +0.0).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_add(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xF, false);
    return v + __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of the wave (all active), the same value in every lane: an inclusive
// scan by DPP moves (row_shr 1, 2, 4, 8 within rows of 16 lanes, then row_bcast:15 into rows
// 1 and 3 and row_bcast:31 into rows 2 and 3) leaves the total in lane 63, which is
// broadcast.  DPP moves cost a few cycles where the LDS permutes of a shuffle butterfly
// cost a round trip through the LDS unit each.
__device__ __forceinline__ double wave_sum(double v) {
    v = dpp_add<0x111, 0xF>(v);
    v = dpp_add<0x112, 0xF>(v);
    v = dpp_add<0x114, 0xF>(v);
    v = dpp_add<0x118, 0xF>(v);
    v = dpp_add<0x142, 0xA>(v);
    v = dpp_add<0x143, 0xC>(v);
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// Attraction of one neighbour pair, in the reference's FP64 operation order:
// f += d * (W * |d|) (:759), f += d * W (:985), f += d * log2(1 + W|d|) (:912).
__device__ __forceinline__ void attract64(int law, double W, double dx, double dy, double &fx, double &fy) {
    double s;
    if (law == LAW_FA) {
        s = W;
    } else {
        const double mag = sqrt(dx * dx + dy * dy);
        s = (law == LAW_FR) ? W * mag : log2(1.0 + W * mag);
    }
    fx += dx * s;
    fy += dy * s;
}

// 1/|d|^2 for a pair with FP64 difference (dx, dy).  REFINE: |d|^2 in FP64, the FP32
// reciprocal (v_rcp_f32) as seed and one FP64 Newton step (relative error ~1e-14).
// Otherwise |d|^2 and the reciprocal in FP32 (relative error ~1e-7).
template <bool REFINE>
__device__ __forceinline__ double inv_r2(const double dx, const double dy) {
    if constexpr (REFINE) {
        const double r2 = fma(dy, dy, fma(dx, dx, (double)kTiny));
        const double s = (double)__builtin_amdgcn_rcpf((float)r2);
        return fma(s, fma(-r2, s, 1.0), s);
    } else {
        const float fx = (float)dx, fy = (float)dy;
        const float r2 = fmaf(fy, fy, fmaf(fx, fx, kTiny));
        return (double)__builtin_amdgcn_rcpf(r2);
    }
}

// One dense repulsion pair: -d/|d|^2 accumulated in FP64.
template <bool REFINE>
__device__ __forceinline__ void pair_rep(const double qx, const double qy, const double px, const double py,
                                         double &ax, double &ay) {
    const double dx = qx - px, dy = qy - py;
    const double inv = inv_r2<REFINE>(dx, dy);
    ax = fma(-dx, inv, ax);
    ay = fma(-dy, inv, ay);
}

// Neighbour correction for one (i, j) pair: removes the dense loop's product (dx * inv with
// the same operands) and adds the attraction.
template <bool REFINE>
__device__ __forceinline__ void correct_pair(int law, double W, double dx, double dy, double &cx, double &cy) {
    const double inv = inv_r2<REFINE>(dx, dy);
    cx += dx * inv;
    cy += dy * inv;
    attract64(law, W, dx, dy, cx, cy);
}

// x_i += (F/|F|) * STEP, algorithms.cpp:775-779 / Coordinate.h:57-65.
__device__ __forceinline__ double2 moved(double2 p, double fx, double fy, double step) {
    const double mag = sqrt(fx * fx + fy * fy);
    p.x = p.x + (fx / mag) * step;
    p.y = p.y + (fy / mag) * step;
    return p;
}

}  // namespace bl
