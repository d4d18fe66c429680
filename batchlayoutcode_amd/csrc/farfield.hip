// Physically consistent far field for the Barnes-Hut layout (no reference counterpart).
//
// The reference's Barnes-Hut (BarnesHut.h:69-86) adds one unit vector per accepted cell,
// whatever the cell holds.  The layouts here use the repulsion law of algo 2
// (algorithms.cpp:764-768: a pair contributes d/|d|^2) for every vertex pair and approximate
// the far field of a cell by its mass-weighted monopole m (c - v)/|c - v|^2 (c = mean of the
// cell's points, m = their count), optionally corrected for the cell's shape: in the frame of
// u = (c - v)/|c - v| and w = u turned by +90 degrees, the cell's exact field is
//     F = (m/rho) ((1 + a) u + b w),     rho = |c - v|,
// and the correction (a, b) is either the second-order multipole term (a = S_uu - S_ww,
// b = -2 S_uw with S the cell's normalised second moments, from 1/conj(w + delta) expanded
// in delta) or the output of the learned model of mlp.h scaled by (side/rho)^2.
//
//   k_far_direct / k_far_top  count, mean, central second moments and first Morton rank of
//                  every cell, bottom-up over the tree of bh.hip
//   k_far_trav     traversal, one wave per 64 Morton-consecutive queries with a shared stack
//                  of (node, lane mask); a cell is accepted when side^2 < theta^2 |c - v|^2.
//                  The learned mode queues accepted cells per wave in LDS and evaluates the
//                  model on full tiles of 64 cells with MFMA
//   k_far_samples  training data: features and exact (a, b) of every accepted cell
//   k_far_exact    exact sum over all points for every query
#include <math.h>

#include <algorithm>
#include <vector>

#include "bh_common.h"
#include "mlp.h"
#include "pair_math.h"

namespace bl {
namespace {

constexpr int WG = 256;
constexpr int STACK = 128;
constexpr int QCAP = 128;  // queued cells per wave (a step adds at most 64)

// One 64-byte record per node, 64-byte aligned, so that a traversal step reads one cache line:
// the summary, and the node's level and children copied from the tree (BHNode.depth, .ch).  A
// summary and a BHNode read together took three lines a step on average.  The node's points
// are Morton ranks [node_lo[v], node_lo[v] + m).
struct __attribute__((aligned(64))) FarCell {
    double cx, cy;          // mean of the cell's points
    double mxx, mxy, myy;   // central second moments, sums over the points
    int m;                  // number of points
    int depth;              // base-4 level of the cell: side = MAXMIN * 2^-depth
    int4 ch;                // children: >= 0 node, -1 empty, <= -2 leaf at Morton rank (-c-2)
};
static_assert(sizeof(FarCell) == 64, "one cache line per node");

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Combines the non-empty quadrants ch of a node (children first): mean by counts, second
// moments by the parallel-axis rule, in quadrant order.
__device__ FarCell far_summary(int4 ch, const double2 *spos, const FarCell *cells) {
    const int c[4] = {ch.x, ch.y, ch.z, ch.w};
    double m = 0.0, sx = 0.0, sy = 0.0;
    double km[4], kx[4], ky[4], kxx[4], kxy[4], kyy[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        km[q] = 0.0;
        kx[q] = ky[q] = kxx[q] = kxy[q] = kyy[q] = 0.0;
        if (c[q] >= 0) {
            const FarCell k = cells[c[q]];
            km[q] = k.m;
            kx[q] = k.cx;
            ky[q] = k.cy;
            kxx[q] = k.mxx;
            kxy[q] = k.mxy;
            kyy[q] = k.myy;
        } else if (c[q] <= -2) {
            const double2 p = spos[-c[q] - 2];
            km[q] = 1.0;
            kx[q] = p.x;
            ky[q] = p.y;
        }
        m += km[q];
        sx += km[q] * kx[q];
        sy += km[q] * ky[q];
    }
    FarCell out;
    out.cx = sx / m;
    out.cy = sy / m;
    double a = 0.0, b = 0.0, d = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double ux = kx[q] - out.cx, uy = ky[q] - out.cy;
        a += kxx[q] + km[q] * ux * ux;
        b += kxy[q] + km[q] * ux * uy;
        d += kyy[q] + km[q] * uy * uy;
    }
    out.mxx = a;
    out.mxy = b;
    out.myy = d;
    out.m = (int)m;
    return out;
}

// Stores cell v with agent-scope stores, which the XCD's L2 writes through to memory.
// The summary fields (not depth and ch, written by k_far_direct).
__device__ __forceinline__ void store_cell_through(FarCell *cells, int v, const FarCell &c) {
    FarCell *d = cells + v;
    __hip_atomic_store(&d->cx, c.cx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&d->cy, c.cy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&d->mxx, c.mxx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&d->mxy, c.mxy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&d->myy, c.myy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&d->m, c.m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The summaries in two launches.  A node of at most FAR_SMALL points is summarised straight
// from its points, Morton ranks [lo, lo + m), in two passes (the mean, then the central second
// moments): one thread per node, no dependence between nodes; the same launch copies every
// node's level and children into its record and counts each larger node's children of more
// than FAR_SMALL points (nlc).  The larger nodes climb, from those
// without large children: a node's summary is stored with agent-scope stores (written through
// to memory) and, once they have completed, a relaxed counter at the parent (BHNode.cnt, zeroed
// by the tree build) counts the arrival; the last child to arrive acquires (invalidating its own
// L2's stale lines) and forms the parent from its children by far_summary.  An acq_rel counter
// would write the whole L2 back at every step.  What does not change in the launch (the
// parent, its children and its nlc) is loaded before the hand-off, so a step waits for the
// children's summaries, the stores and the counter only (k_far_top 145 -> 137 us on the
// 1.66M-vertex Delaunay graph's layout at FAR_SMALL 32; the climb's nine levels cost about
// 15 us each).
constexpr int FAR_SMALL = 64;  // Delaunay graph, per iteration: 16: 251 us, 32: 204, 64: 179, 128: 191

__global__ __launch_bounds__(WG) void k_far_direct(const BHNode *__restrict__ nodes, const double2 *__restrict__ spos,
                                                   const int *__restrict__ npts, const int *__restrict__ node_lo,
                                                   const int *__restrict__ misc, FarCell *__restrict__ cells,
                                                   int *__restrict__ nlc) {
    const int v = blockIdx.x * WG + threadIdx.x;
    if (misc[0] || v >= misc[2]) return;
    const int m = npts[v];
    const int4 ch = nodes[v].ch;
    const int depth = nodes[v].depth;
    if (m > FAR_SMALL) {
        const int c[4] = {ch.x, ch.y, ch.z, ch.w};
        int k = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) k += (c[q] >= 0 && npts[c[q]] > FAR_SMALL) ? 1 : 0;
        nlc[v] = k;
        cells[v].depth = depth;
        cells[v].ch = ch;
        return;
    }
    const int lo = node_lo[v];
    double sx = 0.0, sy = 0.0;
    for (int k = 0; k < m; ++k) {
        const double2 p = spos[lo + k];
        sx += p.x;
        sy += p.y;
    }
    FarCell out;
    out.cx = sx / m;
    out.cy = sy / m;
    double a = 0.0, b = 0.0, d = 0.0;
    for (int k = 0; k < m; ++k) {
        const double2 p = spos[lo + k];
        const double ux = p.x - out.cx, uy = p.y - out.cy;
        a += ux * ux;
        b += ux * uy;
        d += uy * uy;
    }
    out.mxx = a;
    out.mxy = b;
    out.myy = d;
    out.m = m;
    out.depth = depth;
    out.ch = ch;
    cells[v] = out;
}

__global__ __launch_bounds__(WG) void k_far_top(BHNode *nodes, const double2 *__restrict__ spos,
                                                const int *__restrict__ npts, const int *__restrict__ misc,
                                                const int *__restrict__ nlc, FarCell *__restrict__ cells) {
    int v = blockIdx.x * WG + threadIdx.x;
    if (misc[0] || v >= misc[2]) return;
    if (npts[v] <= FAR_SMALL || nlc[v] != 0) return;
    int4 ch = nodes[v].ch;
    for (;;) {
        const int u = nodes[v].parent;
        int need = 0;
        int4 chu = make_int4(-1, -1, -1, -1);
        if (u >= 0) {
            need = nlc[u];
            chu = nodes[u].ch;
        }
        store_cell_through(cells, v, far_summary(ch, spos, cells));
        if (u < 0) break;
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        const int old = __hip_atomic_fetch_add(&nodes[u].cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old != need - 1) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        v = u;
        ch = chu;
    }
}

// Features of an accepted cell seen from v; (dx, dy) = c - v, r2 = |c - v|^2, D = side:
// [D/rho, S_uu, S_uw, S_ww, S_uu - S_ww, log2(m)/16, 1/m, (D/rho)^2], S = M2 / (m rho^2) in the
// (u, w) frame.
__device__ __forceinline__ void far_features(const FarCell &cl, double dx, double dy, double r2, double D, float *f) {
    const double rho = sqrt(r2);
    const double ux = dx / rho, uy = dy / rho, wx = -uy, wy = ux;
    const double m = (double)cl.m;
    const double s = 1.0 / (m * r2);
    const double suu = (ux * ux * cl.mxx + 2.0 * ux * uy * cl.mxy + uy * uy * cl.myy) * s;
    const double suw = (ux * wx * cl.mxx + (ux * wy + uy * wx) * cl.mxy + uy * wy * cl.myy) * s;
    const double sww = (wx * wx * cl.mxx + 2.0 * wx * wy * cl.mxy + wy * wy * cl.myy) * s;
    const double t = D / rho;
    f[0] = (float)t;
    f[1] = (float)suu;
    f[2] = (float)suw;
    f[3] = (float)sww;
    f[4] = (float)(suu - sww);
    f[5] = (float)(log2(m) / 16.0);
    f[6] = (float)(1.0 / m);
    f[7] = (float)(t * t);
}

// The same features in FP32 with hardware reciprocal, reciprocal square root and log2 (about
// 1 ulp each): the learned traversal's model takes them as BF16 inputs, so FP32 is ample,
// and FP64 divisions, square root and log2 cost every wave in which any lane accepts a cell.
// Training samples (k_far_samples) use far_features; the quadrupole term has its own closed
// form in d (k_far_trav).
__device__ __forceinline__ void far_features32(const FarCell &cl, double dx, double dy, double r2, double D,
                                               float *f) {
    const float irho = __builtin_amdgcn_rsqf((float)r2);
    const float ux = (float)dx * irho, uy = (float)dy * irho, wx = -uy, wy = ux;
    const float m = (float)cl.m;
    const float s = __builtin_amdgcn_rcpf(m * (float)r2);
    const float mxx = (float)cl.mxx, mxy = (float)cl.mxy, myy = (float)cl.myy;
    const float suu = (ux * ux * mxx + 2.0f * ux * uy * mxy + uy * uy * myy) * s;
    const float suw = (ux * wx * mxx + (ux * wy + uy * wx) * mxy + uy * wy * myy) * s;
    const float sww = (wx * wx * mxx + 2.0f * wx * wy * mxy + wy * wy * myy) * s;
    const float t = (float)D * irho;
    f[0] = t;
    f[1] = suu;
    f[2] = suw;
    f[3] = sww;
    f[4] = suu - sww;
    f[5] = __builtin_amdgcn_logf(m) * (1.0f / 16.0f);
    f[6] = __builtin_amdgcn_rcpf(m);
    f[7] = t * t;
}

enum TravMode : int { TRAV_MONO = 0, TRAV_QUAD = 1, TRAV_LEARNED = 2 };
constexpr size_t QOFF = (sizeof(MlpLds) + 15) & ~(size_t)15;

// Per-wave queue of accepted cells waiting for the model (learned mode).
struct FarQueue {
    float feat[QCAP][MLP_IN];
    float mono[QCAP][2];  // the cell's monopole m (c - v)/|c - v|^2
    float corr[64][2];    // model correction of the evaluated tile
};
constexpr size_t LEARNED_LDS = QOFF + sizeof(FarQueue) * (WG / 64);

// Evaluates the first 64 queued cells and adds (f7 y0) mono + (f7 y1) mono_perp to their
// owners, each lane summing its own cells in queue order: `mine` has bit s set when the
// lane queued slot s (a lane visits only its own slots, not all 64).
__device__ void far_flush(const MlpLds &L, FarQueue &Q, int count, int lane, uint64_t mine, double &fx,
                          double &fy) {
    auto feat = [&](int s, int f) { return s < count ? Q.feat[s][f] : 0.0f; };
    f32x4 Y[4];
    mlp_forward<false>(L, feat, Y, lane);  // K = 16: fewer registers, higher occupancy
    if (lane < 16) {
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
            const int s = sb * 16 + lane;
            const float t2 = Q.feat[s][7];
            const float a = t2 * Y[sb][0], b = t2 * Y[sb][1];
            Q.corr[s][0] = a * Q.mono[s][0] - b * Q.mono[s][1];
            Q.corr[s][1] = a * Q.mono[s][1] + b * Q.mono[s][0];
        }
    }
    __builtin_amdgcn_wave_barrier();
    for (uint64_t m = mine; m; m &= m - 1) {
        const int s = __builtin_ctzll(m);
        fx += (double)Q.corr[s][0];
        fy += (double)Q.corr[s][1];
    }
    __builtin_amdgcn_wave_barrier();
}

// Repulsion F(v) = sum over accepted cells of their far field + sum over the single points met
// of (p - v)/|p - v|^2.  Opening test and monopole terms are FP64-grade (FP64 differences,
// FP32-seeded reciprocal with an FP64 Newton step); sums in FP64.
template <int MODE>
__global__ __launch_bounds__(WG) void k_far_trav(const double2 *__restrict__ qpos, const int *__restrict__ qid,
                                                 int nq, const double2 *__restrict__ spos,
                                                 const BHNode *__restrict__ nodes, const FarCell *__restrict__ cells,
                                                 const int *__restrict__ misc, const double *__restrict__ frame,
                                                 double theta, const float *__restrict__ model,
                                                 double2 *__restrict__ out, unsigned long long *ncells,
                                                 const int *use_model, int *model_iters) {
    __shared__ int stk_node[WG / 64][STACK];
    __shared__ uint64_t stk_mask[WG / 64][STACK];
    // learned mode: the model and the per-wave queues in dynamic LDS (LEARNED_LDS bytes)
    extern __shared__ __align__(16) unsigned char dyn[];
    MlpLds *L = reinterpret_cast<MlpLds *>(dyn);
    FarQueue *Qs = reinterpret_cast<FarQueue *>(dyn + QOFF);
    if (misc[0]) return;
    // learned mode: the judge's verdict (use_model, nullable: always the model); when it rejected
    // the model the cells get the analytic second-order correction instead
    const bool learned = MODE == TRAV_LEARNED && (use_model == nullptr || *use_model != 0);
    const bool quad = MODE == TRAV_QUAD || (MODE == TRAV_LEARNED && !learned);
    if (MODE == TRAV_LEARNED && model_iters && learned && blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(model_iters, 1);
    if (learned) {
        mlp_load(*L, model, threadIdx.x, WG);
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k = (xcd_tile(blockIdx.x, gridDim.x) * (WG / 64) + w) * 64 + lane;
    const bool valid = k < nq && qid[k] >= 0;  // a negative id: an idle lane (the judge's sample)
    const double2 p = qpos[valid ? k : nq - 1];
    const double maxmin = frame[2], th2 = theta * theta;
    double fx = 0.0, fy = 0.0;
    const int root = misc[1];
    const uint64_t m0 = __ballot(valid);
    if (m0 == 0) return;
    if (root < 0) {  // a single point
        if (valid) {
            const double dx = spos[0].x - p.x, dy = spos[0].y - p.y;
            const double r2 = dx * dx + dy * dy;
            out[qid[k]] = r2 > 0.0 ? make_double2(dx / r2, dy / r2) : make_double2(0.0, 0.0);
        }
        return;
    }
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    unsigned accepted = 0;  // cells this wave's lanes accepted (wave-uniform)
    int qn = 0;  // queued cells of this wave (learned mode)
    uint64_t mlo = 0, mhi = 0;  // this lane's queue slots 0..63 and 64..127 (learned mode)
    if (lane == 0) {
        stk_node[w][0] = root;
        stk_mask[w][0] = m0;
    }
    int sp = 1;
    __builtin_amdgcn_wave_barrier();
    while (sp > 0) {
        --sp;
        const int v = uni(stk_node[w][sp]);
        const uint64_t m = stk_mask[w][sp];
        __builtin_amdgcn_wave_barrier();
        const bool act = (m >> lane) & 1ull;
        const FarCell cl = cells[v];  // one line: the summary, the level and the children
        const double dx = cl.cx - p.x, dy = cl.cy - p.y;
        const double r2 = dx * dx + dy * dy;
        const double D = ldexp(maxmin, -cl.depth);
        const bool acc = D * D < th2 * r2;
        if (ncells) accepted += (unsigned)__popcll(__ballot(act && acc));
        if (act && acc) {
            const double ir2 = inv_r2<true>(dx, dy);
            const double s = (double)cl.m * ir2;
            fx += dx * s;
            fy += dy * s;
            if (quad) {
                // the second-order term (a, b) = (S_uu - S_ww, -2 S_uw) applied to the monopole
                // m d / |d|^2: in d itself the mass cancels and it is |d|^-6 (A d + B d_perp),
                // A = (dx^2 - dy^2)(Mxx - Myy) + 4 dx dy Mxy, B = -2 (dx dy (Myy - Mxx) +
                // (dx^2 - dy^2) Mxy): FP64 products only, no square root, division or log
                const double ex = dx * dx - dy * dy, xy = dx * dy;
                const double A = fma(ex, cl.mxx - cl.myy, 4.0 * xy * cl.mxy);
                const double Bq = -2.0 * fma(xy, cl.myy - cl.mxx, ex * cl.mxy);
                const double i6 = ir2 * ir2 * ir2;
                fx = fma(i6, fma(A, dx, -Bq * dy), fx);
                fy = fma(i6, fma(A, dy, Bq * dx), fy);
            }
        }
        if (MODE == TRAV_LEARNED && learned) {
            const uint64_t accm = __ballot(act && acc);
            if (accm) {
                if (act && acc) {
                    const int slot = qn + __popcll(accm & lt);
                    FarQueue &Q = Qs[w];
                    far_features32(cl, dx, dy, r2, D, Q.feat[slot]);
                    const double s = (double)cl.m * inv_r2<true>(dx, dy);
                    Q.mono[slot][0] = (float)(dx * s);
                    Q.mono[slot][1] = (float)(dy * s);
                    if (slot < 64) mlo |= 1ull << slot; else mhi |= 1ull << (slot - 64);
                }
                qn += __popcll(accm);
                __builtin_amdgcn_wave_barrier();
                if (qn >= 64) {
                    FarQueue &Q = Qs[w];
                    far_flush(*L, Q, 64, lane, mlo, fx, fy);
                    // move the remaining qn - 64 (< 64) entries to the front
                    const int rest = qn - 64;
                    if (lane < rest) {
#pragma unroll
                        for (int f = 0; f < MLP_IN; ++f) Q.feat[lane][f] = Q.feat[64 + lane][f];
                        Q.mono[lane][0] = Q.mono[64 + lane][0];
                        Q.mono[lane][1] = Q.mono[64 + lane][1];
                    }
                    qn = rest;
                    mlo = mhi;
                    mhi = 0;
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
        const uint64_t rej = __ballot(act && !acc);
        if (rej) {
            const int c[4] = {cl.ch.x, cl.ch.y, cl.ch.z, cl.ch.w};
#pragma unroll
            for (int q = 3; q >= 0; --q) {
                if (c[q] >= 0) {
                    if (lane == 0) {
                        stk_node[w][sp] = c[q];
                        stk_mask[w][sp] = rej;
                    }
                    ++sp;
                } else if (c[q] <= -2 && ((rej >> lane) & 1ull)) {
                    const double2 l = spos[-c[q] - 2];
                    const double lx = l.x - p.x, ly = l.y - p.y;
                    if (lx * lx + ly * ly > 0.0) {
                        const double s = inv_r2<true>(lx, ly);
                        fx = fma(lx, s, fx);
                        fy = fma(ly, s, fy);
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (MODE == TRAV_LEARNED && learned) {
        if (qn > 0) far_flush(*L, Qs[w], qn, lane, mlo, fx, fy);
    }
    if (valid) out[qid[k]] = make_double2(fx, fy);
    if (ncells && lane == 0) atomicAdd(ncells, (unsigned long long)accepted);
}

// Training samples: for query k, its accepted cells (at most cap) with their features and
// the exact (a, b) of F_cell = (m/rho)((1 + a) u + b w), the exact field summed over the cell's
// points in FP64.  Rows k*cap .. k*cap + count[k] - 1, in traversal order.
__global__ __launch_bounds__(WG) void k_far_samples(const double2 *__restrict__ qpos, int nq,
                                                    const double2 *__restrict__ spos,
                                                    const BHNode *__restrict__ nodes, const int *__restrict__ node_lo,
                                                    const FarCell *__restrict__ cells, const int *__restrict__ misc,
                                                    const double *__restrict__ frame, double theta, int cap,
                                                    float *__restrict__ feat, float *__restrict__ targ,
                                                    int *__restrict__ count) {
    __shared__ int stk_node[WG / 64][STACK];
    __shared__ uint64_t stk_mask[WG / 64][STACK];
    if (misc[0]) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k = (blockIdx.x * (WG / 64) + w) * 64 + lane;
    const bool valid = k < nq;
    const double2 p = qpos[valid ? k : nq - 1];
    const double maxmin = frame[2], th2 = theta * theta;
    const int root = misc[1];
    const uint64_t m0 = __ballot(valid);
    if (m0 == 0) return;
    int cnt = 0;
    if (root >= 0) {
        if (lane == 0) {
            stk_node[w][0] = root;
            stk_mask[w][0] = m0;
        }
        int sp = 1;
        __builtin_amdgcn_wave_barrier();
        while (sp > 0) {
            --sp;
            const int v = uni(stk_node[w][sp]);
            const uint64_t m = stk_mask[w][sp];
            __builtin_amdgcn_wave_barrier();
            const bool act = (m >> lane) & 1ull;
            const FarCell cl = cells[v];
            const BHNode nd = nodes[v];
            const double dx = cl.cx - p.x, dy = cl.cy - p.y;
            const double r2 = dx * dx + dy * dy;
            const double D = ldexp(maxmin, -nd.depth);
            const bool acc = D * D < th2 * r2;
            const bool take = act && acc && cnt < cap;
            if (__ballot(take)) {
                double ex = 0.0, ey = 0.0;
                const int lo = uni(node_lo[v]), hi = lo + uni(cl.m);
                for (int j = lo; j < hi; ++j) {  // same points for every lane: broadcast loads
                    const double2 q = spos[j];
                    const double qx = q.x - p.x, qy = q.y - p.y;
                    const double q2 = qx * qx + qy * qy;
                    if (take && q2 > 0.0) {
                        ex += qx / q2;
                        ey += qy / q2;
                    }
                }
                if (take) {
                    const size_t row = (size_t)k * cap + cnt;
                    far_features(cl, dx, dy, r2, D, feat + row * MLP_IN);
                    const double rho = sqrt(r2), ux = dx / rho, uy = dy / rho, sc = rho / (double)cl.m;
                    targ[2 * row] = (float)((ex * ux + ey * uy) * sc - 1.0);
                    targ[2 * row + 1] = (float)((-ex * uy + ey * ux) * sc);
                    ++cnt;
                }
            }
            const uint64_t rej = __ballot(act && !acc);
            if (rej) {
                const int c[4] = {nd.ch.x, nd.ch.y, nd.ch.z, nd.ch.w};
#pragma unroll
                for (int q = 3; q >= 0; --q) {
                    if (c[q] >= 0) {
                        if (lane == 0) {
                            stk_node[w][sp] = c[q];
                            stk_mask[w][sp] = rej;
                        }
                        ++sp;
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    if (valid) count[k] = cnt;
}

// Exact F(v) = sum_j (x_j - v)/|x_j - v|^2 over all points (the self pair and coincident
// points give 0).  Point j is read with a wave-uniform index, i.e. by scalar loads.
typedef const __attribute__((address_space(4))) double cdouble;

__global__ __launch_bounds__(WG) void k_far_exact(const double2 *__restrict__ qpos, const int *__restrict__ qid,
                                                  int nq, const double2 *__restrict__ pts, int n,
                                                  double2 *__restrict__ out) {
    const int k = blockIdx.x * WG + threadIdx.x;
    const double2 p = qpos[k < nq ? k : nq - 1];
    cdouble *cp = (cdouble *)pts;
    double ax = 0.0, ay = 0.0;
    int j = 0;
    for (; j + 8 <= n; j += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) pair_rep<true>(cp[2 * (j + u)], cp[2 * (j + u) + 1], p.x, p.y, ax, ay);
    }
    for (; j < n; ++j) pair_rep<true>(cp[2 * j], cp[2 * j + 1], p.x, p.y, ax, ay);
    if (k < nq) out[qid ? qid[k] : k] = make_double2(-ax, -ay);  // pair_rep sums -d/|d|^2
}

int launch_trav(BHWork &w, const double2 *qpos, const int *qid, int nq, double theta, int mode, const float *model,
                double2 *out, hipStream_t s, unsigned long long *ncells = nullptr, const int *use_model = nullptr,
                int *model_iters = nullptr) {
    const int g64 = ((nq + 63) / 64 + WG / 64 - 1) / (WG / 64);
    const BHNode *nodes = w.node_child.as<BHNode>();
    const FarCell *cells = w.far_cells.as<FarCell>();
    const int *misc = w.misc.as<int>();
    const double2 *spos = w.spos.as<double2>();
    const double *frame = w.frame.as<double>();
    if (mode == FAR_MONO)
        hipLaunchKernelGGL(k_far_trav<TRAV_MONO>, dim3(g64), dim3(WG), 0, s, qpos, qid, nq, spos, nodes, cells, misc,
                           frame, theta, model, out, ncells, use_model, model_iters);
    else if (mode == FAR_QUAD)
        hipLaunchKernelGGL(k_far_trav<TRAV_QUAD>, dim3(g64), dim3(WG), 0, s, qpos, qid, nq, spos, nodes, cells, misc,
                           frame, theta, model, out, ncells, use_model, model_iters);
    else if (mode == FAR_LEARNED && model)
        hipLaunchKernelGGL(k_far_trav<TRAV_LEARNED>, dim3(g64), dim3(WG), LEARNED_LDS, s, qpos, qid, nq, spos, nodes, cells,
                           misc, frame, theta, model, out, ncells, use_model, model_iters);
    else
        return BL_E_ARG;
    BL_HIP(hipGetLastError());
    return BL_OK;
}

struct HostRun {
    hipStream_t s = nullptr;
    BHWork w;
    DevBuf pts, qs, qi, res, mdl, f, t, c;
    ~HostRun() {
        if (s) (void)hipStreamSynchronize(s);
        w.release();
        DevBuf *all[] = {&pts, &qs, &qi, &res, &mdl, &f, &t, &c};
        for (DevBuf *b : all) b->release();
        if (s) (void)hipStreamDestroy(s);
    }
    int status() {
        int st = 0;
        BL_HIP(hipMemcpyAsync(&st, w.misc.as<int>(), sizeof(int), hipMemcpyDeviceToHost, s));
        BL_HIP(hipStreamSynchronize(s));
        return (st & ST_DUPKEY) ? BL_E_DUPKEY : BL_OK;
    }
};

// ---- the learned far field's judge -------------------------------------------------------
// Every few iterations the learned field is compared with the exact repulsion on a sample of
// JUDGE_S vertices at the iteration-start positions, next to the analytic second-order field;
// the traversals of the iteration then use the model only while its median relative force
// error on the sample is at most the analytic term's.
constexpr int JUDGE_S = 256;   // sample vertices
constexpr int JUDGE_P = 1024;  // j chunks of the exact sum (four workgroups per CU)

// layout of BHWork::judge
struct JudgeBuf {
    double2 q[JUDGE_S];            // sample positions
    // the traversals' queries: sample vertex k in lane 0 of wave k (output slot k), the other
    // lanes idle (id -1).  The sample is spread over the layout, so a wave of 64 different
    // sample vertices would walk the union of 64 unrelated paths one dependent node load at a
    // time (about 5 ms per traversal on 1.7M vertices); one vertex per wave walks one path, the
    // waves in parallel.  With the vertex in all 64 lanes the learned traversal queued every
    // accepted cell 64 times (0.9 ms against 0.06 ms for the analytic one).
    double2 qw[JUDGE_S * 64];
    int qwid[JUDGE_S * 64];
    double2 part[JUDGE_P * JUDGE_S];  // chunk-major: part[c * JUDGE_S + k]
    double2 field[2 * JUDGE_S];    // learned, then analytic, at the sample
    int use_model;                 // the verdict read by the traversals
    int model_iters;               // traversals that used the model
    float err[2];                  // median relative error of the learned and the analytic field
};

__global__ __launch_bounds__(JUDGE_S) void k_judge_sample(const double2 *__restrict__ pos, int n, int S, int offset,
                                                           JudgeBuf *J) {
    const int k = threadIdx.x;
    if (k >= S) return;
    const long long id = (long long)k * n / S + offset;
    const double2 q = pos[id < n ? id : n - 1];
    J->q[k] = q;
    for (int l = 0; l < 64; ++l) {
        J->qw[64 * k + l] = q;
        J->qwid[64 * k + l] = l == 0 ? k : -1;
    }
}

// part[c * S + k] = sum over points of chunk c of -(x_j - q_k)/|x_j - q_k|^2 (FP64-grade terms)
__global__ __launch_bounds__(JUDGE_S) void k_judge_exact(const double2 *__restrict__ pts, int n, int S,
                                                          JudgeBuf *__restrict__ J) {
    typedef const __attribute__((address_space(4))) double cdbl;
    const int k = threadIdx.x, c = blockIdx.x;
    const int ch = (n + JUDGE_P - 1) / JUDGE_P;
    const int j0 = min(n, c * ch), j1 = min(n, j0 + ch);
    const double2 q = J->q[k < S ? k : 0];
    cdbl *cp = (cdbl *)pts;
    double ax = 0.0, ay = 0.0;
    int j = j0;
    for (; j + 8 <= j1; j += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) pair_rep<true>(cp[2 * (j + u)], cp[2 * (j + u) + 1], q.x, q.y, ax, ay);
    }
    for (; j < j1; ++j) pair_rep<true>(cp[2 * j], cp[2 * j + 1], q.x, q.y, ax, ay);
    if (k < S) J->part[c * JUDGE_S + k] = make_double2(ax, ay);
}

// Medians by rank (ties broken by index) and the verdict.
__global__ __launch_bounds__(JUDGE_S) void k_judge_decide(int S, JudgeBuf *J) {
    __shared__ float e[2][JUDGE_S];
    const int k = threadIdx.x;
    if (k < S) {
        double ex = 0.0, ey = 0.0;
        for (int c = 0; c < JUDGE_P; ++c) {  // chunk order; the threads read consecutive entries
            ex -= J->part[c * JUDGE_S + k].x;
            ey -= J->part[c * JUDGE_S + k].y;
        }
        const double m = sqrt(ex * ex + ey * ey);
        for (int t = 0; t < 2; ++t) {
            const double2 f = J->field[t * JUDGE_S + k];
            const double dx = f.x - ex, dy = f.y - ey;
            e[t][k] = m > 0.0 ? (float)(sqrt(dx * dx + dy * dy) / m) : 0.0f;
        }
    }
    __syncthreads();
    if (k < S) {
        for (int t = 0; t < 2; ++t) {
            int rank = 0;
            for (int j = 0; j < S; ++j) rank += (e[t][j] < e[t][k]) || (e[t][j] == e[t][k] && j < k);
            if (rank == S / 2) J->err[t] = e[t][k];
        }
    }
    __syncthreads();
    if (k == 0) J->use_model = J->err[0] <= J->err[1] ? 1 : 0;
}

}  // namespace

static int ensure_far(BHWork &w, int n) {
    if (w.far_cells.ensure(sizeof(FarCell) * (size_t)std::max(n, 1)) || w.far_nlc.ensure(4 * (size_t)std::max(n, 1)))
        return BL_E_NOMEM;
    return BL_OK;
}

// cell summaries of the tree already built in w
int far_cells_of(BHWork &w, int n, hipStream_t s) {
    int rc = ensure_far(w, n);
    if (rc) return rc;
    if (n >= 2) {
        // the arrival counters (BHNode.cnt) were zeroed by the tree build, which forms no centroids
        hipLaunchKernelGGL(k_far_direct, dim3((n + WG - 1) / WG), dim3(WG), 0, s, w.node_child.as<BHNode>(),
                           w.spos.as<double2>(), w.node_cnt.as<int>(), w.node_lo.as<int>(), w.misc.as<int>(),
                           w.far_cells.as<FarCell>(), w.far_nlc.as<int>());
        hipLaunchKernelGGL(k_far_top, dim3((n + WG - 1) / WG), dim3(WG), 0, s, w.node_child.as<BHNode>(),
                           w.spos.as<double2>(), w.node_cnt.as<int>(), w.misc.as<int>(), w.far_nlc.as<int>(),
                           w.far_cells.as<FarCell>());
        BL_HIP(hipGetLastError());
    }
    return BL_OK;
}

int far_judge(BHWork &w, const double2 *d_pos, int n, double theta, const float *model, int it, hipStream_t s,
              int *verdict) {
    if (w.judge.ensure(sizeof(JudgeBuf))) return BL_E_NOMEM;
    JudgeBuf *J = w.judge.as<JudgeBuf>();
    const int S = std::min(JUDGE_S, n);
    const int stride = std::max(1, n / S);
    const int offset = (int)(((long long)it * 7919) % stride);
    hipLaunchKernelGGL(k_judge_sample, dim3(1), dim3(JUDGE_S), 0, s, d_pos, n, S, offset, J);
    hipLaunchKernelGGL(k_judge_exact, dim3(JUDGE_P), dim3(JUDGE_S), 0, s, d_pos, n, S, J);
    BL_HIP(hipGetLastError());
    int rc = launch_trav(w, J->qw, J->qwid, 64 * S, theta, FAR_LEARNED, model, J->field, s);
    if (rc) return rc;
    rc = launch_trav(w, J->qw, J->qwid, 64 * S, theta, FAR_QUAD, model, J->field + JUDGE_S, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_judge_decide, dim3(1), dim3(JUDGE_S), 0, s, S, J);
    BL_HIP(hipGetLastError());
    // the host launches the traversal the verdict names: the analytic kernel alone holds fewer
    // registers than the learned one with its analytic branch (one wait per judged iteration)
    BL_HIP(hipMemcpyAsync(verdict, &J->use_model, sizeof(int), hipMemcpyDeviceToHost, s));
    BL_HIP(hipStreamSynchronize(s));
    return BL_OK;
}

int far_judge_reset(BHWork &w, hipStream_t s) {
    if (w.judge.ensure(sizeof(JudgeBuf))) return BL_E_NOMEM;
    JudgeBuf *J = w.judge.as<JudgeBuf>();
    const int init[2] = {1, 0};  // use the model until the first verdict; no traversal counted yet
    BL_HIP(hipMemcpyAsync(&J->use_model, init, sizeof(init), hipMemcpyHostToDevice, s));
    BL_HIP(hipMemsetAsync(J->err, 0, sizeof(J->err), s));
    return BL_OK;
}

int far_judge_read(BHWork &w, int *model_iters, float *err, hipStream_t s) {
    JudgeBuf *J = w.judge.as<JudgeBuf>();
    BL_HIP(hipMemcpyAsync(model_iters, &J->model_iters, sizeof(int), hipMemcpyDeviceToHost, s));
    BL_HIP(hipMemcpyAsync(err, J->err, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
    return BL_OK;
}

int far_repulsion(BHWork &w, const double2 *d_pos, int n, double theta, int mode, const float *model,
                  hipStream_t s, unsigned long long *ncells, bool judged) {
    if (mode == FAR_EXACT) {
        // by vertex id, straight from the positions
        hipLaunchKernelGGL(k_far_exact, dim3((n + WG - 1) / WG), dim3(WG), 0, s, d_pos, nullptr, n, d_pos, n,
                           w.rep.as<double2>());
        BL_HIP(hipGetLastError());
        return BL_OK;
    }
    // the tree of d_pos and its cell summaries (far_cells_of) are already in w; queries in
    // Morton order
    JudgeBuf *J = judged ? w.judge.as<JudgeBuf>() : nullptr;
    return launch_trav(w, w.spos.as<double2>(), w.idx_a.as<int>(), n, theta, mode, model, w.rep.as<double2>(), s,
                       ncells, J ? &J->use_model : nullptr, J ? &J->model_iters : nullptr);
}

// Host entry: far-field forces of nq query points against the n points xy.
int far_forces_host(int32_t n, const double *xy, double theta, int32_t mode, const float *model, int32_t nq,
                    const double *q, double *out) {
    HostRun G;
    BL_HIP(hipStreamCreateWithFlags(&G.s, hipStreamNonBlocking));
    if (G.pts.ensure(16 * (size_t)n) || G.qs.ensure(16 * (size_t)std::max(nq, 1)) ||
        G.qi.ensure(4 * (size_t)std::max(nq, 1)) || G.res.ensure(16 * (size_t)std::max(nq, 1)) ||
        G.mdl.ensure(4 * (size_t)MLP_NPARAM))
        return BL_E_NOMEM;
    if (nq == 0) return BL_OK;
    BL_HIP(hipMemcpyAsync(G.pts.p, xy, 16 * (size_t)n, hipMemcpyHostToDevice, G.s));
    BL_HIP(hipMemcpyAsync(G.qs.p, q, 16 * (size_t)nq, hipMemcpyHostToDevice, G.s));
    std::vector<int> ids(nq);
    for (int i = 0; i < nq; ++i) ids[i] = i;
    BL_HIP(hipMemcpyAsync(G.qi.p, ids.data(), 4 * (size_t)nq, hipMemcpyHostToDevice, G.s));
    if (mode == FAR_EXACT) {
        hipLaunchKernelGGL(k_far_exact, dim3((nq + WG - 1) / WG), dim3(WG), 0, G.s, G.qs.as<double2>(), G.qi.as<int>(),
                           nq, G.pts.as<double2>(), n, G.res.as<double2>());
        BL_HIP(hipGetLastError());
    } else {
        if (mode == FAR_LEARNED) {
            if (!model) return BL_E_ARG;
            BL_HIP(hipMemcpyAsync(G.mdl.p, model, 4 * (size_t)MLP_NPARAM, hipMemcpyHostToDevice, G.s));
        }
        int rc = bh_build_tree(G.w, G.pts.as<double2>(), n, G.s);
        if (rc) return rc;
        rc = far_cells_of(G.w, n, G.s);
        if (rc) return rc;
        rc = launch_trav(G.w, G.qs.as<double2>(), G.qi.as<int>(), nq, theta, mode, G.mdl.as<float>(),
                         G.res.as<double2>(), G.s);
        if (rc) return rc;
        rc = G.status();
        if (rc) return rc;
    }
    BL_HIP(hipMemcpyAsync(out, G.res.p, 16 * (size_t)nq, hipMemcpyDeviceToHost, G.s));
    BL_HIP(hipStreamSynchronize(G.s));
    return BL_OK;
}

// Host entry: training samples of nq query points (cap cells each) against the n points xy.
int far_samples_host(int32_t n, const double *xy, double theta, int32_t nq, const double *q, int32_t cap,
                     float *feat, float *targ, int32_t *count) {
    HostRun G;
    BL_HIP(hipStreamCreateWithFlags(&G.s, hipStreamNonBlocking));
    const size_t rows = (size_t)nq * (size_t)cap;
    if (G.pts.ensure(16 * (size_t)n) || G.qs.ensure(16 * (size_t)std::max(nq, 1)) ||
        G.f.ensure(4 * MLP_IN * std::max<size_t>(rows, 1)) || G.t.ensure(8 * std::max<size_t>(rows, 1)) ||
        G.c.ensure(4 * (size_t)std::max(nq, 1)))
        return BL_E_NOMEM;
    if (nq == 0) return BL_OK;
    BL_HIP(hipMemcpyAsync(G.pts.p, xy, 16 * (size_t)n, hipMemcpyHostToDevice, G.s));
    BL_HIP(hipMemcpyAsync(G.qs.p, q, 16 * (size_t)nq, hipMemcpyHostToDevice, G.s));
    int rc = bh_build_tree(G.w, G.pts.as<double2>(), n, G.s);
    if (rc) return rc;
    rc = far_cells_of(G.w, n, G.s);
    if (rc) return rc;
    BL_HIP(hipMemsetAsync(G.c.p, 0, 4 * (size_t)nq, G.s));
    const int g64 = ((nq + 63) / 64 + WG / 64 - 1) / (WG / 64);
    hipLaunchKernelGGL(k_far_samples, dim3(g64), dim3(WG), 0, G.s, G.qs.as<double2>(), nq, G.w.spos.as<double2>(),
                       G.w.node_child.as<BHNode>(), G.w.node_lo.as<int>(), G.w.far_cells.as<FarCell>(),
                       G.w.misc.as<int>(), G.w.frame.as<double>(), theta, cap, G.f.as<float>(), G.t.as<float>(),
                       G.c.as<int>());
    BL_HIP(hipGetLastError());
    rc = G.status();
    if (rc) return rc;
    BL_HIP(hipMemcpyAsync(feat, G.f.p, 4 * MLP_IN * rows, hipMemcpyDeviceToHost, G.s));
    BL_HIP(hipMemcpyAsync(targ, G.t.p, 8 * rows, hipMemcpyDeviceToHost, G.s));
    BL_HIP(hipMemcpyAsync(count, G.c.p, 4 * (size_t)nq, hipMemcpyDeviceToHost, G.s));
    BL_HIP(hipStreamSynchronize(G.s));
    return BL_OK;
}

}  // namespace bl
