// Physically consistent far field for the Barnes-Hut layout (no reference counterpart).
//
// The reference's Barnes-Hut (BarnesHut.h:69-86) adds one unit vector per accepted cell,
// whatever the cell holds.  The layouts here use the repulsion law of algo 2
// (algorithms.cpp:764-768: a pair contributes d/|d|^2) for every vertex pair and approximate
// the far field of a cell by its mass-weighted monopole m (c - v)/|c - v|^2 (c = mean of the
// cell's points, m = their count), or by the monopole corrected by a learned model of the
// cell's shape.  The exact sum over all points is the ground truth of both.
//
//   k_far_moments  count, mean, central second moments and first Morton rank of every cell,
//                  bottom-up over the tree of bh.hip (last child to arrive computes the parent)
//   k_far_trav     traversal, one wave per 64 Morton-consecutive queries with a shared stack
//                  of (node, lane mask); a cell is accepted when side^2 < theta^2 |c - v|^2
//   k_far_exact    exact sum over all points for every query
#include <math.h>

#include <algorithm>

#include "bh_common.h"
#include "pair_math.h"

namespace bl {
namespace {

constexpr int WG = 256;
constexpr int STACK = 128;

struct FarCell {
    double cx, cy;          // mean of the cell's points
    double mxx, mxy, myy;   // central second moments, sums over the points
    int m;                  // number of points
    int lo;                 // first Morton rank; the points are ranks [lo, lo + m)
};

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Combines the non-empty quadrants of node v (children first): mean by counts, second
// moments by the parallel-axis rule, in quadrant order.
__device__ void far_summary(const BHNode *nodes, const double2 *spos, FarCell *cells, int v) {
    const int4 ch = nodes[v].ch;
    const int c[4] = {ch.x, ch.y, ch.z, ch.w};
    double m = 0.0, sx = 0.0, sy = 0.0;
    int lo = 0x7fffffff;
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
            lo = min(lo, k.lo);
        } else if (c[q] <= -2) {
            const double2 p = spos[-c[q] - 2];
            km[q] = 1.0;
            kx[q] = p.x;
            ky[q] = p.y;
            lo = min(lo, -c[q] - 2);
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
    out.lo = lo;
    cells[v] = out;
}

__global__ __launch_bounds__(WG) void k_far_moments(const BHNode *__restrict__ nodes, const double2 *__restrict__ spos,
                                                    const int *__restrict__ misc, FarCell *__restrict__ cells,
                                                    int *__restrict__ arrive) {
    int v = blockIdx.x * WG + threadIdx.x;
    if (misc[0] || v >= misc[2]) return;
    if (nodes[v].nint != 0) return;
    for (;;) {
        far_summary(nodes, spos, cells, v);
        const int u = nodes[v].parent;
        if (u < 0) break;
        const int old = __hip_atomic_fetch_add(&arrive[u], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old != nodes[u].nint - 1) break;
        v = u;
    }
}

// Repulsion F(v) = sum over accepted cells of m (c - v)/|c - v|^2 + sum over the single points
// met of (p - v)/|p - v|^2.  The opening test and the terms are FP64-grade (FP64 differences,
// FP32-seeded reciprocal with an FP64 Newton step); sums in FP64.
__global__ __launch_bounds__(WG) void k_far_trav(const double2 *__restrict__ qpos, const int *__restrict__ qid,
                                                 int nq, const double2 *__restrict__ spos,
                                                 const BHNode *__restrict__ nodes, const FarCell *__restrict__ cells,
                                                 const int *__restrict__ misc, const double *__restrict__ frame,
                                                 double theta, double2 *__restrict__ out) {
    __shared__ int stk_node[WG / 64][STACK];
    __shared__ uint64_t stk_mask[WG / 64][STACK];
    if (misc[0]) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k = (blockIdx.x * (WG / 64) + w) * 64 + lane;
    const bool valid = k < nq;
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
        if (act && acc) {
            const double s = (double)cl.m * inv_r2<true>(dx, dy);
            fx = fma(dx, s, fx);
            fy = fma(dy, s, fy);
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
    if (valid) out[qid[k]] = make_double2(fx, fy);
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

int ensure_far(BHWork &w, int n) {
    if (w.far_cells.ensure(sizeof(FarCell) * (size_t)std::max(n, 1)) || w.far_arrive.ensure(4 * (size_t)std::max(n, 1)))
        return BL_E_NOMEM;
    return BL_OK;
}

}  // namespace

int far_repulsion(BHWork &w, const double2 *d_pos, int n, double theta, int mode, const float *model,
                  hipStream_t s) {
    (void)model;
    const int g64 = ((n + 63) / 64 + WG / 64 - 1) / (WG / 64);
    if (mode == FAR_EXACT) {
        // by vertex id, straight from the positions
        hipLaunchKernelGGL(k_far_exact, dim3((n + WG - 1) / WG), dim3(WG), 0, s, d_pos, nullptr, n, d_pos, n,
                           w.rep.as<double2>());
        BL_HIP(hipGetLastError());
        return BL_OK;
    }
    if (mode != FAR_MONO) return BL_E_ARG;
    // the tree of d_pos is already in w (bh_build_tree)
    int rc = ensure_far(w, n);
    if (rc) return rc;
    BHNode *nodes = w.node_child.as<BHNode>();
    FarCell *cells = w.far_cells.as<FarCell>();
    const int *misc = w.misc.as<int>();
    if (n >= 2) {
        BL_HIP(hipMemsetAsync(w.far_arrive.p, 0, 4 * (size_t)n, s));
        hipLaunchKernelGGL(k_far_moments, dim3((n + WG - 1) / WG), dim3(WG), 0, s, nodes, w.spos.as<double2>(), misc,
                           cells, w.far_arrive.as<int>());
    }
    hipLaunchKernelGGL(k_far_trav, dim3(g64), dim3(WG), 0, s, w.spos.as<double2>(), w.idx_a.as<int>(), n,
                       w.spos.as<double2>(), nodes, cells, misc, w.frame.as<double>(), theta, w.rep.as<double2>());
    BL_HIP(hipGetLastError());
    return BL_OK;
}

// Host entry: far-field forces of nq query points against the n points xy.
int far_forces_host(int32_t n, const double *xy, double theta, int32_t mode, int32_t nq, const double *q,
                    double *out) {
    hipStream_t s = nullptr;
    BL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct Guard {
        hipStream_t s;
        BHWork w;
        DevBuf pts, qs, qi, res;
        ~Guard() {
            (void)hipStreamSynchronize(s);
            w.release();
            pts.release();
            qs.release();
            qi.release();
            res.release();
            (void)hipStreamDestroy(s);
        }
    } G{s, {}, {}, {}, {}, {}};
    if (G.pts.ensure(16 * (size_t)n) || G.qs.ensure(16 * (size_t)std::max(nq, 1)) ||
        G.qi.ensure(4 * (size_t)std::max(nq, 1)) || G.res.ensure(16 * (size_t)std::max(nq, 1)))
        return BL_E_NOMEM;
    BL_HIP(hipMemcpyAsync(G.pts.p, xy, 16 * (size_t)n, hipMemcpyHostToDevice, s));
    if (nq == 0) return BL_OK;
    BL_HIP(hipMemcpyAsync(G.qs.p, q, 16 * (size_t)nq, hipMemcpyHostToDevice, s));
    std::vector<int> ids(nq);
    for (int i = 0; i < nq; ++i) ids[i] = i;
    BL_HIP(hipMemcpyAsync(G.qi.p, ids.data(), 4 * (size_t)nq, hipMemcpyHostToDevice, s));
    if (mode == FAR_EXACT) {
        hipLaunchKernelGGL(k_far_exact, dim3((nq + WG - 1) / WG), dim3(WG), 0, s, G.qs.as<double2>(), G.qi.as<int>(),
                           nq, G.pts.as<double2>(), n, G.res.as<double2>());
        BL_HIP(hipGetLastError());
    } else if (mode == FAR_MONO) {
        int rc = bh_build_tree(G.w, G.pts.as<double2>(), n, s);
        if (rc) return rc;
        rc = ensure_far(G.w, n);
        if (rc) return rc;
        const int *misc = G.w.misc.as<int>();
        if (n >= 2) {
            BL_HIP(hipMemsetAsync(G.w.far_arrive.p, 0, 4 * (size_t)n, s));
            hipLaunchKernelGGL(k_far_moments, dim3((n + WG - 1) / WG), dim3(WG), 0, s, G.w.node_child.as<BHNode>(),
                               G.w.spos.as<double2>(), misc, G.w.far_cells.as<FarCell>(), G.w.far_arrive.as<int>());
        }
        const int g64 = ((nq + 63) / 64 + WG / 64 - 1) / (WG / 64);
        hipLaunchKernelGGL(k_far_trav, dim3(g64), dim3(WG), 0, s, G.qs.as<double2>(), G.qi.as<int>(), nq,
                           G.w.spos.as<double2>(), G.w.node_child.as<BHNode>(), G.w.far_cells.as<FarCell>(), misc,
                           G.w.frame.as<double>(), theta, G.res.as<double2>());
        BL_HIP(hipGetLastError());
        int status = 0;
        BL_HIP(hipMemcpyAsync(&status, misc, sizeof(int), hipMemcpyDeviceToHost, s));
        BL_HIP(hipStreamSynchronize(s));
        if (status & ST_DUPKEY) return BL_E_DUPKEY;
    } else {
        return BL_E_ARG;
    }
    BL_HIP(hipMemcpyAsync(out, G.res.p, 16 * (size_t)nq, hipMemcpyDeviceToHost, s));
    BL_HIP(hipStreamSynchronize(s));
    return BL_OK;
}

}  // namespace bl
