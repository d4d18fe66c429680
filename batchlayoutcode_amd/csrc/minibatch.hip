// Minibatch force step of the BatchLayout cache-blocking family on CDNA4 (gfx950).
//
// Reference: sample/algorithms.cpp:715-794 (algo 2), :795-871 (converged variant),
// :874-945 (algo 3, LinLog), :947-1018 (algo 8).  One iteration visits the vertices
// in ceil(n/B) consecutive batches (Gauss-Seidel across batches); inside a batch
// every vertex i sees the positions from the start of the batch step (Jacobi):
//
//   F_i = sum_{j in N(i)} A(x_j - x_i)  -  sum_{j not in N(i), |x_j-x_i|>0} (x_j - x_i)/|x_j - x_i|^2
//   x_i += STEP * F_i/|F_i|,   ENERGY += |F_i|^2,   STEP *= decay after each iteration
//
// where N(i) is the set of column ids the reference's merge pointer meets (:758-762)
// and A is the attraction law (W|d| d, W d, or log2(1+W|d|) d).
//
// Two kernels per batch step:
//   k_pairs32 / k_pairs64  all-pairs repulsion for the batch against all n vertices,
//                          j split into chunks over workgroups, one partial per chunk
//   k_move32  / k_move64   sum of the partials per vertex, unit-step move, |F|^2
// and one energy reduction per iteration.
//
// FP32 mode (the fast path): each lane owns one batch vertex; the j loop is
// wave-uniform, so x_j arrives through scalar loads and every pair costs
// 2 FP64 subtracts + 2 converts, 2 FMA for |d|^2, one v_rcp_f32 and 2 FMA.  Sums run in
// FP32 over 64-j tiles and in FP64 across tiles and chunks.  Neighbours are not
// branched on in the dense loop: one extra workgroup per four batch vertices adds,
// in FP64, the attraction of every neighbour and removes the repulsion term the
// dense loop counted for it (the exact FP64 value of the FP32 product it added).
//
// FP64 mode (reference-grade): the reference's own operations, in its order.  The
// j range is cut into the reference's 512-wide blocks (blocky, :717); each block's
// partial is summed in j order with the merge-pointer branch, and the partials are
// added in block order, so the result reproduces the strict-IEEE reference build.
#include <math.h>

#include "bl_internal.h"

namespace bl {
namespace {

constexpr int WG = 256;
// |d|^2 floor of the FP32 path: the self pair and coincident vertices give
// 0 * (1/kTiny) = 0, as the reference's `dist > 0` test skips them.
constexpr float kTiny = 1e-30f;
constexpr int kBlockY = 512;  // the reference's j block, algorithms.cpp:717

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
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

// Read-only view of the positions in the constant address space: with a wave-uniform
// index the compiler fetches x_j with scalar loads (one fetch per wave, no VGPRs).
typedef const __attribute__((address_space(4))) double cdouble;

__device__ __forceinline__ void pair32(const double qx, const double qy, const double px, const double py, float &tx,
                                       float &ty) {
    const float dx = (float)(qx - px);
    const float dy = (float)(qy - py);
    const float r2 = fmaf(dy, dy, fmaf(dx, dx, kTiny));
    const float inv = __builtin_amdgcn_rcpf(r2);
    tx = fmaf(-dx, inv, tx);
    ty = fmaf(-dy, inv, ty);
}

// Dense FP32 repulsion + neighbour correction.  grid = (G + ncorr, rows).
// Workgroups x < G: batch slots [y*IPW, y*IPW+IPW) against j chunk x (JS j-subgroups).
// Workgroups x >= G (row 0 only): one wave per batch slot, neighbour correction.
// part[slot * P + x] receives the chunk partial; part[slot * P + P-1] the correction.
template <int JS>
__global__ __launch_bounds__(WG) void k_pairs32(const double2 *__restrict__ pos, int n, int b0, int bc,
                                                int chunk, int G, double2 *__restrict__ part, int P,
                                                const int *__restrict__ erp, const int *__restrict__ eci,
                                                const float *__restrict__ ew, double W, int law) {
    __shared__ double2 red[WG];
    const int tid = threadIdx.x;
    if ((int)blockIdx.x >= G) {
        if (blockIdx.y != 0) return;
        const int slot = ((int)blockIdx.x - G) * (WG / 64) + tid / 64;
        const int lane = tid & 63;
        if (slot >= bc) return;
        const int i = b0 + slot;
        const double2 pi = pos[i];
        double cx = 0.0, cy = 0.0;
        const int kend = erp[i + 1];
        for (int k = erp[i] + lane; k < kend; k += 64) {
            const int j = eci[k];
            const double2 pj = pos[j];
            const double dx = pj.x - pi.x, dy = pj.y - pi.y;
            // the dense loop added fma(-dx32, inv32, acc): remove exactly -dx32*inv32
            const float dx32 = (float)dx, dy32 = (float)dy;
            const float r2 = fmaf(dy32, dy32, fmaf(dx32, dx32, kTiny));
            const float inv = __builtin_amdgcn_rcpf(r2);
            cx += (double)dx32 * (double)inv;
            cy += (double)dy32 * (double)inv;
            attract64(law, ew ? W * (double)ew[k] : W, dx, dy, cx, cy);
        }
        cx = wave_sum(cx);
        cy = wave_sum(cy);
        if (lane == 0) part[(size_t)slot * P + (P - 1)] = make_double2(cx, cy);
        return;
    }
    constexpr int IPW = WG / JS;
    const int li = tid % IPW;
    const int js = __builtin_amdgcn_readfirstlane(tid / IPW);  // IPW is a multiple of 64
    cdouble *cpos = (cdouble *)pos;
    const int slot = (int)blockIdx.y * IPW + li;
    const int i = b0 + (slot < bc ? slot : bc - 1);
    const double px = pos[i].x, py = pos[i].y;
    const int jb = (int)blockIdx.x * chunk;
    const int je = min(n, jb + chunk);
    const int sub = (chunk + JS - 1) / JS;
    const int j0 = min(je, jb + js * sub);
    const int j1 = min(je, j0 + sub);
    double ax = 0.0, ay = 0.0;
    for (int t0 = j0; t0 < j1; t0 += 64) {
        const int t1 = min(j1, t0 + 64);
        float tx = 0.f, ty = 0.f;
        int u = t0;
        for (; u + 8 <= t1; u += 8) {
#pragma unroll
            for (int k = 0; k < 8; ++k) pair32(cpos[2 * (u + k)], cpos[2 * (u + k) + 1], px, py, tx, ty);
        }
        for (; u < t1; ++u) pair32(cpos[2 * u], cpos[2 * u + 1], px, py, tx, ty);
        ax += (double)tx;
        ay += (double)ty;
    }
    if constexpr (JS > 1) {
        red[tid] = make_double2(ax, ay);
        __syncthreads();
        if (js == 0) {
#pragma unroll
            for (int s = 1; s < JS; ++s) {
                ax += red[s * IPW + li].x;
                ay += red[s * IPW + li].y;
            }
        }
    }
    if (js == 0 && slot < bc) part[(size_t)slot * P + blockIdx.x] = make_double2(ax, ay);
}

// FP64 reference-order pair kernel.  grid = (ceil(n/512), rows); block x covers the
// reference's j block [512x, 512x+512), one lane per batch vertex.
__global__ __launch_bounds__(WG) void k_pairs64(const double2 *__restrict__ pos, int n, int b0, int bc,
                                                double2 *__restrict__ part, int P, const int *__restrict__ erp,
                                                const int *__restrict__ eci, const float *__restrict__ ew,
                                                double W, int law) {
    const int slot = (int)blockIdx.y * WG + threadIdx.x;
    if (slot >= bc) return;
    const int i = b0 + slot;
    const int jb = (int)blockIdx.x * kBlockY;
    const int je = min(n, jb + kBlockY);
    const double vx = pos[i].x, vy = pos[i].y;
    int lo = erp[i], hi = erp[i + 1];
    const int kend = hi;
    while (lo < hi) {  // first neighbour >= jb (the merge pointer's state at the block start)
        const int m = (lo + hi) >> 1;
        if (eci[m] < jb) lo = m + 1; else hi = m;
    }
    int k = lo;
    int nb = k < kend ? eci[k] : -1;
    double fx = 0.0, fy = 0.0;
    for (int u = jb; u < je; ++u) {
        const double2 q = pos[u];
        const double dx = q.x - vx, dy = q.y - vy;
        if (u == nb) {  // :758-762
            attract64(law, ew ? W * (double)ew[k] : W, dx, dy, fx, fy);
            ++k;
            nb = k < kend ? eci[k] : -1;
        } else {  // :764-768
            const double dist = dx * dx + dy * dy;
            if (dist > 0) {
                const double inv = 1.0 / dist;
                fx = fx - dx * inv;
                fy = fy - dy * inv;
            }
        }
    }
    part[(size_t)slot * P + blockIdx.x] = make_double2(fx, fy);
}

// x_i += (F/|F|) * STEP and |F|^2, algorithms.cpp:775-779 / Coordinate.h:57-65.
__device__ __forceinline__ void move_one(double2 *pos, int i, double fx, double fy, double step, double *eterm) {
    const double mag = sqrt(fx * fx + fy * fy);
    double2 p = pos[i];
    p.x = p.x + (fx / mag) * step;
    p.y = p.y + (fy / mag) * step;
    pos[i] = p;
    eterm[i] = fx * fx + fy * fy;
}

// FP32 mode: one workgroup per batch vertex sums its P partials (FP64, fixed order:
// strided per thread, then a fixed tree), so every partial row is read by 256 lanes at once.
__global__ __launch_bounds__(WG) void k_move32(const double2 *__restrict__ part, int P, double2 *__restrict__ pos,
                                               int b0, int bc, double step, double *__restrict__ eterm) {
    __shared__ double2 red[WG / 64];
    const int slot = (int)blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (slot >= bc) return;
    const double2 *row = part + (size_t)slot * P;
    double sx = 0.0, sy = 0.0;
    for (int p = tid; p < P; p += WG) {
        const double2 v = row[p];
        sx += v.x;
        sy += v.y;
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (lane == 0) red[w] = make_double2(sx, sy);
    __syncthreads();
    if (tid == 0) {
        const double fx = (red[0].x + red[1].x) + (red[2].x + red[3].x);
        const double fy = (red[0].y + red[1].y) + (red[2].y + red[3].y);
        move_one(pos, b0 + slot, fx, fy, step, eterm);
    }
}

// FP64 mode: prevCoordinates[i] += f per j block in block order (:771), then the move.
__global__ __launch_bounds__(WG) void k_move64(const double2 *__restrict__ part, int P, double2 *__restrict__ pos,
                                               int b0, int bc, double step, double *__restrict__ eterm) {
    const int slot = (int)blockIdx.x * WG + threadIdx.x;
    if (slot >= bc) return;
    const double2 *row = part + (size_t)slot * P;
    double sx = 0.0, sy = 0.0;
    for (int p = 0; p < P; ++p) {
        sx += row[p].x;
        sy += row[p].y;
    }
    move_one(pos, b0 + slot, sx, sy, step, eterm);
}

}  // namespace

// ---- energy reductions (shared with the Barnes-Hut path) -------------------------------
// Fixed-order tree sum of |F_i|^2 (FP32 mode).
__global__ __launch_bounds__(1024) void k_energy_tree(const double *__restrict__ eterm, int n, double *out) {
    __shared__ double s[1024];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) a += eterm[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// Sequential sum in vertex order, ENERGY += |F_i|^2 as the reference (FP64 mode).
__global__ void k_energy_seq(const double *__restrict__ eterm, int n, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double e = 0.0;
    for (int i = 0; i < n; ++i) e += eterm[i];
    *out = e;
}

int launch_energy(const double *eterm, int n, double *out, int precision, hipStream_t s) {
    if (precision == BL_PREC_FP64)
        hipLaunchKernelGGL(k_energy_seq, dim3(1), dim3(64), 0, s, eterm, n, out);
    else
        hipLaunchKernelGGL(k_energy_tree, dim3(1), dim3(1024), 0, s, eterm, n, out);
    BL_HIP(hipGetLastError());
    return BL_OK;
}

int run_minibatch(bl_graph *g, double *d_xy, const bl_params *p, int law, double decay, double step0, int iters,
                  bool converged, bl_stats *st) {
    const int n = g->n;
    const int B = p->batch;
    const int nb = (n + B - 1) / B;
    const int bmax = B < n ? B : n;
    const bool fp64 = p->precision == BL_PREC_FP64;
    hipStream_t s = g->stream;
    double2 *pos = reinterpret_cast<double2 *>(d_xy);
    const int *erp = g->eff_rowptr.as<int>();
    const int *eci = g->eff_colids.as<int>();
    const float *ew = g->has_w ? g->eff_w.as<float>() : nullptr;

    // launch geometry
    int JS = 1, rows = 1, G = 1, chunk = n, P = 1, ncorr = 0;
    if (fp64) {
        G = (n + kBlockY - 1) / kBlockY;
        P = G;
        rows = (bmax + WG - 1) / WG;
    } else {
        JS = bmax <= 64 ? 4 : (bmax <= 128 ? 2 : 1);
        const int ipw = WG / JS;
        rows = (bmax + ipw - 1) / ipw;
        const int target = (2048 + rows - 1) / rows;  // ~2048 workgroups per step
        const int quant = 8 * JS;
        chunk = (n + target - 1) / target;
        chunk = ((chunk + quant - 1) / quant) * quant;
        if (chunk < 4 * quant) chunk = 4 * quant;
        G = (n + chunk - 1) / chunk;
        P = G + 1;
        ncorr = (bmax + 3) / 4;
    }
    if (g->part.ensure(sizeof(double2) * (size_t)bmax * P)) return BL_E_NOMEM;
    if (g->eterm.ensure(sizeof(double) * (size_t)n)) return BL_E_NOMEM;
    if (g->energy.ensure(sizeof(double) * 4)) return BL_E_NOMEM;
    double2 *part = g->part.as<double2>();
    double *eterm = g->eterm.as<double>();
    double *energy = g->energy.as<double>();

    Prof prof;
    prof.on = p->profile != 0;
    prof.s = s;
    double STEP = step0;
    double E = 0.0, E0 = 1.79769313486231570815e+308;
    int loop = 0;
    for (;;) {
        if (converged) {
            if (!(fabs(E0 - E) / E0 > p->engt)) break;  // algorithms.cpp:818
        } else if (loop >= iters) {
            break;
        }
        E0 = E;
        for (int b = 0; b < nb; ++b) {
            const int b0 = b * B;
            const int bc = (n - b0) < B ? (n - b0) : B;
            if (fp64) {
                prof.begin(0);
                hipLaunchKernelGGL(k_pairs64, dim3(G, (bc + WG - 1) / WG), dim3(WG), 0, s, pos, n, b0, bc, part, P,
                                   erp, eci, ew, p->weight, law);
                prof.end();
                prof.begin(1);
                hipLaunchKernelGGL(k_move64, dim3((bc + WG - 1) / WG), dim3(WG), 0, s, part, P, pos, b0, bc, STEP,
                                   eterm);
                prof.end();
            } else {
                const int ipw = WG / JS;
                const dim3 grid(G + (bc + 3) / 4, (bc + ipw - 1) / ipw);
                prof.begin(0);
                if (JS == 4)
                    hipLaunchKernelGGL(k_pairs32<4>, grid, dim3(WG), 0, s, pos, n, b0, bc, chunk, G, part, P, erp,
                                       eci, ew, p->weight, law);
                else if (JS == 2)
                    hipLaunchKernelGGL(k_pairs32<2>, grid, dim3(WG), 0, s, pos, n, b0, bc, chunk, G, part, P, erp,
                                       eci, ew, p->weight, law);
                else
                    hipLaunchKernelGGL(k_pairs32<1>, grid, dim3(WG), 0, s, pos, n, b0, bc, chunk, G, part, P, erp,
                                       eci, ew, p->weight, law);
                prof.end();
                prof.begin(1);
                hipLaunchKernelGGL(k_move32, dim3(bc), dim3(WG), 0, s, part, P, pos, b0, bc, STEP, eterm);
                prof.end();
            }
        }
        (void)ncorr;
        prof.begin(2);
        int rc = launch_energy(eterm, n, energy, p->precision, s);
        prof.end();
        if (rc) return rc;
        prof.flush();
        STEP = STEP * decay;  // :781 (0.999), :934 / :1007 (0.99)
        ++loop;
        if (converged) {
            BL_HIP(hipMemcpyAsync(&E, energy, sizeof(double), hipMemcpyDeviceToHost, s));
            BL_HIP(hipStreamSynchronize(s));
        }
    }
    BL_HIP(hipGetLastError());
    if (!converged && loop > 0) {
        BL_HIP(hipMemcpyAsync(&E, energy, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    BL_HIP(hipStreamSynchronize(s));
    prof.report(st);
    if (st) {
        st->energy = loop > 0 ? E : (law == LAW_FR ? 0.0 : 1.79769313486231570815e+308);
        st->iters_done = loop;
        st->step_final = STEP;
    }
    return BL_OK;
}

}  // namespace bl
