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
// FP32 mode (the fast path).  Per pair: the FP64 difference d = x_j - x_i, |d|^2 and
// 1/|d|^2 in FP32 (v_rcp_f32), and the term -d/|d|^2 accumulated in FP64 (one FMA per
// component with the FP64 difference).  The dense loop runs over all j; neighbours are not
// branched on: a correction adds, in FP64, each neighbour's attraction and removes the
// repulsion product the dense loop accumulated for it (the same operands, so it cancels).
//   * k_pairs32 + k_move32 (default): two launches per batch step.  The pair kernel's
//     chunk workgroups write each batch vertex's chunk partials; its correction waves take
//     the neighbour corrections in pieces of at most CORR_PIECE entries (a hub's
//     correction is spread over many waves); k_move32 adds them in a fixed order and moves.
//   * With BL_GRAPH=1, one iteration's launches (every batch step, the energy, STEP *= decay
//     on the device) are captured once as a hipGraph and replayed, one host launch per
//     iteration (measured no faster on this GPU, so not the default).
//
// FP64 mode (reference-grade): the reference's own operations, in its order.  The
// j range is cut into the reference's 512-wide blocks (blocky, :717); each block's
// partial is summed in j order with the merge-pointer branch, and the partials are
// added in block order, so the result reproduces the strict-IEEE reference build.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include <hip/hip_ext.h>

#include "bl_internal.h"
#include "pair_math.h"

namespace bl {
namespace {

constexpr int WG = 256;
constexpr int kBlockY = 512;  // the reference's j block, algorithms.cpp:717
constexpr int UNROLL = 16;

// Read-only view of the positions in the constant address space: with a wave-uniform
// index the compiler fetches x_j with scalar loads (one fetch per wave, no VGPRs).  Only
// used by kernels during which the positions do not change.
typedef const __attribute__((address_space(4))) double cdouble;

// x_i += (F/|F|) * STEP and |F|^2, algorithms.cpp:775-779 / Coordinate.h:57-65.
__device__ __forceinline__ void move_one(double2 *pos, int i, double fx, double fy, double step, double *eterm) {
    pos[i] = moved(pos[i], fx, fy, step);
    eterm[i] = fx * fx + fy * fy;
}

// ---------------------------------------------------------------------------------------
// Per-step kernels: the FP32 fallback and the FP64 reference-order path

// Dense FP32 repulsion + neighbour correction (one launch per batch step).
// grid = (G + ncorr, rows), 256 threads.  Workgroups x < G: batch slots [y*IPW, y*IPW+IPW)
// against j chunk x, split over JS = 256/IPW j-subgroups added in LDS in subgroup order.
// Workgroups x >= G (row 0 only): one wave per correction piece of the batch (pieces
// [pc0, pc0+npc): at most CORR_PIECE neighbour entries of one vertex, so a hub's correction
// is spread over many waves instead of holding up the step).
// part[slot * P + x] receives the chunk partial; part[slot * P + G + k] piece k's correction.
template <int JS, bool REFINE>
__global__ __launch_bounds__(WG) void k_pairs32(const double2 *__restrict__ pos, int n, int b0, int bc, int chunk,
                                                int G, double2 *__restrict__ part, int P,
                                                const int *__restrict__ erp, const int *__restrict__ eci,
                                                const float *__restrict__ ew, double W, int law,
                                                const int *__restrict__ cpre, const int *__restrict__ cpv, int pc0,
                                                int npc) {
    constexpr int IPW = WG / JS;
    __shared__ double2 red[WG];
    const int tid = threadIdx.x;
    if ((int)blockIdx.x >= G) {
        if (blockIdx.y != 0) return;
        const int w = ((int)blockIdx.x - G) * (WG / 64) + tid / 64;
        const int lane = tid & 63;
        if (w >= npc) return;
        const int pc = pc0 + w;
        const int i = cpv[pc], k = pc - cpre[i];
        const double2 pi = pos[i];
        double cx = 0.0, cy = 0.0;
        const int kbeg = erp[i] + k * CORR_PIECE;
        const int kend = min(erp[i + 1], kbeg + CORR_PIECE);
        for (int e = kbeg + lane; e < kend; e += 64) {
            const double2 pj = pos[eci[e]];
            correct_pair<REFINE>(law, ew ? W * (double)ew[e] : W, pj.x - pi.x, pj.y - pi.y, cx, cy);
        }
        cx = wave_sum(cx);
        cy = wave_sum(cy);
        if (lane == 0) part[(size_t)(i - b0) * P + G + k] = make_double2(cx, cy);
        return;
    }
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
    int u = j0;
    for (; u + UNROLL <= j1; u += UNROLL) {
#pragma unroll
        for (int k = 0; k < UNROLL; ++k)
            pair_rep<REFINE>(cpos[2 * (u + k)], cpos[2 * (u + k) + 1], px, py, ax, ay);
    }
    for (; u < j1; ++u) pair_rep<REFINE>(cpos[2 * u], cpos[2 * u + 1], px, py, ax, ay);
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

// FP32 path: one workgroup per batch vertex sums its G chunk partials and its correction
// pieces (FP64, fixed order: strided per thread, then a fixed tree).
__global__ __launch_bounds__(WG) void k_move32(const double2 *__restrict__ part, int P, int G,
                                               const int *__restrict__ cpre, double2 *__restrict__ pos, int b0,
                                               int bc, const double *__restrict__ stepp, double *__restrict__ eterm) {
    __shared__ double2 red[WG / 64];
    const int slot = (int)blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (slot >= bc) return;
    const int i = b0 + slot;
    const double2 *row = part + (size_t)slot * P;
    // Loads that wait for nothing are issued together: the vertex's position, its number of
    // correction pieces, and the G chunk partials, which every row has whatever that number;
    // the pieces follow in the same per-thread order (p = tid, tid + WG, ...).
    const double2 x0 = pos[i];
    const int pe = G + (cpre[i + 1] - cpre[i]);
    double sx = 0.0, sy = 0.0;
    int p = tid;
    for (; p < G; p += WG) {
        const double2 v = row[p];
        sx += v.x;
        sy += v.y;
    }
    for (; p < pe; p += WG) {
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
        pos[i] = moved(x0, fx, fy, *stepp);
        eterm[i] = fx * fx + fy * fy;
    }
}

// FP64 mode: prevCoordinates[i] += f per j block in block order (:771), then the move.
__global__ __launch_bounds__(WG) void k_move64(const double2 *__restrict__ part, int P, double2 *__restrict__ pos,
                                               int b0, int bc, const double *__restrict__ stepp,
                                               double *__restrict__ eterm) {
    const int slot = (int)blockIdx.x * WG + threadIdx.x;
    if (slot >= bc) return;
    const double2 *row = part + (size_t)slot * P;
    double sx = 0.0, sy = 0.0;
    for (int p = 0; p < P; ++p) {
        sx += row[p].x;
        sy += row[p].y;
    }
    move_one(pos, b0 + slot, sx, sy, *stepp, eterm);
}

// STEP lives on the device so that one iteration's launches are the same every iteration (and
// can be replayed as a captured graph): set once, then multiplied after each iteration with the
// host's FP64 operation.
__global__ void k_step_set(double *stepp, double v) {
    if (threadIdx.x == 0) *stepp = v;
}
__global__ void k_step_decay(double *stepp, double decay) {
    if (threadIdx.x == 0) *stepp = *stepp * decay;  // :781 (0.999), :934 / :1007 (0.99)
}

}  // namespace

// ---- the non-finite guard -------------------------------------------------------------
__global__ void k_zero_int(int *p) {
    if (threadIdx.x == 0) *p = 0;
}

// One vector atomic per wave that holds a non-finite coordinate.
__global__ __launch_bounds__(256) void k_count_nonfinite(const double2 *__restrict__ pos, int n, int *count) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    bool bad = false;
    if (i < n) {
        const double2 p = pos[i];
        bad = !isfinite(p.x) || !isfinite(p.y);
    }
    const uint64_t m = __ballot(bad);
    if (m && (threadIdx.x & 63) == (int)(__builtin_ctzll(m))) atomicAdd(count, __popcll(m));
}

int count_nonfinite(const double2 *pos, int n, int *d_count, hipStream_t s, int *h_count) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, d_count);
    if (n > 0) hipLaunchKernelGGL(k_count_nonfinite, dim3((n + 255) / 256), dim3(256), 0, s, pos, n, d_count);
    BL_HIP(hipGetLastError());
    BL_HIP(hipMemcpyAsync(h_count, d_count, sizeof(int), hipMemcpyDeviceToHost, s));
    BL_HIP(hipStreamSynchronize(s));
    return BL_OK;
}

// ---- energy reductions (shared with the Barnes-Hut path) -------------------------------
// Fixed-order tree sum of |F_i|^2 (FP32 mode).
// Energy sum_i |F_i|^2 in two fixed-order stages: ENERGY_PARTS workgroups each add a
// contiguous share (threads strided, then a wave and workgroup reduction), then one wave adds
// the shares in order.  Deterministic; out[0] = the sum, out[4 ..] = the shares.
constexpr int ENERGY_PARTS = 256;

__global__ __launch_bounds__(256) void k_energy_part(const double *__restrict__ eterm, int n, double *out) {
    __shared__ double red[4];
    const int ch = (n + ENERGY_PARTS - 1) / ENERGY_PARTS;
    const int i0 = blockIdx.x * ch, i1 = min(n, i0 + ch);
    double a = 0.0, b = 0.0;
    int i = i0 + (int)threadIdx.x;
    for (; i + 256 < i1; i += 512) {
        a += eterm[i];
        b += eterm[i + 256];
    }
    if (i < i1) a += eterm[i];
    a = wave_sum(a + b);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) out[4 + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(64) void k_energy_final(double *out) {
    double a = 0.0;
    for (int k = threadIdx.x; k < ENERGY_PARTS; k += 64) a += out[4 + k];
    a = wave_sum(a);
    if (threadIdx.x == 0) out[0] = a;
}

// Sequential sum in vertex order, ENERGY += |F_i|^2 as the reference (FP64 mode).
__global__ void k_energy_seq(const double *__restrict__ eterm, int n, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double e = 0.0;
    for (int i = 0; i < n; ++i) e += eterm[i];
    *out = e;
}

// out: 4 + ENERGY_PARTS doubles (the sum in out[0])
int launch_energy(const double *eterm, int n, double *out, int precision, hipStream_t s) {
    if (precision == BL_PREC_FP64)
        hipLaunchKernelGGL(k_energy_seq, dim3(1), dim3(64), 0, s, eterm, n, out);
    else {
        hipLaunchKernelGGL(k_energy_part, dim3(ENERGY_PARTS), dim3(256), 0, s, eterm, n, out);
        hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(64), 0, s, out);
    }
    BL_HIP(hipGetLastError());
    return BL_OK;
}

// ---- the static workspace (bl_run_device_static) ------------------------------------------
// One array per GPU, held until the process ends.  On the host its name stands for a host-side
// placeholder, not device memory, so host code uses only the device address that
// hipGetSymbolAddress returns, looked up once per device with that device current.
static __device__ __align__(256) unsigned char scratch[BL_STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;

static int static_scratch(int device, unsigned char **out) {
    static std::mutex mu;
    static void *addr[MAX_DEVICES] = {};
    if (device < 0 || device >= MAX_DEVICES) return BL_E_DEVICE;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = -1;
        BL_HIP(hipGetDevice(&cur));
        if (cur != device) BL_HIP(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return device_error(e, "hipGetSymbolAddress(scratch)", __LINE__);
        if (!p) return BL_E_DEVICE;
        addr[device] = p;
    }
    *out = static_cast<unsigned char *>(addr[device]);
    return BL_OK;
}

static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

int run_minibatch(bl_graph *g, double *d_xy, const bl_params *p, int law, double decay, double step0, int iters,
                  bool converged, bl_stats *st, bool static_ws) {
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
    int JS = 1, G = 1, chunk = n, P = 1;
    size_t part_elems = 0;
    if (fp64) {
        G = (n + kBlockY - 1) / kBlockY;
        P = G;
        part_elems = (size_t)bmax * P;
    } else {
        // 64 batch slots x 4 j-subgroups per workgroup: at least ~1024 workgroups (4 per CU),
        // and j chunks of at most ~CHUNK_MAX positions, whose scalar loads the rows of a chunk
        // then mostly share in the scalar cache
        constexpr int CHUNK_MAX = 1024;
        JS = 4;
        const int ipw = WG / JS;
        const int rows = (bmax + ipw - 1) / ipw;
        const int target = std::max(std::max(1, 1024 / rows), (n + CHUNK_MAX - 1) / CHUNK_MAX);
        const int quant = UNROLL * JS;
        chunk = (n + target - 1) / target;
        chunk = ((chunk + quant - 1) / quant) * quant;
        G = (n + chunk - 1) / chunk;
        P = G + g->corr_max_pieces;
        part_elems = (size_t)bmax * P;
    }
    double2 *part;
    double *eterm, *energy;
    if (static_ws) {
        // [energy shares | partial forces | energy terms]; every word a step reads is written
        // earlier in the same step, so the contents of the previous call do not matter
        static_assert(sizeof(double) * (4 + ENERGY_PARTS) <= 4096, "energy shares exceed their part");
        const size_t off_part = 4096;
        const size_t off_eterm = off_part + round256(sizeof(double2) * part_elems);
        const size_t need = off_eterm + round256(sizeof(double) * (size_t)n);
        if (need > BL_STATIC_SCRATCH_BYTES) return BL_E_SCRATCH;
        unsigned char *ws = nullptr;
        const int rc = static_scratch(g->device, &ws);
        if (rc) return rc;
        energy = reinterpret_cast<double *>(ws);
        part = reinterpret_cast<double2 *>(ws + off_part);
        eterm = reinterpret_cast<double *>(ws + off_eterm);
    } else {
        if (g->part.ensure(sizeof(double2) * part_elems)) return BL_E_NOMEM;
        if (g->eterm.ensure(sizeof(double) * (size_t)n)) return BL_E_NOMEM;
        if (g->energy.ensure(sizeof(double) * (4 + ENERGY_PARTS))) return BL_E_NOMEM;
        part = g->part.as<double2>();
        eterm = g->eterm.as<double>();
        energy = g->energy.as<double>();
    }

    Prof prof;
    prof.on = p->profile != 0;
    prof.s = s;
    double *d_step = energy + 1;  // energy[0] is the sum, energy[4 ..] its shares
    hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, s, d_step, step0);
    // One iteration's launches: every batch step (pairs, then the move), the energy, STEP *= decay.
    auto emit_iteration = [&]() -> int {
        for (int b = 0; b < nb; ++b) {
            const int b0 = b * B;
            const int bc = (n - b0) < B ? (n - b0) : B;
            if (fp64) {
                prof.begin(0);
                hipLaunchKernelGGL(k_pairs64, dim3(G, (bc + WG - 1) / WG), dim3(WG), 0, s, pos, n, b0, bc, part, P,
                                   erp, eci, ew, p->weight, law);
                prof.end();
                prof.begin(1);
                hipLaunchKernelGGL(k_move64, dim3((bc + WG - 1) / WG), dim3(WG), 0, s, part, P, pos, b0, bc, d_step,
                                   eterm);
                prof.end();
            } else {
                const int ipw = WG / JS;
                const int pc0 = g->h_cpre[b0], npc = g->h_cpre[b0 + bc] - pc0;
                const dim3 grid(G + (npc + 3) / 4, (bc + ipw - 1) / ipw);
                const int *cpre = g->cpre.as<int>(), *cpv = g->cpv.as<int>();
                if (prof.on) {  // the dispatch records its own start and end
                    hipEvent_t e0, e1;
                    prof.ext(0, &e0, &e1);
                    if (p->precision == BL_PREC_FP32)
                        hipExtLaunchKernelGGL((k_pairs32<4, false>), grid, dim3(WG), 0, s, e0, e1, 0, pos, n, b0, bc,
                                              chunk, G, part, P, erp, eci, ew, p->weight, law, cpre, cpv, pc0, npc);
                    else
                        hipExtLaunchKernelGGL((k_pairs32<4, true>), grid, dim3(WG), 0, s, e0, e1, 0, pos, n, b0, bc,
                                              chunk, G, part, P, erp, eci, ew, p->weight, law, cpre, cpv, pc0, npc);
                    prof.ext(1, &e0, &e1);
                    hipExtLaunchKernelGGL(k_move32, dim3(bc), dim3(WG), 0, s, e0, e1, 0, part, P, G, cpre, pos, b0, bc,
                                          d_step, eterm);
                } else {
                    if (p->precision == BL_PREC_FP32)
                        hipLaunchKernelGGL((k_pairs32<4, false>), grid, dim3(WG), 0, s, pos, n, b0, bc, chunk, G, part,
                                           P, erp, eci, ew, p->weight, law, cpre, cpv, pc0, npc);
                    else
                        hipLaunchKernelGGL((k_pairs32<4, true>), grid, dim3(WG), 0, s, pos, n, b0, bc, chunk, G, part,
                                           P, erp, eci, ew, p->weight, law, cpre, cpv, pc0, npc);
                    hipLaunchKernelGGL(k_move32, dim3(bc), dim3(WG), 0, s, part, P, G, cpre, pos, b0, bc, d_step,
                                       eterm);
                }
            }
        }
        prof.begin(2);
        const int rc = launch_energy(eterm, n, energy, p->precision, s);
        prof.end();
        hipLaunchKernelGGL(k_step_decay, dim3(1), dim3(64), 0, s, d_step, decay);
        return rc;
    };
    // BL_GRAPH=1 (and no per-launch timing): one iteration is captured once as a graph and
    // replayed, one host launch per iteration instead of two per batch step.  Not the default:
    // measured on MI355X (ROCm 7) the replay was no faster (64x64 grid 96 vs 95 us per iteration,
    // 32x32 grid 32 vs 28 us, 250k vertices 28.6 vs 28.0 ms): the GPU's own dispatch of
    // dependent small kernels, not the host's launch calls, sets the pace.  The static-workspace
    // path creates no graph (instantiation allocates).
    struct Exec {
        hipGraphExec_t e = nullptr;
        ~Exec() {
            if (e) (void)hipGraphExecDestroy(e);
        }
    } ex;
    const char *genv = getenv("BL_GRAPH");
    const bool want_graph = !prof.on && !static_ws && (converged || iters >= 2) && genv && !strcmp(genv, "1");
    if (want_graph && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        const int rc = emit_iteration();
        hipGraph_t gr = nullptr;
        const hipError_t ee = hipStreamEndCapture(s, &gr);
        if (rc == BL_OK && ee == hipSuccess && gr) {
            if (hipGraphInstantiate(&ex.e, gr, nullptr, nullptr, 0) != hipSuccess) ex.e = nullptr;
        }
        if (gr) (void)hipGraphDestroy(gr);
        (void)hipGetLastError();  // a failed capture leaves the launches to the loop below
    }
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
        if (ex.e) {
            BL_HIP(hipGraphLaunch(ex.e, s));
        } else {
            const int rc = emit_iteration();
            if (rc) return rc;
            prof.flush();
        }
        STEP = STEP * decay;  // the host's copy (the device multiplies its own the same way)
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
        if (static_ws) {  // the guard's counter in the energy block's spare word
            int nf = 0;
            const int rc = count_nonfinite(pos, n, reinterpret_cast<int *>(energy + 3), s, &nf);
            if (rc) return rc;
            st->nonfinite = nf;
        }
    }
    return BL_OK;
}

}  // namespace bl
