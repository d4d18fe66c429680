// Many small graphs in one persistent launch (bl_run_batched).
//
// Each graph follows the single-graph minibatch algorithm (algorithms.cpp:715-794 for algo 2,
// :947-1018 for algo 8): ceil(n/B) batches per iteration, Jacobi inside a batch, Gauss-Seidel
// across batches, x_i += STEP*F/|F|, STEP *= decay, with the pair arithmetic and neighbour
// correction of pair_math.h.  Workgroups of 1024 threads pull graphs from an atomic work
// queue (largest n^2 first) and run all of a graph's iterations alone: the graph's FP64
// positions sit in LDS when they fit (read by every lane with the same address, i.e.
// broadcast) and are read from global memory otherwise.  A graph never shares data with
// another workgroup, so the only cross-workgroup traffic is the queue counter.
//
// Inside a batch step IPW batch slots (a power of two, 64..1024, per graph) are served at once
// by JS = 1024 / IPW j-subgroups, in rounds of IPW slots.  The j range is cut into NCH = 16
// fixed chunks whose partials are combined by one fixed pairwise tree, ((c0 + c1) + (c2 + c3))
// + ..., whatever the shape: a subgroup sums its 16 / JS consecutive chunks as the subtree they
// form, and the subgroups' subtrees are combined as the rest of the tree.  So every shape gives
// the same bits, and the shape is free to be tuned (run_batched measures the candidates per
// size class of the graphs); the neighbour correction follows.  Results do not depend on
// timing either.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <vector>

#include "bl_internal.h"
#include "pair_math.h"

namespace bl {
namespace {

constexpr int BT = 1024;       // threads per workgroup
constexpr int MAXB = 1024;     // largest batch size of the batched path
constexpr int NCH = 16;        // fixed j chunks of the dense sum (16 = BT / the smallest IPW)

struct BatchedArgs {
    const int *order;          // graph ids, largest first
    const int *ipw_of;         // batch slots served at once, per graph (power of two, 64..1024)
    int count;
    const int *voff;           // vertex offsets [count+1]
    const int *erp;            // effective-neighbour rowptr over all vertices (global entry ids)
    const int *eci;            // effective neighbours (graph-local vertex ids)
    double2 *pos;              // all positions
    int B, iters, law, refine, lds_cap;
    double step0, decay, W;
    double *energies;
    long long *finish;         // wall-clock tick at each graph's completion
    long long *wg_start;       // wall-clock tick at each workgroup's start
    int *queue;
};

typedef __attribute__((address_space(3))) const double lds_double;

// Dense repulsion of position pi against positions [j0, j1) of `src` (x, y interleaved, in LDS
// or global memory; with an LDS-typed pointer the loads are LDS broadcasts instead of generic
// loads).
template <bool REFINE, class Src>
__device__ __forceinline__ void dense_range(Src src, int j0, int j1, double2 pi, double &ax, double &ay) {
    int u = j0;
    for (; u + 8 <= j1; u += 8) {  // eight positions in flight per load batch
        double qx[8], qy[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            qx[k] = src[2 * (u + k)];
            qy[k] = src[2 * (u + k) + 1];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) pair_rep<REFINE>(qx[k], qy[k], pi.x, pi.y, ax, ay);
    }
    for (; u < j1; ++u) pair_rep<REFINE>(src[2 * u], src[2 * u + 1], pi.x, pi.y, ax, ay);
}

// Pairwise combination in a fixed tree: leaves arrive left to right, level l holds the pending
// left subtree of 2^l leaves; after 2^L leaves (L <= 4) level L holds their sum,
// ((v0 + v1) + (v2 + v3)) + ...  Named registers and uniform branches (the leaf count k is the
// same in every lane), no indexed array that would go to scratch memory.
struct PairTree {
    double x0, y0, x1, y1, x2, y2, x3, y3, x4, y4;
    int k = 0;
    __device__ __forceinline__ void add(double vx, double vy) {
        if (k & 1) {  // close the subtrees this leaf completes, left operand first
            vx = x0 + vx, vy = y0 + vy;
            if (k & 2) {
                vx = x1 + vx, vy = y1 + vy;
                if (k & 4) {
                    vx = x2 + vx, vy = y2 + vy;
                    if (k & 8) vx = x3 + vx, vy = y3 + vy;
                }
            }
        }
        if (!(k & 1)) x0 = vx, y0 = vy;
        else if (!(k & 2)) x1 = vx, y1 = vy;
        else if (!(k & 4)) x2 = vx, y2 = vy;
        else if (!(k & 8)) x3 = vx, y3 = vy;
        else x4 = vx, y4 = vy;
        ++k;
    }
    __device__ __forceinline__ double2 total(int L) const {
        return L == 0 ? make_double2(x0, y0) : L == 1 ? make_double2(x1, y1) : L == 2 ? make_double2(x2, y2)
               : L == 3 ? make_double2(x3, y3) : make_double2(x4, y4);
    }
};

template <bool REFINE>
__device__ void layout_graph(const BatchedArgs &a, int g, double2 *P, double2 *red, double2 *fb, double *eb) {
    const int tid = threadIdx.x;
    const int v0 = __builtin_amdgcn_readfirstlane(a.voff[g]);
    const int n = __builtin_amdgcn_readfirstlane(a.voff[g + 1]) - v0;
    const bool in_lds = n <= a.lds_cap;
    double2 *gp = in_lds ? P : a.pos + v0;
    if (in_lds) {
        for (int k = tid; k < n; k += BT) P[k] = a.pos[v0 + k];
        __syncthreads();
    }
    const int B = a.B;
    const int ipw = __builtin_amdgcn_readfirstlane(a.ipw_of[g]);
    const int JS = BT / ipw, cpt = NCH / JS;  // subgroups, chunks per subgroup
    const int lcpt = __builtin_ctz(cpt), ljs = __builtin_ctz(JS);
    const int li = tid % ipw, js = tid / ipw;
    const int sub = (n + NCH - 1) / NCH;      // chunk c: [c * sub, (c + 1) * sub) within [0, n)
    const int nb = (n + B - 1) / B;
    double step = a.step0, E = 0.0;
    for (int it = 0; it < a.iters; ++it) {
        E = 0.0;
        for (int b = 0; b < nb; ++b) {
            const int b0 = b * B, bc = min(B, n - b0);
            for (int r0 = 0; r0 < bc; r0 += ipw) {  // rounds of ipw slots
                const int slot = r0 + li;
                const bool act = slot < bc;
                const double2 pi = gp[b0 + (act ? slot : 0)];
                PairTree t;
                for (int c = js * cpt; c < (js + 1) * cpt; ++c) {  // this subgroup's subtree
                    const int j0 = min(n, c * sub), j1 = min(n, j0 + sub);
                    double cx = 0.0, cy = 0.0;
                    if (in_lds)
                        dense_range<REFINE>((lds_double *)P, j0, j1, pi, cx, cy);
                    else
                        dense_range<REFINE>(reinterpret_cast<const double *>(gp), j0, j1, pi, cx, cy);
                    t.add(cx, cy);
                }
                red[tid] = t.total(lcpt);
                __syncthreads();
                if (js == 0 && act) {
                    PairTree u;  // the subgroups' subtrees, in order: the top of the tree
                    for (int s = 0; s < JS; ++s) u.add(red[s * ipw + li].x, red[s * ipw + li].y);
                    const double2 tot = u.total(ljs);
                    const double ax = tot.x, ay = tot.y;
                    const int i = v0 + b0 + slot;
                    double cx = 0.0, cy = 0.0;
                    for (int e = a.erp[i]; e < a.erp[i + 1]; ++e) {
                        const double2 pj = gp[a.eci[e]];
                        correct_pair<REFINE>(a.law, a.W, pj.x - pi.x, pj.y - pi.y, cx, cy);
                    }
                    fb[slot] = make_double2(ax + cx, ay + cy);
                }
                __syncthreads();
            }
            // move the batch (every force of the batch is formed: Jacobi inside the batch)
            double e2 = 0.0;
            for (int slot = tid; slot < bc; slot += BT) {
                const double2 f = fb[slot];
                gp[b0 + slot] = moved(gp[b0 + slot], f.x, f.y, step);
                e2 += f.x * f.x + f.y * f.y;
            }
            // energy of the batch: wave sums, then the 16 wave sums in order (one barrier, which
            // also makes the moves visible to the next batch)
            e2 = wave_sum(e2);
            if ((tid & 63) == 0) eb[tid >> 6] = e2;
            __syncthreads();
            double eb_sum = 0.0;
#pragma unroll
            for (int q = 0; q < BT / 64; ++q) eb_sum += eb[q];
            E += eb_sum;  // batches added in order
        }
        step = step * a.decay;
    }
    if (in_lds) {
        for (int k = tid; k < n; k += BT) a.pos[v0 + k] = P[k];
    }
    // every thread holds the same E (the batch sums come from the same wave sums); all of them
    // store it so that no per-lane branch sits next to the queue loop's back edge
    a.energies[g] = E;
    a.finish[g] = wall_clock64();
    __syncthreads();
}

__global__ __launch_bounds__(BT) void k_batched(BatchedArgs a) {
    extern __shared__ double2 P[];  // positions of the current graph when it fits
    __shared__ double2 red[BT];
    __shared__ double2 fb[MAXB];
    __shared__ double eb[BT / 64];
    __shared__ int slot;
    const int tid = threadIdx.x;
    if (tid == 0) a.wg_start[blockIdx.x] = wall_clock64();
    for (;;) {
        if (tid == 0) slot = atomicAdd(a.queue, 1);
        __syncthreads();
        // The slot is the same for every thread; readfirstlane makes that visible to the
        // compiler.  Left as a per-lane LDS value, the loop exit looks divergent and the
        // loop may be restructured so that lane 0 (which advances the queue) leaves the
        // per-graph loop while the other lanes of its wave repeat the same graph forever.
        const int q = __builtin_amdgcn_readfirstlane(slot);
        __syncthreads();
        if (q >= a.count) break;
        const int g = __builtin_amdgcn_readfirstlane(a.order[q]);
        if (a.refine)
            layout_graph<true>(a, g, P, red, fb, eb);
        else
            layout_graph<false>(a, g, P, red, fb, eb);
    }
}

// The fixed rule: every batch slot served in one round (the next power of two >= min(B, n)).
int ipw_rule(int n, int B) {
    const int bmax = std::min(B, n);
    int ipw = 64;
    while (ipw < bmax && ipw < BT) ipw <<= 1;
    return ipw;
}

// Tuned shapes per (size class, B, precision, law, positions in LDS), for the process.
struct TuneKey {
    int cls, B, refine, law, lds;
    bool operator<(const TuneKey &o) const {
        return std::tie(cls, B, refine, law, lds) < std::tie(o.cls, o.B, o.refine, o.law, o.lds);
    }
};
std::mutex tune_mu;
std::map<TuneKey, int> tuned;

}  // namespace

int run_batched(int32_t count, const int32_t *voff, const int64_t *eoff, const int32_t *rowptr, const int32_t *colids,
                double *xy, const bl_params *p, double *energies, double *finish_ms, bl_stats *st) {
    if (p->batch > MAXB) return BL_E_ARG;
    const int N = voff[count];
    // effective neighbours per graph, each graph's own entry range being its array end
    std::vector<int32_t> erp(N + 1, 0), eci;
    std::vector<float> unused;
    eci.reserve((size_t)eoff[count]);
    for (int32_t k = 0; k < count; ++k)
        effective_neighbors(rowptr, colids, nullptr, voff[k], voff[k + 1], eoff[k + 1], erp, eci, unused, nullptr);
    // largest graphs first
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const long long nx = voff[x + 1] - voff[x], ny = voff[y + 1] - voff[y];
        return nx * nx > ny * ny;
    });

    int dev = 0, cus = 0, khz = 0, maxlds = 0;
    BL_HIP(hipGetDevice(&dev));
    BL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BL_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
    const int static_lds = (int)(sizeof(double2) * (BT + MAXB) + sizeof(double) * (BT / 64) + 64);
    int maxn = 0;
    for (int k = 0; k < count; ++k) maxn = std::max(maxn, voff[k + 1] - voff[k]);
    const int cap = std::max(0, std::min(maxn, (maxlds - static_lds) / (int)sizeof(double2)));
    const size_t dyn = sizeof(double2) * (size_t)cap;
    if (dyn > 0) BL_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_batched),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    int per_cu = 0;
    BL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batched, BT, dyn));
    const int grid = std::max(1, std::min(count, std::max(1, per_cu) * cus));

    hipStream_t s = nullptr;
    BL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct Buffers {
        void *p = nullptr;
        hipStream_t s = nullptr;
        ~Buffers() {
            if (s) (void)hipStreamSynchronize(s);
            if (p) (void)hipFree(p);
            if (s) (void)hipStreamDestroy(s);
        }
    } buf;
    buf.s = s;
    const size_t b_order = 4 * (size_t)count, b_voff = 4 * (size_t)(count + 1), b_erp = 4 * (size_t)(N + 1),
                 b_eci = 4 * std::max<size_t>(1, eci.size()), b_pos = 16 * (size_t)N, b_en = 8 * (size_t)count,
                 b_fin = 8 * (size_t)count, b_wg = 8 * (size_t)grid, b_q = 64, b_ipw = 4 * (size_t)count;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t total = al(b_order) + al(b_voff) + al(b_erp) + al(b_eci) + al(b_pos) + al(b_en) + al(b_fin) +
                         al(b_wg) + al(b_q) + 2 * al(b_ipw);
    BL_HIP(hipMalloc(&buf.p, total));
    char *c = static_cast<char *>(buf.p);
    int *d_order = reinterpret_cast<int *>(c); c += al(b_order);
    int *d_voff = reinterpret_cast<int *>(c); c += al(b_voff);
    int *d_erp = reinterpret_cast<int *>(c); c += al(b_erp);
    int *d_eci = reinterpret_cast<int *>(c); c += al(b_eci);
    double2 *d_pos = reinterpret_cast<double2 *>(c); c += al(b_pos);
    double *d_en = reinterpret_cast<double *>(c); c += al(b_en);
    long long *d_fin = reinterpret_cast<long long *>(c); c += al(b_fin);
    long long *d_wg = reinterpret_cast<long long *>(c); c += al(b_wg);
    int *d_q = reinterpret_cast<int *>(c); c += al(b_q);
    int *d_ipw = reinterpret_cast<int *>(c); c += al(b_ipw);
    int *d_torder = reinterpret_cast<int *>(c);  // a size class's sample graphs (tuning)
    BL_HIP(hipMemcpyAsync(d_order, order.data(), b_order, hipMemcpyHostToDevice, s));
    BL_HIP(hipMemcpyAsync(d_voff, voff, b_voff, hipMemcpyHostToDevice, s));
    BL_HIP(hipMemcpyAsync(d_erp, erp.data(), b_erp, hipMemcpyHostToDevice, s));
    if (!eci.empty()) BL_HIP(hipMemcpyAsync(d_eci, eci.data(), 4 * eci.size(), hipMemcpyHostToDevice, s));
    BL_HIP(hipMemcpyAsync(d_pos, xy, b_pos, hipMemcpyHostToDevice, s));
    BL_HIP(hipMemsetAsync(d_q, 0, b_q, s));

    BatchedArgs a;
    a.order = d_order;
    a.ipw_of = d_ipw;
    a.count = count;
    a.voff = d_voff;
    a.erp = d_erp;
    a.eci = d_eci;
    a.pos = d_pos;
    a.B = p->batch;
    a.iters = p->iters;
    a.law = p->algo == BL_ALGO_FA ? LAW_FA : LAW_FR;
    a.refine = p->precision == BL_PREC_MIXED ? 1 : 0;
    a.lds_cap = cap;
    a.step0 = p->step0;
    a.decay = p->algo == BL_ALGO_FA ? 0.99 : 0.999;
    a.W = p->weight;
    a.energies = d_en;
    a.finish = d_fin;
    a.wg_start = d_wg;
    a.queue = d_q;
    hipEvent_t e0, e1;
    BL_HIP(hipEventCreate(&e0));
    BL_HIP(hipEventCreate(&e1));
    struct Events {
        hipEvent_t *a, *b;
        ~Events() {
            (void)hipEventDestroy(*a);
            (void)hipEventDestroy(*b);
        }
    } evs{&e0, &e1};

    // The shape of each graph: the fixed rule, or, for runs with enough work, the fastest of the
    // candidate shapes per size class (graphs of 2^c .. 2^(c+1) - 1 vertices), measured once per
    // process on up to TUNE_SAMPLE graphs of the class, one iteration each, on a copy of the
    // positions.  All shapes give the same bits.
    std::vector<int> ipw(count);
    for (int k = 0; k < count; ++k) ipw[k] = ipw_rule(voff[k + 1] - voff[k], p->batch);
    double work = 0.0;
    for (int k = 0; k < count; ++k) work += (double)(voff[k + 1] - voff[k]) * (voff[k + 1] - voff[k]);
    work *= std::max(1, p->iters);
    const char *tenv = getenv("BL_TUNE");
    const bool tune = !(tenv && !strcmp(tenv, "0")) && (work >= 2e9 || (tenv && !strcmp(tenv, "1")));
    if (tune) {
        constexpr int TUNE_SAMPLE = 24;
        std::map<int, std::vector<int>> members;  // class -> graph ids in queue order
        for (int q = 0; q < count; ++q) {
            const int k = order[q], nk = voff[k + 1] - voff[k];
            if (nk >= 2) members[31 - __builtin_clz((unsigned)nk)].push_back(k);
        }
        void *tpos = nullptr;
        struct Free {
            void **p;
            ~Free() {
                if (*p) (void)hipFree(*p);
            }
        } free_tpos{&tpos};
        for (auto &cm : members) {
            const std::vector<int> &ids = cm.second;
            int nmax = 0;
            for (int k : ids) nmax = std::max(nmax, voff[k + 1] - voff[k]);
            const TuneKey key{cm.first, p->batch, a.refine, a.law, nmax <= cap ? 1 : 0};
            int best = -1;
            {
                std::lock_guard<std::mutex> lock(tune_mu);
                auto it = tuned.find(key);
                if (it != tuned.end()) best = it->second;
            }
            const int top = ipw_rule(nmax, p->batch);
            if (best < 0 && top > 64) {
                if (!tpos) {
                    BL_HIP(hipMalloc(&tpos, b_pos));
                    BL_HIP(hipMemcpyAsync(tpos, d_pos, b_pos, hipMemcpyDeviceToDevice, s));
                }
                const int m = std::min((int)ids.size(), TUNE_SAMPLE);
                BL_HIP(hipMemcpyAsync(d_torder, ids.data(), 4 * (size_t)m, hipMemcpyHostToDevice, s));
                BatchedArgs t = a;
                t.order = d_torder;
                t.count = m;
                t.iters = 1;
                t.pos = static_cast<double2 *>(tpos);
                float best_ms = 0.f;
                for (int cand = 64; cand <= top; cand <<= 1) {
                    std::vector<int> all(count, cand);
                    BL_HIP(hipMemcpyAsync(d_ipw, all.data(), b_ipw, hipMemcpyHostToDevice, s));
                    float ms = 0.f;
                    for (int rep = 0; rep < 2; ++rep) {  // the second launch is timed
                        BL_HIP(hipMemsetAsync(d_q, 0, b_q, s));
                        BL_HIP(hipEventRecord(e0, s));
                        hipLaunchKernelGGL(k_batched, dim3(std::min(m, grid)), dim3(BT), dyn, s, t);
                        BL_HIP(hipGetLastError());
                        BL_HIP(hipEventRecord(e1, s));
                        BL_HIP(hipEventSynchronize(e1));
                        BL_HIP(hipEventElapsedTime(&ms, e0, e1));
                    }
                    if (best < 0 || ms < best_ms) best = cand, best_ms = ms;
                }
                std::lock_guard<std::mutex> lock(tune_mu);
                tuned[key] = best;
            }
            if (best > 0)
                for (int k : ids) ipw[k] = std::min(best, ipw_rule(voff[k + 1] - voff[k], p->batch));
        }
    }
    if (const char *force = getenv("BL_BATCHED_IPW")) {  // one shape for every graph (tests)
        const int f = atoi(force);
        if (f >= 64 && f <= BT && (f & (f - 1)) == 0)
            for (int k = 0; k < count; ++k) ipw[k] = std::min(f, ipw_rule(voff[k + 1] - voff[k], p->batch));
    }
    BL_HIP(hipMemcpyAsync(d_ipw, ipw.data(), b_ipw, hipMemcpyHostToDevice, s));
    BL_HIP(hipMemsetAsync(d_q, 0, b_q, s));
    BL_HIP(hipEventRecord(e0, s));
    hipLaunchKernelGGL(k_batched, dim3(grid), dim3(BT), dyn, s, a);
    BL_HIP(hipGetLastError());
    BL_HIP(hipEventRecord(e1, s));
    std::vector<double> en(count);
    std::vector<long long> fin(count), wg(grid);
    int nonfinite = 0;  // the guard's counter in the queue block's spare words
    int rc = count_nonfinite(d_pos, N, d_q + 8, s, &nonfinite);
    if (rc) return rc;
    BL_HIP(hipMemcpyAsync(xy, d_pos, b_pos, hipMemcpyDeviceToHost, s));
    BL_HIP(hipMemcpyAsync(en.data(), d_en, b_en, hipMemcpyDeviceToHost, s));
    BL_HIP(hipMemcpyAsync(fin.data(), d_fin, b_fin, hipMemcpyDeviceToHost, s));
    BL_HIP(hipMemcpyAsync(wg.data(), d_wg, b_wg, hipMemcpyDeviceToHost, s));
    BL_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (energies) std::copy(en.begin(), en.end(), energies);
    if (finish_ms) {
        const long long t0 = *std::min_element(wg.begin(), wg.end());
        for (int k = 0; k < count; ++k) finish_ms[k] = (double)(fin[k] - t0) / (double)khz;
    }
    if (st) {
        memset(st, 0, sizeof(*st));
        st->energy = std::accumulate(en.begin(), en.end(), 0.0);
        st->iters_done = p->iters;
        st->seconds = ms / 1e3;
        double sstep = p->step0;
        for (int it = 0; it < p->iters; ++it) sstep *= a.decay;
        st->step_final = sstep;
        st->kernel_ms[0] = ms;
        st->kernel_launches[0] = 1;
        st->nonfinite = nonfinite;
    }
    return BL_OK;
}

}  // namespace bl
