// The lagged batch walk of the Barnes-Hut family (algorithms.cpp:1050-1066: per batch, the
// attraction of every CSR entry with live positions minus the repulsion, then the sequential
// move), for batches of at most LAG_T vertices whose rows are short.
//
// The reference spreads each batch over its threads; a batch's attraction reads the positions its
// neighbours have at that moment, so batch b depends on every earlier batch that holds one of its
// neighbours.  On graphs with unordered ids those neighbours are spread over all earlier batches,
// and only a few lie in the last D-1 batches.  The walk therefore splits each vertex's "early"
// neighbours (those in earlier batches, which have moved) into
//   settled  in batches <= b - D: final once batch b - D has moved, and
//   recent   in batches b - D + 1 .. b - 1,
// (the "late" ones, in batch b or later, still at their iteration-start positions, are summed
// beforehand by k_late_attract into the repulsion rep').
//
//   workgroup 0, the walker: the one-workgroup row walk (LAG_T computing threads, one per batch
//     vertex, and LAG_T loading threads that stage the next batch in LDS), with
//        f_i = pre_i + sum_{recent} attraction,   pre_i = sum_{settled} attraction - rep'_i;
//   workgroups 1..H, the helpers: helper h takes batches h-1, h-1+H, ...; for batch b it waits
//     until the walker has published batches [0, b - D + 1) (batches it needs are final in
//     memory), computes pre_i of the batch's vertices into memory and flags the batch with the
//     launch's generation number.
// The walker never waits for a helper: three batches ahead it reads the batch's flag; a flagged
// batch is staged with its pre_i and recent entries only, an unflagged one with all its early
// entries, and the walker forms pre_i itself with the same FP64 operations in the same (CSR)
// order, so the layout is the same whichever workgroup formed it.  Helpers stop at a deadline or
// when the walker has finished, so the launch ends whatever the scheduling.
//
// Publishing: after the stores of batch c have completed, one loading thread of the walker
// releases done = c + 1 every LAG_RK batches (an agent-scope release writes the XCD's L2 back);
// a helper acquires it before gathering.  A helper's flag is released after its pre_i stores;
// the walker's loading waves acquire before loading a flagged batch's pre_i.
#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "bh_common.h"

namespace bl {
namespace {

constexpr int LAG_T = 256;                 // batch vertices; walker computing / loading threads
constexpr int LAG_CAP = 2560;              // staged CSR entries per batch
constexpr int LAG_S = LAG_CAP / LAG_T;     // staged entries per loading thread
constexpr int LAG_MAXDEG = 64;             // longest row for which this walk is chosen

struct LagArgs {
    const int *erp, *esp, *eci;  // early entries of vertex i: [erp[i], esp[i]) settled, [esp[i], erp[i+1]) recent
    const float *ew;
    const int *rrp, *rci;        // recent entries only
    const float *rw;
    double W;
    double2 *pos;
    const double2 *rep;          // repulsion minus the late attraction (k_late_attract)
    double2 *pre;                // pre_i written by the helpers
    int *ready;                  // ready[b] == gen: pre of batch b is in memory
    int *done;                   // batches [0, *done) of this launch are final in memory
    int gen;
    int n, B, nb, D, RK;
    double step;
    double *eterm;
    const int *status;
    long long ticks;             // a helper's longest wait
};

constexpr size_t lag_lds(bool has_w) {
    return sizeof(double2) * (3 * LAG_T + 2 * LAG_CAP + 4 * LAG_T) + sizeof(double) * 2 * LAG_T +
           sizeof(int) * (2 * LAG_CAP + 6 * LAG_T + 4) + (has_w ? sizeof(float) * 2 * LAG_CAP : 0);
}

// Publishes *flag = v to other workgroups after every store of this thread (and, through a
// preceding workgroup barrier, of its workgroup): a release fence, whose L2 write-back is waited
// for explicitly, then a relaxed store.  (The fused release store was compiled without that wait
// in the walker's loop, so the flag could overtake the write-back.)
__device__ __forceinline__ void publish(int *flag, int v) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __hip_atomic_store(flag, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One attraction term of vertex v toward pj, d * (Wk |d|) (algorithms.cpp:1055-1058), added to
// (ax, ay).  The walker and the helpers both use it, so their pre_i agree bit for bit.
__device__ __forceinline__ void attr_add(double2 pj, double2 v, double Wk, double &ax, double &ay) {
    const double dx = pj.x - v.x, dy = pj.y - v.y;
    const double sc = Wk * sqrt(dx * dx + dy * dy);
    ax = fma(dx, sc, ax);
    ay = fma(dy, sc, ay);
}

template <bool HAS_W>
__device__ void lag_helper(const LagArgs &a, int hid, int nhelp) {
    __shared__ int ok;
    const int t = threadIdx.x;
    for (int b = hid; b < a.nb; b += nhelp) {
        __syncthreads();  // ok of the previous batch has been read
        const int need = b - a.D + 1;  // batches [0, need) must be final
        if (need > 0) {
            if (t == 0) {
                const long long t0 = wall_clock64();
                int good = -1;  // the published count, or -1: stop
                for (;;) {
                    const int d = __hip_atomic_load(a.done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (d >= need) {
                        good = d;
                        break;
                    }
                    if (__hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                        wall_clock64() - t0 > a.ticks)
                        break;
                    __builtin_amdgcn_s_sleep(1);
                }
                ok = good;
            }
            __syncthreads();
            if (ok < 0) return;  // the walker forms the remaining batches itself
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            // the walker reads a batch's flag three batches ahead: one it has (nearly) reached is
            // formed by the walker
            if (ok > b - 3) continue;
        }
        const int b0 = b * a.B, bc = min(a.B, a.n - b0);
        if (t < bc) {
            const int i = b0 + t;
            const double2 v = a.pos[i];
            const int e0 = a.erp[i], e1 = a.esp[i];
            double ax = 0.0, ay = 0.0;
            for (int e = e0; e < e1; e += 4) {
                int j[4];
                double wk[4];
                double2 pj[4];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int k = min(e + g, e1 - 1);
                    j[g] = a.eci[k];
                    wk[g] = HAS_W ? a.W * (double)a.ew[k] : a.W;
                }
#pragma unroll
                for (int g = 0; g < 4; ++g) pj[g] = a.pos[j[g]];
#pragma unroll
                for (int g = 0; g < 4; ++g)
                    if (e + g < e1) attr_add(pj[g], v, wk[g], ax, ay);
            }
            const double2 r = a.rep[i];
            a.pre[i] = make_double2(ax - r.x, ay - r.y);
        }
        // every wave's stores have reached the L2 before thread 0's release writes it back
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (t == 0) publish(a.ready + b, a.gen);
    }
}

template <bool HAS_W>
__device__ void lag_walker(const LagArgs &a, unsigned char *raw) {
    constexpr int T = LAG_T, S = LAG_S, C = LAG_CAP;
    double2 *ring = reinterpret_cast<double2 *>(raw);       // [3][T] moved positions of batch c (slot c % 3)
    double2 *spos = ring + 3 * T;                           // [2][C] staged neighbour positions
    double2 *sv = spos + 2 * C;                             // [2][T] staged own positions
    double2 *sr = sv + 2 * T;                               // [2][T] staged pre_i (flagged) or rep'_i
    double *ering = reinterpret_cast<double *>(sr + 2 * T); // [2][T] energy terms
    int *scid = reinterpret_cast<int *>(ering + 2 * T);     // [2][C] staged column ids
    int *srs = scid + 2 * C;                                // [2][T] first staged entry of the row
    int *ssp = srs + 2 * T;                                 // [2][T] first recent entry
    int *sre = ssp + 2 * T;                                 // [2][T] end of the row
    int *sflag = sre + 2 * T;                               // [2] the staged batch was flagged
    int *sdec = sflag + 2;                                  // [2] flags of batches b+2, b+3 (slot batch & 1)
    float *sw = reinterpret_cast<float *>(sflag + 4);       // [2][C] staged weights (HAS_W)
    const int n = a.n, B = a.B, nb = a.nb;
    if (threadIdx.x < T) {  // ---- computing threads ----
        const int t = threadIdx.x;
        __syncthreads();  // batch 0 staged
        for (int b = 0; b < nb; ++b) {
            const int b0 = b * B, b1 = min(n, b0 + B);
            const int m1 = b0 - B, m2 = b0 - 2 * B;  // first vertices of batches b-1 and b-2
            const double2 *ring1 = ring + ((b + 2) % 3) * T, *ring2 = ring + ((b + 1) % 3) * T;
            const int cur = b & 1;
            const int i = b0 + t;
            if (i < b1) {
                const int E0 = srs[cur * T];
                const int rs = srs[cur * T + t], sp = ssp[cur * T + t], re = sre[cur * T + t];
                const double2 v = sv[cur * T + t], r = sr[cur * T + t];
                double px, py;
                if (sflag[cur]) {
                    px = r.x;
                    py = r.y;
                } else {  // settled neighbours are final in memory (batches <= b - D < b - 2)
                    double ax = 0.0, ay = 0.0;
                    for (int e = rs; e < sp; ++e) {
                        const int k = cur * C + e - E0;
                        attr_add(spos[k], v, HAS_W ? a.W * (double)sw[k] : a.W, ax, ay);
                    }
                    px = ax - r.x;
                    py = ay - r.y;
                }
                for (int e = sp; e < re; ++e) {  // recent neighbours: the last two batches from the ring
                    const int k = cur * C + e - E0;
                    const int j = scid[k];
                    double2 pj = spos[k];
                    if (j >= m2 && j < b0) pj = j >= m1 ? ring1[j - m1] : ring2[j - m2];
                    attr_add(pj, v, HAS_W ? a.W * (double)sw[k] : a.W, px, py);
                }
                const double mag = sqrt(px * px + py * py);  // x_i += (F/|F|) * STEP, :1062-1066
                double2 q = v;
                q.x = q.x + (px / mag) * a.step;
                q.y = q.y + (py / mag) * a.step;
                ring[(b % 3) * T + t] = q;
                ering[cur * T + t] = px * px + py * py;
            }
            __syncthreads();  // end of batch b
        }
        __syncthreads();  // the loading threads' last stores are complete
        return;
    }
    // ---- loading threads ----
    const int t = threadIdx.x - T;
    // uniform values loaded with vector loads (a scalar load shares its wait counter with LDS)
    auto vload = [&](const int *p, int idx) {
        int v;
        __asm__("v_mov_b32 %0, %1" : "=v"(v) : "s"(idx));
        return p[v];
    };
    struct Range {
        int e0, e1, r0, r1;  // the batch's entries in the early and in the recent lists
    };
    auto range_of = [&](int bb) {
        const int ia = min(bb * B, n), ie = min(n, bb * B + B);
        Range R;
        R.e0 = vload(a.erp, ia);
        R.e1 = vload(a.erp, ie);
        R.r0 = vload(a.rrp, ia);
        R.r1 = vload(a.rrp, ie);
        return R;
    };
    auto flag_of = [&](int bb) {
        if (bb >= nb) return 0;
        int idx;
        __asm__("v_mov_b32 %0, %1" : "=v"(idx) : "s"(bb));
        return (int)(__hip_atomic_load(a.ready + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == a.gen);
    };
    // this thread's staged entries of a batch: [x0, x1) of the list (recent or early)
    auto load_cids = [&](bool flagged, int x0, int x1, int *c, float *w) {
        x0 = __builtin_amdgcn_readfirstlane(x0);
        x1 = __builtin_amdgcn_readfirstlane(x1);
        const int *ci = flagged ? a.rci : a.eci;
        const float *wi = flagged ? a.rw : a.ew;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int k = max(0, min(x0 + t + s * T, x1 - 1));
            c[s] = ci[k];
            if (HAS_W) w[s] = wi[k];
        }
    };
    int cA[S], cB[S];
    float wA[S], wB[S];
    double pAx[S], pAy[S];
    // flags of batches b+1 and b+2 (the path each is staged with).  A batch's flag is read once, by
    // loading thread 0, and reaches the other loading waves through LDS after a barrier: lanes that
    // read it themselves could see it change while a helper publishes and stage one batch two ways.
    bool fA, fB;
    Range R2, R3;        // entry ranges of batches b+2, b+3
    // stage batch bb (flag fl) into half `half`
    auto stage = [&](int bb, int half, bool fl) {
        const int i = min(bb * B + t, n - 1);
        int rs, sp, re;
        if (fl) {
            rs = a.rrp[i];
            re = a.rrp[i + 1];
            sp = rs;
        } else {
            rs = a.erp[i];
            sp = a.esp[i];
            re = a.erp[i + 1];
        }
        const double2 v = a.pos[i], r = fl ? a.pre[i] : a.rep[i];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int k = half * C + t + s * T;
            spos[k] = make_double2(pAx[s], pAy[s]);
            scid[k] = cA[s];
            if (HAS_W) sw[k] = wA[s];
        }
        srs[half * T + t] = rs;
        ssp[half * T + t] = sp;
        sre[half * T + t] = re;
        sv[half * T + t] = v;
        sr[half * T + t] = r;
        if (t == 0) sflag[half] = fl;
    };
    auto gather = [&]() {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double2 q = a.pos[cA[s]];
            pAx[s] = q.x;
            pAy[s] = q.y;
        }
    };
    int helped = 0;  // batches staged with the helpers' sums (loading thread 0)
    {
        // batches 0, 1 and 2 are formed by the walker
        const Range R0 = range_of(0), R1 = range_of(1);
        R2 = range_of(2);
        fA = false;
        if (t == 0) sdec[0] = 0;  // batch 2
        load_cids(false, R0.e0, R0.e1, cA, wA);
        gather();
        stage(0, 0, false);
        load_cids(false, R1.e0, R1.e1, cA, wA);
    }
    __syncthreads();  // batch 0 staged
    for (int b = 0; b < nb; ++b) {
        const bool f2 = sdec[b & 1] != 0;  // flag of batch b+2, read during batch b-1
        // acquire before reading anything a helper wrote (batch b+1 was flagged when its flag
        // was read; every load of the previous batch has completed, so the fence waits for none)
        if (fA) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        // the stores of batches [0, b-1) completed before the last barrier: publish them
        if (t == 0 && b >= 2 && (b - 1) % a.RK == 0) publish(a.done, b - 1);
        if (b > 0) {  // the moves of batch b-1 to memory
            const int c0 = (b - 1) * B;
            if (t < B && c0 + t < n) {
                a.pos[c0 + t] = ring[((b + 2) % 3) * T + t];
                a.eterm[c0 + t] = ering[((b + 1) & 1) * T + t];
            }
        }
        gather();  // batch b+1: positions of its staged entries
        fB = f2;   // batch b+2: path, column ids
        load_cids(fB, fB ? R2.r0 : R2.e0, fB ? R2.r1 : R2.e1, cB, wB);
        if (t == 0) sdec[(b + 1) & 1] = flag_of(b + 3);  // batch b+3: flag, ranges
        R3 = range_of(b + 3);
        stage(b + 1, (b + 1) & 1, fA);
        helped += (fA && b + 1 < nb) ? 1 : 0;
        // the stores of batch b-1 are complete before the barrier (gathers after it read them)
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();  // end of batch b
#pragma unroll
        for (int s = 0; s < S; ++s) {
            cA[s] = cB[s];
            if (HAS_W) wA[s] = wB[s];
        }
        fA = fB;
        R2 = R3;
    }
    // the moves of the last batch, then every batch is final: the helpers may stop
    const int c0 = (nb - 1) * B;
    if (t < B && c0 + t < n) {
        a.pos[c0 + t] = ring[((nb - 1) % 3) * T + t];
        a.eterm[c0 + t] = ering[((nb - 1) & 1) * T + t];
    }
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (t == 0) {
        a.done[1] = helped;
        publish(a.done, nb);
    }
}

template <bool HAS_W>
__global__ __launch_bounds__(2 * LAG_T) void k_walk_lag(LagArgs a) {
    extern __shared__ __align__(16) unsigned char lag_raw[];
    if (__builtin_amdgcn_readfirstlane(*a.status)) return;
    if (blockIdx.x == 0)
        lag_walker<HAS_W>(a, lag_raw);
    else
        lag_helper<HAS_W>(a, (int)blockIdx.x - 1, (int)gridDim.x - 1);
}

}  // namespace

// Whether the lagged walk takes batches of B vertices of graph g with lag D, and if so the
// settled / recent split of the early entries on the device (cached per batch size and lag).
bool lag_walk_fits(bl_graph *g, int B, int D) {
    if (g->lag_batch == B && g->lag_d == D) return g->lag_ok;
    g->lag_batch = B;
    g->lag_d = D;
    g->lag_ok = false;
    const int n = g->n;
    const int nb = (n + B - 1) / B;
    if (!(B <= LAG_T && D >= 3 && nb >= 2 * D && g->nnz > 0 && g->max_row <= LAG_MAXDEG &&
          (int)g->h_rowptr.size() == n + 1 && (int64_t)g->h_colids.size() == g->nnz))
        return false;
    const bool hw = !g->h_w.empty();
    std::vector<int32_t> erp(n + 1, 0), esp(n, 0), rrp(n + 1, 0), eci, rci;
    std::vector<float> ew, rw;
    eci.reserve(g->nnz / 2 + 1);
    for (int i = 0; i < n; ++i) {
        const int bi = i / B;
        for (int pass = 0; pass < 2; ++pass) {  // settled first, then recent, each in CSR order
            if (pass == 1) esp[i] = (int32_t)eci.size();
            for (int e = g->h_rowptr[i]; e < g->h_rowptr[i + 1]; ++e) {
                const int bj = g->h_colids[e] / B;
                if (bj >= bi) continue;  // late: summed by k_late_attract
                const bool settled = bj <= bi - D;
                if (settled != (pass == 0)) continue;
                eci.push_back(g->h_colids[e]);
                if (hw) ew.push_back(g->h_w[e]);
                if (!settled) {
                    rci.push_back(g->h_colids[e]);
                    if (hw) rw.push_back(g->h_w[e]);
                }
            }
        }
        erp[i + 1] = (int32_t)eci.size();
        rrp[i + 1] = (int32_t)rci.size();
    }
    for (int b0 = 0; b0 < n; b0 += B)
        if (erp[std::min(n, b0 + B)] - erp[b0] > LAG_CAP) return false;
    // the loading threads read index 0 of an empty list (clamped loads): keep one valid entry
    if (eci.empty()) eci.push_back(0), ew.push_back(0.0f);
    if (rci.empty()) rci.push_back(0), rw.push_back(0.0f);
    if (g->lag_erp.ensure(4 * (size_t)(n + 1)) || g->lag_esp.ensure(4 * (size_t)n) ||
        g->lag_rrp.ensure(4 * (size_t)(n + 1)) || g->lag_eci.ensure(4 * eci.size()) ||
        g->lag_rci.ensure(4 * rci.size()) || g->lag_pre.ensure(16 * (size_t)n) ||
        g->lag_ready.ensure(4 * (size_t)nb + 256))
        return false;
    if (hw && (g->lag_ew.ensure(4 * ew.size()) || g->lag_rw.ensure(4 * rw.size()))) return false;
    if (hipMemcpy(g->lag_erp.p, erp.data(), 4 * (size_t)(n + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lag_esp.p, esp.data(), 4 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lag_rrp.p, rrp.data(), 4 * (size_t)(n + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lag_eci.p, eci.data(), 4 * eci.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lag_rci.p, rci.data(), 4 * rci.size(), hipMemcpyHostToDevice) != hipSuccess)
        return false;
    if (hw && (hipMemcpy(g->lag_ew.p, ew.data(), 4 * ew.size(), hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(g->lag_rw.p, rw.data(), 4 * rw.size(), hipMemcpyHostToDevice) != hipSuccess))
        return false;
    // flags hold launch generations (never 0); done at the end of the ready array
    if (hipMemset(g->lag_ready.p, 0, 4 * (size_t)nb + 256) != hipSuccess) return false;
    g->lag_gen = 0;
    g->lag_ok = true;
    return true;
}

size_t lag_walk_lds(bool has_w) { return lag_lds(has_w); }

// the count of batches the last lagged walk staged with the helpers' sums
int *lag_walk_helped(bl_graph *g) { return g->lag_ready.as<int>() + (g->n + g->lag_batch - 1) / g->lag_batch + 33; }

// One iteration's walk (after k_late_attract has folded the late attraction into the repulsion).
int launch_walk_lag(bl_graph *g, double2 *pos, const bl_params *p, double step, int helpers, int rk, hipStream_t s) {
    const int n = g->n, B = p->batch;
    LagArgs a;
    a.erp = g->lag_erp.as<int>();
    a.esp = g->lag_esp.as<int>();
    a.eci = g->lag_eci.as<int>();
    a.ew = g->has_w ? g->lag_ew.as<float>() : nullptr;
    a.rrp = g->lag_rrp.as<int>();
    a.rci = g->lag_rci.as<int>();
    a.rw = g->has_w ? g->lag_rw.as<float>() : nullptr;
    a.W = p->weight;
    a.pos = pos;
    a.rep = g->bh.rep.as<double2>();
    a.pre = g->lag_pre.as<double2>();
    a.nb = (n + B - 1) / B;
    a.ready = g->lag_ready.as<int>();
    a.done = a.ready + a.nb + 32;  // its own cache line
    a.gen = ++g->lag_gen;
    if (a.gen == 0x7fffffff) return BL_E_DEVICE;  // generations exhausted (2^31 launches)
    a.n = n;
    a.B = B;
    a.D = g->lag_d;
    a.RK = std::max(1, rk);
    a.step = step;
    a.eterm = g->eterm.as<double>();
    a.status = g->bh.misc.as<int>();
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device) != hipSuccess || khz <= 0)
        khz = 100000;
    a.ticks = (long long)khz * 1000;  // 1 s
    BL_HIP(hipMemsetAsync(a.done, 0, 2 * sizeof(int), s));
    const size_t lds = lag_lds(g->has_w);
    if (g->has_w)
        hipLaunchKernelGGL(k_walk_lag<true>, dim3(1 + helpers), dim3(2 * LAG_T), lds, s, a);
    else
        hipLaunchKernelGGL(k_walk_lag<false>, dim3(1 + helpers), dim3(2 * LAG_T), lds, s, a);
    BL_HIP(hipGetLastError());
    return BL_OK;
}

int lag_walk_prepare(bool has_w) {
    const int lb = (int)lag_lds(has_w);
    BL_HIP(hipFuncSetAttribute(has_w ? reinterpret_cast<const void *>(k_walk_lag<true>)
                                     : reinterpret_cast<const void *>(k_walk_lag<false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, lb));
    return BL_OK;
}

}  // namespace bl
