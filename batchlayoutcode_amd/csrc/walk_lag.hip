// The lagged batch walk of the Barnes-Hut family (algorithms.cpp:1050-1066: per batch, the
// attraction of every CSR entry with live positions minus the repulsion, then the sequential
// move), for batches of at most LAG_T vertices whose rows are short.
//
// A batch's attraction reads the positions its neighbours have at that moment, so batch b
// depends on every earlier batch that holds one of its neighbours.  On graphs with unordered ids
// those neighbours are spread over all earlier batches and few lie in the last D - 1.  Each
// vertex's "early" neighbours (in earlier batches, moved this iteration) are therefore split into
//   settled  in batches <= b - D, final once batch b - D has moved, and
//   recent   in batches b - D + 1 .. b - 1 (at most LAG_RS per vertex);
// the "late" ones (batch b or later, still at their iteration-start positions) are summed
// beforehand by k_late_attract into the repulsion: rep'_i = rep_i - sum_late attraction.
//
//   f_i = pre_i + sum_{recent} attraction,     pre_i = sum_{settled} attraction - rep'_i
//
// Workgroups 1..H, the helpers: helper h takes batches h-1, h-1+H, ...; it loads what does not
// change (its vertices' positions, rep', settled column ids) at once, waits until the walker has
// published batches [0, b - D + 1), gathers the settled positions, writes pre_i of the batch to
// memory and flags the batch with the launch's generation number.
//
// What one workgroup writes for another (moved positions, pre_i) is stored with agent-scope
// stores, which the XCD's L2 writes through to memory: once such a store has completed
// (s_waitcnt), a relaxed flag store issued after it publishes it, and a reader on another XCD
// that sees the flag and then acquires (invalidating its own L2's stale lines) reads the new
// value.  A release fence instead writes back the whole L2 and took about 0.6 us, every second
// batch, on the walker's critical path.
//
// Workgroup 0, the walker, never waits for a helper:
//   computing waves (LAG_T threads, thread t owns vertex b0 + t of every batch) read only LDS: a
//     staged batch (own position, pre_i or rep'_i, the ring slots of the recent neighbours) and a
//     ring of the last D batches' moved positions, where every recent neighbour is found.  A
//     batch without a helper's flag forms pre_i itself from memory, with the same FP64 operations
//     in the same (CSR) order as a helper, so the layout is the same whichever formed it.  Moved
//     positions and energy terms go straight to memory;
//   loading waves stage batch b + 3 during batch b with direct global-to-LDS loads (4 slots), and
//     wait only for the loads of batch b + 1, so three batches of loads are in flight;
//   a control wave reads the helpers' flags two batches at a time, LAG_LEAD batches ahead,
//     acquires, and records for each batch, before the batch is staged, whether its pre_i is used;
//     every batch it publishes the walker's progress to the helpers (batches whose stores
//     completed: the computing waves wait for the stores of three batches back).
// The computing waves' arithmetic is FP64 throughout; |d| and the unit vector F/|F| come from
// the hardware reciprocal square root refined by two Newton steps (a few ulp from the IEEE
// square root and division the row walk uses).
// Barriers are raw s_barrier with explicit LDS waits (a workgroup barrier would also wait for the
// loads in flight).  Helpers stop at a deadline or when the walker has finished, so the launch
// ends whatever the scheduling.
#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "bh_common.h"

namespace bl {
namespace {

typedef __attribute__((address_space(3))) void lds_void;
typedef int v4i __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ T bits_as(v4i r) {
    static_assert(sizeof(T) == sizeof(v4i), "16-byte words only");
    T x;
    __builtin_memcpy(&x, &r, sizeof(T));
    return x;
}

// LDS reads issued back to back and waited for in the same asm block (one round trip; the
// outputs are early-clobber, so no read lands in a register a later read takes its address from)
__device__ __forceinline__ void lds_read_stage(const void *f, const void *a, const void *b, const void *c, int &fl,
                                               v4i &ra, v4i &rb, v4i &rc) {
    __asm__ volatile(
        "ds_read_b32 %0, %4\n\tds_read_b128 %1, %5\n\tds_read_b128 %2, %6\n\tds_read_b128 %3, %7\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(fl), "=&v"(ra), "=&v"(rb), "=&v"(rc)
        : "v"((const lds_void *)f), "v"((const lds_void *)a), "v"((const lds_void *)b), "v"((const lds_void *)c)
        : "memory");
}

__device__ __forceinline__ void lds_read4(const void *a, const void *b, const void *c, const void *d, v4i &ra, v4i &rb,
                                          v4i &rc, v4i &rd) {
    __asm__ volatile(
        "ds_read_b128 %0, %4\n\tds_read_b128 %1, %5\n\tds_read_b128 %2, %6\n\tds_read_b128 %3, %7\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(ra), "=&v"(rb), "=&v"(rc), "=&v"(rd)
        : "v"((const lds_void *)a), "v"((const lds_void *)b), "v"((const lds_void *)c), "v"((const lds_void *)d)
        : "memory");
}

constexpr int LAG_T = 256;              // batch vertices; the walker's computing and loading threads
constexpr int LAG_WG = 2 * LAG_T + 64;  // + one control wave
constexpr int LAG_K = 4;                // stage slots: batch b + 3 is loaded during batch b
constexpr int LAG_DMAX = 27;            // largest lag (ring slots of moved positions)
constexpr int LAG_RS = 4;               // recent neighbours per vertex (ring offsets in an int4)
constexpr int LAG_LEAD = 6;             // a batch's flag is read LAG_LEAD or LAG_LEAD + 1 batches ahead
constexpr int LAG_DEC = 16;             // decision ring (batch % LAG_DEC)
constexpr int LAG_MAXDEG = 64;          // longest row for which this walk is chosen
constexpr int LAG_HS = 8;               // settled entries loaded together (helpers: before the wait)

struct LagArgs {
    const int *erp, *esp, *eci;  // early entries of vertex i: [erp[i], esp[i]) settled, then recent
    const float *ew;
    const int4 *rofs;            // ring offsets of vertex i's recent neighbours (-1: none), CSR order
    const float4 *rw;            // their weights
    double W;
    double2 *pos;
    const double2 *rep;          // repulsion minus the late attraction (k_late_attract)
    double2 *pre;                // pre_i written by the helpers
    int *ready;                  // ready[b] == gen: pre of batch b is in memory
    int *done;                   // batches [0, *done) of this launch are final in memory; done[1]: helped
    int gen;
    int n, B, nb, D;
    double step;
    double *eterm;
    const int *status;
    long long ticks;             // a helper's longest wait
};

constexpr size_t LDS_MAX = 160 * 1024;  // LDS of one workgroup on gfx950

size_t lag_lds_bytes(bool has_w, int D) {
    return sizeof(double2) * (size_t)D * LAG_T + (size_t)LAG_K * LAG_T * (16 + 16 + 16 + (has_w ? 16 : 0)) +
           sizeof(int) * (LAG_DEC + 4);
}

// The largest lag whose ring fits the LDS beside the stage (27 without weights, 23 with).
int lag_dmax(bool has_w) {
    int D = LAG_DMAX;
    while (D > 0 && lag_lds_bytes(has_w, D) > LDS_MAX) --D;
    return D;
}

// Publishes *flag = v to other workgroups after every store of this thread (and, through a
// preceding workgroup barrier, of its workgroup): a release fence, whose L2 write-back is waited
// for explicitly, then a relaxed store.
__device__ __forceinline__ void publish(int *flag, int v) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __hip_atomic_store(flag, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The walker's barrier: its LDS accesses complete, then s_barrier (no wait for memory loads).
__device__ __forceinline__ void wbar() {
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

// Column ids (and weights) of the first LAG_HS settled entries [e0, e1) (clamped loads).
template <bool HAS_W>
__device__ __forceinline__ void settled_ids(const LagArgs &a, int e0, int e1, int *c, float *w) {
#pragma unroll
    for (int g = 0; g < LAG_HS; ++g) {
        const int k = max(0, min(e0 + g, e1 - 1));  // an empty row reads entry 0
        c[g] = a.eci[k];
        w[g] = HAS_W ? a.ew[k] : 0.0f;
    }
}

// pre_i = sum over the settled entries [e0, e1) in CSR order - r, the first LAG_HS column ids
// and weights given in c / w.
template <bool HAS_W>
__device__ __forceinline__ double2 settled_sum(const LagArgs &a, int e0, int e1, const int *c, const float *w,
                                               double2 v, double2 r) {
    double ax = 0.0, ay = 0.0;
    double2 pj[LAG_HS];
#pragma unroll
    for (int g = 0; g < LAG_HS; ++g) pj[g] = a.pos[c[g]];
#pragma unroll
    for (int g = 0; g < LAG_HS; ++g)
        if (e0 + g < e1) attr_add(pj[g], v, HAS_W ? a.W * (double)w[g] : a.W, ax, ay);
    for (int e = e0 + LAG_HS; e < e1; ++e)  // long rows: the rest one by one
        attr_add(a.pos[a.eci[e]], v, HAS_W ? a.W * (double)a.ew[e] : a.W, ax, ay);
    return make_double2(ax - r.x, ay - r.y);
}

template <bool HAS_W>
__device__ void lag_helper(const LagArgs &a, int hid, int nhelp) {
    __shared__ int ok;
    const int t = threadIdx.x;
    for (int b = hid; b < a.nb; b += nhelp) {
        __syncthreads();  // ok of the previous batch has been read
        const int b0 = b * a.B, bc = min(a.B, a.n - b0);
        const int i = b0 + min(t, bc - 1);
        // what does not change during the walk, before the wait; batch b has not moved (the
        // walker decides whether to use this batch's sums several batches before it)
        const double2 v = a.pos[i];
        const double2 r = a.rep[i];
        const int e0 = a.erp[i], e1 = a.esp[i];
        int c[LAG_HS];
        float w[LAG_HS];
        settled_ids<HAS_W>(a, e0, e1, c, w);
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
            // the walker reads a batch's flag at least LAG_LEAD batches ahead: a batch it has
            // (nearly) reached is formed by the walker
            if (ok > b - LAG_LEAD - 2) continue;
        }
        if (t < bc) {  // agent-scope stores: in memory once complete
            const double2 f = settled_sum<HAS_W>(a, e0, e1, c, w, v, r);
            __hip_atomic_store(&a.pre[i].x, f.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&a.pre[i].y, f.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // every wave's stores have completed before thread 0 flags the batch
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (t == 0) __hip_atomic_store(a.ready + b, a.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <bool HAS_W>
__device__ void lag_walker(const LagArgs &a, unsigned char *raw) {
    constexpr int T = LAG_T, K = LAG_K;
    const int n = a.n, B = a.B, nb = a.nb, D = a.D;
    double2 *ring = reinterpret_cast<double2 *>(raw);     // [D][T] moved positions of batch c (slot c % D)
    double2 *sv = ring + (size_t)D * T;                   // [K][T] staged own positions
    double2 *sp = sv + K * T;                             // [K][T] staged pre_i (flagged) or rep'_i
    int4 *so = reinterpret_cast<int4 *>(sp + K * T);      // [K][T] staged ring offsets of the recent neighbours
    float4 *sw = reinterpret_cast<float4 *>(so + K * T);  // [K][T] their weights (HAS_W)
    int *dec = reinterpret_cast<int *>(HAS_W ? (void *)(sw + K * T) : (void *)sw);  // [LAG_DEC] flags used
    int *shelp = dec + LAG_DEC;                           // batches formed from the helpers' sums
    const int tid = threadIdx.x;
    if (tid < T) {  // ---- computing waves ----
        const int t = tid;
        int helped = 0;
        wbar();  // batch 0 staged (and the decisions cleared)
        for (int b = 0; b < nb; ++b) {
            const int s = (b % K) * T + t;
            const int b0 = b * B, i = b0 + t;
            // one LDS round trip: the batch's flag and this thread's whole stage
            int fl;
            v4i w0, w1, w2, w3 = {0, 0, 0, 0};
            lds_read_stage(dec + b % LAG_DEC, sv + s, sp + s, so + s, fl, w0, w1, w2);
            if (HAS_W) w3 = reinterpret_cast<const v4i *>(sw)[s];  // before the first use of w0..w2
            const double2 v = bits_as<double2>(w0);
            double2 p = bits_as<double2>(w1);
            const int4 ro = bits_as<int4>(w2);
            const float4 rw = bits_as<float4>(w3);
            const bool flagged = __builtin_amdgcn_readfirstlane(fl) != 0;
            helped += flagged;
            // the last batch's threads past n compute on stale stage words and store nothing
            const bool live = i < min(n, b0 + B);
            if (live && !flagged) {  // the helper was late (or the batch is one of the first): form pre_i
                const int e0 = a.erp[i], e1 = a.esp[i];
                int c[LAG_HS];
                float w[LAG_HS];
                settled_ids<HAS_W>(a, e0, e1, c, w);
                p = settled_sum<HAS_W>(a, e0, e1, c, w, v, p);
            }
            const int roff[LAG_RS] = {ro.x, ro.y, ro.z, ro.w};
            const float rwk[LAG_RS] = {rw.x, rw.y, rw.z, rw.w};
            // a second round trip: every ring slot this vertex may need (clamped offsets)
            double2 rj[LAG_RS];
            {
                static_assert(LAG_RS == 4, "four ring reads");
                v4i q0, q1, q2, q3;
                lds_read4(ring + max(roff[0], 0), ring + max(roff[1], 0), ring + max(roff[2], 0),
                          ring + max(roff[3], 0), q0, q1, q2, q3);
                rj[0] = bits_as<double2>(q0);
                rj[1] = bits_as<double2>(q1);
                rj[2] = bits_as<double2>(q2);
                rj[3] = bits_as<double2>(q3);
            }
#pragma unroll
            for (int k = 0; k < LAG_RS; ++k)  // recent neighbours, from the ring, in CSR order
                if (roff[k] >= 0) attr_add(rj[k], v, HAS_W ? a.W * (double)rwk[k] : a.W, p.x, p.y);
            // x_i += (F/|F|) * STEP (:1062-1066), the unit vector F * rsqrt2(|F|^2)
            const double m2 = p.x * p.x + p.y * p.y;
            const double r = rsqrt2(m2) * a.step;
            double2 q;
            q.x = fma(p.x, r, v.x);
            q.y = fma(p.y, r, v.y);
            if (live) {
                ring[(b % D) * T + t] = q;
                // agent-scope stores are written through to memory, so the helpers (on other
                // XCDs) read them once they have completed: no L2 write-back is needed to publish
                __hip_atomic_store(&a.pos[i].x, q.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&a.pos[i].y, q.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                a.eterm[i] = m2;
            }
            // the stores of batch b-3 are complete (three per batch, in order; the control wave
            // publishes them after a barrier)
            __asm__ volatile("s_waitcnt vmcnt(9)" ::: "memory");
            wbar();  // end of batch b
        }
        if (t == 0) *shelp = helped;
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        wbar();
        return;
    }
    if (tid < 2 * T) {  // ---- loading waves: batch b + 3 during batch b ----
        const int t = tid - T, w = t >> 6;
        constexpr int OPS = HAS_W ? 4 : 3;  // loads per batch and wave
        // one batch into its slot; pre_i if the control wave recorded a usable flag, else rep'_i
        auto issue = [&](int x, bool use_dec) {
            const int slot = (x % K) * T + w * 64;  // this wave's 64 vertices of the slot
            const int i = min(x * B + t, n - 1);
            int f = 0;
            if (use_dec) {  // an LDS read the compiler does not order after the loads in flight
                const lds_void *dp = (const lds_void *)(dec + x % LAG_DEC);
                __asm__ volatile("ds_read_b32 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(f) : "v"(dp) : "memory");
            }
            __builtin_amdgcn_global_load_lds((const void *)(a.pos + i), (lds_void *)(sv + slot), 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(f ? a.pre + i : a.rep + i), (lds_void *)(sp + slot), 16,
                                             0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(a.rofs + i), (lds_void *)(so + slot), 16, 0, 0);
            if (HAS_W) __builtin_amdgcn_global_load_lds((const void *)(a.rw + i), (lds_void *)(sw + slot), 16, 0, 0);
        };
        // the loads of batch x have landed, given that batches x+1 .. last were issued after it
        auto land = [&](int x, int last) {
            const int younger = max(0, last - x);
            if (younger >= 2) {
                if (OPS == 4) __asm__ volatile("s_waitcnt vmcnt(8)" ::: "memory");
                else __asm__ volatile("s_waitcnt vmcnt(6)" ::: "memory");
            } else if (younger == 1) {
                if (OPS == 4) __asm__ volatile("s_waitcnt vmcnt(4)" ::: "memory");
                else __asm__ volatile("s_waitcnt vmcnt(3)" ::: "memory");
            } else {
                __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
            }
        };
        // prologue: batches 0..2 unflagged (before the control wave's first barrier)
        for (int x = 0; x < K - 1 && x < nb; ++x) issue(x, false);
        land(0, min(K - 2, nb - 1));
        wbar();
        for (int b = 0; b < nb; ++b) {
            const int x = b + K - 1;  // slot (b-1) % K: the computing waves left it at the last barrier
            if (x < nb) issue(x, true);
            if (b + 1 < nb) land(b + 1, min(x, nb - 1));
            wbar();  // end of batch b
        }
        wbar();
        return;
    }
    // ---- control wave ----
    const int l = tid - 2 * T;
    if (l < LAG_DEC) dec[l] = 0;  // the first batches are formed by the walker
    int val = 0;
    wbar();
    for (int b = 0; b < nb; ++b) {
        if ((b & 1) == 0) {
            if (b >= 2) {  // the flags read two batches ago, of batches b - 2 + LAG_LEAD + l
                const int xx = b - 2 + LAG_LEAD + l;
                const bool good = l < 2 && xx < nb && val == a.gen;
                if (__ballot(good)) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                // batch xx is staged during batch xx - 3 >= b + 1, after this batch's barrier
                if (l < 2 && xx < nb) dec[xx % LAG_DEC] = good ? 1 : 0;
            }
            val = 0;
            const int y = b + LAG_LEAD + l;
            if (l < 2 && y < nb) val = __hip_atomic_load(a.ready + y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // batches [0, b - 3): the computing waves completed batch b - 4's stores before the last
        // barrier; they are agent-scope stores, in memory once complete, so a relaxed store
        // publishes them (a release fence would write the L2 back: 0.6 us)
        if (l == 0 && b >= 4) __hip_atomic_store(a.done, b - 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        wbar();  // end of batch b
    }
    wbar();
    if (l == 0) {
        a.done[1] = *shelp;
        publish(a.done, nb);  // every batch is final: the helpers may stop
    }
}

template <bool HAS_W>
__global__ __launch_bounds__(LAG_WG) void k_walk_lag(LagArgs a) {
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
    if (!(B <= LAG_T && D >= 4 && D <= lag_dmax(!g->h_w.empty()) && nb >= 2 * D && nb >= LAG_LEAD + 4 && g->nnz > 0 &&
          g->max_row <= LAG_MAXDEG && (int)g->h_rowptr.size() == n + 1 && (int64_t)g->h_colids.size() == g->nnz))
        return false;
    const bool hw = !g->h_w.empty();
    std::vector<int32_t> erp(n + 1, 0), esp(n, 0), eci;
    std::vector<float> ew;
    std::vector<int32_t> rofs(4 * (size_t)n, -1);
    std::vector<float> rw(hw ? 4 * (size_t)n : 0, 0.0f);
    eci.reserve(g->nnz / 2 + 1);
    for (int i = 0; i < n; ++i) {
        const int bi = i / B;
        int nr = 0;
        for (int pass = 0; pass < 2; ++pass) {  // settled first, then recent, each in CSR order
            if (pass == 1) esp[i] = (int32_t)eci.size();
            for (int e = g->h_rowptr[i]; e < g->h_rowptr[i + 1]; ++e) {
                const int j = g->h_colids[e], bj = j / B;
                if (bj >= bi) continue;  // late: summed by k_late_attract
                const bool settled = bj <= bi - D;
                if (settled != (pass == 0)) continue;
                eci.push_back(j);
                if (hw) ew.push_back(g->h_w[e]);
                if (!settled) {
                    if (nr == LAG_RS) return false;  // more recent neighbours than the stage holds
                    rofs[4 * (size_t)i + nr] = (bj % D) * LAG_T + (j - bj * B);
                    if (hw) rw[4 * (size_t)i + nr] = g->h_w[e];
                    ++nr;
                }
            }
        }
        erp[i + 1] = (int32_t)eci.size();
    }
    // the clamped loads read index 0 of an empty list: keep one valid entry
    if (eci.empty()) eci.push_back(0), ew.push_back(0.0f);
    if (g->lag_erp.ensure(4 * (size_t)(n + 1)) || g->lag_esp.ensure(4 * (size_t)n) ||
        g->lag_eci.ensure(4 * eci.size()) || g->lag_rofs.ensure(16 * (size_t)n) || g->lag_pre.ensure(16 * (size_t)n) ||
        g->lag_ready.ensure(4 * (size_t)nb + 256))
        return false;
    if (hw && (g->lag_ew.ensure(4 * ew.size()) || g->lag_rw.ensure(16 * (size_t)n))) return false;
    if (hipMemcpy(g->lag_erp.p, erp.data(), 4 * (size_t)(n + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lag_esp.p, esp.data(), 4 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lag_eci.p, eci.data(), 4 * eci.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lag_rofs.p, rofs.data(), 16 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
        return false;
    if (hw && (hipMemcpy(g->lag_ew.p, ew.data(), 4 * ew.size(), hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(g->lag_rw.p, rw.data(), 16 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess))
        return false;
    // flags hold launch generations (never 0); done (and the helped count) after the flags
    if (hipMemset(g->lag_ready.p, 0, 4 * (size_t)nb + 256) != hipSuccess) return false;
    g->lag_gen = 0;
    g->lag_ok = true;
    return true;
}

// The lag the walk takes by default for batches of B vertices (0: the walk does not fit): the
// largest that the LDS holds, at most half the batches, and at most the smallest distance, over
// the vertices with five or more neighbours in earlier batches, from a vertex's batch to the
// batch of its fifth-nearest such neighbour (so no vertex has more than LAG_RS = 4 recent
// neighbours).  Cached per batch size.
int lag_walk_choose(bl_graph *g, int B) {
    if (g->lag_auto_batch == B) return g->lag_auto_d;
    g->lag_auto_batch = B;
    g->lag_auto_d = 0;
    const int n = g->n;
    const int nb = (n + B - 1) / B;
    if (B > LAG_T || (int)g->h_rowptr.size() != n + 1 || g->max_row > LAG_MAXDEG) return 0;
    int D = std::min(lag_dmax(!g->h_w.empty()), nb / 2);
    static_assert(LAG_RS == 4, "the fifth-nearest gap bounds the lag");
    for (int i = 0; i < n && D >= 4; ++i) {
        const int bi = i / B;
        int gap[LAG_RS + 1], ng = 0;  // the five smallest gaps bi - bj >= 1, ascending
        for (int e = g->h_rowptr[i]; e < g->h_rowptr[i + 1]; ++e) {
            const int k = bi - g->h_colids[e] / B;
            if (k < 1 || (ng == LAG_RS + 1 && k >= gap[LAG_RS])) continue;
            int m = ng < LAG_RS + 1 ? ng++ : LAG_RS;
            while (m > 0 && gap[m - 1] > k) gap[m] = gap[m - 1], --m;
            gap[m] = k;
        }
        if (ng == LAG_RS + 1) D = std::min(D, gap[LAG_RS]);  // recent: gaps 1 .. D - 1
    }
    g->lag_auto_d = D >= 4 ? D : 0;
    return g->lag_auto_d;
}

// the count of batches the last lagged walk staged with the helpers' sums
int *lag_walk_helped(bl_graph *g) { return g->lag_ready.as<int>() + (g->n + g->lag_batch - 1) / g->lag_batch + 33; }

// One iteration's walk (after k_late_attract has folded the late attraction into the repulsion).
int launch_walk_lag(bl_graph *g, double2 *pos, const bl_params *p, double step, int helpers, hipStream_t s) {
    const int n = g->n, B = p->batch;
    LagArgs a;
    a.erp = g->lag_erp.as<int>();
    a.esp = g->lag_esp.as<int>();
    a.eci = g->lag_eci.as<int>();
    a.ew = g->has_w ? g->lag_ew.as<float>() : nullptr;
    a.rofs = g->lag_rofs.as<int4>();
    a.rw = g->has_w ? g->lag_rw.as<float4>() : nullptr;
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
    a.step = step;
    a.eterm = g->eterm.as<double>();
    a.status = g->bh.misc.as<int>();
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device) != hipSuccess || khz <= 0)
        khz = 100000;
    a.ticks = (long long)khz * 1000;  // 1 s
    BL_HIP(hipMemsetAsync(a.done, 0, 2 * sizeof(int), s));
    const size_t lds = lag_lds_bytes(g->has_w, a.D);
    if (g->has_w)
        hipLaunchKernelGGL(k_walk_lag<true>, dim3(1 + helpers), dim3(LAG_WG), lds, s, a);
    else
        hipLaunchKernelGGL(k_walk_lag<false>, dim3(1 + helpers), dim3(LAG_WG), lds, s, a);
    BL_HIP(hipGetLastError());
    return BL_OK;
}

int lag_walk_prepare(bool has_w) {
    const int lb = (int)lag_lds_bytes(has_w, lag_dmax(has_w));
    BL_HIP(hipFuncSetAttribute(has_w ? reinterpret_cast<const void *>(k_walk_lag<true>)
                                     : reinterpret_cast<const void *>(k_walk_lag<false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, lb));
    return BL_OK;
}

}  // namespace bl
