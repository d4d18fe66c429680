// The lagged batch walk of the Barnes-Hut family (algorithms.cpp:1050-1066: per batch, the
// attraction of every CSR entry with live positions minus the repulsion, then the sequential
// move), for batches of at most 256 vertices whose rows are short.
//
// Groups.  Consecutive batches with no edge between them (no vertex of a later one has a
// neighbour in an earlier one) may move in one step: each vertex's force reads the same
// positions whether its group's other batches moved before or after it.  On large graphs with
// unordered ids most pairs of consecutive batches are such (80% on the 1.66M-vertex Delaunay
// graph at -batch 256), so the walk takes the batches in groups of up to T / B consecutive
// batches, T = 512 vertices, where that makes a fifth fewer steps or better, and one batch a
// step (T = 256) otherwise.  Below, "group" is the walk's step; the layout is the same bits
// either way.
//
// A group's attraction reads the positions its neighbours have at that moment, so group g
// depends on every earlier group that holds one of its neighbours.  On graphs with unordered
// ids those neighbours are spread over all earlier groups and few lie in the last D - 1.  Each
// vertex's "early" neighbours (in earlier batches, moved this iteration) are therefore split
// into
//   settled  in groups <= g - D, final once group g - D has moved, and
//   recent   in groups g - D + 1 .. g - 1 (at most LAG_RS per vertex);
// the "late" ones (the vertex's batch or later, still at their iteration-start positions) are
// summed beforehand by k_late_attract into the repulsion: rep'_i = rep_i - sum_late attraction.
//
//   f_i = pre_i + sum_{recent} attraction,     pre_i = sum_{settled} attraction - rep'_i
//
// Workgroups 1..H, the helpers: helper h takes groups h-1, h-1+H, ...; it loads what does not
// change (its vertices' positions, rep', settled column ids) at once, waits until the walker has
// published groups [0, g - D + 1), gathers the settled positions, writes pre_i of the group to
// memory and flags the group with the launch's generation number.
//
// What one workgroup writes for another (moved positions, pre_i) is stored with agent-scope
// stores, which the XCD's L2 writes through to memory: once such a store has completed
// (s_waitcnt), a relaxed flag store issued after it publishes it, and a reader on another XCD
// that sees the flag and then acquires (invalidating its own L2's stale lines) reads the new
// value.  A release fence instead writes back the whole L2 and took about 0.6 us, every second
// batch, on the walker's critical path.
//
// Workgroup 0, the walker, never waits for a helper:
//   computing waves (T threads, thread t owns vertex t of every group) read only LDS: a staged
//     group (its first vertex and size, own position, pre_i or rep'_i, the ring slots of the
//     recent neighbours, the ring slot of the vertex's own moved position) and a ring of the
//     moved positions that later groups read as recent (the host assigns the slots, so that no
//     slot is overwritten while a group within the lag still reads it).  A group without a
//     helper's flag forms pre_i itself from memory, with the same FP64 operations in the same
//     (CSR) order as a helper, so the layout is the same whichever formed it.  Moved positions
//     and energy terms go straight to memory, three stores per thread and group (threads past
//     the group's end store to a sink);
//   loading waves stage group g + 3 during group g with direct global-to-LDS loads (4 slots), and
//     wait only for the loads of group g + 1, so three groups of loads are in flight;
//   a control wave reads the helpers' flags two groups at a time, LAG_LEAD groups ahead,
//     acquires, and records for each group, before the group is staged, whether its pre_i is used;
//     every group it publishes the walker's progress to the helpers (groups whose stores
//     completed: the computing waves wait for the stores of three groups back).
// The computing waves' arithmetic is FP64 throughout; |d| and the unit vector F/|F| come from
// the hardware reciprocal square root refined by one Newton step (about 20 ulp from the IEEE
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
typedef int v2i __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ T bits_as(v4i r) {
    static_assert(sizeof(T) == sizeof(v4i), "16-byte words only");
    T x;
    __builtin_memcpy(&x, &r, sizeof(T));
    return x;
}

// LDS reads issued back to back and waited for in the same asm block (one round trip; the
// outputs are early-clobber, so no read lands in a register a later read takes its address from):
// the group's flag and header, and this thread's own position, pre_i or rep'_i, recent ring
// slots and own ring slot
__device__ __forceinline__ void lds_read_stage(const void *f, const void *h, const void *a, const void *b,
                                               const void *c, const void *d, int &fl, v2i &rh, v4i &ra, v4i &rb,
                                               v4i &rc, int &rd) {
    __asm__ volatile(
        "ds_read_b32 %0, %6\n\tds_read_b64 %1, %7\n\tds_read_b128 %2, %8\n\tds_read_b128 %3, %9\n\t"
        "ds_read_b128 %4, %10\n\tds_read_b32 %5, %11\n\ts_waitcnt lgkmcnt(0)"
        : "=&v"(fl), "=&v"(rh), "=&v"(ra), "=&v"(rb), "=&v"(rc), "=&v"(rd)
        : "v"((const lds_void *)f), "v"((const lds_void *)h), "v"((const lds_void *)a), "v"((const lds_void *)b),
          "v"((const lds_void *)c), "v"((const lds_void *)d)
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

constexpr int LAG_K = 4;        // stage slots: group g + 3 is loaded during group g
constexpr int LAG_DMAX = 48;    // largest lag (groups)
constexpr int LAG_RS = 4;       // recent neighbours per vertex (ring slots in an int4)
constexpr int LAG_LEAD = 6;     // a group's flag is read LAG_LEAD or LAG_LEAD + 1 groups ahead
constexpr int LAG_DEC = 16;     // decision ring (group % LAG_DEC)
constexpr int LAG_MAXDEG = 64;  // longest row for which this walk is chosen
constexpr int LAG_HS = 8;       // settled entries loaded together (helpers: before the wait)
constexpr int LAG_BMAX = 256;   // largest batch

// The walker of group width T: T computing threads, LW loading waves (chunk c of 64 vertices of
// a staged group is loaded by wave c % LW), one control wave.
template <int T>
struct LagShape {
    static_assert(T == 256 || T == 512, "group widths 256 and 512");
    static constexpr int LW = T == 256 ? 4 : 7;
    static constexpr int WG = T + 64 * LW + 64;
    static constexpr int CH = T / 64;
};

struct LagArgs {
    const int *erp, *esp, *eci;  // early entries of vertex i: [erp[i], esp[i]) settled, then recent
    const float *ew;
    const int4 *rofs;            // ring slots of vertex i's recent neighbours (-1: none), CSR order
    const float4 *rw;            // their weights
    const int *ring_of;          // ring slot of vertex i's moved position, or -1
    const int *gstart;           // groups: vertices [gstart[g], gstart[g + 1])
    double W;
    double2 *pos;
    const double2 *rep;          // repulsion minus the late attraction (k_late_attract)
    double2 *pre;                // pre_i written by the helpers
    double2 *sink;               // stores of the threads past a group's end
    int *ready;                  // ready[g] == gen: pre of group g is in memory
    int *done;                   // groups [0, *done) of this launch are final in memory; done[1]: helped
    int gen;
    int n, ng, D, rcap;
    double step;
    double *eterm;
    const int *status;
    long long ticks;             // a helper's longest wait
};

constexpr size_t LDS_MAX = 160 * 1024 - 64;  // LDS of one workgroup on gfx950, less the helpers' static word

// LDS of the walker: the ring, then per stage slot the own positions, pre_i / rep'_i, recent
// ring slots, (weights), own ring slots, then the slot headers and the decision ring.
size_t lag_stage_bytes(bool has_w, int T) {
    return (size_t)LAG_K * T * (16 + 16 + 16 + (has_w ? 16 : 0) + 4) + 8 * LAG_K + sizeof(int) * (LAG_DEC + 4);
}
int lag_rcap(bool has_w, int T) { return (int)((LDS_MAX - lag_stage_bytes(has_w, T)) / 16); }
size_t lag_lds_bytes(bool has_w, int T, int rcap) { return 16 * (size_t)rcap + lag_stage_bytes(has_w, T); }

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
    for (int g = hid; g < a.ng; g += nhelp) {
        __syncthreads();  // ok of the previous group has been read
        const int b0 = a.gstart[g], gc = a.gstart[g + 1] - b0;
        const int i = b0 + min(t, gc - 1);
        // what does not change during the walk, before the wait; group g has not moved (the
        // walker decides whether to use this group's sums several groups before it)
        const double2 v = a.pos[i];
        const double2 r = a.rep[i];
        const int e0 = a.erp[i], e1 = a.esp[i];
        int c[LAG_HS];
        float w[LAG_HS];
        settled_ids<HAS_W>(a, e0, e1, c, w);
        const int need = g - a.D + 1;  // groups [0, need) must be final
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
            if (ok < 0) return;  // the walker forms the remaining groups itself
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            // the walker reads a group's flag at least LAG_LEAD groups ahead: a group it has
            // (nearly) reached is formed by the walker
            if (ok > g - LAG_LEAD - 2) continue;
        }
        if (t < gc) {  // agent-scope stores: in memory once complete
            const double2 f = settled_sum<HAS_W>(a, e0, e1, c, w, v, r);
            __hip_atomic_store(&a.pre[i].x, f.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&a.pre[i].y, f.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // every wave's stores have completed before thread 0 flags the group
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (t == 0) __hip_atomic_store(a.ready + g, a.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <bool HAS_W, int T>
__device__ void lag_walker(const LagArgs &a, unsigned char *raw) {
    constexpr int K = LAG_K, LW = LagShape<T>::LW, CH = LagShape<T>::CH;
    const int n = a.n, ng = a.ng;
    double2 *ring = reinterpret_cast<double2 *>(raw);     // [rcap] moved positions read as recent
    double2 *sv = ring + a.rcap;                           // [K][T] staged own positions
    double2 *sp = sv + K * T;                             // [K][T] staged pre_i (flagged) or rep'_i
    int4 *so = reinterpret_cast<int4 *>(sp + K * T);      // [K][T] staged ring slots of the recent neighbours
    float4 *sw = reinterpret_cast<float4 *>(so + K * T);  // [K][T] their weights (HAS_W)
    int *sr = reinterpret_cast<int *>(HAS_W ? (void *)(sw + K * T) : (void *)sw);  // [K][T] own ring slots
    int2 *hdr = reinterpret_cast<int2 *>(sr + K * T);     // [K] the staged group's first vertex and size
    int *dec = reinterpret_cast<int *>(hdr + K);          // [LAG_DEC] flags used
    int *shelp = dec + LAG_DEC;                           // groups formed from the helpers' sums
    const int tid = threadIdx.x;
    if (tid < T) {  // ---- computing waves ----
        const int t = tid;
        int helped = 0;
        wbar();  // group 0 staged (and the decisions cleared)
        for (int g = 0; g < ng; ++g) {
            const int s = (g % K) * T + t;
            // one LDS round trip: the group's flag and header and this thread's whole stage
            int fl, ro;
            v2i hh;
            v4i w0, w1, w2, w3 = {0, 0, 0, 0};
            lds_read_stage(dec + g % LAG_DEC, hdr + g % K, sv + s, sp + s, so + s, sr + s, fl, hh, w0, w1, w2, ro);
            if (HAS_W) w3 = reinterpret_cast<const v4i *>(sw)[s];  // before the first use of w0..w2
            const int b0 = hh.x, gc = hh.y;
            const double2 v = bits_as<double2>(w0);
            double2 p = bits_as<double2>(w1);
            const int4 rs = bits_as<int4>(w2);
            const float4 rw = bits_as<float4>(w3);
            const bool flagged = __builtin_amdgcn_readfirstlane(fl) != 0;
            helped += flagged;
            // threads past the group's end compute on stale stage words and store to the sink
            const bool live = t < gc;
            const int i = b0 + t;
            if (live && !flagged) {  // the helper was late (or the group is one of the first): form pre_i
                const int e0 = a.erp[i], e1 = a.esp[i];
                int c[LAG_HS];
                float w[LAG_HS];
                settled_ids<HAS_W>(a, e0, e1, c, w);
                p = settled_sum<HAS_W>(a, e0, e1, c, w, v, p);
            }
            const int roff[LAG_RS] = {rs.x, rs.y, rs.z, rs.w};
            const float rwk[LAG_RS] = {rw.x, rw.y, rw.z, rw.w};
            // a second round trip: every ring slot this vertex may need (clamped slots)
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
            // agent-scope stores are written through to memory, so the helpers (on other XCDs)
            // read them once they have completed: no L2 write-back is needed to publish
            double2 *pd = live ? a.pos + i : a.sink + t;
            double *ed = live ? a.eterm + i : reinterpret_cast<double *>(a.sink + T) + t;
            __hip_atomic_store(&pd->x, q.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&pd->y, q.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            *ed = m2;
            if (live && ro >= 0) ring[ro] = q;
            // the stores of group g-3 are complete (three per group, in order; the control wave
            // publishes them after a barrier)
            __asm__ volatile("s_waitcnt vmcnt(9)" ::: "memory");
            wbar();  // end of group g
        }
        if (t == 0) *shelp = helped;
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        wbar();
        return;
    }
    if (tid < T + 64 * LW) {  // ---- loading waves: group g + 3 during group g ----
        const int lw = (tid - T) >> 6, lane = tid & 63;
        constexpr int OPS = HAS_W ? 5 : 4;  // loads per chunk
        const bool two = lw < CH - LW;      // this wave loads two chunks of a group (T = 512: wave 0)
        // one group into its slot; pre_i if the control wave recorded a usable flag, else rep'_i.
        // The group's bounds come from a scalar load waited for in the same asm block (gstart is
        // read-only in the kernel): as a vector load it would count among the loads in flight
        // that land() waits for, and its wait would take them all.
        auto issue = [&](int x, bool use_dec) {
            unsigned long long gg;
            {
                const int *gp = a.gstart + __builtin_amdgcn_readfirstlane(x);
                __asm__ volatile("s_load_dwordx2 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=s"(gg) : "s"(gp) : "memory");
            }
            const int b0 = (int)(unsigned)gg, gc = (int)(unsigned)(gg >> 32) - b0;
            int f = 0;
            if (use_dec) {  // an LDS read the compiler does not order after the loads in flight
                const lds_void *dp = (const lds_void *)(dec + x % LAG_DEC);
                __asm__ volatile("ds_read_b32 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(f) : "v"(dp) : "memory");
            }
#pragma unroll
            for (int c0 = 0; c0 < (CH + LW - 1) / LW; ++c0) {
                const int c = lw + c0 * LW;
                if (c >= CH) break;                      // wave-uniform
                const int slot = (x % K) * T + c * 64;  // this chunk's 64 vertices of the slot
                const int i = min(b0 + c * 64 + lane, n - 1);
                __builtin_amdgcn_global_load_lds((const void *)(a.pos + i), (lds_void *)(sv + slot), 16, 0, 0);
                __builtin_amdgcn_global_load_lds((const void *)(f ? a.pre + i : a.rep + i), (lds_void *)(sp + slot),
                                                 16, 0, 0);
                __builtin_amdgcn_global_load_lds((const void *)(a.rofs + i), (lds_void *)(so + slot), 16, 0, 0);
                __builtin_amdgcn_global_load_lds((const void *)(a.ring_of + i), (lds_void *)(sr + slot), 4, 0, 0);
                if (HAS_W)
                    __builtin_amdgcn_global_load_lds((const void *)(a.rw + i), (lds_void *)(sw + slot), 16, 0, 0);
            }
            // the header by an asm store: a compiled LDS store after the direct-to-LDS loads would
            // make the compiler wait for all of them first
            if (lw == 0 && lane == 0) {
                const unsigned long long hv = (unsigned long long)(unsigned)b0 | ((unsigned long long)(unsigned)gc << 32);
                __asm__ volatile("ds_write_b64 %0, %1" : : "v"((lds_void *)(hdr + x % K)), "v"(hv) : "memory");
            }
        };
        // the loads of group x have landed, given that groups x+1 .. last were issued after it
        // (OPS loads per chunk, one or two chunks per group)
        auto land = [&](int x, int last) {
            const int younger = max(0, last - x);
            if (younger >= 2) {
                if (two) {
                    if (OPS == 5) __asm__ volatile("s_waitcnt vmcnt(20)" ::: "memory");
                    else __asm__ volatile("s_waitcnt vmcnt(16)" ::: "memory");
                } else {
                    if (OPS == 5) __asm__ volatile("s_waitcnt vmcnt(10)" ::: "memory");
                    else __asm__ volatile("s_waitcnt vmcnt(8)" ::: "memory");
                }
            } else if (younger == 1) {
                if (two) {
                    if (OPS == 5) __asm__ volatile("s_waitcnt vmcnt(10)" ::: "memory");
                    else __asm__ volatile("s_waitcnt vmcnt(8)" ::: "memory");
                } else {
                    if (OPS == 5) __asm__ volatile("s_waitcnt vmcnt(5)" ::: "memory");
                    else __asm__ volatile("s_waitcnt vmcnt(4)" ::: "memory");
                }
            } else {
                __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
            }
        };
        // prologue: groups 0..2 unflagged (before the control wave's first barrier)
        for (int x = 0; x < K - 1 && x < ng; ++x) issue(x, false);
        land(0, min(K - 2, ng - 1));
        wbar();
        for (int g = 0; g < ng; ++g) {
            const int x = g + K - 1;  // slot (g-1) % K: the computing waves left it at the last barrier
            if (x < ng) issue(x, true);
            if (g + 1 < ng) land(g + 1, min(x, ng - 1));
            wbar();  // end of group g
        }
        wbar();
        return;
    }
    // ---- control wave ----
    const int l = tid - T - 64 * LW;
    if (l < LAG_DEC) dec[l] = 0;  // the first groups are formed by the walker
    int val = 0;
    wbar();
    for (int g = 0; g < ng; ++g) {
        if ((g & 1) == 0) {
            if (g >= 2) {  // the flags read two groups ago, of groups g - 2 + LAG_LEAD + l
                const int xx = g - 2 + LAG_LEAD + l;
                const bool good = l < 2 && xx < ng && val == a.gen;
                if (__ballot(good)) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                // group xx is staged during group xx - 3 >= g + 1, after this group's barrier
                if (l < 2 && xx < ng) dec[xx % LAG_DEC] = good ? 1 : 0;
            }
            val = 0;
            const int y = g + LAG_LEAD + l;
            if (l < 2 && y < ng) val = __hip_atomic_load(a.ready + y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // groups [0, g - 3): the computing waves completed group g - 4's stores before the last
        // barrier; they are agent-scope stores, in memory once complete, so a relaxed store
        // publishes them (a release fence would write the L2 back: 0.6 us)
        if (l == 0 && g >= 4) __hip_atomic_store(a.done, g - 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        wbar();  // end of group g
    }
    wbar();
    if (l == 0) {
        a.done[1] = *shelp;
        publish(a.done, ng);  // every group is final: the helpers may stop
    }
}

template <bool HAS_W, int T>
__global__ __launch_bounds__(LagShape<T>::WG) void k_walk_lag(LagArgs a) {
    extern __shared__ __align__(16) unsigned char lag_raw[];
    if (__builtin_amdgcn_readfirstlane(*a.status)) return;
    if (blockIdx.x == 0)
        lag_walker<HAS_W, T>(a, lag_raw);
    else
        lag_helper<HAS_W>(a, (int)blockIdx.x - 1, (int)gridDim.x - 1);
}

// The walk's groups for batches of B vertices and group width T: consecutive batches, at most
// T / B of them, a batch joining the group before it when none of its vertices has a neighbour
// in that group.  Returns the groups' first vertices, then n.
std::vector<int32_t> lag_groups(const bl_graph *g, int B, int T) {
    const int n = g->n, nb = (n + B - 1) / B, gmax = std::max(1, T / B);
    std::vector<int32_t> gs{0};
    int cur0 = 0, size = 1;  // the open group: first vertex and batches
    for (int b = 1; b < nb; ++b) {
        bool join = size < gmax;
        for (int i = b * B; join && i < std::min(n, b * B + B); ++i)
            for (int e = g->h_rowptr[i]; e < g->h_rowptr[i + 1]; ++e) {
                const int j = g->h_colids[e];
                if (j >= cur0 && j < b * B) {
                    join = false;
                    break;
                }
            }
        if (join) {
            ++size;
        } else {
            gs.push_back(b * B);
            cur0 = b * B;
            size = 1;
        }
    }
    gs.push_back(n);
    return gs;
}

// The group width for batches of B: 512 when its groups are at most four fifths as many as the
// batches, else 256 (one batch a step for B > 128).
int lag_width(const bl_graph *g, int B, std::vector<int32_t> *groups) {
    const int nb = (g->n + B - 1) / B;
    std::vector<int32_t> gs = lag_groups(g, B, 512);
    int T = 512;
    if (5 * (int64_t)(gs.size() - 1) > 4 * (int64_t)nb) {
        T = 256;
        gs = lag_groups(g, B, 256);
    }
    if (groups) *groups = std::move(gs);
    return T;
}

// Group index of every vertex.
std::vector<int32_t> group_of(const std::vector<int32_t> &gs, int n) {
    std::vector<int32_t> go(n);
    for (size_t k = 0; k + 1 < gs.size(); ++k)
        for (int i = gs[k]; i < gs[k + 1]; ++i) go[i] = (int32_t)k;
    return go;
}

// Whether the moved positions that later groups read as recent at lag D fit a ring of rcap
// slots, taken in group order: the vertices read (exported) by the groups up to D - 1 later get
// consecutive slots (mod rcap), and the slots from group g - D + 1's first to group g's last
// may number at most rcap.  Fills ring_of (slot or -1) when given.
bool lag_ring_fits(const bl_graph *g, const std::vector<int32_t> &gs, const std::vector<int32_t> &go, int B, int D,
                   int rcap, std::vector<int32_t> *ring_of) {
    const int n = g->n, ng = (int)gs.size() - 1;
    std::vector<unsigned char> ex(n, 0);
    for (int i = 0; i < n; ++i) {
        const int gi = go[i], b0 = (i / B) * B;
        for (int e = g->h_rowptr[i]; e < g->h_rowptr[i + 1]; ++e) {
            const int j = g->h_colids[e];
            if (j < b0 && go[j] > gi - D) ex[j] = 1;  // recent
        }
    }
    std::vector<int64_t> bofs(ng + 1, 0);
    for (int k = 0; k < ng; ++k) {
        int64_t c = 0;
        for (int i = gs[k]; i < gs[k + 1]; ++i) c += ex[i];
        bofs[k + 1] = bofs[k] + c;
    }
    for (int k = 0; k < ng; ++k)
        if (bofs[k + 1] - bofs[std::max(0, k - D + 1)] > rcap) return false;
    if (ring_of) {
        ring_of->assign(n, -1);
        for (int k = 0; k < ng; ++k) {
            int64_t r = bofs[k];
            for (int i = gs[k]; i < gs[k + 1]; ++i)
                if (ex[i]) (*ring_of)[i] = (int32_t)(r++ % rcap);
        }
    }
    return true;
}

}  // namespace

// Whether the lagged walk takes batches of B vertices of graph g with lag D (groups), and if so
// the groups, the settled / recent split of the early entries and the ring slots on the device
// (cached per batch size and lag).
bool lag_walk_fits(bl_graph *g, int B, int D) {
    if (g->lag_batch == B && g->lag_d == D) return g->lag_ok;
    g->lag_batch = B;
    g->lag_d = D;
    g->lag_ok = false;
    const int n = g->n;
    if (!(B <= LAG_BMAX && D >= 4 && D <= LAG_DMAX && g->nnz > 0 && g->max_row <= LAG_MAXDEG &&
          (int)g->h_rowptr.size() == n + 1 && (int64_t)g->h_colids.size() == g->nnz))
        return false;
    std::vector<int32_t> gs;
    const int T = lag_width(g, B, &gs);
    const int ng = (int)gs.size() - 1;
    if (!(ng >= 2 * D && ng >= LAG_LEAD + 4)) return false;
    const bool hw = !g->h_w.empty();
    const int rcap = lag_rcap(hw, T);
    const std::vector<int32_t> go = group_of(gs, n);
    std::vector<int32_t> ring_of;
    if (!lag_ring_fits(g, gs, go, B, D, rcap, &ring_of)) return false;
    std::vector<int32_t> erp(n + 1, 0), esp(n, 0), eci;
    std::vector<float> ew;
    std::vector<int32_t> rofs(4 * (size_t)n, -1);
    std::vector<float> rw(hw ? 4 * (size_t)n : 0, 0.0f);
    eci.reserve(g->nnz / 2 + 1);
    for (int i = 0; i < n; ++i) {
        const int gi = go[i], b0 = (i / B) * B;
        int nr = 0;
        for (int pass = 0; pass < 2; ++pass) {  // settled first, then recent, each in CSR order
            if (pass == 1) esp[i] = (int32_t)eci.size();
            for (int e = g->h_rowptr[i]; e < g->h_rowptr[i + 1]; ++e) {
                const int j = g->h_colids[e];
                if (j >= b0) continue;  // late: summed by k_late_attract
                const bool settled = go[j] <= gi - D;
                if (settled != (pass == 0)) continue;
                eci.push_back(j);
                if (hw) ew.push_back(g->h_w[e]);
                if (!settled) {
                    if (nr == LAG_RS) return false;  // more recent neighbours than the stage holds
                    rofs[4 * (size_t)i + nr] = ring_of[j];
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
        g->lag_ready.ensure(4 * (size_t)ng + 256) || g->lag_ring_of.ensure(4 * (size_t)n) ||
        g->lag_gstart.ensure(4 * (size_t)(ng + 1)) || g->lag_sink.ensure(16 * 2 * 512))
        return false;
    if (hw && (g->lag_ew.ensure(4 * ew.size()) || g->lag_rw.ensure(16 * (size_t)n))) return false;
    if (hipMemcpy(g->lag_erp.p, erp.data(), 4 * (size_t)(n + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lag_esp.p, esp.data(), 4 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lag_eci.p, eci.data(), 4 * eci.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lag_rofs.p, rofs.data(), 16 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lag_ring_of.p, ring_of.data(), 4 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->lag_gstart.p, gs.data(), 4 * (size_t)(ng + 1), hipMemcpyHostToDevice) != hipSuccess)
        return false;
    if (hw && (hipMemcpy(g->lag_ew.p, ew.data(), 4 * ew.size(), hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(g->lag_rw.p, rw.data(), 16 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess))
        return false;
    // flags hold launch generations (never 0); done (and the helped count) after the flags
    if (hipMemset(g->lag_ready.p, 0, 4 * (size_t)ng + 256) != hipSuccess) return false;
    g->lag_gen = 0;
    g->lag_width = T;
    g->lag_ng = ng;
    g->lag_rcap = rcap;
    g->lag_ok = true;
    return true;
}

// The lag the walk takes by default for batches of B vertices (0: the walk does not fit), in
// groups: at most LAG_DMAX and half the groups; at most the smallest distance, over the vertices
// with five or more neighbours in earlier batches, from a vertex's group to the group of its
// fifth-nearest such neighbour (so no vertex has more than LAG_RS = 4 recent neighbours); and
// the largest below those whose recent neighbours' moved positions fit the ring.  Cached per
// batch size.
int lag_walk_choose(bl_graph *g, int B) {
    if (g->lag_auto_batch == B) return g->lag_auto_d;
    g->lag_auto_batch = B;
    g->lag_auto_d = 0;
    const int n = g->n;
    if (B > LAG_BMAX || (int)g->h_rowptr.size() != n + 1 || g->max_row > LAG_MAXDEG) return 0;
    std::vector<int32_t> gs;
    const int T = lag_width(g, B, &gs);
    const int ng = (int)gs.size() - 1;
    const std::vector<int32_t> go = group_of(gs, n);
    int D = std::min(LAG_DMAX, ng / 2);
    static_assert(LAG_RS == 4, "the fifth-nearest gap bounds the lag");
    for (int i = 0; i < n && D >= 4; ++i) {
        const int gi = go[i], b0 = (i / B) * B;
        int gap[LAG_RS + 1], cnt = 0;  // the five smallest gaps gi - gj >= 1, ascending
        for (int e = g->h_rowptr[i]; e < g->h_rowptr[i + 1]; ++e) {
            const int j = g->h_colids[e];
            if (j >= b0) continue;  // late
            const int k = gi - go[j];
            if (cnt == LAG_RS + 1 && k >= gap[LAG_RS]) continue;
            int m = cnt < LAG_RS + 1 ? cnt++ : LAG_RS;
            while (m > 0 && gap[m - 1] > k) gap[m] = gap[m - 1], --m;
            gap[m] = k;
        }
        if (cnt == LAG_RS + 1) D = std::min(D, gap[LAG_RS]);  // recent: gaps 1 .. D - 1
    }
    const int rcap = lag_rcap(!g->h_w.empty(), T);
    if (D >= 4 && !lag_ring_fits(g, gs, go, B, D, rcap, nullptr)) {  // the largest D that fits
        int lo = 3, hi = D;  // fits(lo) or lo == 3; !fits(hi)
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (lag_ring_fits(g, gs, go, B, mid, rcap, nullptr)) lo = mid;
            else hi = mid;
        }
        D = lo;
    }
    g->lag_auto_d = D >= 4 ? D : 0;
    return g->lag_auto_d;
}

// the count of groups the last lagged walk staged with the helpers' sums
int *lag_walk_helped(bl_graph *g) { return g->lag_ready.as<int>() + g->lag_ng + 33; }

// One iteration's walk (after k_late_attract has folded the late attraction into the repulsion).
int launch_walk_lag(bl_graph *g, double2 *pos, const bl_params *p, double step, int helpers, hipStream_t s) {
    LagArgs a;
    a.erp = g->lag_erp.as<int>();
    a.esp = g->lag_esp.as<int>();
    a.eci = g->lag_eci.as<int>();
    a.ew = g->has_w ? g->lag_ew.as<float>() : nullptr;
    a.rofs = g->lag_rofs.as<int4>();
    a.rw = g->has_w ? g->lag_rw.as<float4>() : nullptr;
    a.ring_of = g->lag_ring_of.as<int>();
    a.gstart = g->lag_gstart.as<int>();
    a.W = p->weight;
    a.pos = pos;
    a.rep = g->bh.rep.as<double2>();
    a.pre = g->lag_pre.as<double2>();
    a.sink = g->lag_sink.as<double2>();
    a.ng = g->lag_ng;
    a.ready = g->lag_ready.as<int>();
    a.done = a.ready + a.ng + 32;  // its own cache line
    a.gen = ++g->lag_gen;
    if (a.gen == 0x7fffffff) return BL_E_DEVICE;  // generations exhausted (2^31 launches)
    a.n = g->n;
    a.D = g->lag_d;
    a.rcap = g->lag_rcap;
    a.step = step;
    a.eterm = g->eterm.as<double>();
    a.status = g->bh.misc.as<int>();
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device) != hipSuccess || khz <= 0)
        khz = 100000;
    a.ticks = (long long)khz * 1000;  // 1 s
    BL_HIP(hipMemsetAsync(a.done, 0, 2 * sizeof(int), s));
    const int T = g->lag_width;
    const size_t lds = lag_lds_bytes(g->has_w, T, a.rcap);
    if (T == 512) {
        if (g->has_w)
            hipLaunchKernelGGL((k_walk_lag<true, 512>), dim3(1 + helpers), dim3(LagShape<512>::WG), lds, s, a);
        else
            hipLaunchKernelGGL((k_walk_lag<false, 512>), dim3(1 + helpers), dim3(LagShape<512>::WG), lds, s, a);
    } else {
        if (g->has_w)
            hipLaunchKernelGGL((k_walk_lag<true, 256>), dim3(1 + helpers), dim3(LagShape<256>::WG), lds, s, a);
        else
            hipLaunchKernelGGL((k_walk_lag<false, 256>), dim3(1 + helpers), dim3(LagShape<256>::WG), lds, s, a);
    }
    BL_HIP(hipGetLastError());
    return BL_OK;
}

int lag_walk_prepare(bool has_w) {
    const int lb = (int)LDS_MAX;  // the most any plan takes
    if (has_w) {
        BL_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_walk_lag<true, 256>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, lb));
        BL_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_walk_lag<true, 512>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, lb));
    } else {
        BL_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_walk_lag<false, 256>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, lb));
        BL_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_walk_lag<false, 512>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, lb));
    }
    return BL_OK;
}

}  // namespace bl
