// The wide lagged batch walk of the Barnes-Hut family (algorithms.cpp:1050-1066: per batch, the
// attraction of every CSR entry with live positions minus the repulsion, then the sequential
// move), for batches of up to 1024 vertices on graphs with long rows (hubs), where the lagged
// walk's one-thread-per-vertex stage (walk_lag.hip) does not fit.
//
// Batch b's attraction reads the positions its neighbours have at that moment: a neighbour in
// an earlier batch has moved, one in batch b or later has not.  With a lag D_b per batch, each
// row's entries are split into
//   near   neighbours in batches b - D_b + 1 .. b - 1, summed by the walker, and
//   far    the rest: "settled" ones in batches <= b - D_b, final once batch b - D_b has moved,
//          and "late" ones in batch b or later, still at their iteration-start positions,
// and   f_i = pre_i + sum_near attraction,   pre_i = sum_far attraction - rep_i.
//
// Workgroups 1..H, the helpers, form pre_i chunk by chunk (a chunk: whole rows of one batch
// with at most WD_CH far entries between them, or one longer row), taking chunks k, k + H, ...
// in batch order: a helper loads the chunk's column ids and every position it addresses, waits
// until the walker has published batches [0, b - D_b + 1), loads the settled positions again,
// sums per row with the entries split evenly over its threads, stores pre_i and counts the
// chunk done for its batch.
//
// Workgroup 0, the walker, takes the batches in order; thread t owns vertex t of each batch.
// What a batch needs (own position, ring slot, the thread's share of the batch's near entries)
// is loaded into registers three batches ahead, and pre_i two batches ahead, once every chunk
// of that batch has been counted (the walker waits there for a late helper).  Near neighbours
// are read from an LDS ring holding the moved positions of the vertices that a later batch
// reads as near; the host assigns ring slots, and limits D_b, so that no slot is overwritten
// while a batch within the lag still reads it.  The per-row near sums use the same entry split
// as the helpers.
//
// Every workgroup is resident (a cooperative launch), so the walker may wait for the helpers
// and the helpers for the walker: the walker publishes batches [0, x - 4) before it waits for
// batch x, and a helper of batch x waits for at most batches [0, x - D_x + 1) with D_x >= 5
// (or none, for the first batches).  Every wait has a deadline and an error flag that ends
// the other waits.  Five batches of at most 1024 vertices fit the ring, so every graph with
// B <= 1024 takes the walk.
//
// What one workgroup writes for another (moved positions, pre_i) is stored through to memory
// (agent-scope stores), so a relaxed flag or count store after the stores have completed
// publishes them, and the reader acquires after seeing it.  The arithmetic is FP64: the
// attraction terms and the move use the hardware reciprocal square root refined by one Newton
// step (attr_add, rsqrt2).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "bh_common.h"

namespace bl {
namespace {

constexpr int WD_T = 1024;           // threads per workgroup; walker thread t owns vertex t of a batch
constexpr int WD_KN = 2;             // near entries of a walker thread loaded ahead
constexpr int WD_KH = 8;             // far entries of a helper thread loaded together
constexpr int WD_DMIN = 5;           // smallest lag of a batch past the fourth (the wait rule above)
constexpr int WD_DMAX = 16;          // default largest lag
constexpr int WD_CH = 8192;          // far entries of a helper chunk (a longer row is a chunk alone)
constexpr int WD_MASK = 0x7fffffff;  // far column id; the sign bit marks a settled neighbour
constexpr size_t WD_LDS = 160 * 1024;
// LDS: vb, out, S, tail (double2[WD_T]); wsum (double2[16]); hrow, trow (int[WD_T]);
// lrp (int[WD_T + 64]); wkf, wkl (int[16]); ctl (int[32]); then the walker's ring
constexpr size_t WD_FIXED = 4 * 16 * WD_T + 16 * 16 + 2 * 4 * WD_T + 4 * (WD_T + 64) + 4 * 64;
static_assert(WD_FIXED % 16 == 0, "the ring is 16-byte aligned");
constexpr int WD_CAP = (int)((WD_LDS - WD_FIXED) / 16);  // ring slots

struct WideArgs {
    const int4 *npa, *npb;  // [nb][WD_T] walker thread t's near run: {E0, 0, fr, code0}, {code1, w0, w1, E}
    const int *nrp;         // [n] vertex i's first near entry, relative to its batch's first (E0)
    const int *near_code;   // ring slot of each near entry's neighbour
    const float *near_w;
    const int *ring_of;     // [n] ring slot of vertex i's moved position, or -1
    const int *nch;         // [nb] helper chunks of batch b
    const int4 *chunks;     // chunk k: {batch, first row, rows, batches [0, w) final first}
    const int *cfr;         // [chunk][WD_T] row (in the chunk) of thread t's first far entry
    const int *far_rp;      // [n + 1] far entries of vertex i (+ one unused entry at the end)
    const int *far_ci;
    const float *far_w;
    double W;
    double2 *pos;
    const double2 *rep;
    double2 *pre;           // [nb][BP] by batch and vertex in the batch
    double2 *sink;          // stores of the last batch's threads past its end
    int *cnt;               // [nb] chunks counted, summed over launches (gen * nch[b]: complete)
    int *done;              // [0] batches [0, done) final; [1] error; [2] batches the walker waited for
    int gen, nchunks;
    int n, B, nb, BP;       // BP: B rounded up to 8 (128-byte rows of pre)
    double step;
    double *eterm;
    const int *status;
    long long ticks;        // longest wait
};

struct WLds {
    double2 *vb, *out, *S, *tail, *wsum, *ring;
    int *hrow, *trow, *lrp, *wkf, *wkl, *ctl;
};

__device__ __forceinline__ WLds wlds(unsigned char *raw) {
    WLds L;
    double2 *d = reinterpret_cast<double2 *>(raw);
    L.vb = d;
    L.out = d + WD_T;
    L.S = d + 2 * WD_T;
    L.tail = d + 3 * WD_T;
    L.wsum = d + 4 * WD_T;
    int *ip = reinterpret_cast<int *>(L.wsum + 16);
    L.hrow = ip;
    L.trow = ip + WD_T;
    L.lrp = ip + 2 * WD_T;
    L.wkf = L.lrp + WD_T + 64;
    L.wkl = L.wkf + 16;
    L.ctl = L.wkl + 16;
    L.ring = reinterpret_cast<double2 *>(raw + WD_FIXED);
    return L;
}

// Acquire after a relaxed load has observed a count: this CU's and this XCD's vector caches
// drop their lines, so later loads read what other XCDs stored through to memory.  An acquire
// fence would first wait for every load and store of the wave in flight: the walker's loads of
// the next batches (acquired before) and its stores.  Only lines of batches not yet acquired
// may still be filled by a load in flight, and no cache line holds two batches' sums (pre rows
// are padded per batch to 128 bytes), so such a line is dropped by its own batch's acquire.
__device__ __forceinline__ void acquire_inv() { __asm__ volatile("buffer_inv sc1" ::: "memory"); }

// Barrier for LDS data: this thread's LDS accesses complete, then s_barrier (a workgroup
// barrier would also wait for the memory loads in flight).
__device__ __forceinline__ void wbar() {
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

// Per-row sums of a list of E entries: row r < nr holds the entries [lrp[r], lrp[r + 1]) (L.lrp,
// local offsets), thread t the run [t c, min(E, t c + c)), whose first entry lies in row fr.
// Within a run a row's terms are added in entry order.  A row inside a run is stored by it; a
// row that continues past a run's end leaves its partial there (tail), and the runs it reaches
// contribute their first-row partials, which a segmented scan over the threads (within each
// wave, then over the 16 wave sums) adds up; the thread holding the row's last entry stores
// tail + scan.  The sums depend only on the list and c.  The run is taken in groups of KP
// entries whose operands reload(k0) loads (the first group is loaded by the caller);
// term(g, k, v_r, ax, ay) adds the g-th entry of the group, entry k of row r, whose own
// position is v_r.  Out: L.out[r] for the rows r < nr with entries (the others are not
// written).  Two barriers.
template <int KP, class Reload, class Term>
__device__ __forceinline__ void seg_sum(const WLds &L, int E, int c, int nr, int fr, Reload reload, Term term) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int s = min(E, t * c), e = min(E, s + c);
    int hr = -1, tr = -1;                           // first row (the scan's key), continuing last row
    double hx = 0.0, hy = 0.0, tx = 0.0, ty = 0.0;  // their partials (scalars: they stay in
                                                    // registers through the lambdas)
    if (s < e) {
        int r = fr;
        bool first = true;
        int rend = L.lrp[r + 1];
        double2 vr = L.vb[r];
        double ax = 0.0, ay = 0.0;
        auto group = [&](int k0) __attribute__((always_inline)) {
#pragma unroll
            for (int g = 0; g < KP; ++g) {
                const int k = k0 + g;
                if (k < e) {
                    while (k >= rend) {  // row r ends inside this run
                        if (first) {
                            hx = ax;
                            hy = ay;
                            first = false;
                        } else {
                            L.out[r] = make_double2(ax, ay);
                        }
                        ++r;
                        rend = L.lrp[r + 1];
                        vr = L.vb[r];
                        ax = ay = 0.0;
                    }
                    term(g, k, vr, ax, ay);
                }
            }
        };
        // the first group outside the loop: a loop whose operands may come from loads of its
        // own iterations makes the compiler wait for the loads in flight where it is entered
        group(s);
        for (int k0 = s + KP; k0 < e; k0 += KP) {
            reload(k0);
            group(k0);
        }
        hr = fr;
        if (first) {
            hx = ax;
            hy = ay;
        } else if (rend <= e) {
            L.out[r] = make_double2(ax, ay);
        } else {
            tr = r;
            tx = ax;
            ty = ay;
        }
    }
    L.trow[t] = tr;
    L.tail[t] = make_double2(tx, ty);
    // segmented inclusive scan of the first-row partials (runs of equal keys are contiguous)
    double2 v = make_double2(hx, hy);
    seg_step<0x111, 0xF>(v, hr);  // row_shr:1
    seg_step<0x112, 0xF>(v, hr);  // row_shr:2
    seg_step<0x114, 0xF>(v, hr);  // row_shr:4
    seg_step<0x118, 0xF>(v, hr);  // row_shr:8
    seg_step<0x142, 0xA>(v, hr);  // row_bcast:15 into rows 1 and 3
    seg_step<0x143, 0xC>(v, hr);  // row_bcast:31 into rows 2 and 3
    if (lane == 63) {
        L.wsum[wv] = v;
        L.wkl[wv] = hr;
    }
    if (lane == 0) L.wkf[wv] = hr;
    wbar();
    // the sum carried into this wave's first run from earlier waves: lane p < 16 takes wave p;
    // a chain of waves continues a run while a wave holds one key only and it equals the key
    // before it; within a chain (numbered by a ballot of the chain starts), an inclusive scan
    // of the wave sums gives the run's sum up to wave p
    {
        const int p = lane & 15;
        double2 ws = L.wsum[p];
        const int kf = L.wkf[p], kl = L.wkl[p];
        const int kprev = __builtin_amdgcn_update_dpp(-2, kl, 0x111, 0xF, 0xF, false);  // wave p - 1's last key
        const bool start = !(p > 0 && kf == kl && kprev == kl);
        const unsigned long long m = __ballot(start && lane < 16);
        const int chain = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0)) +
                          (start ? 1 : 0);
        seg_step<0x111, 0xF>(ws, chain);
        seg_step<0x112, 0xF>(ws, chain);
        seg_step<0x114, 0xF>(ws, chain);
        seg_step<0x118, 0xF>(ws, chain);
        if (wv > 0 && hr >= 0) {
            const int klq = __builtin_amdgcn_readlane(kl, wv - 1);
            const int kfq = __builtin_amdgcn_readlane(kf, wv);
            if (klq == kfq && hr == kfq) {
                const double cx = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(ws.x), wv - 1),
                                                   __builtin_amdgcn_readlane(__double2loint(ws.x), wv - 1));
                const double cy = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(ws.y), wv - 1),
                                                   __builtin_amdgcn_readlane(__double2loint(ws.y), wv - 1));
                v.x = cx + v.x;
                v.y = cy + v.y;
            }
        }
    }
    // the thread holding the last entry of its first row stores that row: the tail partial of
    // the run where the row began (if it began inside a run), plus the scan
    if (hr >= 0 && L.lrp[hr + 1] <= e) {
        const int ta = L.lrp[hr] / c;
        double2 o = v;
        if (ta != t && L.trow[ta] == hr) o = add2(L.tail[ta], v);
        L.out[hr] = o;
    }
    wbar();
}

template <bool HAS_W>
__device__ void wide_walker(const WideArgs &a, const WLds &L) {
    const int t = threadIdx.x;
    const int n = a.n, B = a.B, nb = a.nb;
    // batch x's registers but pre_i, loaded three batches ahead; pre_i is loaded two batches
    // ahead, once every chunk of the batch has been counted
    struct Regs {
        double2 v;
        int ro, nrp;
        int4 pa, pb;
    };
    // batch x's chunk count and number of chunks (thread 0 reads them), loaded a batch before
    // they are read and before the loads issued in between, so that reading them waits for
    // nothing issued later; every thread loads them (one coalesced access), so that all
    // threads issue the same loads and stores and the compiler's waits for the registers count
    // exactly what each thread issued
    int fcnt = 0, fnch = 0, waited = 0;
    auto flag_load = [&](int x) {
        const int xc = min(x, nb - 1);
        fcnt = __hip_atomic_load(a.cnt + xc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fnch = __hip_atomic_load(a.nch + xc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __asm__ volatile("" ::: "memory");
    };
    // every chunk of batch x counted (then acquire), or false: stop (error flag or deadline);
    // ends with a barrier, which is also the end of the batch before
    auto flag_wait = [&](int x) -> bool {
        if (t == 0) {
            int ok = 1;
            if (x < nb) {
                const int target = a.gen * fnch;
                if (fcnt < target) {
                    ++waited;
                    const long long t0 = wall_clock64();
                    while (__hip_atomic_load(a.cnt + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                        if (__hip_atomic_load(a.done + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                            wall_clock64() - t0 > a.ticks) {
                            ok = 0;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                if (ok)
                    acquire_inv();
                else
                    __hip_atomic_store(a.done + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            L.ctl[0] = ok;
        }
        wbar();
        return __builtin_amdgcn_readfirstlane(L.ctl[0]) != 0;
    };
    // batch x's vertex of this thread (past the last batch: the last batch's, never used)
    auto vtx = [&](int x, int &xb, int &b0) {
        xb = min(x, nb - 1);
        b0 = xb * B;
        return min(b0 + t, min(n, b0 + B) - 1);
    };
    auto load = [&](Regs &R, int x) {
        int xb, b0;
        const int i = vtx(x, xb, b0);
        R.v = a.pos[i];
        R.ro = a.ring_of[i];
        R.nrp = a.nrp[i];
        R.pa = a.npa[(size_t)xb * WD_T + t];
        R.pb = a.npb[(size_t)xb * WD_T + t];
    };
    auto load_pre = [&](double2 &P, int x) {
        int xb, b0;
        const int i = vtx(x, xb, b0);
        P = a.pre[(size_t)xb * a.BP + (i - b0)];
    };
    auto step = [&](int b, const Regs &R, double2 P) {
        const int b0 = b * B, bc = min(B, n - b0);
        const bool live = t < bc;
        const int E = __builtin_amdgcn_readfirstlane(R.pb.w);
        L.vb[t] = R.v;
        L.lrp[t] = live ? R.nrp : E;
        if (t == 0) L.lrp[WD_T] = E;
        wbar();
        // every thread has read registers loaded after the barrier that followed its stores of
        // batch b - 3, so those stores are complete (vector memory operations complete in issue
        // order): batches [0, b - 2) are final
        if (t == 0 && b >= 3) __hip_atomic_store(a.done, b - 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        double2 f = P;
        if (E > 0) {
            const int c = (E + WD_T - 1) / WD_T;
            const int E0 = R.pa.x;
            int code[WD_KN] = {R.pa.w, R.pb.x};
            float wk[WD_KN] = {__int_as_float(R.pb.y), __int_as_float(R.pb.z)};
            auto reload = [&](int k0) {
#pragma unroll
                for (int g = 0; g < WD_KN; ++g) {
                    const int kk = E0 + min(k0 + g, E - 1);
                    code[g] = a.near_code[kk];
                    if (HAS_W) wk[g] = a.near_w[kk];
                }
            };
            auto term = [&](int g, int, double2 vr, double &ax, double &ay) {
                attr_add(L.ring[code[g]], vr, HAS_W ? a.W * (double)wk[g] : a.W, ax, ay);
            };
            seg_sum<WD_KN>(L, E, c, bc, R.pa.z, reload, term);
            if (L.lrp[t + 1] > L.lrp[t]) {  // rows without near entries: out is not written
                const double2 o = L.out[t];
                f.x += o.x;
                f.y += o.y;
            }
        }
        // x_i += (F/|F|) * STEP (:1062-1066), the unit vector F * rsqrt2(|F|^2)
        const double m2 = f.x * f.x + f.y * f.y;
        const double rs = rsqrt2(m2) * a.step;
        double2 q;
        q.x = fma(f.x, rs, R.v.x);
        q.y = fma(f.y, rs, R.v.y);
        // three stores per thread and batch in every case (threads past the batch's end store
        // to the sink)
        double2 *pd = live ? a.pos + b0 + t : a.sink + t;
        double *ed = live ? a.eterm + b0 + t : reinterpret_cast<double *>(a.sink + WD_T) + t;
        store_through(pd, q);
        *ed = m2;
        if (live && R.ro >= 0) L.ring[R.ro] = q;
    };
    // the end of batch x - 2: batch x's sums, the next flags, pre_i of x and the rest of x + 1
    Regs R0, R1, R2;
    double2 P0, P1, P2;
    flag_load(0);
    if (!flag_wait(0)) return;
    flag_load(1);
    load_pre(P0, 0);
    load(R0, 0);
    load(R1, 1);
    if (!flag_wait(1)) return;
    flag_load(2);
    load_pre(P1, 1);
    load(R2, 2);
    for (int b = 0; b < nb; b += 3) {
        step(b, R0, P0);
        if (!flag_wait(b + 2)) return;
        flag_load(b + 3);
        load_pre(P2, b + 2);
        load(R0, b + 3);
        if (b + 1 >= nb) break;
        step(b + 1, R1, P1);
        if (!flag_wait(b + 3)) return;
        flag_load(b + 4);
        load_pre(P0, b + 3);
        load(R1, b + 4);
        if (b + 2 >= nb) break;
        step(b + 2, R2, P2);
        if (!flag_wait(b + 4)) return;
        flag_load(b + 5);
        load_pre(P1, b + 4);
        load(R2, b + 5);
    }
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    wbar();
    if (t == 0) {
        __hip_atomic_store(a.done + 2, waited, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(a.done, nb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <bool HAS_W>
__device__ void wide_helper(const WideArgs &a, const WLds &L, int hid, int nh) {
    const int t = threadIdx.x;
    for (int k = hid; k < a.nchunks; k += nh) {
        const int4 ck = a.chunks[k];
        const int row0 = ck.y, nr = ck.z, need = ck.w;
        const int E0 = a.far_rp[row0], E = a.far_rp[row0 + nr] - E0;
        const int c = max(1, (E + WD_T - 1) / WD_T);
        const int s = min(E, t * c), e = min(E, s + c);
        const int fr = a.cfr[(size_t)k * WD_T + t];
        const int ir = row0 + min(t, nr - 1);
        const double2 v = a.pos[ir];
        const double2 rr = a.rep[ir];
        const int lo = a.far_rp[ir] - E0;
        int id[WD_KH];
        float wf[WD_KH];
        double2 pj[WD_KH];
        // column ids and weights of the entries k0 .. k0 + WD_KH - 1 of the run (clamped; an
        // empty chunk at the end of the list reads the unused last entry), then the positions
        auto load_group = [&](int k0) {
#pragma unroll
            for (int g = 0; g < WD_KH; ++g) {
                const int kk = E0 + max(0, min(k0 + g, e - 1));
                id[g] = a.far_ci[kk];
                wf[g] = HAS_W ? a.far_w[kk] : 1.0f;
            }
#pragma unroll
            for (int g = 0; g < WD_KH; ++g) pj[g] = a.pos[id[g] & WD_MASK];
        };
        load_group(s);  // late neighbours' positions are final; settled ones are loaded again
        if (need > 0) {
            if (t == 0) {
                int ok = 1;
                const long long t0 = wall_clock64();
                while (__hip_atomic_load(a.done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                    if (__hip_atomic_load(a.done + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                        wall_clock64() - t0 > a.ticks) {
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                if (!ok) __hip_atomic_store(a.done + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                L.ctl[1] = ok;
            }
            wbar();
            if (!__builtin_amdgcn_readfirstlane(L.ctl[1])) return;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#pragma unroll
            for (int g = 0; g < WD_KH; ++g)
                if (id[g] < 0) pj[g] = a.pos[id[g] & WD_MASK];
        }
        L.vb[t] = v;
        L.lrp[t] = t < nr ? lo : E;
        if (t == 0) L.lrp[WD_T] = E;
        wbar();
        auto term = [&](int g, int, double2 vr, double &ax, double &ay) {
            attr_add(pj[g], vr, HAS_W ? a.W * (double)wf[g] : a.W, ax, ay);
        };
        seg_sum<WD_KH>(L, E, c, nr, fr, load_group, term);
        if (t < nr) {
            const double2 o = L.lrp[t + 1] > L.lrp[t] ? L.out[t] : make_double2(0.0, 0.0);
            store_through(a.pre + (size_t)ck.x * a.BP + (row0 - ck.x * a.B) + t, make_double2(o.x - rr.x, o.y - rr.y));
        }
        // every wave's stores have completed before thread 0 counts the chunk
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        wbar();
        if (t == 0) __hip_atomic_fetch_add(a.cnt + ck.x, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <bool HAS_W>
__global__ __launch_bounds__(WD_T) void k_walk_wide(WideArgs a) {
    extern __shared__ __align__(16) unsigned char wd_raw[];
    if (__builtin_amdgcn_readfirstlane(*a.status)) return;
    const WLds L = wlds(wd_raw);
    if (blockIdx.x == 0)
        wide_walker<HAS_W>(a, L);
    else
        wide_helper<HAS_W>(a, L, (int)blockIdx.x - 1, (int)gridDim.x - 1);
}

template <class T>
int upload(bl::DevBuf &d, const std::vector<T> &h) {
    const size_t bytes = sizeof(T) * std::max<size_t>(h.size(), 1);
    if (d.ensure(bytes)) return BL_E_NOMEM;
    if (!h.empty() && hipMemcpy(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
        return BL_E_DEVICE;
    return BL_OK;
}

// Thread t's run of a list of E entries split into runs of c: its first entry's row among the
// rows whose (local) offsets are off[0..nr] (0 when the run is empty).
void first_rows(const int32_t *off, int nr, int E, int c, int *out) {
    int r = 0;
    for (int t = 0; t < WD_T; ++t) {
        const int s = std::min(E, t * c);
        if (s >= E) {
            out[t] = 0;
            continue;
        }
        while (r < nr - 1 && off[r + 1] <= s) ++r;
        out[t] = r;
    }
}

}  // namespace

// The plan of the wide walk for batches of B vertices (cached per batch size and largest lag):
//   1. the lag D_b of each batch: the largest lag (BL_WIDE_D, default WD_DMAX), reduced where
//      the ring cannot hold the moved positions the batch reads as near: the vertices a later
//      batch reads as near ("exported") take consecutive ring slots in batch order (slot =
//      running count mod WD_CAP), and batch b's near reads reach back to batch b - D_b + 1, so
//      the slots from that batch's first to batch b's last must number at most WD_CAP.
//      Reducing a lag only removes exports, so a few passes settle it.  A batch whose lag
//      falls below min(WD_DMIN, b + 1) declines the walk;
//   2. near entries (CSR order within a row) with their ring slots, the walker's per-thread
//      tables, and each exported vertex's slot;
//   3. far entries (CSR order within a row, settled ones marked), cut into chunks of whole rows.
bool wide_walk_fits(bl_graph *g, int B) {
    int dmax = WD_DMAX;
    if (const char *v = getenv("BL_WIDE_D")) dmax = atoi(v);
    if (g->wide_batch == B && g->wide_dmax == dmax) return g->wide_ok;
    g->wide_batch = B;
    g->wide_dmax = dmax;
    g->wide_ok = false;
    const int n = g->n;
    if (B < 64 || B > WD_T || dmax < WD_DMIN || n < 8 * B || g->nnz <= 0 || g->nnz >= WD_MASK ||
        (int)g->h_rowptr.size() != n + 1 || (int64_t)g->h_colids.size() != g->nnz)
        return false;
    const int nb = (n + B - 1) / B;
    const int32_t *rp = g->h_rowptr.data(), *ci = g->h_colids.data();
    const bool hw = !g->h_w.empty();
    // 1. lags and ring slots
    std::vector<int> D(nb), ex(nb), bofs(nb + 1, 0);
    for (int b = 0; b < nb; ++b) D[b] = std::min(dmax, b + 1);
    std::vector<unsigned char> exported(n);
    for (int pass = 0;; ++pass) {
        std::fill(exported.begin(), exported.end(), 0);
        for (int i = 0; i < n; ++i) {
            const int bi = i / B, lo = bi - D[bi] + 1;
            for (int e = rp[i]; e < rp[i + 1]; ++e) {
                const int bj = ci[e] / B;
                if (bj >= lo && bj < bi) exported[ci[e]] = 1;
            }
        }
        std::fill(ex.begin(), ex.end(), 0);
        for (int i = 0; i < n; ++i) ex[i / B] += exported[i];
        for (int b = 0; b < nb; ++b) bofs[b + 1] = bofs[b] + ex[b];
        bool changed = false;
        for (int b = 0; b < nb; ++b) {
            int lo = b - D[b] + 1;
            while (bofs[b] + ex[b] - bofs[lo] > WD_CAP) ++lo;
            if (b - lo + 1 < D[b]) {
                D[b] = b - lo + 1;
                changed = true;
            }
        }
        if (!changed) break;
        if (pass == 64) return false;
    }
    for (int b = 0; b < nb; ++b)
        if (D[b] < std::min(WD_DMIN, b + 1)) return false;
    std::vector<int32_t> ring_of(n, -1);
    for (int b = 0; b < nb; ++b) {
        int r = bofs[b];
        for (int i = b * B; i < std::min(n, b * B + B); ++i)
            if (exported[i]) ring_of[i] = (r++) % WD_CAP;
    }
    // 2-3. near and far entries
    std::vector<int32_t> near_rp(n + 1, 0), near_code, far_rp(n + 1, 0), far_ci;
    std::vector<float> near_w, far_w;
    near_code.reserve(g->nnz / 16 + 1);
    far_ci.reserve(g->nnz + 1);
    if (hw) far_w.reserve(g->nnz + 1);
    std::vector<unsigned char> settled_row(n, 0);
    for (int i = 0; i < n; ++i) {
        const int bi = i / B, lo = bi - D[bi] + 1;
        for (int e = rp[i]; e < rp[i + 1]; ++e) {
            const int j = ci[e], bj = j / B;
            if (bj >= lo && bj < bi) {
                near_code.push_back(ring_of[j]);
                if (hw) near_w.push_back(g->h_w[e]);
            } else {
                const bool settled = bj < lo;
                far_ci.push_back(settled ? (int32_t)((uint32_t)j | 0x80000000u) : j);
                if (hw) far_w.push_back(g->h_w[e]);
                settled_row[i] |= settled;
            }
        }
        near_rp[i + 1] = (int32_t)near_code.size();
        far_rp[i + 1] = (int32_t)far_ci.size();
    }
    // one unused entry at the end of each list: the clamped loads of empty runs read it
    near_code.push_back(0);
    far_ci.push_back(0);
    if (hw) near_w.push_back(0.0f), far_w.push_back(0.0f);
    // the walker's tables: per vertex its near offset within the batch; per batch and thread
    // {E0, 0, first row, first slot}, {second slot, weights, E}
    std::vector<int32_t> nrp(n);
    std::vector<int4> npa((size_t)nb * WD_T), npb((size_t)nb * WD_T);
    std::vector<int32_t> off(WD_T + 1), frs(WD_T);
    for (int b = 0; b < nb; ++b) {
        const int b0 = b * B, bc = std::min(B, n - b0);
        const int E0 = near_rp[b0], E = near_rp[b0 + bc] - E0;
        const int c = std::max(1, (E + WD_T - 1) / WD_T);
        for (int r = 0; r <= bc; ++r) off[r] = near_rp[b0 + r] - E0;
        for (int r = 0; r < bc; ++r) nrp[b0 + r] = off[r];
        first_rows(off.data(), bc, E, c, frs.data());
        for (int t = 0; t < WD_T; ++t) {
            const int s = std::min(E, t * c), e = std::min(E, s + c);
            const int c0 = s < e ? near_code[E0 + s] : 0, c1 = s + 1 < e ? near_code[E0 + s + 1] : 0;
            const float w0 = hw && s < e ? near_w[E0 + s] : 0.0f, w1 = hw && s + 1 < e ? near_w[E0 + s + 1] : 0.0f;
            int wb0, wb1;
            memcpy(&wb0, &w0, 4);
            memcpy(&wb1, &w1, 4);
            npa[(size_t)b * WD_T + t] = make_int4(E0, 0, frs[t], c0);
            npb[(size_t)b * WD_T + t] = make_int4(c1, wb0, wb1, E);
        }
    }
    // the helpers' chunks, in batch order; a chunk with settled entries waits for batches
    // [0, b - D_b + 1)
    std::vector<int4> chunks;
    std::vector<int32_t> nch(nb, 0), cfr;
    auto emit = [&](int b, int r0, int r1) {
        bool settled = false;
        for (int i = r0; i < r1; ++i) settled |= settled_row[i] != 0;
        const int need = settled ? std::max(0, b - D[b] + 1) : 0;
        chunks.push_back(make_int4(b, r0, r1 - r0, need));
        ++nch[b];
        const int E0 = far_rp[r0], E = far_rp[r1] - E0;
        const int c = std::max(1, (E + WD_T - 1) / WD_T);
        for (int r = 0; r <= r1 - r0; ++r) off[r] = far_rp[r0 + r] - E0;
        first_rows(off.data(), r1 - r0, E, c, frs.data());
        cfr.insert(cfr.end(), frs.begin(), frs.end());
    };
    for (int b = 0; b < nb; ++b) {
        const int b0 = b * B, b1 = std::min(n, b0 + B);
        int r0 = b0;
        int64_t acc = 0;
        for (int r = b0; r < b1; ++r) {
            const int fd = far_rp[r + 1] - far_rp[r];
            if (r > r0 && acc + fd > WD_CH) {
                emit(b, r0, r);
                r0 = r;
                acc = 0;
            }
            acc += fd;
        }
        emit(b, r0, b1);
    }
    int maxnch = 0;
    for (int b = 0; b < nb; ++b) maxnch = std::max(maxnch, nch[b]);
    if (upload(g->wide_near_rp, nrp) || upload(g->wide_near_code, near_code) || upload(g->wide_npa, npa) ||
        upload(g->wide_npb, npb) || upload(g->wide_ring_of, ring_of) || upload(g->wide_far_rp, far_rp) ||
        upload(g->wide_far_ci, far_ci) || upload(g->wide_chunks, chunks) || upload(g->wide_cfr, cfr) ||
        upload(g->wide_nch, nch))
        return false;
    if (hw && (upload(g->wide_near_w, near_w) || upload(g->wide_far_w, far_w))) return false;
    if (g->wide_pre.ensure(sizeof(double2) * (size_t)nb * ((B + 7) & ~7)) || g->wide_sink.ensure(sizeof(double2) * 2 * WD_T) ||
        g->wide_cnt.ensure(sizeof(int) * ((size_t)nb + 64)))
        return false;
    if (hipMemset(g->wide_cnt.p, 0, sizeof(int) * ((size_t)nb + 64)) != hipSuccess) return false;
    g->wide_gen = 0;
    g->wide_nchunks = (int)chunks.size();
    g->wide_maxnch = maxnch;
    g->wide_ok = true;
    return true;
}

// Clears the chunk counts (after a walk that stopped early, whose counts are incomplete).
int wide_walk_reset(bl_graph *g) {
    if (!g->wide_ok) return BL_OK;
    const int nb = (g->n + g->wide_batch - 1) / g->wide_batch;
    BL_HIP(hipMemset(g->wide_cnt.p, 0, sizeof(int) * ((size_t)nb + 64)));
    g->wide_gen = 0;
    return BL_OK;
}

int launch_walk_wide(bl_graph *g, double2 *pos, const bl_params *p, double step, int helpers, hipStream_t s) {
    const int n = g->n, B = p->batch;
    const int nb = (n + B - 1) / B;
    // counts are gen * chunks of a batch: start again from zero before they could overflow
    if ((int64_t)(g->wide_gen + 1) * std::max(1, g->wide_maxnch) >= (int64_t)1 << 30) {
        if (int rc = wide_walk_reset(g)) return rc;
    }
    WideArgs a;
    a.npa = g->wide_npa.as<int4>();
    a.npb = g->wide_npb.as<int4>();
    a.nrp = g->wide_near_rp.as<int>();
    a.near_code = g->wide_near_code.as<int>();
    a.near_w = g->has_w ? g->wide_near_w.as<float>() : nullptr;
    a.ring_of = g->wide_ring_of.as<int>();
    a.nch = g->wide_nch.as<int>();
    a.chunks = g->wide_chunks.as<int4>();
    a.cfr = g->wide_cfr.as<int>();
    a.far_rp = g->wide_far_rp.as<int>();
    a.far_ci = g->wide_far_ci.as<int>();
    a.far_w = g->has_w ? g->wide_far_w.as<float>() : nullptr;
    a.W = p->weight;
    a.pos = pos;
    a.rep = g->bh.rep.as<double2>();
    a.pre = g->wide_pre.as<double2>();
    a.sink = g->wide_sink.as<double2>();
    a.cnt = g->wide_cnt.as<int>();
    a.done = a.cnt + nb + 32;  // its own cache line
    a.gen = ++g->wide_gen;
    a.nchunks = g->wide_nchunks;
    a.n = n;
    a.B = B;
    a.nb = nb;
    a.BP = (B + 7) & ~7;
    a.step = step;
    a.eterm = g->eterm.as<double>();
    a.status = g->bh.misc.as<int>();
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device) != hipSuccess || khz <= 0)
        khz = 100000;
    a.ticks = (long long)khz * 1000;  // 1 s
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device) != hipSuccess || cus < 2)
        return BL_E_DEVICE;
    // one workgroup per CU (the LDS of each): the walker and at most cus - 1 helpers
    const int grid = 1 + std::max(1, std::min(std::min(helpers, cus - 1), g->wide_nchunks));
    BL_HIP(hipMemsetAsync(a.done, 0, 3 * sizeof(int), s));
    void *args[] = {&a};
    const void *fn = g->has_w ? reinterpret_cast<const void *>(k_walk_wide<true>)
                              : reinterpret_cast<const void *>(k_walk_wide<false>);
    const hipError_t e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(WD_T), args, WD_LDS, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return BL_E_DEVICE;
    }
    return BL_OK;
}

int wide_walk_read(bl_graph *g, int *err, int *waited, hipStream_t s) {
    const int nb = (g->n + g->wide_batch - 1) / g->wide_batch;
    const int *done = g->wide_cnt.as<int>() + nb + 32;
    BL_HIP(hipMemcpyAsync(err, done + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    BL_HIP(hipMemcpyAsync(waited, done + 2, sizeof(int), hipMemcpyDeviceToHost, s));
    return BL_OK;
}

int wide_walk_prepare(bool has_w) {
    BL_HIP(hipFuncSetAttribute(has_w ? reinterpret_cast<const void *>(k_walk_wide<true>)
                                     : reinterpret_cast<const void *>(k_walk_wide<false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)WD_LDS));
    return BL_OK;
}

}  // namespace bl
