// Internal declarations shared by the HIP translation units of libbatchlayout.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "batchlayout.h"

#define BL_HIP(call)                                                      \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) return bl::device_error(e_, #call, __LINE__); \
    } while (0)

namespace bl {

int device_error(hipError_t e, const char *what, int line);

// Attraction laws of the minibatch family (the only difference between algos 2, 3, 8).
enum Law : int { LAW_FR = 2, LAW_LINLOG = 3, LAW_FA = 8 };

// Device buffer that grows on demand and is freed with its owner.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t want);
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Barnes-Hut workspace (sized to n on first use).
struct BHWork {
    DevBuf keys_a, keys_b, idx_a, idx_b;     // radix sort ping-pong (uint64 keys, int32 ids)
    DevBuf hist;                             // per-block digit histograms
    DevBuf spos;                             // positions in Morton order (double2)
    DevBuf frame;                            // minx, miny, maxmin (+ block partials)
    DevBuf isrep, nodeid;                    // boundary representative flags and node ids
    DevBuf node_child;                       // nodes (BHNode)
    DevBuf node_cnt, node_lo;                // points of each node: Morton ranks [lo, lo + cnt)
    DevBuf scan_tmp;
    DevBuf rep;                              // repulsion result per vertex (double2)
    DevBuf misc;                             // counters / status words
    DevBuf far_cells, far_nlc;            // physically consistent far field: cell summaries
    DevBuf judge;                            // learned far field: the judge's sample, partials, flag
    DevBuf gwalk;                            // multi-workgroup walk: partials, barrier, error flag
    int gwalk_used = 0;                      // the last run launched it
    int32_t n_alloc = 0;
    void release();
};

}  // namespace bl

struct bl_graph {
    int32_t n = 0;
    int64_t nnz = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    // CSR as given (algo 4 attraction walks every entry, algorithms.cpp:1055-1058)
    bl::DevBuf rowptr, colids, w_edge;
    bool has_w = false;
    std::vector<int32_t> h_rowptr, h_colids;  // host copy of the CSR (walk selection and split)
    std::vector<float> h_w;
    int32_t max_row = 0;            // longest CSR row
    int32_t row_walk_batch = 0;     // batch size row_walk_ok and the split were computed for
    bool row_walk_ok = false;
    // the CSR split for the one-workgroup row walk at batch size row_walk_batch: "early"
    // entries (neighbour in an earlier batch, moved during the iteration before the row's
    // batch) and "late" ones (neighbour in the same or a later batch: its iteration-start
    // position), each in CSR order
    bl::DevBuf early_rp, early_ci, early_w, late_rp, late_ci, late_w;
    // the multi-workgroup walk's row lookup at batch size grid_rows_batch: for every batch,
    // workgroup and thread, the row of the thread's first entry, then the first and last
    // rows of the workgroup's share
    int32_t grid_rows_batch = 0;
    bl::DevBuf grid_rows;
    // the lagged walk (walk_lag.hip) at batch size lag_batch and lag lag_d (groups of lag_width
    // vertices at most, lag_ng of them, starting at lag_gstart): each vertex's early entries
    // split into settled (at least lag_d groups back) and recent ones, the recent ones' ring slots
    // and weights (4 per vertex), each vertex's own ring slot (lag_rcap slots), the helpers'
    // per-vertex sums and per-group flags (launch generations)
    int32_t lag_batch = 0, lag_d = 0, lag_gen = 0, lag_width = 0, lag_ng = 0, lag_rcap = 0;
    int32_t lag_auto_batch = 0, lag_auto_d = 0;  // the default lag for batch size lag_auto_batch
    bool lag_ok = false;
    bl::DevBuf lag_erp, lag_esp, lag_eci, lag_ew, lag_rofs, lag_rw, lag_pre, lag_ready, lag_ring_of, lag_gstart,
        lag_sink;
    // the wide lagged walk (walk_wide.hip) at batch size wide_batch and largest lag wide_dmax:
    // each row's entries split into near ones (the walker's, with ring slots) and far ones (the
    // helpers', in chunks of whole rows), the walker's per-thread tables, the helpers' sums and
    // per-batch chunk counts (summed over launches: wide_gen * chunks when a batch is complete)
    int32_t wide_batch = 0, wide_dmax = 0, wide_gen = 0, wide_nchunks = 0, wide_maxnch = 0, wide_cap = 0;
    bool wide_ok = false;
    bl::DevBuf wide_near_rp, wide_near_code, wide_near_w, wide_npa, wide_npb, wide_ring_of, wide_far_rp,
        wide_far_ci, wide_far_w, wide_chunks, wide_cfr, wide_nch, wide_pre, wide_cnt, wide_sink;
    // the neighbours the reference's merge pointer actually meets (algorithms.cpp:758-762),
    // as a CSR of strictly increasing column ids per row (+ their weights)
    bl::DevBuf eff_rowptr, eff_colids, eff_w;
    int32_t eff_max_deg = 0;
    // neighbour-correction pieces of the all-pairs step: vertex i's merge-pointer neighbours
    // in runs of at most CORR_PIECE, pieces [cpre[i], cpre[i+1]); cpv[p] is piece p's vertex
    bl::DevBuf cpre, cpv;
    std::vector<int32_t> h_cpre;
    int32_t corr_max_pieces = 1;
    // workspace
    bl::DevBuf pos, part, eterm, energy, scratch, guard;
    bl::BHWork bh;
    bl::DevBuf far_model;  // learned far-field parameters (bl_graph_set_far_model)
    bool has_far_model = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace bl {

// Optional per-launch timing with HIP events on the launch stream (bl_params.profile).
struct Prof {
    bool on = false;
    hipStream_t s = nullptr;
    std::vector<hipEvent_t> ev;
    std::vector<int> cls;
    size_t used = 0;
    double ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    void begin(int c) {
        if (!on) return;
        if (2 * (used + 1) > ev.size()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { on = false; return; }
            ev.push_back(a);
            ev.push_back(b);
            cls.push_back(c);
        }
        cls[used] = c;
        (void)hipEventRecord(ev[2 * used], s);
    }
    void end() {
        if (!on) return;
        (void)hipEventRecord(ev[2 * used + 1], s);
        ++used;
    }
    // Events for one launch made with hipExtLaunchKernelGGL, which records them in the
    // dispatch itself (start and end of the kernel, without the gaps of separate event
    // records).  Sets *a, *b to null when profiling is off.
    void ext(int c, hipEvent_t *a, hipEvent_t *b) {
        *a = *b = nullptr;
        if (!on) return;
        if (2 * (used + 1) > ev.size()) {
            hipEvent_t x, y;
            if (hipEventCreate(&x) != hipSuccess || hipEventCreate(&y) != hipSuccess) { on = false; return; }
            ev.push_back(x);
            ev.push_back(y);
            cls.push_back(c);
        }
        cls[used] = c;
        *a = ev[2 * used];
        *b = ev[2 * used + 1];
        ++used;
    }
    // waits for the recorded launches and folds their durations into ms/cnt
    void flush() {
        if (!on || used == 0) return;
        (void)hipEventSynchronize(ev[2 * used - 1]);
        for (size_t k = 0; k < used; ++k) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]) == hipSuccess) {
                ms[cls[k]] += t;
                cnt[cls[k]] += 1;
            }
        }
        used = 0;
    }
    void report(bl_stats *st) {
        flush();
        if (!st) return;
        for (int c = 0; c < 8; ++c) {
            st->kernel_ms[c] += ms[c];
            st->kernel_launches[c] += cnt[c];
        }
    }
    ~Prof() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

// Neighbours that the reference's merge pointer (algorithms.cpp:758-762) meets, for rows
// [i0, i1) of a CSR whose entries end at `end` (the array end of that graph): the pointer
// starts at rowptr[i], matches when j == colids[ptr] and advances only while
// ptr < rowptr[i+1]-1, so with j ascending the matched set is the longest strictly
// increasing prefix of the row; an empty row keeps the pointer on the next row's first
// entry, which then acts as one neighbour (an empty row at the end meets nothing).
// Appends to eci (column ids), ew (weights, when w is given) and sets erp[i+1] for each row
// (erp[i0] must hold the running count already).
constexpr int CORR_PIECE = 256;  // neighbour entries per correction wave

inline void effective_neighbors(const int32_t *rowptr, const int32_t *colids, const float *w, int32_t i0, int32_t i1,
                                int64_t end, std::vector<int32_t> &erp, std::vector<int32_t> &eci,
                                std::vector<float> &ew, int32_t *max_deg) {
    for (int32_t i = i0; i < i1; ++i) {
        const int64_t s = rowptr[i], e = rowptr[i + 1];
        int32_t cnt = 0;
        if (s < e) {
            int32_t last = colids[s];
            eci.push_back(last);
            if (w) ew.push_back(w[s]);
            ++cnt;
            for (int64_t k = s + 1; k < e && colids[k] > last; ++k) {
                last = colids[k];
                eci.push_back(last);
                if (w) ew.push_back(w[k]);
                ++cnt;
            }
        } else if (s < end) {
            eci.push_back(colids[s]);
            if (w) ew.push_back(w[s]);
            ++cnt;
        }
        erp[i + 1] = erp[i] + cnt;
        if (max_deg && cnt > *max_deg) *max_deg = cnt;
    }
}

// minibatch.hip
// The number of the n positions that are not finite, counted on the device into *d_count (one
// int of device memory) and returned in *h_count (waits for the stream).
int count_nonfinite(const double2 *pos, int n, int *d_count, hipStream_t s, int *h_count);
// static_ws: the workspace in the library's static device array instead of the graph's buffers
// (allocates nothing; BL_E_SCRATCH when it does not fit)
int run_minibatch(bl_graph *g, double *d_xy, const bl_params *p, int law, double decay, double step0,
                  int iters, bool converged, bl_stats *st, bool static_ws = false);
// bh.hip
// far_mode 0: the reference's unit-vector law; else a FarMode of bh_common.h
int run_bh(bl_graph *g, double *d_xy, const bl_params *p, double step0, int iters, double theta,
           bl_stats *st, int far_mode = 0);
// farfield.hip
int far_forces_host(int32_t n, const double *xy, double theta, int32_t mode, const float *model, int32_t nq,
                    const double *q, double *out);
int far_samples_host(int32_t n, const double *xy, double theta, int32_t nq, const double *q, int32_t cap,
                     float *feat, float *targ, int32_t *count);
// mlp.hip
int far_model_grad_host(const float *params, int32_t n, const float *feat, const float *targ, float *loss,
                        float *grad);
int far_model_eval_host(const float *params, int32_t n, const float *feat, float *out);
int far_model_train_host(float *params, int32_t n, const float *feat, const float *targ, int32_t epochs,
                         int32_t batch, float lr, uint64_t seed, float *losses);
int bh_morton_host(int32_t n, const double *xy, uint32_t *xs, uint32_t *ys, double *maxmin);
int bh_tree_host(int32_t n, const double *xy, int32_t *perm, int32_t max_nodes, int32_t *nnodes,
                 int32_t *root, double *diam, double *cx, double *cy, int32_t *child);
int bh_forces_host(int32_t n, const double *xy, double theta, int32_t nq, const double *q, double *out,
                   int32_t precision);
// batched.hip
int run_batched(int32_t count, const int32_t *voff, const int64_t *eoff, const int32_t *rowptr,
                const int32_t *colids, double *xy, const bl_params *p, double *energies, double *finish_ms,
                bl_stats *st);

}  // namespace bl
