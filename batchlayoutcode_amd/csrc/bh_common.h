// Quadtree layout shared by the Barnes-Hut translation units (bh.hip builds the tree;
// farfield.hip reads it).
#pragma once

#include "bl_internal.h"

namespace bl {

// One node per cell where the Morton-sorted points first split (single-child chains of the
// uncompressed quadtree collapsed into their lowest node).
struct BHNode {
    double2 c;     // unweighted centroid of the non-empty quadrants (BarnesHut.h:197-217)
    int4 ch;       // children: >= 0 node, -1 empty, <= -2 leaf at Morton rank (-c-2)
    int depth;     // base-4 level of the cell: side = MAXMIN * 2^-depth
    int nint;      // number of internal children
    int parent;
    int cnt;       // bottom-up arrival counter
};

// 1/sqrt(x) for x > 0: the hardware estimate (v_rsq_f64, relative error about 2^-25) refined by
// one Newton step y <- y (3/2 - (x/2) y^2), within about 20 ulp of the correctly rounded value
// (a second step, a few ulp, made no difference to the walks on the large graphs and cost the
// one-batch-a-step walk 10%); five dependent operations where an IEEE square root and a
// division take about twenty-five.
// x = 0 gives NaN (0 * inf), as the reference's 0 / 0 does.
__device__ __forceinline__ double rsqrt2(double x) {
    const double h = 0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

// The tile of workgroup b of nb.  The dispatcher deals workgroups to the 8 XCDs in turn
// (workgroup b to XCD b % 8), so tile = start(b % 8) + b / 8 gives each XCD one contiguous
// eighth of the tiles, and the workgroups in flight on an XCD take neighbouring tiles, whose
// data (for Morton-ordered queries: the same subtrees) then shares that XCD's L2.  A bijection
// of [0, nb) whatever the placement, so results do not depend on it.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int x = b & 7, r = b >> 3, q = nb >> 3, rem = nb & 7;
    return x * q + min(x, rem) + r;
}

// Stores for another workgroup with agent-scope stores, which the XCD's L2 writes through to
// memory: once they have completed (s_waitcnt), a relaxed flag or counter published after them
// needs no L2 write-back, and a reader that acquired after seeing it reads the new values.
__device__ __forceinline__ void store_through(double2 *p, double2 v) {
    __hip_atomic_store(&p->x, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&p->y, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_through(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double2 add2(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// One step of a segmented inclusive scan over the wave by a DPP move: each lane takes the
// value and key of the lane CTRL names (row_shr:n within rows of 16 lanes; row_bcast:15 /
// row_bcast:31, the last lane of the previous row / of the first two rows, into the rows
// of ROW_MASK) and adds the value when the keys are equal.  Lanes without a source lane
// see the key -2, which no lane holds.  After the steps 1, 2, 4, 8, bcast15, bcast31 lane l
// holds the sum of the values of its run (the lanes before it with its key) up to l:
// runs are contiguous, so a run that reaches back into an earlier row of 16 contains that
// row's last lane.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void seg_step(double2 &v, int key) {
    const int ks = __builtin_amdgcn_update_dpp(-2, key, CTRL, ROW_MASK, 0xF, false);
    const int xl = __builtin_amdgcn_update_dpp(0, __double2loint(v.x), CTRL, ROW_MASK, 0xF, false);
    const int xh = __builtin_amdgcn_update_dpp(0, __double2hiint(v.x), CTRL, ROW_MASK, 0xF, false);
    const int yl = __builtin_amdgcn_update_dpp(0, __double2loint(v.y), CTRL, ROW_MASK, 0xF, false);
    const int yh = __builtin_amdgcn_update_dpp(0, __double2hiint(v.y), CTRL, ROW_MASK, 0xF, false);
    if (ks == key) {
        v.x += __hiloint2double(xh, xl);
        v.y += __hiloint2double(yh, yl);
    }
}

// One attraction term of vertex v toward pj, d * (Wk |d|) (algorithms.cpp:1055-1058), added to
// (ax, ay), with |d| = |d|^2 / sqrt(|d|^2) from rsqrt2 (coincident or nearly coincident vertices,
// |d| < 2^-450, add nothing, where the reference adds at most 2^-900 W).  The walks with helper
// workgroups use it in the walker and in the helpers alike, so their sums agree bit for bit.
__device__ __forceinline__ void attr_add(double2 pj, double2 v, double Wk, double &ax, double &ay) {
    const double dx = pj.x - v.x, dy = pj.y - v.y;
    const double d2 = dx * dx + dy * dy;
    const double sc = d2 > 0x1p-900 ? Wk * (d2 * rsqrt2(d2)) : 0.0;
    ax = fma(dx, sc, ax);
    ay = fma(dy, sc, ay);
}

// status word bits (misc[0])
constexpr int ST_DUPKEY = 1;

// Builds the tree of the n points d_pos into w: w.keys_a / w.idx_a sorted keys and vertex
// ids, w.spos positions in Morton order, w.node_child nodes, w.frame = {minX, minY, MAXMIN},
// misc[1] = root node (-1 for a single point), misc[2] = node count.  The nodes' unweighted
// centroids (BHNode.c, algo 4's traversal) are not formed: the far field summarises the cells
// itself (far_cells_of).
int bh_build_tree(BHWork &w, const double2 *d_pos, int n, hipStream_t s);

// The lagged batch walk (walk_lag.hip): lag_walk_fits builds the settled / recent split for
// batch size B and lag D (false: the walk does not take this graph and batch size);
// launch_walk_lag runs one iteration's walk with `helpers` helper workgroups (after
// k_late_attract); lag_walk_prepare sets the kernels' LDS limit.
int lag_walk_choose(bl_graph *g, int B);
bool lag_walk_fits(bl_graph *g, int B, int D);
int launch_walk_lag(bl_graph *g, double2 *pos, const bl_params *p, double step, int helpers, hipStream_t s);
int lag_walk_prepare(bool has_w);
int *lag_walk_helped(bl_graph *g);

// The wide lagged batch walk (walk_wide.hip): wide_walk_fits builds the near / far split and
// the helpers' chunks for batch size B (largest lag from BL_WIDE_D; false: the walk does not
// take this graph and batch size); launch_walk_wide runs one iteration's walk (no pre-pass)
// with up to `helpers` helper workgroups; wide_walk_read queues the copies of the error flag
// and of the number of batches the walker waited for (valid after the stream synchronizes);
// wide_walk_prepare sets the kernels' LDS limit.
bool wide_walk_fits(bl_graph *g, int B);
int launch_walk_wide(bl_graph *g, double2 *pos, const bl_params *p, double step, int helpers, hipStream_t s);
int wide_walk_read(bl_graph *g, int *err, int *waited, hipStream_t s);
int wide_walk_reset(bl_graph *g);
int wide_walk_prepare(bool has_w);

// Repulsion of every vertex for the physically consistent layouts (farfield.hip): into
// w.rep (by vertex id).  FAR_EXACT sums over all points; the other modes read the tree of
// d_pos already built in w by bh_build_tree.  FAR_LEARNED needs model (device, MLP_NPARAM
// floats).  The other modes need the cell summaries of that tree first (far_cells_of).  ncells
// (device, nullable) is incremented by the number of cells the traversal accepts.
enum FarMode : int { FAR_MONO = 1, FAR_EXACT = 2, FAR_LEARNED = 3, FAR_QUAD = 4 };
// Once per tree build: the climb consumes the arrival counters (BHNode.cnt) the build zeroed.
int far_cells_of(BHWork &w, int n, hipStream_t s);
// judged (FAR_LEARNED): the traversal follows the verdict of the last far_judge (the model, or
// the analytic second-order correction when the model lost)
int far_repulsion(BHWork &w, const double2 *d_pos, int n, double theta, int mode, const float *model,
                  hipStream_t s, unsigned long long *ncells, bool judged);
// The learned field's judge: the exact repulsion at a sample of the positions d_pos (of the tree
// in w), against the learned and the analytic second-order fields there; sets the verdict on
// the device and returns it in *verdict (1: the model; waits for the stream).  far_judge_reset
// starts a run (model on, counters cleared); far_judge_read queues the copies of the number of
// traversals that used the model and the last medians.
int far_judge(BHWork &w, const double2 *d_pos, int n, double theta, const float *model, int it, hipStream_t s,
              int *verdict);
int far_judge_reset(BHWork &w, hipStream_t s);
int far_judge_read(BHWork &w, int *model_iters, float *err, hipStream_t s);

}  // namespace bl
