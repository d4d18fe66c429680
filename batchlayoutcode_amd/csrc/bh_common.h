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

// status word bits (misc[0])
constexpr int ST_DUPKEY = 1;

// Builds the tree of the n points d_pos into w: w.keys_a / w.idx_a sorted keys and vertex
// ids, w.spos positions in Morton order, w.node_child nodes, w.frame = {minX, minY, MAXMIN},
// misc[1] = root node (-1 for a single point), misc[2] = node count.  The nodes' unweighted
// centroids (BHNode.c, algo 4's traversal) are not formed: the far field summarises the cells
// itself (far_cells_of).
int bh_build_tree(BHWork &w, const double2 *d_pos, int n, hipStream_t s);

// Repulsion of every vertex for the physically consistent layouts (farfield.hip): into
// w.rep (by vertex id).  FAR_EXACT sums over all points; the other modes read the tree of
// d_pos already built in w by bh_build_tree.  FAR_LEARNED needs model (device, MLP_NPARAM
// floats).
// The lagged batch walk (walk_lag.hip): lag_walk_fits builds the settled / recent split for
// batch size B and lag D (false: the walk does not take this graph and batch size);
// launch_walk_lag runs one iteration's walk with `helpers` helper workgroups (after
// k_late_attract); lag_walk_prepare sets the kernels' LDS limit.
int lag_walk_choose(bl_graph *g, int B);
bool lag_walk_fits(bl_graph *g, int B, int D);
int launch_walk_lag(bl_graph *g, double2 *pos, const bl_params *p, double step, int helpers, hipStream_t s);
int lag_walk_prepare(bool has_w);
int *lag_walk_helped(bl_graph *g);

// The other modes need the cell summaries of that tree first (far_cells_of).  ncells
// (device, nullable) is incremented by the number of cells the traversal accepts.
enum FarMode : int { FAR_MONO = 1, FAR_EXACT = 2, FAR_LEARNED = 3, FAR_QUAD = 4 };
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
