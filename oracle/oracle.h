/* CPU oracle for the BatchLayout hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A clean-room FP64 restatement, in C with OpenMP, of the reference's layout
 * algorithms (sample/algorithms.cpp), Morton ordering (sample/MortonCode.h),
 * quadtree and Barnes-Hut force (sample/BarnesHut.h) and .mtx -> CSR ingest
 * (IO.h, CSC.h, CSR.h).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or the timed CPU
 * baseline; the product library never links or calls it.
 *
 * Every function names the reference lines it restates.  How each is pinned
 * (golden fixtures written by the reference itself, closed forms, brute force)
 * is listed in tests/test_oracle_golden.py.
 *
 * Status codes: 0 ok; negative = error (see ORC_E*).
 */
#ifndef BL_ORACLE_H
#define BL_ORACLE_H
#include <stdint.h>

#define ORC_OK 0
#define ORC_EIO (-1)        /* cannot open / parse */
#define ORC_EFORMAT (-2)    /* malformed .mtx (entry count, range, non-square) */
#define ORC_EDUPKEY (-3)    /* two points share a Morton cell (reference recurses forever) */
#define ORC_ECAP (-4)       /* caller buffer too small */
#define ORC_EARG (-5)

/* attraction laws of the minibatch family */
#define ORC_LAW_FR 2        /* d * (W*|d|)        algorithms.cpp:759   (algo 2) */
#define ORC_LAW_LINLOG 3    /* d * log2(1+W*|d|)  algorithms.cpp:912   (algo 3) */
#define ORC_LAW_FA 8        /* d * W              algorithms.cpp:985   (algo 8) */

#ifdef __cplusplus
extern "C" {
#endif

/* .mtx -> CSR exactly as IO.h:60-155 + CSC.h:147-188 + CSR.h:155-186.
 * Arrays are malloc'ed; release with orc_free. */
int orc_mtx_read(const char *path, int32_t *n, int64_t *nnz, int32_t **rowptr, int32_t **colids);
void orc_free(void *p);

/* Minibatch layout (algorithms.cpp:715-794 / :874-945 / :947-1018; converged
 * variant :795-871 when converged != 0).  xy: n x 2 row-major FP64, updated in
 * place.  step0 is the STEP of the first iteration; STEP *= decay after each.
 * w_edge (nullable) scales W per CSR entry (extension; the reference has none). */
int orc_minibatch(int law, int32_t n, const int32_t *rowptr, const int32_t *colids, const float *w_edge,
                  double *xy, int iters, int batch, double weight, double step0, double decay,
                  int converged, double engt, int threads, double *energy, int *iters_done);

/* Barnes-Hut layout, algorithms.cpp:1020-1083. */
int orc_bh_layout(int32_t n, const int32_t *rowptr, const int32_t *colids, const float *w_edge,
                  double *xy, int iters, int batch, double theta, double weight, double step0,
                  int threads, double *energy);

/* Algo 6 (algorithms.cpp:1192-1215): (int)(0.8*iters) BH iterations at theta 1.2,
 * then the rest as algo 2 with STEP = pow(0.999, 4*rest). */
int orc_hybrid_layout(int32_t n, const int32_t *rowptr, const int32_t *colids, double *xy, int iters,
                      int batch, double weight, int threads, double *energy);

/* Morton scales (MortonCode.h:42-67): xs[i] = xscale(x_i), ys[i] = yscale(y_i);
 * *maxmin = MAXMIN (gridLength(1)). */
int orc_morton(int32_t n, const double *xy, uint32_t *xs, uint32_t *ys, double *maxmin);

/* Uncompressed quadtree (BarnesHut.h:114-220), nodes in the reference's
 * pre-order.  perm[k] = index of the k-th point in Morton order.
 * child[4*v+q]: >=0 node id, -1 empty quadrant, <=-2 leaf (point -c-2). */
int orc_bh_tree(int32_t n, const double *xy, int32_t *perm, int32_t max_nodes, int32_t *nnodes,
                int32_t *parent, int32_t *quad, double *diam, double *cx, double *cy, int32_t *child);

/* calcRepForce (BarnesHut.h:69-86, :221-225) for nq query points against the
 * tree of the n points xy. */
int orc_bh_forces(int32_t n, const double *xy, double theta, int32_t nq, const double *q, double *out,
                  int threads);

/* Exact repulsion sum_j (x_j - x_i)/|x_j - x_i|^2 over all j with |d|>0, for a
 * sample of targets (the physical law of algo 2's non-neighbour term). */
int orc_exact_repulsion(int32_t n, const double *xy, int32_t nq, const int32_t *targets, double *out,
                        int threads);


/* Physically consistent Barnes-Hut (no reference counterpart; the baseline for the learned
 * far field): for each query v, sum over the compressed quadtree of the Morton-sorted points
 * (cells where the points first split, side MAXMIN * 2^-d) of m * (c - v)/|c - v|^2 for every
 * cell with side^2 < theta^2 |c - v|^2 (c = mean of the cell's points, m = count), else
 * recursing; single points add (p - v)/|p - v|^2 when |p - v| > 0.  theta = 0 gives the
 * exact sum. */
int orc_far_forces(int32_t n, const double *xy, double theta, int32_t nq, const double *q, double *out,
                   int threads);

/* The cells accepted for one query (qx, qy) by the traversal above, in depth-first quadrant
 * order: feat[8k..8k+8) the learned far-field features, targ[2k..2k+2) the exact-sum targets
 * (a, b) with F_cell = (m/rho)((1 + a) u + b w), cell[2k] = first Morton rank, cell[2k+1] = m.
 * *force = the monopole force of the whole traversal. */
int orc_far_samples(int32_t n, const double *xy, double theta, double qx, double qy, int32_t cap, double *feat,
                    double *targ, int32_t *cell, int32_t *count, double *force);

/* The algo-4 iteration with orc_far_forces as the repulsion (theta = 0: exact repulsion). */
int orc_far_layout(int32_t n, const int32_t *rowptr, const int32_t *colids, const float *w_edge, double *xy,
                   int iters, int batch, double theta, double weight, double step0, int threads, double *energy);

#ifdef __cplusplus
}
#endif
#endif
