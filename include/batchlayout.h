/* batchlayout.h -- C ABI of the MI355X-native BatchLayout library (libbatchlayout.so).
 *
 * The reference (khaled-rahman/BatchLayoutCode) is a CLI program with no library
 * or FFI.  Its in-process seam is the `algorithms` class (sample/algorithms.h:50-85):
 * a constructor that copies the CSR graph (sample/algorithms.cpp:3-15) and one
 * method per layout algorithm that mutates the coordinates in place and returns
 * {ENERGY, seconds}.  This header is that seam as plain C: a graph handle
 * replaces the constructor, bl_run() replaces the layout methods, and the .mtx /
 * -initf / output-file helpers replace IO.h and algorithms.cpp's file code.
 * Every entry point names the reference interface it replaces.
 *
 * Conventions: positions are n x 2 row-major FP64 (x0,y0,x1,y1,...), the same
 * bytes as the reference's Coordinate<double> array (sample/Coordinate.h:26-85).
 * CSR arrays are 0-based int32 with ascending column ids per row, as built by
 * IO.h:60-155 -> CSC.h:147-188 -> CSR.h:155-186.  All functions return BL_OK (0)
 * or a negative BL_E* status; none exits the process or asserts (the reference
 * calls exit(1)/assert instead).  Calls on different graph handles may run
 * concurrently from different host threads; one handle is used by one thread
 * at a time.
 */
#ifndef BATCHLAYOUT_H
#define BATCHLAYOUT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------- */
#define BL_OK 0
#define BL_E_ARG (-1)        /* invalid argument (null pointer, n <= 0, batch <= 0, bad algo ...) */
#define BL_E_IO (-2)         /* file cannot be opened / read / written */
#define BL_E_FORMAT (-3)     /* malformed .mtx (count mismatch, index out of range, non-square) */
#define BL_E_DEVICE (-4)     /* HIP runtime error or no usable gfx950 device */
#define BL_E_DUPKEY (-5)     /* Barnes-Hut: two vertices share a Morton cell (the reference recurses forever) */
#define BL_E_NOMEM (-6)      /* host or device allocation failed */
#define BL_E_GRAPH (-7)      /* CSR fails validation (rowptr not monotone, colid out of range) */
#define BL_E_SCRATCH (-8)    /* bl_run_device_static: the workspace exceeds the static device array */

/* ---- algorithms (the reference's -algo values, Test/BatchLayout.cpp:132-174) ---- */
#define BL_ALGO_CACHE 2      /* cacheBlockingminiBatchForceDirectedAlgorithm, algorithms.cpp:715-794 */
#define BL_ALGO_LINLOG 3     /* LinLogcacheBlockingminiBatchForceDirectedAlgorithm, :874-945 */
#define BL_ALGO_BH 4         /* BarnesHutApproximation, :1020-1083 */
#define BL_ALGO_BHCACHE 6    /* approxCacheBlockBH (80 % BH, then 20 % algo 2), :1192-1215 */
#define BL_ALGO_FA 8         /* FAcacheBlockingminiBatchForceDirectedAlgorithm, :947-1018 */
/* Physically consistent Barnes-Hut layouts (no reference counterpart).  The iteration of
 * algo 4 (repulsion from the iteration-start positions, attraction d * W|d| over every CSR
 * entry with live positions, sequential batches, STEP *= 0.999) with the repulsion law of
 * algo 2 for every pair, d/|d|^2 (algorithms.cpp:764-768), instead of algo 4's unit vectors: */
#define BL_ALGO_FAR_MONO 40  /* cells accepted by side^2 < bht^2 |c - v|^2 add their
                                mass-weighted monopole m (c - v)/|c - v|^2 */
#define BL_ALGO_FAR_EXACT 41 /* the exact sum over all vertices (O(n^2) per iteration) */
#define BL_ALGO_FAR_LEARNED 42 /* monopoles corrected by the learned far-field model set with
                                  bl_graph_set_far_model (BL_E_ARG without one).  Judged at run
                                  time: every 10 iterations from the first, the exact repulsion of a
                                  sample of 256 vertices decides whether the model (median relative
                                  force error at most the analytic term's) or the analytic
                                  second-order correction (algo 43) serves the next iterations */
#define BL_ALGO_FAR_QUAD 43  /* monopoles corrected by the analytic second-order multipole term */

/* ---- precision of the pair arithmetic -------------------------------------------- */
#define BL_PREC_MIXED 0      /* default: FP64 differences and sums; 1/|d|^2 seeded by the FP32
                                reciprocal and refined by one FP64 Newton step (FP64-grade terms) */
#define BL_PREC_FP64 1       /* reference-grade path: the reference's FP64 operations in its order */
#define BL_PREC_FP32 2       /* fastest: |d|^2 and 1/|d|^2 in FP32, FP64 differences and sums */

typedef struct bl_graph bl_graph;

/* Layout parameters.  Defaults (bl_params_default) are the reference CLI defaults,
 * Test/BatchLayout.cpp:57-58: -weight 1.0 -bht 1.2 -engt 0.01 -batch 256 -iter 600. */
typedef struct {
    int32_t algo;       /* BL_ALGO_* */
    int32_t iters;      /* -iter */
    int32_t batch;      /* -batch (minibatch size B) */
    int32_t converged;  /* algo 2 only: run until |E0-E|/E0 <= engt, ignoring iters
                           (cacheBlockingminiBatchForceDirectedAlgorithmConverged, :795-871;
                           the CLI selects it when -engt != 0.01, BatchLayout.cpp:147-151) */
    double bht;         /* -bht, Barnes-Hut opening threshold theta */
    double engt;        /* -engt */
    double weight;      /* -weight, the global attraction scale W */
    double step0;       /* STEP of the first iteration (1.0 as in the reference; a resumed
                           or teacher-forced run passes the STEP it continues from) */
    int32_t precision;  /* BL_PREC_* (the Barnes-Hut path treats MIXED like FP32) */
    int32_t profile;    /* nonzero: time every launch with HIP events on the run's stream and
                           report per-kernel-class totals in bl_stats (adds a sync per iteration) */
} bl_params;

typedef struct {
    double energy;        /* sum_i |F_i|^2 of the last iteration (the reference's ENERGY) */
    int32_t iters_done;   /* iterations run (differs from iters in converged mode) */
    int32_t status;       /* BL_OK, or the first error met */
    double seconds;       /* wall time of the layout loop (the reference's end - start) */
    double step_final;    /* STEP after the last iteration */
    /* with bl_params.profile: GPU time (ms) and launch count per kernel class:
       [0] pair forces / tree traversal, [1] reduce+move / the batch walk,
       [2] energy reduction, [3] tree build (bounding box, keys, sort, nodes, centroids),
       [4] far-field cell summaries (algos 40, 42, 43), [5] the batch walk's pre-pass
       (attraction of the neighbours that have not moved yet), [6] the learned far field's judge,
       [7] unused */
    double kernel_ms[8];
    int32_t kernel_launches[8];
    /* Barnes-Hut family (algos 4, 6, 40-43): the batch walk of the last iteration, one of BL_WALK_*
       (0 for the all-pairs algorithms) */
    int32_t walk;
    /* algos 40, 42, 43: cells the far-field traversal accepted in the last iteration (with the
       learned model, the number of model evaluations); 0 otherwise */
    int64_t far_cells;
    /* algo 42: iterations whose traversal used the learned model (the others fell back to the
       analytic second-order correction), and the median relative force errors of the learned and
       of the analytic field against the exact sum on the judge's last sample */
    int32_t far_model_iters;
    float far_judge_err[2];
    /* BL_WALK_LAG: batches of the last iteration whose sums over long-moved neighbours came from the
       helper workgroups (the walker formed the others itself), and the lag D it ran with.
       BL_WALK_WIDE: batches of the last iteration whose helper sums were ready when the walker
       reached them (it waited for the others), and the largest lag */
    int32_t walk_helped;
    int32_t walk_lag;
    /* vertices whose final position is not finite (NaN or infinite).  The reference's move
       divides by |F| (Coordinate.h:63-65), so a vertex with F = 0 becomes NaN and the NaN
       spreads to its neighbours; the layout keeps that behaviour and reports it here */
    int32_t nonfinite;
} bl_stats;

/* The batch walks of the Barnes-Hut family (bl_stats.walk); the library picks one per graph and
 * batch size (tests may force one with the environment variable BL_WALK=grid|wide|single|pf|row) */
#define BL_WALK_NONE 0
#define BL_WALK_ROW_ORDER 1  /* FP64 mode: the reference's row order, one workgroup */
#define BL_WALK_PF 2         /* one workgroup, a batch's CSR entries split evenly over its threads */
#define BL_WALK_SPLIT 3      /* one workgroup, batches larger than the prefetching walk takes */
#define BL_WALK_ROW 4        /* one workgroup, one thread per vertex, short rows */
#define BL_WALK_GRID 5       /* several co-resident workgroups, one grid barrier per batch */
#define BL_WALK_LAG 6        /* one walking workgroup (one thread per vertex, short rows) and helper
                                workgroups that sum the neighbours moved at least D batches earlier
                                (D = 20 by default; chosen when B <= 256, every row has at most 64
                                entries and no vertex has more than 4 neighbours in the D - 1
                                batches before its own) */
#define BL_WALK_WIDE 7       /* one walking workgroup that sums only the neighbours moved in the
                                last D_b - 1 batches (D_b <= 16 per batch, from an LDS ring) and
                                helper workgroups that sum the others in chunks of rows; chosen
                                for B <= 1024 where batches hold many CSR entries (long rows) */

/* Fills *p with the reference CLI defaults. */
int bl_params_default(bl_params *p);

/* Replaces the algorithms constructor (algorithms.cpp:3-15), which deep-copies the CSR.
 * Copies rowptr[n+1], colids[nnz] and optional per-entry weights w_edge[nnz] (nullable;
 * an extension: the reference multiplies every attraction by the global W only) to the
 * device.  Validates the CSR (BL_E_GRAPH).  Rows need not be sorted or duplicate-free:
 * the reference's neighbour test (algorithms.cpp:758-762) is reproduced for any input. */
int bl_graph_create(int32_t n, int64_t nnz, const int32_t *rowptr, const int32_t *colids,
                    const float *w_edge, bl_graph **out);
int bl_graph_destroy(bl_graph *g);
int bl_graph_size(const bl_graph *g, int32_t *n, int64_t *nnz);

/* Replaces the layout methods (algorithms.h:67-80): runs the layout on host positions
 * xy_inout[2n] (updated in place), synchronously.  Keeps no pointer to caller memory. */
int bl_run(bl_graph *g, double *xy_inout, const bl_params *p, bl_stats *st);

/* Same, on positions already resident in device memory (d_xy: 2n doubles allocated by
 * the caller with hipMalloc); `stream` is a hipStream_t or NULL for the library's own.
 * Returns after the layout has finished. */
int bl_run_device(bl_graph *g, double *d_xy, const bl_params *p, bl_stats *st, void *stream);

/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * recorded under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * bl_run_device for the minibatch family (algos 2, 3, 8; profile 0, otherwise BL_E_ARG) whose
 * workspace (per-batch partial forces, per-vertex energy terms, energy shares) lives in one
 * static device array of BL_STATIC_SCRATCH_BYTES per GPU, which the GPU holds until the process
 * ends, instead of the graph's buffers: the call allocates nothing (the graph handle, created
 * beforehand, holds the CSR).  The workspace takes 4096 + 16 * min(batch, n) * P + 8 * n bytes
 * (each part rounded up to 256), P = the j chunks of one batch step (at most
 * max(256, n / 1024 + 1) in the FP32 and mixed modes, n / 512 rounded up in FP64 mode) plus,
 * outside FP64 mode, the correction pieces of the longest row (one per 256 entries); batches of
 * 256 fit up to about 2.5M vertices (2M in FP64 mode).  More than the array holds returns
 * BL_E_SCRATCH before anything runs.  Every call on one GPU uses that GPU's array and leaves
 * its contents behind (nothing in it needs to start at zero): calls that overlap on one GPU,
 * from different threads or streams, race. */
#define BL_STATIC_SCRATCH_BYTES (32u << 20)
int bl_run_device_static(bl_graph *g, double *d_xy, const bl_params *p, bl_stats *st, void *stream);

/* ---- host I/O: the reference's file formats ------------------------------------- */
/* .mtx -> CSR, replacing SetInputMatricesAsCSR (sample/commonutility.h:44-53) =
 * ReadASCII (IO.h:60-155) + CSC(Triple*) (CSC.h:147-188) + CSR(const CSC&) (CSR.h:155-186).
 * *rowptr and *colids are allocated with malloc; release them with bl_free. */
int bl_mtx_read(const char *path, int32_t *n, int64_t *nnz, int32_t **rowptr, int32_t **colids);
void bl_free(void *p);

/* -initf reader, replacing algorithms::fileInitialization (algorithms.cpp:98-119): one
 * "x y id" line per vertex, in vertex order (the id column is ignored).  Reads at most n
 * lines (the reference overflows its array on more); *lines_read reports how many. */
int bl_initf_read(const char *path, int32_t n, double *xy, int32_t *lines_read);

/* Output writer, replacing algorithms::writeToFile (algorithms.cpp:1244-1256):
 * "x<TAB>y<TAB>id" with 1-based ids and 6 significant digits (default ostream). */
int bl_output_write(const char *path, int32_t n, const double *xy);

/* ---- Barnes-Hut pieces (for tooling and parity checks) -------------------------- */
/* Morton scales of MortonCode.h:42-67 on the device: xs[i] = xscale(x_i), ys[i] =
 * yscale(y_i), *maxmin = MAXMIN.  Host buffers. */
int bl_bh_morton(int32_t n, const double *xy, uint32_t *xs, uint32_t *ys, double *maxmin);

/* Sort permutation and compressed quadtree of BarnesHut's constructor and build_tree
 * (BarnesHut.h:114-220), built on the device.  perm[k] = vertex at Morton rank k.
 * The tree is the reference's quadtree with single-child chains collapsed into their
 * lowest node (chain nodes share one centroid, so the traversal is unchanged).
 * child[4v+q]: >= 0 node, -1 empty quadrant, <= -2 leaf holding vertex (-c-2).
 * diam[v] is the node's cell diameter; *root is the root node id (or -1 when n == 1).
 * Arrays hold max_nodes entries (n-1 always suffices). */
int bl_bh_tree(int32_t n, const double *xy, int32_t *perm, int32_t max_nodes, int32_t *nnodes,
               int32_t *root, double *diam, double *cx, double *cy, int32_t *child);

/* calcRepForce (BarnesHut.h:69-86, :221-225) on the device: for each query q_k, the sum of
 * unit vectors toward accepted cells and leaves of the tree of the n points xy. */
int bl_bh_forces(int32_t n, const double *xy, double theta, int32_t nq, const double *q,
                 double *out, int32_t precision);

/* Repulsion of the physically consistent layouts (no reference counterpart) on the device:
 * out[k] = approximation of sum_j (x_j - q_k)/|x_j - q_k|^2 over the n points xy, on the
 * compressed quadtree of xy (single points exactly; a cell is accepted when its side^2 <
 * theta^2 |c - q|^2, c = mean of its m points).  Seen from q, an accepted cell's exact field is
 * (m/rho)((1 + a) u + b w) with rho = |c - q|, u = (c - q)/rho, w = u turned by +90 degrees:
 *   BL_FAR_MONO     a = b = 0 (mass-weighted monopole)
 *   BL_FAR_QUAD     a = S_uu - S_ww, b = -2 S_uw (second-order multipole; S = the cell's central
 *                   second moments / (m rho^2) in the (u, w) frame)
 *   BL_FAR_LEARNED  (a, b) = (side/rho)^2 * model(features), model = BL_FAR_MODEL_PARAMS floats
 *   BL_FAR_EXACT    the exact sum (theta ignored) */
#define BL_FAR_MONO 1
#define BL_FAR_EXACT 2
#define BL_FAR_LEARNED 3
#define BL_FAR_QUAD 4
int bl_far_forces(int32_t n, const double *xy, double theta, int32_t mode, const float *model, int32_t nq,
                  const double *q, double *out);

/* ---- learned far field (no reference counterpart) ----------------------------------- */
/* The model: 8 features -> 64 -> 64 -> 64 -> 2, GELU (tanh form), evaluated with BF16 weights
 * and activations on the matrix cores, FP32 accumulation.  Parameters (FP32, this order):
 * W1[64][8] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[2][64] b4[2].
 * Features of a cell seen from q: [side/rho, S_uu, S_uw, S_ww, S_uu - S_ww, log2(m)/16, 1/m,
 * (side/rho)^2]; the model's outputs are (a, b) / (side/rho)^2. */
#define BL_FAR_MODEL_PARAMS 9026
#define BL_FAR_FEATURES 8

/* Training data: for each of the nq query points, the cells accepted by the traversal at
 * theta (at most cap, in traversal order): feat[(k*cap + j)*8 ..] and the exact (a, b) in
 * targ[(k*cap + j)*2 ..] (the exact field summed over the cell's points in FP64); count[k] =
 * number of rows of query k. */
int bl_far_samples(int32_t n, const double *xy, double theta, int32_t nq, const double *q, int32_t cap,
                   float *feat, float *targ, int32_t *count);

/* Glorot-uniform weights drawn from `seed`, zero biases, output layer scaled by 0.1. */
int bl_far_model_init(float *params, uint64_t seed);
/* out[2k..2k+2) = model(feat[8k..8k+8)) on the device. */
int bl_far_model_eval(const float *params, int32_t n, const float *feat, float *out);
/* Loss L = (1/n) sum_k |model(feat_k) - targ_k|^2 and its gradient (BL_FAR_MODEL_PARAMS floats),
 * by the device's forward and backward kernels. */
int bl_far_model_grad(const float *params, int32_t n, const float *feat, const float *targ, float *loss,
                      float *grad);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the device for `epochs` passes over a shuffled
 * order of the n samples (shuffle drawn from seed), `batch` samples per step; params updated in
 * place; losses[e] (nullable) = mean loss of epoch e.  Deterministic for a given seed. */
int bl_far_model_train(float *params, int32_t n, const float *feat, const float *targ, int32_t epochs,
                       int32_t batch, float lr, uint64_t seed, float *losses);
/* Attaches a copy of the model to the graph for BL_ALGO_FAR_LEARNED (NULL detaches). */
int bl_graph_set_far_model(bl_graph *g, const float *params);

/* ---- batched small graphs -------------------------------------------------------- */
/* Lays out `count` independent graphs in one persistent launch (no counterpart in the
 * reference, which lays out one graph per process; each graph follows the same algorithm as
 * bl_run).  Graph k has vertices [voff[k], voff[k+1]) and CSR entries [eoff[k], eoff[k+1])
 * of the concatenated arrays (voff[0] == eoff[0] == 0); rowptr is global
 * (rowptr[voff[k]] == eoff[k]) and colids are local to the graph.  Offsets and CSR are
 * validated on the host before any device work (BL_E_ARG / BL_E_GRAPH).  Every graph runs p->algo (2 or 8) with p->batch/iters/weight in
 * precision BL_PREC_MIXED or BL_PREC_FP32; energies[k] (nullable) receives each graph's final
 * energy and finish_ms[k] (nullable) the time from the launch start to the graph's
 * completion.  Graphs of up to BL_BATCHED_MAX_N vertices. */
#define BL_BATCHED_MAX_N 8192
int bl_run_batched(int32_t count, const int32_t *voff, const int64_t *eoff, const int32_t *rowptr,
                   const int32_t *colids, double *xy_inout, const bl_params *p, double *energies,
                   double *finish_ms, bl_stats *st);

/* Library/device information: writes a short NUL-terminated description. */
int bl_device_info(char *buf, int32_t len);

#ifdef __cplusplus
}
#endif
#endif /* BATCHLAYOUT_H */
