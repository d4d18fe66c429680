/* CPU oracle for the BatchLayout hot path -- TEST INFRASTRUCTURE ONLY.
 * See oracle.h for the contract.  Built strict (-ffp-contract=off, no
 * -ffast-math) so every FP64 operation rounds as the reference's strict-IEEE
 * build (reference Makefile:11) does. */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* .mtx ingest: IO.h:60-155 (ReadASCII), CSC.h:147-188, CSR.h:155-186.       */
/* ------------------------------------------------------------------------ */
typedef struct { int32_t r, c; } ent_t;

static int ent_cmp(const void *a, const void *b) {
    const ent_t *x = (const ent_t *)a, *y = (const ent_t *)b;
    if (x->r != y->r) return x->r < y->r ? -1 : 1;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    return 0;
}

int orc_mtx_read(const char *path, int32_t *n_out, int64_t *nnz_out, int32_t **rowptr_out,
                 int32_t **colids_out) {
    FILE *f = fopen(path, "r");
    if (!f) return ORC_EIO;
    char *line = NULL;
    size_t cap = 0;
    ssize_t len;
    int symmetric = 0;
    long m = -1, ncol = -1, nnz = -1;
    /* header: every line starting with '%'; "symmetric" anywhere turns on mirroring (IO.h:67-75) */
    while ((len = getline(&line, &cap, f)) >= 0) {
        if (line[0] == '%') {
            if (strstr(line, "symmetric")) symmetric = 1;
            continue;
        }
        if (sscanf(line, "%ld %ld %ld", &m, &ncol, &nnz) != 3) { free(line); fclose(f); return ORC_EFORMAT; }
        break;
    }
    if (m <= 0 || ncol != m || nnz < 0) { free(line); fclose(f); return ORC_EFORMAT; }
    long want = symmetric ? 2 * nnz : nnz; /* IO.h:90-92 */
    ent_t *e = (ent_t *)malloc(sizeof(ent_t) * (size_t)(want > 0 ? want : 1));
    long cnz = 0;
    while (cnz < want && (len = getline(&line, &cap, f)) >= 0) {
        char *p = line, *q;
        long r = strtol(p, &q, 10);
        if (q == p) { free(e); free(line); fclose(f); return ORC_EFORMAT; }
        p = q;
        long c = strtol(p, &q, 10);
        if (q == p) { free(e); free(line); fclose(f); return ORC_EFORMAT; }
        r--, c--; /* 1-based -> 0-based, IO.h:120-121 */
        if (r < 0 || r >= m || c < 0 || c >= m) { free(e); free(line); fclose(f); return ORC_EFORMAT; }
        if (symmetric) {
            if (r == c) { want -= 2; continue; } /* diagonal dropped, IO.h:131-134 */
            e[cnz].r = (int32_t)r; e[cnz].c = (int32_t)c; cnz++;
            e[cnz].r = (int32_t)c; e[cnz].c = (int32_t)r; cnz++; /* mirror, IO.h:124-129 */
        } else {
            e[cnz].r = (int32_t)r; e[cnz].c = (int32_t)c; cnz++; /* self-loops kept */
        }
    }
    free(line);
    fclose(f);
    if (cnz != want) { free(e); return ORC_EFORMAT; } /* assert(cnz == nnz), IO.h:142 */
    /* CSC sorts rows within each column (CSC.h:173-185); CSR then walks columns in
     * order (CSR.h:175-183): the result is every row's column ids ascending, duplicates
     * kept adjacent.  That is the (row, col) lexicographic order. */
    qsort(e, (size_t)cnz, sizeof(ent_t), ent_cmp);
    int32_t *rowptr = (int32_t *)calloc((size_t)m + 1, sizeof(int32_t));
    int32_t *colids = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cnz > 0 ? cnz : 1));
    for (long k = 0; k < cnz; k++) { rowptr[e[k].r + 1]++; colids[k] = e[k].c; }
    for (long i = 0; i < m; i++) rowptr[i + 1] += rowptr[i];
    free(e);
    *n_out = (int32_t)m;
    *nnz_out = cnz;
    *rowptr_out = rowptr;
    *colids_out = colids;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Minibatch force family: algorithms.cpp:715-794 (algo 2), :795-871         */
/* (converged), :874-945 (algo 3), :947-1018 (algo 8).                       */
/* ------------------------------------------------------------------------ */
static inline void attract(int law, double W, double dx, double dy, double *fx, double *fy) {
    double mag = sqrt(dx * dx + dy * dy);
    double s;
    if (law == ORC_LAW_FR) s = W * mag;               /* :759 */
    else if (law == ORC_LAW_FA) s = W;                /* :985 */
    else s = log2(1.0 + W * mag);                     /* :912 */
    *fx += dx * s;
    *fy += dy * s;
}

int orc_minibatch(int law, int32_t n, const int32_t *rowptr, const int32_t *colids, const float *w_edge,
                  double *xy, int iters, int batch, double weight, double step0, double decay,
                  int converged, double engt, int threads, double *energy, int *iters_done) {
    if (n <= 0 || batch <= 0 || !xy) return ORC_EARG;
    if (law != ORC_LAW_FR && law != ORC_LAW_FA && law != ORC_LAW_LINLOG) return ORC_EARG;
    const int BX = 2, BY = 512; /* blockx, blocky, algorithms.cpp:717 */
    const int64_t nnz = rowptr[n];
    double *prev = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    int64_t *kidx = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    if (threads > 0) omp_set_num_threads(threads);
    double STEP = step0, E = 0.0, E0 = 1.79769313486231570815e+308; /* DBL_MAX, :806-807 */
    int loop = 0;
    const int nb = (int)ceil(1.0 * n / batch);
    for (;;) {
        if (converged) { if (!(fabs(E0 - E) / E0 > engt)) break; } /* :818 */
        else if (loop >= iters) break;
        E0 = E;
        E = 0.0;
        for (int32_t k = 0; k < n; k++) { kidx[k] = rowptr[k]; prev[2 * k] = prev[2 * k + 1] = 0.0; }
        for (int b = 0; b < nb; b++) {
            const int64_t i0 = (int64_t)b * batch, i1 = (int64_t)(b + 1) * batch;
#pragma omp parallel for schedule(static)
            for (int64_t i = i0; i < i1; i += BX) {
                if (i >= n) continue;
                for (int64_t j = 0; j < n; j += BY) {
                    for (int bi = 0; bi < BX && i + bi < i1; bi++) {
                        const int64_t v = i + bi;
                        if (v >= n) break;
                        double fx = 0.0, fy = 0.0;
                        const double vx = xy[2 * v], vy = xy[2 * v + 1];
                        for (int64_t bj = 0; bj < BY && j + bj < n; bj++) {
                            const int64_t u = j + bj;
                            /* merge pointer over v's sorted colids (:758-762); an empty
                             * last row would read colids[nnz]: treated as no neighbour */
                            const int64_t kp = kidx[v];
                            const int32_t cand = kp < nnz ? colids[kp] : -1;
                            const double dx = xy[2 * u] - vx, dy = xy[2 * u + 1] - vy;
                            if (u == cand) {
                                double W = weight * (w_edge ? (double)w_edge[kp] : 1.0);
                                attract(law, W, dx, dy, &fx, &fy);
                                if (kp < (int64_t)rowptr[v + 1] - 1) kidx[v]++;
                            } else {
                                const double dist = dx * dx + dy * dy;
                                if (dist > 0) { /* :764-768 */
                                    const double inv = 1.0 / dist;
                                    fx = fx - dx * inv;
                                    fy = fy - dy * inv;
                                }
                            }
                        }
                        prev[2 * v] += fx; /* :771 */
                        prev[2 * v + 1] += fy;
                    }
                }
            }
            for (int64_t v = i0; v < i1 && v < n; v++) { /* sequential move, :775-779 */
                const double fx = prev[2 * v], fy = prev[2 * v + 1];
                const double mag = sqrt(fx * fx + fy * fy);
                xy[2 * v] = xy[2 * v] + (fx / mag) * STEP;
                xy[2 * v + 1] = xy[2 * v + 1] + (fy / mag) * STEP;
                E += fx * fx + fy * fy;
            }
        }
        STEP = STEP * decay; /* :781, :934, :1007 */
        loop++;
    }
    free(prev);
    free(kidx);
    if (energy) *energy = E;
    if (iters_done) *iters_done = loop;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Morton ordering, MortonCode.h:42-97.                                      */
/* ------------------------------------------------------------------------ */
typedef struct { double minx, miny, maxmin; } frame_t;

static frame_t frame_of(int32_t n, const double *xy) {
    double mnx = xy[0], mxx = xy[0], mny = xy[1], mxy = xy[1];
    for (int32_t i = 1; i < n; i++) {
        double x = xy[2 * i], y = xy[2 * i + 1];
        if (x < mnx) mnx = x;
        if (x > mxx) mxx = x;
        if (y < mny) mny = y;
        if (y > mxy) mxy = y;
    }
    frame_t fr;
    fr.minx = mnx;
    fr.miny = mny;
    double dy = mxy - mny, dx = mxx - mnx;
    fr.maxmin = (dy < dx) ? dx : dy; /* std::max(dy, dx), :51 */
    return fr;
}

/* xscale/yscale, MortonCode.h:62-67: (uint32_t)(0xffffffff*(v-min)/MAXMIN) */
static inline uint32_t scale_of(double v, double mn, double maxmin) {
    double t = 4294967295.0 * (v - mn);
    t = t / maxmin;
    return (uint32_t)t;
}

/* comparator MortonCode.h:83-97 == unsigned order of the interleave with the
 * y bit above the x bit at every level */
static inline uint64_t spread(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}
static inline uint64_t morton_key(uint32_t xs, uint32_t ys) { return (spread(ys) << 1) | spread(xs); }

int orc_morton(int32_t n, const double *xy, uint32_t *xs, uint32_t *ys, double *maxmin) {
    if (n <= 0) return ORC_EARG;
    frame_t fr = frame_of(n, xy);
    for (int32_t i = 0; i < n; i++) {
        xs[i] = scale_of(xy[2 * i], fr.minx, fr.maxmin);
        ys[i] = scale_of(xy[2 * i + 1], fr.miny, fr.maxmin);
    }
    if (maxmin) *maxmin = fr.maxmin;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Quadtree, BarnesHut.h:114-220; traversal :69-86.                          */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t key; int32_t idx; } kv_t;
static int kv_cmp(const void *a, const void *b) {
    const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->idx > y->idx) - (x->idx < y->idx);
}

typedef struct {
    int32_t n, cap, cnt, err;
    const uint64_t *key; /* sorted keys */
    const double *sx;    /* sorted coordinates, 2 per point */
    const int32_t *perm;
    int32_t *parent, *quad, *child;
    double *diam, *cx, *cy;
} tree_t;

static int32_t lower_digit(const tree_t *T, int32_t l, int32_t r, int level, int q) {
    /* first position in [l,r) whose digit at `level` is >= q (lower_bound, :143-146) */
    while (l < r) {
        int32_t m = l + (r - l) / 2;
        int d = (int)((T->key[m] >> (2 * level)) & 3u);
        if (d < q) l = m + 1; else r = m;
    }
    return l;
}

static int32_t build_rec(tree_t *T, int32_t l, int32_t r, int level, double dia, int32_t par, int32_t qd) {
    if (level < 0) { T->err = ORC_EDUPKEY; return -1; }
    if (T->cnt >= T->cap) { T->err = ORC_ECAP; return -1; }
    const int32_t id = T->cnt++;
    T->parent[id] = par;
    T->quad[id] = qd;
    T->diam[id] = dia;
    int32_t b[5];
    b[0] = l;
    b[4] = r;
    for (int q = 1; q < 4; q++) b[q] = lower_digit(T, l, r, level, q);
    double sx = 0.0, sy = 0.0, count = 0.0;
    for (int q = 0; q < 4; q++) {
        const int32_t c = b[q + 1] - b[q];
        if (c > 1) { /* make_child, :163-181 */
            int32_t ch = build_rec(T, b[q], b[q + 1], level - 1, dia / 2.0, id, q);
            if (T->err) return -1;
            T->child[4 * id + q] = ch;
            sx += T->cx[ch];
            sy += T->cy[ch];
            count++;
        } else if (c == 1) { /* leaf, :183-189 */
            T->child[4 * id + q] = -(T->perm[b[q]] + 2);
            sx += T->sx[2 * b[q]];
            sy += T->sx[2 * b[q] + 1];
            count++;
        } else {
            T->child[4 * id + q] = -1;
        }
    }
    if (count > 0) { sx /= count; sy /= count; } /* unweighted mean, :197-217 */
    T->cx[id] = sx;
    T->cy[id] = sy;
    return id;
}

typedef struct {
    int32_t n;
    int32_t *perm, *parent, *quad, *child;
    double *diam, *cx, *cy, *sx;
    uint64_t *key;
    int32_t nnodes;
} bhtree_t;

static void bhtree_free(bhtree_t *t) {
    free(t->perm); free(t->parent); free(t->quad); free(t->child);
    free(t->diam); free(t->cx); free(t->cy); free(t->sx); free(t->key);
}

static int bhtree_build(int32_t n, const double *xy, bhtree_t *out) {
    memset(out, 0, sizeof(*out));
    if (n <= 0) return ORC_EARG;
    frame_t fr = frame_of(n, xy);
    kv_t *kv = (kv_t *)malloc(sizeof(kv_t) * (size_t)n);
    for (int32_t i = 0; i < n; i++) {
        kv[i].key = morton_key(scale_of(xy[2 * i], fr.minx, fr.maxmin), scale_of(xy[2 * i + 1], fr.miny, fr.maxmin));
        kv[i].idx = i;
    }
    qsort(kv, (size_t)n, sizeof(kv_t), kv_cmp);
    out->n = n;
    out->perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    out->key = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    out->sx = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int32_t k = 0; k < n; k++) {
        out->perm[k] = kv[k].idx;
        out->key[k] = kv[k].key;
        out->sx[2 * k] = xy[2 * kv[k].idx];
        out->sx[2 * k + 1] = xy[2 * kv[k].idx + 1];
    }
    free(kv);
    for (int32_t k = 1; k < n; k++)
        if (out->key[k] == out->key[k - 1]) { return ORC_EDUPKEY; }
    /* uncompressed tree: <= n-1 branching nodes plus single-child chain nodes; the
     * capacity starts at 4n and doubles when a build runs out */
    int32_t cap = 4 * n + 64;
    for (;;) {
        tree_t T;
        memset(&T, 0, sizeof(T));
        T.n = n; T.cap = cap; T.key = out->key; T.sx = out->sx; T.perm = out->perm;
        T.parent = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
        T.quad = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
        T.child = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)cap);
        T.diam = (double *)malloc(sizeof(double) * (size_t)cap);
        T.cx = (double *)malloc(sizeof(double) * (size_t)cap);
        T.cy = (double *)malloc(sizeof(double) * (size_t)cap);
        build_rec(&T, 0, n, 31, fr.maxmin, -1, -1); /* root: level 1u<<31, gridLength(1), :128 */
        if (T.err == ORC_ECAP) {
            free(T.parent); free(T.quad); free(T.child); free(T.diam); free(T.cx); free(T.cy);
            cap *= 2;
            continue;
        }
        out->parent = T.parent; out->quad = T.quad; out->child = T.child;
        out->diam = T.diam; out->cx = T.cx; out->cy = T.cy; out->nnodes = T.cnt;
        return T.err;
    }
}

/* calc_resulsive_force: depth-first, quadrants 0..3; leaf coordinates are read
 * through the point index from the positions the tree was built on */
static void force_rec2(const bhtree_t *t, const double *xy, int32_t v, double px, double py, double theta,
                       double *fx, double *fy) {
    const double dx = t->cx[v] - px, dy = t->cy[v] - py;
    const double dist = sqrt(dx * dx + dy * dy);
    if (t->diam[v] / dist < theta) { /* accepted cell: one unit vector, :72-73 */
        const double inv = 1.0 / dist;
        *fx = *fx + dx * inv;
        *fy = *fy + dy * inv;
        return;
    }
    for (int q = 0; q < 4; q++) {
        const int32_t c = t->child[4 * v + q];
        if (c >= 0) {
            force_rec2(t, xy, c, px, py, theta, fx, fy);
        } else if (c <= -2) { /* :78-83 */
            const int32_t p = -c - 2;
            const double lx = xy[2 * p] - px, ly = xy[2 * p + 1] - py;
            const double den = sqrt(lx * lx + ly * ly);
            if (den > 0) {
                const double inv = 1.0 / den;
                *fx = *fx + lx * inv;
                *fy = *fy + ly * inv;
            }
        }
    }
}

int orc_bh_tree(int32_t n, const double *xy, int32_t *perm, int32_t max_nodes, int32_t *nnodes,
                int32_t *parent, int32_t *quad, double *diam, double *cx, double *cy, int32_t *child) {
    bhtree_t t;
    int rc = bhtree_build(n, xy, &t);
    if (rc == ORC_OK) {
        if (perm) memcpy(perm, t.perm, sizeof(int32_t) * (size_t)n);
        if (nnodes) *nnodes = t.nnodes;
        if (t.nnodes > max_nodes) rc = ORC_ECAP;
        else {
            memcpy(parent, t.parent, sizeof(int32_t) * (size_t)t.nnodes);
            memcpy(quad, t.quad, sizeof(int32_t) * (size_t)t.nnodes);
            memcpy(diam, t.diam, sizeof(double) * (size_t)t.nnodes);
            memcpy(cx, t.cx, sizeof(double) * (size_t)t.nnodes);
            memcpy(cy, t.cy, sizeof(double) * (size_t)t.nnodes);
            memcpy(child, t.child, sizeof(int32_t) * 4 * (size_t)t.nnodes);
        }
    } else if (perm && t.perm) {
        memcpy(perm, t.perm, sizeof(int32_t) * (size_t)n);
    }
    bhtree_free(&t);
    return rc;
}

int orc_bh_forces(int32_t n, const double *xy, double theta, int32_t nq, const double *q, double *out,
                  int threads) {
    bhtree_t t;
    int rc = bhtree_build(n, xy, &t);
    if (rc != ORC_OK) { bhtree_free(&t); return rc; }
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t i = 0; i < nq; i++) {
        double fx = 0.0, fy = 0.0;
        force_rec2(&t, xy, 0, q[2 * i], q[2 * i + 1], theta, &fx, &fy);
        out[2 * i] = fx;
        out[2 * i + 1] = fy;
    }
    bhtree_free(&t);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Barnes-Hut layout, algorithms.cpp:1020-1083.                              */
/* ------------------------------------------------------------------------ */
static int bh_iterations(int32_t n, const int32_t *rowptr, const int32_t *colids, const float *w_edge,
                         double *xy, int iters, int batch, double theta, double weight, double step0,
                         double *energy) {
    double *prev = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *temp = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double STEP = step0, E = 0.0;
    const int nb = (int)ceil(1.0 * n / batch);
    int rc = ORC_OK;
    for (int loop = 0; loop < iters; loop++) {
        E = 0.0;
        memcpy(temp, xy, sizeof(double) * 2 * (size_t)n); /* tree from iteration-start positions, :1041-1047 */
        bhtree_t t;
        rc = bhtree_build(n, temp, &t);
        if (rc != ORC_OK) { bhtree_free(&t); break; }
        for (int b = 0; b < nb; b++) {
            const int64_t i0 = (int64_t)b * batch, i1 = (int64_t)(b + 1) * batch;
#pragma omp parallel for schedule(static)
            for (int64_t i = i0; i < i1; i++) {
                if (i >= n) continue;
                double fx = 0.0, fy = 0.0;
                const double vx = xy[2 * i], vy = xy[2 * i + 1];
                for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) { /* every CSR entry, :1055-1058 */
                    const int32_t j = colids[k];
                    const double dx = xy[2 * j] - vx, dy = xy[2 * j + 1] - vy;
                    const double W = weight * (w_edge ? (double)w_edge[k] : 1.0);
                    const double s = W * sqrt(dx * dx + dy * dy);
                    fx = fx + dx * s;
                    fy = fy + dy * s;
                }
                double rx = 0.0, ry = 0.0;
                force_rec2(&t, temp, 0, vx, vy, theta, &rx, &ry); /* calcRepForce(live x_i), :1059 */
                prev[2 * i] = fx - rx;
                prev[2 * i + 1] = fy - ry;
            }
            for (int64_t i = i0; i < i1 && i < n; i++) {
                const double fx = prev[2 * i], fy = prev[2 * i + 1];
                const double mag = sqrt(fx * fx + fy * fy);
                xy[2 * i] = xy[2 * i] + (fx / mag) * STEP;
                xy[2 * i + 1] = xy[2 * i + 1] + (fy / mag) * STEP;
                E += fx * fx + fy * fy;
            }
        }
        bhtree_free(&t);
        STEP = STEP * 0.999; /* :1070 */
    }
    free(prev);
    free(temp);
    if (energy) *energy = E;
    return rc;
}

int orc_bh_layout(int32_t n, const int32_t *rowptr, const int32_t *colids, const float *w_edge,
                  double *xy, int iters, int batch, double theta, double weight, double step0,
                  int threads, double *energy) {
    if (n <= 0 || batch <= 0) return ORC_EARG;
    if (threads > 0) omp_set_num_threads(threads);
    return bh_iterations(n, rowptr, colids, w_edge, xy, iters, batch, theta, weight, step0, energy);
}

int orc_hybrid_layout(int32_t n, const int32_t *rowptr, const int32_t *colids, double *xy, int iters,
                      int batch, double weight, int threads, double *energy) {
    if (n <= 0 || batch <= 0) return ORC_EARG;
    if (threads > 0) omp_set_num_threads(threads);
    const int approx = (int)(iters * 0.8); /* :1193 */
    double e = 0.0;
    int rc = bh_iterations(n, rowptr, colids, NULL, xy, approx, batch, 1.2, weight, 1.0, &e); /* :1205 */
    if (rc != ORC_OK) return rc;
    const int rest = iters - approx;
    const double step = pow(0.999, 4 * rest); /* flag==1 branch, :736-737 */
    return orc_minibatch(ORC_LAW_FR, n, rowptr, colids, NULL, xy, rest, batch, weight, step, 0.999, 0, 0.0,
                         threads, energy, NULL);
}

int orc_exact_repulsion(int32_t n, const double *xy, int32_t nq, const int32_t *targets, double *out,
                        int threads) {
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 16)
    for (int32_t a = 0; a < nq; a++) {
        const int32_t i = targets[a];
        const double vx = xy[2 * i], vy = xy[2 * i + 1];
        double fx = 0.0, fy = 0.0;
        for (int32_t j = 0; j < n; j++) {
            const double dx = xy[2 * j] - vx, dy = xy[2 * j + 1] - vy;
            const double d2 = dx * dx + dy * dy;
            if (d2 > 0) {
                fx += dx / d2;
                fy += dy / d2;
            }
        }
        out[2 * a] = fx;
        out[2 * a + 1] = fy;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Physically consistent Barnes-Hut (no reference counterpart): the algo-2   */
/* repulsion law d/|d|^2 with a mass-weighted monopole per accepted cell, on */
/* the compressed quadtree of the Morton-sorted points (a cell exists where  */
/* its points first split; its side is MAXMIN * 2^-d at base-4 digit d).     */
/* ------------------------------------------------------------------------ */
typedef struct {
    const uint64_t *key; /* sorted keys */
    const double *sx;    /* sorted coordinates */
    double maxmin, theta2;
    /* sample output (orc_far_samples) */
    int32_t cap, cnt;
    double *feat, *targ;
    int32_t *cell;
} far_t;

/* first base-4 digit (0 = top) where the keys of [l, r) differ, r - l >= 2 */
static int split_digit(const far_t *F, int32_t l, int32_t r) {
    const uint64_t x = F->key[l] ^ F->key[r - 1];
    return __builtin_clzll(x) / 2;
}

static int32_t lower_q(const far_t *F, int32_t l, int32_t r, int d, int q) {
    while (l < r) {
        int32_t m = l + (r - l) / 2;
        int dig = (int)((F->key[m] >> (62 - 2 * d)) & 3u);
        if (dig < q) l = m + 1; else r = m;
    }
    return l;
}

/* summary of the points [l, r): count, mean, central second moments (sums) */
static void cell_summary(const far_t *F, int32_t l, int32_t r, double *cx, double *cy, double *mxx, double *mxy,
                         double *myy) {
    double sx = 0.0, sy = 0.0;
    for (int32_t k = l; k < r; k++) { sx += F->sx[2 * k]; sy += F->sx[2 * k + 1]; }
    const double m = (double)(r - l);
    *cx = sx / m;
    *cy = sy / m;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int32_t k = l; k < r; k++) {
        const double ux = F->sx[2 * k] - *cx, uy = F->sx[2 * k + 1] - *cy;
        a += ux * ux; b += ux * uy; c += uy * uy;
    }
    *mxx = a; *mxy = b; *myy = c;
}

/* learned far-field features of an accepted cell seen from v (rotated into the frame of
 * u = (c - v)/|c - v|, w = u rotated by +90 degrees), and the targets (a, b) with
 * F_exact = (m/rho) * ((1 + a) u + b w). */
static void far_features(double m, double D, double rx, double ry, double mxx, double mxy, double myy, double *f) {
    const double rho = sqrt(rx * rx + ry * ry);
    const double ux = rx / rho, uy = ry / rho, wx = -uy, wy = ux;
    const double s = 1.0 / (m * rho * rho);
    const double suu = (ux * ux * mxx + 2.0 * ux * uy * mxy + uy * uy * myy) * s;
    const double suw = (ux * wx * mxx + (ux * wy + uy * wx) * mxy + uy * wy * myy) * s;
    const double sww = (wx * wx * mxx + 2.0 * wx * wy * mxy + wy * wy * myy) * s;
    f[0] = D / rho;
    f[1] = suu;
    f[2] = suw;
    f[3] = sww;
    f[4] = suu - sww;
    f[5] = log2(m) / 16.0;
    f[6] = 1.0 / m;
    f[7] = (D / rho) * (D / rho);
}

static void far_rec(far_t *F, int32_t l, int32_t r, double vx, double vy, int mode, double *fx, double *fy) {
    const int d = split_digit(F, l, r);
    const double D = ldexp(F->maxmin, -d);
    double cx, cy, mxx, mxy, myy;
    cell_summary(F, l, r, &cx, &cy, &mxx, &mxy, &myy);
    const double rx = cx - vx, ry = cy - vy;
    const double r2 = rx * rx + ry * ry;
    if (D * D < F->theta2 * r2) {
        const double m = (double)(r - l);
        if (mode == 1 && F->cnt < F->cap) { /* training sample: exact sum over the cell */
            double ex = 0.0, ey = 0.0;
            for (int32_t k = l; k < r; k++) {
                const double dx = F->sx[2 * k] - vx, dy = F->sx[2 * k + 1] - vy;
                const double q2 = dx * dx + dy * dy;
                if (q2 > 0) { ex += dx / q2; ey += dy / q2; }
            }
            double *f = F->feat + 8 * F->cnt;
            far_features(m, D, rx, ry, mxx, mxy, myy, f);
            const double rho = sqrt(r2), ux = rx / rho, uy = ry / rho;
            const double sc = rho / m;
            F->targ[2 * F->cnt] = (ex * ux + ey * uy) * sc - 1.0;
            F->targ[2 * F->cnt + 1] = (-ex * uy + ey * ux) * sc;
            F->cell[2 * F->cnt] = l;
            F->cell[2 * F->cnt + 1] = r - l;
            F->cnt++;
        }
        *fx += m * rx / r2;
        *fy += m * ry / r2;
        return;
    }
    int32_t b[5];
    b[0] = l;
    b[4] = r;
    for (int q = 1; q < 4; q++) b[q] = lower_q(F, l, r, d, q);
    for (int q = 0; q < 4; q++) {
        if (b[q + 1] - b[q] >= 2) {
            far_rec(F, b[q], b[q + 1], vx, vy, mode, fx, fy);
        } else if (b[q + 1] - b[q] == 1) {
            const double dx = F->sx[2 * b[q]] - vx, dy = F->sx[2 * b[q] + 1] - vy;
            const double q2 = dx * dx + dy * dy;
            if (q2 > 0) { *fx += dx / q2; *fy += dy / q2; }
        }
    }
}

static int far_prepare(int32_t n, const double *xy, uint64_t **key, double **sx, double *maxmin) {
    frame_t fr = frame_of(n, xy);
    kv_t *kv = (kv_t *)malloc(sizeof(kv_t) * (size_t)n);
    *key = (uint64_t *)malloc(8 * (size_t)n);
    *sx = (double *)malloc(16 * (size_t)n);
    for (int32_t i = 0; i < n; i++) {
        kv[i].key = morton_key(scale_of(xy[2 * i], fr.minx, fr.maxmin), scale_of(xy[2 * i + 1], fr.miny, fr.maxmin));
        kv[i].idx = i;
    }
    qsort(kv, (size_t)n, sizeof(kv_t), kv_cmp);
    int dup = 0;
    for (int32_t k = 0; k < n; k++) {
        (*key)[k] = kv[k].key;
        (*sx)[2 * k] = xy[2 * kv[k].idx];
        (*sx)[2 * k + 1] = xy[2 * kv[k].idx + 1];
        if (k > 0 && kv[k].key == kv[k - 1].key) dup = 1;
    }
    free(kv);
    *maxmin = fr.maxmin;
    return dup ? ORC_EDUPKEY : ORC_OK;
}

int orc_far_forces(int32_t n, const double *xy, double theta, int32_t nq, const double *q, double *out,
                   int threads) {
    if (n < 1 || nq < 0) return ORC_EARG;
    uint64_t *key;
    double *sx, maxmin;
    int rc = far_prepare(n, xy, &key, &sx, &maxmin);
    if (rc == ORC_OK) {
        if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 16)
        for (int32_t a = 0; a < nq; a++) {
            far_t F = {key, sx, maxmin, theta * theta, 0, 0, NULL, NULL, NULL};
            double fx = 0.0, fy = 0.0;
            if (n >= 2) {
                far_rec(&F, 0, n, q[2 * a], q[2 * a + 1], 0, &fx, &fy);
            } else {
                const double dx = sx[0] - q[2 * a], dy = sx[1] - q[2 * a + 1];
                const double q2 = dx * dx + dy * dy;
                if (q2 > 0) { fx = dx / q2; fy = dy / q2; }
            }
            out[2 * a] = fx;
            out[2 * a + 1] = fy;
        }
    }
    free(key);
    free(sx);
    return rc;
}

int orc_far_samples(int32_t n, const double *xy, double theta, double qx, double qy, int32_t cap, double *feat,
                    double *targ, int32_t *cell, int32_t *count, double *force) {
    if (n < 2 || cap < 0) return ORC_EARG;
    uint64_t *key;
    double *sx, maxmin;
    int rc = far_prepare(n, xy, &key, &sx, &maxmin);
    if (rc == ORC_OK) {
        far_t F = {key, sx, maxmin, theta * theta, cap, 0, feat, targ, cell};
        double fx = 0.0, fy = 0.0;
        far_rec(&F, 0, n, qx, qy, 1, &fx, &fy);
        *count = F.cnt;
        if (force) { force[0] = fx; force[1] = fy; }
    }
    free(key);
    free(sx);
    return rc;
}

/* Layout with the physically consistent repulsion: the algo-4 iteration (repulsion from the
 * iteration-start positions, attraction d * W|d| over every CSR entry with live positions,
 * sequential batches, STEP *= 0.999) with orc_far_forces in place of calcRepForce. */
int orc_far_layout(int32_t n, const int32_t *rowptr, const int32_t *colids, const float *w_edge, double *xy,
                   int iters, int batch, double theta, double weight, double step0, int threads, double *energy) {
    if (n <= 0 || batch <= 0) return ORC_EARG;
    if (threads > 0) omp_set_num_threads(threads);
    double *rep = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *prev = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *start = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double STEP = step0, E = 0.0;
    const int nb = (int)ceil(1.0 * n / batch);
    int rc = ORC_OK;
    for (int loop = 0; loop < iters && rc == ORC_OK; loop++) {
        E = 0.0;
        memcpy(start, xy, sizeof(double) * 2 * (size_t)n);
        rc = orc_far_forces(n, start, theta, n, start, rep, threads);
        if (rc != ORC_OK) break;
        for (int b = 0; b < nb; b++) {
            const int64_t i0 = (int64_t)b * batch, i1 = (int64_t)(b + 1) * batch;
#pragma omp parallel for schedule(static)
            for (int64_t i = i0; i < i1; i++) {
                if (i >= n) continue;
                double fx = 0.0, fy = 0.0;
                const double vx = xy[2 * i], vy = xy[2 * i + 1];
                for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
                    const int32_t j = colids[k];
                    const double dx = xy[2 * j] - vx, dy = xy[2 * j + 1] - vy;
                    const double W = weight * (w_edge ? (double)w_edge[k] : 1.0);
                    const double s = W * sqrt(dx * dx + dy * dy);
                    fx = fx + dx * s;
                    fy = fy + dy * s;
                }
                prev[2 * i] = fx - rep[2 * i];
                prev[2 * i + 1] = fy - rep[2 * i + 1];
            }
            for (int64_t i = i0; i < i1 && i < n; i++) {
                const double fx = prev[2 * i], fy = prev[2 * i + 1];
                const double mag = sqrt(fx * fx + fy * fy);
                xy[2 * i] = xy[2 * i] + (fx / mag) * STEP;
                xy[2 * i + 1] = xy[2 * i + 1] + (fy / mag) * STEP;
                E += fx * fx + fy * fy;
            }
        }
        STEP = STEP * 0.999;
    }
    free(rep);
    free(prev);
    free(start);
    if (energy) *energy = E;
    return rc;
}
