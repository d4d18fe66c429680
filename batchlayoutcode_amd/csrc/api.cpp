// C ABI of libbatchlayout.so: graph handles, parameter defaults, run dispatch.
// Replaces the reference's algorithms constructor (sample/algorithms.cpp:3-15) and the
// per-algorithm methods (sample/algorithms.h:67-80); see include/batchlayout.h.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "bh_common.h"
#include "mlp.h"

namespace bl {

int device_error(hipError_t e, const char *what, int line) {
    fprintf(stderr, "batchlayout: HIP error '%s' (%d) in %s [line %d]\n", hipGetErrorString(e), (int)e, what, line);
    return BL_E_DEVICE;
}

int DevBuf::ensure(size_t want) {
    if (want <= bytes && p) return 0;
    release();
    if (want == 0) return 0;
    if (hipMalloc(&p, want) != hipSuccess) {
        p = nullptr;
        bytes = 0;
        return BL_E_NOMEM;
    }
    bytes = want;
    return 0;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

void BHWork::release() {
    DevBuf *all[] = {&keys_a, &keys_b, &idx_a, &idx_b, &hist, &spos, &frame, &isrep, &nodeid,
                     &node_child, &node_cnt, &node_lo, &scan_tmp,
                     &rep, &misc, &far_cells, &far_nlc, &judge, &gwalk};
    for (DevBuf *b : all) b->release();
    n_alloc = 0;
}

// The code objects are built for gfx950 only: refuse anything else up front.
static int check_device_once() {
    static std::once_flag once;
    static int status = BL_E_DEVICE;
    std::call_once(once, [] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
            fprintf(stderr, "batchlayout: no HIP device visible\n");
            return;
        }
        hipDeviceProp_t prop;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return;
        if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            fprintf(stderr, "batchlayout: device %d is %s; this build targets gfx950 only\n", dev, prop.gcnArchName);
            return;
        }
        status = BL_OK;
    });
    return status;
}

}  // namespace bl

using namespace bl;

extern "C" int bl_params_default(bl_params *p) {
    if (!p) return BL_E_ARG;
    memset(p, 0, sizeof(*p));
    p->algo = BL_ALGO_CACHE;
    p->iters = 600;
    p->batch = 256;
    p->converged = 0;
    p->bht = 1.2;
    p->engt = 0.01;
    p->weight = 1.0;
    p->step0 = 1.0;
    p->precision = BL_PREC_MIXED;
    return BL_OK;
}

extern "C" int bl_graph_create(int32_t n, int64_t nnz, const int32_t *rowptr, const int32_t *colids,
                               const float *w_edge, bl_graph **out) {
    if (!out) return BL_E_ARG;
    *out = nullptr;
    if (n <= 0 || nnz < 0 || !rowptr || (nnz > 0 && !colids)) return BL_E_ARG;
    if (nnz > INT32_MAX) return BL_E_ARG;
    // CSR validation (the reference trusts its own builder; the ABI does not)
    if (rowptr[0] != 0 || rowptr[n] != nnz) return BL_E_GRAPH;
    for (int32_t i = 0; i < n; ++i)
        if (rowptr[i + 1] < rowptr[i]) return BL_E_GRAPH;
    for (int64_t k = 0; k < nnz; ++k)
        if (colids[k] < 0 || colids[k] >= n) return BL_E_GRAPH;
    int rc = check_device_once();
    if (rc) return rc;

    std::vector<int32_t> erp(n + 1, 0), eci;
    std::vector<float> ew;
    eci.reserve((size_t)nnz);
    int32_t maxdeg = 0;
    effective_neighbors(rowptr, colids, w_edge, 0, n, nnz, erp, eci, ew, &maxdeg);

    bl_graph *g = new (std::nothrow) bl_graph();
    if (!g) return BL_E_NOMEM;
    g->n = n;
    g->nnz = nnz;
    g->has_w = w_edge != nullptr;
    g->eff_max_deg = maxdeg;
    g->h_rowptr.assign(rowptr, rowptr + n + 1);
    g->h_colids.assign(colids, colids + nnz);
    if (w_edge) g->h_w.assign(w_edge, w_edge + nnz);
    for (int32_t i = 0; i < n; ++i) g->max_row = std::max(g->max_row, rowptr[i + 1] - rowptr[i]);
    auto fail = [&](int code) {
        bl_graph_destroy(g);
        return code;
    };
    if (hipGetDevice(&g->device) != hipSuccess) return fail(BL_E_DEVICE);
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) return fail(BL_E_DEVICE);
    const size_t nnzb = (size_t)(nnz > 0 ? nnz : 1);
    const size_t effb = eci.empty() ? 1 : eci.size();
    if (g->rowptr.ensure(sizeof(int32_t) * (n + 1)) || g->colids.ensure(sizeof(int32_t) * nnzb) ||
        g->eff_rowptr.ensure(sizeof(int32_t) * (n + 1)) || g->eff_colids.ensure(sizeof(int32_t) * effb))
        return fail(BL_E_NOMEM);
    if (w_edge && (g->w_edge.ensure(sizeof(float) * nnzb) || g->eff_w.ensure(sizeof(float) * effb)))
        return fail(BL_E_NOMEM);
    hipStream_t s = g->stream;
    if (hipMemcpyAsync(g->rowptr.p, rowptr, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
        (nnz > 0 && hipMemcpyAsync(g->colids.p, colids, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s) != hipSuccess) ||
        hipMemcpyAsync(g->eff_rowptr.p, erp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
        (!eci.empty() && hipMemcpyAsync(g->eff_colids.p, eci.data(), sizeof(int32_t) * eci.size(),
                                        hipMemcpyHostToDevice, s) != hipSuccess))
        return fail(BL_E_DEVICE);
    if (w_edge && ((nnz > 0 && hipMemcpyAsync(g->w_edge.p, w_edge, sizeof(float) * nnz, hipMemcpyHostToDevice, s) !=
                                   hipSuccess) ||
                   (!ew.empty() &&
                    hipMemcpyAsync(g->eff_w.p, ew.data(), sizeof(float) * ew.size(), hipMemcpyHostToDevice, s) !=
                        hipSuccess)))
        return fail(BL_E_DEVICE);
    // correction pieces (one wave each in the all-pairs step)
    g->h_cpre.assign((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i) {
        const int32_t np = std::max(1, (erp[i + 1] - erp[i] + CORR_PIECE - 1) / CORR_PIECE);
        g->h_cpre[i + 1] = g->h_cpre[i] + np;
        g->corr_max_pieces = std::max(g->corr_max_pieces, np);
    }
    {
        std::vector<int32_t> pv((size_t)g->h_cpre[n]);
        for (int32_t i = 0; i < n; ++i)
            for (int32_t q = g->h_cpre[i]; q < g->h_cpre[i + 1]; ++q) pv[q] = i;
        if (g->cpre.ensure(sizeof(int32_t) * (n + 1)) || g->cpv.ensure(sizeof(int32_t) * pv.size()))
            return fail(BL_E_NOMEM);
        if (hipMemcpyAsync(g->cpre.p, g->h_cpre.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s) !=
                hipSuccess ||
            hipMemcpyAsync(g->cpv.p, pv.data(), sizeof(int32_t) * pv.size(), hipMemcpyHostToDevice, s) != hipSuccess)
            return fail(BL_E_DEVICE);
    }
    if (hipStreamSynchronize(s) != hipSuccess) return fail(BL_E_DEVICE);
    *out = g;
    return BL_OK;
}

extern "C" int bl_graph_destroy(bl_graph *g) {
    if (!g) return BL_E_ARG;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    DevBuf *all[] = {&g->rowptr, &g->colids, &g->w_edge, &g->eff_rowptr, &g->eff_colids, &g->eff_w,
                     &g->pos, &g->part, &g->eterm, &g->energy, &g->scratch, &g->guard, &g->far_model, &g->cpre, &g->cpv,
                     &g->early_rp, &g->early_ci, &g->early_w, &g->late_rp, &g->late_ci, &g->late_w,
                     &g->grid_rows, &g->lag_erp, &g->lag_esp, &g->lag_eci, &g->lag_ew, &g->lag_rofs, &g->lag_rw,
                     &g->lag_pre, &g->lag_ready, &g->lag_ring_of, &g->lag_gstart, &g->lag_sink, &g->wide_near_rp, &g->wide_near_code, &g->wide_near_w,
                     &g->wide_npa, &g->wide_npb, &g->wide_ring_of, &g->wide_far_rp, &g->wide_far_ci, &g->wide_far_w,
                     &g->wide_chunks, &g->wide_cfr, &g->wide_nch, &g->wide_pre, &g->wide_cnt, &g->wide_sink};
    for (DevBuf *b : all) b->release();
    g->bh.release();
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return BL_OK;
}

extern "C" int bl_graph_size(const bl_graph *g, int32_t *n, int64_t *nnz) {
    if (!g) return BL_E_ARG;
    if (n) *n = g->n;
    if (nnz) *nnz = g->nnz;
    return BL_OK;
}

static int validate_params(const bl_params *p) {
    if (!p) return BL_E_ARG;
    if (p->batch <= 0 || p->iters < 0) return BL_E_ARG;
    if (p->precision != BL_PREC_MIXED && p->precision != BL_PREC_FP64 && p->precision != BL_PREC_FP32)
        return BL_E_ARG;
    switch (p->algo) {
        case BL_ALGO_CACHE: case BL_ALGO_LINLOG: case BL_ALGO_BH: case BL_ALGO_BHCACHE: case BL_ALGO_FA:
        case BL_ALGO_FAR_MONO: case BL_ALGO_FAR_EXACT: case BL_ALGO_FAR_LEARNED: case BL_ALGO_FAR_QUAD:
            return BL_OK;
        default:
            return BL_E_ARG;
    }
}

// Runs one layout on device-resident positions.  Dispatch mirrors Test/BatchLayout.cpp:145-170.
static int run_on_device(bl_graph *g, double *d_xy, const bl_params *p, bl_stats *st) {
    bl_stats local;
    memset(&local, 0, sizeof(local));
    int rc = BL_OK;
    switch (p->algo) {
        case BL_ALGO_CACHE:
            rc = run_minibatch(g, d_xy, p, LAW_FR, 0.999, p->step0, p->iters, p->converged != 0, &local);
            break;
        case BL_ALGO_LINLOG:
            rc = run_minibatch(g, d_xy, p, LAW_LINLOG, 0.99, p->step0, p->iters, false, &local);
            break;
        case BL_ALGO_FA:
            rc = run_minibatch(g, d_xy, p, LAW_FA, 0.99, p->step0, p->iters, false, &local);
            break;
        case BL_ALGO_BH:
            rc = run_bh(g, d_xy, p, p->step0, p->iters, p->bht, &local);
            break;
        case BL_ALGO_BHCACHE: {
            // approxCacheBlockBH (algorithms.cpp:1192-1215): (int)(0.8*iters) BH iterations at
            // theta 1.2 (the -bht flag is ignored), then the rest as algo 2 starting from
            // STEP = pow(0.999, 4*rest) (the flag == 1 branch, :736-737).
            const int approx = (int)(p->iters * 0.8);
            rc = run_bh(g, d_xy, p, p->step0, approx, 1.2, &local);
            if (rc == BL_OK) {
                const int rest = p->iters - approx;
                rc = run_minibatch(g, d_xy, p, LAW_FR, 0.999, pow(0.999, 4 * rest), rest, false, &local);
                local.iters_done = p->iters;
            }
            break;
        }
        case BL_ALGO_FAR_MONO:
            rc = run_bh(g, d_xy, p, p->step0, p->iters, p->bht, &local, FAR_MONO);
            break;
        case BL_ALGO_FAR_EXACT:
            rc = run_bh(g, d_xy, p, p->step0, p->iters, 0.0, &local, FAR_EXACT);
            break;
        case BL_ALGO_FAR_QUAD:
            rc = run_bh(g, d_xy, p, p->step0, p->iters, p->bht, &local, FAR_QUAD);
            break;
        case BL_ALGO_FAR_LEARNED:
            rc = g->has_far_model ? run_bh(g, d_xy, p, p->step0, p->iters, p->bht, &local, FAR_LEARNED) : BL_E_ARG;
            break;
        default:
            rc = BL_E_ARG;
    }
    if (rc == BL_OK) {  // the non-finite guard
        if (g->guard.ensure(64)) {
            rc = BL_E_NOMEM;
        } else {
            int nf = 0;
            rc = count_nonfinite(reinterpret_cast<const double2 *>(d_xy), g->n, g->guard.as<int>(), g->stream, &nf);
            local.nonfinite = nf;
        }
    }
    local.status = rc;
    if (st) *st = local;
    return rc;
}

extern "C" int bl_run_device(bl_graph *g, double *d_xy, const bl_params *p, bl_stats *st, void *stream) {
    if (!g || !d_xy) return BL_E_ARG;
    int rc = validate_params(p);
    if (rc) return rc;
    hipStream_t own = g->stream;
    if (stream) g->stream = static_cast<hipStream_t>(stream);
    const auto t0 = std::chrono::steady_clock::now();
    rc = run_on_device(g, d_xy, p, st);
    if (hipStreamSynchronize(g->stream) != hipSuccess && rc == BL_OK) rc = BL_E_DEVICE;
    const auto t1 = std::chrono::steady_clock::now();
    g->stream = own;
    if (st) {
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->status = rc;
    }
    return rc;
}

// The minibatch family with its workspace in the static device array: no allocation, no
// profiling events (both are refused up front rather than done quietly).
extern "C" int bl_run_device_static(bl_graph *g, double *d_xy, const bl_params *p, bl_stats *st, void *stream) {
    if (!g || !d_xy) return BL_E_ARG;
    int rc = validate_params(p);
    if (rc) return rc;
    if (p->profile) return BL_E_ARG;
    int law;
    double decay;
    switch (p->algo) {
        case BL_ALGO_CACHE: law = LAW_FR; decay = 0.999; break;
        case BL_ALGO_LINLOG: law = LAW_LINLOG; decay = 0.99; break;
        case BL_ALGO_FA: law = LAW_FA; decay = 0.99; break;
        default: return BL_E_ARG;
    }
    bl_stats local;
    memset(&local, 0, sizeof(local));
    hipStream_t own = g->stream;
    if (stream) g->stream = static_cast<hipStream_t>(stream);
    const auto t0 = std::chrono::steady_clock::now();
    rc = run_minibatch(g, d_xy, p, law, decay, p->step0, p->iters, p->algo == BL_ALGO_CACHE && p->converged != 0,
                       &local, true);
    if (hipStreamSynchronize(g->stream) != hipSuccess && rc == BL_OK) rc = BL_E_DEVICE;
    const auto t1 = std::chrono::steady_clock::now();
    g->stream = own;
    local.seconds = std::chrono::duration<double>(t1 - t0).count();
    local.status = rc;
    if (st) *st = local;
    return rc;
}

extern "C" int bl_run(bl_graph *g, double *xy_inout, const bl_params *p, bl_stats *st) {
    if (!g || !xy_inout) return BL_E_ARG;
    int rc = validate_params(p);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * 2 * (size_t)g->n;
    if (g->pos.ensure(bytes)) return BL_E_NOMEM;
    BL_HIP(hipMemcpyAsync(g->pos.p, xy_inout, bytes, hipMemcpyHostToDevice, g->stream));
    const auto t0 = std::chrono::steady_clock::now();
    rc = run_on_device(g, g->pos.as<double>(), p, st);
    if (hipStreamSynchronize(g->stream) != hipSuccess && rc == BL_OK) rc = BL_E_DEVICE;
    const auto t1 = std::chrono::steady_clock::now();
    if (rc == BL_OK) {
        BL_HIP(hipMemcpyAsync(xy_inout, g->pos.p, bytes, hipMemcpyDeviceToHost, g->stream));
        BL_HIP(hipStreamSynchronize(g->stream));
    }
    if (st) {
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        st->status = rc;
    }
    return rc;
}

extern "C" int bl_bh_morton(int32_t n, const double *xy, uint32_t *xs, uint32_t *ys, double *maxmin) {
    if (n <= 0 || !xy || !xs || !ys) return BL_E_ARG;
    int rc = check_device_once();
    if (rc) return rc;
    return bh_morton_host(n, xy, xs, ys, maxmin);
}

extern "C" int bl_bh_tree(int32_t n, const double *xy, int32_t *perm, int32_t max_nodes, int32_t *nnodes,
                          int32_t *root, double *diam, double *cx, double *cy, int32_t *child) {
    if (n <= 0 || !xy || !perm || !nnodes || !root) return BL_E_ARG;
    int rc = check_device_once();
    if (rc) return rc;
    return bh_tree_host(n, xy, perm, max_nodes, nnodes, root, diam, cx, cy, child);
}

extern "C" int bl_bh_forces(int32_t n, const double *xy, double theta, int32_t nq, const double *q, double *out,
                            int32_t precision) {
    if (n <= 0 || !xy || nq < 0 || (nq > 0 && (!q || !out))) return BL_E_ARG;
    int rc = check_device_once();
    if (rc) return rc;
    return bh_forces_host(n, xy, theta, nq, q, out, precision);
}

static int far_mode_of(int32_t mode) {
    switch (mode) {
        case BL_FAR_MONO: return FAR_MONO;
        case BL_FAR_EXACT: return FAR_EXACT;
        case BL_FAR_QUAD: return FAR_QUAD;
        case BL_FAR_LEARNED: return FAR_LEARNED;
        default: return 0;
    }
}

extern "C" int bl_far_forces(int32_t n, const double *xy, double theta, int32_t mode, const float *model, int32_t nq,
                             const double *q, double *out) {
    if (n <= 0 || !xy || nq < 0 || (nq > 0 && (!q || !out)) || theta < 0.0) return BL_E_ARG;
    const int fm = far_mode_of(mode);
    if (!fm || (fm == FAR_LEARNED && !model)) return BL_E_ARG;
    int rc = check_device_once();
    if (rc) return rc;
    return far_forces_host(n, xy, theta, fm, model, nq, q, out);
}

extern "C" int bl_far_samples(int32_t n, const double *xy, double theta, int32_t nq, const double *q, int32_t cap,
                              float *feat, float *targ, int32_t *count) {
    if (n < 2 || !xy || nq < 0 || cap < 1 || theta <= 0.0 || (nq > 0 && (!q || !feat || !targ || !count)))
        return BL_E_ARG;
    int rc = check_device_once();
    if (rc) return rc;
    return far_samples_host(n, xy, theta, nq, q, cap, feat, targ, count);
}

// Glorot-uniform weights from xorshift64* (seeded), zero biases; the output layer scaled by
// 0.1 so that an untrained model starts near the plain monopole.
extern "C" int bl_far_model_init(float *params, uint64_t seed) {
    if (!params) return BL_E_ARG;
    uint64_t st = seed ? seed : 0x9E3779B97F4A7C15ull;
    auto unif = [&]() {
        st ^= st >> 12;
        st ^= st << 25;
        st ^= st >> 27;
        return (double)((st * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int rows, int cols, double gain) {
        const double a = gain * sqrt(6.0 / (double)(rows + cols));
        for (int k = 0; k < rows * cols; ++k) params[off + k] = (float)((2.0 * unif() - 1.0) * a);
    };
    memset(params, 0, sizeof(float) * BL_FAR_MODEL_PARAMS);
    fill(MLP_W1, MLP_H, MLP_IN, 1.0);
    fill(MLP_W2, MLP_H, MLP_H, 1.0);
    fill(MLP_W3, MLP_H, MLP_H, 1.0);
    fill(MLP_W4, MLP_OUT, MLP_H, 0.1);
    return BL_OK;
}

extern "C" int bl_far_model_eval(const float *params, int32_t n, const float *feat, float *out) {
    if (!params || n < 0 || (n > 0 && (!feat || !out))) return BL_E_ARG;
    if (n == 0) return BL_OK;
    int rc = check_device_once();
    if (rc) return rc;
    return far_model_eval_host(params, n, feat, out);
}

extern "C" int bl_far_model_grad(const float *params, int32_t n, const float *feat, const float *targ, float *loss,
                                 float *grad) {
    if (!params || n <= 0 || !feat || !targ) return BL_E_ARG;
    int rc = check_device_once();
    if (rc) return rc;
    return far_model_grad_host(params, n, feat, targ, loss, grad);
}

extern "C" int bl_far_model_train(float *params, int32_t n, const float *feat, const float *targ, int32_t epochs,
                                  int32_t batch, float lr, uint64_t seed, float *losses) {
    if (!params || n <= 0 || !feat || !targ || epochs < 0 || batch <= 0 || !(lr > 0.0f)) return BL_E_ARG;
    int rc = check_device_once();
    if (rc) return rc;
    return far_model_train_host(params, n, feat, targ, epochs, batch, lr, seed, losses);
}

extern "C" int bl_graph_set_far_model(bl_graph *g, const float *params) {
    if (!g) return BL_E_ARG;
    if (!params) {
        g->has_far_model = false;
        return BL_OK;
    }
    if (g->far_model.ensure(sizeof(float) * BL_FAR_MODEL_PARAMS)) return BL_E_NOMEM;
    if (hipMemcpy(g->far_model.p, params, sizeof(float) * BL_FAR_MODEL_PARAMS, hipMemcpyHostToDevice) != hipSuccess)
        return BL_E_DEVICE;
    g->has_far_model = true;
    return BL_OK;
}

extern "C" int bl_run_batched(int32_t count, const int32_t *voff, const int64_t *eoff, const int32_t *rowptr,
                              const int32_t *colids, double *xy_inout, const bl_params *p, double *energies,
                              double *finish_ms, bl_stats *st) {
    if (count <= 0 || !voff || !eoff || !rowptr || !xy_inout || !p) return BL_E_ARG;
    int rc = validate_params(p);
    if (rc) return rc;
    if (p->algo != BL_ALGO_CACHE && p->algo != BL_ALGO_FA) return BL_E_ARG;
    if (p->precision == BL_PREC_FP64 || p->converged) return BL_E_ARG;
    if (voff[0] != 0 || eoff[0] != 0) return BL_E_ARG;
    for (int32_t k = 0; k < count; ++k)  // offsets first, so the loop below reads only inside the arrays
        if (voff[k + 1] <= voff[k] || voff[k + 1] - voff[k] > BL_BATCHED_MAX_N || eoff[k + 1] < eoff[k])
            return BL_E_ARG;
    if (eoff[count] > INT32_MAX || (eoff[count] > 0 && !colids)) return BL_E_ARG;
    for (int32_t k = 0; k < count; ++k) {  // per-graph CSR validation
        const int32_t n = voff[k + 1] - voff[k];
        if (n <= 0 || n > BL_BATCHED_MAX_N) return BL_E_ARG;
        if (rowptr[voff[k]] != eoff[k] || rowptr[voff[k + 1]] != eoff[k + 1]) return BL_E_GRAPH;
        for (int32_t i = voff[k]; i < voff[k + 1]; ++i)
            if (rowptr[i + 1] < rowptr[i]) return BL_E_GRAPH;
        for (int64_t e = eoff[k]; e < eoff[k + 1]; ++e)
            if (colids[e] < 0 || colids[e] >= n) return BL_E_GRAPH;
    }
    rc = check_device_once();
    if (rc) return rc;
    return run_batched(count, voff, eoff, rowptr, colids, xy_inout, p, energies, finish_ms, st);
}

extern "C" int bl_device_info(char *buf, int32_t len) {
    if (!buf || len <= 0) return BL_E_ARG;
    int dev = 0, count = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || hipGetDevice(&dev) != hipSuccess ||
        hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        snprintf(buf, (size_t)len, "no HIP device");
        return BL_E_DEVICE;
    }
    snprintf(buf, (size_t)len, "%s %s CUs=%d clock=%dkHz mem=%zuMB", prop.name, prop.gcnArchName,
             prop.multiProcessorCount, prop.clockRate, (size_t)(prop.totalGlobalMem >> 20));
    return BL_OK;
}
