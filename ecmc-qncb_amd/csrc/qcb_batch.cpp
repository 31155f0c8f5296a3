// qcb_batch.cpp -- many independent lattices (uneven sizes, own beta and
// Philox key each) advanced by one launch per (mu, parity) half-sweep, per RK3
// stage and per reduction; a tuned launch schedule; hipGraph replay of sweeps.
//
// The heatbath can split the lattices into groups, each with its own
// descriptor table and stream: a group's half-sweeps depend only on its own
// earlier half-sweeps, so while one group's launch drains its last blocks
// the next group's launch fills the idle compute units.  The groups fork from
// the context stream and join back into it through events, which a graph
// capture records as parallel branches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/qcb.h"
#include "qcb_host.hpp"
#include "qcb_internal.hpp"

using namespace qcb;

int enqueue_flow_step(qcb_lattice* lat, double eps);

struct qcb_batch {
    qcb_ctx* ctx = nullptr;
    int prec = QCB_F64;
    int n = 0;
    std::vector<qcb_lattice*> lats;
    void* d_desc = nullptr;
    uint32_t* d_blk[kNumFam] = {};
    uint32_t nb[kNumFam] = {};
    double* d_partial = nullptr;     // 2 * nb[kFamClover] (>= 2 * nb[kFamSite])
    double* d_meas = nullptr;        // measurement buffer (grown on demand)
    size_t meas_cap = 0;
    uint32_t* d_sweep = nullptr;     // device sweep counter for graph replay
    unsigned long long* d_counter = nullptr;   // n exhausted-draw counts, one per lattice
    // launch schedule (autotuned)
    int order[8] = {0, 1, 2, 3, 4, 5, 6, 7};   // half-sweep order, hs = 2*mu + parity
    uint32_t hb_grid = 0;                       // 0: one block per tile; else persistent grid
    // stream groups of the heatbath: lattices grp_idx[g], their descriptors and
    // block prefixes on the device, and the stream of each group (group 0 runs
    // on the context stream)
    int n_groups = 1;
    std::vector<std::vector<int>> grp_idx;
    std::vector<void*> grp_desc;
    std::vector<uint32_t*> grp_blk;
    std::vector<uint32_t> grp_nb;
    std::vector<hipStream_t> grp_stream;
    std::vector<hipEvent_t> grp_event;   // [0] fork, [g] join of group g
    int tile_caps[3] = {8, 4, 4};        // work-order tile caps in (y, z, t)
    // the flow's launch sequence, captured at its first call and replayed while
    // eps, the step count and every buffer address stay the same
    struct FlowGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        double eps = 0.0;
        int n_steps = -1;
        std::vector<const void*> key;
    } flow_graph[2];                     // [0] measured flow, [1] steps only
    // captured sweeps
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_sweeps = 0;
};

namespace {

template <class T>
int upload_groups(qcb_batch* b) {
    qcb_ctx* c = b->ctx;
    for (size_t g = 0; g < b->grp_idx.size(); ++g) {
        std::vector<LatDev<T>> d;
        for (int i : b->grp_idx[g]) d.push_back(make_desc<T>(b->lats[i]));
        QCB_HIP(c, hipMemcpyAsync(b->grp_desc[g], d.data(), d.size() * sizeof(d[0]), hipMemcpyHostToDevice, c->stream));
        QCB_HIP(c, hipStreamSynchronize(c->stream));
    }
    return QCB_OK;
}

int batch_upload(qcb_batch* b) {
    qcb_ctx* c = b->ctx;
    if (b->n_groups > 1) {
        const int rc = b->prec == QCB_F64 ? upload_groups<double>(b) : upload_groups<float>(b);
        if (rc) return rc;
    }
    if (b->prec == QCB_F64) {
        std::vector<LatDev<double>> d(b->n);
        for (int i = 0; i < b->n; ++i) d[i] = make_desc<double>(b->lats[i]);
        QCB_HIP(c, hipMemcpyAsync(b->d_desc, d.data(), d.size() * sizeof(d[0]), hipMemcpyHostToDevice, c->stream));
        QCB_HIP(c, hipStreamSynchronize(c->stream));
    } else {
        std::vector<LatDev<float>> d(b->n);
        for (int i = 0; i < b->n; ++i) d[i] = make_desc<float>(b->lats[i]);
        QCB_HIP(c, hipMemcpyAsync(b->d_desc, d.data(), d.size() * sizeof(d[0]), hipMemcpyHostToDevice, c->stream));
        QCB_HIP(c, hipStreamSynchronize(c->stream));
    }
    return QCB_OK;
}

Launch batch_launch(const qcb_batch* b, int fam) {
    Launch g;
    g.lats = b->d_desc;
    g.blk = b->d_blk[fam];
    g.n_lat = b->n;
    g.nblocks = b->nb[fam];
    g.stream = b->ctx->stream;
    return g;
}

void free_groups(qcb_batch* b) {
    for (void* d : b->grp_desc) (void)hipFree(d);
    for (uint32_t* d : b->grp_blk) (void)hipFree(d);
    for (size_t g = 1; g < b->grp_stream.size(); ++g) (void)hipStreamDestroy(b->grp_stream[g]);
    for (hipEvent_t e : b->grp_event) (void)hipEventDestroy(e);
    b->grp_desc.clear();
    b->grp_blk.clear();
    b->grp_stream.clear();
    b->grp_event.clear();
    b->grp_nb.clear();
    b->grp_idx.clear();
}

void free_flow_graphs(qcb_batch* b) {
    for (auto& fg : b->flow_graph) {
        if (fg.exec) (void)hipGraphExecDestroy(fg.exec);
        if (fg.graph) (void)hipGraphDestroy(fg.graph);
        fg = qcb_batch::FlowGraph{};
    }
}

void batch_free(qcb_batch* b) {
    free_groups(b);
    free_flow_graphs(b);
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (qcb_lattice* l : b->lats)
        if (l) qcb_lattice_destroy(l);
    (void)hipFree(b->d_desc);
    for (int f = 0; f < kNumFam; ++f) (void)hipFree(b->d_blk[f]);
    (void)hipFree(b->d_partial);
    (void)hipFree(b->d_meas);
    (void)hipFree(b->d_sweep);
    (void)hipFree(b->d_counter);
}

int ensure_meas(qcb_batch* b, size_t n) {
    if (n <= b->meas_cap) return QCB_OK;
    (void)hipFree(b->d_meas);
    b->d_meas = nullptr;
    QCB_HIP(b->ctx, hipMalloc((void**)&b->d_meas, n * sizeof(double)));
    b->meas_cap = n;
    return QCB_OK;
}

// Splits the lattices into n groups of about equal work (greedy by size) with
// their own descriptor tables and streams.
int set_groups(qcb_batch* b, int n) {
    qcb_ctx* c = b->ctx;
    QCB_HIP(c, hipStreamSynchronize(c->stream));
    free_groups(b);
    n = std::max(1, std::min(n, b->n));
    b->n_groups = n;
    if (n == 1) return QCB_OK;
    std::vector<int> ids(b->n);
    for (int i = 0; i < b->n; ++i) ids[i] = i;
    std::stable_sort(ids.begin(), ids.end(), [&](int x, int y) { return b->lats[x]->Vh > b->lats[y]->Vh; });
    b->grp_idx.assign(n, {});
    std::vector<uint64_t> load(n, 0);
    for (int i : ids) {
        const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->grp_idx[g].push_back(i);
        load[g] += b->lats[i]->Vh;
    }
    for (auto& v : b->grp_idx) std::sort(v.begin(), v.end());
    b->grp_desc.assign(n, nullptr);
    b->grp_blk.assign(n, nullptr);
    b->grp_nb.assign(n, 0);
    b->grp_stream.assign(n, nullptr);
    b->grp_event.assign(n, nullptr);
    b->grp_stream[0] = c->stream;
    for (int g = 0; g < n; ++g) {
        const auto& idx = b->grp_idx[g];
        QCB_HIP(c, hipMalloc(&b->grp_desc[g], idx.size() * sizeof(LatDev<double>)));
        std::vector<uint32_t> pre(idx.size() + 1, 0);
        for (size_t k = 0; k < idx.size(); ++k) pre[k + 1] = pre[k] + fam_blocks(*b->lats[idx[k]], kFamHalf);
        b->grp_nb[g] = pre.back();
        QCB_HIP(c, hipMalloc((void**)&b->grp_blk[g], pre.size() * sizeof(uint32_t)));
        QCB_HIP(c, hipMemcpy(b->grp_blk[g], pre.data(), pre.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        if (g > 0) QCB_HIP(c, hipStreamCreateWithFlags(&b->grp_stream[g], hipStreamNonBlocking));
        QCB_HIP(c, hipEventCreateWithFlags(&b->grp_event[g], hipEventDisableTiming));
    }
    return b->prec == QCB_F64 ? upload_groups<double>(b) : upload_groups<float>(b);
}

// Work-order tiles of every lattice from the caps (largest divisor of each
// extent not above its cap), visited y-fastest.
int apply_tiles(qcb_batch* b) {
    for (qcb_lattice* l : b->lats) {
        int t[4] = {0, divisor_le(l->L[1], b->tile_caps[0]), divisor_le(l->L[2], b->tile_caps[1]),
                    divisor_le(l->L[3], b->tile_caps[2])};
        for (int d = 1; d < 4; ++d) l->tile[d] = t[d];
        l->morton = 0;   // y-fastest, as default_tiling
        const int rc = upload_desc(l);
        if (rc) return rc;
    }
    return batch_upload(b);
}

template <class T>
void launch_sweeps_on(const Launch& g, const qcb_batch* b, int n_hits, int n_sweeps, uint32_t sweep_idx,
                      const uint32_t* base) {
    for (int s = 0; s < n_sweeps; ++s)
        for (int k = 0; k < 8; ++k)
            launch_hb_half<T>(g, b->order[k] >> 1, b->order[k] & 1, n_hits, sweep_idx + (uint32_t)s, kHbPhilox, nullptr,
                              0, base, b->hb_grid);
}

int enqueue_batch_sweeps(qcb_batch* b, int n_hits, int n_sweeps, uint32_t sweep_idx, const uint32_t* base) {
    qcb_ctx* c = b->ctx;
    if (b->n_groups <= 1) {
        Launch g = batch_launch(b, kFamHalf);
        if (b->prec == QCB_F64) launch_sweeps_on<double>(g, b, n_hits, n_sweeps, sweep_idx, base);
        else launch_sweeps_on<float>(g, b, n_hits, n_sweeps, sweep_idx, base);
        QCB_HIP(c, hipGetLastError());
        return QCB_OK;
    }
    // fork the groups from the context stream, join them back into it
    QCB_HIP(c, hipEventRecord(b->grp_event[0], c->stream));
    for (int g = 1; g < b->n_groups; ++g) QCB_HIP(c, hipStreamWaitEvent(b->grp_stream[g], b->grp_event[0], 0));
    for (int g = 0; g < b->n_groups; ++g) {
        Launch gl;
        gl.lats = b->grp_desc[g];
        gl.blk = b->grp_blk[g];
        gl.n_lat = (int)b->grp_idx[g].size();
        gl.nblocks = b->grp_nb[g];
        gl.stream = b->grp_stream[g];
        if (b->prec == QCB_F64) launch_sweeps_on<double>(gl, b, n_hits, n_sweeps, sweep_idx, base);
        else launch_sweeps_on<float>(gl, b, n_hits, n_sweeps, sweep_idx, base);
    }
    QCB_HIP(c, hipGetLastError());
    for (int g = 1; g < b->n_groups; ++g) {
        QCB_HIP(c, hipEventRecord(b->grp_event[g], b->grp_stream[g]));
        QCB_HIP(c, hipStreamWaitEvent(c->stream, b->grp_event[g], 0));
    }
    return QCB_OK;
}

// Brings every lattice's descriptor to the context's attempt cap.
int batch_sync_cap(qcb_batch* b) {
    if (b->lats[0]->kp_cap == b->ctx->kp_cap) return QCB_OK;
    for (qcb_lattice* l : b->lats) {
        l->kp_cap = b->ctx->kp_cap;
        int rc = upload_desc(l);
        if (rc) return rc;
    }
    return batch_upload(b);
}

int zero_counters(qcb_batch* b) {
    QCB_HIP(b->ctx, hipMemsetAsync(b->d_counter, 0, (size_t)b->n * sizeof(unsigned long long), b->ctx->stream));
    return QCB_OK;
}

// Every address a flow launch sequence bakes into its kernel arguments.
std::vector<const void*> flow_key(const qcb_batch* b) {
    std::vector<const void*> k = {b->d_meas, b->d_partial, b->d_desc};
    for (const uint32_t* p : b->d_blk) k.push_back(p);
    for (const qcb_lattice* l : b->lats) {
        k.push_back(l->U);
        for (void* f : l->F) k.push_back(f);
        k.push_back(l->d_desc);
        k.push_back(l->d_blk[kFamLink]);
    }
    return k;
}

// Replays fg if it was captured for these arguments and addresses, else
// captures `enqueue` on the context stream into it and launches it.
template <class F>
int run_flow_graph(qcb_batch* b, qcb_batch::FlowGraph& fg, double eps, int n_steps, F enqueue) {
    qcb_ctx* c = b->ctx;
    std::vector<const void*> key = flow_key(b);
    if (!(fg.exec && fg.eps == eps && fg.n_steps == n_steps && fg.key == key)) {
        if (fg.exec) QCB_HIP(c, hipGraphExecDestroy(fg.exec));
        if (fg.graph) QCB_HIP(c, hipGraphDestroy(fg.graph));
        fg = qcb_batch::FlowGraph{};
        QCB_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue();
        const hipError_t e = hipStreamEndCapture(c->stream, &fg.graph);
        if (rc) return rc;
        if (e != hipSuccess) return qcb_set_error(c, QCB_EDEVICE, "flow graph capture failed: %s", hipGetErrorString(e));
        QCB_HIP(c, hipGraphInstantiate(&fg.exec, fg.graph, nullptr, nullptr, 0));
        fg.eps = eps;
        fg.n_steps = n_steps;
        fg.key = key;
    }
    QCB_HIP(c, hipGraphLaunch(fg.exec, c->stream));
    return QCB_OK;
}

}  // namespace

extern "C" int qcb_batch_create(qcb_ctx* ctx, int n, const int* dims, int prec, const double* betas,
                                const uint64_t* seeds, qcb_batch** out) {
    if (!ctx || n < 1 || !dims || !betas || !seeds || !out) return QCB_EINVAL;
    *out = nullptr;
    for (int i = 0; i < n; ++i)
        if (!(betas[i] > 0)) return qcb_set_error(ctx, QCB_EINVAL, "lattice %d: need beta > 0", i);
    qcb_batch* b = new qcb_batch();
    b->ctx = ctx;
    b->prec = prec;
    b->n = n;
    b->lats.assign(n, nullptr);
    int rc;
    if (hipMalloc((void**)&b->d_counter, (size_t)n * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(b->d_counter, 0, (size_t)n * sizeof(unsigned long long)) != hipSuccess) {
        batch_free(b);
        delete b;
        return qcb_set_error(ctx, QCB_ENOMEM, "cannot allocate the batch's counters");
    }
    for (int i = 0; i < n; ++i) {
        rc = qcb_lattice_create(ctx, dims + 4 * i, prec, &b->lats[i]);
        if (rc) {
            batch_free(b);
            delete b;
            return rc;
        }
        b->lats[i]->beta = betas[i];
        b->lats[i]->seed = seeds[i];
        b->lats[i]->kp_exh = b->d_counter + i;
        if ((rc = upload_desc(b->lats[i]))) {
            batch_free(b);
            delete b;
            return rc;
        }
    }
    QCB_HIP(ctx, hipMalloc(&b->d_desc, (size_t)n * sizeof(LatDev<double>)));
    for (int f = 0; f < kNumFam; ++f) {
        std::vector<uint32_t> pre(n + 1, 0);
        for (int i = 0; i < n; ++i) pre[i + 1] = pre[i] + fam_blocks(*b->lats[i], f);
        b->nb[f] = pre[n];
        QCB_HIP(ctx, hipMalloc((void**)&b->d_blk[f], pre.size() * sizeof(uint32_t)));
        QCB_HIP(ctx, hipMemcpy(b->d_blk[f], pre.data(), pre.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    QCB_HIP(ctx, hipMalloc((void**)&b->d_partial, 2 * sizeof(double) * b->nb[kFamClover]));
    QCB_HIP(ctx, hipMalloc((void**)&b->d_sweep, sizeof(uint32_t)));
    QCB_HIP(ctx, hipMemset(b->d_sweep, 0, sizeof(uint32_t)));
    rc = batch_upload(b);
    if (rc) {
        batch_free(b);
        delete b;
        return rc;
    }
    ctx->n_live++;
    *out = b;
    return QCB_OK;
}

extern "C" int qcb_batch_destroy(qcb_batch* b) {
    if (!b) return QCB_EINVAL;
    for (qcb_lattice* l : b->lats)
        if (l && l->n_users > 0)
            return qcb_set_error(b->ctx, QCB_ESTATE, "event-chain handles still use a lattice of this batch");
    (void)hipStreamSynchronize(b->ctx->stream);
    b->ctx->n_live--;
    batch_free(b);
    delete b;
    return QCB_OK;
}

extern "C" int qcb_batch_lattice(qcb_batch* b, int i, qcb_lattice** out) {
    if (!b || !out || i < 0 || i >= b->n) return QCB_EINVAL;
    *out = b->lats[i];
    return QCB_OK;
}

extern "C" int qcb_batch_hot(qcb_batch* b) {
    if (!b) return QCB_EINVAL;
    Launch g = batch_launch(b, kFamLink);
    if (b->prec == QCB_F64) launch_hot<double>(g); else launch_hot<float>(g);
    QCB_HIP(b->ctx, hipGetLastError());
    return QCB_OK;
}

extern "C" int qcb_batch_hb_sweeps(qcb_batch* b, int n_hits, int n_sweeps, uint32_t sweep_idx) {
    if (!b || n_hits < 1 || n_hits > 4095 || n_sweeps < 0) return QCB_EINVAL;
    int rc;
    if ((rc = batch_sync_cap(b)) || (rc = zero_counters(b))) return rc;
    if ((rc = enqueue_batch_sweeps(b, n_hits, n_sweeps, sweep_idx, nullptr))) return rc;
    return check_kp_counters(b->ctx, b->ctx->stream, b->d_counter, b->n, b->ctx->kp_cap);
}

extern "C" int qcb_batch_reunitarize(qcb_batch* b) {
    if (!b) return QCB_EINVAL;
    Launch g = batch_launch(b, kFamLink);
    if (b->prec == QCB_F64) launch_reunit<double>(g); else launch_reunit<float>(g);
    QCB_HIP(b->ctx, hipGetLastError());
    return QCB_OK;
}

extern "C" int qcb_batch_plaquette(qcb_batch* b, double* out) {
    if (!b || !out) return QCB_EINVAL;
    int rc = ensure_meas(b, (size_t)b->n);
    if (rc) return rc;
    Launch g = batch_launch(b, kFamSite);
    if (b->prec == QCB_F64) {
        launch_plaquette<double>(g, b->d_partial);
        launch_finalize_plaq<double>(g, b->d_partial, b->d_meas, 1);
    } else {
        launch_plaquette<float>(g, b->d_partial);
        launch_finalize_plaq<float>(g, b->d_partial, b->d_meas, 1);
    }
    QCB_HIP(b->ctx, hipGetLastError());
    QCB_HIP(b->ctx, hipMemcpyAsync(out, b->d_meas, b->n * sizeof(double), hipMemcpyDeviceToHost, b->ctx->stream));
    QCB_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
    return QCB_OK;
}

// Flow every lattice of the batch from its links: measure (Q, E/V) at
// t = 0, eps, ..., n_steps*eps (measure-then-step, plus the final state).
// tQE[(i*(n_steps+1) + s)*3 + {0,1,2}] = (t, Q, E_mean) of lattice i.  Every
// call counts its flow steps from zero, so calls with the same eps and step
// count issue the same launches: they replay one captured graph.
extern "C" int qcb_batch_flow(qcb_batch* b, double eps, int n_steps, double* tQE) {
    if (!b || n_steps < 0 || !tQE) return QCB_EINVAL;
    qcb_ctx* c = b->ctx;
    const int nm = n_steps + 1;
    int rc = ensure_meas(b, (size_t)b->n * nm * 2);
    if (rc) return rc;
    for (qcb_lattice* l : b->lats) {
        if ((rc = ensure_flow_buffers(l))) return rc;
        l->flow_start = 0;
        l->flow_active = true;
    }
    if ((rc = batch_upload(b))) return rc;
    const bool f64 = b->prec == QCB_F64;
    auto enqueue = [&]() -> int {
        Launch gl = batch_launch(b, kFamFlow), gs = batch_launch(b, kFamClover);
        double* pq = b->d_partial;
        double* pe = b->d_partial + b->nb[kFamClover];
        // copy links into the flow fields of every lattice: one launch per lattice
        for (qcb_lattice* l : b->lats) {
            Launch one = lat_launch(l, kFamLink);
            if (f64) launch_copy_field<double>(one, (const double*)l->U, (double*)l->F[0], kLinkOrder, kFlowOrder);
            else launch_copy_field<float>(one, (const float*)l->U, (float*)l->F[0], kLinkOrder, kFlowOrder);
        }
        for (int s = 0; s < nm; ++s) {
            if (f64) {
                launch_clover<double>(gs, 1, (uint32_t)s, pq, pe);
                launch_finalize_qe<double>(gs, pq, pe, b->d_meas + 2 * s, 2 * nm);
            } else {
                launch_clover<float>(gs, 1, (uint32_t)s, pq, pe);
                launch_finalize_qe<float>(gs, pq, pe, b->d_meas + 2 * s, 2 * nm);
            }
            if (s == n_steps) break;
            for (int st = 1; st <= 3; ++st) {
                if (f64) launch_flow_stage<double>(gl, st, (uint32_t)s, eps);
                else launch_flow_stage<float>(gl, st, (uint32_t)s, eps);
            }
        }
        QCB_HIP(c, hipGetLastError());
        return QCB_OK;
    };
    if ((rc = run_flow_graph(b, b->flow_graph[0], eps, n_steps, enqueue))) return rc;
    std::vector<double> h((size_t)b->n * nm * 2);
    QCB_HIP(c, hipMemcpyAsync(h.data(), b->d_meas, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    QCB_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<double> tv(nm);
    double t = 0.0;   // flow time accumulated as the reference does
    for (int s = 0; s < nm; ++s, t += eps) tv[s] = t;
    for (int i = 0; i < b->n; ++i) {
        b->lats[i]->gstep = (uint32_t)n_steps;
        if ((rc = upload_desc(b->lats[i]))) return rc;   // single-lattice calls see the flow state
        for (int s = 0; s < nm; ++s) {
            tQE[((size_t)i * nm + s) * 3] = tv[s];
            tQE[((size_t)i * nm + s) * 3 + 1] = h[(size_t)i * nm * 2 + 2 * s];
            tQE[((size_t)i * nm + s) * 3 + 2] = h[(size_t)i * nm * 2 + 2 * s + 1];
        }
    }
    return QCB_OK;
}

// qcb_flow_steps for every lattice of the batch: each lattice's links are
// copied into its flow fields, n_steps RK3 steps run with one launch per stage
// for all lattices, and the flowed fields are copied back into the links; no
// measurement.  The flowed field stays the lattice's flow state, as after
// qcb_flow_steps.  Replayed from a captured graph like qcb_batch_flow.
extern "C" int qcb_batch_flow_steps(qcb_batch* b, double eps, int n_steps) {
    if (!b || n_steps < 0) return QCB_EINVAL;
    qcb_ctx* c = b->ctx;
    int rc;
    for (qcb_lattice* l : b->lats) {
        if ((rc = ensure_flow_buffers(l))) return rc;
        l->flow_start = 0;
        l->flow_active = true;
    }
    if ((rc = batch_upload(b))) return rc;
    const bool f64 = b->prec == QCB_F64;
    auto enqueue = [&]() -> int {
        for (qcb_lattice* l : b->lats) {
            Launch one = lat_launch(l, kFamLink);
            if (f64) launch_copy_field<double>(one, (const double*)l->U, (double*)l->F[0], kLinkOrder, kFlowOrder);
            else launch_copy_field<float>(one, (const float*)l->U, (float*)l->F[0], kLinkOrder, kFlowOrder);
        }
        Launch gl = batch_launch(b, kFamFlow);
        for (int s = 0; s < n_steps; ++s)
            for (int st = 1; st <= 3; ++st) {
                if (f64) launch_flow_stage<double>(gl, st, (uint32_t)s, eps);
                else launch_flow_stage<float>(gl, st, (uint32_t)s, eps);
            }
        for (qcb_lattice* l : b->lats) {
            Launch one = lat_launch(l, kFamLink);
            const void* cur = l->F[(uint32_t)n_steps & 1u];
            if (f64) launch_copy_field<double>(one, (const double*)cur, (double*)l->U, kFlowOrder, kLinkOrder);
            else launch_copy_field<float>(one, (const float*)cur, (float*)l->U, kFlowOrder, kLinkOrder);
        }
        QCB_HIP(c, hipGetLastError());
        return QCB_OK;
    };
    if ((rc = run_flow_graph(b, b->flow_graph[1], eps, n_steps, enqueue))) return rc;
    for (qcb_lattice* l : b->lats) {
        l->gstep = (uint32_t)n_steps;
        if ((rc = upload_desc(l))) return rc;   // single-lattice calls see the flow state
    }
    return QCB_OK;
}

extern "C" int qcb_batch_set_schedule(qcb_batch* b, const int order[8], uint32_t hb_grid) {
    if (!b || !order || hb_grid % 8) return QCB_EINVAL;
    int seen = 0;
    for (int k = 0; k < 8; ++k) {
        if (order[k] < 0 || order[k] > 7 || (seen >> order[k]) & 1) return QCB_EINVAL;
        seen |= 1 << order[k];
    }
    std::memcpy(b->order, order, sizeof b->order);
    b->hb_grid = hb_grid;
    return QCB_OK;
}

extern "C" int qcb_batch_get_schedule(qcb_batch* b, int order[8], uint32_t* hb_grid) {
    if (!b || !order || !hb_grid) return QCB_EINVAL;
    std::memcpy(order, b->order, sizeof b->order);
    *hb_grid = b->hb_grid;
    return QCB_OK;
}

extern "C" int qcb_batch_set_layout(qcb_batch* b, int n_streams, const int tile_caps[3]) {
    if (!b || !tile_caps || n_streams < 1 || n_streams > 8) return QCB_EINVAL;
    for (int d = 0; d < 3; ++d)
        if (tile_caps[d] < 1) return qcb_set_error(b->ctx, QCB_EINVAL, "tile caps must be >= 1");
    std::memcpy(b->tile_caps, tile_caps, sizeof b->tile_caps);
    int rc = apply_tiles(b);
    if (rc) return rc;
    return n_streams == b->n_groups ? QCB_OK : set_groups(b, n_streams);
}

extern "C" int qcb_batch_get_layout(qcb_batch* b, int* n_streams, int tile_caps[3]) {
    if (!b || !n_streams || !tile_caps) return QCB_EINVAL;
    *n_streams = b->n_groups;
    std::memcpy(tile_caps, b->tile_caps, sizeof b->tile_caps);
    return QCB_OK;
}

// Times candidate schedules (half-sweep order x persistent grid x stream
// groups, then work-order tile caps) on a scratch copy of every lattice's
// links and keeps the fastest.  The links themselves are not modified.
// report (optional) receives one line per candidate.
static int autotune_on_scratch(qcb_batch* b, int n_hits, int reps, char* report, size_t report_len);

extern "C" int qcb_batch_autotune(qcb_batch* b, int n_hits, int reps, char* report, size_t report_len) {
    if (!b || n_hits < 1 || reps < 1) return QCB_EINVAL;
    int rc;
    // tune on scratch copies: swap U with F[0] after copying, and swap back on
    // every return path so the lattices keep their own links
    for (qcb_lattice* l : b->lats)
        if ((rc = ensure_flow_buffers(l))) return rc;
    for (qcb_lattice* l : b->lats) {
        Launch one = lat_launch(l, kFamLink);
        if (b->prec == QCB_F64) launch_copy_field<double>(one, (const double*)l->U, (double*)l->F[0]);
        else launch_copy_field<float>(one, (const float*)l->U, (float*)l->F[0]);
        std::swap(l->U, l->F[0]);
        l->flow_active = false;
    }
    rc = hipGetLastError() == hipSuccess ? autotune_on_scratch(b, n_hits, reps, report, report_len)
                                         : qcb_set_error(b->ctx, QCB_EDEVICE, "autotune: copying the links failed");
    for (qcb_lattice* l : b->lats) std::swap(l->U, l->F[0]);
    const int rc2 = batch_upload(b);
    return rc ? rc : rc2;
}

// The timing part of qcb_batch_autotune, run while every lattice's U is a
// scratch copy of its links.
static int autotune_on_scratch(qcb_batch* b, int n_hits, int reps, char* report, size_t report_len) {
    qcb_ctx* c = b->ctx;
    int rc;
    if ((rc = batch_upload(b))) return rc;
    int dev = 0, ncu = 0;
    QCB_HIP(c, hipGetDevice(&dev));
    QCB_HIP(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int orders[2][8] = {{0, 1, 2, 3, 4, 5, 6, 7}, {0, 2, 4, 6, 1, 3, 5, 7}};
    std::vector<uint32_t> grids = {0};
    for (int per : {2, 4, 8}) {
        uint32_t g = (uint32_t)(ncu * per) & ~7u;
        if (g && g < b->nb[kFamHalf]) grids.push_back(g);
    }
    std::vector<int> groups = {1};
    for (int n : {2, 4})
        if (n <= b->n) groups.push_back(n);
    const int tiles[5][3] = {{8, 4, 4}, {4, 4, 4}, {4, 2, 2}, {8, 8, 8}, {2, 2, 2}};
    std::string rep;
    auto time_one = [&](float* ms) -> int {
        int r = enqueue_batch_sweeps(b, n_hits, 1, 0x7fff0000u, nullptr);   // warm-up
        if (r) return r;
        QCB_HIP(c, hipEventRecord(c->ev0, c->stream));
        if ((r = enqueue_batch_sweeps(b, n_hits, reps, 0x7fff0001u, nullptr))) return r;
        QCB_HIP(c, hipEventRecord(c->ev1, c->stream));
        QCB_HIP(c, hipEventSynchronize(c->ev1));
        QCB_HIP(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
        *ms /= reps;
        char line[192];
        std::snprintf(line, sizeof line, "order=%s grid=%u streams=%d tiles=%dx%dx%d ms_per_sweep=%.4f\n",
                      b->order[1] == 1 ? "mu-major" : "parity-major", b->hb_grid, b->n_groups, b->tile_caps[0],
                      b->tile_caps[1], b->tile_caps[2], *ms);
        rep += line;
        return QCB_OK;
    };
    // 1: half-sweep order x persistent grid x stream groups at the current tiles
    float best = 1e30f;
    int best_o = 0, best_n = b->n_groups;
    uint32_t best_g = 0;
    for (int n : groups) {
        if ((rc = set_groups(b, n))) return rc;
        for (int o = 0; o < 2; ++o)
            for (uint32_t g : grids) {
                std::memcpy(b->order, orders[o], sizeof b->order);
                b->hb_grid = g;
                float ms = 0.f;
                if ((rc = time_one(&ms))) return rc;
                if (ms < best) {
                    best = ms;
                    best_o = o;
                    best_g = g;
                    best_n = n;
                }
            }
    }
    std::memcpy(b->order, orders[best_o], sizeof b->order);
    b->hb_grid = best_g;
    if ((rc = set_groups(b, best_n))) return rc;
    // 2: work-order tiles with that schedule
    int best_t[3];
    std::memcpy(best_t, b->tile_caps, sizeof best_t);
    for (const auto& t : tiles) {
        if (std::equal(t, t + 3, best_t)) continue;   // timed in step 1
        std::memcpy(b->tile_caps, t, sizeof b->tile_caps);
        if ((rc = apply_tiles(b))) return rc;
        float ms = 0.f;
        if ((rc = time_one(&ms))) return rc;
        if (ms < best) {
            best = ms;
            std::memcpy(best_t, t, sizeof best_t);
        }
    }
    std::memcpy(b->tile_caps, best_t, sizeof b->tile_caps);
    if ((rc = apply_tiles(b))) return rc;
    if (report && report_len) {
        std::snprintf(report, report_len, "%s", rep.c_str());
    }
    return QCB_OK;
}

// Captures sweeps_per_graph sweeps (8 half-sweep launches each, with the
// Philox sweep index read from a device counter) plus the counter update into
// a hipGraph; qcb_batch_graph_launch replays it.  Sweep indices continue from
// first_sweep across replays.
extern "C" int qcb_batch_graph_capture(qcb_batch* b, int n_hits, int sweeps_per_graph, uint32_t first_sweep) {
    if (!b || n_hits < 1 || sweeps_per_graph < 1) return QCB_EINVAL;
    qcb_ctx* c = b->ctx;
    if (b->graph_exec) {
        QCB_HIP(c, hipGraphExecDestroy(b->graph_exec));
        b->graph_exec = nullptr;
    }
    if (b->graph) {
        QCB_HIP(c, hipGraphDestroy(b->graph));
        b->graph = nullptr;
    }
    int rc0 = batch_sync_cap(b);
    if (rc0) return rc0;
    QCB_HIP(c, hipMemcpyAsync(b->d_sweep, &first_sweep, sizeof first_sweep, hipMemcpyHostToDevice, c->stream));
    QCB_HIP(c, hipStreamSynchronize(c->stream));
    QCB_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_batch_sweeps(b, n_hits, sweeps_per_graph, 0, b->d_sweep);
    launch_bump(b->d_sweep, (uint32_t)sweeps_per_graph, c->stream);
    hipError_t e = hipStreamEndCapture(c->stream, &b->graph);
    if (rc) return rc;
    if (e != hipSuccess) return qcb_set_error(c, QCB_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
    QCB_HIP(c, hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    b->graph_sweeps = sweeps_per_graph;
    return QCB_OK;
}

extern "C" int qcb_batch_graph_launch(qcb_batch* b, int n_replays) {
    if (!b || n_replays < 0) return QCB_EINVAL;
    if (!b->graph_exec) return qcb_set_error(b->ctx, QCB_ESTATE, "no captured graph");
    int rc = zero_counters(b);
    if (rc) return rc;
    for (int r = 0; r < n_replays; ++r) QCB_HIP(b->ctx, hipGraphLaunch(b->graph_exec, b->ctx->stream));
    return check_kp_counters(b->ctx, b->ctx->stream, b->d_counter, b->n, b->lats[0]->kp_cap);
}
