// qcb_api.cpp -- C-ABI over the gfx950 kernels: contexts, single lattices,
// device buffers and error reporting.  See include/qcb.h for the contract.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/qcb.h"
#include "qcb_internal.hpp"
#include "qcb_host.hpp"

using namespace qcb;

namespace {
// A stream-ordered device temporary, released on every return path: the free
// is queued behind whatever work was enqueued on the stream before it.
struct DevTmp {
    qcb_ctx* c;
    void* p = nullptr;
    explicit DevTmp(qcb_ctx* ctx) : c(ctx) {}
    DevTmp(const DevTmp&) = delete;
    DevTmp& operator=(const DevTmp&) = delete;
    ~DevTmp() {
        if (p) (void)hipFreeAsync(p, c->stream);
    }
    template <class T>
    int alloc(size_t bytes, T** out) {
        hipError_t e = hipMallocAsync(&p, bytes, c->stream);
        if (e != hipSuccess) {
            p = nullptr;
            return qcb_set_error(c, QCB_ENOMEM, "cannot allocate a %zu-byte temporary: %s", bytes, hipGetErrorString(e));
        }
        *out = static_cast<T*>(p);
        return QCB_OK;
    }
};
}  // namespace

// Static device scratch of the entry points that allocate nothing: 1 MiB, the
// (t, Q, E) triples of up to kScratchDoubles / 3 flow measurements.  Each
// device holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns on that device (ctx->scratch).
constexpr size_t kScratchDoubles = (size_t)1 << 17;
static __device__ __align__(256) double scratch[kScratchDoubles];

// ------------------------------------------------------------ context -----

int qcb_set_error(qcb_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

extern "C" int qcb_ctx_create(int device, qcb_ctx** out) {
    if (!out) return QCB_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return QCB_EDEVICE;
    if (device < 0 || device >= n) return QCB_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return QCB_EDEVICE;
    qcb_ctx* c = new qcb_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipGetSymbolAddress((void**)&c->scratch, HIP_SYMBOL(scratch)) != hipSuccess || !c->scratch) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        if (c->ev0) (void)hipEventDestroy(c->ev0);
        if (c->ev1) (void)hipEventDestroy(c->ev1);
        delete c;
        return QCB_EDEVICE;
    }
    *out = c;
    return QCB_OK;
}

extern "C" int qcb_ctx_destroy(qcb_ctx* ctx) {
    if (!ctx) return QCB_EINVAL;
    if (ctx->n_live > 0)
        return qcb_set_error(ctx, QCB_ESTATE, "%d lattice/batch handles still alive on this context", ctx->n_live);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return QCB_OK;
}

thread_local std::string g_free_error;   // failures of context-free calls

int qcb_set_free_error(int code, const std::string& msg) {
    g_free_error = msg;
    return code;
}

extern "C" const char* qcb_last_error(qcb_ctx* ctx) { return ctx ? ctx->err.c_str() : g_free_error.c_str(); }

extern "C" int qcb_ctx_synchronize(qcb_ctx* ctx) {
    if (!ctx) return QCB_EINVAL;
    QCB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QCB_OK;
}

extern "C" int qcb_ctx_set_kp_budget(qcb_ctx* ctx, uint32_t max_attempts) {
    if (!ctx) return QCB_EINVAL;
    if (max_attempts < 1 || max_attempts > (1u << 20))
        return qcb_set_error(ctx, QCB_EINVAL, "Kennedy-Pendleton attempt cap must be in 1..2^20");
    ctx->kp_cap = max_attempts;
    return QCB_OK;
}

extern "C" int qcb_timing_begin(qcb_ctx* ctx) {
    if (!ctx) return QCB_EINVAL;
    QCB_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    return QCB_OK;
}

extern "C" int qcb_timing_end(qcb_ctx* ctx, double* ms) {
    if (!ctx || !ms) return QCB_EINVAL;
    QCB_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    QCB_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float f = 0.f;
    QCB_HIP(ctx, hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
    *ms = f;
    return QCB_OK;
}

// ------------------------------------------------------------ lattice -----

extern "C" int qcb_lattice_create(qcb_ctx* ctx, const int dims[4], int prec, qcb_lattice** out) {
    if (!ctx || !dims || !out) return QCB_EINVAL;
    *out = nullptr;
    if (prec != QCB_F64 && prec != QCB_F32) return qcb_set_error(ctx, QCB_EINVAL, "prec must be QCB_F64 or QCB_F32");
    for (int d = 0; d < 4; ++d)
        if (dims[d] < 4 || dims[d] % 2)
            return qcb_set_error(ctx, QCB_EINVAL, "extent %d of direction %d must be even and >= 4", dims[d], d);
    uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2] * dims[3];
    if (V * 4 >= (1ull << 31)) return qcb_set_error(ctx, QCB_EINVAL, "lattice too large (V=%llu)", (unsigned long long)V);
    QCB_HIP(ctx, hipSetDevice(ctx->device));
    qcb_lattice* l = new qcb_lattice();
    l->ctx = ctx;
    l->prec = prec;
    for (int d = 0; d < 4; ++d) l->L[d] = dims[d];
    l->V = (uint32_t)V;
    l->Vh = (uint32_t)(V / 2);
    l->S = (l->Vh + 63u) & ~63u;
    l->esize = prec == QCB_F64 ? 8 : 4;
    default_tiling(dims, l->tile, &l->morton);
    int rc = lattice_alloc(l);
    if (rc != QCB_OK) {
        lattice_free(l);
        delete l;
        return rc;
    }
    rc = qcb_cold(l);
    if (rc != QCB_OK) {
        lattice_free(l);
        delete l;
        return rc;
    }
    ctx->n_live++;
    *out = l;
    return QCB_OK;
}

extern "C" int qcb_lattice_destroy(qcb_lattice* lat) {
    if (!lat) return QCB_EINVAL;
    if (lat->n_users > 0)
        return qcb_set_error(lat->ctx, QCB_ESTATE, "%d event-chain handles still use this lattice", lat->n_users);
    (void)hipSetDevice(lat->ctx->device);
    (void)hipStreamSynchronize(lat->ctx->stream);
    lat->ctx->n_live--;
    lattice_free(lat);
    delete lat;
    return QCB_OK;
}

extern "C" int qcb_lattice_set_tiling(qcb_lattice* lat, int ty, int tz, int tt, int morton) {
    if (!lat) return QCB_EINVAL;
    const int t[4] = {0, ty, tz, tt};
    for (int d = 1; d < 4; ++d) {
        if (t[d] < 1 || lat->L[d] % t[d])
            return qcb_set_error(lat->ctx, QCB_EINVAL, "tile extent %d does not divide %d", t[d], lat->L[d]);
        const int n = lat->L[d] / t[d];
        if (morton && (n & (n - 1)))
            return qcb_set_error(lat->ctx, QCB_EINVAL, "Z-order needs power-of-two tile counts (%d)", n);
    }
    for (int d = 1; d < 4; ++d) lat->tile[d] = t[d];
    lat->morton = morton ? 1 : 0;
    return upload_desc(lat);
}

extern "C" int qcb_lattice_bytes(qcb_lattice* lat, uint64_t* bytes) {
    if (!lat || !bytes) return QCB_EINVAL;
    *bytes = field_bytes(lat) + (lat->F[0] ? flow_bytes(lat) : 0);
    return QCB_OK;
}

extern "C" int qcb_lattice_set_links(qcb_lattice* lat, const double* aos) {
    if (!lat || !aos) return QCB_EINVAL;
    qcb_ctx* ctx = lat->ctx;
    size_t n = (size_t)lat->V * 72;
    double* d = nullptr;
    DevTmp tmp(ctx);
    if (int rc = tmp.alloc(n * sizeof(double), &d)) return rc;
    QCB_HIP(ctx, hipMemcpyAsync(d, aos, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    Launch g = lat_launch(lat, kFamLink);
    if (lat->prec == QCB_F64) launch_aos_to_soa<double>(g, d, (double*)lat->U);
    else launch_aos_to_soa<float>(g, d, (float*)lat->U);
    QCB_HIP(ctx, hipGetLastError());
    QCB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QCB_OK;
}

static int download_field(qcb_lattice* lat, const void* field, double* aos, int order) {
    qcb_ctx* ctx = lat->ctx;
    size_t n = (size_t)lat->V * 72;
    double* d = nullptr;
    DevTmp tmp(ctx);
    if (int rc = tmp.alloc(n * sizeof(double), &d)) return rc;
    Launch g = lat_launch(lat, kFamLink);
    if (lat->prec == QCB_F64) launch_soa_to_aos<double>(g, (const double*)field, d, order);
    else launch_soa_to_aos<float>(g, (const float*)field, d, order);
    QCB_HIP(ctx, hipGetLastError());
    QCB_HIP(ctx, hipMemcpyAsync(aos, d, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    QCB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QCB_OK;
}

extern "C" int qcb_lattice_get_links(qcb_lattice* lat, double* aos) {
    if (!lat || !aos) return QCB_EINVAL;
    return download_field(lat, lat->U, aos, kLinkOrder);
}

extern "C" int qcb_cold(qcb_lattice* lat) {
    if (!lat) return QCB_EINVAL;
    Launch g = lat_launch(lat, kFamLink);
    if (lat->prec == QCB_F64) launch_cold<double>(g); else launch_cold<float>(g);
    QCB_HIP(lat->ctx, hipGetLastError());
    return QCB_OK;
}

extern "C" int qcb_hot(qcb_lattice* lat, uint64_t seed) {
    if (!lat) return QCB_EINVAL;
    lat->seed = seed;
    int rc = upload_desc(lat);
    if (rc) return rc;
    Launch g = lat_launch(lat, kFamLink);
    if (lat->prec == QCB_F64) launch_hot<double>(g); else launch_hot<float>(g);
    QCB_HIP(lat->ctx, hipGetLastError());
    return QCB_OK;
}

// ---------------------------------------------------------- observables ---

int enqueue_plaquette(qcb_lattice* lat, int which, double* d_out) {
    Launch g = lat_launch(lat, kFamSite);
    if (lat->prec == QCB_F64) {
        launch_plaquette<double>(g, lat->d_partial, which, lat->gstep);
        launch_finalize_plaq<double>(g, lat->d_partial, d_out, 1);
    } else {
        launch_plaquette<float>(g, lat->d_partial, which, lat->gstep);
        launch_finalize_plaq<float>(g, lat->d_partial, d_out, 1);
    }
    QCB_HIP(lat->ctx, hipGetLastError());
    return QCB_OK;
}

int enqueue_clover(qcb_lattice* lat, int which, double* d_out) {
    Launch g = lat_launch(lat, kFamClover);
    double* pq = lat->d_partial;
    double* pe = lat->d_partial + lat->nb[kFamClover];
    if (lat->prec == QCB_F64) {
        launch_clover<double>(g, which, lat->gstep, pq, pe);
        launch_finalize_qe<double>(g, pq, pe, d_out, 2);
    } else {
        launch_clover<float>(g, which, lat->gstep, pq, pe);
        launch_finalize_qe<float>(g, pq, pe, d_out, 2);
    }
    QCB_HIP(lat->ctx, hipGetLastError());
    return QCB_OK;
}

extern "C" int qcb_plaquette(qcb_lattice* lat, double* p) {
    if (!lat || !p) return QCB_EINVAL;
    int rc = enqueue_plaquette(lat, 0, lat->d_out);
    if (rc) return rc;
    QCB_HIP(lat->ctx, hipMemcpyAsync(p, lat->d_out, sizeof(double), hipMemcpyDeviceToHost, lat->ctx->stream));
    QCB_HIP(lat->ctx, hipStreamSynchronize(lat->ctx->stream));
    return QCB_OK;
}

extern "C" int qcb_clover_qe(qcb_lattice* lat, double* Q, double* E) {
    if (!lat || !Q || !E) return QCB_EINVAL;
    int rc = enqueue_clover(lat, 0, lat->d_out);
    if (rc) return rc;
    double h[2];
    QCB_HIP(lat->ctx, hipMemcpyAsync(h, lat->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost, lat->ctx->stream));
    QCB_HIP(lat->ctx, hipStreamSynchronize(lat->ctx->stream));
    *Q = h[0];
    *E = h[1];
    return QCB_OK;
}

extern "C" int qcb_staples(qcb_lattice* lat, double* out) {
    if (!lat || !out) return QCB_EINVAL;
    qcb_ctx* ctx = lat->ctx;
    size_t n = (size_t)lat->V * 72;
    double* d = nullptr;
    DevTmp tmp(ctx);
    if (int rc = tmp.alloc(n * sizeof(double), &d)) return rc;
    Launch g = lat_launch(lat, kFamLink);
    if (lat->prec == QCB_F64) launch_staples<double>(g, d);
    else launch_staples<float>(g, d);
    QCB_HIP(ctx, hipGetLastError());
    QCB_HIP(ctx, hipMemcpyAsync(out, d, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    QCB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QCB_OK;
}

// ------------------------------------------------------------- heatbath ---

int enqueue_hb_sweeps(qcb_lattice* lat, int n_hits, int n_sweeps, uint32_t sweep_idx) {
    Launch g = lat_launch(lat, kFamHalf);
    for (int s = 0; s < n_sweeps; ++s)
        for (int mu = 0; mu < 4; ++mu)
            for (int p = 0; p < 2; ++p) {
                if (lat->prec == QCB_F64)
                    launch_hb_half<double>(g, mu, p, n_hits, sweep_idx + (uint32_t)s, kHbPhilox, nullptr, 0);
                else
                    launch_hb_half<float>(g, mu, p, n_hits, sweep_idx + (uint32_t)s, kHbPhilox, nullptr, 0);
            }
    QCB_HIP(lat->ctx, hipGetLastError());
    return QCB_OK;
}

extern "C" int qcb_hb_sweeps(qcb_lattice* lat, double beta, int n_hits, int n_sweeps, uint64_t seed, uint32_t sweep_idx) {
    if (!lat) return QCB_EINVAL;
    if (n_hits < 1 || n_hits > 4095 || n_sweeps < 0 || !(beta > 0))
        return qcb_set_error(lat->ctx, QCB_EINVAL, "need beta > 0, 1 <= n_hits < 4096, n_sweeps >= 0");
    qcb_ctx* ctx = lat->ctx;
    if (lat->beta != beta || lat->seed != seed || lat->kp_cap != ctx->kp_cap) {
        lat->beta = beta;
        lat->seed = seed;
        lat->kp_cap = ctx->kp_cap;
        int rc = upload_desc(lat);
        if (rc) return rc;
    }
    QCB_HIP(ctx, hipMemsetAsync(lat->kp_exh, 0, sizeof(unsigned long long), ctx->stream));
    int rc = enqueue_hb_sweeps(lat, n_hits, n_sweeps, sweep_idx);
    if (rc) return rc;
    return check_kp_counters(ctx, ctx->stream, lat->kp_exh, 1, lat->kp_cap);
}

extern "C" int qcb_hb_sweep_injected(qcb_lattice* lat, double beta, int n_hits, const double* u, int K, int64_t* n_exh) {
    if (!lat || !u) return QCB_EINVAL;
    if (n_hits < 1 || n_hits > 4095 || K < 1 || K > 64 || !(beta > 0))
        return qcb_set_error(lat->ctx, QCB_EINVAL, "need beta > 0, 1 <= n_hits < 4096, 1 <= K <= 64");
    qcb_ctx* ctx = lat->ctx;
    lat->beta = beta;
    int rc = upload_desc(lat);
    if (rc) return rc;
    size_t n = (size_t)8 * n_hits * 3 * (4 * K + 2) * lat->Vh;
    double* d = nullptr;
    DevTmp tmp(ctx);
    if (int rc2 = tmp.alloc(n * sizeof(double), &d)) return rc2;
    QCB_HIP(ctx, hipMemcpyAsync(d, u, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    QCB_HIP(ctx, hipMemsetAsync(lat->kp_exh, 0, sizeof(unsigned long long), ctx->stream));
    Launch g = lat_launch(lat, kFamHalf);
    for (int mu = 0; mu < 4; ++mu)
        for (int p = 0; p < 2; ++p) {
            // the kernel indexes its half-sweep block with hs = 2*mu+p from the buffer start
            if (lat->prec == QCB_F64) launch_hb_half<double>(g, mu, p, n_hits, 0, kHbInjected, d, K);
            else launch_hb_half<float>(g, mu, p, n_hits, 0, kHbInjected, d, K);
        }
    QCB_HIP(ctx, hipGetLastError());
    unsigned long long h = 0;
    QCB_HIP(ctx, hipMemcpyAsync(&h, lat->kp_exh, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    QCB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (n_exh) *n_exh = (int64_t)h;
    return QCB_OK;
}

extern "C" int qcb_reunitarize(qcb_lattice* lat) {
    if (!lat) return QCB_EINVAL;
    Launch g = lat_launch(lat, kFamLink);
    if (lat->prec == QCB_F64) launch_reunit<double>(g); else launch_reunit<float>(g);
    QCB_HIP(lat->ctx, hipGetLastError());
    return QCB_OK;
}

extern "C" int qcb_translate(qcb_lattice* lat, const int s[4]) {
    if (!lat || !s) return QCB_EINVAL;
    int sh[4];
    for (int d = 0; d < 4; ++d) sh[d] = ((s[d] % lat->L[d]) + lat->L[d]) % lat->L[d];
    if (!sh[0] && !sh[1] && !sh[2] && !sh[3]) return QCB_OK;
    int rc = ensure_flow_buffers(lat);
    if (rc) return rc;
    Launch g = lat_launch(lat, kFamLink);
    if (lat->prec == QCB_F64) launch_translate<double>(g, sh, (double*)lat->F[0]);
    else launch_translate<float>(g, sh, (float*)lat->F[0]);
    QCB_HIP(lat->ctx, hipGetLastError());
    std::swap(lat->U, lat->F[0]);
    lat->flow_active = false;
    return upload_desc(lat);
}

// ----------------------------------------------------------------- flow ---

int enqueue_flow_step(qcb_lattice* lat, double eps) {
    Launch g = lat_launch(lat, kFamFlow);
    for (int st = 1; st <= 3; ++st) {
        if (lat->prec == QCB_F64) launch_flow_stage<double>(g, st, lat->gstep, eps);
        else launch_flow_stage<float>(g, st, lat->gstep, eps);
    }
    QCB_HIP(lat->ctx, hipGetLastError());
    lat->gstep++;
    return QCB_OK;
}

int flow_begin(qcb_lattice* lat) {
    int rc = ensure_flow_buffers(lat);
    if (rc) return rc;
    lat->flow_start = lat->gstep;
    lat->flow_active = true;
    rc = upload_desc(lat);
    if (rc) return rc;
    Launch g = lat_launch(lat, kFamLink);
    if (lat->prec == QCB_F64) launch_copy_field<double>(g, (const double*)lat->U, (double*)lat->F[0], kLinkOrder, kFlowOrder);
    else launch_copy_field<float>(g, (const float*)lat->U, (float*)lat->F[0], kLinkOrder, kFlowOrder);
    QCB_HIP(lat->ctx, hipGetLastError());
    return QCB_OK;
}

void* flow_current(qcb_lattice* lat) { return lat->F[(lat->gstep - lat->flow_start) & 1u]; }

extern "C" int qcb_flow_steps(qcb_lattice* lat, double eps, int n) {
    if (!lat || n < 0) return QCB_EINVAL;
    int rc = flow_begin(lat);
    if (rc) return rc;
    for (int s = 0; s < n; ++s)
        if ((rc = enqueue_flow_step(lat, eps))) return rc;
    Launch g = lat_launch(lat, kFamLink);
    if (lat->prec == QCB_F64) launch_copy_field<double>(g, (const double*)flow_current(lat), (double*)lat->U, kFlowOrder, kLinkOrder);
    else launch_copy_field<float>(g, (const float*)flow_current(lat), (float*)lat->U, kFlowOrder, kLinkOrder);
    QCB_HIP(lat->ctx, hipGetLastError());
    return QCB_OK;
}

extern "C" int qcb_flow_copy(qcb_lattice* lat) {
    if (!lat) return QCB_EINVAL;
    return flow_begin(lat);
}

extern "C" int qcb_flow_rk3(qcb_lattice* lat, double eps, int n) {
    if (!lat || n < 0) return QCB_EINVAL;
    if (!lat->flow_active) return qcb_set_error(lat->ctx, QCB_ESTATE, "qcb_flow_copy must come first");
    for (int s = 0; s < n; ++s) {
        int rc = enqueue_flow_step(lat, eps);
        if (rc) return rc;
    }
    return QCB_OK;
}

extern "C" int qcb_flow_qe(qcb_lattice* lat, double* Q, double* E) {
    if (!lat || !Q || !E) return QCB_EINVAL;
    if (!lat->flow_active) return qcb_set_error(lat->ctx, QCB_ESTATE, "qcb_flow_copy must come first");
    int rc = enqueue_clover(lat, 1, lat->d_out);
    if (rc) return rc;
    double h[2];
    QCB_HIP(lat->ctx, hipMemcpyAsync(h, lat->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost, lat->ctx->stream));
    QCB_HIP(lat->ctx, hipStreamSynchronize(lat->ctx->stream));
    *Q = h[0];
    *E = h[1];
    return QCB_OK;
}

extern "C" int qcb_flow_plaquette(qcb_lattice* lat, double* p) {
    if (!lat || !p) return QCB_EINVAL;
    if (!lat->flow_active) return qcb_set_error(lat->ctx, QCB_ESTATE, "qcb_flow_copy must come first");
    int rc = enqueue_plaquette(lat, 1, lat->d_out);
    if (rc) return rc;
    QCB_HIP(lat->ctx, hipMemcpyAsync(p, lat->d_out, sizeof(double), hipMemcpyDeviceToHost, lat->ctx->stream));
    QCB_HIP(lat->ctx, hipStreamSynchronize(lat->ctx->stream));
    return QCB_OK;
}

extern "C" int qcb_flow_get_links(qcb_lattice* lat, double* aos) {
    if (!lat || !aos) return QCB_EINVAL;
    if (!lat->flow_active) return qcb_set_error(lat->ctx, QCB_ESTATE, "qcb_flow_copy must come first");
    return download_field(lat, flow_current(lat), aos, kFlowOrder);
}

extern "C" int qcb_topo_charge_flowed(qcb_lattice* lat, double eps, int n_steps_gf, int n_rk, double* tQE,
                                      double* final_tQE) {
    if (!lat || n_steps_gf < 0 || n_rk < 0 || (n_steps_gf > 0 && !tQE)) return QCB_EINVAL;
    qcb_ctx* ctx = lat->ctx;
    int rc = flow_begin(lat);
    if (rc) return rc;
    double* d = nullptr;
    size_t nmeas = (size_t)n_steps_gf + 1;
    DevTmp tmp(ctx);
    if ((rc = tmp.alloc(nmeas * 2 * sizeof(double), &d))) return rc;
    std::vector<double> tv(nmeas);
    double t = 0.0;
    for (int s = 0; s < n_steps_gf; ++s) {
        tv[s] = t;
        if ((rc = enqueue_clover(lat, 1, d + 2 * s))) return rc;
        for (int r = 0; r < n_rk; ++r) {
            if ((rc = enqueue_flow_step(lat, eps))) return rc;
            t += eps;
        }
    }
    tv[n_steps_gf] = t;
    if (final_tQE && (rc = enqueue_clover(lat, 1, d + 2 * n_steps_gf))) return rc;
    std::vector<double> h(nmeas * 2);
    QCB_HIP(ctx, hipMemcpyAsync(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    QCB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int s = 0; s < n_steps_gf; ++s) {
        tQE[3 * s] = tv[s];
        tQE[3 * s + 1] = h[2 * s];
        tQE[3 * s + 2] = h[2 * s + 1];
    }
    if (final_tQE) {
        final_tQE[0] = tv[n_steps_gf];
        final_tQE[1] = h[2 * n_steps_gf];
        final_tQE[2] = h[2 * n_steps_gf + 1];
    }
    return QCB_OK;
}

// The same measurement schedule with nothing allocated: the flow buffers
// must exist already (any earlier flow call on the lattice creates them) and
// the per-measurement results are staged in the static scratch array.
extern "C" int qcb_topo_charge_flowed_noalloc(qcb_lattice* lat, double eps, int n_steps_gf, int n_rk, double* tQE,
                                              double* final_tQE) {
    if (!lat || n_steps_gf < 0 || n_rk < 0 || (n_steps_gf > 0 && !tQE)) return QCB_EINVAL;
    qcb_ctx* ctx = lat->ctx;
    const size_t nmeas = (size_t)n_steps_gf + 1;
    if (3 * nmeas > kScratchDoubles)
        return qcb_set_error(ctx, QCB_EINVAL, "at most %zu measurements per call", kScratchDoubles / 3 - 1);
    if (!lat->F[0])
        return qcb_set_error(ctx, QCB_ESTATE, "the lattice has no flow buffers yet: one flow call (e.g. "
                                              "qcb_flow_copy) creates them");
    int rc = flow_begin(lat);   // buffers present: copies the links, allocates nothing
    if (rc) return rc;
    double* d = ctx->scratch;   // measurement s at d[3s .. 3s+2]: Q, E/V written at 3s+1, 3s+2
    for (int s = 0; s < n_steps_gf; ++s) {
        if ((rc = enqueue_clover(lat, 1, d + 3 * s + 1))) return rc;
        for (int r = 0; r < n_rk; ++r)
            if ((rc = enqueue_flow_step(lat, eps))) return rc;
    }
    if (final_tQE && (rc = enqueue_clover(lat, 1, d + 3 * n_steps_gf + 1))) return rc;
    if (n_steps_gf > 0)
        QCB_HIP(ctx, hipMemcpyAsync(tQE, d, 3 * (size_t)n_steps_gf * sizeof(double), hipMemcpyDeviceToHost,
                                    ctx->stream));
    if (final_tQE)
        QCB_HIP(ctx, hipMemcpyAsync(final_tQE, d + 3 * n_steps_gf, 3 * sizeof(double), hipMemcpyDeviceToHost,
                                    ctx->stream));
    QCB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // flow times accumulated as the reference does (t += eps per RK3 step)
    double t = 0.0;
    for (int s = 0; s < n_steps_gf; ++s) {
        tQE[3 * s] = t;
        for (int r = 0; r < n_rk; ++r) t += eps;
    }
    if (final_tQE) final_tQE[0] = t;
    return QCB_OK;
}

// ---------------------------------------------------------- math probes ---

extern "C" int qcb_math_probe(qcb_ctx* ctx, int prec, int op, int n, const double* in, double param, int aux, double* out) {
    if (!ctx || !in || !out || n < 0 || op < 0 || op > 5) return QCB_EINVAL;
    if ((op == 3 || op == 4) && (aux < 1 || aux > 64)) return QCB_EINVAL;
    if (op == 5)
        for (int t = 0; t < n; ++t)
            if (!(in[t] >= 0.0 && in[t] <= 4294967295.0 && in[t] == (double)(uint64_t)in[t])) return QCB_EINVAL;
    if (n == 0) return QCB_OK;
    size_t win = op <= 2 ? 18 : op == 3 ? (size_t)(1 + 4 * aux + 2) : op == 4 ? (size_t)(36 + 3 * (4 * aux + 2)) : 1;
    size_t wout = op == 3 ? 5 : op == 5 ? 1 : 18;
    double *di = nullptr, *dout = nullptr;
    DevTmp tin(ctx), tout(ctx);
    if (int rc = tin.alloc(win * n * sizeof(double), &di)) return rc;
    if (int rc = tout.alloc(wout * n * sizeof(double), &dout)) return rc;
    QCB_HIP(ctx, hipMemcpyAsync(di, in, win * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (prec == QCB_F64) launch_math_probe<double>(op, n, di, param, aux, dout, ctx->stream);
    else launch_math_probe<float>(op, n, di, param, aux, dout, ctx->stream);
    QCB_HIP(ctx, hipGetLastError());
    QCB_HIP(ctx, hipMemcpyAsync(out, dout, wout * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    QCB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QCB_OK;
}
