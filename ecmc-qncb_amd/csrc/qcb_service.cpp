// qcb_service.cpp -- continuous-batching flow-measurement service.
//
// A request asks for Q(t), E(t) of a freshly generated lattice: hot start
// from its seed, n_therm heatbath sweeps at its beta, then Wilson flow to
// t_max with a measurement before every RK3 step and after the last one.
// Requests wait in a FIFO queue and are admitted (first fit, later requests
// may pass a blocked head) when their four link-sized fields fit in a
// fixed HBM arena; once admitted they stay resident until done.  Every
// service iteration advances ALL resident requests by one unit of work in
// batched launches: one heatbath sweep for every thermalising lattice (on
// one stream) and one measurement + RK3 step for every flowing lattice (on
// a second stream), so new work joins the device while older work runs.
// A request's numbers depend only on its own parameters (per-request Philox
// key and sweep count, per-lattice blocks and fixed-order sums), so they are
// bitwise the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "../../include/qcb.h"
#include "qcb_host.hpp"
#include "qcb_internal.hpp"

using namespace qcb;

namespace {

struct Req {
    uint64_t id = 0;
    qcb_lattice lat;            // host view: extents, tiling, arena pointers
    double t_max = 0.0;
    int n_steps = 0;
    double t_arrival = 0, t_admit = 0, t_done = 0;
    int state = 0;              // 0 queued, 1 thermalising, 2 flowing, 3 done
    bool fresh = false;         // hot start pending
    int sweeps_done = 0;
    int flow_step = 0;
    double t_flow = 0.0;
    size_t off = 0, bytes = 0;  // arena slice: links, flow fields, then the exhausted-draw counter
    unsigned long long kp_exhausted = 0;
    std::vector<double> tQE;
};

// First-fit free list over [0, size) with coalescing.
struct Arena {
    size_t size = 0;
    std::map<size_t, size_t> free_;   // offset -> length
    void init(size_t n) {
        size = n;
        free_.clear();
        free_[0] = n;
    }
    bool alloc(size_t n, size_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                *off = it->first;
                size_t rest = it->second - n, at = it->first + n;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t n) {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t free_bytes() const {
        size_t s = 0;
        for (auto& kv : free_) s += kv.second;
        return s;
    }
};

// Device-side descriptor set for one batched launch.
struct DevSet {
    void* d_desc = nullptr;
    uint32_t* d_blk[3] = {nullptr, nullptr, nullptr};
    uint32_t nb[3] = {0, 0, 0};
    int cap = 0, n = 0;
};

}  // namespace

struct qcb_service {
    qcb_ctx* ctx = nullptr;
    int prec = QCB_F32;
    int n_hits = 2, n_therm = 50, reunit_every = 4;
    double eps = 0.02;
    char* arena_base = nullptr;
    Arena arena;
    std::deque<Req*> queue;
    std::vector<Req*> active;
    std::deque<Req*> done;
    DevSet set_hot, set_hb, set_ru, set_init, set_fl, set_fs;
    double* d_partial = nullptr;
    size_t partial_cap = 0;
    double* d_res = nullptr;
    size_t res_cap = 0;
    hipStream_t s_hb = nullptr, s_fl = nullptr;
    uint32_t gstep = 0;
    std::chrono::steady_clock::time_point t0;
    uint64_t n_iter = 0;
};

namespace {

double now_s(const qcb_service* s) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

int devset_fill(qcb_service* s, DevSet& d, const std::vector<Req*>& reqs, hipStream_t st) {
    qcb_ctx* c = s->ctx;
    d.n = (int)reqs.size();
    if (d.n == 0) return QCB_OK;
    if (d.n > d.cap) {
        (void)hipFree(d.d_desc);
        for (int f = 0; f < 3; ++f) (void)hipFree(d.d_blk[f]);
        d.cap = std::max(d.n, 2 * d.cap);
        QCB_HIP(c, hipMalloc(&d.d_desc, (size_t)d.cap * sizeof(LatDev<double>)));
        for (int f = 0; f < 3; ++f) QCB_HIP(c, hipMalloc((void**)&d.d_blk[f], (size_t)(d.cap + 1) * sizeof(uint32_t)));
    }
    std::vector<uint32_t> pre[3];
    for (int f = 0; f < 3; ++f) {
        pre[f].assign(d.n + 1, 0);
        for (int i = 0; i < d.n; ++i) pre[f][i + 1] = pre[f][i] + fam_blocks(reqs[i]->lat.Vh, f);
        d.nb[f] = pre[f][d.n];
        QCB_HIP(c, hipMemcpyAsync(d.d_blk[f], pre[f].data(), pre[f].size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    }
    if (s->prec == QCB_F64) {
        std::vector<LatDev<double>> h(d.n);
        for (int i = 0; i < d.n; ++i) h[i] = make_desc<double>(&reqs[i]->lat);
        QCB_HIP(c, hipMemcpyAsync(d.d_desc, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice, st));
        QCB_HIP(c, hipStreamSynchronize(st));   // host staging goes out of scope
    } else {
        std::vector<LatDev<float>> h(d.n);
        for (int i = 0; i < d.n; ++i) h[i] = make_desc<float>(&reqs[i]->lat);
        QCB_HIP(c, hipMemcpyAsync(d.d_desc, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice, st));
        QCB_HIP(c, hipStreamSynchronize(st));
    }
    return QCB_OK;
}

Launch set_launch(const DevSet& d, int fam, hipStream_t st) {
    Launch g;
    g.lats = d.d_desc;
    g.blk = d.d_blk[fam];
    g.n_lat = d.n;
    g.nblocks = d.nb[fam];
    g.stream = st;
    return g;
}

void devset_free(DevSet& d) {
    (void)hipFree(d.d_desc);
    for (int f = 0; f < 3; ++f) (void)hipFree(d.d_blk[f]);
}

// links, the flow fields and a 256-byte tail for the request's counter
uint64_t req_bytes(const qcb_lattice& l) { return field_bytes(&l) + flow_bytes(&l) + 256; }

}  // namespace

extern "C" int qcb_service_create(qcb_ctx* ctx, int prec, uint64_t arena_bytes, int n_hits, int n_therm,
                                  int reunit_every, double eps, qcb_service** out) {
    if (!ctx || !out || (prec != QCB_F32 && prec != QCB_F64) || n_hits < 1 || n_therm < 0 || eps <= 0 ||
        arena_bytes < (1u << 20))
        return QCB_EINVAL;
    *out = nullptr;
    qcb_service* s = new qcb_service();
    s->ctx = ctx;
    s->prec = prec;
    s->n_hits = n_hits;
    s->n_therm = n_therm;
    s->reunit_every = reunit_every;
    s->eps = eps;
    s->t0 = std::chrono::steady_clock::now();
    arena_bytes &= ~(uint64_t)4095;
    if (hipMalloc((void**)&s->arena_base, arena_bytes) != hipSuccess) {
        delete s;
        return qcb_set_error(ctx, QCB_ENOMEM, "cannot allocate a %llu-byte arena", (unsigned long long)arena_bytes);
    }
    s->arena.init(arena_bytes);
    QCB_HIP(ctx, hipStreamCreateWithFlags(&s->s_hb, hipStreamNonBlocking));
    QCB_HIP(ctx, hipStreamCreateWithFlags(&s->s_fl, hipStreamNonBlocking));
    ctx->n_live++;
    *out = s;
    return QCB_OK;
}

extern "C" int qcb_service_destroy(qcb_service* s) {
    if (!s) return QCB_EINVAL;
    (void)hipStreamSynchronize(s->s_hb);
    (void)hipStreamSynchronize(s->s_fl);
    for (Req* r : s->queue) delete r;
    for (Req* r : s->active) delete r;
    for (Req* r : s->done) delete r;
    devset_free(s->set_hot);
    devset_free(s->set_hb);
    devset_free(s->set_ru);
    devset_free(s->set_init);
    devset_free(s->set_fl);
    devset_free(s->set_fs);
    (void)hipFree(s->d_partial);
    (void)hipFree(s->d_res);
    (void)hipFree(s->arena_base);
    (void)hipStreamDestroy(s->s_hb);
    (void)hipStreamDestroy(s->s_fl);
    s->ctx->n_live--;
    delete s;
    return QCB_OK;
}

extern "C" double qcb_service_clock(qcb_service* s) { return s ? now_s(s) : 0.0; }

extern "C" int qcb_service_submit(qcb_service* s, uint64_t id, const int dims[4], double beta, uint64_t seed,
                                  double t_max, double t_arrival) {
    if (!s || !dims) return QCB_EINVAL;
    if (!(beta > 0) || !(t_max >= 0) || !(t_max / s->eps <= 1e8))
        return qcb_set_error(s->ctx, QCB_EINVAL, "need beta > 0 and 0 <= t_max <= 1e8 eps");
    for (int d = 0; d < 4; ++d)
        if (dims[d] < 4 || dims[d] % 2) return qcb_set_error(s->ctx, QCB_EINVAL, "extents must be even and >= 4");
    if ((uint64_t)dims[0] * dims[1] * dims[2] * dims[3] * 4 >= (1ull << 31))
        return qcb_set_error(s->ctx, QCB_EINVAL, "lattice too large");
    Req* r = new Req();
    r->id = id;
    qcb_lattice& l = r->lat;
    l.ctx = s->ctx;
    l.prec = s->prec;
    for (int d = 0; d < 4; ++d) l.L[d] = dims[d];
    uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2] * dims[3];
    l.V = (uint32_t)V;
    l.Vh = (uint32_t)(V / 2);
    l.S = (l.Vh + 63u) & ~63u;
    l.esize = s->prec == QCB_F64 ? 8 : 4;
    default_tiling(dims, l.tile, &l.morton);   // the same tiling as a lattice created alone
    l.beta = beta;
    l.seed = seed;
    r->t_max = t_max;
    r->n_steps = (int)std::llround(t_max / s->eps);
    r->t_arrival = t_arrival;
    r->bytes = req_bytes(l);
    if (r->bytes > s->arena.size) {
        delete r;
        return qcb_set_error(s->ctx, QCB_ENOMEM, "request needs more HBM than the whole arena");
    }
    s->queue.push_back(r);
    return QCB_OK;
}

// One service iteration.  Returns the number of requests still queued or
// resident in *pending.
extern "C" int qcb_service_step(qcb_service* s, int* pending) {
    if (!s) return QCB_EINVAL;
    qcb_ctx* c = s->ctx;
    int rc;
    // 1. admission by footprint (first fit over the FIFO queue)
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        Req* r = *it;
        size_t off;
        if (s->arena.alloc(r->bytes, &off)) {
            r->off = off;
            char* base = s->arena_base + off;
            const uint64_t fb = field_bytes(&r->lat);
            r->lat.U = base;
            for (int k = 0; k < 3; ++k) r->lat.F[k] = base + (k + 1) * fb;   // F[2] holds force_bytes
            r->lat.kp_exh = (unsigned long long*)(base + fb + flow_bytes(&r->lat));
            r->lat.kp_cap = c->kp_cap;
            QCB_HIP(c, hipMemsetAsync(r->lat.kp_exh, 0, sizeof(unsigned long long), s->s_hb));
            r->state = 1;
            r->fresh = true;
            r->t_admit = now_s(s);
            s->active.push_back(r);
            it = s->queue.erase(it);
        } else {
            ++it;
        }
    }
    // 2. partition the resident work (every list is known before any launch,
    //    so all descriptor uploads happen first and the two streams overlap)
    std::vector<Req*> hot, hb, ru, init, fl, fstep;
    for (Req* r : s->active) {
        if (r->state == 1) {
            if (r->fresh) hot.push_back(r);
            const bool sweeps = r->sweeps_done < s->n_therm;
            if (sweeps) {
                hb.push_back(r);
                if (s->reunit_every > 0 && (r->sweeps_done + 1) % s->reunit_every == 0) ru.push_back(r);
            }
            if (r->sweeps_done + (sweeps ? 1 : 0) >= s->n_therm) init.push_back(r);
        } else if (r->state == 2) {
            fl.push_back(r);
            if (r->flow_step < r->n_steps) fstep.push_back(r);
        }
    }
    for (Req* r : hb) r->lat.sweep_off = (uint32_t)r->sweeps_done;
    if ((rc = devset_fill(s, s->set_hot, hot, s->s_hb))) return rc;
    if ((rc = devset_fill(s, s->set_hb, hb, s->s_hb))) return rc;
    if ((rc = devset_fill(s, s->set_ru, ru, s->s_hb))) return rc;
    if ((rc = devset_fill(s, s->set_init, init, s->s_hb))) return rc;
    if ((rc = devset_fill(s, s->set_fl, fl, s->s_fl))) return rc;
    if ((rc = devset_fill(s, s->set_fs, fstep, s->s_fl))) return rc;
    if (!fl.empty()) {
        const size_t need_p = 2 * (size_t)s->set_fl.nb[kFamSite];
        if (need_p > s->partial_cap) {
            (void)hipFree(s->d_partial);
            s->partial_cap = std::max(need_p, 2 * s->partial_cap);
            QCB_HIP(c, hipMalloc((void**)&s->d_partial, s->partial_cap * sizeof(double)));
        }
        if (2 * fl.size() > s->res_cap) {
            (void)hipFree(s->d_res);
            s->res_cap = std::max(2 * fl.size(), 2 * s->res_cap);
            QCB_HIP(c, hipMalloc((void**)&s->d_res, s->res_cap * sizeof(double)));
        }
    }
    const bool f64 = s->prec == QCB_F64;
    // 3. thermalising lattices on s_hb: hot start, one sweep, reunitarisation, flow start
    if (!hot.empty()) {
        Launch g = set_launch(s->set_hot, kFamLink, s->s_hb);
        if (f64) launch_hot<double>(g); else launch_hot<float>(g);
        for (Req* r : hot) r->fresh = false;
    }
    if (!hb.empty()) {
        Launch g = set_launch(s->set_hb, kFamHalf, s->s_hb);
        for (int hs = 0; hs < 8; ++hs) {
            if (f64) launch_hb_half<double>(g, hs >> 1, hs & 1, s->n_hits, 0, kHbPhilox, nullptr, 0);
            else launch_hb_half<float>(g, hs >> 1, hs & 1, s->n_hits, 0, kHbPhilox, nullptr, 0);
        }
        for (Req* r : hb) r->sweeps_done++;
    }
    if (!ru.empty()) {
        Launch g = set_launch(s->set_ru, kFamLink, s->s_hb);
        if (f64) launch_reunit<double>(g); else launch_reunit<float>(g);
    }
    if (!init.empty()) {
        Launch g = set_launch(s->set_init, kFamLink, s->s_hb);
        if (f64) launch_flow_init<double>(g); else launch_flow_init<float>(g);
    }
    // 4. flowing lattices on s_fl: measure every one, step those not finished
    if (!fl.empty()) {
        Launch gs = set_launch(s->set_fl, kFamSite, s->s_fl);
        double* pq = s->d_partial;
        double* pe = s->d_partial + s->set_fl.nb[kFamSite];
        if (f64) {
            launch_clover<double>(gs, 1, s->gstep, pq, pe);
            launch_finalize_qe<double>(gs, pq, pe, s->d_res, 2);
        } else {
            launch_clover<float>(gs, 1, s->gstep, pq, pe);
            launch_finalize_qe<float>(gs, pq, pe, s->d_res, 2);
        }
    }
    if (!fstep.empty()) {
        Launch gl = set_launch(s->set_fs, kFamLink, s->s_fl);
        for (int st = 1; st <= 3; ++st) {
            if (f64) launch_flow_stage<double>(gl, st, s->gstep, s->eps);
            else launch_flow_stage<float>(gl, st, s->gstep, s->eps);
        }
    }
    QCB_HIP(c, hipGetLastError());
    std::vector<double> res(2 * fl.size());
    if (!fl.empty())
        QCB_HIP(c, hipMemcpyAsync(res.data(), s->d_res, res.size() * sizeof(double), hipMemcpyDeviceToHost, s->s_fl));
    QCB_HIP(c, hipStreamSynchronize(s->s_hb));
    QCB_HIP(c, hipStreamSynchronize(s->s_fl));
    const double t = now_s(s);
    // 5. bookkeeping: record measurements, retire finished requests
    for (size_t i = 0; i < fl.size(); ++i) {
        Req* r = fl[i];
        r->tQE.push_back(r->t_flow);
        r->tQE.push_back(res[2 * i]);
        r->tQE.push_back(res[2 * i + 1]);
        if (r->flow_step == r->n_steps) {
            r->state = 3;
            r->t_done = t;
        } else {
            r->flow_step++;
            r->t_flow += s->eps;
        }
    }
    s->gstep++;
    for (Req* r : init) {   // measured first at the next iteration's gstep
        r->state = 2;
        r->lat.flow_start = s->gstep;
    }
    for (auto it = s->active.begin(); it != s->active.end();) {
        Req* r = *it;
        if (r->state == 3) {
            QCB_HIP(c, hipMemcpy(&r->kp_exhausted, r->lat.kp_exh, sizeof(unsigned long long), hipMemcpyDeviceToHost));
            s->arena.release(r->off, r->bytes);
            s->done.push_back(r);
            it = s->active.erase(it);
        } else {
            ++it;
        }
    }
    s->n_iter++;
    if (pending) *pending = (int)(s->queue.size() + s->active.size());
    return QCB_OK;
}

// Pops one finished request.  tQE (capacity cap triples) receives (t, Q,
// E_mean) at t = 0, eps, ..., n_steps*eps; times are on the service clock.
// Returns QCB_ESTATE when nothing has finished.
extern "C" int qcb_service_poll(qcb_service* s, uint64_t* id, double* tQE, int cap, int* n_meas, double* times3) {
    if (!s || !id || !n_meas) return QCB_EINVAL;
    if (s->done.empty()) return QCB_ESTATE;
    Req* r = s->done.front();
    int n = (int)(r->tQE.size() / 3);
    if (tQE && cap < n) return qcb_set_error(s->ctx, QCB_EINVAL, "result needs %d triples", n);
    s->done.pop_front();
    *id = r->id;
    *n_meas = n;
    if (tQE) std::memcpy(tQE, r->tQE.data(), r->tQE.size() * sizeof(double));
    if (times3) {
        times3[0] = r->t_arrival;
        times3[1] = r->t_admit;
        times3[2] = r->t_done;
    }
    const unsigned long long exh = r->kp_exhausted;
    delete r;
    if (exh)
        return qcb_set_error(s->ctx, QCB_EBUDGET,
                             "request %llu: %llu Kennedy-Pendleton draws ran out of attempts during thermalisation "
                             "(results returned)",
                             (unsigned long long)*id, exh);
    return QCB_OK;
}

// queued, resident, finished-not-polled, free arena bytes, iterations
extern "C" int qcb_service_status(qcb_service* s, int* queued, int* resident, int* finished, uint64_t* free_bytes,
                                  uint64_t* iterations) {
    if (!s) return QCB_EINVAL;
    if (queued) *queued = (int)s->queue.size();
    if (resident) *resident = (int)s->active.size();
    if (finished) *finished = (int)s->done.size();
    if (free_bytes) *free_bytes = s->arena.free_bytes();
    if (iterations) *iterations = s->n_iter;
    return QCB_OK;
}
