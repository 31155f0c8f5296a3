// qcb_service.cpp -- continuous-batching flow-measurement service.
//
// A request asks for Q(t), E(t) of a freshly generated lattice: hot start
// from its seed, n_therm heatbath sweeps at its beta, then Wilson flow to
// t_max with a measurement before every RK3 step and after the last one.
// Requests wait in a FIFO queue and are admitted (first fit, later requests
// may pass a blocked head) when their footprint fits in a fixed HBM arena:
// the links, the flow's two link fields and its force field, the request's
// exhausted-draw counter and its (Q, E) series.  Every service iteration advances every resident request by
// one unit of work in batched launches: one heatbath sweep for each
// thermalising lattice on one stream, one measurement + RK3 step for each
// flowing lattice on a second stream, so new work joins the device while
// older work runs.
//
// The host never waits for an iteration it has just queued.  A request's
// device descriptor is fixed from admission: its Philox sweep index is the
// service iteration minus its admission iteration, its flow step the
// iteration minus its first flow iteration, and the kernels gate the hot
// start, reunitarisation, flow start and the flow's end per lattice from
// those counts (qcb::Gate, LatDev::flow_limit).  The two descriptor sets are
// uploaded again only when a request joins or leaves them, from pinned
// staging alternating between two buffers.  Measurements land in the
// request's own result area (LatDev::qe_out); a finished request's series
// comes back in one asynchronous copy into a pinned pool, and its completion
// is seen through an event query.  The host runs at most kDepth iterations
// ahead of the device.  A request's numbers depend only on its own
// parameters (its Philox key and sweep count, per-lattice blocks and
// fixed-order sums), so they are bitwise the same alone and in any mix.
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

constexpr int kDepth = 3;                          // iterations the host may run ahead
constexpr size_t kPoolBytes = (size_t)64 << 20;    // pinned pool for finished series
constexpr size_t kAlign = 256;

size_t align_up(size_t n) { return (n + kAlign - 1) & ~(kAlign - 1); }

enum Phase { kQueued, kTherm, kFlow, kDraining };

struct Req {
    uint64_t id = 0;
    qcb_lattice lat;            // host view: extents, tiling, arena pointers, schedule offsets
    int n_steps = 0;
    double t_arrival = 0, t_admit = 0, t_done = 0;
    Phase phase = kQueued;
    uint64_t it_admit = 0, it_flow = 0;
    size_t off = 0, bytes = 0;  // arena slice
    size_t tail_off = 0;        // [counter, pad, (Q, E) x (n_steps + 1)] inside the slice
    size_t pool_off = 0;        // the tail's copy in the pinned pool
    bool pooled = false;
    hipEvent_t done_ev = nullptr;
    unsigned long long kp_exhausted = 0;
    std::vector<double> tQE;
    size_t tail_bytes() const { return 16 + 16 * (size_t)(n_steps + 1); }
};

// Whether a gated launch over a set would update any of its lattices, by the
// kernels' own test (qcb::gate_open).  A launch whose gates are all closed is
// skipped: each of its blocks would still look up its lattice's descriptor
// before returning, ~35 us per launch with a few resident lattices.
bool any_gate_open(const std::vector<Req*>& set, const Gate& g) {
    for (const Req* r : set)
        if (g.mod == 0u || (g.sweep + r->lat.sweep_off) % g.mod == g.res) return true;
    return false;
}

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

// Descriptors and block prefix sums of one stream's lattices, on the device.
struct DevSet {
    void* d_desc = nullptr;
    uint32_t* d_blk[kNumFam] = {};
    uint32_t nb[kNumFam] = {};
    int n = 0, cap = 0;
    char* stage[2] = {nullptr, nullptr};     // pinned staging, used alternately
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    int next = 0;
    bool dirty = true;
};

}  // namespace

struct qcb_service {
    qcb_ctx* ctx = nullptr;
    int prec = QCB_F32;
    int n_hits = 2, n_therm = 50, reunit_every = 4;
    double eps = 0.02;
    char* arena_base = nullptr;
    Arena arena;
    char* pool_base = nullptr;   // pinned
    Arena pool;
    std::deque<Req*> queue, done;
    std::vector<Req*> therm, flow, draining;
    DevSet set_therm, set_flow;
    double* d_partial = nullptr;   // clover partials of the flowing set
    size_t partial_cap = 0;
    hipStream_t s_hb = nullptr, s_fl = nullptr;
    hipEvent_t ev_hb[kDepth] = {}, ev_fl[kDepth] = {};
    std::chrono::steady_clock::time_point t0;
    uint64_t n_iter = 0;
};

namespace {

double now_s(const qcb_service* s) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

size_t set_stage_bytes(int cap) { return (size_t)cap * sizeof(LatDev<double>) + kNumFam * (size_t)(cap + 1) * 4; }

void devset_free(DevSet& d) {
    (void)hipFree(d.d_desc);
    for (int f = 0; f < kNumFam; ++f) (void)hipFree(d.d_blk[f]);
    for (int k = 0; k < 2; ++k) {
        if (d.stage[k]) (void)hipHostFree(d.stage[k]);
        if (d.stage_ev[k]) (void)hipEventDestroy(d.stage_ev[k]);
        d.stage[k] = nullptr;
        d.stage_ev[k] = nullptr;
    }
    d.d_desc = nullptr;
    for (int f = 0; f < kNumFam; ++f) d.d_blk[f] = nullptr;
    d.cap = 0;
}

// Everything queued so far has finished (before memory in use is replaced).
int drain_device(qcb_service* s) {
    QCB_HIP(s->ctx, hipStreamSynchronize(s->s_hb));
    QCB_HIP(s->ctx, hipStreamSynchronize(s->s_fl));
    return QCB_OK;
}

template <class T>
void stage_descs(char* dst, const std::vector<Req*>& reqs) {
    LatDev<T>* d = reinterpret_cast<LatDev<T>*>(dst);
    for (size_t i = 0; i < reqs.size(); ++i) d[i] = make_desc<T>(&reqs[i]->lat);
}

// Uploads the set's membership on its stream (ordered after the work already
// queued there, which reads the previous contents).
int devset_upload(qcb_service* s, DevSet& d, const std::vector<Req*>& reqs, hipStream_t st) {
    qcb_ctx* c = s->ctx;
    d.dirty = false;
    d.n = (int)reqs.size();
    if (d.n == 0) {
        for (int f = 0; f < kNumFam; ++f) d.nb[f] = 0;
        return QCB_OK;
    }
    if (d.n > d.cap) {   // rare: replace the arrays once nothing reads them
        int rc = drain_device(s);
        if (rc) return rc;
        devset_free(d);
        d.cap = std::max({d.n, 2 * d.cap, 64});
        QCB_HIP(c, hipMalloc(&d.d_desc, (size_t)d.cap * sizeof(LatDev<double>)));
        for (int f = 0; f < kNumFam; ++f) QCB_HIP(c, hipMalloc((void**)&d.d_blk[f], (size_t)(d.cap + 1) * 4));
        for (int k = 0; k < 2; ++k) {
            QCB_HIP(c, hipHostMalloc((void**)&d.stage[k], set_stage_bytes(d.cap), hipHostMallocDefault));
            QCB_HIP(c, hipEventCreateWithFlags(&d.stage_ev[k], hipEventDisableTiming));
        }
    }
    const int k = d.next;
    d.next ^= 1;
    QCB_HIP(c, hipEventSynchronize(d.stage_ev[k]));   // the last copy out of this buffer is done
    char* h = d.stage[k];
    if (s->prec == QCB_F64) stage_descs<double>(h, reqs);
    else stage_descs<float>(h, reqs);
    uint32_t* pre = reinterpret_cast<uint32_t*>(h + (size_t)d.cap * sizeof(LatDev<double>));
    for (int f = 0; f < kNumFam; ++f) {
        uint32_t* p = pre + (size_t)f * (d.cap + 1);
        p[0] = 0;
        for (int i = 0; i < d.n; ++i) p[i + 1] = p[i] + fam_blocks(reqs[i]->lat, f);
        d.nb[f] = p[d.n];
        QCB_HIP(c, hipMemcpyAsync(d.d_blk[f], p, (size_t)(d.n + 1) * 4, hipMemcpyHostToDevice, st));
    }
    QCB_HIP(c, hipMemcpyAsync(d.d_desc, h, (size_t)d.n * sizeof(LatDev<double>), hipMemcpyHostToDevice, st));
    QCB_HIP(c, hipEventRecord(d.stage_ev[k], st));
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

// links, the flow fields, then the tail: counter, pad, (Q, E) per measurement
uint64_t req_bytes(const qcb_lattice& l, int n_steps) {
    return align_up(field_bytes(&l) + flow_bytes(&l)) + align_up(16 + 16 * (size_t)(n_steps + 1));
}

// The tail of a finished request is on the host: fill its series.
void unpack(qcb_service* s, Req* r, const char* tail) {
    std::memcpy(&r->kp_exhausted, tail, sizeof(unsigned long long));
    const double* qe = reinterpret_cast<const double*>(tail + 16);
    r->tQE.resize(3 * (size_t)(r->n_steps + 1));
    double t = 0.0;   // flow time accumulated as the reference does
    for (int k = 0; k <= r->n_steps; ++k) {
        r->tQE[3 * k] = t;
        r->tQE[3 * k + 1] = qe[2 * k];
        r->tQE[3 * k + 2] = qe[2 * k + 1];
        t += s->eps;
    }
}

void finish(qcb_service* s, Req* r) {
    s->arena.release(r->off, r->bytes);
    if (r->pooled) s->pool.release(r->pool_off, align_up(r->tail_bytes()));
    r->t_done = now_s(s);
    s->done.push_back(r);
}

// Moves drained requests whose copy has completed to the done queue; with
// block set, first waits for the oldest one.
int collect(qcb_service* s, bool block) {
    qcb_ctx* c = s->ctx;
    if (block && !s->draining.empty()) QCB_HIP(c, hipEventSynchronize(s->draining.front()->done_ev));
    for (auto it = s->draining.begin(); it != s->draining.end();) {
        Req* r = *it;
        const hipError_t e = hipEventQuery(r->done_ev);
        if (e == hipErrorNotReady) {
            ++it;
            continue;
        }
        if (e != hipSuccess) return qcb_set_error(c, QCB_EDEVICE, "service event: %s", hipGetErrorString(e));
        unpack(s, r, s->pool_base + r->pool_off);
        (void)hipEventDestroy(r->done_ev);
        r->done_ev = nullptr;
        finish(s, r);
        it = s->draining.erase(it);
    }
    return QCB_OK;
}

// The final measurement of r is queued on s_fl: bring its tail home.
int start_drain(qcb_service* s, Req* r) {
    qcb_ctx* c = s->ctx;
    const char* tail = s->arena_base + r->off + r->tail_off;
    if (s->pool.alloc(align_up(r->tail_bytes()), &r->pool_off)) {
        r->pooled = true;
        QCB_HIP(c, hipMemcpyAsync(s->pool_base + r->pool_off, tail, r->tail_bytes(), hipMemcpyDeviceToHost, s->s_fl));
        QCB_HIP(c, hipEventCreateWithFlags(&r->done_ev, hipEventDisableTiming));
        QCB_HIP(c, hipEventRecord(r->done_ev, s->s_fl));
        r->phase = kDraining;
        s->draining.push_back(r);
        return QCB_OK;
    }
    // the pool is full (very long series): copy synchronously
    std::vector<char> h(r->tail_bytes());
    QCB_HIP(c, hipStreamSynchronize(s->s_fl));
    QCB_HIP(c, hipMemcpy(h.data(), tail, h.size(), hipMemcpyDeviceToHost));
    unpack(s, r, h.data());
    finish(s, r);
    return QCB_OK;
}

}  // namespace

extern "C" int qcb_service_create(qcb_ctx* ctx, int prec, uint64_t arena_bytes, int n_hits, int n_therm,
                                  int reunit_every, double eps, qcb_service** out) {
    if (!ctx || !out || (prec != QCB_F32 && prec != QCB_F64) || n_hits < 1 || n_therm < 0 || eps <= 0 ||
        arena_bytes < (1u << 20) || reunit_every < 0)
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
    arena_bytes &= ~(uint64_t)(kAlign - 1);
    auto fail = [&](int code, const char* what) {
        for (int k = 0; k < kDepth; ++k) {
            if (s->ev_hb[k]) (void)hipEventDestroy(s->ev_hb[k]);
            if (s->ev_fl[k]) (void)hipEventDestroy(s->ev_fl[k]);
        }
        if (s->s_hb) (void)hipStreamDestroy(s->s_hb);
        if (s->s_fl) (void)hipStreamDestroy(s->s_fl);
        (void)hipFree(s->d_partial);
        if (s->pool_base) (void)hipHostFree(s->pool_base);
        (void)hipFree(s->arena_base);
        delete s;
        return qcb_set_error(ctx, code, "%s", what);
    };
    if (hipMalloc((void**)&s->arena_base, arena_bytes) != hipSuccess)
        return fail(QCB_ENOMEM, "cannot allocate the service's HBM arena");
    s->arena.init(arena_bytes);
    if (hipHostMalloc((void**)&s->pool_base, kPoolBytes, hipHostMallocDefault) != hipSuccess)
        return fail(QCB_ENOMEM, "cannot allocate the service's pinned result pool");
    s->pool.init(kPoolBytes);
    // clover partials for as many site blocks as the arena can hold (a block of
    // 128 cb sites of the smallest request stores at least 48 KB of links)
    s->partial_cap = 2 * (arena_bytes / 49152 + 4096);
    if (hipMalloc((void**)&s->d_partial, s->partial_cap * sizeof(double)) != hipSuccess)
        return fail(QCB_ENOMEM, "cannot allocate the service's reduction buffer");
    if (hipStreamCreateWithFlags(&s->s_hb, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->s_fl, hipStreamNonBlocking) != hipSuccess)
        return fail(QCB_EDEVICE, "cannot create the service's streams");
    for (int k = 0; k < kDepth; ++k)
        if (hipEventCreateWithFlags(&s->ev_hb[k], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_fl[k], hipEventDisableTiming) != hipSuccess)
            return fail(QCB_EDEVICE, "cannot create the service's events");
    ctx->n_live++;
    *out = s;
    return QCB_OK;
}

extern "C" int qcb_service_destroy(qcb_service* s) {
    if (!s) return QCB_EINVAL;
    (void)hipStreamSynchronize(s->s_hb);
    (void)hipStreamSynchronize(s->s_fl);
    for (Req* r : s->draining)
        if (r->done_ev) (void)hipEventDestroy(r->done_ev);
    for (auto* list : {&s->therm, &s->flow, &s->draining})
        for (Req* r : *list) delete r;
    for (Req* r : s->queue) delete r;
    for (Req* r : s->done) delete r;
    devset_free(s->set_therm);
    devset_free(s->set_flow);
    for (int k = 0; k < kDepth; ++k) {
        (void)hipEventDestroy(s->ev_hb[k]);
        (void)hipEventDestroy(s->ev_fl[k]);
    }
    (void)hipFree(s->d_partial);
    (void)hipHostFree(s->pool_base);
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
    r->n_steps = (int)std::llround(t_max / s->eps);
    r->t_arrival = t_arrival;
    r->bytes = req_bytes(l, r->n_steps);
    if (r->bytes > s->arena.size) {
        delete r;
        return qcb_set_error(s->ctx, QCB_ENOMEM, "request needs more HBM than the whole arena");
    }
    s->queue.push_back(r);
    return QCB_OK;
}

// One iteration: retire drained requests, admit, upload changed sets, queue
// the iteration's work on both streams, advance the schedule.  Returns the
// number of requests still queued or resident in *pending.
extern "C" int qcb_service_step(qcb_service* s, int* pending) {
    if (!s) return QCB_EINVAL;
    qcb_ctx* c = s->ctx;
    const bool f64 = s->prec == QCB_F64;
    const int slot = (int)(s->n_iter % kDepth);
    int rc;
    // 0. keep the host at most kDepth iterations ahead of the device
    if (s->n_iter >= (uint64_t)kDepth) {
        QCB_HIP(c, hipEventSynchronize(s->ev_hb[slot]));
        QCB_HIP(c, hipEventSynchronize(s->ev_fl[slot]));
    }
    // 1. finished requests; with nothing on the device but copies, wait for
    //    the oldest (queued requests may be waiting for its arena slice)
    if ((rc = collect(s, s->therm.empty() && s->flow.empty()))) return rc;
    // 2. admission by footprint (first fit over the FIFO queue)
    const uint32_t it = (uint32_t)s->n_iter;
    for (auto q = s->queue.begin(); q != s->queue.end();) {
        Req* r = *q;
        size_t off;
        if (!s->arena.alloc(r->bytes, &off)) {
            ++q;
            continue;
        }
        r->off = off;
        char* base = s->arena_base + off;
        const uint64_t fb = field_bytes(&r->lat);
        r->lat.U = base;
        for (int k = 0; k < 3; ++k) r->lat.F[k] = base + (k + 1) * fb;   // F[2] holds force_bytes
        r->tail_off = align_up(fb + flow_bytes(&r->lat));
        r->lat.kp_exh = (unsigned long long*)(base + r->tail_off);
        r->lat.qe_out = (double*)(base + r->tail_off + 16);
        r->lat.kp_cap = c->kp_cap;
        r->lat.sweep_off = 0u - it;   // the kernels' sweep index is it + sweep_off = sweeps done
        r->lat.flow_limit = (uint32_t)r->n_steps;
        QCB_HIP(c, hipMemsetAsync(r->lat.kp_exh, 0, sizeof(unsigned long long), s->s_hb));
        r->phase = kTherm;
        r->it_admit = s->n_iter;
        r->t_admit = now_s(s);
        s->therm.push_back(r);
        s->set_therm.dirty = true;
        q = s->queue.erase(q);
    }
    // 3. descriptor sets whose membership changed
    if (s->set_therm.dirty && (rc = devset_upload(s, s->set_therm, s->therm, s->s_hb))) return rc;
    if (s->set_flow.dirty && (rc = devset_upload(s, s->set_flow, s->flow, s->s_fl))) return rc;
    // 4. thermalising lattices: hot start at their first iteration, one sweep,
    //    reunitarisation on cadence, flow start after their last sweep
    if (!s->therm.empty()) {
        const uint32_t nt = (uint32_t)s->n_therm;
        const Launch gl = set_launch(s->set_therm, kFamLink, s->s_hb);
        const Gate first{it, nt ? nt : 1u, 0u}, last{it, nt ? nt : 1u, nt ? nt - 1u : 0u};
        if (any_gate_open(s->therm, first)) {
            if (f64) launch_hot<double>(gl, first); else launch_hot<float>(gl, first);
        }
        if (nt) {
            const Launch gh = set_launch(s->set_therm, kFamHalf, s->s_hb);
            for (int hs = 0; hs < 8; ++hs) {
                if (f64) launch_hb_half<double>(gh, hs >> 1, hs & 1, s->n_hits, it, kHbPhilox, nullptr, 0);
                else launch_hb_half<float>(gh, hs >> 1, hs & 1, s->n_hits, it, kHbPhilox, nullptr, 0);
            }
            const Gate ru{it, (uint32_t)s->reunit_every, (uint32_t)s->reunit_every - 1u};
            if (s->reunit_every > 0 && any_gate_open(s->therm, ru)) {
                if (f64) launch_reunit<double>(gl, ru); else launch_reunit<float>(gl, ru);
            }
        }
        if (any_gate_open(s->therm, last)) {
            if (f64) launch_flow_init<double>(gl, last); else launch_flow_init<float>(gl, last);
        }
    }
    QCB_HIP(c, hipEventRecord(s->ev_hb[slot], s->s_hb));
    // 5. flowing lattices: measure every one, step those with steps left; a
    //    lattice that started flowing this iteration was prepared on s_hb one
    //    iteration ago
    if (!s->flow.empty()) {
        if (s->n_iter > 0) QCB_HIP(c, hipStreamWaitEvent(s->s_fl, s->ev_hb[(s->n_iter - 1) % kDepth], 0));
        const Launch gs = set_launch(s->set_flow, kFamClover, s->s_fl);
        if (2 * (size_t)s->set_flow.nb[kFamClover] > s->partial_cap) {   // rare: more small requests than assumed
            if ((rc = drain_device(s))) return rc;
            (void)hipFree(s->d_partial);
            s->partial_cap = 4 * (size_t)s->set_flow.nb[kFamClover];
            QCB_HIP(c, hipMalloc((void**)&s->d_partial, s->partial_cap * sizeof(double)));
        }
        double* pq = s->d_partial;
        double* pe = s->d_partial + s->set_flow.nb[kFamClover];
        if (f64) {
            launch_clover<double>(gs, 1, it, pq, pe);
            launch_finalize_qe<double>(gs, pq, pe, nullptr, 0, it);
        } else {
            launch_clover<float>(gs, 1, it, pq, pe);
            launch_finalize_qe<float>(gs, pq, pe, nullptr, 0, it);
        }
        const Launch gl = set_launch(s->set_flow, kFamFlow, s->s_fl);
        for (int st = 1; st <= 3; ++st) {
            if (f64) launch_flow_stage<double>(gl, st, it, s->eps);
            else launch_flow_stage<float>(gl, st, it, s->eps);
        }
    }
    QCB_HIP(c, hipGetLastError());
    // 6. the schedule: requests at their final measurement drain, those at
    //    their last sweep flow from the next iteration on
    for (auto f = s->flow.begin(); f != s->flow.end();) {
        Req* r = *f;
        if (s->n_iter - r->it_flow == (uint64_t)r->n_steps) {
            if ((rc = start_drain(s, r))) return rc;
            f = s->flow.erase(f);
            s->set_flow.dirty = true;
        } else {
            ++f;
        }
    }
    QCB_HIP(c, hipEventRecord(s->ev_fl[slot], s->s_fl));
    for (auto t = s->therm.begin(); t != s->therm.end();) {
        Req* r = *t;
        if (s->n_therm == 0 || s->n_iter - r->it_admit == (uint64_t)s->n_therm - 1) {
            r->phase = kFlow;
            r->it_flow = s->n_iter + 1;
            r->lat.flow_start = it + 1u;
            s->flow.push_back(r);
            t = s->therm.erase(t);
            s->set_therm.dirty = s->set_flow.dirty = true;
        } else {
            ++t;
        }
    }
    s->n_iter++;
    if (pending)
        *pending = (int)(s->queue.size() + s->therm.size() + s->flow.size() + s->draining.size());
    return QCB_OK;
}

// Pops one finished request.  tQE (capacity cap triples) receives (t, Q,
// E_mean) at t = 0, eps, ..., n_steps*eps; times are on the service clock.
// Returns QCB_ESTATE when nothing has finished.
extern "C" int qcb_service_poll(qcb_service* s, uint64_t* id, double* tQE, int cap, int* n_meas, double* times3) {
    if (!s || !id || !n_meas) return QCB_EINVAL;
    int rc = collect(s, false);
    if (rc) return rc;
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

// queued, resident (thermalising, flowing or being copied back),
// finished-not-polled, free arena bytes, iterations
extern "C" int qcb_service_status(qcb_service* s, int* queued, int* resident, int* finished, uint64_t* free_bytes,
                                  uint64_t* iterations) {
    if (!s) return QCB_EINVAL;
    if (queued) *queued = (int)s->queue.size();
    if (resident) *resident = (int)(s->therm.size() + s->flow.size() + s->draining.size());
    if (finished) *finished = (int)s->done.size();
    if (free_bytes) *free_bytes = s->arena.free_bytes();
    if (iterations) *iterations = s->n_iter;
    return QCB_OK;
}
