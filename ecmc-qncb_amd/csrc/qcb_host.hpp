// qcb_host.hpp -- host-side handle types of the C-ABI (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/qcb.h"
#include "qcb_internal.hpp"

struct qcb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    int n_live = 0;   // lattices and batches created on this context and not yet destroyed
    uint32_t kp_cap = 1024;   // Kennedy-Pendleton attempts per draw for Philox heatbath sweeps
    double* scratch = nullptr;   // this device's address of the library's static scratch array
};

namespace qcb {
enum Family { kFamHalf = 0, kFamSite = 1, kFamLink = 2, kFamFlow = 3, kFamClover = 4 };
constexpr int kNumFam = 5;
}

struct qcb_lattice {
    qcb_ctx* ctx = nullptr;
    int prec = QCB_F64;
    int L[4] = {0, 0, 0, 0};
    uint32_t V = 0, Vh = 0, S = 0;
    int esize = 8;
    void* U = nullptr;
    void* F[3] = {nullptr, nullptr, nullptr};
    void* d_desc = nullptr;
    uint32_t* d_blk[qcb::kNumFam] = {};
    uint32_t nb[qcb::kNumFam] = {};
    double* d_partial = nullptr;   // 2 * nb[kFamClover] doubles (>= 2 * nb[kFamSite])
    double* d_out = nullptr;       // 8 doubles of results
    unsigned long long* d_counter = nullptr;
    double beta = 0.0;
    uint64_t seed = 0;
    uint32_t gstep = 0, flow_start = 0;
    bool flow_active = false;   // flow buffers hold a GradientFlow state (qcb_flow_copy)
    int tile[4] = {0, 1, 1, 1};  // work-order tile extents (tile[0] unused: whole x rows)
    uint32_t sweep_off = 0;      // per-lattice offset of the Philox sweep index
    int morton = 0;              // Z-order over tiles (needs power-of-two tile counts)
    uint32_t kp_cap = 1024;      // attempt cap in the uploaded descriptor
    unsigned long long* kp_exh = nullptr;   // where the heatbath counts exhausted draws (d_counter by default)
    int n_users = 0;             // event-chain handles reading this lattice (it must outlive them)
    uint32_t flow_limit = 0xFFFFFFFFu;   // RK3 steps after flow_start (the service's per-request t_max)
    double* qe_out = nullptr;    // per-step measurement slots (the service's result area)
};

int qcb_set_error(qcb_ctx* ctx, int code, const char* fmt, ...);

#define QCB_HIP(ctx, expr)                                                                                  \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            return qcb_set_error((ctx), QCB_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                 __FILE__, __LINE__);                                                       \
    } while (0)

namespace qcb {

// Bytes of one link field (8 half-site classes of S links): 18 reals per link
// in FP64, 12 in FP32 (rows 0 and 1; the kernels rebuild row 2).
inline uint64_t field_bytes(const qcb_lattice* l) {
    return (uint64_t)8 * (l->esize == 4 ? 12 : 18) * l->S * (uint64_t)l->esize;
}
// Bytes of the flow's force accumulator: 8 reals per link (a traceless
// anti-Hermitian matrix).
inline uint64_t force_bytes(const qcb_lattice* l) { return (uint64_t)8 * 8 * l->S * (uint64_t)l->esize; }
// Flow buffers: two link fields (W double buffer) and the force accumulator.
inline uint64_t flow_bytes(const qcb_lattice* l) { return 2 * field_bytes(l) + force_bytes(l); }

// Default work-order tile extent along one direction: largest divisor <= 4.
inline int default_tile(int L) {
    for (int t = 4; t > 1; --t)
        if (L % t == 0) return t;
    return 1;
}

// Largest divisor of L that is <= cap.
inline int divisor_le(int L, int cap) {
    for (int t = cap; t > 1; --t)
        if (L % t == 0) return t;
    return 1;
}

// Default tiling: 8 x 4 x 4 sites in (y, z, t) (largest divisors), tiles
// visited y-fastest.  Z-order over the tiles (qcb_lattice_set_tiling) keeps
// neighbouring tiles a little closer in time but costs every thread ~54
// instructions of bit decoding; at 32^4 the y-fastest order is 1.4% faster
// for the flow stage and 1.3% for the clover, and equal for the heatbath.
inline void default_tiling(const int L[4], int tile[4], int* morton) {
    tile[0] = 0;
    tile[1] = divisor_le(L[1], 8);
    tile[2] = divisor_le(L[2], 4);
    tile[3] = divisor_le(L[3], 4);
    *morton = 0;
}

// Blocks one lattice contributes to a launch of the given kernel family.
inline uint32_t fam_blocks(const qcb_lattice& l, int fam) {
    const uint32_t per = fam == kFamHalf   ? kHalfSites
                         : fam == kFamSite ? kSiteSites
                         : fam == kFamFlow ? flow_block_sites((uint32_t)l.L[0] / 2u)
                         : fam == kFamClover ? 4u * flow_block_sites((uint32_t)l.L[0] / 2u)
                                             : kLinkSites;
    return (l.Vh + per - 1) / per;
}

// The multiplier of FastDiv for divisor d >= 1.
inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    FastDiv f;
    f.m = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
    f.s1 = l ? 1u : 0u;
    f.s2 = l ? l - 1 : 0u;
    return f;
}

template <class T>
inline LatDev<T> make_desc(const qcb_lattice* l) {
    LatDev<T> d{};
    d.U = (T*)l->U;
    for (int k = 0; k < 3; ++k) d.F[k] = (T*)l->F[k];
    for (int k = 0; k < 4; ++k) d.L[k] = l->L[k];
    d.Lxh = l->L[0] / 2;
    d.Vh = l->Vh;
    d.S = l->S;
    d.beta = l->beta;
    d.key0 = (uint32_t)l->seed;
    d.key1 = (uint32_t)(l->seed >> 32);
    d.flow_start = l->flow_start;
    d.sweep_off = l->sweep_off;
    d.TL[0] = l->L[0];
    d.NT[0] = 1;
    for (int k = 1; k < 4; ++k) {
        d.TL[k] = l->tile[k];
        d.NT[k] = l->L[k] / l->tile[k];
    }
    d.tvol = (uint32_t)(l->L[0] / 2) * l->tile[1] * l->tile[2] * l->tile[3];
    d.morton = l->morton;
    d.kp_cap = l->kp_cap;
    d.kp_exh = l->kp_exh;
    d.flow_limit = l->flow_limit;
    d.fsb = flow_block_sites((uint32_t)l->L[0] / 2u);
    d.qe_out = l->qe_out;
    d.tbits[0] = 0;
    for (int k = 1; k < 4; ++k) {
        int b = 0;
        while ((1 << b) < d.NT[k]) ++b;
        d.tbits[k] = b;
    }
    d.dv_tvol = make_fastdiv(d.tvol);
    d.dv_lxh = make_fastdiv((uint32_t)d.Lxh);
    d.dv_t1 = make_fastdiv((uint32_t)d.TL[1]);
    d.dv_t2 = make_fastdiv((uint32_t)d.TL[2]);
    d.dv_nt1 = make_fastdiv((uint32_t)d.NT[1]);
    d.dv_nt2 = make_fastdiv((uint32_t)d.NT[2]);
    return d;
}

inline int upload_desc(qcb_lattice* l) {
    if (l->prec == QCB_F64) {
        LatDev<double> d = make_desc<double>(l);
        QCB_HIP(l->ctx, hipMemcpyAsync(l->d_desc, &d, sizeof d, hipMemcpyHostToDevice, l->ctx->stream));
    } else {
        LatDev<float> d = make_desc<float>(l);
        QCB_HIP(l->ctx, hipMemcpyAsync(l->d_desc, &d, sizeof d, hipMemcpyHostToDevice, l->ctx->stream));
    }
    // the host copy is a stack temporary: wait for the copy before returning
    QCB_HIP(l->ctx, hipStreamSynchronize(l->ctx->stream));
    return QCB_OK;
}

inline Launch lat_launch(const qcb_lattice* l, int fam) {
    Launch g;
    g.lats = l->d_desc;
    g.blk = l->d_blk[fam];
    g.n_lat = 1;
    g.nblocks = l->nb[fam];
    g.stream = l->ctx->stream;
    return g;
}

inline int lattice_alloc(qcb_lattice* l) {
    qcb_ctx* c = l->ctx;
    if (hipMalloc(&l->U, field_bytes(l)) != hipSuccess)
        return qcb_set_error(c, QCB_ENOMEM, "cannot allocate %llu bytes of links", (unsigned long long)field_bytes(l));
    QCB_HIP(c, hipMalloc(&l->d_desc, sizeof(LatDev<double>)));
    for (int f = 0; f < kNumFam; ++f) {
        l->nb[f] = fam_blocks(*l, f);
        uint32_t h[2] = {0, l->nb[f]};
        QCB_HIP(c, hipMalloc((void**)&l->d_blk[f], sizeof h));
        QCB_HIP(c, hipMemcpy(l->d_blk[f], h, sizeof h, hipMemcpyHostToDevice));
    }
    QCB_HIP(c, hipMalloc((void**)&l->d_partial, 2 * sizeof(double) * l->nb[kFamClover]));
    QCB_HIP(c, hipMalloc((void**)&l->d_out, 8 * sizeof(double)));
    QCB_HIP(c, hipMalloc((void**)&l->d_counter, sizeof(unsigned long long)));
    QCB_HIP(c, hipMemset(l->d_counter, 0, sizeof(unsigned long long)));
    l->kp_exh = l->d_counter;
    l->kp_cap = c->kp_cap;
    return upload_desc(l);
}

inline int ensure_flow_buffers(qcb_lattice* l) {
    if (l->F[0]) return QCB_OK;
    for (int k = 0; k < 3; ++k) {
        const uint64_t n = k < 2 ? field_bytes(l) : force_bytes(l);
        if (hipMalloc(&l->F[k], n) != hipSuccess)
            return qcb_set_error(l->ctx, QCB_ENOMEM, "cannot allocate flow buffer %d (%llu bytes)", k,
                                 (unsigned long long)n);
    }
    return upload_desc(l);
}

// Reads n exhaustion counters after the work before it on `st` (synchronises)
// and reports QCB_EBUDGET naming the first lattice whose count is non-zero.
inline int check_kp_counters(qcb_ctx* c, hipStream_t st, const unsigned long long* d, int n, uint32_t cap) {
    std::vector<unsigned long long> h((size_t)n, 0ull);
    QCB_HIP(c, hipMemcpyAsync(h.data(), d, (size_t)n * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    QCB_HIP(c, hipStreamSynchronize(st));
    unsigned long long tot = 0;
    int first = -1;
    for (int i = 0; i < n; ++i) {
        tot += h[i];
        if (h[i] && first < 0) first = i;
    }
    if (tot)
        return qcb_set_error(c, QCB_EBUDGET,
                             "%llu Kennedy-Pendleton draws ran out of their %u attempts (first in lattice %d); "
                             "those SU(2) subgroup updates were skipped",
                             tot, cap, first);
    return QCB_OK;
}

inline void lattice_free(qcb_lattice* l) {
    (void)hipFree(l->U);
    for (int k = 0; k < 3; ++k) (void)hipFree(l->F[k]);
    (void)hipFree(l->d_desc);
    for (int f = 0; f < kNumFam; ++f) (void)hipFree(l->d_blk[f]);
    (void)hipFree(l->d_partial);
    (void)hipFree(l->d_out);
    (void)hipFree(l->d_counter);
}

}  // namespace qcb
