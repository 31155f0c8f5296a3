// qcb_kernels.hip -- gfx950 kernels of the heatbath / Wilson flow / observable
// path, templated on the link type (float or double).  Every kernel runs over
// a batch of lattices described by LatDev entries; a block belongs to exactly
// one lattice, so per-lattice results do not depend on what else is batched.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "qcb_internal.hpp"
#include "qcb_layout.hpp"
#include "qcb_rng.hpp"
#include "qcb_su3.hpp"

namespace qcb {

namespace {

template <class T>
struct BlockCtx {
    const LatDev<T>* L;
    uint32_t local;   // block index within its lattice
    uint32_t tile;    // global tile index (owner of partial slot)
};

template <class T>
__device__ __forceinline__ BlockCtx<T> block_ctx(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                 int n_lat, uint32_t nblocks) {
    BlockCtx<T> c;
    c.tile = xcd_tile(blockIdx.x, nblocks);
    int l = find_lat(blk, n_lat, c.tile);
    c.L = lats + l;
    c.local = c.tile - blk[l];
    return c;
}

// --- block reduction of doubles (wave64 shuffles, then LDS across 4 waves) ---
// Makes this wavefront's earlier LDS writes visible to its own later LDS
// reads from other lanes: every LDS operation of the wavefront has
// completed (lgkmcnt 0) and the compiler moves no memory access across.
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int k = 0; k < kBlock / 64; ++k) r += lds[k];
    }
    __syncthreads();
    return r;   // valid on thread 0
}

// Occupancy floor of the heavy kernels: with no floor the scheduler hoists
// every neighbour load and the double kernels fall to one wave per SIMD.  The
// clover is the exception: with a floor of 2 the FP32 kernel takes 146 VGPRs
// and runs 3 waves per SIMD (a floor of 4 spills to scratch and is 15%
// slower); FP64 runs fastest at 1 (256 VGPRs, no scratch; at 2 it spills).
#ifndef QCB_HB_WAVES_F32
#define QCB_HB_WAVES_F32 3
#endif
#ifndef QCB_FLOW_WAVES_F32
#define QCB_FLOW_WAVES_F32 4
#endif
#ifndef QCB_HB_WAVES_F64
#define QCB_HB_WAVES_F64 2
#endif
#ifndef QCB_FLOW_WAVES_F64
#define QCB_FLOW_WAVES_F64 2
#endif
template <class T> struct MinWaves;
template <> struct MinWaves<float> { static constexpr int hb = QCB_HB_WAVES_F32, flow = QCB_FLOW_WAVES_F32, clover = 2; };
template <> struct MinWaves<double> { static constexpr int hb = QCB_HB_WAVES_F64, flow = QCB_FLOW_WAVES_F64, clover = 1; };

// ---------------------------------------------------------------- fills ----

template <class T>
__global__ __launch_bounds__(kBlock) void k_cold(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                 int n_lat, uint32_t nblocks) {
    BlockCtx<T> bc = block_ctx(lats, blk, n_lat, nblocks);
    const LatDev<T>& L = *bc.L;
    const uint32_t i = bc.local * kLinkSites + (threadIdx.x & (kLinkSites - 1));
    const uint32_t hs = threadIdx.x / kLinkSites;
    if (i >= L.Vh) return;
    M3<T> m;
    m3_identity(m);
    st_link(L.U, L.S, (int)(hs >> 1), (int)(hs & 1), i, m);
}

template <class T>
__device__ __forceinline__ bool gate_open(const Gate& g, const LatDev<T>& L) {
    return g.mod == 0u || (g.sweep + L.sweep_off) % g.mod == g.res;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_hot(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                int n_lat, uint32_t nblocks, Gate gate) {
    BlockCtx<T> bc = block_ctx(lats, blk, n_lat, nblocks);
    const LatDev<T>& L = *bc.L;
    if (!gate_open(gate, L)) return;
    const uint32_t i = bc.local * kLinkSites + (threadIdx.x & (kLinkSites - 1));
    const uint32_t hs = threadIdx.x / kLinkSites;
    if (i >= L.Vh) return;
    int mu = (int)(hs >> 1), p = (int)(hs & 1);
    Coords c = cb_coords(L, i, p);
    uint32_t lex = (uint32_t)c.x + (uint32_t)L.L[0] * ((uint32_t)c.y + (uint32_t)L.L[1] * ((uint32_t)c.z + (uint32_t)L.L[2] * (uint32_t)c.t));
    const M3<double> q = haar_su3_philox(L.key0, L.key1, lex, (uint32_t)mu, 1u);
    M3<T> out;
#pragma unroll
    for (int k = 0; k < 18; ++k) out.v[k] = (T)q.v[k];
    st_link(L.U, L.S, mu, p, i, out);
}

// ------------------------------------------------------- AoS <-> SoA -------

template <class T>
__global__ __launch_bounds__(kBlock) void k_aos_to_soa(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                       int n_lat, uint32_t nblocks, const double* __restrict__ aos,
                                                       T* __restrict__ dst) {
    BlockCtx<T> bc = block_ctx(lats, blk, n_lat, nblocks);
    const LatDev<T>& L = *bc.L;
    const uint32_t i = bc.local * kLinkSites + (threadIdx.x & (kLinkSites - 1));
    const uint32_t hs = threadIdx.x / kLinkSites;
    if (i >= L.Vh) return;
    int mu = (int)(hs >> 1), p = (int)(hs & 1);
    Coords c = cb_coords(L, i, p);
    size_t lex = (size_t)c.x + (size_t)L.L[0] * ((size_t)c.y + (size_t)L.L[1] * ((size_t)c.z + (size_t)L.L[2] * (size_t)c.t));
    const double* src = aos + (lex * 4 + (size_t)mu) * 18;
    M3<T> m;
#pragma unroll
    for (int k = 0; k < 18; ++k) m.v[k] = (T)src[k];
    st_link(dst, L.S, mu, p, i, m);
}

template <class T, int LAY>
__global__ __launch_bounds__(kBlock) void k_soa_to_aos(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                       int n_lat, uint32_t nblocks, const T* __restrict__ src,
                                                       double* __restrict__ aos) {
    BlockCtx<T> bc = block_ctx(lats, blk, n_lat, nblocks);
    const LatDev<T>& L = *bc.L;
    const uint32_t i = bc.local * kLinkSites + (threadIdx.x & (kLinkSites - 1));
    const uint32_t hs = threadIdx.x / kLinkSites;
    if (i >= L.Vh) return;
    int mu = (int)(hs >> 1), p = (int)(hs & 1);
    Coords c = cb_coords(L, i, p);
    size_t lex = (size_t)c.x + (size_t)L.L[0] * ((size_t)c.y + (size_t)L.L[1] * ((size_t)c.z + (size_t)L.L[2] * (size_t)c.t));
    double* d = aos + (lex * 4 + (size_t)mu) * 18;
    M3<T> m = ld_link<T, LAY>(src, L.S, mu, p, i);
#pragma unroll
    for (int k = 0; k < 18; ++k) d[k] = (double)m.v[k];
}

// ILDG staging: links [lo, lo + n) in file order (lex site x fastest, then
// mu; lex = 2 cb + x-parity bit, so cb = lex / 2) as big-endian doubles, one
// thread per link (reference src/io/ildg.cpp:87-106).  The 18 doubles of a
// link are 144 contiguous bytes; byte order is swapped on the device so the
// host only streams the buffer to the file.
template <class T>
__global__ __launch_bounds__(kBlock) void k_soa_to_be(const LatDev<T>* __restrict__ lats, const T* __restrict__ src,
                                                      uint64_t* __restrict__ out, uint64_t lo, uint32_t n) {
    const LatDev<T>& L = lats[0];
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint64_t link = lo + t;
    const uint32_t lex = (uint32_t)(link >> 2);
    const int mu = (int)(link & 3u);
    const Coords c = cb_coords(L, lex >> 1, 0);
    const int x = (int)(lex % (uint32_t)L.L[0]);
    const int p = (x + c.y + c.z + c.t) & 1;
    M3<T> m = ld_link(src, L.S, mu, p, lex >> 1);
    uint64_t* d = out + (size_t)t * 18;
#pragma unroll
    for (int k = 0; k < 18; ++k) d[k] = __builtin_bswap64(__double_as_longlong((double)m.v[k]));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_be_to_soa(const LatDev<T>* __restrict__ lats, const uint64_t* __restrict__ in,
                                                      T* __restrict__ dst, uint64_t lo, uint32_t n) {
    const LatDev<T>& L = lats[0];
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint64_t link = lo + t;
    const uint32_t lex = (uint32_t)(link >> 2);
    const int mu = (int)(link & 3u);
    const Coords c = cb_coords(L, lex >> 1, 0);
    const int x = (int)(lex % (uint32_t)L.L[0]);
    const int p = (x + c.y + c.z + c.t) & 1;
    M3<T> m;
    const uint64_t* s = in + (size_t)t * 18;
#pragma unroll
    for (int k = 0; k < 18; ++k) m.v[k] = (T)__longlong_as_double(__builtin_bswap64(s[k]));
    st_link(dst, L.S, mu, p, lex >> 1, m);
}

template <class T, int FROM, int TO>
__global__ __launch_bounds__(kBlock) void k_copy_field(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                       int n_lat, uint32_t nblocks, const T* __restrict__ src,
                                                       T* __restrict__ dst) {
    BlockCtx<T> bc = block_ctx(lats, blk, n_lat, nblocks);
    const LatDev<T>& L = *bc.L;
    const uint32_t i = bc.local * kLinkSites + (threadIdx.x & (kLinkSites - 1));
    const uint32_t hs = threadIdx.x / kLinkSites;
    if (i >= L.Vh) return;
    st_link<T, TO>(dst, L.S, (int)(hs >> 1), (int)(hs & 1), i, ld_link<T, FROM>(src, L.S, (int)(hs >> 1), (int)(hs & 1), i));
}

// Flow start for every lattice of a batch: F[0] <- U (GradientFlow::copy).
template <class T>
__global__ __launch_bounds__(kBlock) void k_flow_init(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                      int n_lat, uint32_t nblocks, Gate gate) {
    BlockCtx<T> bc = block_ctx(lats, blk, n_lat, nblocks);
    const LatDev<T>& L = *bc.L;
    if (!gate_open(gate, L)) return;
    const uint32_t i = bc.local * kLinkSites + (threadIdx.x & (kLinkSites - 1));
    const uint32_t hs = threadIdx.x / kLinkSites;
    if (i >= L.Vh) return;
    st_link<T, kFlowOrder>(L.F[0], L.S, (int)(hs >> 1), (int)(hs & 1), i, ld_link(L.U, L.S, (int)(hs >> 1), (int)(hs & 1), i));
}

// U'(x) = U(x - s): each thread writes one destination link.
template <class T>
__global__ __launch_bounds__(kBlock) void k_translate(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                      int n_lat, uint32_t nblocks, int s0, int s1, int s2, int s3,
                                                      T* __restrict__ dst) {
    BlockCtx<T> bc = block_ctx(lats, blk, n_lat, nblocks);
    const LatDev<T>& L = *bc.L;
    const uint32_t i = bc.local * kLinkSites + (threadIdx.x & (kLinkSites - 1));
    const uint32_t hs = threadIdx.x / kLinkSites;
    if (i >= L.Vh) return;
    int mu = (int)(hs >> 1), p = (int)(hs & 1);
    Coords c = cb_coords(L, i, p);
    int x = c.x - s0; x += (x < 0) ? L.L[0] : 0;
    int y = c.y - s1; y += (y < 0) ? L.L[1] : 0;
    int z = c.z - s2; z += (z < 0) ? L.L[2] : 0;
    int t = c.t - s3; t += (t < 0) ? L.L[3] : 0;
    int ps = (x + y + z + t) & 1;
    st_link(dst, L.S, mu, p, i, ld_link(L.U, L.S, mu, ps, cb_index(L, x, y, z, t)));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_reunit(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                   int n_lat, uint32_t nblocks, Gate gate) {
    BlockCtx<T> bc = block_ctx(lats, blk, n_lat, nblocks);
    const LatDev<T>& L = *bc.L;
    if (!gate_open(gate, L)) return;
    const uint32_t i = bc.local * kLinkSites + (threadIdx.x & (kLinkSites - 1));
    const uint32_t hs = threadIdx.x / kLinkSites;
    if (i >= L.Vh) return;
    int mu = (int)(hs >> 1), p = (int)(hs & 1);
    M3<T> m = ld_link(L.U, L.S, mu, p, i);
    M3<double> d;
#pragma unroll
    for (int k = 0; k < 18; ++k) d.v[k] = (double)m.v[k];
    reunitarize_d(d);
#pragma unroll
    for (int k = 0; k < 18; ++k) m.v[k] = (T)d.v[k];
    st_link(L.U, L.S, mu, p, i, m);
}

// ------------------------------------------------------------ plaquette ----

// Per-site sum over mu<nu of Re Tr(U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag)/3
// (reference src/observables/observables_mpi.cpp:8-30), FP64 accumulation,
// of the links (which = 0) or of the lattice's current flow field (which = 1).
template <class T, int LAY>
__global__ __launch_bounds__(kBlock) void k_plaquette(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                      int n_lat, uint32_t nblocks, uint32_t gstep,
                                                      double* __restrict__ partial) {
    __shared__ double lds[kBlock / 64];
    BlockCtx<T> bc = block_ctx(lats, blk, n_lat, nblocks);
    const LatDev<T>& L = *bc.L;
    const uint32_t qs = bc.local * kSiteSites + (threadIdx.x & (kSiteSites - 1));
    const int p = (int)(threadIdx.x / kSiteSites);
    double sum = 0.0;
    if (qs < L.Vh) {
        Coords c = tiled_coords(L, qs, p);
        const uint32_t i = cb_index(L, c.x, c.y, c.z, c.t);
        const T* f = LAY == kLinkOrder ? L.U : L.F[(gstep - L.flow_start) & 1u];
#pragma unroll
        for (int mu = 0; mu < 3; ++mu) {
            M3<T> umu = ld_link<T, LAY>(f, L.S, mu, p, i);
            const uint32_t xmu = nb_cb(L, c, mu, +1);
#pragma unroll
            for (int nu = mu + 1; nu < 4; ++nu) {
                const uint32_t xnu = nb_cb(L, c, nu, +1);
                M3<T> a = mul_nn(umu, ld_link<T, LAY>(f, L.S, nu, p ^ 1, xmu));
                M3<T> b = mul_nn(ld_link<T, LAY>(f, L.S, nu, p, i), ld_link<T, LAY>(f, L.S, mu, p ^ 1, xnu));
                sum += (double)retrace_nd(a, b) / 3.0;
            }
        }
    }
    double r = block_sum(sum, lds);
    if (threadIdx.x == 0) partial[bc.tile] = r;
}

// --------------------------------------------------------------- clover ----

// F_{mu nu}(x) = (C - C^dag)/8 of the four-leaf clover, not made traceless
// (reference src/observables/observables_mpi.cpp:61-85).
// a b (or a b^dag) with the SU(3) factor a read from memory: an FP32 link
// gives its two stored rows, rows 0 and 1 of the product follow and row 2 is
// rebuilt; FP64 forms the full product
template <class T, bool DAG, int LAY>
__device__ __forceinline__ M3<T> mul_ld(const T* __restrict__ f, uint32_t S, int mu, int p, uint32_t i,
                                        const M3<T>& b) {
    if constexpr (LinkSlots<T>::n == 6) {
        const M3<T> a = ld_link_rows01<LAY>(f, S, mu, p, i);
        M3<T> t = DAG ? mul_nd<T, 2>(a, b) : mul_nn<T, 2>(a, b);
        su3_third_row(t);
        return t;
    } else {
        const M3<T> a = ld_link<T, LAY>(f, S, mu, p, i);
        return DAG ? mul_nd(a, b) : mul_nn(a, b);
    }
}

// The leaves of plane (MU, NU) at x one at a time (reference leaf order
// src/observables/observables_mpi.cpp:73-81), for the clover's shared planes
// (k_clover); clover_F below forms all four in one piece.
template <class T, int LAY>
struct CloverLeaves {
    const LatDev<T>& L;
    const T* __restrict__ f;
    Coords c;
    int p, MU, NU;
    uint32_t i;
    // Q1 = U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
    __device__ __forceinline__ M3<T> q1() const {
        const int q = p ^ 1;
        M3<T> a = mul_ld<T, false, LAY>(f, L.S, MU, p, i, ld_link<T, LAY>(f, L.S, NU, q, nb_cb(L, c, MU, +1)));
        M3<T> b = mul_ld<T, false, LAY>(f, L.S, NU, p, i, ld_link<T, LAY>(f, L.S, MU, q, nb_cb(L, c, NU, +1)));
        return mul_nd(a, b);
    }
    // C += Q2 = U_nu(x) U_mu(x-mu+nu)^dag U_nu(x-mu)^dag U_mu(x-mu)
    __device__ __forceinline__ void q2(M3<T>& C) const {
        const int q = p ^ 1;
        const uint32_t xmmu = nb_cb(L, c, MU, -1);
        M3<T> a = mul_ld<T, true, LAY>(f, L.S, NU, p, i, ld_link<T, LAY>(f, L.S, MU, p, nb_cb(L, c, MU, -1, NU, +1)));
        M3<T> b = mul_dn(ld_link<T, LAY>(f, L.S, MU, q, xmmu), ld_link<T, LAY>(f, L.S, NU, q, xmmu));
        acc_nd(C, a, b);
    }
    // C += Q3 = U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu)
    __device__ __forceinline__ void q3(M3<T>& C) const {
        const int q = p ^ 1;
        const uint32_t xmmumnu = nb_cb(L, c, MU, -1, NU, -1);
        M3<T> a = mul_ld<T, false, LAY>(f, L.S, NU, p, xmmumnu, ld_link<T, LAY>(f, L.S, MU, q, nb_cb(L, c, MU, -1)));
        M3<T> b = mul_ld<T, false, LAY>(f, L.S, MU, p, xmmumnu, ld_link<T, LAY>(f, L.S, NU, q, nb_cb(L, c, NU, -1)));
        acc_dn(C, a, b);
    }
    // C += Q4 = U_nu(x-nu)^dag U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^dag
    __device__ __forceinline__ void q4(M3<T>& C) const {
        const int q = p ^ 1;
        const uint32_t xmnu = nb_cb(L, c, NU, -1);
        M3<T> a = mul_ld<T, false, LAY>(f, L.S, MU, q, xmnu, ld_link<T, LAY>(f, L.S, NU, p, nb_cb(L, c, MU, +1, NU, -1)));
        M3<T> b = mul_dn(ld_link<T, LAY>(f, L.S, NU, q, xmnu), a);
        if constexpr (LinkSlots<T>::n == 6) {
            // columns 0, 1 of b U_mu(x)^dag from the stored rows of U_mu(x)
            M3<T> u = mul_part<T, false, true, 3, 2>(b, ld_link_rows01<LAY>(f, L.S, MU, p, i));
            su3_third_col(u);
            m3_add(C, u);
        } else {
            acc_nd(C, b, ld_link<T, LAY>(f, L.S, MU, p, i));
        }
    }
};

// C - C^dag of a sum of leaves C, by its 9 reals (anti-Hermitian; not made
// traceless, as the reference's F is not)
template <class T>
__device__ __forceinline__ A9<T> ah_part(const M3<T>& C) {
    A9<T> h;
    constexpr int pi[3] = {0, 0, 1}, pj[3] = {1, 2, 2};
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        h.o[2 * e] = C.re(pi[e], pj[e]) - C.re(pj[e], pi[e]);
        h.o[2 * e + 1] = C.im(pi[e], pj[e]) + C.im(pj[e], pi[e]);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) h.d[k] = T(2) * C.im(k, k);
    return h;
}

// V^dag H V for an anti-Hermitian H given by its 9 reals: the result is
// anti-Hermitian, so only its upper triangle and imaginary diagonal are formed
template <class T>
__device__ __forceinline__ A9<T> ah_conj(const M3<T>& V, const A9<T>& h) {
    const M3<T> B = mul_nn(a9_full(h), V);
    A9<T> r;
    constexpr int pi[3] = {0, 0, 1}, pj[3] = {1, 2, 2};
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const int a = pi[e], b = pj[e];
        T re = T(0), im = T(0);
#pragma unroll
        for (int k = 0; k < 3; ++k) {   // sum_k conj(V_ka) B_kb
            re = fma(V.re(k, a), B.re(k, b), re);
            re = fma(V.im(k, a), B.im(k, b), re);
            im = fma(V.re(k, a), B.im(k, b), im);
            im = fma(-V.im(k, a), B.re(k, b), im);
        }
        r.o[2 * e] = re;
        r.o[2 * e + 1] = im;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        T im = T(0);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            im = fma(V.re(k, a), B.im(k, a), im);
            im = fma(-V.im(k, a), B.re(k, a), im);
        }
        r.d[a] = im;
    }
    return r;
}

// F_mu,nu(x) = (C - C^dag) / 8 with C the sum of the four leaves
// (reference src/observables/observables_mpi.cpp:61-85).
template <class T, int LAY>
__device__ __forceinline__ M3<T> clover_F(const LatDev<T>& L, const T* __restrict__ f, Coords c, int p, uint32_t i,
                                          const int MU, const int NU) {
    const uint32_t S = L.S;
    const int q = p ^ 1;
    const uint32_t xpmu = nb_cb(L, c, MU, +1), xpnu = nb_cb(L, c, NU, +1);
    const uint32_t xmmu = nb_cb(L, c, MU, -1), xmnu = nb_cb(L, c, NU, -1);
    const uint32_t xmmupnu = nb_cb(L, c, MU, -1, NU, +1), xmmumnu = nb_cb(L, c, MU, -1, NU, -1);
    const uint32_t xpmumnu = nb_cb(L, c, MU, +1, NU, -1);
    M3<T> C;
    {   // U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
        M3<T> a = mul_ld<T, false, LAY>(f, S, MU, p, i, ld_link<T, LAY>(f, S, NU, q, xpmu));
        M3<T> b = mul_ld<T, false, LAY>(f, S, NU, p, i, ld_link<T, LAY>(f, S, MU, q, xpnu));
        C = mul_nd(a, b);
    }
    __builtin_amdgcn_sched_barrier(0);   // one leaf's loads at a time: keeps the registers down
    {   // U_nu(x) U_mu(x-mu+nu)^dag U_nu(x-mu)^dag U_mu(x-mu)
        M3<T> a = mul_ld<T, true, LAY>(f, S, NU, p, i, ld_link<T, LAY>(f, S, MU, p, xmmupnu));
        M3<T> b = mul_dn(ld_link<T, LAY>(f, S, MU, q, xmmu), ld_link<T, LAY>(f, S, NU, q, xmmu));
        acc_nd(C, a, b);
    }
    __builtin_amdgcn_sched_barrier(0);
    {   // U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu)
        M3<T> a = mul_ld<T, false, LAY>(f, S, NU, p, xmmumnu, ld_link<T, LAY>(f, S, MU, q, xmmu));
        M3<T> b = mul_ld<T, false, LAY>(f, S, MU, p, xmmumnu, ld_link<T, LAY>(f, S, NU, q, xmnu));
        acc_dn(C, a, b);
    }
    __builtin_amdgcn_sched_barrier(0);
    {   // U_nu(x-nu)^dag U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^dag
        M3<T> a = mul_ld<T, false, LAY>(f, S, MU, q, xmnu, ld_link<T, LAY>(f, S, NU, p, xpmumnu));
        M3<T> b = mul_dn(ld_link<T, LAY>(f, S, NU, q, xmnu), a);
        if constexpr (LinkSlots<T>::n == 6) {
            // columns 0, 1 of b U_mu(x)^dag from the stored rows of U_mu(x)
            M3<T> u = mul_part<T, false, true, 3, 2>(b, ld_link_rows01<LAY>(f, S, MU, p, i));
            su3_third_col(u);
            m3_add(C, u);
        } else {
            acc_nd(C, b, ld_link<T, LAY>(f, S, MU, p, i));
        }
        // C now holds all four leaves
    }
    M3<T> F;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            F.re(a, b) = T(0.125) * (C.re(a, b) - C.re(b, a));
            F.im(a, b) = T(0.125) * (C.im(a, b) + C.im(b, a));
        }
    return F;
}

template <class T>
__device__ __forceinline__ double sqnorm(const M3<T>& a) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 18; ++k) s += (double)a.v[k] * (double)a.v[k];
    return s;
}

// q_x = Re[Tr F01F23 - Tr F02F13 + Tr F03F12]/(4 pi^2), e_x = 1/2 sum ||F||^2
// (reference src/observables/observables_mpi.cpp:88-107).
//
// In a plane (mu, nu), mu < nu, two of x's leaves are the leaves of its
// neighbour y = x - mu conjugated by V = U_mu(y): Q2(x) = V^dag Q1(y) V and
// Q3(x) = V^dag Q4(y) V, so with H = (Q1 + Q4) - (Q1 + Q4)^dag, the
// anti-Hermitian part that F needs,
//   8 F_mu,nu(x) = H(x) + V^dag H(x - mu) V.
// A wavefront holds up to 32 cb sites of both parities (lanes 0-31 parity
// 0, 32-63 parity 1, the same cb sites), and a block four wavefronts' sites.
// They are whole x rows when Lx/2 divides 32, and also when Lx/2 = 3, 5, 6,
// 10, 15, whose wavefronts hold 30 sites (flow_block_sites; the work order
// runs x fastest over whole rows), and consecutive y rows of a tile follow
// each other.
// Every site then forms Q1 + Q4 of the planes (0, nu) and (1, 2), (1, 3) and
// publishes their H in LDS; after a barrier a site takes F of the planes
// (0, nu) from its x-neighbour's H (always in the block), and F of (1, 2),
// (1, 3) from its y-neighbour's H when every site of its wavefront has that
// neighbour in the block (checked against the cb indices the block
// publishes), else from its own leaves 2 and 3.  Plane (2, 3) forms its four
// leaves.  Of the 24 leaves 14 are formed (18 in the wavefronts without
// their y-neighbours), for five (three) conjugations.
template <class T, int LAY>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(MinWaves<T>::clover))) void k_clover(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                   int n_lat, uint32_t nblocks, uint32_t gstep,
                                                   double* __restrict__ pq, double* __restrict__ pe) {
    constexpr int kShared = 5;   // planes with H in LDS: (0,1), (0,2), (0,3), (1,2), (1,3)
    __shared__ double lds[kBlock / 64];
    __shared__ T hx[2 * kShared * 9 * kSiteSites];   // [(parity * 5 + plane) * 9 + k][site]
    __shared__ uint32_t cbs[kSiteSites];             // cb index of each slot (both parities share it)
    BlockCtx<T> bc = block_ctx(lats, blk, n_lat, nblocks);
    const LatDev<T>& L = *bc.L;
    // each wavefront holds L.fsb cb sites (of both parities): whole x rows
    // when Lx/2 divides 32 or when whole rows fill 30 of its 32 site lanes
    // (the other lanes idle), else 32 sites of partial rows; LDS slot s
    // keeps 32 slots per wavefront
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t s = (threadIdx.x >> 6) * 32u + (lane & 31u);
    const uint32_t qs = bc.local * 4u * L.fsb + (threadIdx.x >> 6) * L.fsb + (lane & 31u);
    const int p = (int)(lane >> 5);
    // rows are whole, so a live site's x-neighbour is live
    const bool live = (lane & 31u) < L.fsb && qs < L.Vh;
    const bool xrows = (L.fsb % (uint32_t)L.Lxh) == 0u;   // uniform per block
    const Coords c = tiled_coords(L, live ? qs : 0u, p);
    const uint32_t i = cb_index(L, c.x, c.y, c.z, c.t);
    const T* f = LAY == kLinkOrder ? L.U : L.F[(gstep - L.flow_start) & 1u];
    constexpr int pmu[kShared] = {0, 0, 0, 1, 1}, pnu[kShared] = {1, 2, 3, 2, 3};
    if (xrows) {
        if (live) {
            if (p == 0) cbs[s] = i;
#pragma unroll 1
            for (int k5 = 0; k5 < kShared; ++k5) {
                const CloverLeaves<T, LAY> lv{L, f, c, p, pmu[k5], pnu[k5], i};
                M3<T> C = lv.q1();
                __builtin_amdgcn_sched_barrier(0);
                lv.q4(C);
                const A9<T> h = ah_part(C);
                T* dst = hx + (uint32_t)((p * kShared + k5) * 9) * kSiteSites + s;
#pragma unroll
                for (int k = 0; k < 6; ++k) dst[k * kSiteSites] = h.o[k];
#pragma unroll
                for (int k = 0; k < 3; ++k) dst[(6 + k) * kSiteSites] = h.d[k];
            }
        }
        __syncthreads();
    }
    // F of shared plane k5 from H(x) and the neighbour's H at slot sn (cb in):
    // (H(x) + V^dag H(x - mu) V) / 8, V = U_mu(x - mu)
    auto shared_F = [&](int k5, uint32_t sn, uint32_t in) {
        const T* own = hx + (uint32_t)((p * kShared + k5) * 9) * kSiteSites + s;
        const T* nbr = hx + (uint32_t)(((p ^ 1) * kShared + k5) * 9) * kSiteSites + sn;
        A9<T> hy, h;
#pragma unroll
        for (int k = 0; k < 6; ++k) { hy.o[k] = nbr[k * kSiteSites]; h.o[k] = own[k * kSiteSites]; }
#pragma unroll
        for (int k = 0; k < 3; ++k) { hy.d[k] = nbr[(6 + k) * kSiteSites]; h.d[k] = own[(6 + k) * kSiteSites]; }
        const A9<T> g = ah_conj(ld_link<T, LAY>(f, L.S, pmu[k5], p ^ 1, in), hy);
        return a9_full(a9_axpby(T(0.125), h, T(0.125), g));
    };
    double q = 0.0, e = 0.0;
    if (live) {
        const uint32_t step0 = cb_step(L, c, 0, -1);   // cb index and block slot of x - e0
        const uint32_t step1 = cb_step(L, c, 1, -1);   // cb index step to x - e1
        // x - e1 in the block for every site of the wavefront: the previous
        // row, Lx/2 slots back, or for a wavefront's first row the last row
        // of the wavefront before, past its 32 - L.fsb idle slots
        const uint32_t yc = (lane & 31u) < (uint32_t)L.Lxh ? 32u - L.fsb : 0u;
        const uint32_t sy = s - (uint32_t)L.Lxh - yc;
        const bool yin = s >= (uint32_t)L.Lxh + yc;
        const bool yok = xrows && yin && cbs[yin ? sy : s] == i + step1;
        const bool yshare = __all(yok);
        double tq = 0.0;
        // the three plane pairs (01,23), (02,13), (03,12) one at a time: a
        // loop the compiler keeps rolled bounds the registers to one pair
#pragma unroll 1
        for (int pr = 0; pr < 3; ++pr) {
            const int nu = pr + 1;                            // (0, nu) and its complement
            const int m2 = pr == 2 ? 1 : (pr == 1 ? 1 : 2), n2 = pr == 0 ? 3 : (pr == 1 ? 3 : 2);
            const M3<T> a = xrows ? shared_F(nu - 1, s + step0, i + step0) : clover_F<T, LAY>(L, f, c, p, i, 0, nu);
            M3<T> b;
            if (xrows && pr > 0) {   // (1, 3) or (1, 2): H(x) in LDS
                const int k5 = pr == 1 ? 4 : 3;
                if (yshare) {
                    b = shared_F(k5, sy, i + step1);
                } else {
                    const CloverLeaves<T, LAY> lv{L, f, c, p, m2, n2, i};
                    M3<T> C;
                    m3_zero(C);
                    lv.q2(C);
                    __builtin_amdgcn_sched_barrier(0);
                    lv.q3(C);
                    const A9<T> h23 = ah_part(C);
                    const T* own = hx + (uint32_t)((p * kShared + k5) * 9) * kSiteSites + s;
                    A9<T> h;
#pragma unroll
                    for (int k = 0; k < 6; ++k) h.o[k] = own[k * kSiteSites];
#pragma unroll
                    for (int k = 0; k < 3; ++k) h.d[k] = own[(6 + k) * kSiteSites];
                    b = a9_full(a9_axpby(T(0.125), h, T(0.125), h23));
                }
            } else {
                b = clover_F<T, LAY>(L, f, c, p, i, m2, n2);
            }
            const double tr = (double)retrace_nn(a, b);
            tq += pr == 1 ? -tr : tr;
            e += sqnorm(a) + sqnorm(b);
        }
        q = tq * (1.0 / (4.0 * M_PI * M_PI));
        e *= 0.5;
    }
    double rq = block_sum(q, lds);
    double re = block_sum(e, lds);
    if (threadIdx.x == 0) {
        pq[bc.tile] = rq;
        pe[bc.tile] = re;
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_finalize_plaq(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                          const double* __restrict__ partial, double* __restrict__ out,
                                                          int ostride) {
    __shared__ double lds[kBlock / 64];
    const int l = blockIdx.x;
    double s = 0.0;
    for (uint32_t b = blk[l] + threadIdx.x; b < blk[l + 1]; b += kBlock) s += partial[b];
    double r = block_sum(s, lds);
    if (threadIdx.x == 0) out[(size_t)l * ostride] = r / (6.0 * 2.0 * (double)lats[l].Vh);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_finalize_qe(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                        const double* __restrict__ pq, const double* __restrict__ pe,
                                                        double* __restrict__ out, int ostride, uint32_t gstep) {
    __shared__ double lds[kBlock / 64];
    const int l = blockIdx.x;
    double s = 0.0, t = 0.0;
    for (uint32_t b = blk[l] + threadIdx.x; b < blk[l + 1]; b += kBlock) {
        s += pq[b];
        t += pe[b];
    }
    double rq = block_sum(s, lds);
    double re = block_sum(t, lds);
    if (threadIdx.x == 0) {
        double* o = lats[l].qe_out ? lats[l].qe_out + 2 * (size_t)(gstep - lats[l].flow_start) : out + (size_t)l * ostride;
        o[0] = rq;
        o[1] = re / (2.0 * (double)lats[l].Vh);
    }
}

// -------------------------------------------------------------- staples ----

// Staple sums are not SU(3), so they go out as 18 doubles per link in the
// host's AoS order (lex site, mu), not through the link storage.
template <class T, int MU>
__device__ __forceinline__ void staple_one(const LatDev<T>& L, double* aos, Coords c, int p, uint32_t i) {
    const M3<T> a = staple<T, MU>(L, L.U, c, p, i);
    const size_t lex = (size_t)c.x + (size_t)L.L[0] * ((size_t)c.y + (size_t)L.L[1] * ((size_t)c.z + (size_t)L.L[2] * (size_t)c.t));
    double* d = aos + (lex * 4 + (size_t)MU) * 18;
#pragma unroll
    for (int k = 0; k < 18; ++k) d[k] = (double)a.v[k];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_staples(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                    int n_lat, uint32_t nblocks, double* __restrict__ aos) {
    BlockCtx<T> bc = block_ctx(lats, blk, n_lat, nblocks);
    const LatDev<T>& L = *bc.L;
    const uint32_t q = bc.local * kLinkSites + (threadIdx.x & (kLinkSites - 1));
    const uint32_t hs = threadIdx.x / kLinkSites;
    if (q >= L.Vh) return;
    int mu = (int)(hs >> 1), p = (int)(hs & 1);
    Coords c = tiled_coords(L, q, p);
    const uint32_t i = cb_index(L, c.x, c.y, c.z, c.t);
    switch (mu) {
        case 0: staple_one<T, 0>(L, aos, c, p, i); break;
        case 1: staple_one<T, 1>(L, aos, c, p, i); break;
        case 2: staple_one<T, 2>(L, aos, c, p, i); break;
        default: staple_one<T, 3>(L, aos, c, p, i); break;
    }
}

// ------------------------------------------------------------- heatbath ----

// Transcendentals of the Kennedy-Pendleton step.  The double path calls the
// accurate libm routines (it is held to 1e-12 against the oracle); the float
// path uses the gfx950 hardware approximations, whose error (~1 ulp of float)
// is far below the statistical resolution of a single draw.  cos2pi(u) and
// sin2pi(u) take the angle in turns, which is what v_cos_f32 / v_sin_f32 take.
struct KpMath {
    __device__ static __forceinline__ double ln(double x) { return log(x); }
    // cos(2 pi u) as cospi(2u): the argument stays exact and the reduction is
    // a quarter-turn split, not the large-argument path of cos
    __device__ static __forceinline__ double cos2pi(double u) { return cospi(2.0 * u); }
    __device__ static __forceinline__ void sincos2pi(double u, double* s, double* c) { sincospi(2.0 * u, s, c); }
    __device__ static __forceinline__ double sq(double x) { return sqrt(x); }
    __device__ static __forceinline__ float ln(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }
    __device__ static __forceinline__ float cos2pi(float u) { return __builtin_amdgcn_cosf(u); }
    __device__ static __forceinline__ void sincos2pi(float u, float* s, float* c) {
        *s = __builtin_amdgcn_sinf(u);
        *c = __builtin_amdgcn_cosf(u);
    }
    __device__ static __forceinline__ float sq(float x) { return __builtin_amdgcn_sqrtf(x); }
    // 1/x, alpha = 2 beta k / 3 and -1/(2 alpha): correctly rounded divisions
    // in double (the oracle's arithmetic), v_rcp_f32 (1 ulp) in float
    __device__ static __forceinline__ double rcp(double x) { return 1.0 / x; }
    __device__ static __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    __device__ static __forceinline__ double alpha(double beta, double k) { return beta * k * 2.0 / 3.0; }
    __device__ static __forceinline__ float alpha(float beta, float k) { return beta * k * (2.0f / 3.0f); }
    __device__ static __forceinline__ double neg_half_over(double a) { return -0.5 / a; }
    __device__ static __forceinline__ float neg_half_over(float a) { return -0.5f * __builtin_amdgcn_rcpf(a); }
    // heatbath uniform of a Philox word: (w + 1/2) 2^-32 in the link precision
    __device__ static __forceinline__ void uni(uint32_t w, double* u) { *u = u32d(w); }
    __device__ static __forceinline__ void uni(uint32_t w, float* u) { *u = u32f(w); }
    template <class T>
    __device__ static __forceinline__ T uni(uint32_t w) {
        T u;
        uni(w, &u);
        return u;
    }
};

// One Kennedy-Pendleton attempt (reference src/heatbath/heatbath.cpp:14-26);
// nhia = -1/(2 alpha).
template <class T>
__device__ __forceinline__ bool kp_try(T nhia, T delta, T r1, T r2, T r3, T& a0) {
    T cs = KpMath::cos2pi(r2);
    T lam2 = nhia * (KpMath::ln(r1) + cs * cs * KpMath::ln(r3));
    if (delta * delta <= T(1) - lam2) {
        a0 = T(1) - T(2) * lam2;
        return true;
    }
    return false;
}

// Sphere draw of the remaining three components (reference src/heatbath/heatbath.cpp:29-35).
template <class T>
__device__ __forceinline__ void kp_sphere(T a0, T uphi, T ucos, T x[4]) {
    T rs = KpMath::sq(T(1) - a0 * a0);
    T ct = T(2) * ucos - T(1);
    T st = KpMath::sq(T(1) - ct * ct);
    T sp, cp;
    KpMath::sincos2pi(uphi, &sp, &cp);
    x[0] = a0;
    x[1] = rs * st * cp;
    x[2] = rs * st * sp;
    x[3] = rs * ct;
}

// Cabibbo-Marinari hit on U with staple A, subgroups (0,1),(1,2),(0,2)
// (reference src/heatbath/heatbath_mpi.cpp:11-34).  Only the four entries of
// W = U A in the subgroup block are formed, and only rows i, j of U change.
// Sphere uniforms (Philox mode): subgroup m = 3h + g uses words 2(m&1) and
// 2(m&1)+1 of the call 0x80000000 at the counter of subgroup m - (m&1), so
// two subgroups share one call; `stash` carries the odd member's two words
// from one subgroup (or hit) to the next.  A draw gets cap attempts in
// Philox mode and K in injected mode; one that runs out leaves the link's
// subgroup unchanged.  Returns the number of such exhausted draws.
template <class T, int MODE>
__device__ __forceinline__ int cm_hit(M3<T>& U, const M3<T>& A, T beta, int hs, int h, uint32_t cb, uint32_t sweep,
                                      uint32_t key0, uint32_t key1, const double* __restrict__ inj, int n_hits, int K,
                                      uint32_t Vh, uint32_t stash[2], uint32_t cap) {
    int exhausted = 0;
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        const int si = g == 1 ? 1 : 0;
        const int sj = g == 0 ? 1 : 2;
        // W_rc = sum_k U_rk A_kc for r, c in {si, sj}
        T w[8];
#pragma unroll
        for (int rr = 0; rr < 2; ++rr)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                const int r = rr ? sj : si, col = cc ? sj : si;
                T re = T(0), im = T(0);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    re = fma(U.re(r, k), A.re(k, col), re);
                    re = fma(-U.im(r, k), A.im(k, col), re);
                    im = fma(U.re(r, k), A.im(k, col), im);
                    im = fma(U.im(r, k), A.re(k, col), im);
                }
                w[(rr * 2 + cc) * 2] = re;
                w[(rr * 2 + cc) * 2 + 1] = im;
            }
        // su2_to_quaternion (reference src/su3/utils.cpp:47-50)
        T q0 = (w[0] + w[6]) * T(0.5), q1 = (w[3] + w[5]) * T(0.5);
        T q2 = (w[2] - w[4]) * T(0.5), q3 = (w[1] - w[7]) * T(0.5);
        T k = KpMath::sq(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
        T ik = KpMath::rcp(k);
        T v0 = q0 * ik, v1 = q1 * ik, v2 = q2 * ik, v3 = q3 * ik;
        const T nhia = KpMath::neg_half_over(KpMath::alpha(beta, k));
        // Rejection loop of KP attempts only; the sphere draw follows once
        // per subgroup, outside the loop, so a wavefront whose lanes accept
        // at different attempts runs it once instead of once per attempt.
        T a0 = T(0);
        bool ok = false;
        const double* u = nullptr;
        const uint32_t cz = hb_ctr_z(hs, h, g);
        uint32_t sw0 = 0, sw1 = 0;
        if (MODE != kHbInjected) {
            if (((3 * h + g) & 1) == 0) {
                U4 os = philox4x32_10(U4{cb, sweep, cz, 0x80000000u}, key0, key1);
                sw0 = os.x;
                sw1 = os.y;
                stash[0] = os.z;
                stash[1] = os.w;
            } else {
                sw0 = stash[0];
                sw1 = stash[1];
            }
        }
        if (MODE == kHbInjected) {
            const int nslot = 4 * K + 2;
            u = inj + ((size_t)((hs * n_hits + h) * 3 + g) * (size_t)nslot) * Vh + cb;
            for (int a = 0; a < K; ++a) {
                if (kp_try<T>(nhia, (T)u[(size_t)(4 * a) * Vh], (T)u[(size_t)(4 * a + 1) * Vh],
                              (T)u[(size_t)(4 * a + 2) * Vh], (T)u[(size_t)(4 * a + 3) * Vh], a0)) {
                    ok = true;
                    break;
                }
            }
        } else {
            for (uint32_t a = 0; a < cap; ++a) {
                U4 o = philox4x32_10(U4{cb, sweep, cz, a}, key0, key1);
                if (kp_try<T>(nhia, KpMath::uni<T>(o.x), KpMath::uni<T>(o.y), KpMath::uni<T>(o.z),
                              KpMath::uni<T>(o.w), a0)) {
                    ok = true;
                    break;
                }
            }
        }
        if (!ok) {
            ++exhausted;
            continue;
        }
        T x[4];
        if (MODE == kHbInjected) {
            kp_sphere<T>(a0, (T)u[(size_t)(4 * K) * Vh], (T)u[(size_t)(4 * K + 1) * Vh], x);
        } else {
            kp_sphere<T>(a0, KpMath::uni<T>(sw0), KpMath::uni<T>(sw1), x);
        }
        // r = mult(x, adj(v)) (reference src/heatbath/heatbath.cpp:39-44, src/su3/utils.cpp:53-63)
        const T p0 = v0, p1 = -v1, p2 = -v2, p3 = -v3;
        T r0 = x[0] * p0 - x[1] * p1 - x[2] * p2 - x[3] * p3;
        T r1 = x[0] * p1 + x[1] * p0 + x[2] * p3 - x[3] * p2;
        T r2 = x[0] * p2 - x[1] * p3 + x[2] * p0 + x[3] * p1;
        T r3 = x[0] * p3 + x[1] * p2 - x[2] * p1 + x[3] * p0;
        // U <- R U with R embedded in (si, sj) (reference src/su3/utils.cpp:33-44)
        // R_ii = r0 + i r3, R_ij = r2 + i r1, R_ji = -r2 + i r1, R_jj = r0 - i r3
#pragma unroll
        for (int col = 0; col < 3; ++col) {
            T uir = U.re(si, col), uii = U.im(si, col), ujr = U.re(sj, col), uji = U.im(sj, col);
            T ni_r = r0 * uir - r3 * uii + (r2 * ujr - r1 * uji);
            T ni_i = r0 * uii + r3 * uir + (r2 * uji + r1 * ujr);
            T nj_r = -r2 * uir - r1 * uii + (r0 * ujr + r3 * uji);
            T nj_i = -r2 * uii + r1 * uir + (r0 * uji - r3 * ujr);
            U.re(si, col) = ni_r;
            U.im(si, col) = ni_i;
            U.re(sj, col) = nj_r;
            U.im(sj, col) = nj_i;
        }
    }
    return exhausted;
}

// One (mu, parity) half-sweep: each thread owns one link U_mu(x), computes its
// staple once and applies n_hits CM hits in registers.

template <class T, int MU, int MODE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(MinWaves<T>::hb))) void k_hb_half(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                    int n_lat, uint32_t nblocks, int p, int n_hits, uint32_t sweep,
                                                    const uint32_t* __restrict__ sweep_base,
                                                    const double* __restrict__ inj, int K) {
    // sweep_base (device counter, advanced by k_bump) lets a captured graph of
    // sweeps be replayed with fresh Philox counters
    if (sweep_base) sweep += *sweep_base;
    // persistent grids stride over the tiles; gridDim.x is a multiple of 8 so
    // a block keeps to its XCD's contiguous tile range
    for (uint32_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const uint32_t tile = xcd_tile(b, nblocks);
        const int l = find_lat(blk, n_lat, tile);
        const LatDev<T>& L = lats[l];
        const uint32_t q = (tile - blk[l]) * kHalfSites + threadIdx.x;
        if (q >= L.Vh) continue;
        Coords c = tiled_coords(L, q, p);
        const uint32_t i = cb_index(L, c.x, c.y, c.z, c.t);
        M3<T> A = staple<T, MU>(L, L.U, c, p, i);
        // no other thread of the half-sweep reads a link of its own class (the
        // staples of (mu, p) use (mu, 1-p) and the other directions), so the
        // updated link is loaded non-temporally, like it is stored
        M3<T> U = ld_link<T, kLinkOrder, true>(L.U, L.S, MU, p, i);
        const T beta = (T)L.beta;
        const uint32_t sw = sweep + L.sweep_off;
        int ex = 0;
        uint32_t stash[2] = {0u, 0u};
        for (int h = 0; h < n_hits; ++h)
            ex += cm_hit<T, MODE>(U, A, beta, 2 * MU + p, h, i, sw, L.key0, L.key1, inj, n_hits, K, L.Vh, stash,
                                  L.kp_cap);
        su3_store_fix(U);
        st_link(L.U, L.S, MU, p, i, U);
        if (ex && L.kp_exh) atomicAdd(L.kp_exh, (unsigned long long)ex);
    }
}

__global__ void k_bump(uint32_t* __restrict__ counter, uint32_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter += n;
}

// --------------------------------------------------------------- flow ------

// Stages 2 and 3 read the force that stage 1 (and 2) stored, Z0 below.  FP64
// issues that load first, so that its latency hides behind the staples (the
// FP64 stage is bound by load latency: 1.5% faster); FP32 issues it right
// after the block barrier, where the phase-1 temporaries are dead and the
// upper staples still cover it (stage 2 2% faster than at the tail, 3%
// faster than first).
template <class T> constexpr bool kEarlyForce = LinkSlots<T>::n == 9;

// One fused Luescher RK3 stage for one link (reference src/flow/gradient_flow.cpp:15-128):
//  Z = P_TA(-W Sigma[W]);
//  stage 1: Zs = Z,                      W' = exp(eps/4 Z) W
//  stage 2: Zs = 8/9 Z - 17/36 Zs,       W' = exp(eps Zs) W
//  stage 3:                              W' = exp(eps (3/4 Z - Zs)) W
// Z and Zs are carried as their 9 independent reals (A9).
template <class T, int STAGE, bool ADD>
__device__ __forceinline__ void flow_tail(const LatDev<T>& L, const M3<T>& W, T* __restrict__ wout,
                                          T* __restrict__ zs, const A9<T>& Z0, int mu, int p, uint32_t i, T eps,
                                          const M3<T>& Sg, const A9<T>& Zx) {
    A9<T> Z = force_ta(W, Sg);
    if constexpr (ADD) Z = a9_axpby(T(1), Z, T(1), Zx);
    A9<T> X;
    if (STAGE == 1) {
        st_a9(zs, L.S, mu, p, i, Z);
        X = a9_scale(T(0.25) * eps, Z);
    } else if (STAGE == 2) {
        A9<T> Z1 = a9_axpby(T(8.0 / 9.0), Z, T(-17.0 / 36.0), Z0);
        st_a9(zs, L.S, mu, p, i, Z1);
        X = a9_scale(eps, Z1);
    } else {
        X = a9_axpby(T(0.75) * eps, Z, -eps, Z0);
    }
    M3<T> E = exp_su3_ah<T, ExpTerms<T>::N>(X);
    M3<T> Wn = mul_nn(E, W);
    // FP32: rows 0, 1 re-orthonormalised once per RK3 step.  The stage-1 and
    // stage-2 fields are read once, by the next stage, whose row-2 rebuild
    // then sees rows orthonormal to the rounding of one product; the error
    // does not carry over from step to step.
    if (STAGE == 3) su3_store_fix(Wn);
    st_link<T, kFlowOrder>(wout, L.S, mu, p, i, Wn);
}

// The block's 256 threads are 32 sites x 8 (mu, parity) classes, so the four
// links of a site sit in four wavefronts of one block.  In plane {a, b}
// (a < b) the product S = U_b(x+a) U_a(x+b)^dag is the start of both upper
// staples through x,
//   Sigma_a += S U_b(x)^dag,   Sigma_b += S^dag U_a(x)^dag,
// so one of the two threads forms it and hands it to the other through LDS.
// The upper staples of the four links of x also share the links U_nu(x)
// themselves; in FP32 each thread publishes its own link in LDS and the other
// three read it there.  A block's 32 sites are whole x rows whenever Lx/2
// divides 32 (the work order runs x fastest over whole rows), and then the
// lower staple of U_mu(x), mu != 0, in the plane (0, mu) comes from the block
// too: with y = x - e0, the x-neighbour in the same row,
//   U_mu(x) U_0(y+mu)^dag U_mu(y)^dag U_0(y) = S_(0,mu)(y) U_mu(y)^dag U_0(y),
// S_(0,mu)(y) = U_mu(y+e0) U_0(y+mu)^dag being plane (0, mu)'s shared product
// at y, so it enters the force as P_TA(-S U_mu(y)^dag U_0(y)) with S from
// LDS (in FP32 the links of y too: no loads; in FP64 two links from the
// field) instead of a three-link staple.  The other lower staples
// U_b(x+a-b)^dag U_a(x-b)^dag U_b(x-b) are formed from the field.
// Phase 1: own link (FP32) and owned S into LDS, then the lower staples from
// the field; barrier; phase 2: the upper staples, the (mu, 0) lower staple
// from LDS where the rows are whole, and the update.  The S owners balance
// phase 1: 2/1/2/1 over mu = 0..3 when every direction forms its three lower
// staples from the field, 0/2/2/2 when mu = 1..3 take their (mu, 0) lower
// staple from LDS and form two while mu = 0 forms three.
__device__ __forceinline__ constexpr int plane_idx(int a, int b) { return a == 0 ? b - 1 : (a == 1 ? b + 1 : 5); }
template <bool XL>
__device__ __forceinline__ constexpr int plane_owner(int pl) {
    if constexpr (XL) return pl == 0 ? 1 : pl == 1 ? 2 : pl == 2 ? 3 : pl == 3 ? 1 : pl == 4 ? 3 : 2;
    return pl == 0 ? 0 : pl == 1 ? 2 : pl == 2 ? 0 : pl == 3 ? 1 : pl == 4 ? 3 : 2;
}
// LDS slot of entry k of plane pl's S at parity p, site s of the block
__device__ __forceinline__ constexpr uint32_t plane_slot(int pl, int p, int k, uint32_t s) {
    return ((uint32_t)(pl * 2 + p) * 18u + (uint32_t)k) * kLinkSites + s;
}

// LDS of one flow-stage block: each plane's upper-staple product S per parity
// and site; in FP32 also the stored rows of every site's four links U_mu(x).
template <class T>
struct FlowLds {
    T s[6 * 2 * 18 * kLinkSites];
};
template <>
struct FlowLds<float> {
    typedef Vec2<float>::type C2;
    float s[6 * 2 * 18 * kLinkSites];
    C2 w[8 * 6 * kLinkSites];   // U_mu(x): [(hs * 6 + entry) * 32 + site], hs = 2 mu + parity
};

__device__ __forceinline__ void lds_put_rows01(Vec2<float>::type* a, int hs, uint32_t s, const M3<float>& m) {
#pragma unroll
    for (int k = 0; k < 6; ++k) a[((uint32_t)hs * 6u + (uint32_t)k) * kLinkSites + s] = {m.v[2 * k], m.v[2 * k + 1]};
}
// rows 0 and 1 of U_mu(x) from LDS (row 2 unset)
__device__ __forceinline__ M3<float> lds_get_rows01(const Vec2<float>::type* a, int hs, uint32_t s) {
    M3<float> m;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const Vec2<float>::type c = a[((uint32_t)hs * 6u + (uint32_t)k) * kLinkSites + s];
        m.v[2 * k] = c.x;
        m.v[2 * k + 1] = c.y;
    }
    return m;
}

// The owned plane's upper-staple start S = U_b(x+a) U_a(x+b)^dag into LDS.
// With whole x rows in the block the owner of a plane (0, b) is b, and
// U_b(x+e0) is the link that the other half of its own wavefront (the other
// parity, same direction) has just published in LDS, so waiting for the
// wavefront's LDS writes (wave_lds_sync) is all the read needs.
template <class T, int MU, int NU, bool XL>
__device__ __forceinline__ void flow_put_plane(const LatDev<T>& L, const T* __restrict__ win, Coords c, int p,
                                               uint32_t s, FlowLds<T>& sh) {
    constexpr int a = MU < NU ? MU : NU, b = MU < NU ? NU : MU;
    constexpr int pl = plane_idx(a, b);
    if constexpr (plane_owner<XL>(pl) == MU) {
        const int q = p ^ 1;
        const uint32_t xb = nb_cb(L, c, b, +1);
        M3<T> t;
        if constexpr (LinkSlots<T>::n == 6) {
            M3<T> u;
            if constexpr (XL && a == 0) {   // FP32 only: the stored rows of the block's links are in LDS
                static_assert(b == MU, "planes (0, b) are owned by b on whole rows");
                wave_lds_sync();
                u = lds_get_rows01(sh.w, MU * 2 + q, s + cb_step(L, c, 0, +1));   // U_b(x+e0)
            } else {
                u = ld_link_rows01<kFlowOrder>(win, L.S, b, q, nb_cb(L, c, a, +1));
            }
            t = mul_nd<T, 2>(u, ld_link<T, kFlowOrder>(win, L.S, a, q, xb));
            su3_third_row(t);
        } else {
            t = mul_nd(ld_link<T, kFlowOrder>(win, L.S, b, q, nb_cb(L, c, a, +1)), ld_link<T, kFlowOrder>(win, L.S, a, q, xb));
        }
#pragma unroll
        for (int k = 0; k < 18; ++k) sh.s[plane_slot(pl, p, k, s)] = t.v[k];
    }
}

// Sigma += the lower staple U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu), with
// U_nu(x-nu) given.
template <class T, int MU, int NU>
__device__ __forceinline__ void flow_lower(const LatDev<T>& L, const T* __restrict__ win, Coords c, int p,
                                           const M3<T>& down, M3<T>& Sg) {
    const int q = p ^ 1;
    const uint32_t xmnu = nb_cb(L, c, NU, -1);
    const uint32_t xmumnu = nb_cb(L, c, MU, +1, NU, -1);
    M3<T> t;
    if constexpr (LinkSlots<T>::n == 6) {
        M3<T> u = ld_link_rows01<kFlowOrder>(win, L.S, MU, q, xmnu);   // U_mu(x-nu)
        t = mul_nn<T, 2>(u, ld_link<T, kFlowOrder>(win, L.S, NU, p, xmumnu));
        su3_third_row(t);
    } else {
        t = mul_nn(ld_link<T, kFlowOrder>(win, L.S, MU, q, xmnu), ld_link<T, kFlowOrder>(win, L.S, NU, p, xmumnu));
    }
    acc_dn(Sg, t, down);
}

// Sigma += the upper staple through x from its plane's S in LDS and U_nu(x).
// In FP32 U_nu(x) is its two stored rows: columns 0 and 1 of S U_nu(x)^dag
// (or S^dag U_nu(x)^dag) need no more, and column 2 of the SU(3) product is
// rebuilt from them -- 72 + 24 multiply-adds and 18 adds instead of a row-2
// rebuild of U_nu(x) and 108.
template <class T, int MU, int NU>
__device__ __forceinline__ void flow_upper(const T* lds, int p, uint32_t s, const M3<T>& unu, M3<T>& Sg) {
    constexpr int pl = plane_idx(MU < NU ? MU : NU, MU < NU ? NU : MU);
    M3<T> t;
#pragma unroll
    for (int k = 0; k < 18; ++k) t.v[k] = lds[plane_slot(pl, p, k, s)];
    if constexpr (LinkSlots<T>::n == 6) {
        M3<T> u = mul_part<T, (MU > NU), true, 3, 2>(t, unu);
        su3_third_col(u);
        m3_add(Sg, u);
    } else if constexpr (MU < NU) {
        acc_nd(Sg, t, unu);
    } else {
        acc_dd(Sg, t, unu);
    }
}

template <class T, int MU, bool XL>
__device__ __forceinline__ void flow_phase1(const LatDev<T>& L, const T* __restrict__ win, Coords c, int p,
                                            uint32_t s, uint32_t i, FlowLds<T>& sh, M3<T>& Sg) {
    if constexpr (LinkSlots<T>::n == 6) lds_put_rows01(sh.w, MU * 2 + p, s, ld_link_rows01<kFlowOrder>(win, L.S, MU, p, i));
    m3_zero(Sg);
    const int q = p ^ 1;
    constexpr int N0 = MU == 0 ? 1 : 0, N1 = MU <= 1 ? 2 : 1, N2 = MU <= 2 ? 3 : 2;
    flow_put_plane<T, MU, N0, XL>(L, win, c, p, s, sh);
    if constexpr (!XL || N0 != 0)   // with XL the (mu, 0) lower staple comes from LDS in phase 2
        flow_lower<T, MU, N0>(L, win, c, p, ld_link<T, kFlowOrder>(win, L.S, N0, q, nb_cb(L, c, N0, -1)), Sg);
    flow_put_plane<T, MU, N1, XL>(L, win, c, p, s, sh);
    flow_lower<T, MU, N1>(L, win, c, p, ld_link<T, kFlowOrder>(win, L.S, N1, q, nb_cb(L, c, N1, -1)), Sg);
    flow_put_plane<T, MU, N2, XL>(L, win, c, p, s, sh);
    flow_lower<T, MU, N2>(L, win, c, p, ld_link<T, kFlowOrder>(win, L.S, N2, q, nb_cb(L, c, N2, -1)), Sg);
}

// mu != 0, whole x rows in the block: the force term of the lower staple of
// U_mu(x) in the plane (mu, 0), P_TA(-S_(0,mu)(y) U_mu(y)^dag U_0(y)) with
// y = x - e0: S from LDS (the row's slots are contiguous, so y's slot is s
// plus the cb step to y), the two links of y from LDS in FP32 (their stored
// rows) and from the field in FP64.
template <class T, int MU>
__device__ __forceinline__ A9<T> flow_xlower(const LatDev<T>& L, const T* __restrict__ win, Coords c, int p,
                                             uint32_t i, uint32_t s, const FlowLds<T>& sh) {
    static_assert(MU != 0, "the (0, 0) plane does not exist");
    const uint32_t step = cb_step(L, c, 0, -1);
    const uint32_t sy = s + step;
    const int q = p ^ 1;
    constexpr int pl = plane_idx(0, MU);
    M3<T> t;
#pragma unroll
    for (int k = 0; k < 18; ++k) t.v[k] = sh.s[plane_slot(pl, q, k, sy)];
    if constexpr (LinkSlots<T>::n == 6) {
        M3<T> r = mul_part<T, false, true, 3, 2>(t, lds_get_rows01(sh.w, MU * 2 + q, sy));   // S U_mu(y)^dag
        su3_third_col(r);
        M3<T> u0 = lds_get_rows01(sh.w, q, sy);   // U_0(y)
        su3_third_row(u0);
        return force_ta(r, u0);
    } else {
        const uint32_t iy = i + step;
        return force_ta(mul_nd(t, ld_link<T, kFlowOrder>(win, L.S, MU, q, iy)), ld_link<T, kFlowOrder>(win, L.S, 0, q, iy));
    }
}

// U_nu(x) for the update of a link of x: FP32 rows 0, 1 from LDS (the whole
// link with ROW2), else the whole link from the field.
template <class T, bool ROW2 = false>
__device__ __forceinline__ M3<T> site_link(const LatDev<T>& L, const T* __restrict__ win, const FlowLds<T>& sh,
                                           int nu, int p, uint32_t i, uint32_t s) {
    if constexpr (LinkSlots<T>::n == 6) {
        M3<T> m = lds_get_rows01(sh.w, nu * 2 + p, s);
        if constexpr (ROW2) su3_third_row(m);
        return m;
    } else {
        return ld_link<T, kFlowOrder>(win, L.S, nu, p, i);
    }
}

template <class T, int MU, int STAGE, bool XL>
__device__ __forceinline__ void flow_phase2(const LatDev<T>& L, const T* __restrict__ win, T* __restrict__ wout,
                                            T* __restrict__ zs, const A9<T>& Z0, Coords c, int p, uint32_t i,
                                            uint32_t s, const FlowLds<T>& sh, M3<T>& Sg, T eps) {
    constexpr int N0 = MU == 0 ? 1 : 0, N1 = MU <= 1 ? 2 : 1, N2 = MU <= 2 ? 3 : 2;
    flow_upper<T, MU, N0>(sh.s, p, s, site_link(L, win, sh, N0, p, i, s), Sg);
    flow_upper<T, MU, N1>(sh.s, p, s, site_link(L, win, sh, N1, p, i, s), Sg);
    flow_upper<T, MU, N2>(sh.s, p, s, site_link(L, win, sh, N2, p, i, s), Sg);
    constexpr bool ADD = XL && MU != 0;
    A9<T> Zx;
    if constexpr (ADD) Zx = flow_xlower<T, MU>(L, win, c, p, i, s, sh);
    flow_tail<T, STAGE, ADD>(L, site_link<T, true>(L, win, sh, MU, p, i, s), wout, zs, Z0, MU, p, i, eps, Sg, Zx);
}

template <class T, int STAGE, bool XL>
__device__ __forceinline__ void flow_stage_body(const LatDev<T>& L, const T* __restrict__ win, T* __restrict__ wout,
                                                T* __restrict__ zs, bool live, int mu, Coords c, int p, uint32_t i,
                                                uint32_t s, FlowLds<T>& sh, T eps) {
    A9<T> Z0;   // the stored force (stages 2, 3)
    if (kEarlyForce<T> && STAGE != 1 && live) Z0 = ld_a9((const T*)zs, L.S, mu, p, i);
    M3<T> Sg;
    if (live) {
        switch (mu) {
            case 0: flow_phase1<T, 0, XL>(L, win, c, p, s, i, sh, Sg); break;
            case 1: flow_phase1<T, 1, XL>(L, win, c, p, s, i, sh, Sg); break;
            case 2: flow_phase1<T, 2, XL>(L, win, c, p, s, i, sh, Sg); break;
            default: flow_phase1<T, 3, XL>(L, win, c, p, s, i, sh, Sg); break;
        }
    }
    __syncthreads();
    if (!kEarlyForce<T> && STAGE != 1 && live) Z0 = ld_a9((const T*)zs, L.S, mu, p, i);
    if (live) {
        switch (mu) {
            case 0: flow_phase2<T, 0, STAGE, XL>(L, win, wout, zs, Z0, c, p, i, s, sh, Sg, eps); break;
            case 1: flow_phase2<T, 1, STAGE, XL>(L, win, wout, zs, Z0, c, p, i, s, sh, Sg, eps); break;
            case 2: flow_phase2<T, 2, STAGE, XL>(L, win, wout, zs, Z0, c, p, i, s, sh, Sg, eps); break;
            default: flow_phase2<T, 3, STAGE, XL>(L, win, wout, zs, Z0, c, p, i, s, sh, Sg, eps); break;
        }
    }
}

template <class T, int STAGE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(MinWaves<T>::flow))) void k_flow_stage(const LatDev<T>* __restrict__ lats, const uint32_t* __restrict__ blk,
                                                       int n_lat, uint32_t nblocks, uint32_t gstep, T eps) {
    __shared__ FlowLds<T> sh;
    BlockCtx<T> bc = block_ctx(lats, blk, n_lat, nblocks);
    const LatDev<T>& L = *bc.L;
    if (gstep - L.flow_start >= L.flow_limit) return;   // the whole block: its lattice has finished flowing
    // a block holds L.fsb sites (whole x rows, or 32 sites of partial rows);
    // site lanes from L.fsb on are idle
    const uint32_t s = threadIdx.x & (kLinkSites - 1);
    const uint32_t q = bc.local * L.fsb + s;
    // wavefront w of the block takes direction (w + block) mod 4, so that
    // each SIMD runs every direction's share of the work
    const uint32_t hs = (((threadIdx.x >> 6) + blockIdx.x) & 3u) * 2u + ((threadIdx.x >> 5) & 1u);
    // every thread reaches the barrier; a site past the end has all four of
    // its links past the end, so no live thread reads what it did not write
    const bool live = s < L.fsb && q < L.Vh;
    const int mu = (int)(hs >> 1), p = (int)(hs & 1);
    const int cur = (int)((gstep - L.flow_start) & 1u);
    const T* win = L.F[(STAGE == 2) ? 1 - cur : cur];
    T* wout = L.F[(STAGE == 2) ? cur : 1 - cur];
    T* zs = L.F[2];
    Coords c = tiled_coords(L, live ? q : 0u, p);
    const uint32_t i = cb_index(L, c.x, c.y, c.z, c.t);
    // whole x rows per block (Lx/2 divides its sites; the rows start at q =
    // 0 and every work-order tile is whole rows): the (mu, 0) lower staples
    // come from LDS; the branch is uniform per block
    if ((L.fsb % (uint32_t)L.Lxh) == 0u)
        flow_stage_body<T, STAGE, true>(L, win, wout, zs, live, mu, c, p, i, s, sh, eps);
    else
        flow_stage_body<T, STAGE, false>(L, win, wout, zs, live, mu, c, p, i, s, sh, eps);
}

// ---------------------------------------------------------- math probes ----

template <class T>
__global__ void k_math_probe(int op, int n, const double* __restrict__ in, double param, int aux,
                             double* __restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (op == 0 || op == 1 || op == 2) {
        M3<T> a;
        for (int k = 0; k < 18; ++k) a.v[k] = (T)in[(size_t)t * 18 + k];
        if (op == 0) {
            // the flow's exponential: only the upper triangle and the imaginary
            // diagonal of the (anti-Hermitian) input are read
            A9<T> z;
            z.o[0] = a.re(0, 1); z.o[1] = a.im(0, 1);
            z.o[2] = a.re(0, 2); z.o[3] = a.im(0, 2);
            z.o[4] = a.re(1, 2); z.o[5] = a.im(1, 2);
            for (int k = 0; k < 3; ++k) z.d[k] = a.im(k, k);
            a = exp_su3_ah<T, ExpTerms<T>::N>(a9_scale((T)param, z));
        } else if (op == 1) {
            proj_lie(a);
        } else {
            M3<double> d;
            for (int k = 0; k < 18; ++k) d.v[k] = (double)a.v[k];
            reunitarize_d(d);
            for (int k = 0; k < 18; ++k) a.v[k] = (T)d.v[k];
        }
        for (int k = 0; k < 18; ++k) out[(size_t)t * 18 + k] = (double)a.v[k];
    } else if (op == 3) {
        const int K = aux, w = 1 + 4 * K + 2;
        const double* u = in + (size_t)t * w;
        T k = (T)u[0];
        const T nhia = KpMath::neg_half_over(KpMath::alpha((T)param, k));
        T a0 = T(0), x[4] = {T(0), T(0), T(0), T(0)};
        int acc = -1;
        for (int a = 0; a < K; ++a)
            if (kp_try<T>(nhia, (T)u[1 + 4 * a], (T)u[2 + 4 * a], (T)u[3 + 4 * a], (T)u[4 + 4 * a], a0)) {
                kp_sphere<T>(a0, (T)u[1 + 4 * K], (T)u[2 + 4 * K], x);
                acc = a;
                break;
            }
        out[(size_t)t * 5] = acc;
        for (int c = 0; c < 4; ++c) out[(size_t)t * 5 + 1 + c] = (double)x[c];
    } else if (op == 4) {
        const int K = aux, nslot = 4 * K + 2, w = 36 + 3 * nslot;
        const double* src = in + (size_t)t * w;
        M3<T> U, A;
        for (int k = 0; k < 18; ++k) { U.v[k] = (T)src[k]; A.v[k] = (T)src[18 + k]; }
        // uniforms laid out [g][slot] with stride 1: reuse the injected path with Vh = 1
        uint32_t stash[2] = {0u, 0u};
        cm_hit<T, kHbInjected>(U, A, (T)param, 0, 0, 0, 0, 0, 0, src + 36, 1, K, 1, stash, (uint32_t)K);
        for (int k = 0; k < 18; ++k) out[(size_t)t * 18 + k] = (double)U.v[k];
    } else if (op == 5) {
        out[t] = (double)KpMath::uni<T>((uint32_t)in[t]);
    }
}

}  // namespace

// ============================================================ launchers ====

#define QCB_ARGS(g) static_cast<const LatDev<T>*>((g).lats), (g).blk, (g).n_lat, (g).nblocks

template <class T>
void launch_cold(const Launch& g) {
    hipLaunchKernelGGL(k_cold<T>, dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g));
}
template <class T>
void launch_hot(const Launch& g, Gate gate) {
    hipLaunchKernelGGL(k_hot<T>, dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), gate);
}
template <class T>
void launch_aos_to_soa(const Launch& g, const double* aos, T* dst) {
    hipLaunchKernelGGL(k_aos_to_soa<T>, dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), aos, dst);
}
template <class T>
void launch_soa_to_aos(const Launch& g, const T* src, double* aos, int order) {
    if (order == kFlowOrder)
        hipLaunchKernelGGL((k_soa_to_aos<T, kFlowOrder>), dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), src, aos);
    else
        hipLaunchKernelGGL((k_soa_to_aos<T, kLinkOrder>), dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), src, aos);
}
template <class T>
void launch_soa_to_be(const Launch& g, const T* src, uint64_t* out, uint64_t lo, uint32_t n) {
    hipLaunchKernelGGL(k_soa_to_be<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, g.stream,
                       static_cast<const LatDev<T>*>(g.lats), src, out, lo, n);
}
template <class T>
void launch_be_to_soa(const Launch& g, const uint64_t* in, T* dst, uint64_t lo, uint32_t n) {
    hipLaunchKernelGGL(k_be_to_soa<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, g.stream,
                       static_cast<const LatDev<T>*>(g.lats), in, dst, lo, n);
}
template <class T>
void launch_copy_field(const Launch& g, const T* src, T* dst, int from, int to) {
    const dim3 gr(g.nblocks), bl(kBlock);
    if (from == kLinkOrder && to == kFlowOrder)
        hipLaunchKernelGGL((k_copy_field<T, kLinkOrder, kFlowOrder>), gr, bl, 0, g.stream, QCB_ARGS(g), src, dst);
    else if (from == kFlowOrder && to == kLinkOrder)
        hipLaunchKernelGGL((k_copy_field<T, kFlowOrder, kLinkOrder>), gr, bl, 0, g.stream, QCB_ARGS(g), src, dst);
    else if (from == kFlowOrder)
        hipLaunchKernelGGL((k_copy_field<T, kFlowOrder, kFlowOrder>), gr, bl, 0, g.stream, QCB_ARGS(g), src, dst);
    else
        hipLaunchKernelGGL((k_copy_field<T, kLinkOrder, kLinkOrder>), gr, bl, 0, g.stream, QCB_ARGS(g), src, dst);
}
template <class T>
void launch_flow_init(const Launch& g, Gate gate) {
    hipLaunchKernelGGL(k_flow_init<T>, dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), gate);
}
template <class T>
void launch_plaquette(const Launch& g, double* partial, int which, uint32_t gstep) {
    if (which)
        hipLaunchKernelGGL((k_plaquette<T, kFlowOrder>), dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g),
                           gstep, partial);
    else
        hipLaunchKernelGGL((k_plaquette<T, kLinkOrder>), dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g),
                           gstep, partial);
}
template <class T>
void launch_clover(const Launch& g, int which, uint32_t gstep, double* pq, double* pe) {
    if (which)
        hipLaunchKernelGGL((k_clover<T, kFlowOrder>), dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), gstep,
                           pq, pe);
    else
        hipLaunchKernelGGL((k_clover<T, kLinkOrder>), dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), gstep,
                           pq, pe);
}
template <class T>
void launch_staples(const Launch& g, double* out) {
    hipLaunchKernelGGL(k_staples<T>, dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), out);
}
template <class T>
void launch_reunit(const Launch& g, Gate gate) {
    hipLaunchKernelGGL(k_reunit<T>, dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), gate);
}
template <class T>
void launch_translate(const Launch& g, const int s[4], T* dst) {
    hipLaunchKernelGGL(k_translate<T>, dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), s[0], s[1], s[2], s[3],
                       dst);
}

template <class T, int MU>
static void hb_mu(const Launch& g, int p, int n_hits, uint32_t sweep, const uint32_t* base, int mode, const double* inj,
                  int K, uint32_t grid) {
    if (mode == kHbInjected)
        hipLaunchKernelGGL((k_hb_half<T, MU, kHbInjected>), dim3(grid), dim3(kBlock), 0, g.stream, QCB_ARGS(g), p,
                           n_hits, sweep, base, inj, K);
    else
        hipLaunchKernelGGL((k_hb_half<T, MU, kHbPhilox>), dim3(grid), dim3(kBlock), 0, g.stream, QCB_ARGS(g), p,
                           n_hits, sweep, base, inj, K);
}

template <class T>
void launch_hb_half(const Launch& g, int mu, int p, int n_hits, uint32_t sweep, int mode, const double* inj, int K,
                    const uint32_t* sweep_base, uint32_t grid) {
    if (grid == 0 || grid > g.nblocks) grid = g.nblocks;
    switch (mu) {
        case 0: hb_mu<T, 0>(g, p, n_hits, sweep, sweep_base, mode, inj, K, grid); break;
        case 1: hb_mu<T, 1>(g, p, n_hits, sweep, sweep_base, mode, inj, K, grid); break;
        case 2: hb_mu<T, 2>(g, p, n_hits, sweep, sweep_base, mode, inj, K, grid); break;
        default: hb_mu<T, 3>(g, p, n_hits, sweep, sweep_base, mode, inj, K, grid); break;
    }
}

void launch_bump(uint32_t* counter, uint32_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_bump, dim3(1), dim3(64), 0, st, counter, n);
}

template <class T>
void launch_flow_stage(const Launch& g, int stage, uint32_t gstep, double eps) {
    if (stage == 1)
        hipLaunchKernelGGL((k_flow_stage<T, 1>), dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), gstep, (T)eps);
    else if (stage == 2)
        hipLaunchKernelGGL((k_flow_stage<T, 2>), dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), gstep, (T)eps);
    else
        hipLaunchKernelGGL((k_flow_stage<T, 3>), dim3(g.nblocks), dim3(kBlock), 0, g.stream, QCB_ARGS(g), gstep, (T)eps);
}

template <class T>
void launch_math_probe(int op, int n, const double* in, double param, int aux, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_math_probe<T>, dim3((n + 63) / 64), dim3(64), 0, st, op, n, in, param, aux, out);
}

template <class T>
void launch_finalize_plaq(const Launch& g, const double* partial, double* out, int ostride) {
    hipLaunchKernelGGL(k_finalize_plaq<T>, dim3(g.n_lat), dim3(kBlock), 0, g.stream,
                       static_cast<const LatDev<T>*>(g.lats), g.blk, partial, out, ostride);
}
template <class T>
void launch_finalize_qe(const Launch& g, const double* pq, const double* pe, double* out, int ostride,
                        uint32_t gstep) {
    hipLaunchKernelGGL(k_finalize_qe<T>, dim3(g.n_lat), dim3(kBlock), 0, g.stream,
                       static_cast<const LatDev<T>*>(g.lats), g.blk, pq, pe, out, ostride, gstep);
}

#define QCB_INSTANTIATE(T)                                                                                     \
    template void launch_cold<T>(const Launch&);                                                              \
    template void launch_hot<T>(const Launch&, Gate);                                                         \
    template void launch_aos_to_soa<T>(const Launch&, const double*, T*);                                     \
    template void launch_soa_to_aos<T>(const Launch&, const T*, double*, int);                                \
    template void launch_copy_field<T>(const Launch&, const T*, T*, int, int);                                \
    template void launch_soa_to_be<T>(const Launch&, const T*, uint64_t*, uint64_t, uint32_t);                 \
    template void launch_be_to_soa<T>(const Launch&, const uint64_t*, T*, uint64_t, uint32_t);                 \
    template void launch_flow_init<T>(const Launch&, Gate);                                                   \
    template void launch_plaquette<T>(const Launch&, double*, int, uint32_t);                                                \
    template void launch_clover<T>(const Launch&, int, uint32_t, double*, double*);                           \
    template void launch_staples<T>(const Launch&, double*);                                                       \
    template void launch_reunit<T>(const Launch&, Gate);                                                      \
    template void launch_translate<T>(const Launch&, const int*, T*);                                         \
    template void launch_hb_half<T>(const Launch&, int, int, int, uint32_t, int, const double*, int,           \
                                    const uint32_t*, uint32_t);                                               \
    template void launch_flow_stage<T>(const Launch&, int, uint32_t, double);                                 \
    template void launch_math_probe<T>(int, int, const double*, double, int, double*, hipStream_t);           \
    template void launch_finalize_plaq<T>(const Launch&, const double*, double*, int);                        \
    template void launch_finalize_qe<T>(const Launch&, const double*, const double*, double*, int, uint32_t);

QCB_INSTANTIATE(float)
QCB_INSTANTIATE(double)

}  // namespace qcb
