// qcb_layout.hpp -- device link storage and periodic even/odd indexing.
//
// HBM layout (one lattice): blocks of 64 sites, complex entries innermost
// across the block (AoSoA), even/odd split,
//   field[((hs*NB + cb/64)*9 + k)*64 + cb%64]   (one complex = 2 reals)
// hs = 2*mu + parity, k = 3*i + j, cb = lex/2 with lex = x + Lx*(y + Ly*(z +
// Lz*t)), parity = (x+y+z+t) & 1, NB = S/64 and S = Vh rounded up to 64.  A
// 64-lane wavefront reading entry k of 64 consecutive cb sites issues one
// fully coalesced 512-byte (float) / 1 KiB (double) load, and the 9 entries
// of one link sit at fixed 512 B / 1 KiB offsets from each other, so a link
// costs one address computation.  The reference keeps an AoS field with a
// one-site halo and a neighbour table (src/gauge/GaugeField.h:27,
// src/geometry/GeometryCB.cpp:10-31); here neighbours are closed-form
// periodic index arithmetic, so the halo exchange of src/mpi/HalosExchange.cpp
// has no counterpart.
//
// Work order: kernels that read neighbours visit sites in a 4D-tiled order
// (tiles of Lx x T1 x T2 x T3 sites, tiles y-fastest), so that the +-t and +-z
// neighbours of most sites were fetched by the same or a recent block and
// are still in L2.  The tiling changes only which thread handles which site,
// never the cb index a site's data or random numbers live at.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qcb_su3.hpp"

namespace qcb {

constexpr int kBlock = 256;          // threads per block of every lattice kernel
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, blocks dispatched round-robin

// n / d for any 32-bit n by a multiplier fixed per divisor (Granlund and
// Montgomery, "Division by invariant integers using multiplication", PLDI
// 1994, sec. 4): with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
// t = mulhi(m, n) and n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0).
// One quarter-rate multiply instead of a reciprocal-and-correct sequence.
struct FastDiv {
    uint32_t m, s1, s2;
};

__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv f) {
    const uint32_t t = __umulhi(n, f.m);
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

// Per-lattice descriptor, resident in device memory; one entry per lattice of
// a batch (a single lattice is a batch of one).
template <class T>
struct LatDev {
    T* U;                 // gauge links
    T* F[3];              // flow buffers: two link fields and one force field
    int L[4];             // x, y, z, t extents (x even)
    int Lxh;              // Lx / 2
    uint32_t Vh;          // V / 2
    uint32_t S;           // component stride (Vh rounded up to 64)
    double beta;
    uint32_t key0, key1;  // Philox key (seed)
    uint32_t flow_start;  // global flow-step index at which this lattice's flow began
    uint32_t sweep_off;   // added to the launch's Philox sweep index (per-lattice sweep count)
    int TL[4];            // tile extents in sites (TL[0] = Lx: whole x rows)
    int NT[4];            // tiles per direction (NT[0] = 1)
    uint32_t tvol;        // cb sites per tile = (Lx/2) * T1 * T2 * T3
    int morton;           // tiles in Z-order over (ty, tz, tt) (NT powers of two) instead of y-fastest
    int tbits[4];         // log2 NT[d] when morton
    uint32_t kp_cap;      // Kennedy-Pendleton attempts per subgroup draw in Philox mode
    unsigned long long* kp_exh;  // device count of draws that ran out of attempts (skipped updates)
    uint32_t flow_limit;  // RK3 steps this lattice takes after flow_start (later stage launches skip it)
    uint32_t fsb;         // cb sites per 32 site lanes of the flow stage and the clover (flow_block_sites)
    double* qe_out;       // if set, a measurement at flow step s writes (Q, E/V) to qe_out[2s], qe_out[2s+1]
    FastDiv dv_tvol, dv_lxh, dv_t1, dv_t2, dv_nt1, dv_nt2;   // by tvol, Lxh, TL[1], TL[2], NT[1], NT[2]
};

// A per-lattice cadence for kernels launched over a set of lattices that are
// each at their own point of a schedule: lattice L takes part iff mod == 0 or
// (sweep + L.sweep_off) % mod == res, i.e. its own sweep count hits the cadence.
struct Gate {
    uint32_t sweep = 0, mod = 0, res = 0;
};

// Threads per block of each kernel family and the work item of a thread:
//   half-sweep: 256 cb sites of one (mu, parity)
//   site kernels: 128 cb sites x 2 parities
//   link kernels: 32 cb sites x 8 (mu, parity) classes
constexpr uint32_t kHalfSites = 256, kSiteSites = 128, kLinkSites = 32;

// cb sites per 32 site lanes (a flow-stage block, a clover wavefront), at
// most kLinkSites: whole x rows of Lx/2 sites when they fill 30 or more of
// the 32 lanes, which lets the flow stage take its (mu, 0) lower staples and
// the clover its shared leaves from LDS -- 32 when Lx/2 divides 32, 30 for
// Lx/2 = 3, 5, 6, 10, 15 (the other lanes idle) -- else 32 sites that need
// not be whole rows.
__host__ __device__ inline uint32_t flow_block_sites(uint32_t lxh) {
    const uint32_t rows = lxh <= kLinkSites ? lxh * (kLinkSites / lxh) : 0u;
    return rows + 2u >= kLinkSites ? rows : kLinkSites;
}

struct Coords {
    int x, y, z, t;
};

template <class T>
__device__ __forceinline__ Coords cb_coords(const LatDev<T>& L, uint32_t i, int p) {
    Coords c;
    uint32_t r = i / (uint32_t)L.Lxh;
    int xh = (int)(i - r * (uint32_t)L.Lxh);
    uint32_t r2 = r / (uint32_t)L.L[1];
    c.y = (int)(r - r2 * (uint32_t)L.L[1]);
    uint32_t r3 = r2 / (uint32_t)L.L[2];
    c.z = (int)(r2 - r3 * (uint32_t)L.L[2]);
    c.t = (int)r3;
    c.x = 2 * xh + ((c.y + c.z + c.t + p) & 1);
    return c;
}

template <class T>
__device__ __forceinline__ uint32_t cb_index(const LatDev<T>& L, int x, int y, int z, int t) {
    return (uint32_t)(x >> 1) + (uint32_t)L.Lxh * ((uint32_t)y + (uint32_t)L.L[1] * ((uint32_t)z + (uint32_t)L.L[2] * (uint32_t)t));
}

// Coordinates of the q-th site of parity p in the tiled work order.
// The divisions use the descriptor's multipliers and the remainders 24-bit
// products (tile counts and tile volumes stay below 2^24).
template <class T>
__device__ __forceinline__ Coords tiled_coords(const LatDev<T>& L, uint32_t q, int p) {
    uint32_t tau = fdiv(q, L.dv_tvol);
    uint32_t r = q - __umul24(tau, L.tvol);
    uint32_t r1 = fdiv(r, L.dv_lxh);
    int xh = (int)(r - __umul24(r1, (uint32_t)L.Lxh));
    uint32_t r2 = fdiv(r1, L.dv_t1);
    int yl = (int)(r1 - __umul24(r2, (uint32_t)L.TL[1]));
    uint32_t r3 = fdiv(r2, L.dv_t2);
    int zl = (int)(r2 - __umul24(r3, (uint32_t)L.TL[2]));
    int tl = (int)r3;
    int ty, tz, tt;
    if (L.morton) {
        // bit b of each tile coordinate, y then z then t, for b = 0, 1, ...
        ty = tz = tt = 0;
        int pos = 0;
        const int maxb = max(L.tbits[1], max(L.tbits[2], L.tbits[3]));
        for (int b = 0; b < maxb; ++b) {
            if (b < L.tbits[1]) ty |= (int)((tau >> pos++) & 1u) << b;
            if (b < L.tbits[2]) tz |= (int)((tau >> pos++) & 1u) << b;
            if (b < L.tbits[3]) tt |= (int)((tau >> pos++) & 1u) << b;
        }
    } else {
        uint32_t u1 = fdiv(tau, L.dv_nt1);
        ty = (int)(tau - __umul24(u1, (uint32_t)L.NT[1]));
        uint32_t u2 = fdiv(u1, L.dv_nt2);
        tz = (int)(u1 - __umul24(u2, (uint32_t)L.NT[2]));
        tt = (int)u2;
    }
    Coords c;
    c.y = (int)__umul24((uint32_t)ty, (uint32_t)L.TL[1]) + yl;
    c.z = (int)__umul24((uint32_t)tz, (uint32_t)L.TL[2]) + zl;
    c.t = (int)__umul24((uint32_t)tt, (uint32_t)L.TL[3]) + tl;
    c.x = 2 * xh + ((c.y + c.z + c.t + p) & 1);
    return c;
}

__device__ __forceinline__ int wrap_up(int v, int n) { return v + 1 == n ? 0 : v + 1; }
__device__ __forceinline__ int wrap_dn(int v, int n) { return v == 0 ? n - 1 : v - 1; }

// Change of the cb index for a step s = +-1 along d from site c, periodic.
// With cb = (x >> 1) + Lxh (y + Ly (z + Lz t)), a step in y, z or t adds
// +-stride_d, less the period L_d stride_d where it wraps, and a step in x
// moves x >> 1 only when it leaves an odd x upwards or an even x downwards.
// The steps are wave-uniform scalars selected by one compare each, so a
// neighbour costs a select and an add instead of a cb_index product chain.
template <class T>
__device__ __forceinline__ uint32_t cb_step(const LatDev<T>& L, Coords c, int d, int s) {
    const uint32_t lxh = (uint32_t)L.Lxh;
    if (d == 0) {
        if (s > 0) return (c.x & 1) ? (c.x == L.L[0] - 1 ? 1u - lxh : 1u) : 0u;
        return (c.x & 1) ? 0u : (c.x == 0 ? lxh - 1u : ~0u);
    }
    uint32_t str = lxh;
    for (int k = 1; k < d; ++k) str *= (uint32_t)L.L[k];
    const uint32_t per = str * (uint32_t)L.L[d];
    const int v = d == 1 ? c.y : (d == 2 ? c.z : c.t);
    if (s > 0) return v == L.L[d] - 1 ? str - per : str;
    return v == 0 ? per - str : 0u - str;
}

// cb index of x + s1*e_d1 + s2*e_d2 (s in {-1,0,+1}; d1 != d2 when both non-zero)
template <class T>
__device__ __forceinline__ uint32_t nb_cb(const LatDev<T>& L, Coords c, int d1, int s1, int d2 = 0, int s2 = 0) {
    uint32_t i = cb_index(L, c.x, c.y, c.z, c.t);
    if (s1 != 0) i += cb_step(L, c, d1, s1);
    if (s2 != 0) i += cb_step(L, c, d2, s2);
    return i;
}

template <class T>
struct alignas(2 * sizeof(T)) Cplx {
    T re, im;
};

// Complex entries stored per link.  FP64 links keep all nine (the oracle's
// arithmetic, bit for bit); FP32 links keep rows 0 and 1 and rebuild row 2 on
// load (su3_third_row), so every FP32 kernel moves 2/3 of the bytes.  An FP32
// lattice therefore holds SU(3) matrices only, as a gauge field does.
template <class T> struct LinkSlots { static constexpr uint32_t n = 9; };
template <> struct LinkSlots<float> { static constexpr uint32_t n = 6; };

// Field pointers live in descriptors (generic pointers); casting them to the
// global address space lets the compiler emit global_load/global_store
// instead of flat accesses, whose waits also block on LDS traffic.
#define QCB_GLOBAL __attribute__((address_space(1)))

template <class T> struct Vec2;
template <> struct Vec2<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct Vec2<double> { typedef double type __attribute__((ext_vector_type(2))); };

// Storage order of a link field, set by its role.  kLinkOrder (the gauge
// links U): one complex entry per slot, entry k of 64 consecutive sites at
// slot offset 64 k.  kFlowOrder (the flow's two link fields F[0], F[1]): FP32
// entries in pairs, one 16-byte slot per pair, so a link is three 16-byte
// loads instead of six 8-byte ones.  The flow and clover kernels issue fewer
// load instructions that way (11% and 9% faster at 32^4), while the heatbath's
// L1 hit rate falls from 0.77 to 0.56 with 16-byte slots and it runs 4-7%
// slower, so the links keep 8-byte slots.  FP64 entries are 16 bytes in both.
constexpr int kLinkOrder = 0, kFlowOrder = 1;
template <class T, int LAY>
struct SlotOf {
    typedef typename Vec2<T>::type type;
    static constexpr uint32_t per = 1;   // complex entries per slot
};
template <>
struct SlotOf<float, kFlowOrder> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr uint32_t per = 2;
};

// Slot index of entry 0 of the link at cb index i of class hs in a field of
// n slots per link; entry j is 64*j further.  Class hs starts at hs * n * S;
// inside it the link sits at (i >> 6) * 64 n + (i & 63) = i + (n - 1)(i & ~63),
// which is below n S <= 9 * 2^28 < 2^32 (S <= 2^28 since 4V < 2^31), so only
// the class offset is 64-bit: 2 (hs >> 1) n S, wave-uniform where the
// direction is, plus n S for an odd class -- a select and an add per link
// instead of two 64-bit multiply-adds.
__device__ __forceinline__ size_t aosoa_slot(uint32_t n, uint32_t S, int hs, uint32_t i) {
    const uint32_t cs = n * S;
    const size_t cls = (size_t)cs * (uint32_t)(hs & ~1) + ((hs & 1) ? (size_t)cs : (size_t)0);
    return cls + (size_t)(i + (n - 1u) * (i & ~63u));
}
template <class T, int LAY = kLinkOrder>
__device__ __forceinline__ size_t link_base(uint32_t S, int hs, uint32_t i) {
    return aosoa_slot(LinkSlots<T>::n / SlotOf<T, LAY>::per, S, hs, i);
}

// The stored complex entries of a link (rows 0-2 in FP64, rows 0-1 in FP32).
template <class T, int LAY, bool NT = false>
__device__ __forceinline__ M3<T> ld_stored(const T* f, uint32_t S, int mu, int p, uint32_t i) {
    typedef SlotOf<T, LAY> SO;
    typedef typename SO::type V;
    constexpr int slots = (int)(LinkSlots<T>::n / SO::per), w = 2 * (int)SO::per;
    const QCB_GLOBAL V* b = (const QCB_GLOBAL V*)f + link_base<T, LAY>(S, mu * 2 + p, i);
    M3<T> m;
#pragma unroll
    for (int j = 0; j < slots; ++j) {
        V c = NT ? __builtin_nontemporal_load(b + 64 * j) : b[64 * j];
#pragma unroll
        for (int e = 0; e < w; ++e) m.v[w * j + e] = c[e];
    }
    return m;
}

template <class T, int LAY = kLinkOrder, bool NT = false>
__device__ __forceinline__ M3<T> ld_link(const T* f, uint32_t S, int mu, int p, uint32_t i) {
    M3<T> m = ld_stored<T, LAY, NT>(f, S, mu, p, i);
    if (LinkSlots<T>::n == 6) su3_third_row(m);
    return m;
}

// Rows 0 and 1 of an FP32 link as stored, row 2 left unset: for a product
// whose row 2 is rebuilt afterwards (mul_nn / mul_nd with R = 2).
template <int LAY = kLinkOrder>
__device__ __forceinline__ M3<float> ld_link_rows01(const float* f, uint32_t S, int mu, int p, uint32_t i) {
    return ld_stored<float, LAY>(f, S, mu, p, i);
}

// Every kernel writes whole fields that no thread of the same launch reads
// back, so stores are non-temporal: they stream through L2 without evicting
// the neighbour links the staples are still reusing.
template <class T, int LAY = kLinkOrder>
__device__ __forceinline__ void st_link(T* f, uint32_t S, int mu, int p, uint32_t i, const M3<T>& m) {
    typedef SlotOf<T, LAY> SO;
    typedef typename SO::type V;
    constexpr int slots = (int)(LinkSlots<T>::n / SO::per), w = 2 * (int)SO::per;
    QCB_GLOBAL V* b = (QCB_GLOBAL V*)f + link_base<T, LAY>(S, mu * 2 + p, i);
#pragma unroll
    for (int j = 0; j < slots; ++j) {
        V v;
#pragma unroll
        for (int e = 0; e < w; ++e) v[e] = m.v[w * j + e];
        __builtin_nontemporal_store(v, b + 64 * j);
    }
}

// Force fields hold traceless anti-Hermitian matrices (A9 in qcb_su3.hpp) by
// their 8 independent reals: the upper triangle (o0..o5) and the imaginary
// diagonal entries d0, d1; d2 = -(d0 + d1) on load.  They sit in 16-byte
// slots in the links' AoSoA-64 order, (o0 o1 o2 o3)(o4 o5 d0 d1) in FP32 and
// (o0 o1)(o2 o3)(o4 o5)(d0 d1) in FP64: 32 / 64 bytes per link, two / four
// loads.  Each force entry is read once, by the thread of its own link, so
// loads as well as stores are non-temporal: the force stream does not evict
// the neighbour links from the caches (FP32 flow stages 2 and 3 3-4% faster
// than with cached loads).
template <class T>
struct ForceSlot {
    typedef T type __attribute__((ext_vector_type(16 / sizeof(T))));
    static constexpr int per = 16 / (int)sizeof(T), n = 8 / per;   // reals per slot, slots per link
};

template <class T>
__device__ __forceinline__ size_t a9_base(uint32_t S, int hs, uint32_t i) {
    return aosoa_slot((uint32_t)ForceSlot<T>::n, S, hs, i);
}

template <class T>
__device__ __forceinline__ A9<T> ld_a9(const T* f, uint32_t S, int mu, int p, uint32_t i) {
    typedef ForceSlot<T> FS;
    typedef typename FS::type V;
    const QCB_GLOBAL V* b = (const QCB_GLOBAL V*)f + a9_base<T>(S, mu * 2 + p, i);
    T r[8];
#pragma unroll
    for (int j = 0; j < FS::n; ++j) {
        const V c = __builtin_nontemporal_load(b + 64 * j);
#pragma unroll
        for (int e = 0; e < FS::per; ++e) r[FS::per * j + e] = c[e];
    }
    A9<T> z;
#pragma unroll
    for (int k = 0; k < 6; ++k) z.o[k] = r[k];
    z.d[0] = r[6];
    z.d[1] = r[7];
    z.d[2] = -(r[6] + r[7]);
    return z;
}

template <class T>
__device__ __forceinline__ void st_a9(T* f, uint32_t S, int mu, int p, uint32_t i, const A9<T>& z) {
    typedef ForceSlot<T> FS;
    typedef typename FS::type V;
    QCB_GLOBAL V* b = (QCB_GLOBAL V*)f + a9_base<T>(S, mu * 2 + p, i);
    const T r[8] = {z.o[0], z.o[1], z.o[2], z.o[3], z.o[4], z.o[5], z.d[0], z.d[1]};
#pragma unroll
    for (int j = 0; j < FS::n; ++j) {
        V v;
#pragma unroll
        for (int e = 0; e < FS::per; ++e) v[e] = r[FS::per * j + e];
        __builtin_nontemporal_store(v, b + 64 * j);
    }
}

// Block index -> work tile such that each XCD owns one contiguous range of
// tiles (blocks are dispatched to XCD b % 8), keeping neighbour rows of the
// lattice in the same L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
    uint32_t q = nb / kNumXcd, r = nb % kNumXcd;
    uint32_t x = b % kNumXcd, j = b / kNumXcd;
    return x * q + (x < r ? x : r) + j;
}

// Lattice owning tile b: blk_begin holds n_lat+1 prefix sums of tile counts.
__device__ __forceinline__ int find_lat(const uint32_t* __restrict__ blk_begin, int n_lat, uint32_t b) {
    int lo = 0, hi = n_lat;  // invariant blk_begin[lo] <= b < blk_begin[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (blk_begin[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

// Staple Sigma_mu(x) = sum_{nu != mu} U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
//                               + U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)
// so that U_mu(x) Sigma_mu(x) is the sum of the six plaquettes through the
// link (reference src/gauge/GaugeField.cpp:74-97).  x has parity p and cb i.
template <class T, int MU>
__device__ __forceinline__ M3<T> staple(const LatDev<T>& L, const T* __restrict__ f, Coords c, int p, uint32_t i) {
    M3<T> A;
    m3_zero(A);
    const uint32_t S = L.S;
    const int q = p ^ 1;
    const uint32_t xmu = nb_cb(L, c, MU, +1);
#pragma unroll
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == MU) continue;
        const uint32_t xnu = nb_cb(L, c, nu, +1);
        const uint32_t xmnu = nb_cb(L, c, nu, -1);
        const uint32_t xmumnu = nb_cb(L, c, MU, +1, nu, -1);
        if constexpr (LinkSlots<T>::n == 6) {
            // FP32: the first factor's stored rows give rows 0, 1 of the
            // product t (an SU(3) matrix), whose row 2 is then rebuilt; the
            // stored rows of U_nu(x) give columns 0, 1 of t U_nu(x)^dag, whose
            // column 2 is rebuilt
            {
                M3<T> a = ld_link_rows01(f, S, nu, q, xmu);   // U_nu(x+mu)
                M3<T> b = ld_link(f, S, MU, q, xnu);          // U_mu(x+nu)
                M3<T> t = mul_nd<T, 2>(a, b);
                su3_third_row(t);
                M3<T> u = mul_part<T, false, true, 3, 2>(t, ld_link_rows01(f, S, nu, p, i));   // U_nu(x)
                su3_third_col(u);
                m3_add(A, u);
            }
            {
                M3<T> a = ld_link_rows01(f, S, MU, q, xmnu);  // U_mu(x-nu)
                M3<T> b = ld_link(f, S, nu, p, xmumnu);       // U_nu(x+mu-nu)
                M3<T> t = mul_nn<T, 2>(a, b);
                su3_third_row(t);
                M3<T> d = ld_link(f, S, nu, q, xmnu);         // U_nu(x-nu)
                acc_dn(A, t, d);
            }
        } else {
            {
                M3<T> a = ld_link(f, S, nu, q, xmu);   // U_nu(x+mu)
                M3<T> b = ld_link(f, S, MU, q, xnu);   // U_mu(x+nu)
                M3<T> t = mul_nd(a, b);
                M3<T> d = ld_link(f, S, nu, p, i);     // U_nu(x)
                acc_nd(A, t, d);
            }
            {
                M3<T> a = ld_link(f, S, MU, q, xmnu);    // U_mu(x-nu)
                M3<T> b = ld_link(f, S, nu, p, xmumnu);  // U_nu(x+mu-nu)
                M3<T> t = mul_nn(a, b);
                M3<T> d = ld_link(f, S, nu, q, xmnu);    // U_nu(x-nu)
                acc_dn(A, t, d);
            }
        }
    }
    return A;
}

}  // namespace qcb
