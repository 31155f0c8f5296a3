// qcb_layout.hpp -- device link storage and periodic even/odd indexing.
//
// HBM layout (one lattice): structure-of-arrays, even/odd split,
//   field[((mu*2 + parity)*18 + comp)*S + cb]
// comp = 2*(3*i + j) + {re,im}, cb = lex/2 with lex = x + Lx*(y + Ly*(z + Lz*t)),
// parity = (x+y+z+t) & 1 and S = Vh rounded up to 64 (each component array
// starts 256-byte aligned in float).  A 64-lane wavefront reading one
// component of 64 consecutive cb sites issues one fully coalesced 256/512-byte
// load.  The reference keeps an AoS field with a one-site halo and a
// neighbour table (src/gauge/GaugeField.h:27, src/geometry/GeometryCB.cpp:10-31);
// here neighbours are closed-form periodic index arithmetic, so the halo
// exchange of src/mpi/HalosExchange.cpp has no counterpart.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qcb_su3.hpp"

namespace qcb {

constexpr int kBlock = 256;          // threads per block of every lattice kernel
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, blocks dispatched round-robin

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
    uint32_t pad;
};

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

__device__ __forceinline__ int wrap_up(int v, int n) { return v + 1 == n ? 0 : v + 1; }
__device__ __forceinline__ int wrap_dn(int v, int n) { return v == 0 ? n - 1 : v - 1; }

// cb index of x + s1*e_d1 + s2*e_d2 (s in {-1,0,+1}; d1 != d2 when both non-zero)
template <class T>
__device__ __forceinline__ uint32_t nb_cb(const LatDev<T>& L, Coords c, int d1, int s1, int d2 = 0, int s2 = 0) {
    int v[4] = {c.x, c.y, c.z, c.t};
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        if (d == d1 && s1 > 0) v[d] = wrap_up(v[d], L.L[d]);
        if (d == d1 && s1 < 0) v[d] = wrap_dn(v[d], L.L[d]);
        if (d == d2 && s2 > 0) v[d] = wrap_up(v[d], L.L[d]);
        if (d == d2 && s2 < 0) v[d] = wrap_dn(v[d], L.L[d]);
    }
    return cb_index(L, v[0], v[1], v[2], v[3]);
}

template <class T>
__device__ __forceinline__ const T* comp_base(const T* f, uint32_t S, int mu, int p) {
    return f + (size_t)((mu * 2 + p) * 18) * S;
}

template <class T>
__device__ __forceinline__ M3<T> ld_link(const T* f, uint32_t S, int mu, int p, uint32_t i) {
    const T* b = comp_base(f, S, mu, p) + i;
    M3<T> m;
#pragma unroll
    for (int k = 0; k < 18; ++k) m.v[k] = b[(size_t)k * S];
    return m;
}

template <class T>
__device__ __forceinline__ void st_link(T* f, uint32_t S, int mu, int p, uint32_t i, const M3<T>& m) {
    T* b = f + (size_t)((mu * 2 + p) * 18) * S + i;
#pragma unroll
    for (int k = 0; k < 18; ++k) b[(size_t)k * S] = m.v[k];
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
    return A;
}

}  // namespace qcb
