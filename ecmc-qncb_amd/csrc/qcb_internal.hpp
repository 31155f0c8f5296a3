// qcb_internal.hpp -- host-side launchers shared by the C-ABI implementation
// and the kernel translation units (not part of the public interface).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qcb_layout.hpp"

namespace qcb {

// Work decomposition of one launch over a batch of lattices.  Each kernel
// family has its own per-lattice unit count: cb sites of one parity (Vh),
// sites (2*Vh) or links (8*Vh); blk holds the n_lat+1 prefix sums of
// ceil(units / kBlock) for that family.
struct Launch {
    const void* lats;       // device LatDev<T>[n_lat]
    const uint32_t* blk;    // device prefix sums for the family being launched
    int n_lat;
    uint32_t nblocks;       // blk[n_lat]
    hipStream_t stream;
};

enum HbMode { kHbPhilox = 0, kHbInjected = 1 };

template <class T> void launch_cold(const Launch& g);
// gate: per-lattice cadence (Gate{} = every lattice of the launch)
template <class T> void launch_hot(const Launch& g, Gate gate = Gate{});   // key from descriptor
template <class T> void launch_aos_to_soa(const Launch& g, const double* aos, T* dst_field);
// ILDG staging of links [lo, lo + n) in file order (single-lattice launches)
template <class T> void launch_soa_to_be(const Launch& g, const T* src, uint64_t* out, uint64_t lo, uint32_t n);
template <class T> void launch_be_to_soa(const Launch& g, const uint64_t* in, T* dst, uint64_t lo, uint32_t n);
// order: the storage order of the field (kLinkOrder for U, kFlowOrder for F[0], F[1])
template <class T> void launch_soa_to_aos(const Launch& g, const T* src_field, double* aos, int order = kLinkOrder);
template <class T> void launch_copy_field(const Launch& g, const T* src, T* dst, int from = kLinkOrder,
                                          int to = kLinkOrder);
template <class T> void launch_flow_init(const Launch& g, Gate gate = Gate{});   // F[0] <- U
// which = 0: the links; 1: each lattice's current flow field at flow step gstep
template <class T> void launch_plaquette(const Launch& g, double* partial, int which = 0, uint32_t gstep = 0);
template <class T> void launch_clover(const Launch& g, int which_field, uint32_t gstep, double* partial_q, double* partial_e);
template <class T> void launch_staples(const Launch& g, double* out_aos);
// sweep_base: optional device counter added to sweep_idx (graph replay);
// grid: 0 = one block per tile, else a persistent grid (a multiple of 8).
// Exhausted Kennedy-Pendleton draws are added to each lattice's kp_exh.
template <class T> void launch_hb_half(const Launch& g, int mu, int parity, int n_hits, uint32_t sweep_idx,
                                       int mode, const double* inj, int K, const uint32_t* sweep_base = nullptr,
                                       uint32_t grid = 0);
void launch_bump(uint32_t* counter, uint32_t n, hipStream_t st);
template <class T> void launch_reunit(const Launch& g, Gate gate = Gate{});
template <class T> void launch_translate(const Launch& g, const int s[4], T* dst_field);
// One RK3 stage on every lattice of the batch.  Field roles follow the
// per-lattice step parity (gstep - flow_start) & 1; which_src = 0 reads U for
// stage 1 of a lattice's first step when flowing the links in place.
template <class T> void launch_flow_stage(const Launch& g, int stage, uint32_t gstep, double eps);
template <class T> void launch_math_probe(int op, int n, const double* in, double param, int aux, double* out,
                                          hipStream_t st);

// Fixed-order per-lattice reductions of per-block partials.
// plaq: out[l*ostride] = sum / (6 V);  qe: out[l*ostride] = Q, out[l*ostride+1] = E / V
template <class T> void launch_finalize_plaq(const Launch& g, const double* partial, double* out, int ostride);
// qe: a lattice with qe_out set writes at its own flow step gstep - flow_start instead
template <class T> void launch_finalize_qe(const Launch& g, const double* pq, const double* pe, double* out,
                                           int ostride, uint32_t gstep = 0);

}  // namespace qcb
