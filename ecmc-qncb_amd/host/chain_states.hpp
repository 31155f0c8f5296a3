// chain_states.hpp -- event-chain states between the device and the state files.
//
// One file per reference rank r, run_name_state/run_name_state<r>.txt, in the
// layout of the reference's save_state / load_state (src/io/io.cpp:160-232):
//   line 1: site mu epsilon par_refresh_site par_refresh_R th_refresh_site
//           th_refresh_R set_counter events lifts reversals initialized
//   lines 2-4: R, one row per line, re im pairs
// where site is the rank-local padded index ((t*Le + z)*Le + y)*Le + x with
// Le = L_core + 2 and the block's sites at 1..L_core.  A fifth line holds the
// chain's Philox position (rng_n set_gen), which the reference does not read.
// Rank r owns the block MPI_Cart_create gives it (row-major, t fastest,
// src/mpi/MpiTopology.cpp:13-18); the device numbers its chains x fastest.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/qcb.h"
#include "run_files.hpp"

namespace qcb_drv {

// LocalChainState (src/ecmc/ecmc_mpi_cb.h:15-35) plus the Philox position.
struct ChainState {
    double R[18] = {};
    double th_refresh_site = 0, th_refresh_R = 0, par_refresh_site = 0, par_refresh_R = 0;
    uint64_t events = 0, lifts = 0, revs = 0, set_counter = 0, set_gen = 0, rng_n = 0;
    uint64_t site = 0;   // global lexicographic site (x fastest)
    int mu = 0, eps = 0, initialized = 0;
};

std::vector<ChainState> read_chains(qcb_ecmc* e, int n);   // qcb_ecmc_get_chains
void write_chains(qcb_ecmc* e, const std::vector<ChainState>& st);

void save_chain_files(const RunFiles& files, const std::vector<ChainState>& st, int L_core, int n_core_dims);
// false when no state files exist (a fresh run); throws std::runtime_error
// on a missing or malformed file.  A file written by the reference (its
// set_counter and events run together, 11 fields on line 1) loads with both
// counters zero.
bool load_chain_files(const RunFiles& files, std::vector<ChainState>& st, int L_core, int n_core_dims);

}  // namespace qcb_drv
