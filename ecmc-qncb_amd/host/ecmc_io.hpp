// ecmc_io.hpp -- run parameters, input parser and output files of the event-chain
// driver, in the reference's formats (src/io/params.h:9-44, src/io/io.cpp:49-76,
// 160-232, 298-349, 538-620).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace qcb_host {

// ECMCParams / RunParamsECB with the reference defaults (src/io/params.h:9-44).
struct ECMCParams {
    double beta = 6.0;
    int N_samples = 10;
    double param_theta_sample = 100;
    double param_theta_refresh_site = 50;
    double param_theta_refresh_R = 15;
    bool poisson = false;
    double epsilon_set = 0.15;
};

struct RunParamsECB {
    int L_core = 6;
    int n_core_dims = 2;
    bool cold_start = true;
    int seed = 123;
    int N_switch_eo = 3;   // declared, never parsed
    int N_shift = 2;
    ECMCParams ecmc_params{};
    int N_therm = 100;
    bool topo = true;
    int N_shift_plaquette = 2;
    int N_shift_topo = 10;
    int N_steps_gf = 10;
    int N_rk_steps = 40;
    std::string run_name = "c";
    std::string run_dir = "data";
    int save_each_shifts = 2;
};

// One chain's LocalChainState (src/ecmc/ecmc_mpi_cb.h:15-35) plus the Philox
// position of its device stream.
struct ChainState {
    double R[18] = {};
    double th_refresh_site = 0, th_refresh_R = 0, par_refresh_site = 0, par_refresh_R = 0;
    uint64_t events = 0, lifts = 0, revs = 0, set_counter = 0, set_gen = 0, rng_n = 0;
    uint64_t site = 0;   // global lex site
    int mu = 0, eps = 0, initialized = 0;
};

namespace io {
// key=value parser (src/io/io.cpp:298-349); throws std::runtime_error if the file cannot be opened
void load_params(const std::string& filename, RunParamsECB& rp);
// parse + resume detection (configuration and seed file present) for one process
bool read_params(RunParamsECB& params, const std::string& input);
void save_params(const RunParamsECB& rp, const std::string& filename, const std::string& dirpath);
// run_name_nbevent.txt lines "events lifts reversals lambda" (src/io/io.cpp:49-76)
void save_event_nb(const std::vector<size_t>& event_nb, const std::vector<size_t>& lift_nb,
                   const std::vector<size_t>& rev_nb, const std::vector<double>& lambda, const std::string& filename,
                   const std::string& dirpath);
// run_name_state/run_name_state<c>.txt (src/io/io.cpp:160-232), one file per chain.
// site is written as the reference's per-rank index ((t*Le + z)*Le + y)*Le + x with
// Le = L_core + 2 and the block's sites at 1..L_core; the Philox position follows
// the matrix R on a fifth line, which the reference's reader does not look at.
void save_states(const std::vector<ChainState>& st, int L_core, int n_core_dims, const std::string& filename,
                 const std::string& dirpath);
// false when the files are absent (a new run); throws on a malformed file.
// A file written by the reference (set_counter and events glued into one
// token, 11 fields on line 1) loads with both counters zeroed.
bool load_states(std::vector<ChainState>& st, int L_core, int n_core_dims, const std::string& filename,
                 const std::string& dirpath);
}  // namespace io

void print_parameters(const RunParamsECB& rp);

}  // namespace qcb_host
