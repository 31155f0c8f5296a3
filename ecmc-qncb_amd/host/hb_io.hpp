// hb_io.hpp -- run parameters, input-file parser and observable/log files of
// the heatbath driver, in the reference's formats (src/io/params.h,
// src/io/io.cpp).  File contents and stdout lines are kept byte-compatible so
// that existing analysis scripts read the new driver's outputs unchanged.
#pragma once
#include <random>
#include <string>
#include <vector>

namespace qcb_host {

// HbParams / RunParamsHbCB with the reference defaults (src/io/params.h:17-23,38-56).
struct HbParams {
    double beta = 6.0;
    int N_samples = 10;
    int N_hits = 1;
    int N_sweeps = 1;
    int N_therm = 100;
};

struct RunParamsHbCB {
    int L_core = 6;
    int n_core_dims = 2;
    bool cold_start = true;
    int N_switch_eo = 3;   // declared, never parsed (src/io/io.cpp:353-399)
    int N_shift = 2;
    int seed = 123;
    HbParams hp{};
    int N_therm = 100;
    bool topo = true;
    int N_shift_plaquette = 2;
    int N_shift_topo = 10;
    int N_steps_gf = 10;
    int N_rk_steps = 40;
    std::string run_name = "c";
    std::string run_dir = "data";
    int save_each_shifts = 2;
    // keys of this driver (ignored by the reference parser, so inputs stay
    // interchangeable): link precision and reunitarisation cadence
    std::string precision = "fp64";   // "fp64" | "fp32"
    int reunit_every = -1;            // sweeps between reunitarisations; -1: 4 for fp32, never for fp64
};

namespace io {
std::string trim(const std::string& s);
// key=value parser (src/io/io.cpp:353-399); throws std::runtime_error if the file cannot be opened
void load_params(const std::string& filename, RunParamsHbCB& rp);
// rank-0 parse + resume detection (src/io/io.cpp:454-536) for one process
bool read_params(RunParamsHbCB& params, const std::string& input);
void save_params(const RunParamsHbCB& rp, const std::string& filename, const std::string& dirpath);
void save_plaquette(const std::vector<double>& data, const std::string& filename, const std::string& dirpath,
                    int precision);
void save_topo(const std::vector<double>& tQE, const std::string& filename, const std::string& dirpath, int precision);
void save_seed(std::mt19937_64& rng, const std::string& filename, const std::string& dirpath, int rank);
void add_shift(int shift, const std::string& filename, const std::string& dirpath);
void add_finished(const std::string& filename, const std::string& dirpath);
std::string format_double(double val, int precision);
}  // namespace io

void print_parameters(const RunParamsHbCB& rp);
void print_time(long elapsed);

}  // namespace qcb_host
