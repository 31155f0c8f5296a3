// run_config.hpp -- the input files of the two drivers and the text they echo.
//
// Input format (the reference's, src/io/io.cpp:298-399): one "key = value"
// per line, blanks and tabs around key and value stripped, lines that start
// with '#' or hold no '=' skipped, unknown keys ignored, the last occurrence
// of a key wins.  Integers and reals convert as std::stoi / std::stod do
// (leading number, rest ignored); a flag is true only for the literal "true".
// The defaults are the reference's (src/io/params.h:9-56).
#pragma once
#include <map>
#include <ostream>
#include <stdexcept>
#include <string>

namespace qcb_drv {

struct InputError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class KeyValues {
public:
    // throws InputError when the file cannot be read
    static KeyValues read(const std::string& path);
    void take(const char* key, int& v) const;
    void take(const char* key, double& v) const;
    void take(const char* key, std::string& v) const;
    void take_flag(const char* key, bool& v) const;
    bool has(const char* key) const { return kv_.count(key) > 0; }
    const std::string& at(const char* key) const { return kv_.at(key); }

private:
    std::map<std::string, std::string> kv_;
};

// What every run has: the lattice, the shift cadence and the output folder.
struct LatticeInput {
    int L_core = 6;        // one block's extent (the reference's per-rank lattice)
    int n_core_dims = 2;   // blocks per direction
    bool cold_start = true;
    int seed = 123;
    int L() const { return L_core * n_core_dims; }
};

struct Cadence {
    int therm = 100;           // N_shift_therm
    int shifts = 2;            // N_shift
    int plaquette_every = 2;   // N_shift_plaquette
    bool topo = true;
    int topo_every = 10;       // N_shift_topo
    int gf_steps = 10;         // N_steps_gf: flow measurements per topology sample
    int rk_steps = 40;         // N_rk_steps: RK3 steps between two of them
    int save_every = 2;        // save_each_shifts
};

struct OutputInput {
    std::string run_name = "c";
    std::string run_dir = "data";
};

struct HeatbathInput {
    LatticeInput lat;
    Cadence cad;
    OutputInput out;
    double beta = 6.0;
    int n_sweeps = 1;
    int n_hits = 1;
    // keys of this driver only (the reference ignores them): link precision
    // and the sweeps between reunitarisations (-1: 4 for fp32, never for fp64)
    std::string precision = "fp64";
    int reunit_every = -1;
};

struct ChainInput {
    LatticeInput lat;
    Cadence cad;
    OutputInput out;
    double beta = 6.0;
    double theta_sample = 100;
    double theta_refresh_site = 50;
    double theta_refresh_R = 15;
    bool poisson = false;
    double epsilon_set = 0.15;
};

// The reference keeps two run parameters that its parsers never read: the
// samples per shift are forced to one and N_switch_eo keeps its default.
constexpr int kSamplesPerShift = 1;
constexpr int kSwitchEoDefault = 3;

HeatbathInput parse_heatbath(const std::string& path);   // throws InputError / std::logic_error
ChainInput parse_chain(const std::string& path);

// The parameter summary printed at the start of a run.
void print_summary(std::ostream& os, const HeatbathInput& in);
void print_summary(std::ostream& os, const ChainInput& in);
// The block appended to run_name_params.txt at every start.
std::string params_log(const HeatbathInput& in);
std::string params_log(const ChainInput& in);

}  // namespace qcb_drv
