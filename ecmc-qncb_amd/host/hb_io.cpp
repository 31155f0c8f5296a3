// hb_io.cpp -- parameter parser and output files of the heatbath driver in the
// reference formats (src/io/io.cpp:20-138,233-287,290-450,725-772).
#include "hb_io.hpp"

#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace fs = std::filesystem;

namespace qcb_host {
namespace io {

namespace {
// Opens run_dir/run_name/<suffix> for appending, creating the folder.
bool open_append(const std::string& filename, const std::string& dirpath, const std::string& suffix,
                 std::ofstream& file, fs::path& filepath) {
    fs::path dir = fs::path(dirpath) / filename;
    try {
        if (!fs::exists(dir)) fs::create_directories(dir);
    } catch (const fs::filesystem_error& e) {
        std::cerr << "Couldn't create folder data : " << e.what() << std::endl;
        return false;
    }
    filepath = dir / (filename + suffix);
    file.open(filepath, std::ios::out | std::ios::app);
    if (!file.is_open()) {
        std::cout << "Could not open file " << filepath << "\n";
        return false;
    }
    return true;
}
}  // namespace

std::string trim(const std::string& s) {
    size_t first = s.find_first_not_of(" \t");
    if (first == std::string::npos) return "";
    size_t last = s.find_last_not_of(" \t");
    return s.substr(first, last - first + 1);
}

void load_params(const std::string& filename, RunParamsHbCB& rp) {
    std::ifstream file(filename);
    if (!file.is_open()) throw std::runtime_error("Can't open file " + filename);
    std::map<std::string, std::string> cfg;
    std::string line;
    while (std::getline(file, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        std::string key, value;
        if (std::getline(is, key, '=') && std::getline(is, value)) cfg[trim(key)] = trim(value);
    }
    auto has = [&](const char* k) { return cfg.count(k) > 0; };
    if (has("L_core")) rp.L_core = std::stoi(cfg["L_core"]);
    if (has("n_core_dims")) rp.n_core_dims = std::stoi(cfg["n_core_dims"]);
    if (has("cold_start")) rp.cold_start = (cfg["cold_start"] == "true");
    if (has("seed")) rp.seed = std::stoi(cfg["seed"]);
    if (has("N_shift")) rp.N_shift = std::stoi(cfg["N_shift"]);
    if (has("beta")) rp.hp.beta = std::stod(cfg["beta"]);
    rp.hp.N_samples = 1;   // one sample per shift (src/io/io.cpp:380-381)
    if (has("N_sweeps")) rp.hp.N_sweeps = std::stoi(cfg["N_sweeps"]);
    if (has("N_hits")) rp.hp.N_hits = std::stoi(cfg["N_hits"]);
    if (has("N_shift_plaquette")) rp.N_shift_plaquette = std::stoi(cfg["N_shift_plaquette"]);
    if (has("N_shift_therm")) rp.N_therm = std::stoi(cfg["N_shift_therm"]);
    if (has("topo")) rp.topo = (cfg["topo"] == "true");
    if (has("N_shift_topo")) rp.N_shift_topo = std::stoi(cfg["N_shift_topo"]);
    if (has("N_steps_gf")) rp.N_steps_gf = std::stoi(cfg["N_steps_gf"]);
    if (has("N_rk_steps")) rp.N_rk_steps = std::stoi(cfg["N_rk_steps"]);
    if (has("run_name")) rp.run_name = cfg["run_name"];
    if (has("run_dir")) rp.run_dir = cfg["run_dir"];
    if (has("save_each_shifts")) rp.save_each_shifts = std::stoi(cfg["save_each_shifts"]);
    // keys of this driver only: checked strictly, since the reference never reads them
    if (has("precision")) {
        rp.precision = cfg["precision"];
        if (rp.precision != "fp64" && rp.precision != "fp32")
            throw std::runtime_error("precision must be fp64 or fp32, not '" + rp.precision + "'");
    }
    if (has("reunit_every")) {
        size_t used = 0;
        rp.reunit_every = std::stoi(cfg["reunit_every"], &used);
        if (used != cfg["reunit_every"].size() || rp.reunit_every < 0)
            throw std::runtime_error("reunit_every must be a count >= 0");
    }
}

bool read_params(RunParamsHbCB& params, const std::string& input) {
    try {
        load_params(input, params);
    } catch (const std::exception& e) {
        std::cerr << "Error reading input : " << e.what() << std::endl;
        std::exit(1);
    }
    fs::path base = fs::path(params.run_dir) / params.run_name;
    fs::path config_file = base / params.run_name;
    fs::path seed_file = base / (params.run_name + "_seed") / (params.run_name + "_seed0.txt");
    bool existing = fs::exists(config_file) && fs::exists(seed_file);
    if (existing) std::cout << "All ranks found existing configuration. Resuming...\n";
    return existing;
}

void save_params(const RunParamsHbCB& rp, const std::string& filename, const std::string& dirpath) {
    std::ofstream file;
    fs::path filepath;
    if (!open_append(filename, dirpath, "_params.txt", file, filepath)) return;
    file << std::boolalpha;
    file << "\n#########################################################\n\n";
    file << "# Lattice params\n";
    file << "L_core=" << rp.L_core << "\n";
    file << "n_core_dims=" << rp.n_core_dims << "\n";
    file << "cold_start=" << rp.cold_start << "\n\n";
    file << "# Run params\n";
    file << "seed=" << rp.seed << "\n";
    file << "N_shift_therm = " << rp.N_therm << "\n";
    file << "N_shift =" << rp.N_shift << "\n";
    file << "# Heatbath params\n";
    file << "beta = " << rp.hp.beta << "\n";
    file << "N_sweeps = " << rp.hp.N_sweeps << "\n";
    file << "N_hits = " << rp.hp.N_hits << "\n\n";
    file << "# Plaquette params\n";
    file << "N_shift_plaquette = " << rp.N_shift_plaquette << "\n\n";
    file << "# Topo params\n";
    file << "topo = " << rp.topo << "\n";
    file << "N_shift_topo = " << rp.N_shift_topo << "\n";
    file << "N_steps_gf = " << rp.N_steps_gf << "\n";
    file << "N_rk_steps = " << rp.N_rk_steps << "\n\n";
    file << "#Save params\n";
    file << "save_each_shifts = " << rp.save_each_shifts << "\n\n";
    file.close();
    std::cout << "Parameters saved in " << filepath << "\n";
}

void save_plaquette(const std::vector<double>& data, const std::string& filename, const std::string& dirpath,
                    int precision) {
    std::ofstream file;
    fs::path filepath;
    if (!open_append(filename, dirpath, "_plaquette.txt", file, filepath)) return;
    file << std::fixed << std::setprecision(precision);
    for (double x : data) file << x << "\n";
    file.close();
    std::cout << "Plaquette written in " << filepath << "\n";
}

void save_topo(const std::vector<double>& tQE, const std::string& filename, const std::string& dirpath,
               int precision) {
    std::ofstream file;
    fs::path filepath;
    if (!open_append(filename, dirpath, "_topo.txt", file, filepath)) return;
    file << std::fixed << std::setprecision(precision);
    for (size_t i = 0; i + 2 < tQE.size(); i += 3) file << tQE[i] << " " << tQE[i + 1] << " " << tQE[i + 2] << "\n";
    file.close();
    std::cout << "Topology written in " << filepath << "\n";
}

void save_seed(std::mt19937_64& rng, const std::string& filename, const std::string& dirpath, int rank) {
    fs::path run_dir = fs::path(dirpath) / filename / (filename + "_seed");
    try {
        if (!fs::exists(run_dir)) fs::create_directories(run_dir);
    } catch (const fs::filesystem_error& e) {
        std::cerr << "Error creating directory structure " << run_dir << " : " << e.what() << std::endl;
        return;
    }
    fs::path filepath = run_dir / (filename + "_seed" + std::to_string(rank) + ".txt");
    std::ofstream file(filepath);
    if (!file.is_open()) {
        std::cout << "Could not open file " << filepath << "\n";
        return;
    }
    file << rng;
    file.close();
    if (rank == 0) std::cout << "Seed saved in " << filepath << "\n";
}

void add_shift(int shift, const std::string& filename, const std::string& dirpath) {
    std::ofstream file;
    fs::path filepath;
    if (!open_append(filename, dirpath, "_params.txt", file, filepath)) return;
    file << "Saved shift " << shift << "\n";
}

void add_finished(const std::string& filename, const std::string& dirpath) {
    std::ofstream file;
    fs::path filepath;
    if (!open_append(filename, dirpath, "_params.txt", file, filepath)) return;
    file << "Saved final state !\n";
}

std::string format_double(double val, int precision) {
    std::stringstream ss;
    ss << std::fixed << std::setprecision(precision) << val;
    return ss.str();
}

}  // namespace io

void print_parameters(const RunParamsHbCB& rp) {
    std::cout << "==========================================" << std::endl;
    std::cout << "Heatbath - MPI No Checkboard" << std::endl;
    std::cout << "==========================================" << std::endl;
    std::cout << "Run name : " + rp.run_name << "\n";
    std::cout << "---Lattice params---\n";
    std::cout << "Total lattice size : " << rp.L_core * rp.n_core_dims << "^4\n";
    std::cout << "Local lattice size : " << rp.L_core << "^4\n\n";
    std::cout << "---Run params---\n";
    std::cout << "Initial seed : " << rp.seed << "\n";
    std::cout << "Thermalization shifts : " << rp.N_therm << "\n";
    std::cout << "Number of shifts : " << rp.N_shift << "\n";
    std::cout << "Save each : " << rp.save_each_shifts << " shifts\n\n";
    std::cout << "---Heatbath params---\n";
    std::cout << "Beta : " << rp.hp.beta << "\n";
    std::cout << "Number of sweeps : " << rp.hp.N_sweeps << "\n";
    std::cout << "Number of hits : " << rp.hp.N_hits << "\n\n";
    std::cout << "---Measures params---\n";
    std::cout << "Number of <P> samples : " << rp.N_shift / rp.N_shift_plaquette << "\n";
    std::cout << "Measure <P> each : " << rp.N_shift_plaquette << " shifts\n";
    std::cout << "Measure topo : " << (rp.topo ? "Yes" : "No") << "\n";
    if (rp.topo) {
        std::cout << "Number of Q samples : " << rp.N_shift / rp.N_shift_topo << "\n";
        std::cout << "Measure Q each : " << rp.N_shift_topo << " shifts\n\n";
        std::cout << "---Gradient Flow---\n";
        std::cout << "Number of RK3 steps per GF step : " << rp.N_rk_steps << "\n";
        std::cout << "Number of GF steps : " << rp.N_steps_gf << "\n";
    }
    std::cout << "==========================================" << std::endl;
}

void print_time(long elapsed) {
    std::cout << "==========================================" << std::endl;
    std::cout << "Elapsed time : " << elapsed << "s\n";
    std::cout << "==========================================" << std::endl;
}

}  // namespace qcb_host
