// ecmc_io.cpp -- parameters and files of the event-chain driver (see ecmc_io.hpp).
#include "ecmc_io.hpp"

#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "hb_io.hpp"

namespace fs = std::filesystem;

namespace qcb_host {
namespace io {

namespace {
bool open_dir_file(const std::string& filename, const std::string& dirpath, const std::string& suffix,
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

fs::path state_path(const std::string& filename, const std::string& dirpath, int r) {
    return fs::path(dirpath) / filename / (filename + "_state") / (filename + "_state" + std::to_string(r) + ".txt");
}

// Rank r of the reference's Cartesian grid has block coordinates in row-major
// order, t fastest (MPI_Cart_create, src/mpi/MpiTopology.cpp:13-18); the device
// numbers its chains with x fastest.
int chain_of_rank(int r, int n) {
    int b[4];
    for (int d = 3; d >= 0; --d) {
        b[d] = r % n;
        r /= n;
    }
    return ((b[3] * n + b[2]) * n + b[1]) * n + b[0];
}
}  // namespace

void load_params(const std::string& filename, RunParamsECB& rp) {
    std::ifstream file(filename);
    if (!file.is_open()) throw std::runtime_error("Impossible d'ouvrir " + filename);
    std::map<std::string, std::string> config;
    std::string line;
    while (std::getline(file, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is_line(line);
        std::string key, value;
        if (std::getline(is_line, key, '=') && std::getline(is_line, value)) config[trim(key)] = trim(value);
    }
    auto has = [&](const char* k) { return config.count(k) > 0; };
    if (has("L_core")) rp.L_core = std::stoi(config["L_core"]);
    if (has("n_core_dims")) rp.n_core_dims = std::stoi(config["n_core_dims"]);
    if (has("cold_start")) rp.cold_start = (config["cold_start"] == "true");
    if (has("seed")) rp.seed = std::stoi(config["seed"]);
    if (has("N_shift")) rp.N_shift = std::stoi(config["N_shift"]);
    if (has("beta")) rp.ecmc_params.beta = std::stod(config["beta"]);
    rp.ecmc_params.N_samples = 1;
    if (has("param_theta_sample")) rp.ecmc_params.param_theta_sample = std::stod(config["param_theta_sample"]);
    if (has("param_theta_refresh_site"))
        rp.ecmc_params.param_theta_refresh_site = std::stod(config["param_theta_refresh_site"]);
    if (has("param_theta_refresh_R")) rp.ecmc_params.param_theta_refresh_R = std::stod(config["param_theta_refresh_R"]);
    if (has("poisson")) rp.ecmc_params.poisson = (config["poisson"] == "true");
    if (has("epsilon_set")) rp.ecmc_params.epsilon_set = std::stod(config["epsilon_set"]);
    if (has("N_shift_plaquette")) rp.N_shift_plaquette = std::stoi(config["N_shift_plaquette"]);
    if (has("N_shift_therm")) rp.N_therm = std::stoi(config["N_shift_therm"]);
    if (has("topo")) rp.topo = (config["topo"] == "true");
    if (has("N_shift_topo")) rp.N_shift_topo = std::stoi(config["N_shift_topo"]);
    if (has("N_steps_gf")) rp.N_steps_gf = std::stoi(config["N_steps_gf"]);
    if (has("N_rk_steps")) rp.N_rk_steps = std::stoi(config["N_rk_steps"]);
    if (has("run_name")) rp.run_name = config["run_name"];
    if (has("run_dir")) rp.run_dir = config["run_dir"];
    if (has("save_each_shifts")) rp.save_each_shifts = std::stoi(config["save_each_shifts"]);
}

bool read_params(RunParamsECB& params, const std::string& input) {
    try {
        load_params(input, params);
    } catch (const std::exception& e) {
        std::cerr << "Error reading input : " << e.what() << std::endl;
        std::exit(1);
    }
    fs::path base = fs::path(params.run_dir) / params.run_name;
    bool existing = fs::exists(base / params.run_name) &&
                    fs::exists(base / (params.run_name + "_seed") / (params.run_name + "_seed0.txt"));
    if (existing) std::cout << "All ranks found existing configuration. Resuming...\n";
    return existing;
}

void save_params(const RunParamsECB& rp, const std::string& filename, const std::string& dirpath) {
    std::ofstream file;
    fs::path filepath;
    if (!open_dir_file(filename, dirpath, "_params.txt", file, filepath)) return;
    file << std::boolalpha;
    file << "\n#########################################################\n\n";
    file << "# Lattice params\n";
    file << "L_core = " << rp.L_core << "\n";
    file << "n_core_dims = " << rp.n_core_dims << "\n";
    file << "cold_start = " << rp.cold_start << "\n\n";
    file << "# Run params\n";
    file << "seed = " << rp.seed << "\n";
    file << "N_shift_therm = " << rp.N_therm << "\n";
    file << "N_shift = " << rp.N_shift << "\n";
    file << "# ECMC params\n";
    file << "beta = " << rp.ecmc_params.beta << "\n";
    file << "theta_sample = " << rp.ecmc_params.param_theta_sample << "\n";
    file << "theta_refresh_site = " << rp.ecmc_params.param_theta_refresh_site << "\n";
    file << "theta_refresh_R = " << rp.ecmc_params.param_theta_refresh_R << "\n";
    file << "epsilon_set = " << rp.ecmc_params.epsilon_set << "\n";
    file << "poisson = " << rp.ecmc_params.poisson << "\n\n";
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

void save_event_nb(const std::vector<size_t>& event_nb, const std::vector<size_t>& lift_nb,
                   const std::vector<size_t>& rev_nb, const std::vector<double>& lambda, const std::string& filename,
                   const std::string& dirpath) {
    std::ofstream file;
    fs::path filepath;
    if (!open_dir_file(filename, dirpath, "_nbevent.txt", file, filepath)) return;
    for (size_t i = 0; i < event_nb.size(); i++)
        file << event_nb[i] << " " << lift_nb[i] << " " << rev_nb[i] << " " << lambda[i] << "\n";
    file.close();
    std::cout << "Number of events written in " << filepath << "\n";
}

void save_states(const std::vector<ChainState>& st, int L_core, int n_core_dims, const std::string& filename,
                 const std::string& dirpath) {
    const int L = L_core * n_core_dims, Le = L_core + 2;
    fs::path dir = fs::path(dirpath) / filename / (filename + "_state");
    try {
        if (!fs::exists(dir)) fs::create_directories(dir);
    } catch (const fs::filesystem_error& e) {
        std::cerr << "Couldn't create folder " << dir << " : " << e.what() << std::endl;
        return;
    }
    for (size_t r = 0; r < st.size(); ++r) {
        const ChainState& s = st[chain_of_rank((int)r, n_core_dims)];
        std::ofstream ofs(state_path(filename, dirpath, (int)r));
        if (!ofs.is_open()) {
            std::cerr << "Error: Could not open state file for writing: " << state_path(filename, dirpath, (int)r)
                      << std::endl;
            return;
        }
        // global lex site -> the reference's padded per-rank index
        uint64_t x = s.site;
        int loc[4];
        for (int d = 0; d < 4; ++d) {
            loc[d] = (int)(x % (uint64_t)L) % L_core + 1;
            x /= (uint64_t)L;
        }
        const uint64_t idx = (((uint64_t)loc[3] * Le + loc[2]) * Le + loc[1]) * Le + loc[0];
        ofs << std::setprecision(17);
        ofs << idx << " " << s.mu << " " << s.eps << " " << s.par_refresh_site << " " << s.par_refresh_R << " "
            << s.th_refresh_site << " " << s.th_refresh_R << " " << s.set_counter << " " << s.events << " " << s.lifts
            << " " << s.revs << " " << s.initialized << "\n";
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) ofs << s.R[(3 * i + j) * 2] << " " << s.R[(3 * i + j) * 2 + 1] << " ";
            ofs << "\n";
        }
        ofs << s.rng_n << " " << s.set_gen << "\n";
    }
    std::cout << "ECMC Chain saved in : " << dir << std::endl;
}

bool load_states(std::vector<ChainState>& st, int L_core, int n_core_dims, const std::string& filename,
                 const std::string& dirpath) {
    const int L = L_core * n_core_dims, Le = L_core + 2;
    if (!fs::exists(state_path(filename, dirpath, 0))) return false;
    for (size_t r = 0; r < st.size(); ++r) {
        std::ifstream ifs(state_path(filename, dirpath, (int)r));
        if (!ifs.is_open()) throw std::runtime_error("missing " + state_path(filename, dirpath, (int)r).string());
        const int c = chain_of_rank((int)r, n_core_dims);
        ChainState& s = st[c];
        uint64_t idx;
        std::string head;
        std::getline(ifs, head);
        std::istringstream hs(head);
        std::vector<std::string> tok;
        for (std::string t; hs >> t;) tok.push_back(t);
        std::istringstream fs_(head);
        fs_ >> idx >> s.mu >> s.eps >> s.par_refresh_site >> s.par_refresh_R >> s.th_refresh_site >> s.th_refresh_R;
        if (tok.size() == 12) {
            fs_ >> s.set_counter >> s.events >> s.lifts >> s.revs >> s.initialized;
        } else if (tok.size() == 11) {
            // written by the reference: set_counter and events glued into one
            // token (src/io/io.cpp:172-175); both restart from zero
            uint64_t glued;
            fs_ >> glued >> s.lifts >> s.revs >> s.initialized;
            s.set_counter = s.events = 0;
        } else {
            fs_.setstate(std::ios::failbit);
        }
        if (!fs_) throw std::runtime_error("malformed " + state_path(filename, dirpath, (int)r).string());
        for (int k = 0; k < 18; ++k) ifs >> s.R[k];
        if (!ifs) throw std::runtime_error("malformed " + state_path(filename, dirpath, (int)r).string());
        if (!(ifs >> s.rng_n >> s.set_gen)) s.rng_n = s.set_gen = 0;
        // padded per-rank index -> global lex site of block c
        int loc[4];
        uint64_t q = idx;
        for (int d = 0; d < 4; ++d) {
            loc[d] = (int)(q % (uint64_t)Le) - 1;
            q /= (uint64_t)Le;
        }
        int b = c;
        uint64_t site = 0, stride = 1;
        for (int d = 0; d < 4; ++d) {
            const int bc = b % n_core_dims;
            b /= n_core_dims;
            if (loc[d] < 0 || loc[d] >= L_core)
                throw std::runtime_error("site outside its block in " + state_path(filename, dirpath, (int)r).string());
            site += (uint64_t)(bc * L_core + loc[d]) * stride;
            stride *= (uint64_t)L;
        }
        s.site = site;
    }
    std::cout << "Successfully loaded ECMC chain state from: " << state_path(filename, dirpath, 0).parent_path()
              << std::endl;
    return true;
}

}  // namespace io

void print_parameters(const RunParamsECB& rp) {
    std::cout << "==========================================" << std::endl;
    std::cout << "ECMC - MPI No Checkboard" << std::endl;
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
    std::cout << "---ECMC params---\n";
    std::cout << "Beta : " << rp.ecmc_params.beta << "\n";
    std::cout << "Theta sample : " << rp.ecmc_params.param_theta_sample << "\n";
    std::cout << "Theta refresh site : " << rp.ecmc_params.param_theta_refresh_site << "\n";
    std::cout << "Theta refresh R : " << rp.ecmc_params.param_theta_refresh_R << "\n";
    std::cout << "Epsilon set : " << rp.ecmc_params.epsilon_set << "\n";
    std::cout << "Poisson law : " << (rp.ecmc_params.poisson ? "Yes" : "No") << "\n\n";
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

}  // namespace qcb_host
