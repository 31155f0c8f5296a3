// run_files.cpp -- the run folder and its files (run_files.hpp).
#include "run_files.hpp"

#include <cstdio>
#include <fstream>
#include <iostream>

namespace fs = std::filesystem;

namespace qcb_drv {

namespace {
// "%.10f" of each value: the reference writes observables with std::fixed and
// ten decimals
std::string fixed10(double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.10f", v);
    return buf;
}
}  // namespace

RunFiles::RunFiles(const std::string& run_dir, const std::string& run_name)
    : folder_(fs::path(run_dir) / run_name), name_(run_name) {}

fs::path RunFiles::generator_state(int rank) const {
    return folder_ / (name_ + "_seed") / (name_ + "_seed" + std::to_string(rank) + ".txt");
}

fs::path RunFiles::chain_state(int rank) const {
    return chain_state_folder() / (name_ + "_state" + std::to_string(rank) + ".txt");
}

bool RunFiles::resumable() const { return fs::exists(configuration()) && fs::exists(generator_state(0)); }

bool RunFiles::make_folder(const fs::path& p) const {
    std::error_code ec;
    fs::create_directories(p, ec);
    if (ec) {
        std::cerr << "Cannot create the folder " << p << ": " << ec.message() << "\n";
        return false;
    }
    return true;
}

bool RunFiles::append(const std::string& suffix, const std::string& text, const char* what) const {
    if (!make_folder(folder_)) return false;
    const fs::path p = folder_ / (name_ + suffix);
    std::ofstream f(p, std::ios::app);
    if (!(f << text)) {
        std::cout << "Could not open file " << p << "\n";
        return false;
    }
    if (what) std::cout << what << " " << p << "\n";
    return true;
}

void RunFiles::append_plaquettes(const std::vector<double>& p) const {
    std::string s;
    for (double v : p) s += fixed10(v) + "\n";
    append("_plaquette.txt", s, "Plaquette written in");
}

void RunFiles::append_flow_series(const std::vector<double>& tQE) const {
    std::string s;
    for (size_t i = 0; i + 2 < tQE.size(); i += 3)
        s += fixed10(tQE[i]) + " " + fixed10(tQE[i + 1]) + " " + fixed10(tQE[i + 2]) + "\n";
    append("_topo.txt", s, "Topology written in");
}

void RunFiles::save_generator(const std::mt19937_64& g, int rank) const {
    const fs::path p = generator_state(rank);
    if (!make_folder(p.parent_path())) return;
    std::ofstream f(p);
    if (!(f << g)) {
        std::cout << "Could not open file " << p << "\n";
        return;
    }
    if (rank == 0) std::cout << "Seed saved in " << p << "\n";
}

bool RunFiles::load_generator(std::mt19937_64& g, int rank) const {
    const fs::path p = generator_state(rank);
    std::ifstream f(p);
    if (!f) {
        std::cerr << "Could not open " << p << "\n";
        return false;
    }
    f >> g;
    return true;
}

}  // namespace qcb_drv
