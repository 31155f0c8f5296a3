// chain_states.cpp -- event-chain state transfer and files (chain_states.hpp).
#include "chain_states.hpp"

#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "qcb_mirror.hpp"

namespace qcb_drv {

namespace {

// The fields of a chain in qcb_ecmc_get_chains / set_chains order
// (QCB_ECMC_CHAIN_DOUBLES per chain): R[18], then these, in order.
void pack(const ChainState& s, double* v) {
    for (int k = 0; k < 18; ++k) v[k] = s.R[k];
    const double rest[14] = {s.th_refresh_site, s.th_refresh_R, s.par_refresh_site, s.par_refresh_R,
                             (double)s.events,  (double)s.lifts, (double)s.revs,    (double)s.set_counter,
                             (double)s.set_gen, (double)s.rng_n, (double)s.site,    (double)s.mu,
                             (double)s.eps,     (double)s.initialized};
    for (int k = 0; k < 14; ++k) v[18 + k] = rest[k];
}

void unpack(const double* v, ChainState& s) {
    for (int k = 0; k < 18; ++k) s.R[k] = v[k];
    s.th_refresh_site = v[18];
    s.th_refresh_R = v[19];
    s.par_refresh_site = v[20];
    s.par_refresh_R = v[21];
    s.events = (uint64_t)v[22];
    s.lifts = (uint64_t)v[23];
    s.revs = (uint64_t)v[24];
    s.set_counter = (uint64_t)v[25];
    s.set_gen = (uint64_t)v[26];
    s.rng_n = (uint64_t)v[27];
    s.site = (uint64_t)v[28];
    s.mu = (int)v[29];
    s.eps = (int)v[30];
    s.initialized = (int)v[31];
}

// Block coordinates of reference rank r (t fastest) -> device chain index (x fastest).
int chain_of_rank(int r, int n) {
    int b[4];
    for (int d = 3; d >= 0; --d, r /= n) b[d] = r % n;   // rank order: t fastest
    return ((b[3] * n + b[2]) * n + b[1]) * n + b[0];    // chain order: x fastest
}

[[noreturn]] void bad_file(const std::filesystem::path& p, const char* why) {
    throw std::runtime_error(std::string(why) + " in " + p.string());
}

}  // namespace

std::vector<ChainState> read_chains(qcb_ecmc* e, int n) {
    std::vector<double> buf((size_t)n * QCB_ECMC_CHAIN_DOUBLES);
    qcb_mirror::check(qcb_ecmc_get_chains(e, buf.data()), "qcb_ecmc_get_chains");
    std::vector<ChainState> st((size_t)n);
    for (int c = 0; c < n; ++c) unpack(buf.data() + (size_t)c * QCB_ECMC_CHAIN_DOUBLES, st[c]);
    return st;
}

void write_chains(qcb_ecmc* e, const std::vector<ChainState>& st) {
    std::vector<double> buf(st.size() * QCB_ECMC_CHAIN_DOUBLES);
    for (size_t c = 0; c < st.size(); ++c) pack(st[c], buf.data() + c * QCB_ECMC_CHAIN_DOUBLES);
    qcb_mirror::check(qcb_ecmc_set_chains(e, buf.data()), "qcb_ecmc_set_chains");
}

void save_chain_files(const RunFiles& files, const std::vector<ChainState>& st, int L_core, int n) {
    const uint64_t L = (uint64_t)L_core * n, Le = (uint64_t)L_core + 2;
    std::error_code ec;
    std::filesystem::create_directories(files.chain_state_folder(), ec);
    if (ec) {
        std::cerr << "Cannot create the folder " << files.chain_state_folder() << ": " << ec.message() << "\n";
        return;
    }
    for (size_t r = 0; r < st.size(); ++r) {
        const ChainState& s = st[chain_of_rank((int)r, n)];
        // global site -> the rank-local padded index
        uint64_t idx = 0, rest = s.site;
        uint64_t local[4];
        for (int d = 0; d < 4; ++d, rest /= L) local[d] = rest % L % (uint64_t)L_core + 1;
        for (int d = 3; d >= 0; --d) idx = idx * Le + local[d];
        std::ostringstream o;
        o << std::setprecision(17) << idx << " " << s.mu << " " << s.eps << " " << s.par_refresh_site << " "
          << s.par_refresh_R << " " << s.th_refresh_site << " " << s.th_refresh_R << " " << s.set_counter << " "
          << s.events << " " << s.lifts << " " << s.revs << " " << s.initialized << "\n";
        for (int row = 0; row < 3; ++row) {
            for (int k = 6 * row; k < 6 * row + 6; ++k) o << s.R[k] << " ";
            o << "\n";
        }
        o << s.rng_n << " " << s.set_gen << "\n";
        std::ofstream f(files.chain_state(r));
        if (!(f << o.str())) {
            std::cerr << "Error: Could not open state file for writing: " << files.chain_state(r) << "\n";
            return;
        }
    }
    std::cout << "ECMC Chain saved in : " << files.chain_state_folder() << "\n";
}

bool load_chain_files(const RunFiles& files, std::vector<ChainState>& st, int L_core, int n) {
    const uint64_t L = (uint64_t)L_core * n, Le = (uint64_t)L_core + 2;
    if (!std::filesystem::exists(files.chain_state(0))) return false;
    for (size_t r = 0; r < st.size(); ++r) {
        const std::filesystem::path p = files.chain_state(r);
        std::ifstream f(p);
        if (!f) bad_file(p, "missing state file");
        std::string head;
        std::getline(f, head);
        std::vector<std::string> tok;
        {
            std::istringstream ts(head);
            for (std::string t; ts >> t;) tok.push_back(t);
        }
        const int c = chain_of_rank((int)r, n);
        ChainState& s = st[c];
        std::istringstream h(head);
        uint64_t idx = 0;
        h >> idx >> s.mu >> s.eps >> s.par_refresh_site >> s.par_refresh_R >> s.th_refresh_site >> s.th_refresh_R;
        if (tok.size() == 12) {
            h >> s.set_counter >> s.events >> s.lifts >> s.revs >> s.initialized;
        } else if (tok.size() == 11) {   // the reference's writer: set_counter and events run together
            uint64_t joined = 0;
            h >> joined >> s.lifts >> s.revs >> s.initialized;
            s.set_counter = s.events = 0;
        } else {
            bad_file(p, "malformed state header");
        }
        if (!h) bad_file(p, "malformed state header");
        for (double& x : s.R) f >> x;
        if (!f) bad_file(p, "malformed matrix R");
        if (!(f >> s.rng_n >> s.set_gen)) s.rng_n = s.set_gen = 0;
        // padded local index -> global site of chain c's block
        uint64_t site = 0, stride = 1;
        for (int d = 0, b = c; d < 4; ++d, b /= n, idx /= Le, stride *= L) {
            const int64_t local = (int64_t)(idx % Le) - 1;
            if (local < 0 || local >= L_core) bad_file(p, "site outside its block");
            site += ((uint64_t)(b % n) * (uint64_t)L_core + (uint64_t)local) * stride;
        }
        s.site = site;
    }
    std::cout << "Successfully loaded ECMC chain state from: " << files.chain_state_folder() << "\n";
    return true;
}

}  // namespace qcb_drv
