// gauge_ecmc_cb.cpp -- the event-chain driver on one MI355X.
//
// Same command line, input keys, stdout lines, output files, ILDG checkpoints,
// chain-state files and resume rule as the reference's ECMC driver
// (src/bin/gauge_ecmc_ncb.cpp:20-300).  The reference runs one chain per MPI
// rank on its L_core^4 block; here the whole (L_core * n_core_dims)^4
// lattice lives on the device and the n_core_dims^4 chains run together
// (qcb_ecmc_*), one per block.  Random shifts keep the host mt19937_64 draws;
// the chains draw from Philox counters keyed by the run's seed, and their
// positions in those streams are saved with the chain states, so a resumed
// run continues every chain's stream.
#include <chrono>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <vector>

#include "ecmc_io.hpp"
#include "hb_io.hpp"
#include "qcb_mirror.hpp"

namespace fs = std::filesystem;
using namespace qcb_mirror;
using qcb_host::ChainState;
using qcb_host::RunParamsECB;

namespace {

// splitmix64 finaliser: the chains' Philox key from the run seed
uint64_t chain_key(int seed) {
    uint64_t z = (uint64_t)(int64_t)seed + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

std::vector<ChainState> get_states(qcb_ecmc* e, int n) {
    std::vector<double> buf((size_t)n * QCB_ECMC_CHAIN_DOUBLES);
    check(qcb_ecmc_get_chains(e, buf.data()), "qcb_ecmc_get_chains");
    std::vector<ChainState> st(n);
    for (int c = 0; c < n; ++c) {
        const double* v = buf.data() + (size_t)c * QCB_ECMC_CHAIN_DOUBLES;
        ChainState& s = st[c];
        for (int k = 0; k < 18; ++k) s.R[k] = v[k];
        s.th_refresh_site = v[18]; s.th_refresh_R = v[19]; s.par_refresh_site = v[20]; s.par_refresh_R = v[21];
        s.events = (uint64_t)v[22]; s.lifts = (uint64_t)v[23]; s.revs = (uint64_t)v[24];
        s.set_counter = (uint64_t)v[25]; s.set_gen = (uint64_t)v[26]; s.rng_n = (uint64_t)v[27];
        s.site = (uint64_t)v[28]; s.mu = (int)v[29]; s.eps = (int)v[30]; s.initialized = (int)v[31];
    }
    return st;
}

void set_states(qcb_ecmc* e, const std::vector<ChainState>& st) {
    std::vector<double> buf(st.size() * QCB_ECMC_CHAIN_DOUBLES);
    for (size_t c = 0; c < st.size(); ++c) {
        double* v = buf.data() + c * QCB_ECMC_CHAIN_DOUBLES;
        const ChainState& s = st[c];
        for (int k = 0; k < 18; ++k) v[k] = s.R[k];
        v[18] = s.th_refresh_site; v[19] = s.th_refresh_R; v[20] = s.par_refresh_site; v[21] = s.par_refresh_R;
        v[22] = (double)s.events; v[23] = (double)s.lifts; v[24] = (double)s.revs;
        v[25] = (double)s.set_counter; v[26] = (double)s.set_gen; v[27] = (double)s.rng_n;
        v[28] = (double)s.site; v[29] = s.mu; v[30] = s.eps; v[31] = s.initialized;
    }
    check(qcb_ecmc_set_chains(e, buf.data()), "qcb_ecmc_set_chains");
}

// the driver zeroes the event, lift and reversal counters after reading them
void reset_counters(qcb_ecmc* e, int n, bool revs) {
    std::vector<ChainState> st = get_states(e, n);
    for (ChainState& s : st) {
        s.events = s.lifts = 0;
        if (revs) s.revs = 0;
    }
    set_states(e, st);
}

void generate_ecmc_cb(const RunParamsECB& rp, bool existing) {
    mpi::MpiTopology topo(rp.n_core_dims);
    GeometryCB geo(rp.L_core, rp.n_core_dims);
    GaugeField field(geo, QCB_F64, 0);
    std::mt19937_64 rng(rp.seed + topo.rank);
    if (!rp.cold_start) field.hot_start(geo, rng, rp.seed);

    qcb_ecmc_params ep{};
    ep.beta = rp.ecmc_params.beta;
    ep.theta_sample = rp.ecmc_params.param_theta_sample;
    ep.theta_refresh_site = rp.ecmc_params.param_theta_refresh_site;
    ep.theta_refresh_R = rp.ecmc_params.param_theta_refresh_R;
    ep.poisson = rp.ecmc_params.poisson ? 1 : 0;
    ep.epsilon_set = rp.ecmc_params.epsilon_set;
    for (int d = 0; d < 4; ++d) ep.block[d] = rp.L_core;
    ep.seed = chain_key(rp.seed);
    qcb_ecmc* ecmc = nullptr;
    check(qcb_ecmc_create(field.lat, &ep, &ecmc), "qcb_ecmc_create");
    int n_chains = 0;
    check(qcb_ecmc_n_chains(ecmc, &n_chains), "qcb_ecmc_n_chains");

    if (existing) {
        read_ildg_clime(rp.run_name, rp.run_dir, field, geo, topo);
        std::vector<ChainState> st(n_chains);
        try {
            if (qcb_host::io::load_states(st, rp.L_core, rp.n_core_dims, rp.run_name, rp.run_dir))
                set_states(ecmc, st);
        } catch (const std::exception& ex) {
            std::cerr << ex.what() << "\n";
            std::exit(1);
        }
        fs::path state_path = fs::path(rp.run_dir) / rp.run_name / (rp.run_name + "_seed") /
                              (rp.run_name + "_seed" + std::to_string(topo.rank) + ".txt");
        std::ifstream ifs(state_path);
        if (ifs.is_open()) ifs >> rng;
        else std::cerr << "Could not open " << state_path << "\n";
    }
    mpi::exchange::exchange_halos_cascade(field, geo, topo);
    mpi::HalosShift halo_shift(geo);
    const double eps = 0.02;
    GradientFlow flow(eps, field, geo);

    std::vector<double> plaquette, tQE_tot, tQE_current;
    std::vector<size_t> event_nb, lift_nb, rev_nb;
    std::vector<double> lambda;

    qcb_host::io::save_params(rp, rp.run_name, rp.run_dir);
    qcb_host::print_parameters(rp);

    std::cout << "\n\n===========================================\n";
    std::cout << "Thermalisation : " << rp.N_therm << " shifts\n";
    std::cout << "===========================================\n";
    for (int i = 0; i < rp.N_therm; i++) {
        std::cout << "\n\n==========" << "(Therm) Shift " << i << "==========\n";
        mpi::shift::random_shift(field, geo, halo_shift, topo, rng);
        check(qcb_ecmc_sample(ecmc, 1), "qcb_ecmc_sample");
        mpi::exchange::exchange_halos_cascade(field, geo, topo);
        if (i % rp.N_shift_plaquette == 0) {
            double p = mpi::observables::mean_plaquette_global(field, geo, topo);
            std::cout << "====== Plaquette ======\n";
            std::cout << "(Therm) Sample " << i / rp.N_shift_plaquette << ", <P> = " << p << "\n";
        }
        reset_counters(ecmc, n_chains, false);
    }

    std::cout << "\n\n===========================================\n";
    std::cout << "Sampling : " << rp.N_shift / rp.N_shift_plaquette << " <P> samples, "
              << rp.N_shift / rp.N_shift_topo << " Q samples\n";
    std::cout << "===========================================\n";
    reset_counters(ecmc, n_chains, false);

    auto save_all = [&](int shift, bool finished) {
        std::cout << "\n\n==========================================\n";
        qcb_host::io::save_plaquette(plaquette, rp.run_name, rp.run_dir, 10);
        if (rp.topo) qcb_host::io::save_topo(tQE_tot, rp.run_name, rp.run_dir, 10);
        qcb_host::io::add_shift(shift, rp.run_name, rp.run_dir);
        if (finished) qcb_host::io::add_finished(rp.run_name, rp.run_dir);
        qcb_host::io::save_event_nb(event_nb, lift_nb, rev_nb, lambda, rp.run_name, rp.run_dir);
        save_ildg_clime(rp.run_name, rp.run_dir, field, geo, topo);
        qcb_host::io::save_seed(rng, rp.run_name, rp.run_dir, topo.rank);
        qcb_host::io::save_states(get_states(ecmc, n_chains), rp.L_core, rp.n_core_dims, rp.run_name, rp.run_dir);
        std::cout << "==========================================\n";
    };

    for (int i = 0; i < rp.N_shift; i++) {
        std::cout << "\n\n=============" << "Shift " << i << "=============\n";
        mpi::shift::random_shift(field, geo, halo_shift, topo, rng);
        check(qcb_ecmc_sample(ecmc, 1), "qcb_ecmc_sample");
        mpi::exchange::exchange_halos_cascade(field, geo, topo);
        if (i % rp.N_shift_plaquette == 0) {
            double p = mpi::observables::mean_plaquette_global(field, geo, topo);
            std::cout << "====== Plaquette ======\n";
            std::cout << "Sample " << i / rp.N_shift_plaquette << ", <P> = " << p << "\n";
            plaquette.emplace_back(p);
            // event counts summed over the chains (the reference's MPI_Reduce over ranks)
            uint64_t g_lifts = 0, g_events = 0, g_revs = 0;
            for (const ChainState& s : get_states(ecmc, n_chains)) {
                g_lifts += s.lifts;
                g_events += s.events;
                g_revs += s.revs;
            }
            const double total_dist = rp.N_shift_plaquette * rp.ecmc_params.param_theta_sample * n_chains;
            const double avg_lambda = total_dist / (double)g_lifts;
            const size_t avg_lift = g_lifts / n_chains, avg_rev = g_revs / n_chains, avg_event = g_events / n_chains;
            lambda.emplace_back(avg_lambda);
            lift_nb.emplace_back(avg_lift);
            event_nb.emplace_back(avg_event);
            rev_nb.emplace_back(avg_rev);
            std::cout << ">>> Average λ = " << avg_lambda << ", Events : " << avg_event
                      << " (Reversibility : " << (double)avg_rev / (double)avg_lift * 100 << "%)\n";
            reset_counters(ecmc, n_chains, true);
        }
        if (rp.topo && (i % rp.N_shift_topo == 0)) {
            std::cout << "====== Topology ======\n";
            tQE_current = mpi::observables::topo_charge_flowed(field, geo, flow, topo, rp.N_steps_gf, rp.N_rk_steps);
            if (tQE_current.size() >= 2)
                std::cout << "Sample " << i / rp.N_shift_topo << ", final Q = " << tQE_current[tQE_current.size() - 2]
                          << "\n";
            tQE_tot.insert(tQE_tot.end(), tQE_current.begin(), tQE_current.end());
        }
        if (i > 0 && i % rp.save_each_shifts == 0) {
            save_all(i, false);
            plaquette.clear();
            event_nb.clear();
            lift_nb.clear();
            rev_nb.clear();
            lambda.clear();
            tQE_tot.clear();
        }
    }
    save_all(rp.N_shift, true);
    check(qcb_ecmc_destroy(ecmc), "qcb_ecmc_destroy");
}

// --print-states: the chain states a resume of this input would load, one
// line per chain (device chain index, rank of the file, global site coords,
// mu, epsilon, counters, Philox position); no device is touched.
int print_states(const RunParamsECB& rp) {
    const int n = rp.n_core_dims * rp.n_core_dims * rp.n_core_dims * rp.n_core_dims;
    const int L = rp.L_core * rp.n_core_dims;
    std::vector<ChainState> st(n);
    try {
        if (!qcb_host::io::load_states(st, rp.L_core, rp.n_core_dims, rp.run_name, rp.run_dir)) {
            std::cerr << "No chain state files for run " << rp.run_name << " in " << rp.run_dir << "\n";
            return 1;
        }
    } catch (const std::exception& ex) {
        std::cerr << ex.what() << "\n";
        return 1;
    }
    std::cout << std::setprecision(17);
    for (int c = 0; c < n; ++c) {
        const ChainState& s = st[c];
        uint64_t x = s.site;
        int co[4];
        for (int d = 0; d < 4; ++d) {
            co[d] = (int)(x % (uint64_t)L);
            x /= (uint64_t)L;
        }
        std::cout << "chain " << c << " site " << co[0] << " " << co[1] << " " << co[2] << " " << co[3] << " mu "
                  << s.mu << " eps " << s.eps << " par_rs " << s.par_refresh_site << " par_rR " << s.par_refresh_R
                  << " th_rs " << s.th_refresh_site << " th_rR " << s.th_refresh_R << " set_counter " << s.set_counter
                  << " events " << s.events << " lifts " << s.lifts << " revs " << s.revs << " init "
                  << s.initialized << " rng_n " << s.rng_n << " set_gen " << s.set_gen << " R";
        for (int k = 0; k < 18; ++k) std::cout << " " << s.R[k];
        std::cout << "\n";
    }
    return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc < 2) {
        std::cerr << "Usage: " << argv[0] << " [--print-params | --print-states] <input_file.txt>" << std::endl;
        return 1;
    }
    RunParamsECB params;
    const std::string mode = argv[1];
    if ((mode == "--print-params" || mode == "--print-states") && argc >= 3) {
        // parse (and read the saved chain states) only: no device, no output files
        try {
            qcb_host::io::load_params(argv[2], params);
        } catch (const std::exception& e) {
            std::cerr << "Error reading input : " << e.what() << std::endl;
            return 1;
        }
        if (mode == "--print-states") return print_states(params);
        qcb_host::print_parameters(params);
        std::cout << "N_samples=" << params.ecmc_params.N_samples << " N_switch_eo=" << params.N_switch_eo
                  << " cold_start=" << (params.cold_start ? "true" : "false") << "\n";
        return 0;
    }
    bool existing = qcb_host::io::read_params(params, argv[1]);
    init_device(0);
    auto start = std::chrono::steady_clock::now();
    generate_ecmc_cb(params, existing);
    check(qcb_ctx_synchronize(device_ctx()), "qcb_ctx_synchronize");
    double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::cout << std::fixed << std::setprecision(4);
    qcb_host::print_time((long)total);
    qcb_ctx_destroy(device_ctx());
    return 0;
}
