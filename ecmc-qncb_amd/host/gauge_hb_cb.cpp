// gauge_hb_cb.cpp -- the heatbath driver on one MI355X.
//
// Same command line, input keys, stdout lines, output files, ILDG
// checkpoints and resume rule as the reference's driver
// (src/bin/gauge_hb_ncb.cpp:18-238); the per-shift work runs on the GPU
// through the qcb C-ABI.  The whole n_core_dims^4 * L_core^4 lattice lives on
// the one device, so there are no ranks, no frozen links and no halos; the
// random shifts are kept (same host mt19937_64 draws, device translate).
// Extra input keys of this driver: precision = fp64 | fp32, reunit_every = n.
#include <chrono>
#include <fstream>
#include <filesystem>
#include <iomanip>
#include <iostream>

#include "hb_io.hpp"
#include "qcb_mirror.hpp"

namespace fs = std::filesystem;
using namespace qcb_mirror;
using qcb_host::RunParamsHbCB;
namespace io = qcb_host::io;

static void generate_hb_cb(const RunParamsHbCB& rp, bool existing) {
    mpi::MpiTopology topo(rp.n_core_dims);
    GeometryCB geo(rp.L_core, rp.n_core_dims);
    const int prec = rp.precision == "fp32" ? QCB_F32 : QCB_F64;
    const int reunit = rp.reunit_every >= 0 ? rp.reunit_every : (prec == QCB_F32 ? 4 : 0);
    GaugeField field(geo, prec, reunit);
    std::mt19937_64 rng(rp.seed + topo.rank);
    if (!rp.cold_start) field.hot_start(geo, rng, rp.seed);

    if (existing) {
        read_ildg_clime(rp.run_name, rp.run_dir, field, geo, topo);
        fs::path state_path = fs::path(rp.run_dir) / rp.run_name / (rp.run_name + "_seed") /
                              (rp.run_name + "_seed" + std::to_string(topo.rank) + ".txt");
        std::ifstream ifs(state_path);
        if (ifs.is_open()) {
            ifs >> rng;
        } else {
            std::cerr << "Could not open " << state_path << "\n";
        }
    }
    mpi::exchange::exchange_halos_cascade(field, geo, topo);

    HbParams hp{rp.hp.beta, rp.hp.N_samples, rp.hp.N_hits, rp.hp.N_sweeps, rp.hp.N_therm};
    mpi::HalosShift halo_shift(geo);
    const double eps = 0.02;
    GradientFlow flow(eps, field, geo);

    std::vector<double> plaquette;
    std::vector<double> tQE_tot, tQE_current;

    if (topo.rank == 0) io::save_params(rp, rp.run_name, rp.run_dir);
    qcb_host::print_parameters(rp);

    if (topo.rank == 0) {
        std::cout << "\n\n===========================================\n";
        std::cout << "Thermalisation : " << rp.N_therm << " shifts\n";
        std::cout << "===========================================\n";
    }
    for (int i = 0; i < rp.N_therm; i++) {
        mpi::shift::random_shift(field, geo, halo_shift, topo, rng);
        mpi::heatbathcb::samples(field, geo, topo, hp, rng);
        mpi::exchange::exchange_halos_cascade(field, geo, topo);
        if (i % rp.N_shift_plaquette == 0) {
            double p = mpi::observables::mean_plaquette_global(field, geo, topo);
            if (topo.rank == 0) {
                std::cout << "====== Plaquette ======\n";
                std::cout << "(Therm) Sample " << i / rp.N_shift_plaquette << ", <P> = " << p << "\n";
            }
        }
    }

    if (topo.rank == 0) {
        std::cout << "\n\n===========================================\n";
        std::cout << "Sampling : " << rp.N_shift / rp.N_shift_plaquette << " <P> samples, "
                  << rp.N_shift / rp.N_shift_topo << " Q samples\n";
        std::cout << "===========================================\n";
    }
    for (int i = 0; i < rp.N_shift; i++) {
        if (topo.rank == 0) std::cout << "\n\n=============" << "Shift " << i << "=============\n";
        std::cout.flush();
        mpi::shift::random_shift(field, geo, halo_shift, topo, rng);
        mpi::heatbathcb::samples(field, geo, topo, hp, rng);
        mpi::exchange::exchange_halos_cascade(field, geo, topo);
        if (i % rp.N_shift_plaquette == 0) {
            double p = mpi::observables::mean_plaquette_global(field, geo, topo);
            if (topo.rank == 0) {
                std::cout << "====== Plaquette ======\n";
                std::cout << "Sample " << i / rp.N_shift_plaquette << ", <P> = " << p << "\n";
            }
            plaquette.emplace_back(p);
        }
        if (rp.topo && (i % rp.N_shift_topo == 0)) {
            if (topo.rank == 0) std::cout << "====== Topology ======\n";
            std::cout.flush();
            tQE_current = mpi::observables::topo_charge_flowed(field, geo, flow, topo, rp.N_steps_gf, rp.N_rk_steps);
            if (topo.rank == 0 && tQE_current.size() >= 2)
                std::cout << "Sample " << i / rp.N_shift_topo << ", final Q = " << tQE_current[tQE_current.size() - 2]
                          << "\n";
            tQE_tot.insert(tQE_tot.end(), tQE_current.begin(), tQE_current.end());
        }
        if (i > 0 && i % rp.save_each_shifts == 0) {
            if (topo.rank == 0) {
                std::cout << "\n\n==========================================\n";
                io::save_plaquette(plaquette, rp.run_name, rp.run_dir, 10);
                if (rp.topo) io::save_topo(tQE_tot, rp.run_name, rp.run_dir, 10);
                io::add_shift(i, rp.run_name, rp.run_dir);
            }
            io::save_seed(rng, rp.run_name, rp.run_dir, topo.rank);
            save_ildg_clime(rp.run_name, rp.run_dir, field, geo, topo);
            if (topo.rank == 0) std::cout << "==========================================\n";
            plaquette.clear();
            tQE_tot.clear();
        }
    }

    if (topo.rank == 0) {
        std::cout << "\n\n==========================================\n";
        io::save_plaquette(plaquette, rp.run_name, rp.run_dir, 10);
        if (rp.topo) io::save_topo(tQE_tot, rp.run_name, rp.run_dir, 10);
        io::add_shift(rp.N_shift, rp.run_name, rp.run_dir);
        io::add_finished(rp.run_name, rp.run_dir);
    }
    io::save_seed(rng, rp.run_name, rp.run_dir, topo.rank);
    save_ildg_clime(rp.run_name, rp.run_dir, field, geo, topo);
    if (topo.rank == 0) std::cout << "==========================================\n";
}

int main(int argc, char* argv[]) {
    if (argc < 2) {
        std::cerr << "Usage: " << argv[0] << " [--print-params] <input_file.txt>" << std::endl;
        return 1;
    }
    RunParamsHbCB params;
    if (std::string(argv[1]) == "--print-params" && argc >= 3) {
        // parse and echo the run parameters only (no device, no files)
        try {
            io::load_params(argv[2], params);
        } catch (const std::exception& e) {
            std::cerr << "Error reading input : " << e.what() << std::endl;
            return 1;
        }
        qcb_host::print_parameters(params);
        std::cout << "precision=" << params.precision << " reunit_every=" << params.reunit_every
                  << " N_samples=" << params.hp.N_samples << " N_switch_eo=" << params.N_switch_eo << "\n";
        return 0;
    }
    bool existing = io::read_params(params, argv[1]);
    init_device(0);
    auto start = std::chrono::steady_clock::now();
    generate_hb_cb(params, existing);
    qcb_ctx_synchronize(device_ctx());
    double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::cout << std::fixed << std::setprecision(4);
    qcb_host::print_time((long)total);
    qcb_ctx_destroy(device_ctx());
    return 0;
}
