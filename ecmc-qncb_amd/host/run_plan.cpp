// run_plan.cpp -- plan construction and the executor (run_plan.hpp).
#include "run_plan.hpp"

#include <filesystem>
#include <iostream>

namespace qcb_drv {

using namespace qcb_mirror;

Plan make_plan(const Cadence& c) {
    Plan p;
    for (int i = 0; i < c.therm; ++i) {
        ShiftTask t;
        t.therm = true;
        t.index = i;
        t.plaquette = i % c.plaquette_every == 0;
        p.therm.push_back(t);
    }
    for (int i = 0; i < c.shifts; ++i) {
        ShiftTask t;
        t.index = i;
        t.plaquette = i % c.plaquette_every == 0;
        t.topo = c.topo && i % c.topo_every == 0;
        t.checkpoint = i > 0 && i % c.save_every == 0;
        p.sampling.push_back(t);
    }
    return p;
}

Runner::Runner(const Cadence& cad, const RunFiles& files, GaugeField& field, std::mt19937_64& rng, Generator& gen)
    : cad_(cad), files_(files), field_(field), rng_(rng), gen_(gen), halo_(field.geo),
      flow_(0.02, field, field.geo) {}   // the flow step is fixed at 0.02, as in the reference drivers

void Runner::run() {
    const char* rule = "===========================================\n";
    files_.append("_params.txt", gen_.params_text(), "Parameters saved in");
    gen_.summary(std::cout);
    const Plan plan = make_plan(cad_);
    std::cout << "\n\n" << rule << "Thermalisation : " << cad_.therm << " shifts\n" << rule;
    for (const ShiftTask& t : plan.therm) shift(t);
    std::cout << "\n\n" << rule << "Sampling : " << cad_.shifts / cad_.plaquette_every << " <P> samples, "
              << cad_.shifts / cad_.topo_every << " Q samples\n" << rule;
    gen_.before_sampling();
    for (const ShiftTask& t : plan.sampling) shift(t);
    checkpoint(cad_.shifts, true);
}

void Runner::shift(const ShiftTask& t) {
    if (!t.therm) {
        std::cout << "\n\n=============Shift " << t.index << "=============\n" << std::flush;
    } else if (gen_.announces_therm_shifts()) {
        std::cout << "\n\n==========(Therm) Shift " << t.index << "==========\n";
    }
    mpi::shift::random_shift(field_, field_.geo, halo_, topo_, rng_);
    gen_.update(rng_);
    if (t.plaquette) measure_plaquette(t);
    if (t.topo) measure_topology(t);
    if (t.checkpoint) {
        checkpoint(t.index, false);
        plaquettes_.clear();
        flow_series_.clear();
        gen_.clear_windows();
    }
    gen_.after_shift(t);
}

void Runner::measure_plaquette(const ShiftTask& t) {
    const double p = mpi::observables::mean_plaquette_global(field_, field_.geo, topo_);
    std::cout << "====== Plaquette ======\n"
              << (t.therm ? "(Therm) " : "") << "Sample " << t.index / cad_.plaquette_every << ", <P> = " << p << "\n";
    if (t.therm) return;
    plaquettes_.push_back(p);
    gen_.after_plaquette(t);
}

void Runner::measure_topology(const ShiftTask& t) {
    std::cout << "====== Topology ======\n" << std::flush;
    const std::vector<double> tQE =
        mpi::observables::topo_charge_flowed(field_, field_.geo, flow_, topo_, cad_.gf_steps, cad_.rk_steps);
    if (tQE.size() >= 2)
        std::cout << "Sample " << t.index / cad_.topo_every << ", final Q = " << tQE[tQE.size() - 2] << "\n";
    flow_series_.insert(flow_series_.end(), tQE.begin(), tQE.end());
}

void Runner::checkpoint(int shift_mark, bool final) {
    std::cout << "\n\n==========================================\n";
    for (Artifact a : gen_.checkpoint_artifacts()) {
        switch (a) {
            case Artifact::Plaquettes:
                files_.append_plaquettes(plaquettes_);
                break;
            case Artifact::FlowSeries:
                if (cad_.topo) files_.append_flow_series(flow_series_);
                break;
            case Artifact::ShiftMark:
                files_.log_line("Saved shift " + std::to_string(shift_mark));
                break;
            case Artifact::FinalMark:
                if (final) files_.log_line("Saved final state !");
                break;
            case Artifact::Generator:
                files_.save_generator(rng_, 0);
                break;
            case Artifact::Configuration: {
                std::error_code ec;
                std::filesystem::create_directories(files_.folder(), ec);
                const std::string path = files_.configuration().string();
                check(qcb_ildg_save(field_.lat, path.c_str()), "qcb_ildg_save");
                std::cout << "Configuration saved in " << files_.configuration() << "\n";
                break;
            }
            default:
                gen_.write(a, files_);
        }
    }
    std::cout << "==========================================\n";
}

}  // namespace qcb_drv
