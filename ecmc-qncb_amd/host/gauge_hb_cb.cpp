// gauge_hb_cb.cpp -- the heatbath driver on one MI355X.
//
// Reads the reference's heatbath input file and writes its stdout lines,
// observable files, ILDG checkpoints and generator state, with its resume
// rule (src/bin/gauge_hb_ncb.cpp).  The whole (L_core * n_core_dims)^4
// lattice lives on the device; the update between two random shifts is
// N_sweeps checkerboard sweeps of N_hits Cabibbo-Marinari hits each, keyed by
// one draw of the host std::mt19937_64(seed) so that generator drives all
// randomness and its saved text state resumes a run.  Extra input keys:
// precision = fp64 | fp32 and reunit_every = n.
//
//   gauge_hb_cb <input.txt>                 run (or resume) the input
//   gauge_hb_cb --print-params <input.txt>  echo the parsed parameters only
#include <chrono>
#include <iostream>
#include <random>

#include "qcb_mirror.hpp"
#include "run_config.hpp"
#include "run_files.hpp"
#include "run_plan.hpp"

using namespace qcb_drv;
namespace qm = qcb_mirror;

namespace {

class HeatbathGenerator final : public Generator {
public:
    HeatbathGenerator(const HeatbathInput& in, qm::GaugeField& field) : in_(in), field_(field) {
        hp_.beta = in.beta;
        hp_.N_samples = kSamplesPerShift;
        hp_.N_hits = in.n_hits;
        hp_.N_sweeps = in.n_sweeps;
    }
    void summary(std::ostream& os) const override { print_summary(os, in_); }
    std::string params_text() const override { return params_log(in_); }
    void update(std::mt19937_64& rng) override {
        qm::mpi::MpiTopology one;
        qm::mpi::heatbathcb::samples(field_, field_.geo, one, hp_, rng);
    }
    std::vector<Artifact> checkpoint_artifacts() const override {
        return {Artifact::Plaquettes, Artifact::FlowSeries, Artifact::ShiftMark,
                Artifact::FinalMark,  Artifact::Generator,  Artifact::Configuration};
    }

private:
    const HeatbathInput& in_;
    qm::GaugeField& field_;
    qm::HbParams hp_;
};

int run(const HeatbathInput& in) {
    const RunFiles files(in.out.run_dir, in.out.run_name);
    const bool resume = files.resumable();
    if (resume) std::cout << "All ranks found existing configuration. Resuming...\n";
    qm::init_device(0);
    const auto t0 = std::chrono::steady_clock::now();
    {
        const int prec = in.precision == "fp32" ? QCB_F32 : QCB_F64;
        const int reunit = in.reunit_every >= 0 ? in.reunit_every : (prec == QCB_F32 ? 4 : 0);
        const qm::GeometryCB geo(in.lat.L_core, in.lat.n_core_dims);
        qm::GaugeField field(geo, prec, reunit);
        std::mt19937_64 rng((uint64_t)(int64_t)in.lat.seed);
        if (!in.lat.cold_start) field.hot_start(geo, rng, in.lat.seed);
        if (resume) {
            qm::check(qcb_ildg_load(field.lat, files.configuration().string().c_str()), "qcb_ildg_load");
            files.load_generator(rng, 0);
        }
        HeatbathGenerator gen(in, field);
        Runner(in.cad, files, field, rng, gen).run();
        qm::check(qcb_ctx_synchronize(qm::device_ctx()), "qcb_ctx_synchronize");
    }
    const long secs = (long)std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "==========================================\nElapsed time : " << secs
              << "s\n==========================================\n";
    qcb_ctx_destroy(qm::device_ctx());
    return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
    const bool echo = argc >= 3 && std::string(argv[1]) == "--print-params";
    if (argc < 2) {
        std::cerr << "Usage: " << argv[0] << " [--print-params] <input_file.txt>\n";
        return 1;
    }
    HeatbathInput in;
    try {
        in = parse_heatbath(argv[echo ? 2 : 1]);
    } catch (const std::exception& e) {
        std::cerr << "Error reading input : " << e.what() << "\n";
        return 1;
    }
    if (echo) {   // no device and no files
        print_summary(std::cout, in);
        std::cout << "precision=" << in.precision << " reunit_every=" << in.reunit_every
                  << " N_samples=" << kSamplesPerShift << " N_switch_eo=" << kSwitchEoDefault << "\n";
        return 0;
    }
    return run(in);
}
