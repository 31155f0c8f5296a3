// gauge_ecmc_cb.cpp -- the event-chain driver on one MI355X.
//
// Reads the reference's ECMC input file and writes its stdout lines,
// observable and event-count files, ILDG checkpoints, generator state and
// per-rank chain-state files, with its resume rule
// (src/bin/gauge_ecmc_ncb.cpp).  The reference runs one chain per MPI rank on
// its L_core^4 block; here the (L_core * n_core_dims)^4 lattice lives on the
// device and the n_core_dims^4 block chains run together (qcb_ecmc_*).  Random
// shifts draw from the host std::mt19937_64(seed) as in the reference; the
// chains draw from Philox counters keyed by the run seed, and each chain's
// position in its stream is saved with its state, so a resume continues it.
//
//   gauge_ecmc_cb <input.txt>                 run (or resume) the input
//   gauge_ecmc_cb --print-params <input.txt>  echo the parsed parameters only
//   gauge_ecmc_cb --print-states <input.txt>  the chain states a resume would load
#include <chrono>
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "chain_states.hpp"
#include "qcb_mirror.hpp"
#include "run_config.hpp"
#include "run_files.hpp"
#include "run_plan.hpp"

using namespace qcb_drv;
namespace qm = qcb_mirror;

namespace {

// splitmix64 finaliser of the run seed: the chains' Philox key
uint64_t chain_key(int seed) {
    uint64_t z = (uint64_t)(int64_t)seed + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Per-measurement chain statistics: the reference's MPI_Reduce sums over
// ranks, here over the chains, then averages per chain.
struct EventWindow {
    std::vector<size_t> events, lifts, revs;
    std::vector<double> lambda;
    void clear() {
        events.clear();
        lifts.clear();
        revs.clear();
        lambda.clear();
    }
};

class ChainGenerator final : public Generator {
public:
    ChainGenerator(const ChainInput& in, qm::GaugeField& field) : in_(in) {
        qcb_ecmc_params p{};
        p.beta = in.beta;
        p.theta_sample = in.theta_sample;
        p.theta_refresh_site = in.theta_refresh_site;
        p.theta_refresh_R = in.theta_refresh_R;
        p.poisson = in.poisson ? 1 : 0;
        p.epsilon_set = in.epsilon_set;
        for (int& b : p.block) b = in.lat.L_core;
        p.seed = chain_key(in.lat.seed);
        qm::check(qcb_ecmc_create(field.lat, &p, &e_), "qcb_ecmc_create");
        qm::check(qcb_ecmc_n_chains(e_, &n_), "qcb_ecmc_n_chains");
    }
    ~ChainGenerator() override { qcb_ecmc_destroy(e_); }
    ChainGenerator(const ChainGenerator&) = delete;
    ChainGenerator& operator=(const ChainGenerator&) = delete;

    // the chain states saved with the configuration (a resume)
    void restore(const RunFiles& files) {
        std::vector<ChainState> st((size_t)n_);
        if (load_chain_files(files, st, in_.lat.L_core, in_.lat.n_core_dims)) write_chains(e_, st);
    }

    void summary(std::ostream& os) const override { print_summary(os, in_); }
    std::string params_text() const override { return params_log(in_); }
    bool announces_therm_shifts() const override { return true; }
    void update(std::mt19937_64&) override { qm::check(qcb_ecmc_sample(e_, 1), "qcb_ecmc_sample"); }
    void before_sampling() override { zero_counts(false); }
    void after_shift(const ShiftTask& t) override {
        if (t.therm) zero_counts(false);
    }
    void after_plaquette(const ShiftTask&) override {
        uint64_t ev = 0, li = 0, re = 0;
        for (const ChainState& s : read_chains(e_, n_)) {
            ev += s.events;
            li += s.lifts;
            re += s.revs;
        }
        // theta travelled by all chains since the last measurement, per lift
        const double lambda = in_.cad.plaquette_every * in_.theta_sample * n_ / (double)li;
        const size_t avg_ev = ev / n_, avg_li = li / n_, avg_re = re / n_;
        win_.events.push_back(avg_ev);
        win_.lifts.push_back(avg_li);
        win_.revs.push_back(avg_re);
        win_.lambda.push_back(lambda);
        std::cout << ">>> Average \xce\xbb = " << lambda << ", Events : " << avg_ev
                  << " (Reversibility : " << (double)avg_re / (double)avg_li * 100 << "%)\n";
        zero_counts(true);
    }
    std::vector<Artifact> checkpoint_artifacts() const override {
        return {Artifact::Plaquettes,  Artifact::FlowSeries,    Artifact::ShiftMark, Artifact::FinalMark,
                Artifact::EventCounts, Artifact::Configuration, Artifact::Generator, Artifact::ChainStates};
    }
    void write(Artifact a, const RunFiles& files) override {
        if (a == Artifact::EventCounts) {
            std::ostringstream o;
            for (size_t i = 0; i < win_.events.size(); ++i)
                o << win_.events[i] << " " << win_.lifts[i] << " " << win_.revs[i] << " " << win_.lambda[i] << "\n";
            files.append("_nbevent.txt", o.str(), "Number of events written in");
        } else if (a == Artifact::ChainStates) {
            save_chain_files(files, read_chains(e_, n_), in_.lat.L_core, in_.lat.n_core_dims);
        }
    }
    void clear_windows() override { win_.clear(); }

private:
    // the event, lift (and reversal) counts restart after each reading
    void zero_counts(bool with_revs) {
        std::vector<ChainState> st = read_chains(e_, n_);
        for (ChainState& s : st) {
            s.events = s.lifts = 0;
            if (with_revs) s.revs = 0;
        }
        write_chains(e_, st);
    }

    const ChainInput& in_;
    qcb_ecmc* e_ = nullptr;
    int n_ = 0;
    EventWindow win_;
};

int run(const ChainInput& in) {
    const RunFiles files(in.out.run_dir, in.out.run_name);
    const bool resume = files.resumable();
    if (resume) std::cout << "All ranks found existing configuration. Resuming...\n";
    qm::init_device(0);
    const auto t0 = std::chrono::steady_clock::now();
    {
        const qm::GeometryCB geo(in.lat.L_core, in.lat.n_core_dims);
        qm::GaugeField field(geo, QCB_F64, 0);
        std::mt19937_64 rng((uint64_t)(int64_t)in.lat.seed);
        if (!in.lat.cold_start) field.hot_start(geo, rng, in.lat.seed);
        ChainGenerator gen(in, field);
        if (resume) {
            qm::check(qcb_ildg_load(field.lat, files.configuration().string().c_str()), "qcb_ildg_load");
            try {
                gen.restore(files);
            } catch (const std::exception& ex) {
                std::cerr << ex.what() << "\n";
                return 1;
            }
            files.load_generator(rng, 0);
        }
        Runner(in.cad, files, field, rng, gen).run();
        qm::check(qcb_ctx_synchronize(qm::device_ctx()), "qcb_ctx_synchronize");
    }
    const long secs = (long)std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "==========================================\nElapsed time : " << secs
              << "s\n==========================================\n";
    qcb_ctx_destroy(qm::device_ctx());
    return 0;
}

// The chain states a resume of this input would load, one line per device
// chain (its block's global site coordinates, mu, epsilon, the counters and
// Philox position, R); no device is touched.
int print_states(const ChainInput& in) {
    const int n_core = in.lat.n_core_dims, L = in.lat.L();
    std::vector<ChainState> st((size_t)n_core * n_core * n_core * n_core);
    try {
        if (!load_chain_files(RunFiles(in.out.run_dir, in.out.run_name), st, in.lat.L_core, n_core)) {
            std::cerr << "No chain state files for run " << in.out.run_name << " in " << in.out.run_dir << "\n";
            return 1;
        }
    } catch (const std::exception& ex) {
        std::cerr << ex.what() << "\n";
        return 1;
    }
    std::cout << std::setprecision(17);
    for (size_t c = 0; c < st.size(); ++c) {
        const ChainState& s = st[c];
        std::cout << "chain " << c << " site";
        for (uint64_t d = 0, x = s.site; d < 4; ++d, x /= (uint64_t)L) std::cout << " " << x % (uint64_t)L;
        std::cout << " mu " << s.mu << " eps " << s.eps << " par_rs " << s.par_refresh_site << " par_rR "
                  << s.par_refresh_R << " th_rs " << s.th_refresh_site << " th_rR " << s.th_refresh_R
                  << " set_counter " << s.set_counter << " events " << s.events << " lifts " << s.lifts << " revs "
                  << s.revs << " init " << s.initialized << " rng_n " << s.rng_n << " set_gen " << s.set_gen << " R";
        for (double x : s.R) std::cout << " " << x;
        std::cout << "\n";
    }
    return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
    const std::string mode = argc >= 3 ? argv[1] : "";
    const bool echo = mode == "--print-params", states = mode == "--print-states";
    if (argc < 2) {
        std::cerr << "Usage: " << argv[0] << " [--print-params | --print-states] <input_file.txt>\n";
        return 1;
    }
    ChainInput in;
    try {
        in = parse_chain(argv[echo || states ? 2 : 1]);
    } catch (const std::exception& e) {
        std::cerr << "Error reading input : " << e.what() << "\n";
        return 1;
    }
    if (states) return print_states(in);
    if (echo) {   // no device and no files
        print_summary(std::cout, in);
        std::cout << "N_samples=" << kSamplesPerShift << " N_switch_eo=" << kSwitchEoDefault
                  << " cold_start=" << (in.lat.cold_start ? "true" : "false") << "\n";
        return 0;
    }
    return run(in);
}
