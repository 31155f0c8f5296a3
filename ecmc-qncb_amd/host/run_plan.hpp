// run_plan.hpp -- the schedule both drivers run, as data plus one executor.
//
// A run is a list of shifts.  Each shift translates the lattice by a random
// vector (src/mpi/Shift.cpp:163-186) and lets the generator update it once;
// the shift's task then says which measurements follow and whether a
// checkpoint is written.  The generator (the heatbath or the event chains)
// supplies the update, its own per-shift bookkeeping and the list of
// checkpoint artifacts it needs, in the order it writes them.  The stdout
// lines, observable files and resume rule are those of the reference's
// drivers (src/bin/gauge_hb_ncb.cpp, src/bin/gauge_ecmc_ncb.cpp).
#pragma once
#include <iosfwd>
#include <random>
#include <string>
#include <vector>

#include "qcb_mirror.hpp"
#include "run_config.hpp"
#include "run_files.hpp"

namespace qcb_drv {

struct ShiftTask {
    bool therm = false;
    int index = 0;            // shift number within its phase
    bool plaquette = false;   // measure <P> after the update
    bool topo = false;        // flow and measure Q, E (sampling only)
    bool checkpoint = false;  // periodic save after the measurements (sampling only)
};

struct Plan {
    std::vector<ShiftTask> therm, sampling;
};
Plan make_plan(const Cadence& c);

enum class Artifact {
    Plaquettes,     // run_name_plaquette.txt
    FlowSeries,     // run_name_topo.txt (when topology is measured)
    ShiftMark,      // "Saved shift n" in run_name_params.txt
    FinalMark,      // "Saved final state !" (final checkpoint only)
    Generator,      // run_name_seed/run_name_seed0.txt
    Configuration,  // the ILDG file
    EventCounts,    // run_name_nbevent.txt
    ChainStates,    // run_name_state/ files
};

class Generator {
public:
    virtual ~Generator() = default;
    virtual void summary(std::ostream& os) const = 0;
    virtual std::string params_text() const = 0;
    // banner line printed before each thermalisation shift (empty: none)
    virtual bool announces_therm_shifts() const { return false; }
    virtual void update(std::mt19937_64& rng) = 0;
    virtual void before_sampling() {}
    virtual void after_plaquette(const ShiftTask&) {}
    virtual void after_shift(const ShiftTask&) {}
    virtual std::vector<Artifact> checkpoint_artifacts() const = 0;
    // writes an artifact only this generator knows (EventCounts, ChainStates)
    virtual void write(Artifact, const RunFiles&) {}
    // a periodic checkpoint has written the observable windows: start new ones
    virtual void clear_windows() {}
};

class Runner {
public:
    Runner(const Cadence& cad, const RunFiles& files, qcb_mirror::GaugeField& field, std::mt19937_64& rng,
           Generator& gen);
    void run();

private:
    void shift(const ShiftTask& t);
    void measure_plaquette(const ShiftTask& t);
    void measure_topology(const ShiftTask& t);
    void checkpoint(int shift_mark, bool final);

    const Cadence& cad_;
    const RunFiles& files_;
    qcb_mirror::GaugeField& field_;
    std::mt19937_64& rng_;
    Generator& gen_;
    qcb_mirror::mpi::MpiTopology topo_;
    qcb_mirror::mpi::HalosShift halo_;
    qcb_mirror::GradientFlow flow_;
    std::vector<double> plaquettes_, flow_series_;
};

}  // namespace qcb_drv
