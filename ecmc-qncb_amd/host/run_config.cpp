// run_config.cpp -- input parsing and the echoed parameter text (run_config.hpp).
#include "run_config.hpp"

#include <fstream>
#include <sstream>

namespace qcb_drv {

namespace {

std::string strip(const std::string& s) {
    const char* blank = " \t";
    const size_t a = s.find_first_not_of(blank);
    if (a == std::string::npos) return std::string();
    return s.substr(a, s.find_last_not_of(blank) - a + 1);
}

// Keys every input shares.
void take_common(const KeyValues& kv, LatticeInput& lat, Cadence& cad, OutputInput& out) {
    kv.take("L_core", lat.L_core);
    kv.take("n_core_dims", lat.n_core_dims);
    kv.take_flag("cold_start", lat.cold_start);
    kv.take("seed", lat.seed);
    kv.take("N_shift", cad.shifts);
    kv.take("N_shift_plaquette", cad.plaquette_every);
    kv.take("N_shift_therm", cad.therm);
    kv.take_flag("topo", cad.topo);
    kv.take("N_shift_topo", cad.topo_every);
    kv.take("N_steps_gf", cad.gf_steps);
    kv.take("N_rk_steps", cad.rk_steps);
    kv.take("run_name", out.run_name);
    kv.take("run_dir", out.run_dir);
    kv.take("save_each_shifts", cad.save_every);
}

// "label : value" lines of the start-of-run summary
struct Summary {
    std::ostream& os;
    template <class V>
    Summary& put(const char* label, const V& v, const char* unit = "") {
        os << label << " : " << v << unit << "\n";
        return *this;
    }
    Summary& text(const char* t) {
        os << t;
        return *this;
    }
};

void summary_head(Summary& s, const char* title, const LatticeInput& lat, const Cadence& cad,
                  const OutputInput& out) {
    const char* rule = "==========================================\n";
    s.text(rule).text(title).text("\n").text(rule);
    s.os << "Run name : " << out.run_name << "\n";
    s.text("---Lattice params---\n");
    s.put("Total lattice size", lat.L(), "^4");
    s.put("Local lattice size", lat.L_core, "^4\n");
    s.text("---Run params---\n");
    s.put("Initial seed", lat.seed).put("Thermalization shifts", cad.therm).put("Number of shifts", cad.shifts);
    s.put("Save each", cad.save_every, " shifts\n");
}

void summary_tail(Summary& s, const Cadence& cad) {
    s.text("---Measures params---\n");
    s.put("Number of <P> samples", cad.shifts / cad.plaquette_every);
    s.put("Measure <P> each", cad.plaquette_every, " shifts");
    s.put("Measure topo", cad.topo ? "Yes" : "No");
    if (cad.topo) {
        s.put("Number of Q samples", cad.shifts / cad.topo_every);
        s.put("Measure Q each", cad.topo_every, " shifts\n");
        s.text("---Gradient Flow---\n");
        s.put("Number of RK3 steps per GF step", cad.rk_steps).put("Number of GF steps", cad.gf_steps);
    }
    s.text("==========================================\n");
}

// The params-file block; `eq` is how each key is joined to its value (the
// two reference drivers space their lines differently and files must match).
struct LogBlock {
    std::ostringstream os;
    LogBlock() { os << std::boolalpha << "\n#########################################################\n\n"; }
    template <class V>
    LogBlock& kv(const char* key, const char* eq, const V& v) {
        os << key << eq << v << "\n";
        return *this;
    }
    LogBlock& text(const char* t) {
        os << t;
        return *this;
    }
};

void log_tail(LogBlock& b, const Cadence& cad) {
    b.text("# Plaquette params\n").kv("N_shift_plaquette", " = ", cad.plaquette_every).text("\n");
    b.text("# Topo params\n").kv("topo", " = ", cad.topo).kv("N_shift_topo", " = ", cad.topo_every);
    b.kv("N_steps_gf", " = ", cad.gf_steps).kv("N_rk_steps", " = ", cad.rk_steps).text("\n");
    b.text("#Save params\n").kv("save_each_shifts", " = ", cad.save_every).text("\n");
}

}  // namespace

KeyValues KeyValues::read(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw InputError("cannot open input file " + path);
    KeyValues r;
    for (std::string line; std::getline(in, line);) {
        if (line.empty() || line[0] == '#') continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos || eq + 1 == line.size()) continue;   // a key needs a value after '='
        r.kv_[strip(line.substr(0, eq))] = strip(line.substr(eq + 1));
    }
    return r;
}

void KeyValues::take(const char* key, int& v) const {
    if (has(key)) v = std::stoi(at(key));
}
void KeyValues::take(const char* key, double& v) const {
    if (has(key)) v = std::stod(at(key));
}
void KeyValues::take(const char* key, std::string& v) const {
    if (has(key)) v = at(key);
}
void KeyValues::take_flag(const char* key, bool& v) const {
    if (has(key)) v = at(key) == "true";
}

HeatbathInput parse_heatbath(const std::string& path) {
    const KeyValues kv = KeyValues::read(path);
    HeatbathInput in;
    take_common(kv, in.lat, in.cad, in.out);
    kv.take("beta", in.beta);
    kv.take("N_sweeps", in.n_sweeps);
    kv.take("N_hits", in.n_hits);
    // this driver's keys are checked strictly: the reference never reads them
    if (kv.has("precision")) {
        in.precision = kv.at("precision");
        if (in.precision != "fp64" && in.precision != "fp32")
            throw InputError("precision must be fp64 or fp32, not '" + in.precision + "'");
    }
    if (kv.has("reunit_every")) {
        const std::string& s = kv.at("reunit_every");
        size_t used = 0;
        in.reunit_every = std::stoi(s, &used);
        if (used != s.size() || in.reunit_every < 0) throw InputError("reunit_every must be a count >= 0");
    }
    return in;
}

ChainInput parse_chain(const std::string& path) {
    const KeyValues kv = KeyValues::read(path);
    ChainInput in;
    take_common(kv, in.lat, in.cad, in.out);
    kv.take("beta", in.beta);
    kv.take("param_theta_sample", in.theta_sample);
    kv.take("param_theta_refresh_site", in.theta_refresh_site);
    kv.take("param_theta_refresh_R", in.theta_refresh_R);
    kv.take_flag("poisson", in.poisson);
    kv.take("epsilon_set", in.epsilon_set);
    return in;
}

void print_summary(std::ostream& os, const HeatbathInput& in) {
    Summary s{os};
    summary_head(s, "Heatbath - MPI No Checkboard", in.lat, in.cad, in.out);
    s.text("---Heatbath params---\n");
    s.put("Beta", in.beta).put("Number of sweeps", in.n_sweeps).put("Number of hits", in.n_hits, "\n");
    summary_tail(s, in.cad);
}

void print_summary(std::ostream& os, const ChainInput& in) {
    Summary s{os};
    summary_head(s, "ECMC - MPI No Checkboard", in.lat, in.cad, in.out);
    s.text("---ECMC params---\n");
    s.put("Beta", in.beta).put("Theta sample", in.theta_sample).put("Theta refresh site", in.theta_refresh_site);
    s.put("Theta refresh R", in.theta_refresh_R).put("Epsilon set", in.epsilon_set);
    s.put("Poisson law", in.poisson ? "Yes" : "No", "\n");
    summary_tail(s, in.cad);
}

std::string params_log(const HeatbathInput& in) {
    LogBlock b;
    b.text("# Lattice params\n").kv("L_core", "=", in.lat.L_core).kv("n_core_dims", "=", in.lat.n_core_dims);
    b.kv("cold_start", "=", in.lat.cold_start).text("\n");
    b.text("# Run params\n").kv("seed", "=", in.lat.seed).kv("N_shift_therm", " = ", in.cad.therm);
    b.kv("N_shift", " =", in.cad.shifts);
    b.text("# Heatbath params\n").kv("beta", " = ", in.beta).kv("N_sweeps", " = ", in.n_sweeps);
    b.kv("N_hits", " = ", in.n_hits).text("\n");
    log_tail(b, in.cad);
    return b.os.str();
}

std::string params_log(const ChainInput& in) {
    LogBlock b;
    b.text("# Lattice params\n").kv("L_core", " = ", in.lat.L_core).kv("n_core_dims", " = ", in.lat.n_core_dims);
    b.kv("cold_start", " = ", in.lat.cold_start).text("\n");
    b.text("# Run params\n").kv("seed", " = ", in.lat.seed).kv("N_shift_therm", " = ", in.cad.therm);
    b.kv("N_shift", " = ", in.cad.shifts);
    b.text("# ECMC params\n").kv("beta", " = ", in.beta).kv("theta_sample", " = ", in.theta_sample);
    b.kv("theta_refresh_site", " = ", in.theta_refresh_site).kv("theta_refresh_R", " = ", in.theta_refresh_R);
    b.kv("epsilon_set", " = ", in.epsilon_set).kv("poisson", " = ", in.poisson).text("\n");
    log_tail(b, in.cad);
    return b.os.str();
}

}  // namespace qcb_drv
