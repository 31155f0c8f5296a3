// qcb_mirror.hpp -- the reference's C++ interface for the heatbath path
// (src/gauge/GaugeField.h, src/flow/gradient_flow.h, src/heatbath/heatbath_mpi.h,
// src/observables/observables_mpi.h, src/mpi/Shift.h, src/mpi/HalosExchange.h,
// src/io/ildg.h) re-expressed over the qcb C-ABI, so a driver written against
// the reference reads the same.  The gauge field lives on the GPU; there is
// one process and no ranks (MpiTopology is rank 0 of 1); halo exchanges are
// periodic index arithmetic inside the kernels and become no-ops here.
// Errors: a failing C-ABI call prints qcb_last_error to stderr and exits with
// status 1, the counterpart of the reference's MPI_Abort.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "../../include/qcb.h"

namespace qcb_mirror {

// HbParams (src/io/params.h:17-23), with the reference defaults.
struct HbParams {
    double beta = 6.0;
    int N_samples = 10;
    int N_hits = 1;
    int N_sweeps = 1;
    int N_therm = 100;
};

// printf("%.*f"): the reference's std::fixed / setprecision text of a value
inline std::string fixed_text(double v, int decimals) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*f", decimals, v);
    return buf;
}

inline qcb_ctx*& device_ctx() {
    static qcb_ctx* c = nullptr;
    return c;
}

inline void check(int rc, const char* what) {
    if (rc != QCB_OK) {
        std::fprintf(stderr, "Fatal: %s failed (%d): %s\n", what, rc, qcb_last_error(device_ctx()));
        std::exit(1);
    }
}

// Opens the process-wide device context (MPI_Init + MpiTopology of the reference).
inline void init_device(int device = 0) {
    if (!device_ctx()) check(qcb_ctx_create(device, &device_ctx()), "qcb_ctx_create");
}

namespace mpi {
struct MpiTopology {
    int rank = 0;
    int size = 1;
    int n_core_dim = 1;
    explicit MpiTopology(int n = 1) : n_core_dim(n) {}
};
}  // namespace mpi

// Lattice extents (GeometryCB, src/geometry/GeometryCB.h): no tables.  The
// reference's GeometryCB is one rank's L_core^4 block of an n_core^4 Cartesian
// grid of ranks; here L is the whole lattice and Lc the block, which bounds
// the random shifts (src/mpi/Shift.cpp:166) and the hot start's rank blocks.
struct GeometryCB {
    int L[4];
    int Lc[4];
    int n_core = 1;
    explicit GeometryCB(int L_) : L{L_, L_, L_, L_}, Lc{L_, L_, L_, L_} {}
    GeometryCB(int L_core, int n_core_dims)
        : L{L_core * n_core_dims, L_core * n_core_dims, L_core * n_core_dims, L_core * n_core_dims},
          Lc{L_core, L_core, L_core, L_core}, n_core(n_core_dims) {}
    GeometryCB(int lx, int ly, int lz, int lt) : L{lx, ly, lz, lt}, Lc{lx, ly, lz, lt} {}
    size_t V() const { return (size_t)L[0] * L[1] * L[2] * L[3]; }
    int n_ranks() const { return n_core * n_core * n_core * n_core; }
    // block coordinates of a rank in MPI_Cart_create's row-major order (t fastest)
    void rank_block(int r, int b[4]) const {
        for (int d = 3; d >= 0; --d) {
            b[d] = r % n_core;
            r /= n_core;
        }
    }
};

// SU(3) draw of the reference's hot start (src/su3/utils.cpp:13-30) on the
// host: 18 normal deviates, Householder QR with Eigen's sign conventions,
// Q / det(Q)^(1/3).  Writes 18 doubles (row-major, re/im).
void random_su3(std::mt19937_64& rng, double* out);

class GaugeField {
public:
    qcb_lattice* lat = nullptr;
    GeometryCB geo;
    int prec;
    long sweeps_done = 0;   // for the reunitarisation cadence
    int reunit_every = 0;

    explicit GaugeField(const GeometryCB& g, int prec_ = QCB_F64, int reunit_every_ = 0)
        : geo(g), prec(prec_), reunit_every(reunit_every_) {
        init_device();
        check(qcb_lattice_create(device_ctx(), g.L, prec, &lat), "qcb_lattice_create");
    }
    ~GaugeField() {
        if (lat) qcb_lattice_destroy(lat);
    }
    GaugeField(const GaugeField&) = delete;
    GaugeField& operator=(const GaugeField&) = delete;

    // GaugeField::hot_start (src/gauge/GaugeField.cpp:10-20): host draws in
    // t,z,y,x,mu order over each rank's block, then one upload.  Rank 0 draws
    // from the caller's mt19937_64 (seeded seed + 0); rank r > 0 from its own
    // mt19937_64(seed + r), as each reference rank seeds its generator.
    void hot_start(const GeometryCB& g, std::mt19937_64& rng, int seed = 0) {
        std::vector<double> h(g.V() * 72);
        for (int r = 0; r < g.n_ranks(); ++r) {
            std::mt19937_64 own((uint64_t)(seed + r));
            std::mt19937_64& gen = r == 0 ? rng : own;
            int b[4];
            g.rank_block(r, b);
            for (int t = 0; t < g.Lc[3]; ++t)
                for (int z = 0; z < g.Lc[2]; ++z)
                    for (int y = 0; y < g.Lc[1]; ++y)
                        for (int x = 0; x < g.Lc[0]; ++x) {
                            const size_t X = b[0] * g.Lc[0] + x, Y = b[1] * g.Lc[1] + y, Z = b[2] * g.Lc[2] + z,
                                         T = b[3] * g.Lc[3] + t;
                            const size_t site = ((T * g.L[2] + Z) * g.L[1] + Y) * g.L[0] + X;
                            for (int mu = 0; mu < 4; ++mu) random_su3(gen, h.data() + 18 * (4 * site + mu));
                        }
        }
        check(qcb_lattice_set_links(lat, h.data()), "qcb_lattice_set_links");
    }
    void cold_start(const GeometryCB&) { check(qcb_cold(lat), "qcb_cold"); }
    void project_field_su3(const GeometryCB&) { check(qcb_reunitarize(lat), "qcb_reunitarize"); }
    std::vector<double> links() const {
        std::vector<double> h(geo.V() * 72);
        check(qcb_lattice_get_links(lat, h.data()), "qcb_lattice_get_links");
        return h;
    }
};

// GradientFlow (src/flow/gradient_flow.h:11-27): field_c lives in the flow
// buffers of the lattice handle.
class GradientFlow {
public:
    double epsilon;
    GaugeField* field;
    GradientFlow(double eps, GaugeField& f, const GeometryCB&) : epsilon(eps), field(&f) {}
    void copy(const GaugeField&) { check(qcb_flow_copy(field->lat), "qcb_flow_copy"); }
    void rk3_step(mpi::MpiTopology&) { check(qcb_flow_rk3(field->lat, epsilon, 1), "qcb_flow_rk3"); }
};

namespace mpi {
namespace exchange {
// Halos are periodic index arithmetic on the device: nothing to exchange.
inline void exchange_halos_cascade(GaugeField&, const GeometryCB&, MpiTopology&) {}
}  // namespace exchange

namespace heatbathcb {
// samples (src/heatbath/heatbath_mpi.cpp:59-69): N_samples x N_sweeps
// checkerboard sweeps on the device.  The Philox key is one draw of the
// caller's mt19937_64 per call, so the host stream, its saved text state and
// resume keep driving all randomness.
inline void samples(GaugeField& field, const GeometryCB&, MpiTopology&, const HbParams& p, std::mt19937_64& rng) {
    const uint64_t key = rng();
    const int n = p.N_samples * p.N_sweeps;
    for (int s = 0; s < n; ++s) {
        check(qcb_hb_sweeps(field.lat, p.beta, p.N_hits, 1, key, (uint32_t)s), "qcb_hb_sweeps");
        field.sweeps_done++;
        if (field.reunit_every > 0 && field.sweeps_done % field.reunit_every == 0)
            check(qcb_reunitarize(field.lat), "qcb_reunitarize");
    }
}
}  // namespace heatbathcb

namespace observables {
inline double mean_plaquette_global(GaugeField& field, const GeometryCB&, MpiTopology&) {
    double p = 0;
    check(qcb_plaquette(field.lat, &p), "qcb_plaquette");
    return p;
}
inline std::pair<double, double> topo_q_e_clover_global(GaugeField& field, const GeometryCB&, MpiTopology&) {
    double q = 0, e = 0;
    check(qcb_clover_qe(field.lat, &q, &e), "qcb_clover_qe");
    return {q, e};
}
// topo_charge_flowed (src/observables/observables_mpi.cpp:153-178), printing
// the same per-step lines on rank 0.
inline std::vector<double> topo_charge_flowed(GaugeField& field, const GeometryCB&, GradientFlow& gf, MpiTopology& topo,
                                              int N_steps_gf, int N_rk_steps) {
    std::vector<double> tQE(3 * (size_t)N_steps_gf);
    if (N_steps_gf > 0)
        check(qcb_topo_charge_flowed(field.lat, gf.epsilon, N_steps_gf, N_rk_steps, tQE.data(), nullptr),
              "qcb_topo_charge_flowed");
    if (topo.rank == 0)
        for (int s = 0; s < N_steps_gf; ++s) {
            double t = tQE[3 * s], q = tQE[3 * s + 1], e = tQE[3 * s + 2];
            std::printf("n = %d, t = %s, Q = %s, t\xc2\xb2" "E = %s\n", s, fixed_text(t, 2).c_str(),
                        fixed_text(q, 6).c_str(), fixed_text(t * t * e, 6).c_str());
        }
    std::fflush(stdout);
    return tQE;
}
}  // namespace observables

struct HalosShift {
    explicit HalosShift(const GeometryCB&) {}
};

namespace shift {
// random_shift (src/mpi/Shift.cpp:163-186): the same four uniform_int draws
// from the caller's mt19937_64, the same stdout line, then a device translate.
inline void random_shift(GaugeField& field, const GeometryCB& geo, HalosShift&, MpiTopology& topo, std::mt19937_64& rng) {
    int s[4];
    for (int d = 0; d < 4; ++d) {
        std::uniform_int_distribution<int> rand_l_shift(0, geo.Lc[d] - 1);
        s[d] = rand_l_shift(rng);
    }
    if (topo.rank == 0) std::printf("Shift (%d, %d, %d, %d)\n", s[0], s[1], s[2], s[3]);
    check(qcb_translate(field.lat, s), "qcb_translate");
}
}  // namespace shift
}  // namespace mpi

// save_ildg_clime / read_ildg_clime (src/io/ildg.cpp:34-237): run_dir/run_name/run_name.
void save_ildg_clime(const std::string& filename, const std::string& dirpath, GaugeField& field, const GeometryCB& geo,
                     const mpi::MpiTopology& topo);
void read_ildg_clime(const std::string& filename, const std::string& dirpath, GaugeField& field, const GeometryCB& geo,
                     mpi::MpiTopology& topo);

}  // namespace qcb_mirror
