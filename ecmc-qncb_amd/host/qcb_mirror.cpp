// qcb_mirror.cpp -- host parts of the reference-interface mirror: the
// reference-style hot start and ILDG checkpoint paths.
#include "qcb_mirror.hpp"

#include <cmath>
#include <complex>
#include <filesystem>
#include <iostream>

namespace fs = std::filesystem;

namespace qcb_mirror {

using cd = std::complex<double>;

// Householder QR of a 3x3 complex Gaussian matrix, Q = H0 H1 H2 with
// H_k = 1 - conj(tau_k) v_k v_k^dag, beta = -sign(Re c0) |x|,
// tau = conj((beta - c0)/beta), v = (1, x_tail/(c0 - beta)) -- the
// conventions of Eigen's HouseholderQR used by the reference.  Entries are
// drawn real part first (the reference leaves the order to the compiler).
void random_su3(std::mt19937_64& rng, double* out) {
    std::normal_distribution<double> gauss(0.0, 1.0);
    cd A[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double re = gauss(rng);
            double im = gauss(rng);
            A[i][j] = cd(re, im);
        }
    cd v[3][3] = {};
    cd tau[3];
    for (int k = 0; k < 3; ++k) {
        double tail = 0.0;
        for (int i = k + 1; i < 3; ++i) tail += std::norm(A[i][k]);
        cd c0 = A[k][k];
        for (int i = 0; i < 3; ++i) v[k][i] = 0.0;
        v[k][k] = 1.0;
        const double tiny = std::numeric_limits<double>::min();
        if (tail <= tiny && c0.imag() * c0.imag() <= tiny) {
            tau[k] = 0.0;
            continue;
        }
        double beta = std::sqrt(std::norm(c0) + tail);
        if (c0.real() >= 0) beta = -beta;
        for (int i = k + 1; i < 3; ++i) v[k][i] = A[i][k] / (c0 - beta);
        tau[k] = std::conj((beta - c0) / beta);
        for (int j = k; j < 3; ++j) {
            cd s = 0.0;
            for (int i = k; i < 3; ++i) s += std::conj(v[k][i]) * A[i][j];
            for (int i = k; i < 3; ++i) A[i][j] -= v[k][i] * (tau[k] * s);
        }
    }
    cd M[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[i][j] = (i == j) ? 1.0 : 0.0;
    for (int k = 2; k >= 0; --k) {
        cd h = std::conj(tau[k]);
        for (int j = 0; j < 3; ++j) {
            cd s = 0.0;
            for (int i = 0; i < 3; ++i) s += std::conj(v[k][i]) * M[i][j];
            for (int i = 0; i < 3; ++i) M[i][j] -= v[k][i] * (h * s);
        }
    }
    cd det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) + M[0][1] * (M[1][2] * M[2][0] - M[1][0] * M[2][2]) +
             M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    cd root = std::pow(det, 1.0 / 3.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            cd q = M[i][j] / root;
            out[(3 * i + j) * 2] = q.real();
            out[(3 * i + j) * 2 + 1] = q.imag();
        }
}

void save_ildg_clime(const std::string& filename, const std::string& dirpath, GaugeField& field, const GeometryCB&,
                     const mpi::MpiTopology& topo) {
    fs::path dir(dirpath);
    fs::path file = dir / filename / filename;
    if (!fs::exists(dir)) fs::create_directories(dir);
    check(qcb_ildg_save(field.lat, file.string().c_str()), "qcb_ildg_save");
    if (topo.rank == 0) std::cout << "Configuration saved in " << file << "\n";
}

void read_ildg_clime(const std::string& filename, const std::string& dirpath, GaugeField& field, const GeometryCB&,
                     mpi::MpiTopology&) {
    fs::path file = fs::path(dirpath) / filename / filename;
    check(qcb_ildg_load(field.lat, file.string().c_str()), "qcb_ildg_load");
}

}  // namespace qcb_mirror
