/*
 * qo_lattice.c -- lattice-level restatement of the reference hot path
 * (TEST INFRASTRUCTURE ONLY; see qo.h for the layout and the halo argument).
 */
#include "qo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the oracle's data-parallel loops (flow, plaquette, clover);
 * results do not depend on it: every global sum runs sequentially. */
void qo_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

void qo_m3_dag(const double* a, double* out);
void qo_kp_mt(double beta, double k, qo_mt64* g, double out[4]);
int qo_kp_try(double alpha, double delta, double r1, double r2, double r3, double* a0);
void qo_kp_sphere(double a0, double u_phi, double u_cos, double out[4]);

#define RE(a, i, j) (a)[((i)*3 + (j))*2]
#define IM(a, i, j) (a)[((i)*3 + (j))*2 + 1]

static inline double* LNK(const qo_lat* l, size_t site, int mu) { return l->U + (site * 4 + (size_t)mu) * 18; }

static inline void coords(const qo_lat* l, size_t s, int c[4]) {
    c[0] = (int)(s % (size_t)l->L[0]); s /= (size_t)l->L[0];
    c[1] = (int)(s % (size_t)l->L[1]); s /= (size_t)l->L[1];
    c[2] = (int)(s % (size_t)l->L[2]);
    c[3] = (int)(s / (size_t)l->L[2]);
}
static inline size_t lex(const qo_lat* l, const int c[4]) {
    return (size_t)c[0] + (size_t)l->L[0] * ((size_t)c[1] + (size_t)l->L[1] * ((size_t)c[2] + (size_t)l->L[2] * (size_t)c[3]));
}
/* periodic neighbour: the one-rank halo of GeometryCB (src/geometry/GeometryCB.cpp:10-31
 * + src/mpi/HalosExchange.cpp) */
static inline size_t nb(const qo_lat* l, size_t s, int mu, int dir) {
    int c[4];
    coords(l, s, c);
    c[mu] = (c[mu] + (dir > 0 ? 1 : l->L[mu] - 1)) % l->L[mu];
    return lex(l, c);
}

int qo_lat_init(qo_lat* l, const int L[4]) {
    for (int d = 0; d < 4; ++d) {
        if (L[d] < 2) return -1;
        l->L[d] = L[d];
    }
    if (L[0] % 2) return -1;
    l->V = (size_t)L[0] * L[1] * L[2] * L[3];
    l->U = (double*)calloc(l->V * 72, sizeof(double));
    if (!l->U) return -1;
    qo_lat_cold(l);
    return 0;
}

void qo_lat_free(qo_lat* l) {
    free(l->U);
    l->U = NULL;
}

void qo_lat_cold(qo_lat* l) {
    memset(l->U, 0, l->V * 72 * sizeof(double));
    for (size_t s = 0; s < l->V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            double* u = LNK(l, s, mu);
            u[0] = u[8] = u[16] = 1.0;
        }
}

/* ---------- hot starts ---------- */

/* Householder QR in the form Eigen's HouseholderQR uses, Q = H0 H1 H2 with
 * H_k = 1 - conj(tau_k) v v^dag, then Q / det(Q)^(1/3) on the principal branch
 * (src/su3/utils.cpp:13-30).
 * Draw order: the reference's Complex(gauss(rng), gauss(rng)) leaves the order
 * of the two draws to the compiler.  imag_first = 0 is a clang-based build
 * (IntelLLVM, the reference cluster's compiler, modules_load.sh:5): arguments
 * left to right, the first draw is the real part.  imag_first = 1 is a GCC
 * build (CMakeLists.txt:97-102 branch): right to left, the first draw is the
 * imaginary part.  Both orders are pinned by compiled probes, the QR
 * conventions by QR uniqueness plus the geometric form of each reflection
 * (tests/test_oracle_su3.py). */
void qo_random_su3_ref_order(qo_mt64* g, double* Q, int imag_first) {
    qo_c A[3][3];
    /* a fresh normal_distribution per call: 18 draws are 9 whole polar pairs */
    double gs[18];
    for (int k = 0; k < 18; k += 2) qo_mt64_normal_pair(g, &gs[k], &gs[k + 1]);
    const int a = imag_first ? 1 : 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            A[i][j].re = gs[2 * (3 * i + j) + a];
            A[i][j].im = gs[2 * (3 * i + j) + 1 - a];
        }
    qo_c v[3][3]; /* v[k][0..] householder vectors, v[k][k] = 1 */
    qo_c tau[3];
    for (int k = 0; k < 3; ++k) {
        double tail = 0.0;
        for (int i = k + 1; i < 3; ++i) tail += A[i][k].re * A[i][k].re + A[i][k].im * A[i][k].im;
        qo_c c0 = A[k][k];
        for (int i = 0; i < 3; ++i) { v[k][i].re = 0; v[k][i].im = 0; }
        v[k][k].re = 1.0;
        if (tail <= 2.2250738585072014e-308 && c0.im * c0.im <= 2.2250738585072014e-308) {
            tau[k].re = 0; tau[k].im = 0;
            continue;
        }
        double beta = sqrt(c0.re * c0.re + c0.im * c0.im + tail);
        if (c0.re >= 0) beta = -beta;
        qo_c den = {c0.re - beta, c0.im};
        double dn = den.re * den.re + den.im * den.im;
        for (int i = k + 1; i < 3; ++i) { /* essential = tail / (c0 - beta) */
            v[k][i].re = (A[i][k].re * den.re + A[i][k].im * den.im) / dn;
            v[k][i].im = (A[i][k].im * den.re - A[i][k].re * den.im) / dn;
        }
        /* tau = conj((beta - c0)/beta) */
        tau[k].re = (beta - c0.re) / beta;
        tau[k].im = c0.im / beta;
        /* apply (1 - tau v v^dag) from the left to columns k..2 of A */
        for (int j = k; j < 3; ++j) {
            qo_c s = {0, 0}; /* v^dag A[:,j] */
            for (int i = k; i < 3; ++i) {
                s.re += v[k][i].re * A[i][j].re + v[k][i].im * A[i][j].im;
                s.im += v[k][i].re * A[i][j].im - v[k][i].im * A[i][j].re;
            }
            qo_c ts = {tau[k].re * s.re - tau[k].im * s.im, tau[k].re * s.im + tau[k].im * s.re};
            for (int i = k; i < 3; ++i) {
                A[i][j].re -= v[k][i].re * ts.re - v[k][i].im * ts.im;
                A[i][j].im -= v[k][i].re * ts.im + v[k][i].im * ts.re;
            }
        }
    }
    /* Q = H0 H1 H2 with H_k = 1 - conj(tau_k) v_k v_k^dag: accumulate from the right */
    qo_c M[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { M[i][j].re = (i == j); M[i][j].im = 0; }
    for (int k = 2; k >= 0; --k) {
        qo_c h = {tau[k].re, -tau[k].im};
        for (int j = 0; j < 3; ++j) {
            qo_c s = {0, 0};
            for (int i = 0; i < 3; ++i) {
                s.re += v[k][i].re * M[i][j].re + v[k][i].im * M[i][j].im;
                s.im += v[k][i].re * M[i][j].im - v[k][i].im * M[i][j].re;
            }
            qo_c hs = {h.re * s.re - h.im * s.im, h.re * s.im + h.im * s.re};
            for (int i = 0; i < 3; ++i) {
                M[i][j].re -= v[k][i].re * hs.re - v[k][i].im * hs.im;
                M[i][j].im -= v[k][i].re * hs.im + v[k][i].im * hs.re;
            }
        }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { RE(Q, i, j) = M[i][j].re; IM(Q, i, j) = M[i][j].im; }
    /* det and principal cube root */
    double dre = 0, dim = 0;
    {
        qo_c (*m)[3] = M;
#define CM(a, b, r, i) do { r = (a).re * (b).re - (a).im * (b).im; i = (a).re * (b).im + (a).im * (b).re; } while (0)
        double r1, i1, r2, i2;
        qo_c c0, c1, c2;
        CM(m[1][1], m[2][2], r1, i1); CM(m[1][2], m[2][1], r2, i2); c0.re = r1 - r2; c0.im = i1 - i2;
        CM(m[1][2], m[2][0], r1, i1); CM(m[1][0], m[2][2], r2, i2); c1.re = r1 - r2; c1.im = i1 - i2;
        CM(m[1][0], m[2][1], r1, i1); CM(m[1][1], m[2][0], r2, i2); c2.re = r1 - r2; c2.im = i1 - i2;
        CM(m[0][0], c0, r1, i1); dre += r1; dim += i1;
        CM(m[0][1], c1, r1, i1); dre += r1; dim += i1;
        CM(m[0][2], c2, r1, i1); dre += r1; dim += i1;
#undef CM
    }
    double mod = pow(sqrt(dre * dre + dim * dim), 1.0 / 3.0);
    double ph = atan2(dim, dre) / 3.0;
    /* divide by mod*e^{i ph} */
    double cr = cos(ph) / mod, ci = -sin(ph) / mod;
    for (int k = 0; k < 9; ++k) {
        double re = Q[2 * k] * cr - Q[2 * k + 1] * ci;
        double im = Q[2 * k] * ci + Q[2 * k + 1] * cr;
        Q[2 * k] = re;
        Q[2 * k + 1] = im;
    }
}

/* the clang / IntelLLVM draw order (the reference cluster's build) */
void qo_random_su3_ref(qo_mt64* g, double* Q) { qo_random_su3_ref_order(g, Q, 0); }

void qo_lat_hot_ref(qo_lat* l, qo_mt64* g) {
    /* interior sites in t,z,y,x order, mu inner (src/gauge/GaugeField.cpp:10-20) */
    for (size_t s = 0; s < l->V; ++s)
        for (int mu = 0; mu < 4; ++mu) qo_random_su3_ref(g, LNK(l, s, mu));
}

/* Device hot start restated: 18 Box-Muller Gaussians from Philox
 * ctr = {lex_site, mu, 1, call}, modified Gram-Schmidt on the columns
 * (Haar on U(3)), then multiplication by det^(-1/3) (principal branch). */
void qo_hot_link_philox(uint64_t seed, uint32_t site, uint32_t mu, double* Q) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    qo_haar_su3_philox(key, site, mu, 1u, Q);
}

/* The same generator on counters {c0, c1, c2, call 0..8}. */
void qo_haar_su3_philox(const uint32_t key[2], uint32_t c0, uint32_t c1, uint32_t c2, double* Q) {
    double gz[18];
    for (uint32_t c = 0; c < 9; ++c) {
        uint32_t ctr[4] = {c0, c1, c2, c}, o[4];
        qo_philox4x32_10(ctr, key, o);
        double u1 = qo_u53(o[0], o[1]), u2 = qo_u53(o[2], o[3]);
        double r = sqrt(-2.0 * log(u1));
        gz[2 * c] = r * cos(2.0 * M_PI * u2);
        gz[2 * c + 1] = r * sin(2.0 * M_PI * u2);
    }
    /* columns: col j entry i = gz[2*(3i+j)], gz[2*(3i+j)+1] */
    qo_c col[3][3];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) { col[j][i].re = gz[2 * (3 * i + j)]; col[j][i].im = gz[2 * (3 * i + j) + 1]; }
    for (int j = 0; j < 3; ++j) {
        for (int p = 0; p < j; ++p) {
            double dr = 0, di = 0; /* <col_p, col_j> */
            for (int i = 0; i < 3; ++i) {
                dr += col[p][i].re * col[j][i].re + col[p][i].im * col[j][i].im;
                di += col[p][i].re * col[j][i].im - col[p][i].im * col[j][i].re;
            }
            for (int i = 0; i < 3; ++i) {
                col[j][i].re -= col[p][i].re * dr - col[p][i].im * di;
                col[j][i].im -= col[p][i].re * di + col[p][i].im * dr;
            }
        }
        double n = 0;
        for (int i = 0; i < 3; ++i) n += col[j][i].re * col[j][i].re + col[j][i].im * col[j][i].im;
        n = 1.0 / sqrt(n);
        for (int i = 0; i < 3; ++i) { col[j][i].re *= n; col[j][i].im *= n; }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { RE(Q, i, j) = col[j][i].re; IM(Q, i, j) = col[j][i].im; }
    /* det by cofactors along row 0 */
    double m[18];
    memcpy(m, Q, sizeof m);
    double a11r = RE(m, 1, 1), a11i = IM(m, 1, 1), a22r = RE(m, 2, 2), a22i = IM(m, 2, 2);
    double a12r = RE(m, 1, 2), a12i = IM(m, 1, 2), a21r = RE(m, 2, 1), a21i = IM(m, 2, 1);
    double a10r = RE(m, 1, 0), a10i = IM(m, 1, 0), a20r = RE(m, 2, 0), a20i = IM(m, 2, 0);
    double c0r = (a11r * a22r - a11i * a22i) - (a12r * a21r - a12i * a21i);
    double c0i = (a11r * a22i + a11i * a22r) - (a12r * a21i + a12i * a21r);
    double c1r = (a12r * a20r - a12i * a20i) - (a10r * a22r - a10i * a22i);
    double c1i = (a12r * a20i + a12i * a20r) - (a10r * a22i + a10i * a22r);
    double c2r = (a10r * a21r - a10i * a21i) - (a11r * a20r - a11i * a20i);
    double c2i = (a10r * a21i + a10i * a21r) - (a11r * a20i + a11i * a20r);
    double dre = RE(m, 0, 0) * c0r - IM(m, 0, 0) * c0i + RE(m, 0, 1) * c1r - IM(m, 0, 1) * c1i +
                 RE(m, 0, 2) * c2r - IM(m, 0, 2) * c2i;
    double dim = RE(m, 0, 0) * c0i + IM(m, 0, 0) * c0r + RE(m, 0, 1) * c1i + IM(m, 0, 1) * c1r +
                 RE(m, 0, 2) * c2i + IM(m, 0, 2) * c2r;
    double ph = -atan2(dim, dre) / 3.0;
    double cr = cos(ph), ci = sin(ph);
    for (int k = 0; k < 9; ++k) {
        double re = Q[2 * k] * cr - Q[2 * k + 1] * ci;
        double im = Q[2 * k] * ci + Q[2 * k + 1] * cr;
        Q[2 * k] = re;
        Q[2 * k + 1] = im;
    }
}

void qo_lat_hot_philox(qo_lat* l, uint64_t seed) {
    for (size_t s = 0; s < l->V; ++s)
        for (int mu = 0; mu < 4; ++mu) qo_hot_link_philox(seed, (uint32_t)s, (uint32_t)mu, LNK(l, s, mu));
}

/* ---------- staple, plaquette, clover ---------- */

static inline void m3_acc(double* acc, const double* a) { for (int k = 0; k < 18; ++k) acc[k] += a[k]; }

/* Sum over nu != mu of U_nu(x+mu) [U_nu(x) U_mu(x+nu)]^dag (forward) and
 * [U_mu(x-nu) U_nu(x+mu-nu)]^dag U_nu(x-nu) (backward), nu ascending,
 * forward before backward (src/gauge/GaugeField.cpp:74-97). */
void qo_staple(const qo_lat* l, size_t x, int mu, double* st) {
    double t1[18], t2[18];
    memset(st, 0, 18 * sizeof(double));
    size_t xmu = nb(l, x, mu, +1);
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        size_t xnu = nb(l, x, nu, +1);
        qo_m3_mul(LNK(l, x, nu), LNK(l, xnu, mu), t1);
        qo_m3_dag(t1, t1);
        qo_m3_mul(LNK(l, xmu, nu), t1, t2);
        m3_acc(st, t2);
        size_t xmunu = nb(l, xmu, nu, -1);
        size_t xmnu = nb(l, x, nu, -1);
        qo_m3_mul(LNK(l, xmnu, mu), LNK(l, xmunu, nu), t1);
        qo_m3_dag(t1, t1);
        qo_m3_mul(t1, LNK(l, xmnu, nu), t2);
        m3_acc(st, t2);
    }
}

/* <P> = sum_{x, mu<nu} Re Tr(U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag)/3 / (6V)
 * (src/observables/observables_mpi.cpp:8-43), sequential sum in t,z,y,x,mu,nu order. */
double qo_plaquette(const qo_lat* l) {
    /* the terms are formed in parallel and summed in the sequential order */
    double* term = (double*)malloc(l->V * 6 * sizeof(double));
#pragma omp parallel for schedule(static)
    for (size_t s = 0; s < l->V; ++s) {
        double a[18], b[18], c[18];
        int k = 0;
        for (int mu = 0; mu < 4; ++mu)
            for (int nu = mu + 1; nu < 4; ++nu) {
                qo_m3_mul(LNK(l, s, mu), LNK(l, nb(l, s, mu, +1), nu), a);
                qo_m3_dag(LNK(l, nb(l, s, nu, +1), mu), b);
                qo_m3_mul(a, b, c);
                qo_m3_dag(LNK(l, s, nu), b);
                qo_m3_mul(c, b, a);
                term[s * 6 + k++] = (a[0] + a[8] + a[16]) / 3.0;
            }
    }
    double sum = 0.0;
    for (size_t k = 0; k < l->V * 6; ++k) sum += term[k];
    free(term);
    return sum / (6.0 * (double)l->V);
}

static void mul4(const double* A, int da, const double* B, int db, const double* C, int dc,
                 const double* D, int dd, double* out) {
    double a[18], b[18], t[18];
    if (da) qo_m3_dag(A, a); else memcpy(a, A, sizeof a);
    if (db) qo_m3_dag(B, b); else memcpy(b, B, sizeof b);
    qo_m3_mul(a, b, t);
    if (dc) qo_m3_dag(C, b); else memcpy(b, C, sizeof b);
    qo_m3_mul(t, b, a);
    if (dd) qo_m3_dag(D, b); else memcpy(b, D, sizeof b);
    qo_m3_mul(a, b, out);
}

/* Four-leaf clover, F = (C - C^dag)/8, not made traceless
 * (src/observables/observables_mpi.cpp:61-85). */
void qo_clover_F(const qo_lat* l, size_t x, int mu, int nu, double* F) {
    size_t xpmu = nb(l, x, mu, +1), xpnu = nb(l, x, nu, +1);
    size_t xmmu = nb(l, x, mu, -1), xmnu = nb(l, x, nu, -1);
    size_t xmmupnu = nb(l, xmmu, nu, +1), xmmumnu = nb(l, xmmu, nu, -1), xpmumnu = nb(l, xpmu, nu, -1);
    double C[18] = {0}, t[18];
    mul4(LNK(l, x, mu), 0, LNK(l, xpmu, nu), 0, LNK(l, xpnu, mu), 1, LNK(l, x, nu), 1, t); m3_acc(C, t);
    mul4(LNK(l, x, nu), 0, LNK(l, xmmupnu, mu), 1, LNK(l, xmmu, nu), 1, LNK(l, xmmu, mu), 0, t); m3_acc(C, t);
    mul4(LNK(l, xmmu, mu), 1, LNK(l, xmmumnu, nu), 1, LNK(l, xmmumnu, mu), 0, LNK(l, xmnu, nu), 0, t); m3_acc(C, t);
    mul4(LNK(l, xmnu, nu), 1, LNK(l, xmnu, mu), 0, LNK(l, xpmumnu, nu), 0, LNK(l, x, mu), 1, t); m3_acc(C, t);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            RE(F, i, j) = 0.125 * (RE(C, i, j) - RE(C, j, i));
            IM(F, i, j) = 0.125 * (IM(C, i, j) + IM(C, j, i));
        }
}

static double retr_mul(const double* a, const double* b) {
    double s = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) s += RE(a, i, k) * RE(b, k, i) - IM(a, i, k) * IM(b, k, i);
    return s;
}
static double sqnorm(const double* a) {
    double s = 0.0;
    for (int k = 0; k < 18; ++k) s += a[k] * a[k];
    return s;
}

static void qe_site(const qo_lat* l, size_t s, double* q, double* e) {
    double F01[18], F02[18], F03[18], F12[18], F13[18], F23[18];
    qo_clover_F(l, s, 0, 1, F01); qo_clover_F(l, s, 0, 2, F02); qo_clover_F(l, s, 0, 3, F03);
    qo_clover_F(l, s, 1, 2, F12); qo_clover_F(l, s, 1, 3, F13); qo_clover_F(l, s, 2, 3, F23);
    double qq = retr_mul(F01, F23) - retr_mul(F02, F13) + retr_mul(F03, F12);
    *e = 0.5 * (sqnorm(F01) + sqnorm(F02) + sqnorm(F03) + sqnorm(F12) + sqnorm(F13) + sqnorm(F23));
    *q = qq * (1.0 / (4.0 * M_PI * M_PI));
}

void qo_clover_qe_site(const qo_lat* l, double* q_site, double* e_site) {
#pragma omp parallel for schedule(static)
    for (size_t s = 0; s < l->V; ++s) qe_site(l, s, &q_site[s], &e_site[s]);
}

/* (Q, E/V) with Q = sum_x q_x, E = sum_x e_x (src/observables/observables_mpi.cpp:88-150);
 * site densities in parallel, sums in the sequential site order. */
void qo_clover_qe(const qo_lat* l, double* Q, double* E_mean) {
    double* qs = (double*)malloc(l->V * 2 * sizeof(double));
    double* es = qs + l->V;
    qo_clover_qe_site(l, qs, es);
    double q = 0.0, e = 0.0;
    for (size_t s = 0; s < l->V; ++s) {
        q += qs[s];
        e += es[s];
    }
    free(qs);
    *Q = q;
    *E_mean = e / (double)l->V;
}

/* ---------- reunitarisation, translation, gauge transformation ---------- */

void qo_reunitarize(qo_lat* l) {
    for (size_t s = 0; s < l->V; ++s)
        for (int mu = 0; mu < 4; ++mu) qo_reunitarize_link(LNK(l, s, mu));
}

/* U'(x) = U(x - s), periodic: one-rank shift_field (src/mpi/Shift.cpp:109-160). */
void qo_translate(qo_lat* l, const int sh[4]) {
    double* out = (double*)malloc(l->V * 72 * sizeof(double));
    for (size_t s = 0; s < l->V; ++s) {
        int c[4];
        coords(l, s, c);
        for (int d = 0; d < 4; ++d) c[d] = ((c[d] - sh[d]) % l->L[d] + l->L[d]) % l->L[d];
        memcpy(out + s * 72, l->U + lex(l, c) * 72, 72 * sizeof(double));
    }
    memcpy(l->U, out, l->V * 72 * sizeof(double));
    free(out);
}

/* U_mu(x) -> G(x) U_mu(x) G(x+mu)^dag (src/gauge/GaugeField.cpp:100-114) with G given. */
void qo_gauge_transform(qo_lat* l, const double* G) {
    double* out = (double*)malloc(l->V * 72 * sizeof(double));
    double t[18], gd[18];
    for (size_t s = 0; s < l->V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            qo_m3_mul(G + s * 18, LNK(l, s, mu), t);
            qo_m3_dag(G + nb(l, s, mu, +1) * 18, gd);
            qo_m3_mul(t, gd, out + (s * 4 + mu) * 18);
        }
    memcpy(l->U, out, l->V * 72 * sizeof(double));
    free(out);
}

/* ---------- gradient flow ---------- */

/* Z = P_TA(-U Sigma) for every link (src/flow/gradient_flow.cpp:15-34). */
static void compute_force(const qo_lat* l, double* F) {
#pragma omp parallel for schedule(static)
    for (size_t s = 0; s < l->V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            double st[18], t[18];
            qo_staple(l, s, mu, st);
            qo_m3_mul(LNK(l, s, mu), st, t);
            for (int k = 0; k < 18; ++k) t[k] = -t[k];
            qo_proj_lie(t);
            memcpy(F + (s * 4 + mu) * 18, t, sizeof t);
        }
}

/* Luescher RK3: W1 = e^{eps/4 Z0} W0; W2 = e^{eps(8/9 Z1 - 17/36 Z0)} W1;
 * W3 = e^{eps(3/4 Z2 - (8/9 Z1 - 17/36 Z0))} W2 (src/flow/gradient_flow.cpp:37-128). */
void qo_flow_rk3_step(qo_lat* l, double eps) {
    size_t n = l->V * 72;
    double* F0 = (double*)malloc(n * sizeof(double));
    double* F1 = (double*)malloc(n * sizeof(double));
    /* every stage reads the forces of the whole field before any link moves,
     * so links are updated in parallel */
    compute_force(l, F0);
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < l->V * 4; ++k) {
        double E[18];
        qo_exp_luscher(F0 + k * 18, 0.25 * eps, E);
        qo_m3_mul(E, l->U + k * 18, l->U + k * 18);
    }
    compute_force(l, F1);
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < l->V * 4; ++k) {
        double Z[18], E[18];
        for (int c = 0; c < 18; ++c)
            Z[c] = (8.0 / 9.0) * eps * F1[k * 18 + c] - (17.0 / 36.0) * eps * F0[k * 18 + c];
        qo_exp_luscher(Z, 1.0, E);
        qo_m3_mul(E, l->U + k * 18, l->U + k * 18);
    }
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < n; ++k) F0[k] = 8.0 / 9.0 * F1[k] - (17.0 / 36.0) * F0[k];
    compute_force(l, F1);
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < l->V * 4; ++k) {
        double Z[18], E[18];
        for (int c = 0; c < 18; ++c) Z[c] = 0.75 * eps * F1[k * 18 + c] - eps * F0[k * 18 + c];
        qo_exp_luscher(Z, 1.0, E);
        qo_m3_mul(E, l->U + k * 18, l->U + k * 18);
    }
    free(F0);
    free(F1);
}

/* Measure-then-step schedule (src/observables/observables_mpi.cpp:153-178):
 * for each of n_steps_gf points record (t, Q, E/V) and then advance
 * n_rk_steps RK3 steps. The input lattice is left unchanged. */
void qo_topo_charge_flowed(const qo_lat* l, double eps, int n_steps_gf, int n_rk_steps, double* tQE) {
    qo_lat c;
    qo_lat_init(&c, l->L);
    memcpy(c.U, l->U, l->V * 72 * sizeof(double));
    double t = 0.0;
    for (int s = 0; s < n_steps_gf; ++s) {
        double Q, E;
        qo_clover_qe(&c, &Q, &E);
        tQE[3 * s] = t;
        tQE[3 * s + 1] = Q;
        tQE[3 * s + 2] = E;
        for (int r = 0; r < n_rk_steps; ++r) {
            qo_flow_rk3_step(&c, eps);
            t += eps;
        }
    }
    qo_lat_free(&c);
}

/* ---------- heatbath ---------- */

/* Cabibbo-Marinari hit with a precomputed staple A; subgroups (0,1),(1,2),(0,2)
 * in that order, U <- R U each time (src/heatbath/heatbath_mpi.cpp:11-34).
 * kp(ctx, beta, k, out) returns 0 on success, -1 when the budget is exhausted. */
typedef int (*kp_fn)(void* ctx, int g, double beta, double k, double out[4]);

static void cm_hit(double* U, const double* A, double beta, kp_fn kp, void* ctx) {
    static const int SI[3] = {0, 1, 0}, SJ[3] = {1, 2, 2};
    double W[18], w[8], q[4], v[4], x[4], r[4], R[18];
    for (int g = 0; g < 3; ++g) {
        int i = SI[g], j = SJ[g];
        qo_m3_mul(U, A, W);
        w[0] = RE(W, i, i); w[1] = IM(W, i, i);
        w[2] = RE(W, i, j); w[3] = IM(W, i, j);
        w[4] = RE(W, j, i); w[5] = IM(W, j, i);
        w[6] = RE(W, j, j); w[7] = IM(W, j, j);
        qo_su2_to_q(w, q);
        /* su2_step (src/heatbath/heatbath.cpp:39-44): k = |w|, v = w/k, r = x v^dag */
        double k = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (int c = 0; c < 4; ++c) v[c] = q[c] / k;
        if (kp(ctx, g, beta, k, x) != 0) continue;
        double vd[4] = {v[0], -v[1], -v[2], -v[3]};
        qo_qmult(x, vd, r);
        qo_su2q_to_su3(r, i, j, R);
        qo_m3_mul(R, U, U);
    }
}

/* reference mode: unbounded draws from mt19937_64 */
static int kp_mt_cb(void* ctx, int g, double beta, double k, double out[4]) {
    (void)g;
    qo_kp_mt(beta, k, (qo_mt64*)ctx, out);
    return 0;
}

/* Frozen links of the one-rank block: coordinate 0 (the reference's x_d == 1)
 * in any direction d != mu (src/geometry/GeometryCB.cpp:33-56). */
long qo_hb_sweep_ref(qo_lat* l, double beta, int n_hits, qo_mt64* g) {
    double A[18];
    long visits = 0;
    for (size_t s = 0; s < l->V; ++s) {
        int c[4];
        coords(l, s, c);
        for (int mu = 0; mu < 4; ++mu) {
            int frozen = 0;
            for (int d = 0; d < 4; ++d)
                if (d != mu && c[d] == 0) frozen = 1;
            if (frozen) continue;
            ++visits;
            for (int h = 0; h < n_hits; ++h) {
                qo_staple(l, s, mu, A);
                cm_hit(LNK(l, s, mu), A, beta, kp_mt_cb, g);
            }
        }
    }
    return visits;
}

typedef struct { const double* u; size_t Vh; int K; size_t cb; size_t base_hs; int n_hits; int h; long exhausted; } inj_ctx;

static int kp_inj_cb(void* vctx, int g, double beta, double k, double out[4]) {
    inj_ctx* c = (inj_ctx*)vctx;
    int nslot = 4 * c->K + 2;
    double u[4 * 64 + 2];
    size_t base = (((c->base_hs * (size_t)c->n_hits + (size_t)c->h) * 3 + (size_t)g) * (size_t)nslot) * c->Vh + c->cb;
    for (int s = 0; s < nslot; ++s) u[s] = c->u[base + (size_t)s * c->Vh];
    if (qo_kp_injected(beta, k, u, c->K, out) < 0) {
        c->exhausted++;
        return -1;
    }
    return 0;
}

/* One Cabibbo-Marinari hit (src/heatbath/heatbath_mpi.cpp:11-34) on U with
 * staple A and injected uniforms u[g*(4K+2) + slot] for subgroup g (the
 * layout of the device probe QCB_PROBE_CM_HIT).  A subgroup whose K
 * attempts all reject is skipped; returns how many were. */
int qo_cm_hit_injected(double* U, const double* A, double beta, const double* u, int K) {
    if (K < 1 || K > 64) return -1;
    inj_ctx ctx = {u, 1, K, 0, 0, 1, 0, 0};
    cm_hit(U, A, beta, kp_inj_cb, &ctx);
    return (int)ctx.exhausted;
}

long qo_hb_sweep_cb_injected(qo_lat* l, double beta, int n_hits, const double* u, int K) {
    if (K < 1 || K > 64) return -1;
    double A[18];
    inj_ctx ctx = {u, l->V / 2, K, 0, 0, n_hits, 0, 0};
    for (int mu = 0; mu < 4; ++mu)
        for (int p = 0; p < 2; ++p) {
            ctx.base_hs = (size_t)(2 * mu + p);
            for (size_t s = 0; s < l->V; ++s) {
                int c[4];
                coords(l, s, c);
                if (((c[0] + c[1] + c[2] + c[3]) & 1) != p) continue;
                ctx.cb = s / 2;
                qo_staple(l, s, mu, A);
                for (int h = 0; h < n_hits; ++h) {
                    ctx.h = h;
                    cm_hit(LNK(l, s, mu), A, beta, kp_inj_cb, &ctx);
                }
            }
        }
    return ctx.exhausted;
}

typedef struct { uint32_t key[2]; uint32_t cb; uint32_t sweep; uint32_t hs; int h; } phx_ctx;

static int kp_phx_cb(void* vctx, int g, double beta, double k, double out[4]) {
    phx_ctx* c = (phx_ctx*)vctx;
    double alpha = beta * k * 2.0 / 3.0, a0;
    uint32_t c2 = (c->hs << 24) | ((uint32_t)c->h << 8) | ((uint32_t)g << 4);
    uint32_t a;
    for (a = 0; a < 1024; ++a) {   /* same attempt cap as the device kernel */
        uint32_t ctr[4] = {c->cb, c->sweep, c2, a}, o[4];
        qo_philox4x32_10(ctr, c->key, o);
        if (qo_kp_try(alpha, qo_u32(o[0]), qo_u32(o[1]), qo_u32(o[2]), qo_u32(o[3]), &a0)) break;
    }
    if (a == 1024) return -1;
    /* sphere: subgroup m = 3h + g reads words 2(m&1), 2(m&1)+1 of the call
     * 0x80000000 at the counter of subgroup m - (m&1) (one call per two subgroups) */
    int m = 3 * c->h + g, m0 = m & ~1, wsel = 2 * (m & 1);
    uint32_t cs2 = (c->hs << 24) | ((uint32_t)(m0 / 3) << 8) | ((uint32_t)(m0 % 3) << 4);
    uint32_t cs[4] = {c->cb, c->sweep, cs2, 0x80000000u}, os[4];
    qo_philox4x32_10(cs, c->key, os);
    qo_kp_sphere(a0, qo_u32(os[wsel]), qo_u32(os[wsel + 1]), out);
    return 0;
}

void qo_hb_sweep_cb_philox(qo_lat* l, double beta, int n_hits, uint64_t seed, uint32_t sweep_idx) {
    double A[18];
    phx_ctx ctx;
    ctx.key[0] = (uint32_t)seed;
    ctx.key[1] = (uint32_t)(seed >> 32);
    ctx.sweep = sweep_idx;
    for (int mu = 0; mu < 4; ++mu)
        for (int p = 0; p < 2; ++p) {
            ctx.hs = (uint32_t)(2 * mu + p);
            for (size_t s = 0; s < l->V; ++s) {
                int c[4];
                coords(l, s, c);
                if (((c[0] + c[1] + c[2] + c[3]) & 1) != p) continue;
                ctx.cb = (uint32_t)(s / 2);
                qo_staple(l, s, mu, A);
                for (int h = 0; h < n_hits; ++h) {
                    ctx.h = h;
                    cm_hit(LNK(l, s, mu), A, beta, kp_phx_cb, &ctx);
                }
            }
        }
}
