/*
 * qo_ecmc.c -- FP64 CPU restatement of the reference's event-chain Monte
 * Carlo (src/ecmc/ecmc_mpi_cb.cpp, src/ecmc/ecmc_mpi_cb.h, the matrix set of
 * src/su3/utils.cpp:66-127).  TEST INFRASTRUCTURE ONLY (see qo.h).
 *
 * One chain per sub-block of b[0] x b[1] x b[2] x b[3] sites of the periodic
 * lattice, the reference's per-rank block: a link (x, mu) is frozen when
 * x_d mod b_d == 0 for some d != mu (src/geometry/GeometryCB.cpp:33-56 with
 * the reference's 1-based "x == 1" face).  Non-frozen links of one block
 * only ever read their own block's links or frozen ones, so the chains are
 * independent.  Two sources of randomness:
 *   - reference mode: std::mt19937_64 and the libstdc++ distributions in the
 *     reference's draw order (one chain, b = L: the one-rank run);
 *   - Philox mode: the counter scheme of the device kernel (qo.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "qo.h"

#define RE(a, i, j) (a)[((i)*3 + (j))*2]
#define IM(a, i, j) (a)[((i)*3 + (j))*2 + 1]
#define LNK(l, s, mu) ((l)->U + ((s)*4 + (size_t)(mu)) * 18)
#define SET_SIZE 10001

static void m3_dag(const double* a, double* out) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            RE(out, i, j) = RE(a, j, i);
            IM(out, i, j) = -IM(a, j, i);
        }
}

static void m3_id(double* a) {
    memset(a, 0, 18 * sizeof(double));
    a[0] = a[8] = a[16] = 1.0;
}

static int dsign(double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }   /* src/su3/utils.h:36-41 */

/* ---------- pieces ---------- */

/* solve_reject_fast (src/ecmc/ecmc_mpi_cb.h:56-94).  Pinned by the defining
 * integral: the cumulative positive part of dS/dtheta = -A sin + eps B cos
 * from 0 to the result equals gamma (tests/test_oracle_ecmc.py). */
double qo_ecmc_reject_fast(double A, double B, double gamma, int epsilon) {
    B = (epsilon == -1) ? -B : B;
    double R = hypot(A, B);
    double invR = 1.0 / R;
    double period = 2.0 * R;
    double discarded = floor(gamma / period);
    gamma -= discarded * period;
    double phi = atan2(-A, B);
    phi += (phi < 0.0) ? (2.0 * M_PI) : 0.0;
    double alpha;
    double p1 = R - A;
    if (phi < (M_PI * 0.5) || phi > (M_PI * 1.5))
        alpha = (gamma > p1) ? (gamma - p1) * invR - 1.0 : (gamma + A) * invR;
    else
        alpha = gamma * invR - 1.0;
    alpha = (alpha > 1.0) ? 1.0 : ((alpha < -1.0) ? -1.0 : alpha);
    double theta = phi + asin(alpha);
    theta += (theta < 0.0) ? (2.0 * M_PI) : 0.0;
    theta -= (theta >= 2.0 * M_PI) ? (2.0 * M_PI) : 0.0;
    return theta + 2.0 * M_PI * discarded;
}

/* solve_reject, the explicit-interval form (src/ecmc/ecmc_mpi_cb.cpp:82-178). */
double qo_ecmc_reject_exact(double A, double B, double gamma, int epsilon) {
    if (epsilon == -1) B = -B;
    double R = sqrt(A * A + B * B);
    double phi = atan2(-A, B);
    if (phi < 0) phi += 2 * M_PI;
    double iv[4] = {0.0, 0.0, 2 * M_PI, 2 * M_PI};
    double reject = 0.0, period, p1, p2;
    int discarded = 0;
    if (phi < M_PI / 2.0 || phi > 3 * M_PI / 2.0) {
        if (phi < M_PI / 2.0) {
            iv[1] = M_PI / 2.0 + phi;
            iv[2] = 3 * M_PI / 2.0 + phi;
        } else {
            iv[1] = -3 * M_PI / 2.0 + phi;
            iv[2] = -M_PI / 2.0 + phi;
        }
        p1 = R * (sin(iv[1] - phi) - sin(iv[0] - phi));
        p2 = R * (sin(iv[3] - phi) - sin(iv[2] - phi));
        period = p1 + p2;
        discarded = (int)floor(gamma / period);
        gamma = gamma - floor(gamma / period) * period;
        int second = gamma > p1;
        if (second) gamma -= p1;
        double lo = second ? iv[2] : iv[0], hi = second ? iv[3] : iv[1];
        double alpha = gamma / R + sin(lo - phi);
        double t1 = fmod(phi + asin(alpha) + 2 * M_PI, 2 * M_PI);
        double t2 = fmod(phi + M_PI - asin(alpha) + 2 * M_PI, 2 * M_PI);
        reject = (t1 < hi && t1 > lo) ? t1 : t2;
    } else {
        iv[0] = -M_PI / 2.0 + phi;
        iv[1] = M_PI / 2.0 + phi;
        period = R * (sin(iv[1] - phi) - sin(iv[0] - phi));
        discarded = (int)floor(gamma / period);
        gamma = gamma - floor(gamma / period) * period;
        double alpha = gamma / R + sin(iv[0] - phi);
        double t1 = fmod(phi + asin(alpha) + 2 * M_PI, 2 * M_PI);
        double t2 = fmod(phi + M_PI - asin(alpha) + 2 * M_PI, 2 * M_PI);
        reject = (t1 < iv[1] && t1 > iv[0]) ? t1 : t2;
    }
    return reject + 2 * M_PI * discarded;
}

/* compute_ds (ecmc_mpi_cb.cpp:263-275): Im[(R^dag Pi R)_00 - (R^dag Pi R)_11].
 * Pinned by the direct matrix product in numpy. */
double qo_ecmc_compute_ds(const double* Pi, const double* R) {
    double M[18];
    qo_m3_mul(Pi, R, M);
    double res_im = 0.0;
    for (int k = 0; k < 3; ++k) {
        /* conj(R_k0) M_k0 - conj(R_k1) M_k1, imaginary part */
        res_im += RE(R, k, 0) * IM(M, k, 0) - IM(R, k, 0) * RE(M, k, 0);
        res_im -= RE(R, k, 1) * IM(M, k, 1) - IM(R, k, 1) * RE(M, k, 1);
    }
    return res_im;
}

/* update (ecmc_mpi_cb.cpp:357-372): U <- R diag(e^{i xi}, e^{-i xi}, 1) R^dag U,
 * xi = eps theta.  Pinned by the direct product in numpy. */
void qo_ecmc_update(double* U, double theta, int epsilon, const double* R) {
    double Rd[18], T[18];
    m3_dag(R, Rd);
    qo_m3_mul(Rd, U, T);
    double xi = epsilon * theta, c = cos(xi), s = sin(xi);
    for (int j = 0; j < 3; ++j) {
        double a = RE(T, 0, j), b = IM(T, 0, j);
        RE(T, 0, j) = a * c - b * s;
        IM(T, 0, j) = a * s + b * c;
        a = RE(T, 1, j);
        b = IM(T, 1, j);
        RE(T, 1, j) = a * c + b * s;
        IM(T, 1, j) = -a * s + b * c;
    }
    qo_m3_mul(R, T, U);
}

/* random_SU3_epsilon (src/su3/utils.cpp:66-103) from nine uniforms in
 * (-1/2, 1/2): M = X01 X02 X12, each X the SU(2) element
 * (sqrt(1 - eps^2), eps r / |r|) embedded in that block. */
void qo_ecmc_random_su3_eps(double eps, const double r[9], double* M) {
    static const int bi[3] = {0, 0, 1}, bj[3] = {1, 2, 2};
    double X[18], T[18];
    m3_id(M);
    for (int f = 0; f < 3; ++f) {
        const double* q = r + 3 * f;
        double norm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        double x[4] = {sqrt(1 - eps * eps), eps * q[0] / norm, eps * q[1] / norm, eps * q[2] / norm};
        qo_su2q_to_su3(x, bi[f], bj[f], X);
        qo_m3_mul(M, X, T);
        memcpy(M, T, sizeof T);
    }
}

/* ---------- geometry of the block decomposition ---------- */

static void gcoords(const qo_lat* l, size_t s, int c[4]) {
    c[0] = (int)(s % (size_t)l->L[0]);
    s /= (size_t)l->L[0];
    c[1] = (int)(s % (size_t)l->L[1]);
    s /= (size_t)l->L[1];
    c[2] = (int)(s % (size_t)l->L[2]);
    c[3] = (int)(s / (size_t)l->L[2]);
}

static size_t glex(const qo_lat* l, const int c[4]) {
    return (size_t)c[0] + (size_t)l->L[0] * ((size_t)c[1] + (size_t)l->L[1] * ((size_t)c[2] + (size_t)l->L[2] * (size_t)c[3]));
}

static size_t gshift(const qo_lat* l, size_t s, int d1, int s1, int d2, int s2) {
    int c[4];
    gcoords(l, s, c);
    if (s1) c[d1] = (c[d1] + s1 + l->L[d1]) % l->L[d1];
    if (s2) c[d2] = (c[d2] + s2 + l->L[d2]) % l->L[d2];
    return glex(l, c);
}

static int frozen(const qo_lat* l, const int b[4], size_t s, int mu) {
    int c[4];
    gcoords(l, s, c);
    for (int d = 0; d < 4; ++d)
        if (d != mu && c[d] % b[d] == 0) return 1;
    return 0;
}

/* staple j of link (x, mu): nu = j/2-th direction != mu ascending; j even the
 * forward staple U_nu(x+mu) [U_nu(x) U_mu(x+nu)]^dag, j odd the backward one
 * [U_mu(x-nu) U_nu(x+mu-nu)]^dag U_nu(x-nu) (ecmc_mpi_cb.cpp:12-36). */
static int dir_of(int mu, int j) {
    int k = j / 2, n = 0;
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        if (n++ == k) return nu;
    }
    return -1;
}

static void staple_j(const qo_lat* l, size_t x, int mu, int j, double* S) {
    int nu = dir_of(mu, j);
    double t[18], td[18];
    if ((j & 1) == 0) {
        qo_m3_mul(LNK(l, x, nu), LNK(l, gshift(l, x, nu, 1, 0, 0), mu), t);
        m3_dag(t, td);
        qo_m3_mul(LNK(l, gshift(l, x, mu, 1, 0, 0), nu), td, S);
    } else {
        size_t xmnu = gshift(l, x, nu, -1, 0, 0);
        qo_m3_mul(LNK(l, xmnu, mu), LNK(l, gshift(l, x, mu, 1, nu, -1), nu), t);
        m3_dag(t, td);
        qo_m3_mul(td, LNK(l, xmnu, nu), S);
    }
}

/* the three other links of plaquette j (GeometryCB.cpp:60-90) */
static void plaquette_links(const qo_lat* l, size_t x, int mu, int j, size_t site[4], int dir[4]) {
    int nu = dir_of(mu, j);
    site[0] = x;
    dir[0] = mu;
    if ((j & 1) == 0) {
        site[1] = gshift(l, x, mu, 1, 0, 0); dir[1] = nu;
        site[2] = gshift(l, x, nu, 1, 0, 0); dir[2] = mu;
        site[3] = x;                         dir[3] = nu;
    } else {
        site[1] = gshift(l, x, mu, 1, nu, -1); dir[1] = nu;
        site[2] = gshift(l, x, nu, -1, 0, 0);  dir[2] = mu;
        site[3] = site[2];                     dir[3] = nu;
    }
}

/* ---------- randomness ---------- */

typedef struct {
    int philox;
    /* reference mode */
    qo_mt64* mt;
    double* set;          /* SET_SIZE matrices */
    /* Philox mode */
    uint32_t key[2];
    uint32_t chain;
    uint64_t n;           /* event index of the chain */
    uint64_t gen;         /* matrix-set generation */
    double eps_set;
} rng_t;

static void ph(const rng_t* r, uint64_t n, uint32_t call, uint32_t o[4]) {
    uint32_t ctr[4] = {r->chain, (uint32_t)n, 0xEC000000u | ((uint32_t)(n >> 32) & 0x00FFFFFFu), call};
    qo_philox4x32_10(ctr, r->key, o);
}

static uint32_t mulhi32(uint32_t w, uint32_t n) { return (uint32_t)(((uint64_t)w * n) >> 32); }

static uint64_t mulhi64(uint64_t x, uint64_t n) {
    /* high word of the 128-bit product */
    uint64_t xl = x & 0xffffffffu, xh = x >> 32, nl = n & 0xffffffffu, nh = n >> 32;
    uint64_t ll = xl * nl, lh = xl * nh, hl = xh * nl, hh = xh * nh;
    uint64_t mid = (ll >> 32) + (lh & 0xffffffffu) + (hl & 0xffffffffu);
    return hh + (lh >> 32) + (hl >> 32) + (mid >> 32);
}

static double exp_ref(qo_mt64* g, double mean) {   /* std::exponential_distribution(1/mean) */
    return -log(1.0 - qo_mt64_canonical(g)) / (1.0 / mean);
}

static void set_regen_ref(rng_t* r) {   /* ecmc_set (utils.cpp:119-127) */
    m3_id(r->set);
    for (int i = 1; i < SET_SIZE; i += 2) {
        double u[9];
        for (int k = 0; k < 9; ++k) u[k] = qo_mt64_canonical(r->mt) * 1.0 + (-0.5);
        qo_ecmc_random_su3_eps(r->eps_set, u, r->set + (size_t)i * 18);
        m3_dag(r->set + (size_t)i * 18, r->set + (size_t)(i + 1) * 18);
    }
}

static void set_elem_philox(const rng_t* r, uint64_t i, double* M) {
    if (i == 0) {
        m3_id(M);
        return;
    }
    uint32_t m = (uint32_t)((i - 1) / 2);
    double u[9], T[18];
    for (uint32_t k = 0; k < 3; ++k) {
        uint32_t ctr[4] = {r->chain, (uint32_t)r->gen, 0xE5000000u | m, k}, o[4];
        qo_philox4x32_10(ctr, r->key, o);
        for (int w = 0; w < 3; ++w) u[3 * k + w] = qo_u32(o[w]) - 0.5;
    }
    qo_ecmc_random_su3_eps(r->eps_set, u, T);
    if (i & 1) memcpy(M, T, sizeof T);
    else m3_dag(T, M);
}

static void draw_site(rng_t* r, const qo_lat* l, const int b[4], const int org[4], size_t* site, int* mu, int* eps,
                      double* th_rs, const qo_ecmc_params* p, int with_exp) {
    int c[4];
    if (!r->philox) {
        /* dist_refresh_site first (when Poisson), then random_site, random_dir, random_eps
         * (ecmc_mpi_cb.cpp:375-384,596-602) */
        if (with_exp && p->poisson) *th_rs = exp_ref(r->mt, p->theta_refresh_site);
        for (int d = 0; d < 4; ++d) c[d] = org[d] + (int)qo_mt64_uniform_int(r->mt, 2, (uint64_t)b[d] - 1) - 1;
        *mu = (int)qo_mt64_uniform_int(r->mt, 0, 3);
        *eps = 2 * (int)qo_mt64_uniform_int(r->mt, 0, 1) - 1;
    } else {
        uint32_t o3[4], o4[4];
        ph(r, r->n, 3, o3);
        ph(r, r->n, 4, o4);
        for (int d = 0; d < 4; ++d) c[d] = org[d] + 1 + (int)mulhi32(o3[d], (uint32_t)(b[d] - 2));
        *mu = (int)(o4[0] >> 30);
        *eps = 2 * (int)(o4[1] >> 31) - 1;
        if (with_exp && p->poisson) *th_rs = -p->theta_refresh_site * log(qo_u53(o4[2], o4[3]));
    }
    *site = glex(l, c);
}

static void draw_R(rng_t* r, double* R, double* th_rR, const qo_ecmc_params* p, int with_exp) {
    if (!r->philox) {
        if (with_exp && p->poisson) *th_rR = exp_ref(r->mt, p->theta_refresh_R);
        qo_random_su3_ref(r->mt, R);
    } else {
        uint32_t o5[4];
        ph(r, r->n, 5, o5);
        if (with_exp && p->poisson) *th_rR = -p->theta_refresh_R * log(qo_u53(o5[0], o5[1]));
        qo_haar_su3_philox(r->key, r->chain, (uint32_t)r->n, 0xED000000u | ((uint32_t)(r->n >> 32) & 0x00FFFFFFu), R);
    }
}

/* ---------- the chain ---------- */

/* lift_improved_fast (ecmc_mpi_cb.cpp:277-354).  Returns the new epsilon and
 * the chosen link through nsite and nmu; R may change. */
static int lift(const qo_lat* l, const int b[4], size_t x, int mu, int j, double* R, rng_t* r,
                const uint32_t o1[4], const uint32_t o2[4], size_t* nsite, int* nmu) {
    size_t ls[4];
    int ld[4];
    plaquette_links(l, x, mu, j, ls, ld);
    const double *U0 = LNK(l, ls[0], ld[0]), *U1 = LNK(l, ls[1], ld[1]), *U2 = LNK(l, ls[2], ld[2]),
                 *U3 = LNK(l, ls[3], ld[3]);
    int fr[4];
    for (int i = 0; i < 4; ++i) fr[i] = frozen(l, b, ls[i], ld[i]);
    double P[4][18], t1[18], t2[18];
    memset(P, 0, sizeof P);
    if ((j & 1) == 0) {
        double U01[18], U32[18], U32d[18], U01d[18];
        qo_m3_mul(U0, U1, U01);
        qo_m3_mul(U3, U2, U32);
        m3_dag(U32, U32d);
        m3_dag(U01, U01d);
        qo_m3_mul(U01, U32d, P[0]);
        if (!fr[1]) { qo_m3_mul(U1, U32d, t1); qo_m3_mul(t1, U0, P[1]); }
        if (!fr[2]) { qo_m3_mul(U2, U01d, t1); qo_m3_mul(t1, U3, P[2]); }
        if (!fr[3]) m3_dag(P[0], P[3]);
    } else {
        double U21[18], U21d[18], T[18], U0d[18], U3d[18];
        qo_m3_mul(U2, U1, U21);
        m3_dag(U21, U21d);
        qo_m3_mul(U0, U21d, T);
        qo_m3_mul(T, U3, P[0]);
        if (!fr[1]) {
            m3_dag(U0, U0d);
            m3_dag(U3, U3d);
            qo_m3_mul(U1, U0d, t1);
            qo_m3_mul(t1, U3d, t2);
            qo_m3_mul(t2, U2, P[1]);
        }
        qo_m3_mul(U3, T, P[3]);
        m3_dag(P[3], P[2]);
    }
    double prob[4], absd[4], sum = 0.0;
    int sgn[4];
    for (int i = 0; i < 4; ++i) {
        prob[i] = fr[i] ? 0.0 : qo_ecmc_compute_ds(P[i], R);
        sgn[i] = dsign(prob[i]);
        prob[i] = fabs(prob[i]);
        absd[i] = prob[i];
        sum += prob[i];
    }
    for (int i = 0; i < 4; ++i) prob[i] /= sum;
    /* selectVariable (ecmc_mpi_cb.cpp:252-260) */
    double u = r->philox ? qo_u32(o1[2]) : qo_mt64_canonical(r->mt);
    int idx = u < prob[0] ? 0 : (u < prob[0] + prob[1] ? 1 : (u < prob[0] + prob[1] + prob[2] ? 2 : 3));
    uint64_t iset;
    double M[18], Rn[18];
    if (r->philox) {
        iset = mulhi64(((uint64_t)o2[0] << 32) | o2[1], SET_SIZE);
        set_elem_philox(r, iset, M);
    } else {
        iset = qo_mt64_uniform_int(r->mt, 0, SET_SIZE - 1);
        memcpy(M, r->set + iset * 18, sizeof M);
    }
    qo_m3_mul(M, R, Rn);
    double dS = sgn[idx] * absd[idx];
    double dSn = qo_ecmc_compute_ds(P[idx], Rn);
    double neps = -sgn[idx];
    if (fabs(dS) < fabs(dSn)) {
        memcpy(R, Rn, sizeof Rn);
        neps = -dsign(dSn);
    } else {
        double ua = r->philox ? qo_u32(o1[3]) : qo_mt64_canonical(r->mt);
        if (ua < fabs(dSn) / fabs(dS)) {
            memcpy(R, Rn, sizeof Rn);
            neps = -dsign(dSn);
        }
    }
    *nsite = ls[idx];
    *nmu = ld[idx];
    return (int)neps;
}

/* sample_persistant (ecmc_mpi_cb.cpp:496-637) for one chain in block org[]. */
static void sample_chain(qo_lat* l, const qo_ecmc_params* p, const int org[4], qo_ecmc_chain* st, rng_t* r) {
    const int* b = p->block;
    const double beta_red = -(p->beta / 3.0);
    if (!st->initialized) {
        double dummy = 0.0;
        r->n = 0;
        draw_site(r, l, b, org, &st->site, &st->mu, &st->eps, &dummy, p, 0);
        if (!r->philox) {
            qo_random_su3_ref(r->mt, st->R);
            st->th_refresh_site = p->poisson ? exp_ref(r->mt, p->theta_refresh_site) : p->theta_refresh_site;
            st->th_refresh_R = p->poisson ? exp_ref(r->mt, p->theta_refresh_R) : p->theta_refresh_R;
        } else {
            uint32_t o4[4], o5[4];
            ph(r, 0, 4, o4);
            ph(r, 0, 5, o5);
            qo_haar_su3_philox(r->key, r->chain, 0u, 0xED000000u, st->R);
            st->th_refresh_site = p->poisson ? -p->theta_refresh_site * log(qo_u53(o4[2], o4[3])) : p->theta_refresh_site;
            st->th_refresh_R = p->poisson ? -p->theta_refresh_R * log(qo_u53(o5[0], o5[1])) : p->theta_refresh_R;
        }
        st->par_refresh_site = st->par_refresh_R = 0.0;
        st->set_counter = st->events = st->lifts = 0;
        st->initialized = 1;
    }
    r->n = st->rng_n;
    r->gen = st->set_gen;
    /* theta_sample, then the matrix set of this call (ecmc_mpi_cb.cpp:534-541) */
    double th_sample;
    if (!r->philox) {
        th_sample = p->poisson ? exp_ref(r->mt, p->theta_sample) : p->theta_sample;
    } else {
        r->n += 1;
        uint32_t o6[4];
        ph(r, r->n, 6, o6);
        th_sample = p->poisson ? -p->theta_sample * log(qo_u53(o6[0], o6[1])) : p->theta_sample;
    }
    r->gen += 1;
    if (!r->philox) set_regen_ref(r);
    double par_sample = 0.0;
    for (;;) {
        if (st->set_counter % 10000 == 0) {
            r->gen += 1;
            if (!r->philox) set_regen_ref(r);
            st->set_counter = 0;
        }
        r->n += 1;
        uint32_t o0[4] = {0}, o1[4] = {0}, o2[4] = {0};
        double gam[6];
        if (r->philox) {
            ph(r, r->n, 0, o0);
            ph(r, r->n, 1, o1);
            ph(r, r->n, 2, o2);
            for (int i = 0; i < 4; ++i) gam[i] = -log(qo_u32(o0[i]));
            gam[4] = -log(qo_u32(o1[0]));
            gam[5] = -log(qo_u32(o1[1]));
        } else {
            for (int i = 0; i < 6; ++i) gam[i] = -log(qo_mt64_canonical(r->mt));
        }
        /* compute_reject_angles_fast (ecmc_mpi_cb.cpp:203-250) */
        double Rd[18], T[18], S[18];
        m3_dag(st->R, Rd);
        qo_m3_mul(Rd, LNK(l, st->site, st->mu), T);
        int j = 0;
        double theta_reject = 0.0;
        for (int i = 0; i < 6; ++i) {
            staple_j(l, st->site, st->mu, i, S);
            double p00r = 0, p00i = 0, p11r = 0, p11i = 0;
            for (int c = 0; c < 3; ++c) {
                double m0r = 0, m0i = 0, m1r = 0, m1i = 0;
                for (int k = 0; k < 3; ++k) {
                    m0r += RE(T, 0, k) * RE(S, k, c) - IM(T, 0, k) * IM(S, k, c);
                    m0i += RE(T, 0, k) * IM(S, k, c) + IM(T, 0, k) * RE(S, k, c);
                    m1r += RE(T, 1, k) * RE(S, k, c) - IM(T, 1, k) * IM(S, k, c);
                    m1i += RE(T, 1, k) * IM(S, k, c) + IM(T, 1, k) * RE(S, k, c);
                }
                p00r += m0r * RE(st->R, c, 0) - m0i * IM(st->R, c, 0);
                p00i += m0r * IM(st->R, c, 0) + m0i * RE(st->R, c, 0);
                p11r += m1r * RE(st->R, c, 1) - m1i * IM(st->R, c, 1);
                p11i += m1r * IM(st->R, c, 1) + m1i * RE(st->R, c, 1);
            }
            double A = (p00r + p11r) * beta_red;
            double B = (p11i - p00i) * beta_red;
            double rej = qo_ecmc_reject_fast(A, B, gam[i], st->eps);
            if (i == 0 || rej < theta_reject) {
                theta_reject = rej;
                j = i;
            }
        }
        if (!(theta_reject >= 0.0)) break;   /* non-finite angle: stop instead of looping */
        double d_sample = th_sample - par_sample;
        double d_rs = st->th_refresh_site - st->par_refresh_site;
        double d_rR = st->th_refresh_R - st->par_refresh_R;
        double step = theta_reject;
        if (d_sample < step) step = d_sample;
        if (d_rs < step) step = d_rs;
        if (d_rR < step) step = d_rR;
        double* U = LNK(l, st->site, st->mu);
        if (step == d_sample) {
            qo_ecmc_update(U, d_sample, st->eps, st->R);
            st->events++;
            st->par_refresh_site += d_sample;
            st->par_refresh_R += d_sample;
            break;
        } else if (step == d_rs) {
            qo_ecmc_update(U, d_rs, st->eps, st->R);
            st->events++;
            par_sample += d_rs;
            st->par_refresh_R += d_rs;
            st->par_refresh_site = 0.0;
            draw_site(r, l, b, org, &st->site, &st->mu, &st->eps, &st->th_refresh_site, p, 1);
        } else if (step == d_rR) {
            qo_ecmc_update(U, d_rR, st->eps, st->R);
            st->events++;
            par_sample += d_rR;
            st->par_refresh_site += d_rR;
            st->par_refresh_R = 0.0;
            draw_R(r, st->R, &st->th_refresh_R, p, 1);
        } else {
            qo_ecmc_update(U, theta_reject, st->eps, st->R);
            st->events++;
            par_sample += theta_reject;
            st->par_refresh_site += theta_reject;
            st->par_refresh_R += theta_reject;
            size_t ns;
            int nm;
            int ne = lift(l, b, st->site, st->mu, j, st->R, r, o1, o2, &ns, &nm);
            st->revs += (ns == st->site) ? 1 : 0;
            st->set_counter++;
            st->lifts++;
            st->site = ns;
            st->mu = nm;
            st->eps = ne;
        }
    }
    st->rng_n = r->n;
    st->set_gen = r->gen;
}

static int block_setup(const qo_lat* l, const qo_ecmc_params* p, int nb[4]) {
    for (int d = 0; d < 4; ++d) {
        if (p->block[d] < 4 || l->L[d] % p->block[d]) return -1;
        nb[d] = l->L[d] / p->block[d];
    }
    return 0;
}

int qo_ecmc_n_chains(const qo_lat* l, const qo_ecmc_params* p) {
    int nb[4];
    if (block_setup(l, p, nb)) return -1;
    return nb[0] * nb[1] * nb[2] * nb[3];
}

/* Every chain of the block decomposition, one sample each, Philox mode. */
int qo_ecmc_sample_philox(qo_lat* l, const qo_ecmc_params* p, qo_ecmc_chain* chains, uint64_t seed) {
    int nb[4];
    if (block_setup(l, p, nb)) return -1;
    int n = nb[0] * nb[1] * nb[2] * nb[3];
    rng_t r;
    memset(&r, 0, sizeof r);
    r.philox = 1;
    r.key[0] = (uint32_t)seed;
    r.key[1] = (uint32_t)(seed >> 32);
    r.eps_set = p->epsilon_set;
    for (int c = 0; c < n; ++c) {
        int bc[4] = {c % nb[0], (c / nb[0]) % nb[1], (c / (nb[0] * nb[1])) % nb[2], c / (nb[0] * nb[1] * nb[2])};
        int org[4];
        for (int d = 0; d < 4; ++d) org[d] = bc[d] * p->block[d];
        r.chain = (uint32_t)c;
        sample_chain(l, p, org, &chains[c], &r);
    }
    return n;
}

/* The reference's one-rank chain (block = whole lattice), mt19937_64 draws. */
int qo_ecmc_sample_ref(qo_lat* l, const qo_ecmc_params* p, qo_ecmc_chain* chain, qo_mt64* g) {
    qo_ecmc_params q = *p;
    for (int d = 0; d < 4; ++d) q.block[d] = l->L[d];
    rng_t r;
    memset(&r, 0, sizeof r);
    r.mt = g;
    r.eps_set = p->epsilon_set;
    r.set = (double*)malloc((size_t)SET_SIZE * 18 * sizeof(double));
    if (!r.set) return -1;
    const int org[4] = {0, 0, 0, 0};
    sample_chain(l, &q, org, chain, &r);
    free(r.set);
    return 0;
}
