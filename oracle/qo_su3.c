/*
 * qo_su3.c -- SU(3)/SU(2) algebra of the oracle (TEST INFRASTRUCTURE ONLY).
 * A 3x3 complex matrix is 18 doubles, row-major, (re,im) interleaved, as in
 * the reference's Eigen RowMajor SU3 (src/su3/utils.h:12).
 */
#include "qo.h"

#include <math.h>
#include <string.h>

#define RE(a, i, j) (a)[((i)*3 + (j))*2]
#define IM(a, i, j) (a)[((i)*3 + (j))*2 + 1]

void qo_m3_mul(const double* a, const double* b, double* out) {
    double t[18];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k < 3; ++k) {
                re += RE(a, i, k) * RE(b, k, j) - IM(a, i, k) * IM(b, k, j);
                im += RE(a, i, k) * IM(b, k, j) + IM(a, i, k) * RE(b, k, j);
            }
            RE(t, i, j) = re;
            IM(t, i, j) = im;
        }
    memcpy(out, t, sizeof t);
}

void qo_m3_dag(const double* a, double* out) {
    double t[18];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            RE(t, i, j) = RE(a, j, i);
            IM(t, i, j) = -IM(a, j, i);
        }
    memcpy(out, t, sizeof t);
}

/* P_TA(A) = (A - A^dag)/2 - Tr[(A - A^dag)/2]/3 * 1   (src/su3/utils.cpp:130-134) */
void qo_proj_lie(double* A) {
    double B[18];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            RE(B, i, j) = 0.5 * (RE(A, i, j) - RE(A, j, i));
            IM(B, i, j) = 0.5 * (IM(A, i, j) + IM(A, j, i));
        }
    double tr_re = (RE(B, 0, 0) + RE(B, 1, 1) + RE(B, 2, 2)) / 3.0;
    double tr_im = (IM(B, 0, 0) + IM(B, 1, 1) + IM(B, 2, 2)) / 3.0;
    for (int i = 0; i < 3; ++i) {
        RE(B, i, i) -= tr_re;
        IM(B, i, i) -= tr_im;
    }
    memcpy(A, B, sizeof B);
}

/* exp(coeff*Z) for Z in su(3): Cayley-Hamilton form q0 + q1 X + q2 X^2 with
 * t = -Re Tr X^2 / 2, d = Re(i det X), Horner over N = 18 Taylor terms
 * (src/su3/utils.cpp:145-193). */
void qo_exp_luscher(const double* Z, double coeff, double* out) {
    double X[18], X2[18];
    for (int k = 0; k < 18; ++k) X[k] = coeff * Z[k];
    qo_m3_mul(X, X, X2);
    double t = -0.5 * (RE(X2, 0, 0) + RE(X2, 1, 1) + RE(X2, 2, 2));
    /* det X by the rule of Sarrus-free cofactor expansion along row 0 */
    qo_c m[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { m[i][j].re = RE(X, i, j); m[i][j].im = IM(X, i, j); }
#define CMUL_RE(a, b) ((a).re * (b).re - (a).im * (b).im)
#define CMUL_IM(a, b) ((a).re * (b).im + (a).im * (b).re)
    qo_c c0, c1, c2;
    c0.re = CMUL_RE(m[1][1], m[2][2]) - CMUL_RE(m[1][2], m[2][1]);
    c0.im = CMUL_IM(m[1][1], m[2][2]) - CMUL_IM(m[1][2], m[2][1]);
    c1.re = CMUL_RE(m[1][2], m[2][0]) - CMUL_RE(m[1][0], m[2][2]);
    c1.im = CMUL_IM(m[1][2], m[2][0]) - CMUL_IM(m[1][0], m[2][2]);
    c2.re = CMUL_RE(m[1][0], m[2][1]) - CMUL_RE(m[1][1], m[2][0]);
    c2.im = CMUL_IM(m[1][0], m[2][1]) - CMUL_IM(m[1][1], m[2][0]);
    double det_im = CMUL_IM(m[0][0], c0) + CMUL_IM(m[0][1], c1) + CMUL_IM(m[0][2], c2);
#undef CMUL_RE
#undef CMUL_IM
    double d = -det_im; /* Re(i * det) */
    const int N = 18;
    double c[19];
    c[0] = 1.0;
    for (int n = 1; n <= N; ++n) c[n] = c[n - 1] / (double)n;
    /* q0,q1,q2 complex; recursion q0' = c_n - i d q2, q1' = q0 - t q2, q2' = q1 */
    double q0r = c[N], q0i = 0.0, q1r = 0.0, q1i = 0.0, q2r = 0.0, q2i = 0.0;
    for (int n = N - 1; n >= 0; --n) {
        double n0r = c[n] + d * q2i;   /* -i d (q2r + i q2i) = d q2i - i d q2r */
        double n0i = -d * q2r;
        double n1r = q0r - t * q2r, n1i = q0i - t * q2i;
        q2r = q1r; q2i = q1i;
        q1r = n1r; q1i = n1i;
        q0r = n0r; q0i = n0i;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double re = q1r * RE(X, i, j) - q1i * IM(X, i, j) + q2r * RE(X2, i, j) - q2i * IM(X2, i, j);
            double im = q1r * IM(X, i, j) + q1i * RE(X, i, j) + q2r * IM(X2, i, j) + q2i * RE(X2, i, j);
            if (i == j) { re += q0r; im += q0i; }
            RE(out, i, j) = re;
            IM(out, i, j) = im;
        }
}

/* Embedding of a0 + i a.sigma into rows/columns (i,j) (src/su3/utils.cpp:33-44). */
void qo_su2q_to_su3(const double q[4], int i, int j, double* out) {
    memset(out, 0, 18 * sizeof(double));
    int k = 3 - i - j;
    RE(out, k, k) = 1.0;
    RE(out, i, i) = q[0]; IM(out, i, i) = q[3];
    RE(out, j, j) = q[0]; IM(out, j, j) = -q[3];
    RE(out, i, j) = q[2]; IM(out, i, j) = q[1];
    RE(out, j, i) = -q[2]; IM(out, j, i) = q[1];
}

/* 2x2 block [[a,b],[c,d]] (4 complex, row-major) -> quaternion (src/su3/utils.cpp:47-50). */
void qo_su2_to_q(const double* w, double q[4]) {
    /* w = {a.re,a.im, b.re,b.im, c.re,c.im, d.re,d.im} */
    q[0] = (w[0] + w[6]) / 2.0;
    q[1] = (w[3] + w[5]) / 2.0;
    q[2] = (w[2] - w[4]) / 2.0;
    q[3] = (w[1] - w[7]) / 2.0;
}

void qo_qmult(const double q[4], const double p[4], double r[4]) {
    double r0 = q[0] * p[0] - q[1] * p[1] - q[2] * p[2] - q[3] * p[3];
    double r1 = q[0] * p[1] + q[1] * p[0] + q[2] * p[3] - q[3] * p[2];
    double r2 = q[0] * p[2] - q[1] * p[3] + q[2] * p[0] + q[3] * p[1];
    double r3 = q[0] * p[3] + q[1] * p[2] - q[2] * p[1] + q[3] * p[0];
    r[0] = r0; r[1] = r1; r[2] = r2; r[3] = r3;
}

/* One Kennedy-Pendleton attempt (src/heatbath/heatbath.cpp:14-26). */
int qo_kp_try(double alpha, double delta, double r1, double r2, double r3, double* a0) {
    double cs = cos(2 * M_PI * r2);
    double lambda_sq = -0.5 / alpha * (log(r1) + cs * cs * log(r3));
    if (delta * delta <= 1.0 - lambda_sq) {
        *a0 = 1.0 - 2.0 * lambda_sq;
        return 1;
    }
    return 0;
}

/* Uniform point on the 2-sphere of radius sqrt(1-a0^2) (src/heatbath/heatbath.cpp:29-35). */
void qo_kp_sphere(double a0, double u_phi, double u_cos, double out[4]) {
    double r_spatial = sqrt(1.0 - a0 * a0);
    double phi = 2.0 * M_PI * u_phi;
    double cos_theta = 2.0 * u_cos - 1.0;
    double sin_theta = sqrt(1.0 - cos_theta * cos_theta);
    out[0] = a0;
    out[1] = r_spatial * sin_theta * cos(phi);
    out[2] = r_spatial * sin_theta * sin(phi);
    out[3] = r_spatial * cos_theta;
}

int qo_kp_injected(double beta, double k, const double* u, int K, double out[4]) {
    double alpha = beta * k * 2.0 / 3.0;
    double a0 = 0.0;
    for (int a = 0; a < K; ++a) {
        if (qo_kp_try(alpha, u[4 * a], u[4 * a + 1], u[4 * a + 2], u[4 * a + 3], &a0)) {
            qo_kp_sphere(a0, u[4 * K], u[4 * K + 1], out);
            return a;
        }
    }
    return -1;
}

/* Reference-order KP drawing from mt19937_64 until acceptance (heatbath.cpp:8-36). */
void qo_kp_mt(double beta, double k, qo_mt64* g, double out[4]) {
    double alpha = beta * k * 2.0 / 3.0;
    double a0;
    for (;;) {
        double delta = qo_mt64_canonical(g);
        double r1 = qo_mt64_canonical(g);
        double r2 = qo_mt64_canonical(g);
        double r3 = qo_mt64_canonical(g);
        if (qo_kp_try(alpha, delta, r1, r2, r3, &a0)) break;
    }
    double u_phi = qo_mt64_canonical(g);
    double u_cos = qo_mt64_canonical(g);
    qo_kp_sphere(a0, u_phi, u_cos, out);
}

/* Gram-Schmidt on columns then c2 = conj(c0 x c1) (src/gauge/GaugeField.cpp:36-58). */
void qo_reunitarize_link(double* U) {
    qo_c c0[3], c1[3], c2[3];
    for (int i = 0; i < 3; ++i) {
        c0[i].re = RE(U, i, 0); c0[i].im = IM(U, i, 0);
        c1[i].re = RE(U, i, 1); c1[i].im = IM(U, i, 1);
    }
    double n0 = sqrt(c0[0].re * c0[0].re + c0[0].im * c0[0].im + c0[1].re * c0[1].re +
                     c0[1].im * c0[1].im + c0[2].re * c0[2].re + c0[2].im * c0[2].im);
    for (int i = 0; i < 3; ++i) { c0[i].re /= n0; c0[i].im /= n0; }
    /* dot = sum conj(c0_i) c1_i */
    double dr = 0.0, di = 0.0;
    for (int i = 0; i < 3; ++i) {
        dr += c0[i].re * c1[i].re + c0[i].im * c1[i].im;
        di += c0[i].re * c1[i].im - c0[i].im * c1[i].re;
    }
    for (int i = 0; i < 3; ++i) {
        c1[i].re -= c0[i].re * dr - c0[i].im * di;
        c1[i].im -= c0[i].re * di + c0[i].im * dr;
    }
    double n1 = sqrt(c1[0].re * c1[0].re + c1[0].im * c1[0].im + c1[1].re * c1[1].re +
                     c1[1].im * c1[1].im + c1[2].re * c1[2].re + c1[2].im * c1[2].im);
    for (int i = 0; i < 3; ++i) { c1[i].re /= n1; c1[i].im /= n1; }
    for (int i = 0; i < 3; ++i) {
        int a = (i + 1) % 3, b = (i + 2) % 3;
        /* (c0_a c1_b - c0_b c1_a), conjugated */
        double re = (c0[a].re * c1[b].re - c0[a].im * c1[b].im) - (c0[b].re * c1[a].re - c0[b].im * c1[a].im);
        double im = (c0[a].re * c1[b].im + c0[a].im * c1[b].re) - (c0[b].re * c1[a].im + c0[b].im * c1[a].re);
        c2[i].re = re;
        c2[i].im = -im;
    }
    for (int i = 0; i < 3; ++i) {
        RE(U, i, 0) = c0[i].re; IM(U, i, 0) = c0[i].im;
        RE(U, i, 1) = c1[i].re; IM(U, i, 1) = c1[i].im;
        RE(U, i, 2) = c2[i].re; IM(U, i, 2) = c2[i].im;
    }
}
