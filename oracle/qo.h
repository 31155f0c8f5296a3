/*
 * qo.h -- FP64 CPU restatement of the reference's hot path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the libqcb HIP library,
 * the gauge_hb_cb and gauge_ecmc_cb drivers, the Python bindings) may link or
 * call this code;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only as the checker.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the reference repository yacobozdalkiran/ECMC-QNCB).  Link storage is the
 * reference's interior AoS order: links[((site*4 + mu)*9 + 3*i + j)*2 + {re,im}],
 * site = x + Lx*(y + Ly*(z + Lz*t)), x fastest (src/gauge/GaugeField.h:27,49-56,
 * src/geometry/GeometryCB.h:24-26).  The one-rank halo of the reference is
 * replaced by periodic indexing; with one rank the two are identical (every
 * halo link a non-frozen staple reads is itself frozen, and every other
 * consumer runs after exchange_halos_cascade).
 *
 * Parity status of each function is stated in its comment: "pinned by ..."
 * names the independent check in tests/ that would fail on a plausible slip.
 */
#ifndef QCB_ORACLE_QO_H
#define QCB_ORACLE_QO_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } qo_c;
typedef struct { qo_c e[3][3]; } qo_m3;

/* ---------------- SU(3) algebra (src/su3/utils.cpp) ---------------- */
void qo_m3_mul(const double* a, const double* b, double* out);        /* a*b           */
void qo_proj_lie(double* A);                                           /* utils.cpp:130-134; pinned by anti-Hermitian/traceless + linear-algebra checks */
void qo_exp_luscher(const double* Z, double coeff, double* out);       /* utils.cpp:145-193; pinned by scipy expm golden vectors */
void qo_su2q_to_su3(const double q[4], int i, int j, double* out);     /* utils.cpp:33-44;   pinned by Pauli-matrix closed form */
void qo_su2_to_q(const double* w2x2, double q[4]);                     /* utils.cpp:47-50;   pinned by round trip */
void qo_qmult(const double q[4], const double p[4], double out[4]);    /* utils.cpp:53-60;   pinned by SU(2) matrix product */
void qo_reunitarize_link(double* U);                                   /* GaugeField.cpp:36-58; pinned by unitarity/det and by fixing SU(3) inputs */

/* Kennedy-Pendleton (src/heatbath/heatbath.cpp:8-36) with explicit uniforms.
 * u holds K attempts of (delta, r1, r2, r3) followed by (u_phi, u_costheta).
 * Returns the 0-based accepted attempt, or -1 when all K attempts reject
 * (then out is left untouched).  Pinned by the closed-form map and a KS test
 * of the a0 marginal against the target density. */
int qo_kp_injected(double beta, double k, const double* u, int K, double out[4]);

/* ---------------- RNG (libstdc++ <random> as used by the reference) ---------------- */
typedef struct { uint64_t mt[312]; int idx; } qo_mt64;
void qo_mt64_seed(qo_mt64* g, uint64_t seed);
uint64_t qo_mt64_next(qo_mt64* g);
double qo_mt64_canonical(qo_mt64* g);                 /* std::generate_canonical<double,53> */
uint64_t qo_mt64_uniform_int(qo_mt64* g, uint64_t a, uint64_t b); /* std::uniform_int_distribution */
void qo_mt64_normal_pair(qo_mt64* g, double* first, double* second); /* std::normal_distribution: two draws */
void qo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double qo_u53(uint32_t hi, uint32_t lo);
double qo_u32(uint32_t w);

/* ---------------- lattice ---------------- */
typedef struct {
    int L[4];          /* x, y, z, t extents (x even) */
    size_t V;
    double* U;         /* V*4*18 doubles */
} qo_lat;

void qo_set_threads(int n);                                             /* OpenMP threads of the per-site loops */
int  qo_lat_init(qo_lat* l, const int L[4]);
void qo_lat_free(qo_lat* l);
void qo_lat_cold(qo_lat* l);                                            /* GaugeField.cpp:23-33 */
void qo_lat_hot_ref(qo_lat* l, qo_mt64* g);                             /* GaugeField.cpp:10-20 + utils.cpp:13-30 (clang/IntelLLVM draw order) */
void qo_lat_hot_philox(qo_lat* l, uint64_t seed);                       /* device hot start restated: Box-Muller + Gram-Schmidt + det^(1/3) */
void qo_haar_su3_philox(const uint32_t key[2], uint32_t c0, uint32_t c1, uint32_t c2, double* Q); /* the same on counters {c0,c1,c2,0..8} */
void qo_random_su3_ref(qo_mt64* g, double* Q);                          /* utils.cpp:13-30 (one matrix), clang/IntelLLVM draw order */
void qo_random_su3_ref_order(qo_mt64* g, double* Q, int imag_first);    /* utils.cpp:13-30 with the draw order of a clang (0) or GCC (1) build */
void qo_staple(const qo_lat* l, size_t site, int mu, double* out);      /* GaugeField.cpp:74-97 */
double qo_plaquette(const qo_lat* l);                                   /* observables_mpi.cpp:8-43 */
void qo_clover_F(const qo_lat* l, size_t site, int mu, int nu, double* F); /* observables_mpi.cpp:61-85 */
void qo_clover_qe(const qo_lat* l, double* Q, double* E_mean);          /* observables_mpi.cpp:88-150 */
void qo_clover_qe_site(const qo_lat* l, double* q_site, double* e_site);/* per-site densities */
void qo_reunitarize(qo_lat* l);                                         /* GaugeField.cpp:61-71 */
void qo_translate(qo_lat* l, const int s[4]);                           /* Shift.cpp:69-160 (one rank) */
void qo_gauge_transform(qo_lat* l, const double* G);                    /* GaugeField.cpp:100-114 with a given G field (V*18) */
void qo_flow_rk3_step(qo_lat* l, double eps);                           /* gradient_flow.cpp:15-128 */
void qo_topo_charge_flowed(const qo_lat* l, double eps, int n_steps_gf, int n_rk_steps,
                           double* tQE);                                 /* observables_mpi.cpp:153-178 */

/* Reference-order heatbath: lexicographic t,z,y,x, mu inner, frozen links of
 * the one-rank block skipped, N_hits CM hits per link, uniforms from mt19937_64
 * (src/heatbath/heatbath_mpi.cpp:37-69, src/geometry/GeometryCB.cpp:33-56).
 * Returns the number of link visits (active links). */
long qo_hb_sweep_ref(qo_lat* l, double beta, int n_hits, qo_mt64* g);

/* Checkerboard order used on the GPU: for mu 0..3, parity 0..1, every site of
 * that parity (links in one class are independent), N_hits CM hits with the
 * same per-hit steps as heatbath_mpi.cpp:11-34 and the staple computed once.
 * Injected uniforms: u[((((hs*n_hits + h)*3 + g)*(4K+2) + slot)*Vh + cb]
 * with hs = 2*mu + parity.  Returns the number of exhausted (link,hit,subgroup)
 * budgets; those leave the link unchanged for that subgroup. */
long qo_hb_sweep_cb_injected(qo_lat* l, double beta, int n_hits, const double* u, int K);
/* One CM hit with injected uniforms u[g*(4K+2) + slot] (heatbath_mpi.cpp:11-34);
 * returns the number of subgroups whose K attempts all rejected. */
int qo_cm_hit_injected(double* U, const double* A, double beta, const double* u, int K);

/* Checkerboard order with Philox4x32-10 uniforms (counter scheme documented
 * in qo_rng.c and in the device header). */
void qo_hb_sweep_cb_philox(qo_lat* l, double beta, int n_hits, uint64_t seed, uint32_t sweep_idx);

/* One-rank reference run (src/bin/gauge_hb_ncb.cpp:18-201) with per-phase timers. */
typedef struct {
    int L; int cold_start; int seed;
    int N_therm; int N_shift; double beta; int N_sweeps; int N_hits;
    int N_shift_plaquette; int topo; int N_shift_topo; int N_steps_gf; int N_rk_steps;
} qo_run_params;
typedef struct {
    double t_hb, t_plaq, t_flow, t_shift;   /* seconds */
    long link_visits;                        /* active links visited by HB */
    long site_steps;                         /* V * RK3 steps */
    int n_plaq; int n_tqe;
} qo_run_stats;
/* plaq_out: N_shift/N_shift_plaquette (+1) values of the sampling loop;
 * tqe_out: 3*N_steps_gf per flow; shifts_out: 4 ints per shift (therm + sampling). */
int qo_ref_run(const qo_run_params* p, double* plaq_out, double* tqe_out, int* shifts_out,
               double* final_links, qo_run_stats* st);

/* ---------------- event-chain Monte Carlo (src/ecmc/ecmc_mpi_cb.cpp) ---------------- */
typedef struct {
    double beta;
    double theta_sample, theta_refresh_site, theta_refresh_R;   /* param_theta_* (params.h:9-17) */
    int poisson;
    double epsilon_set;
    int block[4];            /* chain sub-block extents (the reference's L_core per rank) */
} qo_ecmc_params;
typedef struct {             /* LocalChainState (ecmc_mpi_cb.h:15-35) plus the Philox position */
    double R[18];
    double th_refresh_site, th_refresh_R, par_refresh_site, par_refresh_R;
    uint64_t events, lifts, revs, set_counter, set_gen, rng_n;
    size_t site;             /* global lex site */
    int mu, eps, initialized;
} qo_ecmc_chain;
/* ecmc_mpi_cb.h:56-94 and ecmc_mpi_cb.cpp:82-178; both pinned by the defining
 * integral (quadrature of the positive part of dS/dtheta between its zeros
 * equals gamma) and by their mutual agreement */
double qo_ecmc_reject_fast(double A, double B, double gamma, int epsilon);
double qo_ecmc_reject_exact(double A, double B, double gamma, int epsilon);
/* ecmc_mpi_cb.cpp:263-275 and :357-372; pinned by numpy matrix expressions
 * Im[(R^dag P R)_00 - (R^dag P R)_11] and R diag(e^{i xi}, e^{-i xi}, 1) R^dag U */
double qo_ecmc_compute_ds(const double* Pi, const double* R);
void qo_ecmc_update(double* U, double theta, int epsilon, const double* R);
/* utils.cpp:66-103; pinned by the embedded-SU(2) product form and unitarity */
void qo_ecmc_random_su3_eps(double eps, const double r[9], double* M);
/* one chain per block; pinned by every chain's site lying in its own block */
int qo_ecmc_n_chains(const qo_lat* l, const qo_ecmc_params* p);
/* One sample (sample_persistant, ecmc_mpi_cb.cpp:496-637) of every chain of
 * the block decomposition with the device's Philox counters:
 *   key = seed; ctr(n, call) = {chain, n_lo, 0xEC000000 | n_hi(24 bits), call}
 *   sample start: n += 1, call 6 -> theta_sample (u53 of w0,w1, Poisson)
 *   event: n += 1; call 0 -> gamma_0..3, call 1 -> gamma_4,5, select, accept;
 *          call 2 -> set index (64-bit w0:w1 times 10001, high word);
 *          refresh site: call 3 -> 4 coordinates 1 + hi32(w * (b-2)), call 4 ->
 *          dir = w0>>30, eps = 2(w1>>31)-1, theta_refresh_site (w2,w3);
 *          refresh R: call 5 -> theta_refresh_R (w0,w1), Haar SU(3) on
 *          {chain, n_lo, 0xED000000 | n_hi, 0..8}
 *   init: the same draws at n = 0; gamma u = u32(w), exponentials -mean log(u53)
 *   matrix set, generation g (+1 per sample call and per 10000 lifts):
 *          element 2m+1 from {chain, g, 0xE5000000 | m, 0..2} (9 words u32 - 1/2),
 *          element 2m+2 its adjoint, element 0 the identity. */
/* Pinned by: frozen links untouched, SU(3) kept, chains of different blocks
 * independent (a perturbation in one block changes no other block), and the
 * device chains' equilibrium plaquette against the device heatbath.  The
 * event sequence itself (which event comes when) has no check independent of
 * this restatement: parity unpinned beyond those invariants. */
int qo_ecmc_sample_philox(qo_lat* l, const qo_ecmc_params* p, qo_ecmc_chain* chains, uint64_t seed);
/* The reference's one-rank chain (the whole lattice one block), mt19937_64;
 * pinned by the same invariants (the draw order of the reference's std
 * distributions is restated from its source: parity unpinned beyond them). */
int qo_ecmc_sample_ref(qo_lat* l, const qo_ecmc_params* p, qo_ecmc_chain* chain, qo_mt64* g);

double qo_now(void);

#ifdef __cplusplus
}
#endif
#endif
