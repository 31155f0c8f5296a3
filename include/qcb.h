/*
 * qcb.h -- C-ABI of the MI355X heatbath / Wilson-flow / observable library
 * (libqcb.so).  Plain pointers and sizes only; no exceptions cross it.
 *
 * Conventions
 *  - Every call returns QCB_OK (0) or a negative QCB_E* code; the message of
 *    the last failure on a context is qcb_last_error(ctx).
 *  - Host link buffers use the reference's interior AoS order with host
 *    endianness: aos[((site*4 + mu)*9 + 3*i + j)*2 + {re,im}], site = x +
 *    Lx*(y + Ly*(z + Lz*t)) (reference src/gauge/GaugeField.h:27,49-56 and the
 *    ILDG site order of src/io/ildg.cpp:87-106).  Host buffers are borrowed
 *    only for the duration of a call.  Device memory belongs to the handle.
 *  - Calls on one context are ordered on its HIP stream and are not
 *    thread-safe; use one context per host thread.  Results returned through
 *    host pointers are complete when the call returns.
 *  - prec: QCB_F64 (links and arithmetic in double) or QCB_F32 (links,
 *    staples and flow stages in float; global sums, flow time and
 *    reunitarisation in double).  An FP32 lattice stores rows 0 and 1 of
 *    each link and rebuilds row 2 as conj(row 0 x row 1), so it holds SU(3)
 *    matrices: qcb_lattice_set_links keeps rows 0 and 1 of its input and
 *    qcb_lattice_get_links returns the rebuilt row 2.
 *  - Lattice extents are periodic, each >= 4, Lx even (even/odd split);
 *    L^3 x T lattices are allowed (the reference is hypercubic only).
 */
#ifndef QCB_H
#define QCB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QCB_OK 0
#define QCB_EINVAL (-1)     /* bad argument */
#define QCB_EDEVICE (-2)    /* HIP runtime / no GPU */
#define QCB_ENOMEM (-3)     /* device allocation failed */
#define QCB_EBUDGET (-4)    /* Kennedy-Pendleton attempt budget exhausted (links left unchanged) */
#define QCB_ESTATE (-5)     /* call not valid in the handle's state */

#define QCB_F64 0
#define QCB_F32 1

typedef struct qcb_ctx qcb_ctx;
typedef struct qcb_lattice qcb_lattice;
typedef struct qcb_batch qcb_batch;

/* ---- context ------------------------------------------------------------ */
/* Replaces the process setup of src/bin/gauge_hb_ncb.cpp:203-226 (MPI_Init and
 * MpiTopology, src/mpi/MpiTopology.cpp:7-24): one GPU, no ranks. */
int qcb_ctx_create(int device, qcb_ctx** out);
/* Fails with QCB_ESTATE (and frees nothing) while lattices or batches created
 * on the context are still alive. */
int qcb_ctx_destroy(qcb_ctx* ctx);
const char* qcb_last_error(qcb_ctx* ctx);
int qcb_ctx_synchronize(qcb_ctx* ctx);
/* qcb_last_error(NULL) reports the last failure of a context-free call. */
/* Attempts a Kennedy-Pendleton draw gets in Philox-mode heatbath sweeps on
 * this context from the next sweep call on (default 1024, 1..2^20).  The
 * reference loops without bound (src/heatbath/heatbath.cpp:14-26); a draw
 * that runs out leaves its link's SU(2) subgroup unchanged, and the sweep call
 * that made it returns QCB_EBUDGET (its other updates stand).  At beta = 6 a
 * draw needs more than 10 attempts about once in 10^9. */
int qcb_ctx_set_kp_budget(qcb_ctx* ctx, uint32_t max_attempts);
/* Device-side time of the last kernel-timing window (see qcb_timing_*), ms. */
int qcb_timing_begin(qcb_ctx* ctx);
int qcb_timing_end(qcb_ctx* ctx, double* ms);

/* ---- lattice ------------------------------------------------------------ */
/* GaugeField(const GeometryCB&) (src/gauge/GaugeField.h:30-43): links start
 * at the identity. dims = {Lx, Ly, Lz, Lt}. */
int qcb_lattice_create(qcb_ctx* ctx, const int dims[4], int prec, qcb_lattice** out);
int qcb_lattice_destroy(qcb_lattice* lat);
/* Device bytes held by the handle (links plus flow buffers once allocated). */
int qcb_lattice_bytes(qcb_lattice* lat, uint64_t* bytes);
/* Work-order tiling of the neighbour-reading kernels: tiles of Lx x ty x tz x tt
 * sites (each must divide its extent), visited y-fastest or, with morton != 0,
 * in Z-order over the tile grid (power-of-two tile counts only).  Changes
 * speed only: every kernel
 * result is independent of it except the last bits of the FP64 global sums,
 * whose block partition follows the tiles.  Default: 8 x 4 x 4 (largest
 * divisors), y-fastest. */
int qcb_lattice_set_tiling(qcb_lattice* lat, int ty, int tz, int tt, int morton);
/* Upload / download of the whole field (src/io/ildg.cpp:87-106,213-232 minus
 * the byte swap, which stays with the caller's file codec). */
int qcb_lattice_set_links(qcb_lattice* lat, const double* aos);
int qcb_lattice_get_links(qcb_lattice* lat, double* aos);

/* GaugeField::cold_start (src/gauge/GaugeField.cpp:23-33). */
int qcb_cold(qcb_lattice* lat);
/* Device hot start: Haar SU(3) per link from Philox (Box-Muller Gaussians,
 * Gram-Schmidt, det^(-1/3)); replaces GaugeField::hot_start
 * (src/gauge/GaugeField.cpp:10-20) where the caller does not need the host
 * mt19937_64 stream to advance. */
int qcb_hot(qcb_lattice* lat, uint64_t seed);

/* ---- observables -------------------------------------------------------- */
/* mpi::observables::mean_plaquette_global (src/observables/observables_mpi.cpp:34-43). */
int qcb_plaquette(qcb_lattice* lat, double* mean_plaquette);
/* mpi::observables::topo_q_e_clover_global (src/observables/observables_mpi.cpp:130-150):
 * Q = total clover charge, E_mean = sum_x e_x / V. */
int qcb_clover_qe(qcb_lattice* lat, double* Q, double* E_mean);
/* GaugeField::compute_staple for every link (src/gauge/GaugeField.cpp:74-97),
 * written to out_aos in the host link order (a checking aid). */
int qcb_staples(qcb_lattice* lat, double* out_aos);

/* ---- heatbath ----------------------------------------------------------- */
/* mpi::heatbathcb::samples (src/heatbath/heatbath_mpi.cpp:59-69) on the GPU:
 * n_sweeps sweeps of 8 (mu, parity) half-sweeps, each link receiving n_hits
 * Cabibbo-Marinari hits (src/heatbath/heatbath_mpi.cpp:11-34) with
 * Kennedy-Pendleton SU(2) draws (src/heatbath/heatbath.cpp:8-44).  Uniforms
 * come from Philox keyed by seed with counter sweep_idx, sweep_idx+1, ...
 * (a Philox stream replaces the std::mt19937_64& argument).  Synchronises;
 * QCB_EBUDGET if any draw ran out of attempts (qcb_ctx_set_kp_budget). */
int qcb_hb_sweeps(qcb_lattice* lat, double beta, int n_hits, int n_sweeps, uint64_t seed,
                  uint32_t sweep_idx);
/* One sweep fed with host-supplied uniforms: u[((((2*mu+p)*n_hits + h)*3 + g)*(4K+2)
 * + slot)*(V/2) + cb], slots 4a..4a+3 = (delta, r1, r2, r3) of attempt a < K,
 * slots 4K, 4K+1 = (phi, cos theta) uniforms.  A (link, hit, subgroup) whose K
 * attempts all reject is left unchanged and counted in *n_exhausted. */
int qcb_hb_sweep_injected(qcb_lattice* lat, double beta, int n_hits, const double* u, int K,
                          int64_t* n_exhausted);
/* GaugeField::project_field_su3 (src/gauge/GaugeField.cpp:61-71), in double. */
int qcb_reunitarize(qcb_lattice* lat);
/* mpi::shift::shift_field (src/mpi/Shift.cpp:109-160), one rank: U'(x) = U(x - s).
 * The caller draws s (random_shift, src/mpi/Shift.cpp:163-186). */
int qcb_translate(qcb_lattice* lat, const int s[4]);

/* ---- Wilson flow -------------------------------------------------------- */
/* n RK3 steps of GradientFlow::rk3_step (src/flow/gradient_flow.cpp:117-128)
 * applied to the lattice's own links. */
int qcb_flow_steps(qcb_lattice* lat, double eps, int n);
/* The GradientFlow object of the reference (src/flow/gradient_flow.h:11-27)
 * held inside the lattice handle: flow_copy = GradientFlow::copy (field_c =
 * field, src/flow/gradient_flow.cpp:131-133); flow_rk3 = n x rk3_step
 * (src/flow/gradient_flow.cpp:117-128) on field_c; flow_qe =
 * topo_q_e_clover_global(field_c); flow_get_links downloads field_c.  The
 * lattice's own links are untouched by these four calls. */
int qcb_flow_copy(qcb_lattice* lat);
int qcb_flow_rk3(qcb_lattice* lat, double eps, int n);
int qcb_flow_qe(qcb_lattice* lat, double* Q, double* E_mean);
/* mean_plaquette_global (src/observables/observables_mpi.cpp:34-43) of field_c. */
int qcb_flow_plaquette(qcb_lattice* lat, double* mean_plaquette);
int qcb_flow_get_links(qcb_lattice* lat, double* aos);
/* mpi::observables::topo_charge_flowed (src/observables/observables_mpi.cpp:153-178):
 * on a copy of the links, for s < n_steps_gf record (t, Q, E_mean) into
 * tQE[3*s..3*s+2] and then advance n_rk_steps RK3 steps; the links are not
 * modified.  If final_QE is non-NULL it receives (t, Q, E_mean) of the state
 * after the last step (the reference never measures it). */
int qcb_topo_charge_flowed(qcb_lattice* lat, double eps, int n_steps_gf, int n_rk_steps,
                           double* tQE, double* final_tQE);
/* No counterpart in the reference: qcb_topo_charge_flowed for callers that
 * cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc).  Same results; it calls no hipMalloc/hipFree.  The lattice's
 * flow buffers must exist (any earlier flow call creates them; QCB_ESTATE
 * otherwise) and the measurement results are staged in a static 1 MiB device
 * array that the library holds per GPU until the process ends, so
 * n_steps_gf + 1 <= 43690 (QCB_EINVAL beyond).  Every call on one GPU shares
 * that array: calls that overlap on a GPU (other threads or contexts) race. */
int qcb_topo_charge_flowed_noalloc(qcb_lattice* lat, double eps, int n_steps_gf, int n_rk_steps,
                                   double* tQE, double* final_tQE);

/* ---- batches of independent lattices ------------------------------------ */
/* n lattices of uneven extents (dims[4*i..4*i+3]), each with its own beta and
 * Philox seed, advanced together: one launch per (mu, parity) half-sweep,
 * per RK3 stage and per reduction covers every lattice (blocks never span
 * two lattices, so each lattice's results are bitwise those of running it
 * alone).  The reference runs one lattice per process; this replaces many
 * concurrent runs of src/bin/gauge_hb_ncb.cpp. */
int qcb_batch_create(qcb_ctx* ctx, int n, const int* dims, int prec, const double* betas,
                     const uint64_t* seeds, qcb_batch** out);
int qcb_batch_destroy(qcb_batch* b);
/* Borrowed handle of lattice i (set/get links, observables); owned by the batch. */
int qcb_batch_lattice(qcb_batch* b, int i, qcb_lattice** out);
int qcb_batch_hot(qcb_batch* b);
/* Synchronises; QCB_EBUDGET (naming the first lattice) if any draw ran out of attempts. */
int qcb_batch_hb_sweeps(qcb_batch* b, int n_hits, int n_sweeps, uint32_t sweep_idx);
int qcb_batch_reunitarize(qcb_batch* b);
int qcb_batch_plaquette(qcb_batch* b, double* out /* n */);
/* Flow all lattices from their links, measuring (t, Q, E_mean) at t = 0, eps,
 * ..., n_steps*eps: tQE[((i*(n_steps+1)) + s)*3 + {0,1,2}].  The launch
 * sequence is captured into a hipGraph at the first call and replayed by later
 * calls with the same eps and n_steps (captured again when either, or a buffer
 * address, changes). */
int qcb_batch_flow(qcb_batch* b, double eps, int n_steps, double* tQE);
/* qcb_flow_steps for every lattice: n RK3 steps of GradientFlow::rk3_step
 * (src/flow/gradient_flow.cpp:117-128) applied to each lattice's own links,
 * one launch per stage for all lattices, no measurement; replayed from a
 * captured graph as qcb_batch_flow is. */
int qcb_batch_flow_steps(qcb_batch* b, double eps, int n_steps);
/* Launch schedule of the heatbath: the order of the 8 half-sweeps (hs = 2*mu +
 * parity; any permutation is a valid Gibbs sweep) and a persistent grid size
 * (0 = one block per tile, else a multiple of 8 blocks striding over tiles). */
int qcb_batch_set_schedule(qcb_batch* b, const int order[8], uint32_t hb_grid);
int qcb_batch_get_schedule(qcb_batch* b, int order[8], uint32_t* hb_grid);
/* The heatbath's launch layout: the lattices split into n_streams groups of
 * about equal work, each on its own stream (a group's tail overlaps the next
 * group's half-sweep), and work-order tiles of each lattice capped at
 * tile_caps = {ty, tz, tt} (largest divisors of the extents not above them).
 * Results are bitwise those of one stream; the tiles change only the last bits
 * of the FP64 global sums (their block partition). */
int qcb_batch_set_layout(qcb_batch* b, int n_streams, const int tile_caps[3]);
int qcb_batch_get_layout(qcb_batch* b, int* n_streams, int tile_caps[3]);
/* Times candidate schedules on scratch copies and keeps the fastest: half-sweep
 * order x persistent grid x stream groups, then tile caps. */
int qcb_batch_autotune(qcb_batch* b, int n_hits, int reps, char* report, size_t report_len);
/* hipGraph of sweeps_per_graph sweeps; replays continue the Philox sweep index. */
int qcb_batch_graph_capture(qcb_batch* b, int n_hits, int sweeps_per_graph, uint32_t first_sweep);
/* Synchronises; QCB_EBUDGET as qcb_batch_hb_sweeps. */
int qcb_batch_graph_launch(qcb_batch* b, int n_replays);

/* ---- flow-measurement service (continuous batching) ---------------------- */
/* A request = a lattice (dims, beta, seed) generated on the device (Philox
 * hot start + n_therm heatbath sweeps, N_hits = n_hits, reunitarised every
 * reunit_every sweeps) and flowed to t_max with (t, Q, E_mean) recorded at
 * t = 0, eps, ..., t_max -- the schedule of topo_charge_flowed
 * (src/observables/observables_mpi.cpp:153-178) plus the final state.
 * Requests are admitted into a fixed arena of arena_bytes of HBM by footprint
 * (four link fields each) and advanced together, one sweep or one flow step
 * per service iteration, in batched launches on two streams.  Results do not
 * depend on what else is resident (bitwise).  No counterpart in the
 * reference, which runs one configuration per MPI job. */
typedef struct qcb_service qcb_service;
int qcb_service_create(qcb_ctx* ctx, int prec, uint64_t arena_bytes, int n_hits, int n_therm, int reunit_every,
                       double eps, qcb_service** out);
int qcb_service_destroy(qcb_service* s);
/* Seconds since the service was created (the clock of arrival/admit/done times). */
double qcb_service_clock(qcb_service* s);
int qcb_service_submit(qcb_service* s, uint64_t id, const int dims[4], double beta, uint64_t seed, double t_max,
                       double t_arrival);
/* One iteration: admit, advance every resident request, retire finished ones. */
int qcb_service_step(qcb_service* s, int* pending);
/* Pops one finished request (QCB_ESTATE if none): tQE holds n_meas triples;
 * times3 = (arrival, admit, done) on the service clock.  QCB_EBUDGET (with
 * every output filled in) if a heatbath draw of its thermalisation ran out of
 * attempts. */
int qcb_service_poll(qcb_service* s, uint64_t* id, double* tQE, int cap, int* n_meas, double* times3);
int qcb_service_status(qcb_service* s, int* queued, int* resident, int* finished, uint64_t* free_bytes,
                       uint64_t* iterations);

/* ---- ILDG / LIME configuration files ----------------------------------- */
/* The file layout of save_ildg_clime / read_ildg_clime (src/io/ildg.cpp:13-237):
 * an "ildg-format" XML record, then "ildg-binary-data" with big-endian FP64
 * links in the host AoS order.  The codec is this library's own (LIME spec);
 * no c-lime.  qcb_ildg_write/read need no GPU; their failure message is
 * qcb_last_error(NULL). */
int qcb_ildg_write(const char* path, const int dims[4], const double* aos);
int qcb_ildg_read(const char* path, const int dims[4], double* aos);
int qcb_ildg_save(qcb_lattice* lat, const char* path);
int qcb_ildg_load(qcb_lattice* lat, const char* path);

/* ---- event-chain Monte Carlo (src/ecmc/ecmc_mpi_cb.cpp, src/bin/gauge_ecmc_ncb.cpp) ----
 * The lattice (FP64) is cut into blocks of block[0..3] sites, the reference's
 * per-rank L_core^4 block with its c = 0 faces frozen
 * (src/geometry/GeometryCB.cpp:33-56): link (x, mu) is frozen when
 * x_d mod block_d == 0 for some d != mu.  One chain per block, all chains
 * at once on the device; each keeps its LocalChainState
 * (src/ecmc/ecmc_mpi_cb.h:15-35) between samples, as sample_persistant does
 * across calls.  Philox4x32-10 counters, key = seed, per chain c and event
 * index n: ctr = {c, n_lo, 0xEC000000 | n_hi, call} (call 0, 1: six
 * gammas, lift select and accept; 2: set index; 3, 4: refresh site, dir,
 * epsilon and its Poisson budget; 5: refresh-R budget; 6: theta_sample),
 * Haar R refreshes on {c, n_lo, 0xED000000 | n_hi, 0..8}, matrix-set element
 * 2m+1 of generation g on {c, g, 0xE5000000 | m, 0..2}. */
typedef struct qcb_ecmc qcb_ecmc;
typedef struct {
    double beta;
    double theta_sample;        /* param_theta_sample (src/io/params.h:9-17) */
    double theta_refresh_site;  /* param_theta_refresh_site */
    double theta_refresh_R;     /* param_theta_refresh_R */
    int poisson;                /* exponential budgets instead of fixed ones */
    double epsilon_set;         /* spread of the near-identity matrix set */
    int block[4];               /* chain block extents (>= 4, dividing the lattice) */
    uint64_t seed;
    uint64_t max_events;        /* per chain and sample call; 0 = 1e9.  Reaching it ends that
                                   chain's sample early and qcb_ecmc_get_chains reports QCB_EBUDGET */
} qcb_ecmc_params;
#define QCB_ECMC_CHAIN_DOUBLES 32 /* R[18], th_refresh_site, th_refresh_R, par_refresh_site,
                                     par_refresh_R, events, lifts, revs, set_counter, set_gen,
                                     rng_n, site (lex), mu, epsilon, initialized */
/* mpi::ecmccb::sample_persistant chain states for every block of lat (chain c
 * is block c in x-fastest block order); lat must outlive the handle.  The
 * handle also holds a chain-local copy of lat's links (72 B per link, FP64),
 * filled and written back by every qcb_ecmc_sample call */
int qcb_ecmc_create(qcb_lattice* lat, const qcb_ecmc_params* p, qcb_ecmc** out);
int qcb_ecmc_destroy(qcb_ecmc* e);
int qcb_ecmc_n_chains(qcb_ecmc* e, int* n);
/* n_samples calls of sample_persistant on every chain (src/ecmc/ecmc_mpi_cb.cpp:496-637) */
int qcb_ecmc_sample(qcb_ecmc* e, int n_samples);
/* chain states, QCB_ECMC_CHAIN_DOUBLES doubles per chain (synchronises) */
int qcb_ecmc_get_chains(qcb_ecmc* e, double* out);
/* restore chain states (io::load_state, src/io/io.cpp:200-232); an
 * initialized chain's site must lie in its own block, its active link (site, mu)
 * off the frozen faces and epsilon = +-1, else QCB_EINVAL and nothing changes */
int qcb_ecmc_set_chains(qcb_ecmc* e, const double* in);

/* ---- device math probes (checking aids for the register-level routines) ---- */
#define QCB_PROBE_EXP 0        /* out = exp(param * in) (utils.cpp:145-193), 18 in / 18 out; in is
                                  traceless anti-Hermitian: its upper triangle and Im diagonal are read */
#define QCB_PROBE_PROJ 1       /* out = P_TA(in) (utils.cpp:130-134) */
#define QCB_PROBE_REUNIT 2     /* out = Gram-Schmidt(in) (GaugeField.cpp:36-58) */
#define QCB_PROBE_KP 3         /* in = {k, u[4K+2]} per item, param = beta, K = (int)aux;
                                  out = {accepted attempt or -1, x0..x3} (heatbath.cpp:8-36) */
#define QCB_PROBE_CM_HIT 4     /* in = {U 18, A 18, u[3*(4K+2)]}, param = beta, aux = K:
                                  out = U after one CM hit (heatbath_mpi.cpp:11-34) */
#define QCB_PROBE_UNIFORM 5    /* in = {w} (a 32-bit Philox word as a double), 1 in / 1 out:
                                  out = the heatbath uniform (w + 1/2) 2^-32 in the precision prec */
int qcb_math_probe(qcb_ctx* ctx, int prec, int op, int n, const double* in, double param, int aux,
                   double* out);

#ifdef __cplusplus
}
#endif
#endif
