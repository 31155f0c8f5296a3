/*
 * qo_refrun.c -- one-rank restatement of the reference HB driver loop with
 * per-phase timers (TEST INFRASTRUCTURE ONLY; it is also the CPU baseline).
 *
 * Follows src/bin/gauge_hb_ncb.cpp:18-201 for n_core_dims = 1:
 *   rng = mt19937_64(seed + rank)                         (:28)
 *   hot start unless cold_start                          (:29-31)
 *   thermalisation: shift -> HB -> (plaquette)            (:89-106)
 *   sampling: shift -> HB -> plaquette (stored) -> flow   (:116-151)
 * Shifts draw s_x,s_y,s_z,s_t from uniform_int_distribution(0, L-1) and
 * translate U'(x) = U(x - s) (src/mpi/Shift.cpp:163-186).  The flow uses the
 * hard-coded eps = 0.02 (src/bin/gauge_hb_ncb.cpp:57).
 */
#include "qo.h"

#include <stdlib.h>
#include <string.h>

int qo_ref_run(const qo_run_params* p, double* plaq_out, double* tqe_out, int* shifts_out,
               double* final_links, qo_run_stats* st) {
    int L[4] = {p->L, p->L, p->L, p->L};
    qo_lat lat;
    if (qo_lat_init(&lat, L) != 0) return -1;
    memset(st, 0, sizeof *st);
    qo_mt64 g;
    qo_mt64_seed(&g, (uint64_t)(int64_t)(p->seed + 0));
    if (!p->cold_start) qo_lat_hot_ref(&lat, &g);
    const double eps = 0.02;
    int nsh = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int n = pass == 0 ? p->N_therm : p->N_shift;
        for (int i = 0; i < n; ++i) {
            double t0 = qo_now();
            int s[4];
            for (int d = 0; d < 4; ++d) s[d] = (int)qo_mt64_uniform_int(&g, 0, (uint64_t)(p->L - 1));
            if (shifts_out) memcpy(shifts_out + 4 * nsh, s, sizeof s);
            ++nsh;
            qo_translate(&lat, s);
            double t1 = qo_now();
            for (int sw = 0; sw < p->N_sweeps; ++sw) st->link_visits += qo_hb_sweep_ref(&lat, p->beta, p->N_hits, &g);
            double t2 = qo_now();
            st->t_shift += t1 - t0;
            st->t_hb += t2 - t1;
            if (i % p->N_shift_plaquette == 0) {
                double pl = qo_plaquette(&lat);
                double t3 = qo_now();
                st->t_plaq += t3 - t2;
                if (pass == 1) {
                    if (plaq_out) plaq_out[st->n_plaq] = pl;
                    st->n_plaq++;
                }
            }
            if (pass == 1 && p->topo && i % p->N_shift_topo == 0) {
                double t4 = qo_now();
                if (tqe_out) qo_topo_charge_flowed(&lat, eps, p->N_steps_gf, p->N_rk_steps, tqe_out + 3 * st->n_tqe);
                st->n_tqe += p->N_steps_gf;
                st->site_steps += (long)lat.V * p->N_steps_gf * p->N_rk_steps;
                st->t_flow += qo_now() - t4;
            }
        }
    }
    if (final_links) memcpy(final_links, lat.U, lat.V * 72 * sizeof(double));
    qo_lat_free(&lat);
    return 0;
}
