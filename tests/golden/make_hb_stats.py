"""Writes tests/golden/hb_plaquette_8x4.json: the mean plaquette of the
reference-method heatbath chain at 8^4, beta = 6.0, N_hits = 2.

The chains are the oracle's restatement of the reference driver
(src/bin/gauge_hb_ncb.cpp:18-201): cold start, mt19937_64(seed), one random
shift then one lexicographic sweep over the non-frozen links per step
(src/heatbath/heatbath_mpi.cpp:37-69), plaquette after every step.  Four
independent seeds; each chain drops its first N_THERM steps and is binned into
N_BINS bins.  The fixture holds the per-chain and combined means and binned
errors, so a device chain can be compared with the reference algorithm's
equilibrium value within its statistical error.

Run from the repository root:  python tests/golden/make_hb_stats.py   (~3 min)
"""
import json
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

L, BETA, N_HITS = 8, 6.0, 2
N_THERM, N_SHIFT, N_BINS = 100, 2000, 20
SEEDS = (12345, 12346, 12347, 12348)


def chain(seed):
    import oracle_lib as o
    r = o.ref_run(L=L, cold_start=1, seed=seed, N_therm=N_THERM, N_shift=N_SHIFT, beta=BETA, N_sweeps=1,
                  N_hits=N_HITS, N_shift_plaquette=1, topo=0, N_shift_topo=10, N_steps_gf=50, N_rk_steps=1)
    p = np.asarray(r["plaq"][1:], dtype=float)  # entry 0 is measured before the first sampling step
    bins = p[: len(p) // N_BINS * N_BINS].reshape(N_BINS, -1).mean(axis=1)
    return dict(seed=seed, n=len(p), mean=float(bins.mean()), err=float(bins.std(ddof=1) / np.sqrt(N_BINS)))


def main():
    import oracle_lib as o
    o.build()
    with Pool(len(SEEDS)) as pool:
        chains = pool.map(chain, SEEDS)
    means = np.array([c["mean"] for c in chains])
    errs = np.array([c["err"] for c in chains])
    w = 1.0 / errs ** 2
    out = dict(lattice=[L] * 4, beta=BETA, n_hits=N_HITS, n_therm=N_THERM, n_shift=N_SHIFT, n_bins=N_BINS,
               method="reference order: frozen links, random shifts, lexicographic sweep, mt19937_64",
               chains=chains, mean=float((w * means).sum() / w.sum()), err=float(1.0 / np.sqrt(w.sum())),
               spread_of_chain_means=float(means.std(ddof=1) / np.sqrt(len(means))))
    with open(os.path.join(HERE, "hb_plaquette_8x4.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
