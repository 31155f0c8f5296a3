#!/usr/bin/env python3
"""Benchmark of the MI355X heatbath / Wilson-flow path.

Workload (one GPU): 32^4 lattice, beta = 6.0, hot start from seed 11, FP32
links and staples with FP64 global sums.  A step is one full heatbath sweep
(8 (mu, parity) half-sweeps, N_hits = 2 Cabibbo-Marinari hits per link) over
the lattice; value = link-updates/s (4V per sweep).  The Wilson-flow rate
(RK3 site-steps/s, eps = 0.02) and the FP64 rates of both are reported
beside it, each timed with HIP events on the library's stream.

cpu_baseline: the FP64 reference-order CPU restatement (frozen links,
lexicographic order, mt19937_64, one core) on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ecmc-qncb_amd"))

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E
HB_BYTES_PER_LINK = 162         # x sizeof(real): R+W U_mu(p), R U_mu(1-p), R 3 U_nu, per link-update
WF_BYTES_PER_SITE = 720         # x sizeof(real): 3 fused RK3 stages with double-buffered W and one Z field
METRIC = "HB link-updates/s & WF site-steps/s (FP32 vs FP64); flow-service p50/p99"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_hb(qcb, ctx, L, prec, steps, warmup, n_hits=2, beta=6.0, seed=11):
    lat = qcb.Lattice(ctx, (L,) * 4, prec)
    lat.hot(seed)
    lat.hb_sweeps(beta, n_hits, warmup, seed=seed, sweep_idx=0)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timing_begin()
    lat.hb_sweeps(beta, n_hits, steps, seed=seed, sweep_idx=warmup)
    dev_ms = ctx.timing_end()
    wall = time.perf_counter() - t0
    plaq = lat.plaquette()
    V = lat.V
    lat.close()
    s = 4 if prec == qcb.F32 else 8
    launch_s = dev_ms * 1e-3 / (8 * steps)
    achieved = HB_BYTES_PER_LINK * s * (V // 2) / launch_s / 1e9
    return dict(link_updates_per_s=4 * V * steps / wall, wall_s=wall, dev_ms=dev_ms,
                ms_per_sweep=wall * 1e3 / steps, avg_launch_us=launch_s * 1e6,
                achieved_gbs=achieved, plaquette=plaq)


def time_wf(qcb, ctx, L, prec, steps, warmup, eps=0.02, seed=11):
    lat = qcb.Lattice(ctx, (L,) * 4, prec)
    lat.hot(seed)
    lat.hb_sweeps(6.0, 2, 4, seed=seed)
    lat.flow_copy()
    lat.flow_rk3(eps, warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.timing_begin()
    lat.flow_rk3(eps, steps)
    dev_ms = ctx.timing_end()
    wall = time.perf_counter() - t0
    V = lat.V
    lat.close()
    s = 4 if prec == qcb.F32 else 8
    launch_s = dev_ms * 1e-3 / (3 * steps)
    achieved = WF_BYTES_PER_SITE / 3 * s * V / launch_s / 1e9
    return dict(site_steps_per_s=V * steps / wall, wall_s=wall, dev_ms=dev_ms,
                avg_stage_launch_us=launch_s * 1e6, achieved_gbs=achieved)


def cpu_baseline(L=32):
    """One reference-order FP64 sweep (N_hits = 2, beta = 6.0) of a hot L^4
    lattice on one core; counts only the non-frozen links it updates."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as qo
    lat = qo.Lattice((L,) * 4)
    lat.hot_philox(11)
    mt = qo.Mt64(11)
    t0 = time.perf_counter()
    visits = lat.sweep_ref(6.0, 2, mt)
    dt = time.perf_counter() - t0
    return dict(value=visits / dt, unit="link-updates/s", cores=1, kind="port",
                sample=f"1 reference-order sweep (frozen links, lexicographic, mt19937_64, N_hits=2, "
                       f"beta=6.0, FP64) of a hot {L}^4 lattice: {visits} active links in {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--flow-steps", type=int, default=10)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "independent lattices shard one-or-more per GPU with no communication", file=sys.stderr)
        sys.exit(2)

    import qcb
    ctx = qcb.Context(0)
    L = args.L
    log(f"[bench] HB FP32 {L}^4: warmup {args.warmup}, timed {args.steps} sweeps")
    hb32 = time_hb(qcb, ctx, L, qcb.F32, args.steps, args.warmup)
    log(f"[bench] HB FP32 {hb32['link_updates_per_s']:.3e} link-updates/s, {hb32['achieved_gbs']:.0f} GB/s")
    log(f"[bench] WF FP32 {L}^4: {args.flow_steps} RK3 steps")
    wf32 = time_wf(qcb, ctx, L, qcb.F32, args.flow_steps, 2)
    log(f"[bench] WF FP32 {wf32['site_steps_per_s']:.3e} site-steps/s, {wf32['achieved_gbs']:.0f} GB/s")
    comp = {"hb_fp32": hb32, "wf_fp32": wf32}
    if not args.no_fp64:
        hb64 = time_hb(qcb, ctx, L, qcb.F64, max(2, args.steps // 2), 1)
        log(f"[bench] HB FP64 {hb64['link_updates_per_s']:.3e} link-updates/s, {hb64['achieved_gbs']:.0f} GB/s")
        wf64 = time_wf(qcb, ctx, L, qcb.F64, max(2, args.flow_steps // 2), 1)
        log(f"[bench] WF FP64 {wf64['site_steps_per_s']:.3e} site-steps/s, {wf64['achieved_gbs']:.0f} GB/s")
        comp.update(hb_fp64=hb64, wf_fp64=wf64)
    ctx.close()

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (reference-order FP64 sweep, 1 core)")
        cpu = cpu_baseline(L)
        log(f"[bench] CPU {cpu['value']:.3e} link-updates/s")

    V = L ** 4
    out = {
        "metric": METRIC,
        "value": hb32["link_updates_per_s"],
        "unit": "link-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hb32["ms_per_sweep"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: device hot start (Haar SU(3), Philox seed 11)",
        "config": {"workload": f"heatbath sweep, {L}^4 periodic lattice, beta=6.0, N_hits=2, FP32 links, FP64 sums",
                   "lattice": [L, L, L, L], "sites": V, "beta": 6.0, "n_hits": 2},
        "roofline": {"bound": "hbm", "achieved": hb32["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": hb32["achieved_gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_hb_half<float> (one (mu,parity) half-sweep)",
                     "bytes_per_launch": HB_BYTES_PER_LINK * 4 * (V // 2)},
        "cpu_baseline": cpu,
        "components": comp,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
