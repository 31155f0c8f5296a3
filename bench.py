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
# clover Q and E per site: 6 planes x 4 leaves x 3 SU(3) products of 216 flops,
# plus the adjoint differences and the three pair traces
CLOVER_FLOPS_PER_SITE = 15100
VALU_PEAK_TFLOPS = 78.6         # MI355X FP64 / unpacked FP32 vector rate (packed FP32 is switched off)


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


def warm_device(qcb, ctx, L, secs=0.5):
    """Heatbath sweeps on an FP32 L^4 lattice for secs of wall time, so that
    the first timed leg starts at the clocks the GPU holds under load (a cold
    device runs its first ~0.1-0.3 s of work ~10% slower)."""
    lat = qcb.Lattice(ctx, (L,) * 4, qcb.F32)
    lat.hot(1)
    t0, k = time.perf_counter(), 0
    while time.perf_counter() - t0 < secs:
        lat.hb_sweeps(6.0, 2, 10, seed=1, sweep_idx=k)
        ctx.sync()
        k += 10
    lat.close()


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


def time_clover(qcb, ctx, L, prec, reps=10, seed=11):
    """clover Q, E of a lattice one heatbath sweep from a hot start: device time
    of one measurement (k_clover, the fixed-order finalisation, the two-double
    readback), averaged over reps"""
    lat = qcb.Lattice(ctx, (L,) * 4, prec)
    lat.hot(seed)
    lat.hb_sweeps(6.0, 2, 1, seed=seed)
    lat.clover_qe()
    ctx.sync()
    ctx.timing_begin()
    for _ in range(reps):
        q, e = lat.clover_qe()
    ms = ctx.timing_end() / reps
    V = lat.V
    lat.close()
    tflops = CLOVER_FLOPS_PER_SITE * V / (ms * 1e-3) / 1e12
    return dict(ms=ms, tflops=tflops, frac_of_valu_peak=tflops / VALU_PEAK_TFLOPS, Q=q, E=e)


def measured_traffic(kernel="k_hb_half<float"):
    """HBM bytes per launch of a kernel from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE scaled by the plain-copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "hb_flow_32x4_rocprof.md")
    if not os.path.exists(path):
        return None
    vals, section = [], False
    for line in open(path):
        if line.startswith("## "):
            section = line.startswith("## HBM bytes")
        if section and line.startswith("| `" + kernel):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            try:
                vals.append(float(cells[3]) * 1e6)
            except (ValueError, IndexError):
                pass
    return sum(vals) / len(vals) if vals else None


def batch64(qcb, ctx, sweeps=20, flow_steps=50):
    """64 lattices, 16 each of L in {8,10,12,16}, beta ~ U[5.8,6.2] (seed 3),
    FP32 links: autotuned schedule, sweeps replayed from a hipGraph, then a
    batched 50-step flow with a measurement per step."""
    import numpy as np
    rng = np.random.default_rng(3)
    dims = [(L,) * 4 for L in (8, 10, 12, 16) for _ in range(16)]
    betas = rng.uniform(5.8, 6.2, size=64)
    b = qcb.Batch(ctx, dims, qcb.F32, betas, list(range(1, 65)))
    b.hot()
    report = b.autotune(2, 2)
    order, grid = b.get_schedule()
    n_streams, tiles = b.get_layout()
    b.graph_capture(2, 10, first_sweep=0)
    b.graph_launch(1)
    ctx.sync()
    reps = max(1, sweeps // 10)
    ctx.timing_begin()
    b.graph_launch(reps)
    ms = ctx.timing_end()
    V = int(sum(np.prod(d) for d in dims))
    P = b.plaquette()
    t0 = time.perf_counter()
    b.flow(0.02, flow_steps)
    dt = time.perf_counter() - t0
    # the same flow without measurements (one launch per stage for all 64
    # lattices), device time from HIP events on the library stream
    ctx.sync()
    ctx.timing_begin()
    b.flow_steps(0.02, flow_steps)
    fms = ctx.timing_end()
    b.close()
    flow_rate = V * flow_steps / (fms * 1e-3)
    return dict(lattices=64, sites=V, link_updates_per_s=4 * V * 10 * reps / (ms * 1e-3),
                ms_per_sweep=ms / (10 * reps),
                schedule={"order": order, "grid": grid, "streams": n_streams, "tile_caps": list(tiles)},
                autotune=report.strip().split("\n"),
                flow_site_steps_per_s_incl_measure=V * flow_steps / dt,
                flow_site_steps_per_s=flow_rate, flow_achieved_gbs=WF_BYTES_PER_SITE * 4 * flow_rate / 1e9,
                mean_plaquette=float(P.mean()))


def serve(qcb, ctx, n_req=60, rate=20.0, n_therm=200, seed=2026):
    """Flow-measurement service under Poisson arrivals: L in {12,16,24},
    T = 2L, beta in {5.9,6.0,6.1}, t_max ~ U[0.5, 2.0], eps = 0.02, FP32 links
    with FP64 sums.  Each request is generated on the device: hot start and
    n_therm = 200 heatbath sweeps (reunitarised every 4), which brings <P>
    at beta = 6 to within ~1e-4 of equilibrium, then its flow.  Latency =
    result ready - scheduled arrival."""
    import numpy as np
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
    Ls = rng.choice([12, 16, 24], size=n_req)
    betas = rng.choice([5.9, 6.0, 6.1], size=n_req)
    tmax = np.round(rng.uniform(0.5, 2.0, size=n_req) / 0.02) * 0.02
    svc = qcb.Service(ctx, qcb.F32, 64 << 30, 2, n_therm, 4, 0.02)
    t0 = svc.clock()
    nxt, done, lat = 0, {}, {}
    while len(done) < n_req:
        now = svc.clock() - t0
        while nxt < n_req and arrivals[nxt] <= now:
            L = int(Ls[nxt])
            svc.submit(nxt, (L, L, L, 2 * L), float(betas[nxt]), 1000 + nxt, float(tmax[nxt]), t0 + arrivals[nxt])
            nxt += 1
        st = svc.status()
        if st["queued"] + st["resident"] == 0:
            time.sleep(max(0.0, min(0.01, arrivals[nxt] - now)) if nxt < n_req else 0.0)
            continue
        svc.step()
        for rid, tqe, (ta, tadm, td) in svc.poll():
            done[rid] = tqe
            lat[rid] = td - ta
    t_end = svc.clock() - t0
    it = svc.status()["iterations"]
    svc.close()
    l = np.array([lat[i] for i in range(n_req)])
    sites = sum(int(L) ** 3 * 2 * int(L) for L in Ls)
    return dict(requests=n_req, rate_per_s=rate, p50_ms=float(np.percentile(l, 50) * 1e3),
                p99_ms=float(np.percentile(l, 99) * 1e3), mean_ms=float(l.mean() * 1e3),
                requests_per_s=n_req / t_end, wall_s=t_end, iterations=int(it), n_therm=n_therm,
                mean_sites=sites / n_req)


def ecmc(qcb, ctx, L=32, block=4, samples=2):
    """Event chains (FP64) on an L^4 lattice cut into block^4 chain blocks,
    the reference's ECMC input-template budgets (theta_sample 1800,
    theta_refresh_site 1500, theta_refresh_R 150, epsilon_set 0.15) at
    beta = 6.0, a random shift before every sample; events/s over all chains."""
    import numpy as np
    lat = qcb.Lattice(ctx, (L,) * 4)
    lat.hot(5)
    lat.hb_sweeps(6.0, 2, 50, seed=5)
    e = qcb.Ecmc(lat, qcb.ecmc_params(beta=6.0, theta_sample=1800.0, theta_refresh_site=1500.0,
                                      theta_refresh_R=150.0, epsilon_set=0.15, block=(block,) * 4, seed=5))
    rng = np.random.default_rng(5)
    lat.translate(rng.integers(0, L, size=4))
    e.sample(1)
    ev = qcb.ECMC_FIELDS.index("events")
    n0 = e.chains()[:, ev].sum()
    ctx.timing_begin()
    for _ in range(samples):
        lat.translate(rng.integers(0, L, size=4))
        e.sample(1)
    ms = ctx.timing_end()
    n1 = e.chains()[:, ev].sum()
    out = dict(lattice=[L] * 4, block=[block] * 4, chains=e.n_chains, events_per_s=(n1 - n0) / (ms * 1e-3),
               ms_per_sample=ms / samples, events_per_chain_per_sample=(n1 - n0) / samples / e.n_chains,
               plaquette=lat.plaquette())
    e.close()
    lat.close()
    return out


def ecmc_cpu_baseline(L=8):
    """The reference's one-rank event chain (whole L^4 lattice one block,
    mt19937_64, FP64) restated by the oracle, one core: events/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as qo
    lat = qo.Lattice((L,) * 4)
    lat.hot_philox(5)
    p = qo.ecmc_params(beta=6.0, theta_sample=1800.0, theta_refresh_site=1500.0, theta_refresh_R=150.0)
    ch = qo.EcmcChain()
    mt = qo.Mt64(12345)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 3.0:
        lat.ecmc_ref(p, ch, mt)
    dt = time.perf_counter() - t0
    return dict(value=ch.events / dt, unit="events/s", cores=1, kind="port",
                sample=f"reference-order event chain on a {L}^4 one-rank lattice: {ch.events} events in {dt:.1f} s")


def precision_report(qcb, ctx, L=32, sweeps=500, flow_steps=150, seed=11, beta=6.0):
    """FP32 links (reunitarised every 4 sweeps) against the FP64 GPU run on
    identical Philox uniforms: |<P>| differences along the chain, then the
    Wilson flow of the final configurations.  'flow_only' flows the FP64
    final configuration in both precisions, isolating the flow's rounding
    from the divergence of the two Markov chains."""
    import numpy as np
    a = qcb.Lattice(ctx, (L,) * 4, qcb.F64)
    b = qcb.Lattice(ctx, (L,) * 4, qcb.F32)
    for x in (a, b):
        x.hot(seed)
    marks = {1, 2, 5, 10, 20, 50, 100, 200, 300, 400, 500}
    chain, pa_eq, pb_eq = [], [], []
    n_eq0 = min(100, sweeps // 5)
    for s_ in range(sweeps):
        a.hb_sweeps(beta, 2, 1, seed=seed, sweep_idx=s_)
        b.hb_sweeps(beta, 2, 1, seed=seed, sweep_idx=s_)
        if (s_ + 1) % 4 == 0:
            b.reunitarize()
        if s_ + 1 in marks or s_ >= n_eq0:
            pa, pb = a.plaquette(), b.plaquette()
            if s_ + 1 in marks:
                chain.append(dict(sweep=s_ + 1, P64=pa, P32=pb, dP=abs(pa - pb)))
            if s_ >= n_eq0:
                pa_eq.append(pa)
                pb_eq.append(pb)

    def mean_err(x, nbin=20):
        x = np.asarray(x)
        bins = x[: len(x) // nbin * nbin].reshape(nbin, -1).mean(1)
        return float(bins.mean()), float(bins.std(ddof=1) / np.sqrt(nbin))

    m64, e64 = mean_err(pa_eq)
    m32, e32 = mean_err(pb_eq)
    equilibrium = dict(sweeps=f"{n_eq0 + 1}..{sweeps}", P64=m64, err64=e64, P32=m32, err32=e32,
                       diff=m32 - m64, sigma=float(np.hypot(e64, e32)))
    # the same configuration measured in both precisions
    c0 = qcb.Lattice(ctx, (L,) * 4, qcb.F32)
    c0.set_links(a.get_links())
    p_meas = dict(P64=a.plaquette(), P32_same_links=c0.plaquette())
    p_meas["dP"] = abs(p_meas["P64"] - p_meas["P32_same_links"])
    c0.close()
    ta = a.topo_charge_flowed(0.02, flow_steps + 1, 1)
    tb = b.topo_charge_flowed(0.02, flow_steps + 1, 1)
    c = qcb.Lattice(ctx, (L,) * 4, qcb.F32)
    c.set_links(a.get_links())
    tc = c.topo_charge_flowed(0.02, flow_steps + 1, 1)
    rows = []
    for t in (0.5, 1.0, 2.0, 3.0):
        i = int(round(t / 0.02))
        if i >= len(ta):
            continue
        rows.append(dict(t=t, Q64=ta[i, 1], E64=ta[i, 2], dQ_chain=abs(tb[i, 1] - ta[i, 1]),
                         dE_rel_chain=abs(tb[i, 2] - ta[i, 2]) / ta[i, 2],
                         dQ_flow_only=abs(tc[i, 1] - ta[i, 1]), dE_rel_flow_only=abs(tc[i, 2] - ta[i, 2]) / ta[i, 2]))
    for x in (a, b, c):
        x.close()
    return dict(lattice=[L] * 4, beta=beta, sweeps=sweeps, flow_steps=flow_steps, chain=chain,
                equilibrium=equilibrium, plaquette_measurement=p_meas, flow=rows)


def ecmc_report(qcb, ctx, L=32, hb_sweeps=3000, samples=1500, seed=29):
    """Equilibrium mean plaquette of the device event chains (4^4 blocks, a
    random shift before every sample) against the device heatbath, both FP64
    at beta = 6.0 on an L^4 lattice; binned errors (20 bins after 10% burn-in)."""
    import numpy as np

    def binned(x, nb=20):
        x = np.asarray(x)
        n = len(x) // nb
        m = x[:n * nb].reshape(nb, n).mean(axis=1)
        return float(m.mean()), float(m.std(ddof=1) / np.sqrt(nb))

    dims = (L,) * 4
    h = qcb.Lattice(ctx, dims)
    h.hot(seed)
    h.hb_sweeps(6.0, 2, 200, seed=seed)
    ph = []
    for s in range(hb_sweeps):
        h.hb_sweeps(6.0, 2, 1, seed=seed, sweep_idx=200 + s)
        ph.append(h.plaquette())
    g = qcb.Lattice(ctx, dims)
    g.set_links(h.get_links())
    h.close()
    e = qcb.Ecmc(g, qcb.ecmc_params(beta=6.0, theta_sample=1000.0, theta_refresh_site=430.0,
                                    theta_refresh_R=51.0, seed=17, block=(4, 4, 4, 4), max_events=10 ** 7))
    rng = np.random.default_rng(19)
    pe = []
    for s in range(samples):
        g.translate(rng.integers(0, L, size=4))
        e.sample(1)
        pe.append(g.plaquette())
        if s % 250 == 0:
            log(f"[ecmc-report] sample {s}")
    e.chains()   # raises if any chain stopped early
    e.close()
    g.close()
    m_hb, e_hb = binned(ph[hb_sweeps // 10:])
    m_ec, e_ec = binned(pe[samples // 10:])
    sigma = float(np.hypot(e_hb, e_ec))
    return dict(lattice=list(dims), beta=6.0, heatbath=dict(sweeps=hb_sweeps, P=m_hb, err=e_hb),
                ecmc=dict(samples=samples, chains=(L // 4) ** 4, P=m_ec, err=e_ec),
                diff=m_ec - m_hb, sigma=sigma, n_sigma=abs(m_ec - m_hb) / sigma)


def wf_cpu_baseline(L=16, steps=2):
    """The reference's RK3 Wilson flow step (FP64, src/flow/gradient_flow.cpp:15-128)
    restated by the oracle, one core, on a hot L^4 lattice: site-steps/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as qo
    lat = qo.Lattice((L,) * 4)
    lat.hot_philox(11)
    t0 = time.perf_counter()
    for _ in range(steps):
        lat.flow_step(0.02)
    dt = time.perf_counter() - t0
    return dict(value=steps * L ** 4 / dt, unit="site-steps/s", cores=1, kind="port",
                sample=f"{steps} RK3 steps (eps = 0.02, FP64) of a hot {L}^4 lattice in {dt:.1f} s")


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
    # ~0.1 s of heatbath sweeps: the GPU reaches its steady clocks under load
    # only after that (10 warmup sweeps time 62 us per half-sweep, 200 time 56 us)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--flow-steps", type=int, default=10)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-ecmc", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--precision-report", action="store_true",
                    help="run the FP32-vs-FP64 comparison (32^4, 500 sweeps, 150 flow steps) and print it")
    ap.add_argument("--serve-requests", type=int, default=60)
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the flow-service leg and print its JSON (for a kernel trace of that leg)")
    ap.add_argument("--ecmc-report", action="store_true",
                    help="compare the event chains' equilibrium plaquette with the heatbath's (32^4, FP64) and print it")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: qcb has no multi-GPU path (asked for --gpus {args.gpus}); "
              "independent lattices shard one-or-more per GPU with no communication", file=sys.stderr)
        sys.exit(2)

    import qcb
    ctx = qcb.Context(0)
    L = args.L
    if args.precision_report:
        print(json.dumps(precision_report(qcb, ctx, L)))
        ctx.close()
        return
    if args.ecmc_report:
        print(json.dumps(ecmc_report(qcb, ctx, L)))
        ctx.close()
        return
    if args.serve_only:
        print(json.dumps(serve(qcb, ctx, args.serve_requests)))
        ctx.close()
        return
    # The other legs run first and the headline last: the GPU then starts the
    # headline's own warmup sweeps already at its steady clocks under load
    # (on a cold device the first ~0.1 s of sweeps run ~10% slower).
    comp = {}
    warm_device(qcb, ctx, L)
    log(f"[bench] WF FP32 {L}^4: {args.flow_steps} RK3 steps")
    wf32 = time_wf(qcb, ctx, L, qcb.F32, args.flow_steps, args.flow_steps)
    log(f"[bench] WF FP32 {wf32['site_steps_per_s']:.3e} site-steps/s, {wf32['achieved_gbs']:.0f} GB/s")
    comp["wf_fp32"] = wf32
    cl = {"flops_per_site": CLOVER_FLOPS_PER_SITE, "valu_peak_tflops": VALU_PEAK_TFLOPS,
          "fp32": time_clover(qcb, ctx, L, qcb.F32)}
    if not args.no_fp64:
        cl["fp64"] = time_clover(qcb, ctx, L, qcb.F64)
    log(f"[bench] clover Q/E {L}^4: FP32 {cl['fp32']['ms']:.3f} ms ({cl['fp32']['tflops']:.1f} TFLOP/s)")
    comp["clover_qe"] = cl
    if not args.no_fp64:
        hb64 = time_hb(qcb, ctx, L, qcb.F64, max(2, args.steps // 2), max(2, args.steps // 2))
        log(f"[bench] HB FP64 {hb64['link_updates_per_s']:.3e} link-updates/s, {hb64['achieved_gbs']:.0f} GB/s")
        wf64 = time_wf(qcb, ctx, L, qcb.F64, max(2, args.flow_steps // 2), max(2, args.flow_steps // 2))
        log(f"[bench] WF FP64 {wf64['site_steps_per_s']:.3e} site-steps/s, {wf64['achieved_gbs']:.0f} GB/s")
        comp.update(hb_fp64=hb64, wf_fp64=wf64)
        # the plain FP64 GPU configuration: 16^4, hot start from seed 7, N_hits = 2
        hbp = time_hb(qcb, ctx, 16, qcb.F64, 40, 4, seed=7)
        wfp = time_wf(qcb, ctx, 16, qcb.F64, 20, 2, seed=7)
        log(f"[bench] plain FP64 16^4: HB {hbp['link_updates_per_s']:.3e} link-updates/s, "
            f"WF {wfp['site_steps_per_s']:.3e} site-steps/s")
        comp["plain_fp64_16x4"] = dict(hb=hbp, wf=wfp)
    if not args.no_batch:
        log("[bench] batch of 64 lattices (L in {8,10,12,16}), autotuned + hipGraph")
        bt = batch64(qcb, ctx)
        log(f"[bench] batch64 {bt['link_updates_per_s']:.3e} link-updates/s")
        comp["batch64_fp32"] = bt
    if not args.no_ecmc:
        log(f"[bench] event chains, {L}^4 in 4^4 blocks")
        ec = ecmc(qcb, ctx, L)
        log(f"[bench] ECMC {ec['events_per_s']:.3e} events/s ({ec['chains']} chains)")
        comp["ecmc_fp64"] = ec
    if not args.no_serve:
        log(f"[bench] flow service: {args.serve_requests} Poisson requests at 20/s")
        sv = serve(qcb, ctx, args.serve_requests)
        log(f"[bench] service p50 {sv['p50_ms']:.1f} ms p99 {sv['p99_ms']:.1f} ms, {sv['requests_per_s']:.2f} req/s")
        comp["flow_service"] = sv
    if not args.no_precision and not args.no_fp64:
        log(f"[bench] precision: FP32 against FP64 on identical uniforms, {L}^4, 500 sweeps + 150 flow steps")
        pr = precision_report(qcb, ctx, L)
        comp["precision_fp32_vs_fp64"] = pr
        log(f"[bench] precision: |dP| same links {pr['plaquette_measurement']['dP']:.1e}, "
            f"equilibrium dP {pr['equilibrium']['diff']:+.1e} ({pr['equilibrium']['sigma']:.1e} sigma unit)")
    log(f"[bench] HB FP32 {L}^4: warmup {args.warmup}, timed {args.steps} sweeps")
    hb32 = time_hb(qcb, ctx, L, qcb.F32, args.steps, args.warmup)
    log(f"[bench] HB FP32 {hb32['link_updates_per_s']:.3e} link-updates/s, {hb32['achieved_gbs']:.0f} GB/s")
    comp["hb_fp32"] = hb32
    if "hb_fp64" in comp:
        comp["fp32_over_fp64_hb"] = hb32["link_updates_per_s"] / comp["hb_fp64"]["link_updates_per_s"]
        comp["fp32_over_fp64_wf"] = wf32["site_steps_per_s"] / comp["wf_fp64"]["site_steps_per_s"]
        comp["fp32_32x4_over_plain_fp64_16x4_hb"] = hb32["link_updates_per_s"] / hbp["link_updates_per_s"]
        comp["fp32_32x4_over_plain_fp64_16x4_wf"] = wf32["site_steps_per_s"] / wfp["site_steps_per_s"]
    ctx.close()

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (reference-order FP64 sweep, 1 core)")
        cpu = cpu_baseline(L)
        log(f"[bench] CPU {cpu['value']:.3e} link-updates/s")
        if not args.no_ecmc:
            comp["ecmc_cpu_baseline"] = ecmc_cpu_baseline()
            log(f"[bench] CPU ECMC {comp['ecmc_cpu_baseline']['value']:.3e} events/s")
        comp["wf_cpu_baseline"] = wf_cpu_baseline()
        log(f"[bench] CPU WF {comp['wf_cpu_baseline']['value']:.3e} site-steps/s")

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
                     "frac": hb32["achieved_gbs"] / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "k_hb_half<float> (one (mu,parity) half-sweep)",
                     "bytes_per_launch": HB_BYTES_PER_LINK * 4 * (V // 2)},
        "cpu_baseline": cpu,
        "components": comp,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
