"""Event chains on the GPU (reference src/ecmc/ecmc_mpi_cb.cpp:12-637) against
the oracle's restatement on the same Philox counters: links and chain states
after a sample, frozen links untouched, SU(3) kept, and the equilibrium
plaquette of the chains against the heatbath's."""
import os

import numpy as np
import pytest

import oracle_lib as qo
import qcb

pytestmark = pytest.mark.gpu

F = {n: k for k, n in enumerate(qcb.ECMC_FIELDS)}


@pytest.fixture(scope="module")
def ctx():
    c = qcb.Context(0)
    yield c
    c.close()


def _pair(ctx, dims, seed):
    g = qcb.Lattice(ctx, dims)
    g.hot(seed)
    o = qo.Lattice(dims)
    o.set_links(g.get_links())
    return g, o


def _compare(chains_gpu, chains_or):
    for c, s in zip(chains_gpu, chains_or):
        assert int(c[F["site"]]) == s.site and int(c[F["mu"]]) == s.mu and int(c[F["eps"]]) == s.eps
        assert int(c[F["events"]]) == s.events and int(c[F["lifts"]]) == s.lifts and int(c[F["revs"]]) == s.revs
        assert int(c[F["set_gen"]]) == s.set_gen and int(c[F["rng_n"]]) == s.rng_n
        assert np.abs(c[:18] - np.array(s.R[:])).max() < 1e-9
        assert abs(c[F["par_refresh_site"]] - s.par_refresh_site) < 1e-8
        assert abs(c[F["par_refresh_R"]] - s.par_refresh_R) < 1e-8
        assert abs(c[F["th_refresh_site"]] - s.th_refresh_site) < 1e-8


# the last case has 6 chains (a wavefront holds 4) and one block along z and t
@pytest.mark.parametrize("poisson,dims,block", [(False, (8, 8, 8, 8), (4, 4, 4, 4)), (True, (8, 8, 8, 8), (4, 4, 8, 4)),
                                                (False, (12, 8, 4, 8), (4, 4, 4, 8))])
def test_chains_match_oracle(ctx, poisson, dims, block):
    g, o = _pair(ctx, dims, 3)
    # budgets with no common multiple: with fixed budgets 12 and 3 the sample
    # event and a refresh-R event coincide exactly, and rounding picks one
    kw = dict(beta=6.0, theta_sample=12.3, theta_refresh_site=5.1, theta_refresh_R=3.7, poisson=poisson,
              block=block)
    e = qcb.Ecmc(g, qcb.ecmc_params(seed=77, max_events=200000, **kw))
    ch = o.ecmc_chains(qo.ecmc_params(**kw))
    e.sample(1)
    assert o.ecmc_philox(qo.ecmc_params(**kw), ch, 77) == e.n_chains
    gc = e.chains()
    _compare(gc, ch)
    assert np.abs(g.get_links() - o.links).max() < 1e-10
    # a second sample continues the same chains
    g.set_links(o.links)
    e.sample(1)
    o.ecmc_philox(qo.ecmc_params(**kw), ch, 77)
    _compare(e.chains(), ch)
    assert np.abs(g.get_links() - o.links).max() < 1e-9
    e.close()


def test_frozen_links_untouched_and_su3(ctx):
    dims = (8, 8, 8, 12)
    g = qcb.Lattice(ctx, dims)
    g.hot(5)
    g.hb_sweeps(6.0, 2, 5, seed=5)
    U0 = g.get_links()
    e = qcb.Ecmc(g, qcb.ecmc_params(seed=3, block=(4, 4, 4, 4), max_events=200000))
    e.sample(4)
    U1 = g.get_links()
    c = np.indices(dims[::-1]).reshape(4, -1)[::-1]
    fr = np.zeros((c.shape[1], 4), dtype=bool)
    for mu in range(4):
        for d in range(4):
            if d != mu:
                fr[:, mu] |= (c[d] % 4) == 0
    assert np.array_equal(U1[fr], U0[fr])
    assert (np.abs(U1[~fr] - U0[~fr]).max(axis=(1, 2)) > 0).mean() > 0.3
    U = U1.reshape(-1, 3, 3)
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-12
    ch = e.chains()
    assert np.all(ch[:, F["initialized"]] == 1) and np.all(ch[:, F["events"]] >= ch[:, F["lifts"]])
    e.close()


def test_event_budget_is_reported(ctx):
    g = qcb.Lattice(ctx, (8, 8, 8, 8))
    g.hot(1)
    e = qcb.Ecmc(g, qcb.ecmc_params(seed=2, theta_sample=1e6, max_events=50))
    e.sample(1)
    with pytest.raises(qcb.QcbError):
        e.chains()
    e.close()
    with pytest.raises(qcb.QcbError):
        qcb.Ecmc(g, qcb.ecmc_params(block=(3, 4, 4, 4)))
    f = qcb.Lattice(ctx, (8, 8, 8, 8), qcb.F32)
    with pytest.raises(qcb.QcbError):
        qcb.Ecmc(f, qcb.ecmc_params())


def test_set_chains_round_trip_and_validation(ctx):
    """Saved chain states restore exactly (io::load_state, src/io/io.cpp:200-232):
    a chain restored into a fresh generator continues as the original does;
    states that would move a link of another block or a frozen link are refused."""
    dims = (8, 8, 8, 8)
    kw = dict(theta_sample=12.3, theta_refresh_site=5.1, theta_refresh_R=3.7, seed=5, max_events=200000)
    g = qcb.Lattice(ctx, dims)
    g.hot(8)
    e = qcb.Ecmc(g, qcb.ecmc_params(**kw))
    e.sample(1)
    st = e.chains()
    U = g.get_links()
    e.sample(1)
    want_links, want_chains = g.get_links(), e.chains()
    f = qcb.Lattice(ctx, dims)
    f.set_links(U)
    e2 = qcb.Ecmc(f, qcb.ecmc_params(**kw))
    e2.set_chains(st)
    assert np.array_equal(e2.chains(), st)
    e2.sample(1)
    assert np.array_equal(f.get_links(), want_links)
    assert np.array_equal(e2.chains(), want_chains)
    bad = st.copy()
    bad[0, F["site"]] = st[1, F["site"]]          # another block's site
    with pytest.raises(qcb.QcbError):
        e2.set_chains(bad)
    bad = st.copy()
    mu = int(bad[0, F["mu"]])
    x = [int(bad[0, F["site"]]) // 8 ** d % 8 for d in range(4)]
    x[(mu + 1) % 4] = 0                            # on a frozen face for this mu
    bad[0, F["site"]] = sum(v * 8 ** d for d, v in enumerate(x))
    with pytest.raises(qcb.QcbError):
        e2.set_chains(bad)
    for field, value in (("eps", 0), ("mu", 1.5), ("mu", 4), ("site", st[0, F["site"]] + 0.5), ("events", -1),
                         ("lifts", 2.0 ** 60), ("rng_n", float("nan"))):
        bad = st.copy()
        bad[0, F[field]] = value
        with pytest.raises(qcb.QcbError):
            e2.set_chains(bad)
    assert np.array_equal(e2.chains(), want_chains)   # refused states change nothing
    e.close()
    e2.close()


def _binned(x, nbin):
    x = np.asarray(x)
    bins = x[: len(x) // nbin * nbin].reshape(nbin, -1).mean(axis=1)
    return bins.mean(), bins.std(ddof=1) / np.sqrt(nbin)


def test_equilibrium_plaquette_matches_heatbath(ctx):
    """Event chains sample the Wilson-action equilibrium.  At 16^4, beta = 6.0,
    with a random shift before every sample (src/bin/gauge_ecmc_ncb.cpp:142-152),
    the mean plaquette of 256 device chains (4^4 blocks) agrees with the
    device heatbath, itself pinned to the reference-method chain
    (test_gpu_physics.py), within 3 combined standard errors.  ECMC samples are
    strongly correlated (each moves a fraction of the links), so the error
    comes from 10 long bins."""
    dims = (16, 16, 16, 16)
    h = qcb.Lattice(ctx, dims)
    h.hot(29)
    h.hb_sweeps(6.0, 2, 200, seed=29)
    ph = []
    for s in range(3000):
        h.hb_sweeps(6.0, 2, 1, seed=29, sweep_idx=200 + s)
        ph.append(h.plaquette())
    m_hb, e_hb = _binned(ph, 20)
    g = qcb.Lattice(ctx, dims)
    g.set_links(h.get_links())
    e = qcb.Ecmc(g, qcb.ecmc_params(beta=6.0, theta_sample=1000.0, theta_refresh_site=430.0,
                                    theta_refresh_R=51.0, seed=17, block=(4, 4, 4, 4), max_events=10 ** 7))
    rng = np.random.default_rng(19)
    p = []
    for s in range(600):
        g.translate(rng.integers(0, dims[0], size=4))
        e.sample(1)
        p.append(g.plaquette())
    e.chains()   # raises if any chain stopped early
    m, err = _binned(p[100:], 10)
    sigma = np.hypot(err, e_hb)
    print(f"ECMC <P> = {m:.6f} +- {err:.6f}, heatbath {m_hb:.6f} +- {e_hb:.6f}, {abs(m - m_hb) / sigma:.2f} sigma")
    assert abs(m - m_hb) < 3 * sigma


def _stopped(err):
    return int(str(err).split("qcb error -4: ")[1].split()[0])


def test_early_stop_status_is_sticky(ctx):
    """A chain that stops early (max_events) in any sample launch is reported
    by the next qcb_ecmc_get_chains, not only when it stops in the last one;
    a report clears it.  Two handles on identical lattices follow the same
    trajectories: reading after each of two samples gives the stop sets S1 and
    S2, reading once after both must give |S1 u S2|."""
    dims, kw = (16, 16, 8, 8), dict(beta=6.0, theta_sample=12.3, theta_refresh_site=5.1, theta_refresh_R=3.7,
                                    poisson=True, block=(4, 4, 4, 4))
    g0 = qcb.Lattice(ctx, dims)
    g0.hot(3)
    e0 = qcb.Ecmc(g0, qcb.ecmc_params(seed=77, **kw))
    e0.sample(1)
    M = int(np.median(e0.chains()[:, F["events"]]))
    e0.close()
    g0.close()
    gA, gB = qcb.Lattice(ctx, dims), qcb.Lattice(ctx, dims)
    gA.hot(3)
    gB.hot(3)
    eA = qcb.Ecmc(gA, qcb.ecmc_params(seed=77, max_events=M, **kw))
    eB = qcb.Ecmc(gB, qcb.ecmc_params(seed=77, max_events=M, **kw))
    eA.sample(2)
    with pytest.raises(qcb.QcbError) as e:
        eA.chains()
    nA = _stopped(e.value)
    eA.chains()                       # reported once, then cleared
    b = []
    for _ in range(2):
        eB.sample(1)
        try:
            eB.chains()
            b.append(0)
        except qcb.QcbError as err:
            assert err.code == qcb.EBUDGET
            b.append(_stopped(err))
    assert np.array_equal(gA.get_links(), gB.get_links())
    assert max(b) <= nA <= b[0] + b[1]
    assert nA > b[1], (nA, b)         # stops of the first sample are not forgotten
    for x in (eA, eB, gA, gB):
        x.close()


def test_lattice_outlives_its_chains(ctx):
    """qcb_lattice_destroy refuses while event chains use the lattice; the
    Python layer closes the chains first."""
    g = qcb.Lattice(ctx, (8, 8, 8, 8))
    e = qcb.Ecmc(g, qcb.ecmc_params(seed=1))
    assert qcb.lib().qcb_lattice_destroy(g.h) == qcb.ESTATE
    e.sample(1)
    g.close()                         # closes e, then the lattice
    assert not e.h and not g.h
    # a context closes chains before lattices whatever the order of its children
    c2 = qcb.Context(0)
    g2 = qcb.Lattice(c2, (8, 8, 8, 8))
    e2 = qcb.Ecmc(g2, qcb.ecmc_params(seed=1))
    c2.close()
    assert not e2.h and not g2.h
