"""ctypes view of the FP64 CPU oracle (oracle/libqcb_oracle.so).

Test infrastructure only: tests, the smoke check and the CPU-baseline leg of
the benchmark load it as the checker.  Links are numpy float64 arrays shaped
(V, 4, 3, 3, 2) in the reference's interior AoS order (x fastest).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "libqcb_oracle.so")

_D = C.POINTER(C.c_double)


class QoLat(C.Structure):
    _fields_ = [("L", C.c_int * 4), ("V", C.c_size_t), ("U", _D)]


class QoMt64(C.Structure):
    _fields_ = [("mt", C.c_uint64 * 312), ("idx", C.c_int)]


class RunParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("L", "cold_start", "seed", "N_therm", "N_shift")] + [
        ("beta", C.c_double)] + [(n, C.c_int) for n in (
            "N_sweeps", "N_hits", "N_shift_plaquette", "topo", "N_shift_topo", "N_steps_gf",
            "N_rk_steps")]


class EcmcParams(C.Structure):
    _fields_ = [("beta", C.c_double), ("theta_sample", C.c_double), ("theta_refresh_site", C.c_double),
                ("theta_refresh_R", C.c_double), ("poisson", C.c_int), ("epsilon_set", C.c_double),
                ("block", C.c_int * 4)]


class EcmcChain(C.Structure):
    _fields_ = [("R", C.c_double * 18), ("th_refresh_site", C.c_double), ("th_refresh_R", C.c_double),
                ("par_refresh_site", C.c_double), ("par_refresh_R", C.c_double)] + [
        (n, C.c_uint64) for n in ("events", "lifts", "revs", "set_counter", "set_gen", "rng_n")] + [
        ("site", C.c_size_t), ("mu", C.c_int), ("eps", C.c_int), ("initialized", C.c_int)]


def ecmc_params(beta=6.0, theta_sample=100.0, theta_refresh_site=50.0, theta_refresh_R=15.0, poisson=False,
                epsilon_set=0.15, block=(4, 4, 4, 4)):
    return EcmcParams(beta, theta_sample, theta_refresh_site, theta_refresh_R, int(poisson), epsilon_set,
                      (C.c_int * 4)(*block))


class RunStats(C.Structure):
    _fields_ = [("t_hb", C.c_double), ("t_plaq", C.c_double), ("t_flow", C.c_double),
                ("t_shift", C.c_double), ("link_visits", C.c_long), ("site_steps", C.c_long),
                ("n_plaq", C.c_int), ("n_tqe", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        sig = {
            "qo_m3_mul": (None, [_D, _D, _D]),
            "qo_set_threads": (None, [C.c_int]),
            "qo_proj_lie": (None, [_D]),
            "qo_exp_luscher": (None, [_D, C.c_double, _D]),
            "qo_su2q_to_su3": (None, [_D, C.c_int, C.c_int, _D]),
            "qo_su2_to_q": (None, [_D, _D]),
            "qo_qmult": (None, [_D, _D, _D]),
            "qo_reunitarize_link": (None, [_D]),
            "qo_kp_injected": (C.c_int, [C.c_double, C.c_double, _D, C.c_int, _D]),
            "qo_mt64_seed": (None, [P(QoMt64), C.c_uint64]),
            "qo_mt64_next": (C.c_uint64, [P(QoMt64)]),
            "qo_mt64_canonical": (C.c_double, [P(QoMt64)]),
            "qo_mt64_uniform_int": (C.c_uint64, [P(QoMt64), C.c_uint64, C.c_uint64]),
            "qo_mt64_normal_pair": (None, [P(QoMt64), _D, _D]),
            "qo_philox4x32_10": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
            "qo_u53": (C.c_double, [C.c_uint32, C.c_uint32]),
            "qo_lat_init": (C.c_int, [P(QoLat), P(C.c_int)]),
            "qo_lat_free": (None, [P(QoLat)]),
            "qo_lat_cold": (None, [P(QoLat)]),
            "qo_random_su3_ref": (None, [P(QoMt64), P(C.c_double)]),
            "qo_random_su3_ref_order": (None, [P(QoMt64), P(C.c_double), C.c_int]),
            "qo_lat_hot_ref": (None, [P(QoLat), P(QoMt64)]),
            "qo_lat_hot_philox": (None, [P(QoLat), C.c_uint64]),
            "qo_hot_link_philox": (None, [C.c_uint64, C.c_uint32, C.c_uint32, _D]),
            "qo_staple": (None, [P(QoLat), C.c_size_t, C.c_int, _D]),
            "qo_plaquette": (C.c_double, [P(QoLat)]),
            "qo_clover_F": (None, [P(QoLat), C.c_size_t, C.c_int, C.c_int, _D]),
            "qo_clover_qe": (None, [P(QoLat), _D, _D]),
            "qo_clover_qe_site": (None, [P(QoLat), _D, _D]),
            "qo_reunitarize": (None, [P(QoLat)]),
            "qo_translate": (None, [P(QoLat), P(C.c_int)]),
            "qo_gauge_transform": (None, [P(QoLat), _D]),
            "qo_flow_rk3_step": (None, [P(QoLat), C.c_double]),
            "qo_topo_charge_flowed": (None, [P(QoLat), C.c_double, C.c_int, C.c_int, _D]),
            "qo_hb_sweep_ref": (C.c_long, [P(QoLat), C.c_double, C.c_int, P(QoMt64)]),
            "qo_hb_sweep_cb_injected": (C.c_long, [P(QoLat), C.c_double, C.c_int, _D, C.c_int]),
            "qo_cm_hit_injected": (C.c_int, [_D, _D, C.c_double, _D, C.c_int]),
            "qo_hb_sweep_cb_philox": (None, [P(QoLat), C.c_double, C.c_int, C.c_uint64, C.c_uint32]),
            "qo_ref_run": (C.c_int, [P(RunParams), _D, _D, P(C.c_int), _D, P(RunStats)]),
            "qo_now": (C.c_double, []),
            "qo_ecmc_reject_fast": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_int]),
            "qo_ecmc_reject_exact": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_int]),
            "qo_ecmc_compute_ds": (C.c_double, [_D, _D]),
            "qo_ecmc_update": (None, [_D, C.c_double, C.c_int, _D]),
            "qo_ecmc_random_su3_eps": (None, [C.c_double, _D, _D]),
            "qo_ecmc_n_chains": (C.c_int, [P(QoLat), P(EcmcParams)]),
            "qo_ecmc_sample_philox": (C.c_int, [P(QoLat), P(EcmcParams), P(EcmcChain), C.c_uint64]),
            "qo_ecmc_sample_ref": (C.c_int, [P(QoLat), P(EcmcParams), P(EcmcChain), P(QoMt64)]),
            "qo_haar_su3_philox": (None, [P(C.c_uint32), C.c_uint32, C.c_uint32, C.c_uint32, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def m3(a) -> np.ndarray:
    """complex (3,3) -> 18 interleaved doubles."""
    a = np.asarray(a, dtype=np.complex128)
    return np.ascontiguousarray(a.view(np.float64).reshape(18))


def c3(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.complex128).reshape(3, 3)


def exp_luscher(Z: np.ndarray, coeff: float) -> np.ndarray:
    out = np.zeros(18)
    lib().qo_exp_luscher(dp(m3(Z)), coeff, dp(out))
    return c3(out)


def proj_lie(A: np.ndarray) -> np.ndarray:
    a = m3(A)
    lib().qo_proj_lie(dp(a))
    return c3(a)


class Lattice:
    """Owns an oracle lattice; .links is a (V,4,3,3) complex view."""

    def __init__(self, dims):
        self.s = QoLat()
        d = (C.c_int * 4)(*dims)
        if lib().qo_lat_init(C.byref(self.s), d) != 0:
            raise ValueError(f"bad dims {dims}")
        self.dims = tuple(dims)
        self.V = int(self.s.V)

    def __del__(self):
        try:
            lib().qo_lat_free(C.byref(self.s))
        except Exception:
            pass

    @property
    def raw(self) -> np.ndarray:
        # the array's base keeps this Lattice alive, so a view taken from a
        # temporary (e.g. f().links.copy()) never outlives the C buffer
        buf = (C.c_double * (self.V * 72)).from_address(C.addressof(self.s.U.contents))
        buf._owner = self
        return np.ctypeslib.as_array(buf)

    @property
    def links(self) -> np.ndarray:
        return self.raw.view(np.complex128).reshape(self.V, 4, 3, 3)

    def set_links(self, links: np.ndarray):
        self.raw[:] = np.ascontiguousarray(links, dtype=np.complex128).view(np.float64).reshape(-1)

    def ref(self):
        return C.byref(self.s)

    def cold(self):
        lib().qo_lat_cold(self.ref())

    def hot_philox(self, seed: int):
        lib().qo_lat_hot_philox(self.ref(), seed)

    def plaquette(self) -> float:
        return lib().qo_plaquette(self.ref())

    def clover_qe(self):
        q, e = C.c_double(), C.c_double()
        lib().qo_clover_qe(self.ref(), C.cast(C.byref(q), _D), C.cast(C.byref(e), _D))
        return q.value, e.value

    def staple(self, site: int, mu: int) -> np.ndarray:
        out = np.zeros(18)
        lib().qo_staple(self.ref(), site, mu, dp(out))
        return c3(out)

    def flow_step(self, eps: float):
        lib().qo_flow_rk3_step(self.ref(), eps)

    def topo_flowed(self, eps, n_steps_gf, n_rk):
        out = np.zeros(3 * n_steps_gf)
        lib().qo_topo_charge_flowed(self.ref(), eps, n_steps_gf, n_rk, dp(out))
        return out.reshape(n_steps_gf, 3)

    def translate(self, s):
        lib().qo_translate(self.ref(), (C.c_int * 4)(*s))

    def reunitarize(self):
        lib().qo_reunitarize(self.ref())

    def gauge_transform(self, G: np.ndarray):
        g = np.ascontiguousarray(G, dtype=np.complex128).view(np.float64).reshape(-1)
        lib().qo_gauge_transform(self.ref(), dp(np.ascontiguousarray(g)))

    def sweep_ref(self, beta, n_hits, mt: "Mt64") -> int:
        return lib().qo_hb_sweep_ref(self.ref(), beta, n_hits, C.byref(mt.s))

    def sweep_injected(self, beta, n_hits, u: np.ndarray, K: int) -> int:
        return lib().qo_hb_sweep_cb_injected(self.ref(), beta, n_hits, dp(u), K)

    def sweep_philox(self, beta, n_hits, seed, sweep_idx):
        lib().qo_hb_sweep_cb_philox(self.ref(), beta, n_hits, seed, sweep_idx)

    def ecmc_philox(self, params: EcmcParams, chains, seed: int) -> int:
        return lib().qo_ecmc_sample_philox(self.ref(), C.byref(params), chains, seed)

    def ecmc_ref(self, params: EcmcParams, chain: EcmcChain, mt: "Mt64") -> int:
        return lib().qo_ecmc_sample_ref(self.ref(), C.byref(params), C.byref(chain), C.byref(mt.s))

    def ecmc_chains(self, params: EcmcParams):
        n = lib().qo_ecmc_n_chains(self.ref(), C.byref(params))
        assert n > 0
        return (EcmcChain * n)()

    def copy(self) -> "Lattice":
        o = Lattice(self.dims)
        o.raw[:] = self.raw
        return o


class Mt64:
    def __init__(self, seed: int):
        self.s = QoMt64()
        lib().qo_mt64_seed(C.byref(self.s), seed & (2**64 - 1))

    def next(self) -> int:
        return lib().qo_mt64_next(C.byref(self.s))

    def random_su3(self, imag_first: bool = False) -> np.ndarray:
        """one SU(3) draw of the reference's hot start (src/su3/utils.cpp:13-30);
        imag_first selects a GCC build's draw order (default: clang/IntelLLVM)"""
        out = np.zeros(18)
        lib().qo_random_su3_ref_order(C.byref(self.s), out.ctypes.data_as(C.POINTER(C.c_double)),
                                      int(imag_first))
        return out.view(np.complex128).reshape(3, 3)

    def canonical(self) -> float:
        return lib().qo_mt64_canonical(C.byref(self.s))

    def uniform_int(self, a, b) -> int:
        return lib().qo_mt64_uniform_int(C.byref(self.s), a, b)

    def normal_pair(self):
        a, b = C.c_double(), C.c_double()
        lib().qo_mt64_normal_pair(C.byref(self.s), C.cast(C.byref(a), _D), C.cast(C.byref(b), _D))
        return a.value, b.value


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().qo_philox4x32_10(c, k, o)
    return list(o)


def n_uniform_slots(K: int) -> int:
    return 4 * K + 2


def ref_run(**kw):
    p = RunParams(**kw)
    n_plaq = kw["N_shift"] // kw["N_shift_plaquette"] + 1
    n_flows = (kw["N_shift"] // kw["N_shift_topo"] + 1) if kw["topo"] else 0
    plaq = np.zeros(max(n_plaq, 1))
    tqe = np.zeros(max(3 * kw["N_steps_gf"] * n_flows, 1))
    shifts = (C.c_int * (4 * (kw["N_therm"] + kw["N_shift"] + 1)))()
    V = kw["L"] ** 4
    final = np.zeros(V * 72)
    st = RunStats()
    rc = lib().qo_ref_run(C.byref(p), dp(plaq), dp(tqe), shifts, dp(final), C.byref(st))
    assert rc == 0
    return dict(plaq=plaq[: st.n_plaq], tqe=tqe[: 3 * st.n_tqe].reshape(-1, 3),
                shifts=np.array(list(shifts))[: 4 * (kw["N_therm"] + kw["N_shift"])].reshape(-1, 4),
                final=final.view(np.complex128).reshape(V, 4, 3, 3), stats=st)
