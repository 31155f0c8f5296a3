"""Python bindings of libqcb.so (the MI355X heatbath / Wilson-flow library).

Thin ctypes layer over include/qcb.h used by the tests, the benchmark and the
flow service.  Links cross the boundary as float64 arrays shaped
(V, 4, 3, 3, 2) in the reference's interior AoS order (x fastest).  The
library must be built (`make -C ecmc-qncb_amd`); there is no CPU fallback:
loading fails loudly if libqcb.so is missing, and creating a context fails if
no GPU is visible.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# QCB_LIBRARY selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("QCB_LIBRARY") or os.path.join(PKG_DIR, "libqcb.so")

F64, F32 = 0, 1
OK, EINVAL, EDEVICE, ENOMEM, EBUDGET, ESTATE = 0, -1, -2, -3, -4, -5
PROBE_EXP, PROBE_PROJ, PROBE_REUNIT, PROBE_KP, PROBE_CM_HIT, PROBE_UNIFORM = 0, 1, 2, 3, 4, 5

_D = C.POINTER(C.c_double)
_VP = C.c_void_p


class QcbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qcb error {code}: {msg}")
        self.code = code


_lib = None


class EcmcParams(C.Structure):
    """qcb_ecmc_params (include/qcb.h); defaults of src/io/params.h:9-17."""
    _fields_ = [("beta", C.c_double), ("theta_sample", C.c_double), ("theta_refresh_site", C.c_double),
                ("theta_refresh_R", C.c_double), ("poisson", C.c_int), ("epsilon_set", C.c_double),
                ("block", C.c_int * 4), ("seed", C.c_uint64), ("max_events", C.c_uint64)]


def ecmc_params(beta=6.0, theta_sample=100.0, theta_refresh_site=50.0, theta_refresh_R=15.0, poisson=False,
                epsilon_set=0.15, block=(4, 4, 4, 4), seed=1, max_events=0) -> EcmcParams:
    return EcmcParams(beta, theta_sample, theta_refresh_site, theta_refresh_R, int(poisson), epsilon_set,
                      (C.c_int * 4)(*block), seed, max_events)


ECMC_CHAIN_DOUBLES = 32
ECMC_FIELDS = [f"R{k}" for k in range(18)] + [
    "th_refresh_site", "th_refresh_R", "par_refresh_site", "par_refresh_R", "events", "lifts", "revs",
    "set_counter", "set_gen", "rng_n", "site", "mu", "eps", "initialized"]

# (name, restype, argtypes) of every entry point declared in include/qcb.h
SIGNATURES = [
    ("qcb_ctx_create", C.c_int, [C.c_int, C.POINTER(_VP)]),
    ("qcb_ctx_destroy", C.c_int, [_VP]),
    ("qcb_ctx_set_kp_budget", C.c_int, [_VP, C.c_uint32]),
    ("qcb_last_error", C.c_char_p, [_VP]),
    ("qcb_ctx_synchronize", C.c_int, [_VP]),
    ("qcb_timing_begin", C.c_int, [_VP]),
    ("qcb_timing_end", C.c_int, [_VP, _D]),
    ("qcb_lattice_create", C.c_int, [_VP, C.POINTER(C.c_int), C.c_int, C.POINTER(_VP)]),
    ("qcb_lattice_destroy", C.c_int, [_VP]),
    ("qcb_lattice_bytes", C.c_int, [_VP, C.POINTER(C.c_uint64)]),
    ("qcb_lattice_set_tiling", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("qcb_lattice_set_links", C.c_int, [_VP, _D]),
    ("qcb_lattice_get_links", C.c_int, [_VP, _D]),
    ("qcb_cold", C.c_int, [_VP]),
    ("qcb_hot", C.c_int, [_VP, C.c_uint64]),
    ("qcb_plaquette", C.c_int, [_VP, _D]),
    ("qcb_clover_qe", C.c_int, [_VP, _D, _D]),
    ("qcb_staples", C.c_int, [_VP, _D]),
    ("qcb_hb_sweeps", C.c_int, [_VP, C.c_double, C.c_int, C.c_int, C.c_uint64, C.c_uint32]),
    ("qcb_hb_sweep_injected", C.c_int, [_VP, C.c_double, C.c_int, _D, C.c_int, C.POINTER(C.c_int64)]),
    ("qcb_reunitarize", C.c_int, [_VP]),
    ("qcb_translate", C.c_int, [_VP, C.POINTER(C.c_int)]),
    ("qcb_flow_steps", C.c_int, [_VP, C.c_double, C.c_int]),
    ("qcb_flow_copy", C.c_int, [_VP]),
    ("qcb_flow_rk3", C.c_int, [_VP, C.c_double, C.c_int]),
    ("qcb_flow_qe", C.c_int, [_VP, _D, _D]),
    ("qcb_flow_plaquette", C.c_int, [_VP, _D]),
    ("qcb_flow_get_links", C.c_int, [_VP, _D]),
    ("qcb_topo_charge_flowed", C.c_int, [_VP, C.c_double, C.c_int, C.c_int, _D, _D]),
    ("qcb_topo_charge_flowed_noalloc", C.c_int, [_VP, C.c_double, C.c_int, C.c_int, _D, _D]),
    ("qcb_batch_create", C.c_int, [_VP, C.c_int, C.POINTER(C.c_int), C.c_int, _D, C.POINTER(C.c_uint64),
                                    C.POINTER(_VP)]),
    ("qcb_batch_destroy", C.c_int, [_VP]),
    ("qcb_batch_lattice", C.c_int, [_VP, C.c_int, C.POINTER(_VP)]),
    ("qcb_batch_hot", C.c_int, [_VP]),
    ("qcb_batch_hb_sweeps", C.c_int, [_VP, C.c_int, C.c_int, C.c_uint32]),
    ("qcb_batch_reunitarize", C.c_int, [_VP]),
    ("qcb_batch_plaquette", C.c_int, [_VP, _D]),
    ("qcb_batch_flow", C.c_int, [_VP, C.c_double, C.c_int, _D]),
    ("qcb_batch_flow_steps", C.c_int, [_VP, C.c_double, C.c_int]),
    ("qcb_batch_set_schedule", C.c_int, [_VP, C.POINTER(C.c_int), C.c_uint32]),
    ("qcb_batch_get_schedule", C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_uint32)]),
    ("qcb_batch_set_layout", C.c_int, [_VP, C.c_int, C.POINTER(C.c_int)]),
    ("qcb_batch_get_layout", C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("qcb_batch_autotune", C.c_int, [_VP, C.c_int, C.c_int, C.c_char_p, C.c_size_t]),
    ("qcb_batch_graph_capture", C.c_int, [_VP, C.c_int, C.c_int, C.c_uint32]),
    ("qcb_batch_graph_launch", C.c_int, [_VP, C.c_int]),
    ("qcb_service_create", C.c_int, [_VP, C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.POINTER(_VP)]),
    ("qcb_service_destroy", C.c_int, [_VP]),
    ("qcb_service_clock", C.c_double, [_VP]),
    ("qcb_service_submit", C.c_int, [_VP, C.c_uint64, C.POINTER(C.c_int), C.c_double, C.c_uint64, C.c_double,
                                      C.c_double]),
    ("qcb_service_step", C.c_int, [_VP, C.POINTER(C.c_int)]),
    ("qcb_service_poll", C.c_int, [_VP, C.POINTER(C.c_uint64), _D, C.c_int, C.POINTER(C.c_int), _D]),
    ("qcb_service_status", C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("qcb_ildg_write", C.c_int, [C.c_char_p, C.POINTER(C.c_int), _D]),
    ("qcb_ildg_read", C.c_int, [C.c_char_p, C.POINTER(C.c_int), _D]),
    ("qcb_ildg_save", C.c_int, [_VP, C.c_char_p]),
    ("qcb_ildg_load", C.c_int, [_VP, C.c_char_p]),
    ("qcb_ecmc_create", C.c_int, [_VP, C.POINTER(EcmcParams), C.POINTER(_VP)]),
    ("qcb_ecmc_destroy", C.c_int, [_VP]),
    ("qcb_ecmc_n_chains", C.c_int, [_VP, C.POINTER(C.c_int)]),
    ("qcb_ecmc_sample", C.c_int, [_VP, C.c_int]),
    ("qcb_ecmc_get_chains", C.c_int, [_VP, _D]),
    ("qcb_ecmc_set_chains", C.c_int, [_VP, _D]),
    ("qcb_math_probe", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, _D, C.c_double, C.c_int, _D]),
]


def lib():
    """Load libqcb.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C ecmc-qncb_amd`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_D)


class Context:
    def __init__(self, device: int = 0):
        self.h = _VP()
        self._children = weakref.WeakSet()   # handles that must die before the context
        rc = lib().qcb_ctx_create(device, C.byref(self.h))
        if rc != OK:
            raise QcbError(rc, "no usable GPU (qcb_ctx_create)")

    def check(self, rc):
        if rc != OK:
            raise QcbError(rc, lib().qcb_last_error(self.h).decode())
        return rc

    def close(self):
        if self.h:
            # event chains read their lattice (the library refuses to destroy a
            # lattice that chains still use), so they go first
            children = list(self._children)
            for child in [c for c in children if isinstance(c, Ecmc)] + [c for c in children if not isinstance(c, Ecmc)]:
                child.close()
            self.check(lib().qcb_ctx_destroy(self.h))
            self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self.check(lib().qcb_ctx_synchronize(self.h))

    def set_kp_budget(self, max_attempts: int):
        """Kennedy-Pendleton attempts per draw for Philox heatbath sweeps (default 1024)."""
        self.check(lib().qcb_ctx_set_kp_budget(self.h, max_attempts))

    def timing_begin(self):
        self.check(lib().qcb_timing_begin(self.h))

    def timing_end(self) -> float:
        ms = C.c_double()
        self.check(lib().qcb_timing_end(self.h, C.byref(ms)))
        return ms.value

    def probe(self, prec: int, op: int, inp: np.ndarray, param: float = 0.0, aux: int = 0) -> np.ndarray:
        inp = np.ascontiguousarray(inp, dtype=np.float64)
        n = inp.shape[0]
        wout = 5 if op == PROBE_KP else 1 if op == PROBE_UNIFORM else 18
        out = np.zeros((n, wout))
        self.check(lib().qcb_math_probe(self.h, prec, op, n, _dp(inp.reshape(n, -1)), param, aux, _dp(out)))
        return out


class Lattice:
    """A device-resident gauge field (GaugeField of the reference)."""

    def __init__(self, ctx: Context, dims, prec: int = F64):
        self.ctx = ctx
        self.dims = tuple(int(d) for d in dims)
        self.prec = prec
        self.V = int(np.prod(self.dims))
        self.h = _VP()
        self._users = weakref.WeakSet()   # event chains on this lattice
        ctx.check(lib().qcb_lattice_create(ctx.h, (C.c_int * 4)(*self.dims), prec, C.byref(self.h)))
        ctx._children.add(self)

    def close(self):
        for u in list(self._users):
            u.close()
        if self.h and self.ctx.h:
            self.ctx.check(lib().qcb_lattice_destroy(self.h))
        self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        return self.ctx.check(rc)

    def set_links(self, links: np.ndarray):
        a = np.ascontiguousarray(links)
        if a.dtype == np.complex128:
            a = a.view(np.float64)
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        if a.size != self.V * 72:
            raise ValueError("links must hold V*4*3*3 complex numbers")
        self._c(lib().qcb_lattice_set_links(self.h, _dp(a)))

    def get_links(self) -> np.ndarray:
        out = np.zeros(self.V * 72)
        self._c(lib().qcb_lattice_get_links(self.h, _dp(out)))
        return out.view(np.complex128).reshape(self.V, 4, 3, 3)

    def set_tiling(self, ty: int, tz: int, tt: int, morton: bool = False):
        self._c(lib().qcb_lattice_set_tiling(self.h, ty, tz, tt, int(morton)))

    def nbytes(self) -> int:
        b = C.c_uint64()
        self._c(lib().qcb_lattice_bytes(self.h, C.byref(b)))
        return b.value

    def cold(self):
        self._c(lib().qcb_cold(self.h))

    def hot(self, seed: int):
        self._c(lib().qcb_hot(self.h, seed))

    def plaquette(self) -> float:
        p = C.c_double()
        self._c(lib().qcb_plaquette(self.h, C.byref(p)))
        return p.value

    def clover_qe(self):
        q, e = C.c_double(), C.c_double()
        self._c(lib().qcb_clover_qe(self.h, C.byref(q), C.byref(e)))
        return q.value, e.value

    def staples(self) -> np.ndarray:
        out = np.zeros(self.V * 72)
        self._c(lib().qcb_staples(self.h, _dp(out)))
        return out.view(np.complex128).reshape(self.V, 4, 3, 3)

    def hb_sweeps(self, beta: float, n_hits: int, n_sweeps: int, seed: int, sweep_idx: int = 0):
        self._c(lib().qcb_hb_sweeps(self.h, beta, n_hits, n_sweeps, seed, sweep_idx))

    def hb_sweep_injected(self, beta: float, n_hits: int, u: np.ndarray, K: int) -> int:
        n = C.c_int64()
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        if u.size != 8 * n_hits * 3 * (4 * K + 2) * (self.V // 2):
            raise ValueError("uniform buffer has the wrong size")
        self._c(lib().qcb_hb_sweep_injected(self.h, beta, n_hits, _dp(u), K, C.byref(n)))
        return n.value

    def reunitarize(self):
        self._c(lib().qcb_reunitarize(self.h))

    def translate(self, s):
        self._c(lib().qcb_translate(self.h, (C.c_int * 4)(*[int(v) for v in s])))

    def flow_steps(self, eps: float, n: int):
        self._c(lib().qcb_flow_steps(self.h, eps, n))

    # GradientFlow mirror: copy -> rk3 steps -> measure, links untouched
    def flow_copy(self):
        self._c(lib().qcb_flow_copy(self.h))

    def flow_rk3(self, eps: float, n: int = 1):
        self._c(lib().qcb_flow_rk3(self.h, eps, n))

    def flow_qe(self):
        q, e = C.c_double(), C.c_double()
        self._c(lib().qcb_flow_qe(self.h, C.byref(q), C.byref(e)))
        return q.value, e.value

    def flow_plaquette(self) -> float:
        p = C.c_double()
        self._c(lib().qcb_flow_plaquette(self.h, C.byref(p)))
        return p.value

    def flow_links(self) -> np.ndarray:
        out = np.zeros(self.V * 72)
        self._c(lib().qcb_flow_get_links(self.h, _dp(out)))
        return out.view(np.complex128).reshape(self.V, 4, 3, 3)

    def ildg_save(self, path: str):
        self._c(lib().qcb_ildg_save(self.h, str(path).encode()))

    def ildg_load(self, path: str):
        self._c(lib().qcb_ildg_load(self.h, str(path).encode()))

    def topo_charge_flowed_noalloc(self, eps: float, n_steps_gf: int, n_rk: int, final: bool = False):
        """qcb_topo_charge_flowed_noalloc: the same series, nothing allocated on the device."""
        tqe = np.zeros(3 * max(n_steps_gf, 1))
        fin = np.zeros(3)
        self._c(lib().qcb_topo_charge_flowed_noalloc(self.h, eps, n_steps_gf, n_rk, _dp(tqe),
                                                     _dp(fin) if final else None))
        t = tqe[: 3 * n_steps_gf].reshape(n_steps_gf, 3)
        return (t, fin) if final else t

    def topo_charge_flowed(self, eps: float, n_steps_gf: int, n_rk: int, final: bool = False):
        tqe = np.zeros(3 * max(n_steps_gf, 1))
        fin = np.zeros(3)
        self._c(lib().qcb_topo_charge_flowed(self.h, eps, n_steps_gf, n_rk, _dp(tqe),
                                             _dp(fin) if final else None))
        t = tqe[: 3 * n_steps_gf].reshape(n_steps_gf, 3)
        return (t, fin) if final else t


class _Borrowed(Lattice):
    """A lattice owned by a Batch (never destroyed through this handle)."""

    def __init__(self, batch, i, handle, dims, prec):
        self.ctx = batch.ctx
        self.dims = tuple(dims)
        self.prec = prec
        self.V = int(np.prod(self.dims))
        self.h = handle

    def close(self):
        self.h = _VP()


class Batch:
    """Independent lattices advanced together by single launches."""

    def __init__(self, ctx: Context, dims_list, prec: int, betas, seeds):
        self.ctx = ctx
        self.prec = prec
        self.dims = [tuple(int(v) for v in d) for d in dims_list]
        n = len(self.dims)
        self.n = n
        flat = (C.c_int * (4 * n))(*[v for d in self.dims for v in d])
        b = np.ascontiguousarray(betas, dtype=np.float64)
        sd = (C.c_uint64 * n)(*[int(x) for x in seeds])
        self.h = _VP()
        ctx.check(lib().qcb_batch_create(ctx.h, n, flat, prec, _dp(b), sd, C.byref(self.h)))
        ctx._children.add(self)

    def close(self):
        if self.h and self.ctx.h:
            lib().qcb_batch_destroy(self.h)
        self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        return self.ctx.check(rc)

    def lattice(self, i) -> Lattice:
        h = _VP()
        self._c(lib().qcb_batch_lattice(self.h, i, C.byref(h)))
        return _Borrowed(self, i, h, self.dims[i], self.prec)

    def hot(self):
        self._c(lib().qcb_batch_hot(self.h))

    def hb_sweeps(self, n_hits, n_sweeps, sweep_idx=0):
        self._c(lib().qcb_batch_hb_sweeps(self.h, n_hits, n_sweeps, sweep_idx))

    def reunitarize(self):
        self._c(lib().qcb_batch_reunitarize(self.h))

    def plaquette(self) -> np.ndarray:
        out = np.zeros(self.n)
        self._c(lib().qcb_batch_plaquette(self.h, _dp(out)))
        return out

    def flow(self, eps, n_steps) -> np.ndarray:
        out = np.zeros(self.n * (n_steps + 1) * 3)
        self._c(lib().qcb_batch_flow(self.h, eps, n_steps, _dp(out)))
        return out.reshape(self.n, n_steps + 1, 3)

    def flow_steps(self, eps, n_steps):
        """qcb_flow_steps on every lattice (no measurement)"""
        self._c(lib().qcb_batch_flow_steps(self.h, eps, n_steps))

    def set_schedule(self, order, grid=0):
        self._c(lib().qcb_batch_set_schedule(self.h, (C.c_int * 8)(*order), grid))

    def get_schedule(self):
        o = (C.c_int * 8)()
        g = C.c_uint32()
        self._c(lib().qcb_batch_get_schedule(self.h, o, C.byref(g)))
        return list(o), g.value

    def set_layout(self, n_streams=1, tile_caps=(8, 4, 4)):
        self._c(lib().qcb_batch_set_layout(self.h, n_streams, (C.c_int * 3)(*tile_caps)))

    def get_layout(self):
        n = C.c_int()
        t = (C.c_int * 3)()
        self._c(lib().qcb_batch_get_layout(self.h, C.byref(n), t))
        return n.value, tuple(t)

    def autotune(self, n_hits=2, reps=2) -> str:
        buf = C.create_string_buffer(4096)
        self._c(lib().qcb_batch_autotune(self.h, n_hits, reps, buf, len(buf)))
        return buf.value.decode()

    def graph_capture(self, n_hits, sweeps_per_graph, first_sweep=0):
        self._c(lib().qcb_batch_graph_capture(self.h, n_hits, sweeps_per_graph, first_sweep))

    def graph_launch(self, n_replays=1):
        self._c(lib().qcb_batch_graph_launch(self.h, n_replays))


class Ecmc:
    """Event chains, one per block of an FP64 lattice (qcb_ecmc_*, include/qcb.h)."""

    def __init__(self, lat: Lattice, params: EcmcParams):
        self.lat = lat
        self.ctx = lat.ctx
        self.h = _VP()
        self.ctx.check(lib().qcb_ecmc_create(lat.h, C.byref(params), C.byref(self.h)))
        if hasattr(lat, "_users"):
            lat._users.add(self)
        n = C.c_int()
        self.ctx.check(lib().qcb_ecmc_n_chains(self.h, C.byref(n)))
        self.n_chains = n.value
        self.ctx._children.add(self)

    def close(self):
        if self.h and self.ctx.h:
            lib().qcb_ecmc_destroy(self.h)
        self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sample(self, n=1):
        self.ctx.check(lib().qcb_ecmc_sample(self.h, n))

    def set_chains(self, states: np.ndarray):
        st = np.ascontiguousarray(states, dtype=np.float64)
        assert st.shape == (self.n_chains, ECMC_CHAIN_DOUBLES)
        self.ctx.check(lib().qcb_ecmc_set_chains(self.h, _dp(st)))

    def chains(self) -> np.ndarray:
        """(n_chains, 32) chain states; columns named by ECMC_FIELDS."""
        out = np.zeros((self.n_chains, ECMC_CHAIN_DOUBLES))
        self.ctx.check(lib().qcb_ecmc_get_chains(self.h, _dp(out)))
        return out


class Service:
    """Continuous-batching flow-measurement service (see include/qcb.h)."""

    def __init__(self, ctx: Context, prec=F32, arena_bytes=16 << 30, n_hits=2, n_therm=50, reunit_every=4,
                 eps=0.02):
        self.ctx = ctx
        self.eps = eps
        self.h = _VP()
        ctx.check(lib().qcb_service_create(ctx.h, prec, int(arena_bytes), n_hits, n_therm, reunit_every, eps,
                                           C.byref(self.h)))
        ctx._children.add(self)

    def close(self):
        if self.h and self.ctx.h:
            lib().qcb_service_destroy(self.h)
        self.h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self) -> float:
        return lib().qcb_service_clock(self.h)

    def submit(self, rid, dims, beta, seed, t_max, t_arrival=None):
        if t_arrival is None:
            t_arrival = self.clock()
        self.ctx.check(lib().qcb_service_submit(self.h, rid, (C.c_int * 4)(*dims), beta, seed, t_max, t_arrival))

    def step(self) -> int:
        p = C.c_int()
        self.ctx.check(lib().qcb_service_step(self.h, C.byref(p)))
        return p.value

    def poll(self, cap=4096):
        """Finished requests as a list of (id, tQE[n,3], (arrival, admit, done)).
        A request whose thermalisation skipped subgroup updates (Kennedy-Pendleton
        budget) raises QcbError(EBUDGET) carrying it as .result and the requests
        popped before it as .done."""
        out = []
        while True:
            rid, n = C.c_uint64(), C.c_int()
            buf = np.zeros(cap * 3)
            times = np.zeros(3)
            rc = lib().qcb_service_poll(self.h, C.byref(rid), _dp(buf), cap, C.byref(n), _dp(times))
            if rc == ESTATE:
                return out
            if rc == EBUDGET:
                err = QcbError(rc, lib().qcb_last_error(self.ctx.h).decode())
                err.result = (rid.value, buf[: 3 * n.value].reshape(n.value, 3), tuple(times))
                err.done = out
                raise err
            self.ctx.check(rc)
            out.append((rid.value, buf[: 3 * n.value].reshape(n.value, 3), tuple(times)))

    def status(self):
        q, r, f = C.c_int(), C.c_int(), C.c_int()
        fb, it = C.c_uint64(), C.c_uint64()
        self.ctx.check(lib().qcb_service_status(self.h, C.byref(q), C.byref(r), C.byref(f), C.byref(fb), C.byref(it)))
        return dict(queued=q.value, resident=r.value, finished=f.value, free_bytes=fb.value, iterations=it.value)
