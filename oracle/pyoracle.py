"""ctypes wrapper of the FP64 host oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py``, always as the checker or the
timed CPU baseline, never by the product library.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "liboracle.so"
BIN = HERE / "build" / "astrosim_oracle"

_lib = None


class _Opts(C.Structure):
    _fields_ = [("dt", C.c_double), ("substeps", C.c_int), ("w_n2a", C.c_double),
                ("w_a2n", C.c_double), ("stp_enabled", C.c_int), ("threads", C.c_int)]


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        raise FileNotFoundError(f"{LIB} missing: run `make -C oracle` (or the package build)")
    lib = C.CDLL(str(LIB))
    sig = {
        "orc_rng_bits": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "orc_rng_uniform": (C.c_double, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "orc_rng_normal": (C.c_double, [C.c_uint64, C.c_uint64]),
        "orc_rng_poisson": (C.c_int, [C.c_uint64, C.c_uint64, C.c_double]),
        "orc_hostrng_normals": (None, [C.c_uint64, C.c_uint64, C.c_int64, C.c_void_p]),
        "orc_exp": (C.c_double, [C.c_double]),
        "orc_log": (C.c_double, [C.c_double]),
        "orc_cos": (C.c_double, [C.c_double]),
        "orc_net_build": (C.c_void_p, [C.c_uint32, C.c_uint64]),
        "orc_net_free": (None, [C.c_void_p]),
        "orc_net_from_arrays": (C.c_void_p, [C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]),
        "orc_net_sizes": (None, [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
        "orc_net_csr": (C.c_uint64, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
        "orc_net_param": (None, [C.c_void_p, C.c_int, C.c_void_p]),
        "orc_net_build_seconds": (C.c_double, [C.c_void_p]),
        "orc_sim_create": (C.c_void_p, [C.c_void_p, C.POINTER(_Opts)]),
        "orc_sim_free": (None, [C.c_void_p]),
        "orc_sim_step": (None, [C.c_void_p, C.c_int64, C.c_void_p]),
        "orc_sim_steps_done": (C.c_int64, [C.c_void_p]),
        "orc_sim_spike_count": (C.c_uint64, [C.c_void_p]),
        "orc_sim_reset_ever_active": (None, [C.c_void_p]),
        "orc_sim_read": (None, [C.c_void_p, C.c_int, C.c_void_p]),
        "orc_sim_read_refractory": (None, [C.c_void_p, C.c_void_p]),
        "orc_sim_read_ever_active": (None, [C.c_void_p, C.c_void_p]),
        "orc_astro_rhs": (None, [C.c_double] * 7 + [C.c_void_p]),
        "orc_astro_advance": (None, [C.c_double] * 7 + [C.c_int, C.c_void_p]),
        "orc_sic_factor": (C.c_double, [C.c_double]),
        "orc_neuron_rhs": (None, [C.c_int] + [C.c_double] * 5 + [C.c_void_p]),
        "orc_neuron_substeps": (C.c_int, [C.c_int, C.c_void_p, C.POINTER(C.c_int32), C.c_double, C.c_double,
                                          C.c_double, C.c_double, C.c_int]),
        "orc_psc_init": (C.c_double, [C.c_int, C.c_int]),
        "orc_stp_on_spike": (C.c_double, [C.c_void_p, C.c_double]),
        "orc_gather": (C.c_double, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]),
        "orc_sim_params": (None, [C.c_void_p, C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


FIELDS = {"Ca": 0, "IP3": 1, "h": 2, "V": 3, "w": 4, "g_ex": 5, "dg_ex": 6, "g_in": 7,
          "dg_in": 8, "I_sic_hold": 9, "stp_x": 10, "stp_u": 11, "stp_t_last": 12}


class OracleNet:
    CSR = {"exc": 0, "inh": 1, "sic": 2, "e2a": 3}
    PARAMS = {"Ca_tot": 0, "IP3_0": 1, "tau_IP3": 2, "delta_IP3": 3, "V_reset": 4, "b": 5}

    def __init__(self, n_total: int, seed: int):
        lib = load()
        self._h = lib.orc_net_build(n_total, seed)
        if not self._h:
            raise ValueError("invalid n_total")
        c = np.zeros(4, dtype=np.uint32)
        p = C.c_double()
        lib.orc_net_sizes(self._h, _p(c), C.byref(p))
        self.n_exc, self.n_inh, self.n_astro, self.n_neurons = map(int, c)
        self.p_primary = p.value
        self.n_total, self.seed = n_total, seed

    @classmethod
    def from_arrays(cls, n_total: int, seed: int, csrs: dict, params: dict) -> "OracleNet":
        """Wraps CSR arrays and parameters built elsewhere (for timing the
        stepper on networks too large for the oracle's own generator)."""
        self = cls.__new__(cls)
        keep = [np.ascontiguousarray(csrs[w][0], dtype=np.uint32) for w in ("exc", "inh", "sic", "e2a")]
        keep += [np.ascontiguousarray(csrs[w][1], dtype=np.uint32) for w in ("exc", "inh", "sic", "e2a")]
        keep += [np.ascontiguousarray(params[w], dtype=np.float64)
                 for w in ("Ca_tot", "IP3_0", "tau_IP3", "delta_IP3", "V_reset", "b")]
        rs = (C.c_void_p * 4)(*[a.ctypes.data for a in keep[0:4]])
        src = (C.c_void_p * 4)(*[a.ctypes.data for a in keep[4:8]])
        par = (C.c_void_p * 6)(*[a.ctypes.data for a in keep[8:14]])
        self._h = load().orc_net_from_arrays(n_total, seed, rs, src, par)
        c = np.zeros(4, dtype=np.uint32)
        p = C.c_double()
        load().orc_net_sizes(self._h, _p(c), C.byref(p))
        self.n_exc, self.n_inh, self.n_astro, self.n_neurons = map(int, c)
        self.p_primary = p.value
        self.n_total, self.seed = n_total, seed
        return self

    def csr(self, which: str):
        lib = load()
        idx = self.CSR[which]
        ne = lib.orc_net_csr(self._h, idx, None, None)
        nt = self.n_astro if which == "e2a" else self.n_neurons
        rs = np.empty(nt + 1, dtype=np.uint32)
        src = np.empty(ne, dtype=np.uint32)
        lib.orc_net_csr(self._h, idx, _p(rs), _p(src) if ne else None)
        return rs, src

    def params(self, which: str) -> np.ndarray:
        idx = self.PARAMS[which]
        out = np.empty(self.n_astro if idx < 4 else self.n_neurons, dtype=np.float64)
        load().orc_net_param(self._h, idx, _p(out))
        return out

    def build_seconds(self) -> float:
        return load().orc_net_build_seconds(self._h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_net_free(self._h)
            self._h = None


class OracleSim:
    def __init__(self, net: OracleNet, dt=0.1, substeps=1, w_n2a=None, w_a2n=None, stp=True, threads=0):
        o = _Opts(dt, substeps, math.nan if w_n2a is None else w_n2a,
                  math.nan if w_a2n is None else w_a2n, 1 if stp else 0, threads)
        self.net = net  # keeps the network alive
        self._h = load().orc_sim_create(net._h, C.byref(o))
        self.n_neurons, self.n_astro = net.n_neurons, net.n_astro

    def step(self, k: int = 1, raster: bool = False):
        if raster:
            out = np.zeros((k, self.n_neurons), dtype=np.uint8)
            load().orc_sim_step(self._h, k, _p(out))
            return out
        load().orc_sim_step(self._h, k, None)
        return None

    @property
    def steps(self) -> int:
        return load().orc_sim_steps_done(self._h)

    def spike_count(self) -> int:
        return load().orc_sim_spike_count(self._h)

    def reset_ever_active(self):
        load().orc_sim_reset_ever_active(self._h)

    def state(self, field: str) -> np.ndarray:
        idx = FIELDS[field]
        out = np.empty(self.n_astro if idx < 3 else self.n_neurons, dtype=np.float64)
        load().orc_sim_read(self._h, idx, _p(out))
        return out

    def refractory(self) -> np.ndarray:
        out = np.empty(self.n_neurons, dtype=np.int32)
        load().orc_sim_read_refractory(self._h, _p(out))
        return out

    def ever_active(self) -> np.ndarray:
        out = np.empty(self.n_astro, dtype=np.uint8)
        load().orc_sim_read_ever_active(self._h, _p(out))
        return out

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_sim_free(self._h)
            self._h = None


def rng_bits(seed, stream, counter):
    return load().orc_rng_bits(seed, stream, counter)


def rng_uniform(seed, stream, counter):
    return load().orc_rng_uniform(seed, stream, counter)


def rng_normal(seed, stream):
    return load().orc_rng_normal(seed, stream)


def rng_poisson(seed, stream, lam):
    return load().orc_rng_poisson(seed, stream, lam)


def hostrng_normals(seed, stream, n):
    out = np.empty(n, dtype=np.float64)
    load().orc_hostrng_normals(seed, stream, n, _p(out))
    return out


# ---- per-cell step arithmetic (known-answer tests) ----

def astro_rhs(Ca_tot, IP3_0, tau_IP3, Ca, IP3, h, noise=0.0):
    """{dCa, dIP3, dh} of the Li-Rinzel right-hand side."""
    out = np.empty(3, dtype=np.float64)
    load().orc_astro_rhs(Ca_tot, IP3_0, tau_IP3, Ca, IP3, h, noise, _p(out))
    return out


def astro_advance(Ca_tot, IP3_0, tau_IP3, delta_IP3, state, ip3_input=0.0, noise=0.0, h_step=0.1, substeps=1):
    """(Ca, IP3, h) after the IP3 jump, `substeps` RK4 steps and the clamps."""
    st = np.array(state, dtype=np.float64)
    load().orc_astro_advance(Ca_tot, IP3_0, tau_IP3, delta_IP3, ip3_input, noise, h_step, substeps, _p(st))
    return st


def sic_factor(ca):
    return load().orc_sic_factor(ca)


def neuron_rhs(excitatory, V, w, g_ex=0.0, g_in=0.0, I_ext=0.0):
    """{dV, dw} of the AdEx right-hand side of the population."""
    out = np.empty(2, dtype=np.float64)
    load().orc_neuron_rhs(1 if excitatory else 0, V, w, g_ex, g_in, I_ext, _p(out))
    return out


def neuron_substeps(excitatory, state, refractory=0, I_ext=0.0, V_reset=-50.0, b=300.0, h_step=0.1, substeps=1):
    """(state {V, w, g_ex, dg_ex, g_in, dg_in}, refractory, spiked) after the substep loop."""
    st = np.array(state, dtype=np.float64)
    r = C.c_int32(refractory)
    sp = load().orc_neuron_substeps(1 if excitatory else 0, _p(st), C.byref(r), I_ext, V_reset, b, h_step, substeps)
    return st, r.value, bool(sp)


def psc_init(excitatory, inhibitory_synapse):
    return load().orc_psc_init(1 if excitatory else 0, 1 if inhibitory_synapse else 0)


def stp_on_spike(state, t_now):
    """(release x*u, new {x, u, t_last})."""
    st = np.array(state, dtype=np.float64)
    out = load().orc_stp_on_spike(_p(st), t_now)
    return out, st


def gather(row_start, src, ring_row, target):
    rs = np.ascontiguousarray(row_start, dtype=np.uint32)
    sr = np.ascontiguousarray(src if len(src) else [0], dtype=np.uint32)
    ring = np.ascontiguousarray(ring_row, dtype=np.float64)
    return load().orc_gather(_p(rs), _p(sr), len(rs) - 1, _p(ring), target)


def sim_params(sim: "OracleSim") -> dict:
    out = np.empty(5, dtype=np.int32)
    load().orc_sim_params(sim._h, _p(out))
    return dict(zip(("d_e", "d_i", "d_a2n", "R", "steps_per_noise"), (int(x) for x in out)))
