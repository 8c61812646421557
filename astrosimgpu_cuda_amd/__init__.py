"""astrosimgpu_cuda_amd -- MI355X-native astrosim_fused step pipeline.

Python face of the C ABI in ``include/astrosim.h`` (``lib/libastrosim.so``).
The objects mirror the reference program's own structure
(src/astrosim_fused.cu): a :class:`Network` is what ``ScalingConfig::make``,
``build_connectivity`` and the population draws produce from ``(--n, --seed)``;
a :class:`Simulation` is the device state plus the time loop, configured by
the reference's remaining flags (``--dt --substeps --block --w-n2a --w-a2n
--no-stp``). Every call goes to the HIP library; there is no fallback, and a
missing library raises :class:`LibraryNotBuilt`.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from pathlib import Path

import numpy as np

__all__ = [
    "AstrosimError", "LibraryNotBuilt", "Network", "Simulation", "lib_path", "cli_path",
    "device_info", "device_memory", "FIELDS", "FP64", "FP32C", "PLAIN", "BLOCKED", "load", "Batch", "Server",
]

PKG = Path(__file__).resolve().parent
FP64, FP32C = 0, 1
PLAIN, BLOCKED = 0, 1
FIELDS = {"Ca": 0, "IP3": 1, "h": 2, "V": 3, "w": 4, "g_ex": 5, "dg_ex": 6, "g_in": 7,
          "dg_in": 8, "I_sic_hold": 9, "stp_x": 10, "stp_u": 11, "stp_t_last": 12}
ASTRO_FIELDS = {"Ca", "IP3", "h"}
_STATUS = {0: "AS_OK", 1: "AS_ERR_INVALID", 2: "AS_ERR_HIP", 3: "AS_ERR_NOMEM",
           4: "AS_ERR_NODEVICE", 5: "AS_ERR_IO", 6: "AS_ERR_RANGE"}


class AstrosimError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS.get(status, status)}: {message}")
        self.status = status


class LibraryNotBuilt(RuntimeError):
    pass


def lib_path() -> Path:
    return PKG / "lib" / "libastrosim.so"


def cli_path() -> Path:
    return PKG / "bin" / "astrosim_fused"


class _Config(C.Structure):
    _fields_ = [("dt", C.c_double), ("substeps", C.c_int32), ("block", C.c_int32),
                ("w_n2a", C.c_double), ("w_a2n", C.c_double), ("stp_enabled", C.c_int32),
                ("precision", C.c_int32), ("method", C.c_int32)]


class _NetInfo(C.Structure):
    _fields_ = [("n_total", C.c_uint32), ("n_exc", C.c_uint32), ("n_inh", C.c_uint32),
                ("n_astro", C.c_uint32), ("n_neurons", C.c_uint32), ("p_primary", C.c_double),
                ("n_exc_edges", C.c_uint64), ("n_inh_edges", C.c_uint64),
                ("n_e2a_edges", C.c_uint64), ("n_sic_edges", C.c_uint64),
                ("build_seconds", C.c_double), ("built_on_gpu", C.c_int32)]


class _BatchCap(C.Structure):
    _fields_ = [("neurons", C.c_uint64), ("astros", C.c_uint64), ("edges", C.c_uint64 * 4),
                ("members", C.c_uint32)]


class _BatchInfo(C.Structure):
    _fields_ = [("global_step", C.c_int64), ("block_steps", C.c_int32), ("active_members", C.c_int32),
                ("block_threads", C.c_int32), ("device_bytes", C.c_uint64), ("free_neurons", C.c_uint64),
                ("bucket_overflows", C.c_uint64), ("chunk_steps", C.c_int32)]


class _ServerConfig(C.Structure):
    _fields_ = [("dt", C.c_double), ("block", C.c_int32), ("hbm_fraction", C.c_double),
                ("fp64_share", C.c_double), ("build_threads", C.c_int32), ("blocks_per_round", C.c_int32),
                ("max_members", C.c_int32), ("setup_where", C.c_int32), ("fail_after_rounds", C.c_int32)]


class _Request(C.Structure):
    _fields_ = [("n_total", C.c_uint32), ("seed", C.c_uint64), ("n_steps", C.c_int64), ("precision", C.c_int32)]


class _RequestResult(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("n_neurons", C.c_uint32), ("n_astro", C.c_uint32),
                ("t_submit", C.c_double), ("t_built", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("spikes", C.c_uint64), ("mean_ca", C.c_double),
                ("active_astro", C.c_uint32), ("t_build_start", C.c_double), ("built_on_gpu", C.c_int32)]


class _ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("p50_latency_s", C.c_double), ("p99_latency_s", C.c_double), ("mean_latency_s", C.c_double),
                ("sims_per_s", C.c_double), ("device_bytes", C.c_uint64), ("peak_active", C.c_uint64),
                ("peak_resident_cells", C.c_uint64), ("capacity_cells", C.c_uint64), ("gpu_busy_s", C.c_double),
                ("cell_steps", C.c_double), ("admit_s", C.c_double), ("retire_s", C.c_double),
                ("dedicated_queues", C.c_int32), ("build_streams", C.c_int32)]


_lib = None

# (name, restype, argtypes) for every symbol include/astrosim.h declares.
SIGNATURES = [
    ("as_config_default", None, [C.POINTER(_Config)]),
    ("as_network_build", C.c_int, [C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("as_network_build_on", C.c_int, [C.c_uint32, C.c_uint64, C.c_int32, C.POINTER(C.c_void_p)]),
    ("as_network_info_get", C.c_int, [C.c_void_p, C.POINTER(_NetInfo)]),
    ("as_network_csr", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    ("as_network_params", C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ("as_network_dump_edges", C.c_int, [C.c_void_p, C.c_char_p]),
    ("as_network_free", None, [C.c_void_p]),
    ("as_sim_footprint", C.c_int, [C.c_void_p, C.POINTER(_Config), C.POINTER(C.c_uint64)]),
    ("as_sim_create", C.c_int, [C.c_void_p, C.POINTER(_Config), C.POINTER(C.c_void_p)]),
    ("as_sim_step", C.c_int, [C.c_void_p, C.c_int64]),
    ("as_sim_sync", C.c_int, [C.c_void_p]),
    ("as_sim_steps", C.c_int64, [C.c_void_p]),
    ("as_sim_spike_count", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    ("as_sim_reset_ever_active", C.c_int, [C.c_void_p]),
    ("as_sim_read_state", C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ("as_sim_read_refractory", C.c_int, [C.c_void_p, C.c_void_p]),
    ("as_sim_read_ever_active", C.c_int, [C.c_void_p, C.c_void_p]),
    ("as_sim_read_spike_counts", C.c_int, [C.c_void_p, C.c_void_p]),
    ("as_sim_read_raster", C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]),
    ("as_sim_read_spikes", C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_uint64,
                                     C.POINTER(C.c_uint64)]),
    ("as_sim_read_spikes_noalloc", C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_uint64,
                                             C.POINTER(C.c_uint64)]),
    ("as_sim_bytes", C.c_uint64, [C.c_void_p]),
    ("as_sim_set_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("as_sim_kernel_time", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    ("as_sim_destroy", None, [C.c_void_p]),
    ("as_batch_capacity_for", C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(_BatchCap)]),
    ("as_batch_footprint", C.c_int, [C.POINTER(_BatchCap), C.POINTER(_Config), C.POINTER(C.c_uint64)]),
    ("as_batch_create", C.c_int, [C.POINTER(_BatchCap), C.POINTER(_Config), C.POINTER(C.c_void_p)]),
    ("as_batch_fits", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]),
    ("as_batch_add", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int32)]),
    ("as_batch_add_many", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int32)]),
    ("as_batch_remove", C.c_int, [C.c_void_p, C.c_int32]),
    ("as_batch_step", C.c_int, [C.c_void_p, C.c_int64]),
    ("as_batch_run_graph", C.c_int, [C.c_void_p, C.c_int64]),
    ("as_batch_autotune", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                    C.POINTER(C.c_int32)]),
    ("as_batch_sync", C.c_int, [C.c_void_p]),
    ("as_batch_get_info", C.c_int, [C.c_void_p, C.POINTER(_BatchInfo)]),
    ("as_batch_member_steps", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("as_batch_spike_count", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64)]),
    ("as_batch_reset_ever_active", C.c_int, [C.c_void_p, C.c_int32]),
    ("as_batch_read_state", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    ("as_batch_read_ever_active", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    ("as_batch_read_raster", C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_int32, C.c_void_p]),
    ("as_batch_read_spike_counts", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    ("as_batch_set_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("as_batch_kernel_time", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    ("as_batch_destroy", None, [C.c_void_p]),
    ("as_server_config_default", None, [C.POINTER(_ServerConfig)]),
    ("as_server_create", C.c_int, [C.POINTER(_ServerConfig), C.POINTER(C.c_void_p)]),
    ("as_server_submit", C.c_int, [C.c_void_p, C.POINTER(_Request), C.POINTER(C.c_int64)]),
    ("as_server_result", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(_RequestResult)]),
    ("as_server_wait_all", C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_int32)]),
    ("as_server_get_stats", C.c_int, [C.c_void_p, C.POINTER(_ServerStats)]),
    ("as_server_destroy", None, [C.c_void_p]),
    ("as_device_info", C.c_int, [C.c_char_p, C.c_int, C.c_char_p, C.c_int]),
    ("as_device_memory", C.c_int, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("as_last_error", C.c_char_p, []),
    ("as_version", C.c_char_p, []),
]


def load():
    """Loads lib/libastrosim.so (built by ``python -m astrosimgpu_cuda_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not path.exists():
        raise LibraryNotBuilt(f"{path} is missing: run `python -m astrosimgpu_cuda_amd.build` first")
    lib = C.CDLL(str(path))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != 0:
        raise AstrosimError(status, load().as_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def device_info() -> tuple[str, str]:
    lib = load()
    name = C.create_string_buffer(256)
    arch = C.create_string_buffer(64)
    _check(lib.as_device_info(name, 256, arch, 64))
    return name.value.decode(), arch.value.decode()


def device_memory() -> tuple[int, int]:
    lib = load()
    f, t = C.c_uint64(), C.c_uint64()
    _check(lib.as_device_memory(C.byref(f), C.byref(t)))
    return f.value, t.value


class Network:
    """Connectivity and per-cell parameters for ``(n_total, seed)`` (host only).

    Reference: ScalingConfig::make (src/astrosim_fused.cu:606-614),
    build_connectivity (:631-735), population draws (:917-933).
    """

    CSR_NAMES = {"exc": 0, "inh": 1, "sic": 2, "e2a": 3}
    PARAM_NAMES = {"Ca_tot": 0, "IP3_0": 1, "tau_IP3": 2, "delta_IP3": 3, "V_reset": 4, "b": 5}

    BUILD_WHERE = {"host": 0, "gpu": 1, "gpu-only": 2}

    def __init__(self, n_total: int, seed: int = 1, setup: str = "host"):
        """``setup``: "host" walks the setup stream sequentially; "gpu" generates
        the same connectivity with the parallel generator on the current GPU
        from N_total = 100000 on (the host walk below that, and where the
        generator declines); "gpu-only" uses the generator for every size it
        accepts and raises otherwise. A GPU-built network's connectivity stays
        in device memory (simulations copy it device to device); csr() copies
        it to the host on first use."""
        lib = load()
        h = C.c_void_p()
        _check(lib.as_network_build_on(C.c_uint32(n_total), C.c_uint64(seed),
                                       C.c_int32(self.BUILD_WHERE[setup]), C.byref(h)))
        self._h = h
        info = _NetInfo()
        _check(lib.as_network_info_get(h, C.byref(info)))
        self.n_total, self.seed = n_total, seed
        self.n_exc, self.n_inh, self.n_astro = info.n_exc, info.n_inh, info.n_astro
        self.n_neurons = info.n_neurons
        self.p_primary = info.p_primary
        self.n_edges = {"exc": info.n_exc_edges, "inh": info.n_inh_edges,
                        "e2a": info.n_e2a_edges, "sic": info.n_sic_edges}
        self.build_seconds = info.build_seconds
        self.built_on_gpu = bool(info.built_on_gpu)

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("network was freed")
        return self._h

    def csr(self, which: str) -> tuple[np.ndarray, np.ndarray]:
        idx = self.CSR_NAMES[which]
        n_t = self.n_astro if which == "e2a" else self.n_neurons
        rs = np.empty(n_t + 1, dtype=np.uint32)
        src = np.empty(self.n_edges[which], dtype=np.uint32)
        _check(load().as_network_csr(self.handle, idx, _ptr(rs), _ptr(src) if src.size else None))
        return rs, src

    def params(self, which: str) -> np.ndarray:
        idx = self.PARAM_NAMES[which]
        out = np.empty(self.n_astro if idx < 4 else self.n_neurons, dtype=np.float64)
        _check(load().as_network_params(self.handle, idx, _ptr(out)))
        return out

    def dump_edges(self, path: str | os.PathLike) -> None:
        _check(load().as_network_dump_edges(self.handle, str(path).encode()))

    def free(self) -> None:
        if self._h is not None:
            load().as_network_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def make_config(dt=0.1, substeps=1, block=128, w_n2a=None, w_a2n=None, stp=True,
                precision="fp64", method="plain") -> _Config:
    cfg = _Config()
    load().as_config_default(C.byref(cfg))
    cfg.dt = dt
    cfg.substeps = substeps
    cfg.block = block
    cfg.w_n2a = math.nan if w_n2a is None else w_n2a
    cfg.w_a2n = math.nan if w_a2n is None else w_a2n
    cfg.stp_enabled = 1 if stp else 0
    cfg.precision = {"fp64": FP64, "fp32c": FP32C}[precision] if isinstance(precision, str) else precision
    cfg.method = {"plain": PLAIN, "blocked": BLOCKED}[method] if isinstance(method, str) else method
    return cfg


class Simulation:
    """Device state and time loop of one network on the GPU.

    Reference: device bring-up and upload (src/astrosim_fused.cu:904-1018),
    the step loop (:1042-1064), diagnostics (:1066-1125).
    """

    def __init__(self, network: Network, **kw):
        lib = load()
        self.cfg = make_config(**kw)
        self.n_neurons, self.n_astro, self.n_exc = network.n_neurons, network.n_astro, network.n_exc
        h = C.c_void_p()
        _check(lib.as_sim_create(network.handle, C.byref(self.cfg), C.byref(h)))
        self._h = h

    @staticmethod
    def footprint(network: Network, **kw) -> int:
        cfg = make_config(**kw)
        out = C.c_uint64()
        _check(load().as_sim_footprint(network.handle, C.byref(cfg), C.byref(out)))
        return out.value

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("simulation was destroyed")
        return self._h

    def step(self, k: int = 1) -> None:
        _check(load().as_sim_step(self.handle, C.c_int64(k)))

    def sync(self) -> None:
        _check(load().as_sim_sync(self.handle))

    @property
    def steps(self) -> int:
        return load().as_sim_steps(self.handle)

    def spike_count(self) -> int:
        out = C.c_uint64()
        _check(load().as_sim_spike_count(self.handle, C.byref(out)))
        return out.value

    def reset_ever_active(self) -> None:
        _check(load().as_sim_reset_ever_active(self.handle))

    def state(self, field: str) -> np.ndarray:
        n = self.n_astro if field in ASTRO_FIELDS else self.n_neurons
        out = np.empty(n, dtype=np.float64)
        _check(load().as_sim_read_state(self.handle, FIELDS[field], _ptr(out)))
        return out

    def refractory(self) -> np.ndarray:
        out = np.empty(self.n_neurons, dtype=np.int32)
        _check(load().as_sim_read_refractory(self.handle, _ptr(out)))
        return out

    def ever_active(self) -> np.ndarray:
        out = np.empty(self.n_astro, dtype=np.uint8)
        _check(load().as_sim_read_ever_active(self.handle, _ptr(out)))
        return out

    def spike_counts(self) -> np.ndarray:
        """Spikes of every neuron since creation."""
        out = np.empty(self.n_neurons, dtype=np.uint32)
        _check(load().as_sim_read_spike_counts(self.handle, _ptr(out)))
        return out

    def raster(self, first_step: int, n_steps: int) -> np.ndarray:
        out = np.empty((n_steps, self.n_neurons), dtype=np.uint8)
        _check(load().as_sim_read_raster(self.handle, C.c_int64(first_step), C.c_int32(n_steps), _ptr(out)))
        return out

    def spikes(self, first_step: int, n_steps: int, cap: int | None = None,
               no_alloc: bool = False) -> tuple[np.ndarray, np.ndarray]:
        """(steps, neurons) of the spikes of held steps [first_step, first_step
        + n_steps), in step then neuron order. With cap=None the buffers grow
        until every event fits; no_alloc uses the allocation-free entry point."""
        fn = load().as_sim_read_spikes_noalloc if no_alloc else load().as_sim_read_spikes
        want = 4096 if cap is None else cap
        while True:
            steps = np.empty(max(1, want), dtype=np.int64)
            neurons = np.empty(max(1, want), dtype=np.uint32)
            count = C.c_uint64()
            _check(fn(self.handle, C.c_int64(first_step), C.c_int32(n_steps), _ptr(steps), _ptr(neurons),
                      C.c_uint64(want), C.byref(count)))
            n = count.value
            if n <= want or cap is not None:
                k = min(n, want)
                return steps[:k].copy(), neurons[:k].copy()
            want = n

    def device_bytes(self) -> int:
        return load().as_sim_bytes(self.handle)

    def set_timing(self, on: bool) -> None:
        _check(load().as_sim_set_timing(self.handle, 1 if on else 0))

    def kernel_time(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_int64()
        _check(load().as_sim_kernel_time(self.handle, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self) -> None:
        if self._h is not None:
            load().as_sim_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many networks advanced together in one launch per block (blocked
    method). Members join and leave at block boundaries; each member's
    results are identical to running it alone."""

    def __init__(self, capacity: "_BatchCap | list[Network]", **kw):
        lib = load()
        kw.setdefault("method", "blocked")
        self.cfg = make_config(**kw)
        cap = capacity if isinstance(capacity, _BatchCap) else Batch.capacity_for(capacity)
        self.capacity = cap
        h = C.c_void_p()
        _check(lib.as_batch_create(C.byref(cap), C.byref(self.cfg), C.byref(h)))
        self._h = h
        self._sizes = {}

    @staticmethod
    def capacity_for(nets: list, members: int | None = None) -> _BatchCap:
        cap = _BatchCap()
        arr = (C.c_void_p * max(1, len(nets)))(*[n.handle.value for n in nets])
        _check(load().as_batch_capacity_for(arr, len(nets), C.byref(cap)))
        if members is not None:
            cap.members = members
        return cap

    @staticmethod
    def footprint(cap: _BatchCap, **kw) -> int:
        kw.setdefault("method", "blocked")
        cfg = make_config(**kw)
        out = C.c_uint64()
        _check(load().as_batch_footprint(C.byref(cap), C.byref(cfg), C.byref(out)))
        return out.value

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("batch was destroyed")
        return self._h

    def fits(self, net: Network) -> bool:
        f = C.c_int32()
        _check(load().as_batch_fits(self.handle, net.handle, C.byref(f)))
        return bool(f.value)

    def add(self, net: Network, n_steps: int) -> int:
        m = C.c_int32()
        _check(load().as_batch_add(self.handle, net.handle, C.c_int64(n_steps), C.byref(m)))
        self._sizes[m.value] = (net.n_neurons, net.n_astro)
        return m.value

    def add_many(self, nets, n_steps) -> list[int]:
        """add() for several networks, uploaded back to back and waited for
        once; ``n_steps`` is one count for all or one per network."""
        nets = list(nets)
        steps = [int(n_steps)] * len(nets) if isinstance(n_steps, int) else [int(s) for s in n_steps]
        if len(steps) != len(nets):
            raise ValueError("one step count per network")
        hs = (C.c_void_p * len(nets))(*[n.handle for n in nets])
        st = (C.c_int64 * len(nets))(*steps)
        ms = (C.c_int32 * len(nets))()
        err = load().as_batch_add_many(self.handle, hs, len(nets), st, ms)
        for m, n in zip(ms, nets):
            if m >= 0:
                self._sizes[m] = (n.n_neurons, n.n_astro)
        _check(err)
        return list(ms)

    def remove(self, member: int) -> None:
        _check(load().as_batch_remove(self.handle, member))

    def step(self, k: int) -> None:
        _check(load().as_batch_step(self.handle, C.c_int64(k)))

    def run_graph(self, blocks: int) -> None:
        _check(load().as_batch_run_graph(self.handle, C.c_int64(blocks)))

    def autotune(self, candidates=(64, 128, 256, 512), blocks_each: int = 2) -> int:
        arr = (C.c_int32 * len(candidates))(*candidates)
        out = C.c_int32()
        _check(load().as_batch_autotune(self.handle, arr, len(candidates), blocks_each, C.byref(out)))
        return out.value

    def sync(self) -> None:
        _check(load().as_batch_sync(self.handle))

    def info(self) -> dict:
        i = _BatchInfo()
        _check(load().as_batch_get_info(self.handle, C.byref(i)))
        return {f: getattr(i, f) for f, _ in _BatchInfo._fields_}

    def member_steps(self, member: int) -> tuple[int, int]:
        d, t = C.c_int64(), C.c_int64()
        _check(load().as_batch_member_steps(self.handle, member, C.byref(d), C.byref(t)))
        return d.value, t.value

    def spike_count(self, member: int) -> int:
        out = C.c_uint64()
        _check(load().as_batch_spike_count(self.handle, member, C.byref(out)))
        return out.value

    def reset_ever_active(self, member: int) -> None:
        _check(load().as_batch_reset_ever_active(self.handle, member))

    def state(self, member: int, field: str) -> np.ndarray:
        nn, na = self._sizes[member]
        out = np.empty(na if field in ASTRO_FIELDS else nn, dtype=np.float64)
        _check(load().as_batch_read_state(self.handle, member, FIELDS[field], _ptr(out)))
        return out

    def ever_active(self, member: int) -> np.ndarray:
        out = np.empty(self._sizes[member][1], dtype=np.uint8)
        _check(load().as_batch_read_ever_active(self.handle, member, _ptr(out)))
        return out

    def raster(self, member: int, first_step: int, n_steps: int) -> np.ndarray:
        out = np.empty((n_steps, self._sizes[member][0]), dtype=np.uint8)
        _check(load().as_batch_read_raster(self.handle, member, C.c_int64(first_step), C.c_int32(n_steps),
                                           _ptr(out)))
        return out

    def spike_counts(self, member: int) -> np.ndarray:
        out = np.empty(self._sizes[member][0], dtype=np.uint32)
        _check(load().as_batch_read_spike_counts(self.handle, member, _ptr(out)))
        return out

    def set_timing(self, on: bool) -> None:
        _check(load().as_batch_set_timing(self.handle, 1 if on else 0))

    def kernel_time(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_int64()
        _check(load().as_batch_kernel_time(self.handle, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self) -> None:
        if self._h is not None:
            load().as_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


REQUEST_STATUS = {0: "building", 1: "queued", 2: "running", 3: "done", 4: "rejected"}


class Server:
    """Continuous-batching simulation server (native, include/astrosim.h):
    requests are built by worker threads (networks from N_total = 100000 on
    with the GPU connectivity generator, kept in device memory; smaller ones
    by the host walk; ``setup_where`` as in include/astrosim.h), admitted into
    resident FP32c/FP64 batches as their footprint fits, advanced in rounds
    of blocks pipelined two deep, and retired when done. Each server stream
    has a hardware queue of its own (created with a CU mask covering every
    CU), so it needs no GPU_MAX_HW_QUEUES setting."""

    def __init__(self, **kw):
        lib = load()
        cfg = _ServerConfig()
        lib.as_server_config_default(C.byref(cfg))
        for k, v in kw.items():
            setattr(cfg, k, v)
        h = C.c_void_p()
        _check(lib.as_server_create(C.byref(cfg), C.byref(h)))
        self._h = h

    def submit(self, n_total: int, seed: int, n_steps: int, precision: str = "fp32c") -> int:
        r = _Request(n_total, seed, n_steps, {"fp64": FP64, "fp32c": FP32C}[precision])
        out = C.c_int64()
        _check(load().as_server_submit(self._h, C.byref(r), C.byref(out)))
        return out.value

    def result(self, rid: int) -> dict:
        r = _RequestResult()
        _check(load().as_server_result(self._h, rid, C.byref(r)))
        d = {f: getattr(r, f) for f, _ in _RequestResult._fields_}
        d["status"] = REQUEST_STATUS[d["status"]]
        return d

    def wait_all(self, timeout_s: float) -> bool:
        done = C.c_int32()
        _check(load().as_server_wait_all(self._h, timeout_s, C.byref(done)))
        return bool(done.value)

    def stats(self) -> dict:
        s = _ServerStats()
        _check(load().as_server_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _ServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            load().as_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_cli(args: list[str], timeout: float | None = None) -> subprocess.CompletedProcess:
    """Runs the ``astrosim_fused`` drop-in binary with the reference's flags."""
    exe = cli_path()
    if not exe.exists():
        raise LibraryNotBuilt(f"{exe} is missing: run `python -m astrosimgpu_cuda_amd.build` first")
    return subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=timeout)
