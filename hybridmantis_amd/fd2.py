"""Python host interface over libfd2.so (ctypes).

Mirrors the reference's optical-transport operator: a parameter set with the
names of hybridMANTIS_input.in, a buffer of (x, y, z, E, history, N) events
handed over in one call, and the reference's tallies back (PRF image,
photons detected per history / PHS, and the /optstats/ counters).  Every
call goes to the HIP library; if libfd2.so has not been built, importing
this module's functions raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .events import EVENT_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
# FD2_LIBRARY selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("FD2_LIBRARY") or os.path.join(HERE, "libfd2.so")

PHOTON_DTYPE = np.dtype([("fate", "<i4"), ("pix", "<i4"), ("x", "<f4"),
                         ("y", "<f4")])
STEP_DTYPE = np.dtype([("pos_in", "<f4", 3), ("dcos", "<f4", 3), ("xc", "<f4"),
                       ("yc", "<f4"), ("pos_out", "<f4", 3), ("code", "<i4"),
                       ("photon", "<i4"), ("decisions", "<u4")])
COUNTER_NAMES = ["generated", "detected", "abs_top", "abs_bulk", "lost",
                 "outofcol", "theta1", "regenerated", "surr_undetected"]

FD2_OK = 0


class Fd2Error(RuntimeError):
    pass


class Params(C.Structure):
    """fd2_params == the reference's COMMON /inputargs/ layout."""
    _fields_ = [(n, C.c_double) for n in (
        "detx", "dety", "detheight", "detradius", "detnC", "detnIC", "dettop",
        "detbulk", "detbeta", "detdmin", "detdmax", "detlboundx", "detlboundy",
        "detuboundx", "detuboundy", "detyield", "detsensorRefl")] + [
        (n, C.c_int32) for n in ("detpixel", "rungpu", "machinenum",
                                 "mynumhist", "minphotons", "maxphotons",
                                 "mynumbins")]


def _tuning_dict(t) -> dict:
    return {name: getattr(t, name) for name, _ in Tuning._fields_}


class Tuning(C.Structure):
    _fields_ = [("chunk", C.c_int32), ("waves_per_cu", C.c_int32),
                ("lds_tile", C.c_int32), ("park_lanes", C.c_int32),
                ("refill_lanes", C.c_int32), ("wall_queue", C.c_int32),
                ("max_chunks", C.c_int64)]


# dict keys (events.BASE_PARAMS) -> fd2_params members
_PARAM_MAP = dict(
    xdet="detx", ydet="dety", H="detheight", R="detradius", n_C="detnC",
    n_IC="detnIC", top_absfrac="dettop", bulk="detbulk", beta="detbeta",
    d_min="detdmin", d_max="detdmax", lbx="detlboundx", lby="detlboundy",
    ubx="detuboundx", uby="detuboundy", yield_per_ev="detyield",
    sensor_refl="detsensorRefl", pixel="detpixel", rungpu="rungpu",
    machine="machinenum", num_hist="mynumhist", min_ph="minphotons",
    max_ph="maxphotons", num_bins="mynumbins")


def to_params(p: dict) -> Params:
    out = Params()
    for k, v in _PARAM_MAP.items():
        if k in p:
            setattr(out, v, p[k])
    return out


_lib = None


def lib() -> C.CDLL:
    """Load libfd2.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise Fd2Error(f"{LIB_PATH} is missing: build it with "
                       "`python -m hybridmantis_amd.build` (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "fd2_create": ([C.POINTER(Params), C.c_int, C.c_int, C.POINTER(P)], C.c_int),
        "fd2_destroy": ([P], None),
        "fd2_set_seed": ([P, C.c_uint32], C.c_int),
        "fd2_set_tuning": ([P, C.POINTER(Tuning)], C.c_int),
        "fd2_get_tuning": ([P, C.POINTER(Tuning)], C.c_int),
        "fd2_submit": ([P, P, C.c_int64], C.c_int),
        "fd2_submit_device": ([P, P, C.c_int64], C.c_int),
        "fd2_sync": ([P], C.c_int),
        "fd2_submit_record": ([P, P, C.c_int64, P], C.c_int),
        "fd2_trace_ranecu": ([P, P, C.c_int64, C.c_int64, C.c_int32, P], C.c_int),
        "fd2_trace_ranecu_faithful": ([P, P, C.c_int64, C.c_int64, C.c_int32, P], C.c_int),
        "fd2_set_next_event": ([P, C.c_int64], C.c_int),
        "fd2_submit_device_noalloc": ([P, P, C.c_int64], C.c_int),
        "fd2_debug_steps": ([P, P, C.c_int64, C.c_int32, P, C.c_int32,
                             C.POINTER(C.c_int32)], C.c_int),
        "fd2_debug_math": ([C.c_int, P, P, C.c_int64, C.c_int], C.c_int),
        "fd2_prf_dims": ([P, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
        "fd2_get_prf": ([P, C.c_int, P], C.c_int),
        "fd2_get_hist_counts": ([P, C.c_int, P], C.c_int),
        "fd2_get_phs": ([P, C.c_int, P], C.c_int),
        "fd2_get_counters": ([P, C.c_int, P], C.c_int),
        "fd2_reset": ([P], C.c_int),
        "fd2_last_kernel_ms": ([P, C.POINTER(C.c_float)], C.c_int),
        "fd2_phs_from_counts": ([P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P], C.c_int),
        "fd2_device_alloc": ([P, C.c_size_t, C.POINTER(P)], C.c_int),
        "fd2_device_free": ([P, P], C.c_int),
        "fd2_copy_to_device": ([P, P, P, C.c_size_t], C.c_int),
        "fd2_strerror": ([C.c_int], C.c_char_p),
        "fd2_graph_create": ([P, P, C.c_int64, C.POINTER(P)], C.c_int),
        "fd2_graph_launch": ([P], C.c_int),
        "fd2_graph_destroy": ([P], None),
        "fd2_autotune": ([P, P, C.c_int64, C.c_int32, P, P, C.c_int32,
                          C.POINTER(C.c_int32)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc: int, what: str) -> None:
    if rc != FD2_OK:
        raise Fd2Error(f"{what}: {lib().fd2_strerror(rc).decode()} ({rc})")


# fd2_debug_math function numbers (FD2_MATH_* in fd2.h)
MATH_FN = {"sin": 0, "cos": 1, "log": 2, "asin": 3, "acos": 4, "sin_turns": 5,
           "cos_turns": 6, "rsqrt": 7, "sqrt_unit": 8, "angle_over_157": 9,
           "acosf_libc": 10, "asinf_libc": 11, "ff_sin": 12, "ff_cos": 13, "ff_log": 14}


def debug_math(fn: str, x, device: int = 0) -> np.ndarray:
    """The transport's device arithmetic on the GPU: f(x) elementwise."""
    xs = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(xs)
    _check(lib().fd2_debug_math(MATH_FN[fn], xs.ctypes.data, y.ctypes.data, xs.size, device),
           "fd2_debug_math")
    return y


def as_events(ev) -> np.ndarray:
    a = np.ascontiguousarray(ev)
    if a.dtype != EVENT_DTYPE:
        out = np.zeros(len(a), EVENT_DTYPE)
        for k in a.dtype.names:
            if k in EVENT_DTYPE.names:
                out[k] = a[k]
        a = out
    return a


def phs_from_counts(hist_counts, min_ph: int, max_ph: int, num_bins: int) -> np.ndarray:
    """The reference's PHS binning rule (hybridMANTIS_cuda_ver1_0.cu:376-399)."""
    h = np.ascontiguousarray(hist_counts, np.uint32)
    out = np.zeros(num_bins, np.int32)
    _check(lib().fd2_phs_from_counts(h.ctypes.data, len(h), min_ph, max_ph,
                                     num_bins, out.ctypes.data), "phs")
    return out


class DeviceBuffer:
    """Device allocation owned through the library (no other GPU runtime)."""

    def __init__(self, tr: "Transport", nbytes: int):
        self._tr = tr
        self.nbytes = nbytes
        p = C.c_void_p()
        _check(lib().fd2_device_alloc(tr._ctx, nbytes, C.byref(p)), "device_alloc")
        self.ptr = p.value

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the device buffer")
        _check(lib().fd2_copy_to_device(self._tr._ctx, self.ptr, a.ctypes.data,
                                        a.nbytes), "copy_to_device")

    def free(self) -> None:
        if self.ptr:
            _check(lib().fd2_device_free(self._tr._ctx, self.ptr), "device_free")
            self.ptr = None


class TuneResult(C.Structure):
    _fields_ = [("tuning", Tuning), ("ms", C.c_float)]


class Graph:
    """fd2_graph: a captured batch, replayed by launch()."""

    def __init__(self, tr: "Transport", buf: DeviceBuffer, n: int):
        self._tr = tr
        self.n = n
        h = C.c_void_p()
        _check(lib().fd2_graph_create(tr._ctx, buf.ptr, n, C.byref(h)), "fd2_graph_create")
        self._h = h

    def launch(self) -> None:
        _check(lib().fd2_graph_launch(self._h), "fd2_graph_launch")

    def close(self) -> None:
        if self._h:
            lib().fd2_graph_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Transport:
    """One GPU context: parameter sets (geometries), tallies, work queue."""

    def __init__(self, params, device: int = 0, seed: int = 271828182,
                 tuning: dict | None = None):
        plist = params if isinstance(params, (list, tuple)) else [params]
        self.params = [dict(p) for p in plist]
        arr = (Params * len(plist))(*[to_params(p) for p in plist])
        ctx = C.c_void_p()
        _check(lib().fd2_create(arr, len(plist), device, C.byref(ctx)), "fd2_create")
        self._ctx = ctx
        _check(lib().fd2_set_seed(ctx, seed & 0xFFFFFFFF), "set_seed")
        if tuning:
            self.set_tuning(**tuning)

    # -- lifecycle
    def close(self) -> None:
        if getattr(self, "_ctx", None):
            lib().fd2_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration
    def set_tuning(self, chunk: int = 0, waves_per_cu: int = 0, lds_tile: int = 0,
                   park_lanes: int = 0, refill_lanes: int = 0, wall_queue: int = 0,
                   max_chunks: int = 0):
        t = Tuning(chunk, waves_per_cu, lds_tile, park_lanes, refill_lanes, wall_queue,
                   max_chunks)
        _check(lib().fd2_set_tuning(self._ctx, C.byref(t)), "set_tuning")

    def tuning(self) -> dict:
        t = Tuning()
        _check(lib().fd2_get_tuning(self._ctx, C.byref(t)), "get_tuning")
        return _tuning_dict(t)

    def set_next_event(self, first: int) -> None:
        """Number the next submitted event `first` (fd2_set_next_event)."""
        _check(lib().fd2_set_next_event(self._ctx, first), "set_next_event")

    def set_seed(self, seed: int) -> None:
        _check(lib().fd2_set_seed(self._ctx, seed & 0xFFFFFFFF), "set_seed")

    # -- transport
    def submit(self, events) -> None:
        ev = as_events(events)
        _check(lib().fd2_submit(self._ctx, ev.ctypes.data, len(ev)), "fd2_submit")

    def submit_device(self, buf: DeviceBuffer, n: int) -> None:
        _check(lib().fd2_submit_device(self._ctx, buf.ptr, n), "fd2_submit_device")

    def submit_device_noalloc(self, buf: DeviceBuffer, n: int) -> None:
        """fd2_submit_device_noalloc: synchronous, allocates nothing."""
        _check(lib().fd2_submit_device_noalloc(self._ctx, buf.ptr, n),
               "fd2_submit_device_noalloc")

    def sync(self) -> None:
        _check(lib().fd2_sync(self._ctx), "fd2_sync")

    def submit_record(self, events) -> np.ndarray:
        ev = as_events(events)
        rec = np.zeros(int(ev["N"].sum()), PHOTON_DTYPE)
        _check(lib().fd2_submit_record(self._ctx, ev.ctypes.data, len(ev),
                                       rec.ctypes.data), "fd2_submit_record")
        return rec

    def trace_ranecu(self, events, event_base: int = 0,
                     seed_input: int = 271828182, record: bool = True,
                     faithful: bool = False):
        """One lane per event on the reference's per-event RANECU stream;
        faithful=True runs it in the reference C path's own arithmetic
        (fd2_trace_ranecu_faithful), else in the production FP32 arithmetic."""
        ev = as_events(events)
        rec = np.zeros(int(ev["N"].sum()), PHOTON_DTYPE) if record else None
        name = "fd2_trace_ranecu_faithful" if faithful else "fd2_trace_ranecu"
        _check(getattr(lib(), name)(self._ctx, ev.ctypes.data, len(ev), event_base,
                                    seed_input,
                                    rec.ctypes.data if rec is not None else None),
               name)
        return rec

    def debug_steps(self, events, max_photons: int, max_steps: int) -> np.ndarray:
        ev = as_events(events)
        out = np.zeros(max_steps, STEP_DTYPE)
        n = C.c_int32()
        _check(lib().fd2_debug_steps(self._ctx, ev.ctypes.data, len(ev), max_photons,
                                     out.ctypes.data, max_steps, C.byref(n)),
               "fd2_debug_steps")
        return out[:n.value]

    def device_buffer(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def upload_events(self, events) -> DeviceBuffer:
        """Events copied to a new device buffer (for submit_device, graph,
        autotune)."""
        ev = as_events(events)
        buf = DeviceBuffer(self, max(ev.nbytes, 1))
        buf.upload(ev)
        buf.n = len(ev)
        return buf

    def graph(self, buf: DeviceBuffer, n: int) -> "Graph":
        """HIP graph of the queue build + transport of n device events."""
        return Graph(self, buf, n)

    def autotune(self, buf: DeviceBuffer, n: int, repeats: int = 2,
                 max_results: int = 64):
        """Pick the fastest launch shape on these events (fd2_autotune); it
        is set on this context.  Returns (best tuning, [(tuning, ms), ...])."""
        best = Tuning()
        res = (TuneResult * max_results)()
        nres = C.c_int32(0)
        _check(lib().fd2_autotune(self._ctx, buf.ptr, n, repeats, C.byref(best), res,
                                  max_results, C.byref(nres)), "fd2_autotune")
        return _tuning_dict(best), [(_tuning_dict(r.tuning), float(r.ms))
                                    for r in res[:nres.value]]

    # -- tallies
    def prf_dims(self, geom: int = 0):
        x, y = C.c_int32(), C.c_int32()
        _check(lib().fd2_prf_dims(self._ctx, geom, C.byref(x), C.byref(y)), "prf_dims")
        return x.value, y.value

    def prf(self, geom: int = 0) -> np.ndarray:
        xd, yd = self.prf_dims(geom)
        out = np.zeros((xd, yd), np.uint64)
        _check(lib().fd2_get_prf(self._ctx, geom, out.ctypes.data), "get_prf")
        return out

    def hist_counts(self, geom: int = 0) -> np.ndarray:
        out = np.zeros(max(1, int(self.params[geom]["num_hist"])), np.uint32)
        _check(lib().fd2_get_hist_counts(self._ctx, geom, out.ctypes.data), "get_hist")
        return out[: int(self.params[geom]["num_hist"])]

    def phs(self, geom: int = 0) -> np.ndarray:
        out = np.zeros(int(self.params[geom]["num_bins"]), np.int32)
        _check(lib().fd2_get_phs(self._ctx, geom, out.ctypes.data), "get_phs")
        return out

    def counters(self, geom: int = 0) -> dict:
        out = np.zeros(len(COUNTER_NAMES), np.uint64)
        _check(lib().fd2_get_counters(self._ctx, geom, out.ctypes.data), "get_counters")
        return dict(zip(COUNTER_NAMES, (int(v) for v in out)))

    def reset(self) -> None:
        _check(lib().fd2_reset(self._ctx), "fd2_reset")

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        _check(lib().fd2_last_kernel_ms(self._ctx, C.byref(ms)), "last_kernel_ms")
        return float(ms.value)


_mlp_sigs_done = False


def _mlp_lib() -> C.CDLL:
    global _mlp_sigs_done
    L = lib()
    if not _mlp_sigs_done:
        P = C.c_void_p
        sig = {
            "fd2_mlp_num_params": ([], C.c_size_t),
            "fd2_mlp_create": ([C.c_int, C.c_uint32, C.POINTER(P)], C.c_int),
            "fd2_mlp_destroy": ([P], None),
            "fd2_mlp_set_params": ([P, P, C.c_size_t], C.c_int),
            "fd2_mlp_get_params": ([P, P, C.c_size_t], C.c_int),
            "fd2_mlp_forward": ([P, P, C.c_int64, P], C.c_int),
            "fd2_mlp_logits": ([P, P, C.c_int64, P, C.c_int], C.c_int),
            "fd2_mlp_gradients": ([P, P, P, C.c_int, P, C.POINTER(C.c_float)], C.c_int),
            "fd2_mlp_set_data": ([P, P, P, C.c_int64], C.c_int),
            "fd2_mlp_train": ([P, C.c_int, C.c_int, C.c_float, C.c_uint32, P,
                               C.POINTER(C.c_float)], C.c_int),
            "fd2_mlp_time_forward": ([P, C.c_int, C.c_int, P], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _mlp_sigs_done = True
    return L


class Mlp:
    """The surrogate network on the GPU (BF16 matrix cores, FP32 master
    weights): 5 features -> [detection probability, 64 radial bins]."""

    FEATURES = 5
    TARGETS = 65

    def __init__(self, device: int = 0, seed: int = 0):
        L = _mlp_lib()
        h = C.c_void_p()
        _check(L.fd2_mlp_create(device, seed & 0xFFFFFFFF, C.byref(h)), "fd2_mlp_create")
        self._h = h
        self.n_params = int(L.fd2_mlp_num_params())

    def close(self):
        if getattr(self, "_h", None):
            _mlp_lib().fd2_mlp_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        _check(_mlp_lib().fd2_mlp_get_params(self._h, out.ctypes.data, self.n_params), "get_params")
        return out

    def set_params(self, p) -> None:
        a = np.ascontiguousarray(p, np.float32)
        _check(_mlp_lib().fd2_mlp_set_params(self._h, a.ctypes.data, a.size), "set_params")

    def forward(self, x) -> np.ndarray:
        a = np.ascontiguousarray(x, np.float32)
        out = np.zeros((len(a), self.TARGETS), np.float32)
        _check(_mlp_lib().fd2_mlp_forward(self._h, a.ctypes.data, len(a), out.ctypes.data),
               "mlp_forward")
        return out

    def logits(self, x, layered: bool = False) -> np.ndarray:
        """Raw output logits [n, 65] (fd2_mlp_logits): the fused inference
        forward, or with layered=True the per-layer GEMMs training uses."""
        a = np.ascontiguousarray(x, np.float32)
        out = np.zeros((len(a), self.TARGETS), np.float32)
        _check(_mlp_lib().fd2_mlp_logits(self._h, a.ctypes.data, len(a), out.ctypes.data,
                                         1 if layered else 0), "mlp_logits")
        return out

    def gradients(self, x, y):
        a = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(y, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        _check(_mlp_lib().fd2_mlp_gradients(self._h, a.ctypes.data, t.ctypes.data, len(a),
                                            g.ctypes.data, C.byref(loss)), "mlp_gradients")
        return float(loss.value), g

    def set_data(self, x, y) -> None:
        self._data_x = np.ascontiguousarray(x, np.float32)
        self._data_y = np.ascontiguousarray(y, np.float32)
        _check(_mlp_lib().fd2_mlp_set_data(self._h, self._data_x.ctypes.data,
                                           self._data_y.ctypes.data, len(self._data_x)),
               "mlp_set_data")

    def time_forward(self, batch: int = 65536, reps: int = 20) -> float:
        """Mean GPU ms of the fused inference forward on a device-resident
        batch (fd2_mlp_time_forward)."""
        ms = C.c_float()
        _check(_mlp_lib().fd2_mlp_time_forward(self._h, batch, reps, C.byref(ms)),
               "mlp_time_forward")
        return float(ms.value)

    def train(self, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0):
        losses = np.zeros(steps, np.float32)
        ms = C.c_float()
        _check(_mlp_lib().fd2_mlp_train(self._h, steps, batch, lr, seed & 0xFFFFFFFF,
                                        losses.ctypes.data, C.byref(ms)), "mlp_train")
        return losses, float(ms.value)


class SurrogateCfg(C.Structure):
    _fields_ = [("surface_band", C.c_float), ("audit_fraction", C.c_float),
                ("h_min", C.c_float), ("h_max", C.c_float),
                ("r_min", C.c_float), ("r_max", C.c_float),
                ("refl_min", C.c_float), ("refl_max", C.c_float),
                ("seed", C.c_uint32), ("center_band", C.c_float), ("det_extent", C.c_float),
                ("reserved", C.c_int32)]


class SurrogateStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("n_events", "n_mc", "n_surrogate", "n_audit",
                                         "photons_mc", "photons_surrogate")] + \
               [(n, C.c_double) for n in ("audit_detected_mc", "audit_detected_pred",
                                          "audit_radial_l1", "ms_total", "ms_mc",
                                          "ms_surrogate")]


SURROGATE_DEFAULTS = dict(surface_band=5.0, audit_fraction=0.01, h_min=100.0, h_max=500.0,
                          r_min=2.55, r_max=5.0, refl_min=0.5, refl_max=0.95, seed=0,
                          center_band=25.0, det_extent=1001.0)


def surrogate_submit(tr: "Transport", net: "Mlp", events, **cfg) -> dict:
    """Transport `events` with the surrogate where it applies and the plain
    Monte Carlo elsewhere (surface band, out-of-range geometry, audit
    slice); tallies accumulate in `tr`.  Returns the routing and audit
    statistics."""
    L = lib()
    f = L.fd2_surrogate_submit
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                  C.POINTER(SurrogateCfg), C.POINTER(SurrogateStats)]
    f.restype = C.c_int
    c = dict(SURROGATE_DEFAULTS)
    c.update(cfg)
    sc = SurrogateCfg(**c)
    st = SurrogateStats()
    ev = as_events(events)
    _check(f(tr._ctx, net._h, ev.ctypes.data, len(ev), C.byref(sc), C.byref(st)),
           "fd2_surrogate_submit")
    return {k: getattr(st, k) for k, _ in SurrogateStats._fields_}


def submit_radial(tr: "Transport", events) -> np.ndarray:
    """Transport events and return each event's 64-bin log-radial histogram
    of detection offsets (u32 [n, 64])."""
    L = lib()
    f = L.fd2_submit_radial
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    f.restype = C.c_int
    ev = as_events(events)
    out = np.zeros((len(ev), 64), np.uint32)
    _check(f(tr._ctx, ev.ctypes.data, len(ev), out.ctypes.data), "fd2_submit_radial")
    return out


# ---- file formats (fd2_io.cpp) ---------------------------------------------------------

def from_params(p: Params) -> dict:
    """fd2_params -> events.params()-style dictionary."""
    return {k: getattr(p, v) for k, v in _PARAM_MAP.items()}


def _io_lib() -> C.CDLL:
    L = lib()
    if not getattr(L, "_io_ready", False):
        P = C.c_void_p
        sig = {
            "fd2_read_input": ([C.c_char_p, C.POINTER(Params), P], C.c_int),
            "fd2_output_names": ([P, C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t], C.c_int),
            "fd2_write_prf": ([C.c_char_p, C.POINTER(Params), P, C.c_int32, C.c_int32,
                               C.c_double], C.c_int),
            "fd2_write_phs": ([C.c_char_p, P, C.c_int32], C.c_int),
            "fd2_write_stats": ([C.c_char_p, P, C.c_int, C.c_double], C.c_int),
            "fd2_read_ede": ([C.c_char_p, P, C.c_int64, C.POINTER(C.c_int64)], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        L._io_ready = True
    return L


def read_input(path: str):
    """hybridMANTIS_input.in -> (parameter dictionary, the ten file-name
    tokens)."""
    p = Params()
    tok = (C.c_char * 160)()
    _check(_io_lib().fd2_read_input(os.fsencode(path), C.byref(p), tok), f"read {path}")
    raw = bytes(tok)
    tokens = [raw[16 * i:16 * i + 16].split(b"\0")[0].decode() for i in range(10)]
    return from_params(p), tokens


def output_names(tokens) -> tuple:
    """(myimage_..., detected_...) file names from the ten tokens."""
    tok = (C.c_char * 160)()
    for i, t in enumerate(tokens):
        b = t.encode()[:15]
        C.memmove(C.addressof(tok) + 16 * i, b, len(b))
    a, b = C.create_string_buffer(512), C.create_string_buffer(512)
    _check(_io_lib().fd2_output_names(tok, a, 512, b, 512), "output names")
    return a.value.decode(), b.value.decode()


def write_prf(path: str, params: dict, prf, n_hist: float) -> None:
    prf = np.ascontiguousarray(prf, np.uint64)
    p = to_params(params)
    _check(_io_lib().fd2_write_prf(os.fsencode(path), C.byref(p), prf.ctypes.data,
                                   prf.shape[0], prf.shape[1], float(n_hist)), f"write {path}")


def write_phs(path: str, phs) -> None:
    phs = np.ascontiguousarray(phs, np.int32)
    _check(_io_lib().fd2_write_phs(os.fsencode(path), phs.ctypes.data, len(phs)),
           f"write {path}")


def write_stats(path, stats, rungpu: int = 1, penoptical_s: float = 0.0) -> None:
    """The reference's console optical-statistics block (fd2_write_stats):
    stats = the 8 optstats_ doubles; path None writes to standard output."""
    st = np.ascontiguousarray(stats, np.float64)
    if st.shape != (8,):
        raise ValueError("stats: the 8 optstats_ values")
    _check(_io_lib().fd2_write_stats(None if path is None else os.fsencode(path),
                                     st.ctypes.data, rungpu, float(penoptical_s)),
           "write stats")


def read_ede(path: str, max_events: int) -> np.ndarray:
    """EDE text records -> events (um, eV, history; N = 0)."""
    ev = np.zeros(max_events, EVENT_DTYPE)
    n = C.c_int64(0)
    _check(_io_lib().fd2_read_ede(os.fsencode(path), ev.ctypes.data, max_events, C.byref(n)),
           f"read {path}")
    return ev[:n.value]
