"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and, where it has
been built in this container, the reference-branch harness
(oracle/_ref/libfdref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FDO_LIBRARY selects another build of the oracle (liboracle_O0.so, the
# unoptimised build used for timing); spawned workers inherit it
LIB_PATH = os.environ.get("FDO_LIBRARY") or os.path.join(HERE, "liboracle.so")
LIB_O0_PATH = os.path.join(HERE, "liboracle_O0.so")
REF_PATH = os.path.join(HERE, "_ref", "libfdref.so")

RNG_RANECU = 0
RNG_COUNTER = 1
MATH_FAITHFUL = 0
MATH_DEVICE = 1
MATH_EMULATED = 2

FATE_NONE, FATE_DETECT, FATE_DETECT_OUT, FATE_ABS_TOP, FATE_ABS_BULK, \
    FATE_LOST, FATE_OUTOFCOL, FATE_THETA1 = range(8)

COUNTER_NAMES = ["generated", "detected", "abs_top", "abs_bulk", "lost",
                 "outofcol", "theta1", "regenerated"]
WORK_NAMES = ["draw", "init", "dist", "nudge", "abslog", "sensor", "top",
              "cct", "wallnorm", "theta", "fresnel", "reflect", "refract",
              "rough", "newnormal", "flight", "reflchk"]


class Geom(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "xdet", "ydet", "H", "R", "n_C", "n_IC", "top_absfrac", "bulk", "beta",
        "d_min", "d_max", "lbx", "lby", "ubx", "uby", "sensor_refl")] + [
        ("pixel", C.c_int), ("xdim", C.c_int), ("ydim", C.c_int)]


EVENT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("E", "<f4"),
                        ("hist", "<i4"), ("N", "<i4"), ("col_dx", "<f4")])
PHOTON_DTYPE = np.dtype([("fate", "<i4"), ("pix", "<i4"), ("x", "<f4"),
                         ("y", "<f4")])


class Tally(C.Structure):
    _fields_ = [("prf", C.POINTER(C.c_uint64)),
                ("hist_count", C.POINTER(C.c_uint32)),
                ("num_hist", C.c_int32),
                ("counters", C.c_uint64 * 8),
                ("work", C.c_uint64 * 17),
                ("photons", C.c_void_p)]


def build(ref: bool = False) -> None:
    """Compile the oracle (and, if asked and the reference tree exists, the
    reference-branch harness) with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so", "liboracle_O0.so"], check=True)
    if ref and os.path.exists("/root/reference/main/kernel_cuda_c_ver1_0.cu"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.fdo_run.argtypes = [C.POINTER(Geom), C.c_void_p, C.c_int64,
                                 C.c_int64, C.c_int, C.c_int, C.c_uint32,
                                 C.POINTER(Tally)]
        _lib.fdo_run.restype = C.c_int
        _lib.fdo_ranecu.argtypes = [C.POINTER(C.c_int32)]
        _lib.fdo_ranecu.restype = C.c_float
        _lib.fdo_init_prng.argtypes = [C.c_int, C.c_int, C.c_int,
                                       C.POINTER(C.c_int32)]
        _lib.fdo_abmodm.argtypes = [C.c_int32, C.c_int32, C.c_int32]
        _lib.fdo_abmodm.restype = C.c_int32
        _lib.fdo_philox4x32_10.argtypes = [C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_uint32)]
        _lib.fdo_phs.argtypes = [C.POINTER(C.c_uint32), C.c_int32, C.c_int,
                                 C.c_int, C.c_int, C.POINTER(C.c_int32)]
        _lib.fdo_phs.restype = C.c_int
        _lib.fdo_retrace_step_f64.argtypes = [
            C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, C.c_float,
            C.POINTER(Geom), C.c_int64, C.POINTER(C.c_double)]
        _lib.fdo_retrace_step_f64.restype = C.c_int
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH)
        _ref.fdref_run.argtypes = [C.POINTER(Geom), C.c_void_p, C.c_int64,
                                   C.c_int64, C.c_uint32,
                                   C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint32), C.c_int32,
                                   C.POINTER(C.c_uint64), C.c_void_p, C.c_int]
        _ref.fdref_run.restype = C.c_int
        _ref.fdref_ranecu.argtypes = [C.POINTER(C.c_int32)]
        _ref.fdref_ranecu.restype = C.c_float
        _ref.fdref_init_prng.argtypes = [C.c_int, C.c_int, C.c_int,
                                         C.POINTER(C.c_int32)]
        _ref.fdref_abmodm.argtypes = [C.c_int32, C.c_int32, C.c_int32]
        _ref.fdref_abmodm.restype = C.c_int32
    return _ref


def make_geom(p: dict) -> Geom:
    """Geom from a parameter dict with the reference's input names
    (hybridMANTIS_input.in order); floats are rounded to float32 the way the
    reference drivers assign the double COMMON values to float locals."""
    g = Geom()
    for k in ("xdet", "ydet", "H", "R", "n_C", "n_IC", "top_absfrac", "bulk",
              "beta", "d_min", "d_max", "lbx", "lby", "ubx", "uby",
              "sensor_refl"):
        setattr(g, k, float(np.float32(p[k])))
    g.pixel = int(p["pixel"])
    lbx, ubx = np.float32(p["lbx"]), np.float32(p["ubx"])
    lby, uby = np.float32(p["lby"]), np.float32(p["uby"])
    # xdim = ceil((ub-lb)/pixel) in float as the drivers compute it
    g.xdim = int(np.ceil(np.float32(ubx - lbx) / np.float32(g.pixel)))
    g.ydim = int(np.ceil(np.float32(uby - lby) / np.float32(g.pixel)))
    return g


def _as_events(ev) -> np.ndarray:
    a = np.ascontiguousarray(ev)
    if a.dtype != EVENT_DTYPE:
        out = np.zeros(len(a), EVENT_DTYPE)
        for k in EVENT_DTYPE.names:
            if k in a.dtype.names:
                out[k] = a[k]
        a = out
    return a


class Result:
    def __init__(self, prf, hist_count, counters, work=None, photons=None):
        self.prf = prf
        self.hist_count = hist_count
        self.counters = counters
        self.work = work
        self.photons = photons

    def counter(self, name):
        return int(self.counters[COUNTER_NAMES.index(name)])


def run(p: dict, events, num_hist: int, rng: int = RNG_COUNTER,
        seed: int = 271828182, event_base: int = 0,
        record: bool = False, math: int = MATH_FAITHFUL, rezero: bool = False) -> Result:
    """Transport `events` (EVENT_DTYPE) through the oracle.  rezero repeats
    the reference C driver's two per-event image clears (timing only)."""
    g = make_geom(p)
    ev = _as_events(events)
    prf = np.zeros((g.xdim, g.ydim), np.uint64)
    hc = np.zeros(max(num_hist, 1), np.uint32)
    t = Tally()
    t.prf = prf.ctypes.data_as(C.POINTER(C.c_uint64))
    t.hist_count = hc.ctypes.data_as(C.POINTER(C.c_uint32))
    t.num_hist = num_hist
    photons = None
    if record:
        photons = np.zeros(int(np.maximum(ev["N"], 0).sum()), PHOTON_DTYPE)
        t.photons = photons.ctypes.data
    lib().fdo_set_image_rezero(int(rezero))
    rc = lib().fdo_run(C.byref(g), ev.ctypes.data, len(ev), event_base, rng,
                       math, seed & 0xFFFFFFFF, C.byref(t))
    lib().fdo_set_image_rezero(0)
    if rc != 0:
        raise RuntimeError(f"fdo_run failed ({rc})")
    return Result(prf, hc[:num_hist], np.array(t.counters[:], np.uint64),
                  np.array(t.work[:], np.uint64), photons)


def run_ref(p: dict, events, num_hist: int, seed: int = 271828182,
            event_base: int = 0, record: bool = False,
            rezero: bool = False) -> Result:
    """Transport `events` through the reference's own C branch (RANECU)."""
    g = make_geom(p)
    ev = _as_events(events)
    prf = np.zeros((g.xdim, g.ydim), np.uint64)
    hc = np.zeros(max(num_hist, 1), np.uint32)
    counters = np.zeros(8, np.uint64)
    photons = None
    pptr = None
    if record:
        photons = np.zeros(int(np.maximum(ev["N"], 0).sum()), PHOTON_DTYPE)
        pptr = photons.ctypes.data
    rc = ref().fdref_run(C.byref(g), ev.ctypes.data, len(ev), event_base,
                         seed & 0xFFFFFFFF,
                         prf.ctypes.data_as(C.POINTER(C.c_uint64)),
                         hc.ctypes.data_as(C.POINTER(C.c_uint32)), num_hist,
                         counters.ctypes.data_as(C.POINTER(C.c_uint64)), pptr,
                         int(rezero))
    if rc != 0:
        raise RuntimeError(f"fdref_run failed ({rc})")
    return Result(prf, hc[:num_hist], counters, None, photons)


def phs(hist_count, min_ph: int, max_ph: int, num_bins: int) -> np.ndarray:
    hc = np.ascontiguousarray(hist_count, np.uint32)
    out = np.zeros(num_bins, np.int32)
    rc = lib().fdo_phs(hc.ctypes.data_as(C.POINTER(C.c_uint32)), len(hc),
                       min_ph, max_ph, num_bins,
                       out.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc != 0:
        raise ValueError("PHS bin size is zero")
    return out


def ranecu_stream(seed2, n: int, use_ref: bool = False) -> np.ndarray:
    s = (C.c_int32 * 2)(*seed2)
    f = ref().fdref_ranecu if use_ref else lib().fdo_ranecu
    return np.array([f(s) for _ in range(n)], np.float32)


def init_prng(batch: int, hpt: int, seed_input: int, use_ref: bool = False):
    s = (C.c_int32 * 2)()
    (ref().fdref_init_prng if use_ref else lib().fdo_init_prng)(
        batch, hpt, seed_input, s)
    return int(s[0]), int(s[1])


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().fdo_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def counter_stream(seed: int, event: int, photon: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float32)
    f = lib().fdo_counter_stream
    f.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, C.c_int64, C.c_void_p]
    f(seed & 0xFFFFFFFF, event & 0xFFFFFFFFFFFFFFFF, photon & 0xFFFFFFFF, n, out.ctypes.data)
    return out


MATH_FN = {"sin": 0, "cos": 1, "log": 2, "asin": 3, "acos": 4,
           "sin_turns": 5, "cos_turns": 6, "rsqrt": 7, "sqrt_unit": 8,
           "angle_over_157": 9,
           "acosf_libc": 10, "asinf_libc": 11, "ff_sin": 12, "ff_cos": 13, "ff_log": 14}


def math_eval(fn: str, x) -> np.ndarray:
    xs = np.ascontiguousarray(x, np.float32)
    y = np.zeros_like(xs)
    f = lib().fdo_math_eval
    f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
    f.restype = C.c_int
    if f(MATH_FN[fn], xs.ctypes.data, y.ctypes.data, len(xs)) != 0:
        raise ValueError(fn)
    return y


def fresnel_probs(t1: float, t2: float, math: int = MATH_FAITHFUL):
    pr, pt = C.c_float(), C.c_float()
    f = lib().fdo_fresnel_probs
    f.argtypes = [C.c_float, C.c_float, C.c_int, C.POINTER(C.c_float),
                  C.POINTER(C.c_float)]
    f(t1, t2, math, C.byref(pr), C.byref(pt))
    return pr.value, pt.value


def emit_dirs(seed: int, n: int, math: int = MATH_FAITHFUL) -> np.ndarray:
    out = np.zeros((n, 3), np.float32)
    f = lib().fdo_emit_dirs
    f.argtypes = [C.c_uint32, C.c_int, C.c_int64, C.c_void_p]
    f(seed & 0xFFFFFFFF, math, n, out.ctypes.data)
    return out


def rough_normals(seed: int, beta: float, n0, n: int,
                  math: int = MATH_FAITHFUL) -> np.ndarray:
    out = np.zeros((n, 3), np.float32)
    v = (C.c_float * 3)(*n0)
    f = lib().fdo_rough_normals
    f.argtypes = [C.c_uint32, C.c_int, C.c_float, C.c_void_p, C.c_int64, C.c_void_p]
    f(seed & 0xFFFFFFFF, math, beta, v, n, out.ctypes.data)
    return out


def retrace_step(pos, dcos, xc, yc, p: dict, forced: int = -1):
    """FP64 re-trace of one in-column step; forced = recorded decisions
    (fd2_step.decisions) to follow the FP32 run's branches."""
    g = make_geom(p)
    pi = (C.c_float * 3)(*[float(v) for v in pos])
    di = (C.c_float * 3)(*[float(v) for v in dcos])
    po = (C.c_double * 3)()
    code = lib().fdo_retrace_step_f64(pi, di, float(xc), float(yc), C.byref(g),
                                      int(forced), po)
    return code, np.array(po[:])


# ---------------------------------------------------------- multi-process

def _slice_worker(args):
    kind, p, ev, num_hist, rng, seed, base, math, rezero, record = args
    import time as _t
    t0 = _t.perf_counter()
    if kind == "ref":
        r = run_ref(p, ev, num_hist, seed=seed, event_base=base, rezero=rezero,
                    record=record)
        work = None
    else:
        r = run(p, ev, num_hist, rng=rng, seed=seed, event_base=base, math=math,
                record=record, rezero=rezero)
        work = r.work
    return (r.prf, r.hist_count, r.counters, work, _t.perf_counter() - t0, r.photons)


def run_parallel(p: dict, events, num_hist: int, workers: int = 8,
                 kind: str = "oracle", rng: int = RNG_COUNTER,
                 seed: int = 271828182, math: int = MATH_FAITHFUL,
                 rezero: bool = False, timeout: float = 600.0, record: bool = False,
                 event_base: int = 0):
    """Run contiguous event slices in separate processes (the transport
    keeps per-process state, like the reference C path) and sum the
    tallies; results equal a single-process run because streams are keyed
    by the global event number.  Returns (Result, per-worker seconds)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor

    ev = _as_events(events)
    n = len(ev)
    workers = max(1, min(workers, n))
    bounds = [n * i // workers for i in range(workers + 1)]
    jobs = [(kind, p, ev[bounds[i]:bounds[i + 1]].copy(), num_hist, rng, seed,
             event_base + bounds[i], math, rezero, record) for i in range(workers)]
    # fresh interpreters (spawn): the caller may hold a GPU context; a dead
    # worker raises BrokenProcessPool instead of hanging
    with ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
        futs = [ex.submit(_slice_worker, j) for j in jobs]
        outs = [f.result(timeout=timeout) for f in futs]
    prf = sum(o[0] for o in outs)
    hc = sum(o[1].astype(np.uint64) for o in outs).astype(np.uint32)
    cnt = sum(o[2] for o in outs)
    work = None if outs[0][3] is None else sum(o[3] for o in outs)
    photons = np.concatenate([o[5] for o in outs]) if record else None
    return Result(prf, hc, cnt, work, photons), [o[4] for o in outs]
