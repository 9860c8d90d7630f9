"""ctypes view of libfd2_fortran.so: the reference's four optical entry
points and their COMMON blocks (include/fd2_fortran.h).

The COMMON blocks are the ones a gfortran penEasy shares with the drivers
(main/kernel_cuda_c_ver1_0.cu:86-234).  From Python they are filled and
read through `commons()`; the entry points take pointers as Fortran passes
them.  Calling order mirrors the tally
(main/hybridMANTIS_tallyEnergyDepositionEvents.f:255-266,
main/hybridMANTIS_penEasy.f:434-436): gpuoptical(99, n) once, then
gpuoptical(100, n) per full buffer and gpuoptical(101, n) for the last one.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import fd2

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FD2_FORTRAN_LIBRARY", os.path.join(HERE, "libfd2_fortran.so"))
BUFSIZE = 2304000  # maxvalopt, tally .f:55-56


class Optical(ctypes.Structure):
    _fields_ = [("xbufopt", ctypes.c_double * BUFSIZE), ("ybufopt", ctypes.c_double * BUFSIZE),
                ("zbufopt", ctypes.c_double * BUFSIZE), ("debufopt", ctypes.c_double * BUFSIZE),
                ("nbufopt", ctypes.c_int32 * BUFSIZE), ("myctropt", ctypes.c_int32),
                ("cpu_num_real", ctypes.c_int32)]


class OptStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("glgen", "gldetect", "glabstop", "glabsbulk",
                                               "gllost", "gloutofcol", "gltheta1", "glgputime")]


class OutputDetPrim(ctypes.Structure):
    _fields_ = [("gldetprimary", ctypes.c_int32 * 1000)]


class OutputImage(ctypes.Structure):
    _fields_ = [("newimageopt", (ctypes.c_uint64 * 501) * 501),
                ("tempimageopt", (ctypes.c_uint64 * 501) * 501)]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        fd2.lib()  # libfd2.so first, from its own path
        if not os.path.exists(LIB_PATH):
            raise fd2.Fd2Error(f"{LIB_PATH} is not built (python -m hybridmantis_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        i32p, f64p = ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)
        L.gpuoptical_.argtypes = [i32p, i32p]
        L.cpuoptical_.argtypes = [i32p]
        L.gpuoptlb_.argtypes = [f64p, i32p]
        L.cpuoptlb_.argtypes = [f64p, i32p]
        for f in (L.gpuoptical_, L.cpuoptical_, L.gpuoptlb_, L.cpuoptlb_):
            f.restype = None
        L.fd2f_poisson.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint32,
                                   ctypes.c_int64, ctypes.c_void_p]
        L.fd2f_poisson.restype = ctypes.c_int
        _lib = L
    return _lib


class Commons:
    """The COMMON blocks, with numpy views of the arrays."""

    def __init__(self, L):
        self.inputargs = fd2.Params.in_dll(L, "inputargs_")
        self.optical = Optical.in_dll(L, "optical_")
        self.optstats = OptStats.in_dll(L, "optstats_")
        self.outputdetprim = OutputDetPrim.in_dll(L, "outputdetprim_")
        self.outputimage = OutputImage.in_dll(L, "outputimage_")
        o = self.optical
        self.x = np.ctypeslib.as_array(o.xbufopt)
        self.y = np.ctypeslib.as_array(o.ybufopt)
        self.z = np.ctypeslib.as_array(o.zbufopt)
        self.de = np.ctypeslib.as_array(o.debufopt)
        self.hist = np.ctypeslib.as_array(o.nbufopt)
        self.phs = np.ctypeslib.as_array(self.outputdetprim.gldetprimary)
        self.image = np.ctypeslib.as_array(self.outputimage.newimageopt)

    def set_params(self, p: dict) -> None:
        """Set /inputargs/ from an events.params() dictionary."""
        src = fd2.to_params(p)
        ctypes.memmove(ctypes.addressof(self.inputargs), ctypes.addressof(src),
                       ctypes.sizeof(fd2.Params))

    def fill(self, events, first: int = 0) -> None:
        """Copy events (x, y, z in um, E in eV, hist) into the buffer at
        `first` in the buffer's units (cm)."""
        n = len(events)
        sl = slice(first, first + n)
        self.x[sl] = np.asarray(events["x"], np.float64) / 1e4
        self.y[sl] = np.asarray(events["y"], np.float64) / 1e4
        self.z[sl] = np.asarray(events["z"], np.float64) / 1e4
        self.de[sl] = events["E"]
        self.hist[sl] = events["hist"]

    def clear_outputs(self) -> None:
        self.image[:] = 0
        self.phs[:] = 0
        ctypes.memset(ctypes.addressof(self.optstats), 0, ctypes.sizeof(OptStats))

    def stats(self) -> dict:
        return {n: getattr(self.optstats, n) for n, _ in OptStats._fields_}


_commons = None


def commons() -> Commons:
    global _commons
    if _commons is None:
        _commons = Commons(lib())
    return _commons


def gpuoptical(penctr: int, myfact: int) -> None:
    a, b = ctypes.c_int32(penctr), ctypes.c_int32(myfact)
    lib().gpuoptical_(ctypes.byref(a), ctypes.byref(b))


def cpuoptical(gpusize: int) -> None:
    a = ctypes.c_int32(gpusize)
    lib().cpuoptical_(ctypes.byref(a))


def gpuoptlb(lbbuf: int) -> float:
    t, n = ctypes.c_double(0.0), ctypes.c_int32(lbbuf)
    lib().gpuoptlb_(ctypes.byref(t), ctypes.byref(n))
    return t.value


def cpuoptlb(cpubufsize: int) -> float:
    t, n = ctypes.c_double(0.0), ctypes.c_int32(cpubufsize)
    lib().cpuoptlb_(ctypes.byref(t), ctypes.byref(n))
    return t.value


def poisson(mu, seed: int, first: int = 0) -> np.ndarray:
    """The entry points' photon-count draws for means `mu` at ordinals
    first, first + 1, ..."""
    mu = np.ascontiguousarray(mu, np.float64)
    out = np.zeros(len(mu), np.int32)
    rc = lib().fd2f_poisson(mu.ctypes.data, len(mu), seed & 0xFFFFFFFF, first, out.ctypes.data)
    if rc:
        raise fd2.Fd2Error(f"fd2f_poisson: {rc}")
    return out


def photon_counts(events, yield_per_ev: float, seed: int, first: int = 0) -> np.ndarray:
    """N of events as the entry points draw them (yield read as float)."""
    mu = np.asarray(events["E"], np.float64) * float(np.float32(yield_per_ev))
    return poisson(mu, seed, first)
