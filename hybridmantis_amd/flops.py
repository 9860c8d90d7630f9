"""Algorithmic FP32 work of the transport, per optical photon.

The transport's cost is a sum over primitives k (an in-column step, a
RoughSurface perturbation, a Fresnel evaluation, ...) of c_k * F_k:

  F_k  FP32 operations of primitive k in the reference's formulation of it
       (main/kernel_cuda_c_ver1_0.cu, C branches: the angle of incidence by
       acos, Snell's law by asin, Fresnel from sines/cosines of the angle
       sums, a logarithm per absorption draw): +, -, *, /, sqrt and each
       elementary function (sin, cos, log, asin, acos) count 1, a fused
       multiply-add counts 2; comparisons, selects and integer work count 0
       except the uniform's int->float scaling (1 per draw).  The kernel
       gets some primitives with fewer operations (Fresnel from the sines
       and cosines of the angles, the logarithm only where it can decide the
       step); the count stays the algorithm's, so achieved FLOP/s compares
       across kernel versions.
  c_k  mean number of times primitive k runs per photon, measured by the
       oracle's work counters on the same workload (they count exactly the
       same primitives in the same control flow).

flops_per_photon() combines them; bench.py reports photons/s x this value
against the MI355X FP32 vector peak.
"""
from __future__ import annotations

# name -> FP32 ops per occurrence (the oracle's work-counter order)
F = {
    "draw": 1,        # (x >> 8) + 1 -> float, * 2^-24
    "init": 16,       # 2u-1, sqrt(1-r^2), u*2pi, sincos, 2 mul, |d| (renorm)
    "dist": 39,       # ray/cylinder quadratic, 2 roots, plane, move, |p-c|
    "nudge": 10,      # step back along -d, new radius
    "abslog": 2,      # log(u) * (-1/mu)
    "sensor": 2,      # plane test, reflectivity draw compare
    "top": 12,        # -|2u-1|, sqrt(1-z^2), u*2pi, sincos, 2 mul, tests
    "cct": 4,         # crosstalk probability (piecewise linear in z)
    "wallnorm": 8,    # inward normal (Xc-x, Yc-y)/|.|
    "theta": 6,       # -d.n and acos
    "fresnel": 33,    # asin(n sin t1), 2 sincos, tan ratios, Pr, Pt, normalise
    "reflect": 20,    # d + 2(-d.n)n, renormalise
    "refract": 30,    # sincos(t2), sin t1, -n cos t2 - (s2/s1)(d + c n), renormalise
    "rough": 35,      # perturbation vector, two normalisations
    "newnormal": 12,  # xy projection + normalise, -d.n, acos
    "flight": 12,     # U(dmin,dmax) flight, distances to top/bottom
    "reflchk": 6,     # acos of reflected d . smooth normal
}

WORK_NAMES = list(F)


def flops_per_photon(work_counts, photons: float) -> float:
    """sum_k c_k F_k from raw work counters (oracle WORK_NAMES order)."""
    tot = 0.0
    for name, count in zip(WORK_NAMES, work_counts):
        tot += F[name] * float(count)
    return tot / float(photons)


def per_photon(work_counts, photons: float) -> dict:
    return {n: float(c) / float(photons) for n, c in zip(WORK_NAMES, work_counts)}
