"""Write the golden fixtures of tests/golden/ from the reference's own C
transport branch.

Runs only where the reference tree is present (it builds oracle/_ref with
`make -C oracle ref`): the reference's main/kernel_cuda_c_ver1_0.cu compiled
as C, driven by oracle/ref_harness.c the way main/hybridMANTIS_c_ver1_0.c
drives it, with a fixed seed_input.  Each case stores its parameters, its
events and what the reference produced: every photon's fate and PRF pixel,
the PRF, the per-history detection counts and the counters.  Also stored:
RANECU streams, init_PRNG seeds and abMODm products from the reference.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from hybridmantis_amd import events as E  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

SEED = 271828182


def example_params(**kw):
    """main/hybridMANTIS_input.in:8-31, with a few histories."""
    p = dict(num_hist=40, min_ph=0, max_ph=1400, num_bins=140, xdet=909.0,
             ydet=909.0, H=150.0, R=5.1, n_C=1.8, n_IC=1.0, top_absfrac=0.1,
             bulk=1e-4, beta=0.2, d_min=1.0, d_max=280.0, lbx=0.0, lby=0.0,
             ubx=909.0, uby=909.0, yield_per_ev=0.055, pixel=9,
             sensor_refl=0.25, rungpu=1, machine=1)
    p.update(kw)
    return p


def cases():
    out = []
    p, ev = E.workload("mono25", scale=0.02)
    out.append(("mono25_small", p, ev, SEED))
    p2 = example_params()
    ev2 = E.generate(p2["num_hist"], p2["H"], 25.0, 150.0, 11, beam=(454.5, 454.5),
                     xdet=909.0, ydet=909.0)
    out.append(("example_input", p2, ev2, 123456789))
    p3 = E.params(num_hist=60, beta=0.0, bulk=0.0, top_absfrac=0.0)
    ev3 = E.generate(60, p3["H"], 20.0, 150.0, 12)
    out.append(("smooth_nobulk_notop", p3, ev3, SEED))
    p4 = E.params(num_hist=60, n_C=1.4, n_IC=1.0, sensor_refl=0.5, top_absfrac=0.3)
    ev4 = E.generate(60, p4["H"], 20.0, 150.0, 13)
    out.append(("low_index_sensor_refl", p4, ev4, SEED))
    p5 = E.params(num_hist=80, xdet=60.0, ydet=60.0, lbx=0.0, lby=0.0, ubx=60.0,
                  uby=60.0, pixel=2)
    ev5 = E.generate(80, p5["H"], 10.0, 150.0, 14, beam=(3.0, 57.0), xdet=60.0,
                     ydet=60.0)
    out.append(("detector_edges", p5, ev5, SEED))
    p6 = E.params(num_hist=6, H=500.0, R=5.0, max_ph=5000, num_bins=1000)
    ev6 = E.generate(6, 500.0, 50.0, float(E.attenuation_length_um(50.0)), 15)
    out.append(("thick_50kev", p6, ev6, SEED))
    # the batched workload's two geometries no other case has: (H, R) =
    # (100, 2.55) and (300, 4.0) um, polychromatic 15-90 keV
    for name, H, R, nh, sd in (("thin100_poly", 100.0, 2.55, 30, 16),
                               ("mid300_r4_poly", 300.0, 4.0, 10, 17)):
        p7 = E.params(num_hist=nh, H=H, R=R, max_ph=5000, num_bins=1000)
        e7 = np.random.default_rng(sd).integers(15, 91, nh).astype(np.float64)
        ev7 = E.generate(nh, H, e7, E.attenuation_length_um(e7), sd)
        out.append((name, p7, ev7, SEED))
    return out


def main() -> None:
    po.build(ref=True)
    if not po.ref_available():
        raise SystemExit("reference harness not built (reference tree absent)")
    manifest = {}
    for name, p, ev, seed in cases():
        nh = p["num_hist"]
        r = po.run_ref(p, ev, nh, seed=seed, record=True)
        nz = np.flatnonzero(r.prf)
        fname = f"ref_{name}.npz"
        np.savez_compressed(
            os.path.join(HERE, fname),
            events=np.asarray(ev)[["x", "y", "z", "E", "hist", "N"]].copy()
            if ev.dtype.names and "geom" in ev.dtype.names else ev,
            fate=r.photons["fate"].astype(np.int8),
            pix=r.photons["pix"].astype(np.int32),
            prf_idx=nz.astype(np.int32),
            prf_val=r.prf.ravel()[nz].astype(np.uint64),
            prf_shape=np.array(r.prf.shape, np.int32),
            hist=r.hist_count.astype(np.uint32),
            counters=r.counters.astype(np.uint64))
        manifest[name] = dict(file=fname, params=p, seed=seed,
                              photons=int(ev["N"].sum()), events=int(len(ev)))
        print(name, len(ev), int(ev["N"].sum()), r.counters[:7].tolist(), flush=True)

    # RNG pieces of the reference
    seeds = [(SEED, b) for b in (0, 1, 2, 1000, 2303999)] + \
            [(123456789, 7), (1, 0), (2147483398, 5)]
    init = [(s, b) + po.init_prng(b, 50000, s, use_ref=True) for s, b in seeds]
    streams = np.stack([po.ranecu_stream(po.init_prng(b, 50000, s, use_ref=True),
                                         1000, use_ref=True) for s, b in seeds])
    rng = np.random.default_rng(5)
    trip = []
    for m in (2147483563, 2147483399):
        for _ in range(200):
            a, s_ = int(rng.integers(1, m)), int(rng.integers(1, m))
            trip.append((m, a, s_, int(po.ref().fdref_abmodm(m, a, s_))))
    np.savez_compressed(os.path.join(HERE, "ref_rng.npz"),
                        init=np.array(init, np.int64), streams=streams,
                        abmodm=np.array(trip, np.int64))
    manifest["_rng"] = dict(file="ref_rng.npz", seeds=seeds)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
