"""Synthetic energy-deposition events and detector parameter sets.

PENELOPE (the reference's x-ray/electron transport, which fills the
/optical/ event buffer in main/hybridMANTIS_tallyEnergyDepositionEvents.f:
176-180) is not distributed with the reference, so events are drawn from a
simple parametric model instead:

  * a history deposits nothing with probability `p_none`;
  * otherwise the interaction depth below the entrance (top, z = +H/2)
    surface is Exponential(mean = attenuation length) truncated to [0, H];
  * the history splits into 1 + Poisson(1.5) events with Dirichlet(1)
    energy fractions and N(0, 3 um) position offsets (clamped inside the
    layer);
  * each event's optical-photon count is round(Normal(yield*E, sqrt(yield*E)))
    with E in keV and yield 55 photons/keV (the reference samples
    Poisson(E*yield), main/hybridMANTIS_c_ver1_0.c:219-220; the normal
    approximation is the workload definition used here).

Events carry positions in micrometres (the drivers multiply the cm buffer by
1e4, main/hybridMANTIS_c_ver1_0.c:212-216), E in eV and the 1-based history
number.
"""
from __future__ import annotations

import numpy as np

EVENT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("E", "<f4"),
                        ("hist", "<i4"), ("N", "<i4"), ("geom", "<u4"),
                        ("col_dx", "<f4")])

# Reference parameter defaults (main/hybridMANTIS_input.in:8-31) adjusted to
# the workloads below: 1001 um detector with the beam at its centre, 501x501
# PRF at 1 um, ideal sensor (reflectivity 0) and top reflectivity 0.9.
BASE_PARAMS = dict(
    num_hist=2000, min_ph=0, max_ph=2000, num_bins=2000,
    xdet=1001.0, ydet=1001.0, H=150.0, R=2.55, n_C=1.8, n_IC=1.0,
    top_absfrac=0.1, bulk=1e-4, beta=0.2, d_min=1.0, d_max=280.0,
    lbx=250.0, lby=250.0, ubx=751.0, uby=751.0, yield_per_ev=0.055, pixel=1,
    sensor_refl=0.0, rungpu=1, machine=1)

BEAM_XY = (500.5, 500.5)

# the four (thickness, radius) geometries of the batched workload
BATCH_GEOMS = [(100.0, 2.55), (150.0, 2.55), (300.0, 4.0), (500.0, 5.0)]


def params(**kw) -> dict:
    p = dict(BASE_PARAMS)
    p.update(kw)
    return p


def attenuation_length_um(e_kev):
    """Linear 60 um at 15 keV to 400 um at 90 keV."""
    return 60.0 + (400.0 - 60.0) * (np.asarray(e_kev, np.float64) - 15.0) / 75.0


def generate(num_hist: int, H: float, energy_kev, atten_um, seed: int,
             beam=BEAM_XY, p_none: float = 0.2, yield_per_kev: float = 55.0,
             sigma_um: float = 3.0, xdet: float = 1001.0, ydet: float = 1001.0,
             geom: int = 0, hist_base: int = 0) -> np.ndarray:
    """Events of `num_hist` pencil-beam histories.

    energy_kev: scalar or per-history array; atten_um: scalar or per-history
    array of attenuation lengths.  Returns EVENT_DTYPE records ordered by
    history."""
    rng = np.random.Generator(np.random.PCG64(seed))
    e = np.broadcast_to(np.asarray(energy_kev, np.float64), (num_hist,))
    lam = np.broadcast_to(np.asarray(atten_um, np.float64), (num_hist,))
    deposits = rng.random(num_hist) >= p_none
    idx = np.nonzero(deposits)[0]
    # truncated exponential depth by inversion: d = -lam ln(1 - u(1-e^{-H/lam}))
    u = rng.random(len(idx))
    d = -lam[idx] * np.log1p(-u * (-np.expm1(-H / lam[idx])))
    k = 1 + rng.poisson(1.5, len(idx))
    nev = int(k.sum())
    owner = np.repeat(np.arange(len(idx)), k)
    w = rng.exponential(1.0, nev)
    wsum = np.zeros(len(idx))
    np.add.at(wsum, owner, w)
    frac = w / wsum[owner]
    off = rng.normal(0.0, sigma_um, (nev, 3))
    ev = np.zeros(nev, EVENT_DTYPE)
    margin = 1e-2
    ev["x"] = np.clip(beam[0] + off[:, 0], margin, xdet - margin)
    ev["y"] = np.clip(beam[1] + off[:, 1], margin, ydet - margin)
    ev["z"] = np.clip(H / 2.0 - d[owner] + off[:, 2], -H / 2.0 + margin,
                      H / 2.0 - margin)
    e_ev = e[idx][owner] * frac
    mu = yield_per_kev * e_ev
    ev["N"] = np.maximum(0, np.rint(rng.normal(mu, np.sqrt(mu)))).astype(np.int32)
    ev["E"] = (e_ev * 1000.0).astype(np.float32)
    ev["hist"] = (hist_base + idx[owner] + 1).astype(np.int32)
    ev["geom"] = geom
    return ev


def workload(name: str, scale: float = 1.0, seed: int | None = None):
    """(params, events) of a named workload.  `scale` shrinks the history
    count (for tests); the event model does not change.

    mono25     2,000 histories at 25 keV, H=150 um, R=2.55 um
    mono25x50  100,000 histories, same generator and geometry
    mono50     1e6 histories at 50 keV, H=500 um, R=5.0 um
    """
    if name in ("mono25", "mono25x50"):
        n = 2000 if name == "mono25" else 100000
        n = max(1, int(round(n * scale)))
        p = params(num_hist=n)
        s = 0x5EED0000 + (0 if name == "mono25" else 1) if seed is None else seed
        ev = generate(n, p["H"], 25.0, 150.0, s)
        return p, ev
    if name == "mono50":
        n = max(1, int(round(1_000_000 * scale)))
        p = params(num_hist=n, H=500.0, R=5.0, max_ph=5000, num_bins=1000)
        s = 0x5EED0004 if seed is None else seed
        ev = generate(n, p["H"], 50.0, float(attenuation_length_um(50.0)), s)
        return p, ev
    raise KeyError(name)


def batched_workload(scale: float = 1.0, seed: int = 0x5EED0002):
    """Four geometries x polychromatic 15-90 keV (1-keV bins), 50,000
    histories each.  Returns (list of params per geometry, events with the
    geometry index in `geom`, histories numbered per geometry)."""
    n = max(1, int(round(50000 * scale)))
    plist, evs = [], []
    rng = np.random.Generator(np.random.PCG64(seed))
    for gi, (H, R) in enumerate(BATCH_GEOMS):
        p = params(num_hist=n, H=H, R=R, max_ph=5000, num_bins=1000)
        e = rng.integers(15, 91, n).astype(np.float64)
        ev = generate(n, H, e, attenuation_length_um(e),
                      int(rng.integers(1 << 31)), geom=gi)
        plist.append(p)
        evs.append(ev)
    return plist, np.concatenate(evs)
