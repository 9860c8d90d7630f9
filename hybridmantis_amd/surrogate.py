"""Surrogate of the depth-dependent optical response.

For a photon emitted at depth d below the top surface, offset col_dx from
its first column's axis, in a layer of thickness H with column radius R and
top reflectivity rho, the network predicts the probability that the photon
is detected and the distribution of its detection point's distance from the
emission point in 64 log-radial bins (0.5-800 um, plus [0, 0.5) and
[800, 1000) um).  Training targets come from the plain Monte Carlo
(fd2_submit_radial): each sample is one event of many photons at a grid
point of (d, col_dx, H, R, rho).

Features: d/H, col_dx/R, H/500 um, R/5 um, rho (the order fd2.h states).
"""
from __future__ import annotations

import os

import numpy as np

from . import events as E
from . import fd2

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "surrogate_mlp.npz")

TRAIN_GEOMS = E.BATCH_GEOMS                       # (H, R)
TRAIN_REFL = np.linspace(0.5, 0.95, 5)            # top reflectivities


def radial_edges() -> np.ndarray:
    """65 bin edges (um) of the network's radial distribution."""
    return np.concatenate([[0.0], 0.5 * 1600.0 ** (np.arange(63) / 62.0), [1000.0]])


def grid_params(geoms=TRAIN_GEOMS, refls=TRAIN_REFL):
    """One parameter set per (geometry, reflectivity); the PRF/PHS tallies of
    these sets are not used."""
    plist = []
    for (H, R) in geoms:
        for rho in refls:
            plist.append(E.params(num_hist=0, H=H, R=R, top_absfrac=float(1.0 - rho),
                                  num_bins=10, max_ph=100))
    return plist


def grid_events(n_depth: int, n_offset: int, photons: int, geoms=TRAIN_GEOMS,
                refls=TRAIN_REFL, depth_margin: float = 0.5):
    """Events at the grid points.  Depths are n_depth points spread over
    (margin, H - margin) um; offsets n_offset points over [0, R)."""
    rows, feats, depth_idx = [], [], []
    for gi, (H, R) in enumerate(geoms):
        depths = np.linspace(depth_margin, H - depth_margin, n_depth)
        offs = np.arange(n_offset) * (R / n_offset)
        for ri, rho in enumerate(refls):
            for di, d in enumerate(depths):
                for o in offs:
                    rows.append((E.BEAM_XY[0], E.BEAM_XY[1], H / 2.0 - d, 0.0, 0, photons,
                                 gi * len(refls) + ri, o))
                    feats.append((d / H, o / R, H / 500.0, R / 5.0, rho))
                    depth_idx.append(di)
    ev = np.array(rows, dtype=E.EVENT_DTYPE)
    return ev, np.array(feats, np.float32), np.array(depth_idx)


def targets_from_radial(rad, photons) -> np.ndarray:
    """[detected fraction, radial distribution] from per-event histograms."""
    rad = np.asarray(rad, np.float64)
    det = rad.sum(1)
    y = np.zeros((len(rad), 65), np.float32)
    y[:, 0] = det / photons
    ok = det > 0
    y[ok, 1:] = rad[ok] / det[ok, None]
    return y


def generate(n_depth=200, n_offset=64, photons=100000, seed=271828182, batch=4000,
             log=print):
    """Training set of n_depth x n_offset x 4 x 5 Monte Carlo samples."""
    ev, x, didx = grid_events(n_depth, n_offset, photons)
    plist = grid_params()
    ys = []
    with fd2.Transport(plist, seed=seed) as tr:
        for s in range(0, len(ev), batch):
            rad = fd2.submit_radial(tr, ev[s:s + batch])
            ys.append(targets_from_radial(rad, photons))
            if log:
                log(f"[surrogate-data] {min(s + batch, len(ev))}/{len(ev)} samples")
    return x, np.concatenate(ys), didx


def train(x, y, depth_idx, steps=3000, batch=4096, lr=2e-3, seed=0, holdout_every=10,
          device=0, log=print):
    """Adam on the GPU kernels; every `holdout_every`-th depth is held out.
    Returns (network, report dict)."""
    hold = (depth_idx % holdout_every) == (holdout_every // 2)
    net = fd2.Mlp(device=device, seed=seed)
    net.set_data(x[~hold], y[~hold])
    losses, total_ms = [], 0.0
    segments = [(0.5, lr), (0.3, lr / 4), (0.2, lr / 16)]
    for frac, rate in segments:
        n = max(1, int(steps * frac))
        l, ms = net.train(n, batch=batch, lr=rate, seed=seed + len(losses))
        losses.append(l)
        total_ms += ms
        if log:
            log(f"[surrogate-train] lr {rate:.2e}: loss {l[0]:.4f} -> {l[-1]:.4f}")
    losses = np.concatenate(losses)
    pred = net.forward(x[hold])
    rep = dict(steps=int(len(losses)), final_loss=float(losses[-20:].mean()),
               train_ms=total_ms,
               holdout_pdet_mae=float(np.abs(pred[:, 0] - y[hold, 0]).mean()),
               holdout_pdet_rel=float(np.abs(pred[:, 0] - y[hold, 0]).sum() / y[hold, 0].sum()),
               holdout_radial_l1=float(np.abs(pred[:, 1:] - y[hold, 1:]).sum(1).mean()))
    # FLOPs: forward 2*(5*256 + 2*256*256 + 256*65) per sample, training ~3x
    fwd = 2.0 * (5 * 256 + 2 * 256 * 256 + 256 * 65)
    rep["train_tflops"] = 3 * fwd * batch * len(losses) / (total_ms * 1e-3) / 1e12
    return net, rep


def save(net: "fd2.Mlp", path: str = WEIGHTS, **meta) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez_compressed(path, params=net.params(), **{k: np.asarray(v) for k, v in meta.items()})


def load(device: int = 0, path: str = WEIGHTS) -> "fd2.Mlp":
    d = np.load(path)
    net = fd2.Mlp(device=device)
    net.set_params(d["params"])
    return net
