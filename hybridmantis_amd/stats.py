"""Statistics of the transport tallies: pulse-height moments, Swank factor,
radial PRF profiles and two-sample comparisons.

The Swank factor of a pulse-height spectrum is I = M1^2 / (M0 M2) with Mn
the n-th moment of the detected-photon count over the histories that
detected at least one photon (the histories the reference's PHS keeps,
main/hybridMANTIS_c_ver1_0.c:354).
"""
from __future__ import annotations

import numpy as np


def phs_moments(hist_counts) -> tuple[float, float, float]:
    c = np.asarray(hist_counts, np.float64)
    c = c[c > 0]
    return float(len(c)), float(c.sum()), float((c * c).sum())


def swank(hist_counts) -> float:
    m0, m1, m2 = phs_moments(hist_counts)
    return m1 * m1 / (m0 * m2) if m0 > 0 and m2 > 0 else float("nan")


def mean_detected(hist_counts) -> float:
    m0, m1, _ = phs_moments(hist_counts)
    return m1 / m0 if m0 > 0 else float("nan")


def chi2_two_sample(h1, h2, min_count: float = 5.0) -> tuple[float, int, float]:
    """Two-sample chi-square for binned counts with unequal totals (bins
    with fewer than `min_count` entries in total are merged into one).
    Returns (chi2, dof, p)."""
    from scipy.stats import chi2 as chi2_dist

    a = np.asarray(h1, np.float64).ravel()
    b = np.asarray(h2, np.float64).ravel()
    tot = a + b
    keep = tot >= min_count
    if (~keep).any():
        a = np.append(a[keep], a[~keep].sum())
        b = np.append(b[keep], b[~keep].sum())
        tot = a + b
        nz = tot > 0
        a, b, tot = a[nz], b[nz], tot[nz]
    else:
        a, b = a[keep], b[keep]
        tot = a + b
    na, nb = a.sum(), b.sum()
    if na == 0 or nb == 0:
        return float("nan"), 0, float("nan")
    k1, k2 = np.sqrt(nb / na), np.sqrt(na / nb)
    x2 = float(np.sum((k1 * a - k2 * b) ** 2 / tot))
    dof = len(a) - 1
    return x2, dof, float(chi2_dist.sf(x2, dof))


def radial_profile(prf, center_px, edges_px) -> np.ndarray:
    """Sum of PRF counts in annuli of pixel-centre radius (pixels)."""
    img = np.asarray(prf, np.float64)
    i, j = np.indices(img.shape)
    r = np.hypot(i + 0.5 - center_px[0], j + 0.5 - center_px[1])
    idx = np.digitize(r.ravel(), edges_px) - 1
    out = np.zeros(len(edges_px) - 1)
    ok = (idx >= 0) & (idx < len(out))
    np.add.at(out, idx[ok], img.ravel()[ok])
    return out


def log_radial_bins(r_min: float = 0.5, r_max: float = 800.0, n: int = 64) -> np.ndarray:
    """Edges of a log-radial profile (default 0.5-800 um) plus an inner bin."""
    return np.concatenate([[0.0], np.geomspace(r_min, r_max, n)])


def l1_distance(p, q) -> float:
    """L1 distance between two distributions normalised to unit sum."""
    a = np.asarray(p, np.float64).ravel()
    b = np.asarray(q, np.float64).ravel()
    return float(np.abs(a / a.sum() - b / b.sum()).sum())
