"""Synthetic event generator (the PENELOPE stand-in)."""
import numpy as np

from hybridmantis_amd import events as E


def test_mono25_statistics():
    p, ev = E.workload("mono25")
    assert p["num_hist"] == 2000
    hist = np.unique(ev["hist"])
    assert hist.min() >= 1 and hist.max() <= 2000
    frac_none = 1 - len(hist) / 2000
    assert 0.17 < frac_none < 0.23
    k = np.bincount(ev["hist"])[hist]
    assert abs(k.mean() - 2.5) < 0.15
    per_hist = np.bincount(ev["hist"], weights=ev["N"])[hist]
    assert abs(per_hist.mean() / 1375.0 - 1) < 0.03  # 55/keV x 25 keV
    assert abs(ev["N"].sum() / 2.2e6 - 1) < 0.05
    assert (ev["z"] > -75).all() and (ev["z"] < 75).all()
    assert (ev["x"] > 0).all() and (ev["x"] < 1001).all()
    # energy conserved per history (25 keV)
    e = np.bincount(ev["hist"], weights=ev["E"])[hist]
    assert np.allclose(e, 25000.0, rtol=1e-4)


def test_depth_distribution_is_truncated_exponential():
    ev = E.generate(20000, 150.0, 25.0, 150.0, 3, sigma_um=0.0, p_none=0.0)
    first = np.unique(ev["hist"], return_index=True)[1]
    d = 75.0 - ev["z"][first].astype(np.float64)
    assert d.min() >= 0 and d.max() <= 150.0
    # mean of Exp(150) truncated to [0,150]: 150 - 150 e^-1/(1-e^-1)
    want = 150.0 - 150.0 * np.exp(-1) / (1 - np.exp(-1))
    assert abs(d.mean() - want) < 1.5


def test_batched_workload_shapes():
    plist, ev = E.batched_workload(scale=0.01)
    assert len(plist) == 4
    assert set(np.unique(ev["geom"])) == {0, 1, 2, 3}
    assert ev["N"].max() <= 90 * 55 * 1.2
    for g, (H, R) in enumerate(E.BATCH_GEOMS):
        assert plist[g]["H"] == H and plist[g]["R"] == R
        z = ev["z"][ev["geom"] == g]
        assert (np.abs(z) < H / 2).all()
