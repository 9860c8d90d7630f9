"""GPU transport (libfd2, HIP on gfx950) against the oracle.

* Trajectory test against the reference itself: the trajectory kernel in
  the reference C path's own arithmetic (fd2_trace_ranecu_faithful: float
  state, the double expressions evaluated with float pairs, the C library's
  acosf/asinf restated; no FP64 on the GPU) on the reference's per-event
  RANECU streams reproduces the compiled reference's fates and pixels
  (tests/golden/ref_*.npz) and the oracle's faithful mode on the full
  2,000-history case; at most 0.1% of photons may differ (a divergence in a
  RANECU event shifts the rest of that event's stream, so only the first
  differing photon of an event is counted).
* Trajectory tests: the GPU and the oracle's FP32 restatement
  (MATH_DEVICE) consume the same uniform streams -- the reference's
  per-event RANECU streams (one lane per event) and the production kernel's
  per-photon counter streams -- and must give every photon the same fate
  and PRF pixel (at most 0.1% of photons may differ; a divergence in a
  RANECU event shifts the rest of that event's stream, so only the first
  differing photon of an event is counted).  The oracle's control flow is
  itself pinned bit-exactly to the reference's C branch
  (tests/test_oracle_golden.py).
* Statistical test: the production kernel against the oracle's faithful
  mode (the C path's float/double arithmetic) on the same events: PRF and
  PHS chi-square p > 0.01, mean detected photons and Swank factor within
  0.5%.
* Determinism, split invariance, FP64 re-trace of positions, edge cases and
  size-independent properties at the full 1.1e8-photon workload.
"""
import os

import numpy as np
import pytest

from hybridmantis_amd import events as E
from hybridmantis_amd import stats

pytestmark = pytest.mark.gpu

CASES = ["mono25_small", "example_input", "smooth_nobulk_notop",
         "low_index_sensor_refl", "detector_edges", "thick_50kev",
         "thin100_poly", "mid300_r4_poly"]


def load_case(golden, name):
    here, man = golden
    m = man[name]
    d = np.load(os.path.join(here, m["file"]))
    return m["params"], d["events"], m["seed"]


def first_divergences(events, a, b):
    """Photons compared up to and including each event's first differing
    photon; returns (differing events, photons compared)."""
    n = np.maximum(events["N"], 0)
    off = np.concatenate([[0], np.cumsum(n)])
    bad, compared = 0, 0
    diff = (a["fate"] != b["fate"]) | (a["pix"] != b["pix"])
    for e in range(len(n)):
        d = diff[off[e]:off[e + 1]]
        if d.any():
            bad += 1
            compared += int(np.argmax(d)) + 1
        else:
            compared += int(n[e])
    return bad, compared


def load_reference_fixture(golden, name):
    here, man = golden
    d = np.load(os.path.join(here, man[name]["file"]))
    prf = np.zeros(int(np.prod(d["prf_shape"])), np.uint64)
    prf[d["prf_idx"]] = d["prf_val"]
    return d, prf.reshape(tuple(d["prf_shape"]))


@pytest.mark.parametrize("name", CASES)
def test_trajectory_faithful_matches_reference(fd2lib, oracle, golden, name):
    """The trajectory test against the reference itself: same
    RANECU streams, the reference's arithmetic; fates and pixels of every
    photon against what the reference's compiled C branch produced."""
    p, ev, seed = load_case(golden, name)
    d, ref_prf = load_reference_fixture(golden, name)
    print(f"[trajectory-faithful] {name}: {len(ev)} events, {int(ev['N'].sum())} photons",
          flush=True)
    with fd2lib.Transport(p) as tr:
        rec = tr.trace_ranecu(ev, seed_input=seed, faithful=True)
        counters = tr.counters()
        prf, hist = tr.prf(), tr.hist_counts()
    ref = {"fate": d["fate"].astype(np.int32), "pix": d["pix"]}
    bad, compared = first_divergences(ev, rec, ref)
    print(f"  events diverging {bad}, photons compared {compared}", flush=True)
    assert bad <= 0.001 * compared, (bad, compared)
    if bad == 0:
        assert np.array_equal(prf, ref_prf)
        assert np.array_equal(hist, d["hist"])
        assert [counters[k] for k in oracle.COUNTER_NAMES[:7]] == d["counters"][:7].tolist()
    # the GPU computes exactly the oracle's emulated arithmetic, bit for bit
    o = oracle.run(p, ev, p["num_hist"], rng=oracle.RNG_RANECU, seed=seed,
                   record=True, math=oracle.MATH_EMULATED)
    for f in ("fate", "pix", "x", "y"):
        assert np.array_equal(rec[f], o.photons[f]), f


@pytest.mark.parametrize("scale", [1.0, 10.0])
def test_trajectory_faithful_full_reference_case(fd2lib, oracle, scale):
    """The reference's CPU-runnable case in full (2,000 histories at 25 keV,
    ~2.2e6 photons, 501x501 1-um PRF), and ten times it, on RANECU streams:
    the faithful GPU kernel against the oracle's faithful mode (the
    reference's float/double arithmetic with the C library, bit-exact
    against the compiled reference on the fixtures), photon by photon.
    First-divergence bar 0.1%; the float-pair evaluation rounds like double
    except within ~2^-46 of a rounding boundary, measured at 2 events in
    2.2e6 photons on the CPU."""
    p, ev = E.workload("mono25", scale=scale)
    seed = 123456789
    print(f"[trajectory-faithful] mono25 {p['num_hist']} histories: {len(ev)} events "
          f"{int(ev['N'].sum())} photons", flush=True)
    with fd2lib.Transport(p) as tr:
        rec = tr.trace_ranecu(ev, seed_input=seed, faithful=True)
        prf, counters = tr.prf(), tr.counters()
    o, _ = oracle.run_parallel(p, ev, p["num_hist"], workers=min(16, os.cpu_count() or 1),
                               rng=oracle.RNG_RANECU, seed=seed, math=oracle.MATH_FAITHFUL,
                               record=True)
    bad, compared = first_divergences(ev, rec, o.photons)
    print(f"  events diverging {bad}, photons compared {compared}", flush=True)
    assert bad <= 0.001 * compared, (bad, compared)
    assert counters["generated"] == int(ev["N"].sum())
    # tallies agree up to the few diverging events' photons
    n_det = int(o.counters[1])
    assert abs(counters["detected"] - n_det) <= 2 * int(ev["N"][:].max()) * bad
    if bad == 0:
        assert np.array_equal(prf, o.prf)


@pytest.mark.parametrize("geom", [0, 1, 2, 3])
def test_trajectory_faithful_batched_geometries(fd2lib, oracle, geom):
    """Each geometry of the batched workload ((H, R) = (100, 2.55),
    (150, 2.55), (300, 4.0), (500, 5.0) um, 15-90 keV), 500 histories, on
    RANECU streams: the faithful GPU kernel against the oracle's faithful
    mode, photon by photon, first-divergence bar 0.1%."""
    plist, ev = E.batched_workload(scale=0.01)
    p = plist[geom]
    sel = ev[ev["geom"] == geom].copy()
    sel["geom"] = 0
    seed = 987654321
    print(f"[trajectory-faithful] geometry {geom}: {len(sel)} events "
          f"{int(sel['N'].sum())} photons", flush=True)
    with fd2lib.Transport(p) as tr:
        rec = tr.trace_ranecu(sel, seed_input=seed, faithful=True)
        prf = tr.prf()
    o, _ = oracle.run_parallel(p, sel, p["num_hist"], workers=min(16, os.cpu_count() or 1),
                               rng=oracle.RNG_RANECU, seed=seed, math=oracle.MATH_FAITHFUL,
                               record=True)
    bad, compared = first_divergences(sel, rec, o.photons)
    print(f"  events diverging {bad}, photons compared {compared}", flush=True)
    assert bad <= 0.001 * compared, (bad, compared)
    if bad == 0:
        assert np.array_equal(prf, o.prf)


@pytest.mark.parametrize("name", CASES)
def test_trajectory_ranecu_matches_oracle(fd2lib, oracle, golden, name):
    p, ev, seed = load_case(golden, name)
    print(f"[trajectory-ranecu] {name}: {len(ev)} events, {int(ev['N'].sum())} photons",
          flush=True)
    with fd2lib.Transport(p) as tr:
        rec = tr.trace_ranecu(ev, seed_input=seed)
        counters = tr.counters()
        prf, hist = tr.prf(), tr.hist_counts()
    o = oracle.run(p, ev, p["num_hist"], rng=oracle.RNG_RANECU, seed=seed,
                   record=True, math=oracle.MATH_DEVICE)
    bad, compared = first_divergences(ev, rec, o.photons)
    assert bad <= 0.001 * compared, (bad, compared)
    if bad == 0:
        assert np.array_equal(prf, o.prf)
        assert np.array_equal(hist, o.hist_count)
        assert [counters[k] for k in oracle.COUNTER_NAMES] == o.counters.tolist()


@pytest.mark.parametrize("name", CASES)
def test_production_kernel_matches_oracle_per_photon(fd2lib, oracle, golden, name):
    p, ev, seed = load_case(golden, name)
    print(f"[trajectory-counter] {name}", flush=True)
    with fd2lib.Transport(p, seed=seed) as tr:
        rec = tr.submit_record(ev)
        prf, hist, counters = tr.prf(), tr.hist_counts(), tr.counters()
    o = oracle.run(p, ev, p["num_hist"], rng=oracle.RNG_COUNTER, seed=seed,
                   record=True, math=oracle.MATH_DEVICE)
    diff = (rec["fate"] != o.photons["fate"]) | (rec["pix"] != o.photons["pix"])
    assert diff.sum() <= 0.001 * len(rec), diff.sum()
    if not diff.any():
        assert np.array_equal(prf, o.prf)
        assert np.array_equal(hist, o.hist_count)
        assert [counters[k] for k in oracle.COUNTER_NAMES] == o.counters.tolist()
    # detection points: identical FP32 values
    det = (rec["fate"] == 1) & ~diff
    assert np.array_equal(rec["x"][det], o.photons["x"][det])


@pytest.mark.parametrize("scale", [1.0, 10.0])
def test_plain_submit_tallies_match_oracle_full_reference_case(fd2lib, oracle, scale):
    """The launch the bench times (plain submit: no records, no radial
    tallies, wave queues, default tuning) on the reference's CPU-runnable
    case in full (2,000 histories at 25 keV, ~2.2e6 photons) and on ten
    times it: PRF, per-history counts and counters equal the oracle's device
    arithmetic on the same streams, bit for bit."""
    p, ev = E.workload("mono25", scale=scale)
    print(f"[plain-tallies] {len(ev)} events {int(ev['N'].sum())} photons", flush=True)
    with fd2lib.Transport(p) as tr:
        tr.submit(ev)
        tr.sync()
        prf, hist, counters = tr.prf(), tr.hist_counts(), tr.counters()
    o, _ = oracle.run_parallel(p, ev, p["num_hist"], workers=min(16, os.cpu_count() or 1),
                               rng=oracle.RNG_COUNTER, seed=271828182, math=oracle.MATH_DEVICE)
    assert [counters[k] for k in oracle.COUNTER_NAMES] == o.counters.tolist()
    assert np.array_equal(hist, o.hist_count)
    assert np.array_equal(prf, o.prf)


@pytest.mark.parametrize("scale", [0.25, 1.0])
def test_statistics_match_faithful_oracle(fd2lib, oracle, scale):
    """500 and 2,000 histories at 25 keV (~5.5e5 and ~2.2e6 photons, the
    latter the reference's CPU-runnable case in full), same events and
    streams; GPU in FP32, oracle in the C path's float/double arithmetic."""
    p, ev = E.workload("mono25", scale=scale)
    nh = p["num_hist"]
    print(f"[statistics] {len(ev)} events {int(ev['N'].sum())} photons", flush=True)
    with fd2lib.Transport(p) as tr:
        tr.submit(ev)
        g_prf, g_hist, g_phs, g_cnt = tr.prf(), tr.hist_counts(), tr.phs(), tr.counters()
    o, _ = oracle.run_parallel(p, ev, nh, workers=min(8, os.cpu_count() or 1),
                               math=oracle.MATH_FAITHFUL)
    o_phs = oracle.phs(o.hist_count, p["min_ph"], p["max_ph"], p["num_bins"])
    # PHS in 40-photon bins
    _, _, p_phs = stats.chi2_two_sample(g_phs.reshape(-1, 40).sum(1),
                                        o_phs.reshape(-1, 40).sum(1))
    assert p_phs > 0.01
    # PRF: radial profile about the beam (pixel 250.5,250.5), 1-um annuli
    edges = np.concatenate([np.arange(0, 20, 1.0), np.geomspace(20, 250, 30)])
    rg = stats.radial_profile(g_prf, (250.5, 250.5), edges)
    ro = stats.radial_profile(o.prf, (250.5, 250.5), edges)
    _, _, p_prf = stats.chi2_two_sample(rg, ro)
    assert p_prf > 0.01
    assert abs(stats.mean_detected(g_hist) / stats.mean_detected(o.hist_count) - 1) < 0.005
    assert abs(stats.swank(g_hist) / stats.swank(o.hist_count) - 1) < 0.005
    assert abs(g_cnt["detected"] / o.counter("detected") - 1) < 0.005
    assert g_cnt["generated"] == o.counter("generated")


@pytest.mark.parametrize("scale", [1.0, 10.0])
def test_statistics_match_reference_path_independent_streams(fd2lib, oracle, scale):
    """2,000 and 20,000 histories at 25 keV (~2.2e6 photons, the reference's
    CPU-runnable case, and ten times that): the production kernel
    (counter-based streams, FP32) against the reference C path itself -- the
    oracle's faithful arithmetic with the reference's per-event RANECU
    streams, pinned bit for bit to the compiled reference -- on the same
    events with independent random streams, so the chi-square p-values are
    those of independent samples.  At 20,000 histories the 0.5% bars on the
    mean detected count and the Swank factor are about 4 standard errors of
    their difference, at 2,000 about 1.4."""
    p, ev = E.workload("mono25", scale=scale)
    nh = p["num_hist"]
    print(f"[statistics-independent] {len(ev)} events {int(ev['N'].sum())} photons", flush=True)
    with fd2lib.Transport(p) as tr:
        tr.submit(ev)
        g_prf, g_hist, g_phs, g_cnt = tr.prf(), tr.hist_counts(), tr.phs(), tr.counters()
    o, _ = oracle.run_parallel(p, ev, nh, workers=min(16, os.cpu_count() or 1),
                               rng=oracle.RNG_RANECU, seed=123456789,
                               math=oracle.MATH_FAITHFUL)
    o_phs = oracle.phs(o.hist_count, p["min_ph"], p["max_ph"], p["num_bins"])
    _, _, p_phs = stats.chi2_two_sample(g_phs.reshape(-1, 40).sum(1),
                                        o_phs.reshape(-1, 40).sum(1))
    edges = np.concatenate([np.arange(0, 20, 1.0), np.geomspace(20, 250, 30)])
    _, _, p_prf = stats.chi2_two_sample(stats.radial_profile(g_prf, (250.5, 250.5), edges),
                                        stats.radial_profile(o.prf, (250.5, 250.5), edges))
    print(f"[statistics-independent] p_phs={p_phs:.3g} p_prf={p_prf:.3g}", flush=True)
    assert p_phs > 0.01 and p_prf > 0.01
    assert abs(stats.mean_detected(g_hist) / stats.mean_detected(o.hist_count) - 1) < 0.005
    assert abs(stats.swank(g_hist) / stats.swank(o.hist_count) - 1) < 0.005
    assert g_cnt["generated"] == o.counter("generated")


def test_deterministic_and_split_invariant(fd2lib):
    p, ev = E.workload("mono25", scale=0.2)
    with fd2lib.Transport(p) as tr:
        tr.submit(ev)
        a = (tr.prf(), tr.hist_counts(), tr.counters())
        tr.reset()
        tr.submit(ev)
        b = (tr.prf(), tr.hist_counts(), tr.counters())
        tr.reset()
        cut = [0, 7, len(ev) // 2, len(ev)]
        for i in range(3):
            tr.submit(ev[cut[i]:cut[i + 1]])
        c = (tr.prf(), tr.hist_counts(), tr.counters())
        tr.reset()
        tr.set_tuning(chunk=16, waves_per_cu=8)
        tr.submit(ev)
        d = (tr.prf(), tr.hist_counts(), tr.counters())
    for x in (b, c, d):
        assert np.array_equal(a[0], x[0])
        assert np.array_equal(a[1], x[1])
        assert a[2] == x[2]


def test_fp64_retrace_of_positions(fd2lib, oracle):
    """Every in-column step of 64 photons: the FP32 position reached equals
    the FP64 re-trace from the same FP32 state within 1e-5 relative."""
    p, ev = E.workload("mono25", scale=0.02)
    big = ev[np.argmax(ev["N"])]
    with fd2lib.Transport(p) as tr:
        steps = tr.debug_steps(np.array([big]), 64, 200000)
    assert len(steps) > 1000
    worst = 0.0
    for s in steps:
        if s["code"] != 0:
            continue  # lost / out of column: no end point to compare
        # forced along the FP32 run's decisions (plane vs wall, each nudge)
        _, pos = oracle.retrace_step(s["pos_in"], s["dcos"], s["xc"], s["yc"], p,
                                     forced=int(s["decisions"]))
        rel = np.abs(pos - s["pos_out"]) / np.maximum(np.abs(pos), 1.0)
        worst = max(worst, float(rel.max()))
    assert worst < 1e-5, worst


def test_edge_cases(fd2lib, oracle):
    p = E.params(num_hist=5)
    with fd2lib.Transport(p) as tr:
        tr.submit(np.zeros(0, E.EVENT_DTYPE))  # empty buffer
        ev = np.zeros(4, E.EVENT_DTYPE)
        ev["x"], ev["y"], ev["z"], ev["hist"] = 500.5, 500.5, 10.0, [1, 2, 3, 4]
        ev["N"] = [0, 0, 1, 0]  # ragged: events without photons
        rec = tr.submit_record(ev)
        assert len(rec) == 1 and rec["fate"][0] != 0
        assert tr.counters()["generated"] == 1
        big = ev[:1].copy()
        big["N"] = 900000  # the reference's max_photon_per_EDE
        tr.reset()
        tr.submit(big)
        c = tr.counters()
        assert c["generated"] == 900000
        assert c["generated"] == sum(c[k] for k in ("detected", "abs_top", "abs_bulk",
                                                    "lost", "outofcol", "theta1"))
        assert tr.hist_counts()[0] == c["detected"]
        big["N"] = 900001
        with pytest.raises(fd2lib.Fd2Error, match="900000"):
            tr.submit(big)
        bad = ev[:1].copy()
        bad["N"], bad["geom"] = 5, 3
        with pytest.raises(fd2lib.Fd2Error):
            tr.submit(bad)
        # an event on the detector's edge: its photons leave the detector
        edge = ev[:1].copy()
        edge["x"], edge["N"] = 0.5, 2000
        tr.reset()
        rec = tr.submit_record(edge)
        o = oracle.run(p, edge, 5, rng=oracle.RNG_COUNTER, math=oracle.MATH_DEVICE,
                       record=True, event_base=0)
        assert (rec["fate"] == 5).mean() > 0.3
        assert np.array_equal(rec["fate"], o.photons["fate"])


def test_many_ragged_events_match_oracle(fd2lib, oracle):
    """5,000 events, 60% of them without photons: the chunk-to-event search
    runs three rounds over a chunk table full of repeated offsets.  At two
    chunk sizes every photon's fate and pixel match the oracle's."""
    p = E.params(num_hist=5000)
    rng = np.random.default_rng(11)
    n = 5000
    ev = np.zeros(n, E.EVENT_DTYPE)
    ev["x"] = 500.5 + rng.normal(0.0, 3.0, n)
    ev["y"] = 500.5 + rng.normal(0.0, 3.0, n)
    ev["z"] = rng.uniform(-70.0, 70.0, n)
    ev["hist"] = np.arange(1, n + 1)
    ev["N"] = np.where(rng.random(n) < 0.6, 0, rng.integers(1, 40, n))
    assert n > 64 * 64  # three search rounds
    o = oracle.run(p, ev, n, rng=oracle.RNG_COUNTER, math=oracle.MATH_DEVICE, record=True)
    for chunk in (16, 128):
        print(f"[ragged] chunk {chunk}", flush=True)
        with fd2lib.Transport(p) as tr:
            tr.set_tuning(chunk=chunk)
            rec = tr.submit_record(ev)
            prf, hist = tr.prf(), tr.hist_counts()
        assert len(rec) == int(ev["N"].sum())
        diff = (rec["fate"] != o.photons["fate"]) | (rec["pix"] != o.photons["pix"])
        assert diff.sum() <= 0.001 * len(rec), diff.sum()
        if not diff.any():
            assert np.array_equal(prf, o.prf)
            assert np.array_equal(hist, o.hist_count)


def test_full_workload_properties(fd2lib):
    """100,000 histories (1.1e8 photons): conservation of photons and
    detections, and bit-identical tallies on a second run."""
    p, ev = E.workload("mono25x50")
    n = int(ev["N"].sum())
    print(f"[full] {len(ev)} events {n} photons", flush=True)
    with fd2lib.Transport(p) as tr:
        tr.submit(ev)
        prf, hist, c = tr.prf(), tr.hist_counts(), tr.counters()
        ms = tr.last_kernel_ms()
        tr.reset()
        tr.submit(ev)
        prf2, hist2, c2 = tr.prf(), tr.hist_counts(), tr.counters()
    print(f"[full] kernel {ms:.1f} ms, {n / ms * 1e3:.3e} photons/s", flush=True)
    assert c["generated"] == n
    assert n == sum(c[k] for k in ("detected", "abs_top", "abs_bulk", "lost",
                                   "outofcol", "theta1"))
    assert int(hist.sum()) == c["detected"]
    assert int(prf.sum()) <= c["detected"]
    assert np.array_equal(prf, prf2) and np.array_equal(hist, hist2) and c == c2
    # the PRF peaks at the beam pixel (500.5 um -> pixel 250)
    i, j = np.unravel_index(np.argmax(prf), prf.shape)
    assert abs(i - 250) <= 2 and abs(j - 250) <= 2


def test_reference_maximum_event_buffer(fd2lib):
    """The reference's largest event buffer, maxvalopt = 2,304,000 events
    (main/kernel_cuda_c_ver1_0.cu:93), in one submit: 2,304,000 events of
    0 to 100 photons (1.04e8 photons, about 10% of the events empty) over
    20,000 histories.  Conservation of photons and detections, and the same
    tallies as the events split over two submits (bit for bit: streams are
    keyed by event number)."""
    n_ev = 2304000
    rng = np.random.default_rng(23)
    p = E.params(num_hist=20000)
    ev = np.zeros(n_ev, E.EVENT_DTYPE)
    ev["x"] = 500.5 + rng.normal(0.0, 3.0, n_ev)
    ev["y"] = 500.5 + rng.normal(0.0, 3.0, n_ev)
    ev["z"] = rng.uniform(-0.5 * p["H"], 0.5 * p["H"], n_ev)
    ev["E"] = 1000.0
    ev["hist"] = np.sort(rng.integers(1, 20001, n_ev))
    ev["N"] = rng.integers(0, 101, n_ev)
    ev["N"][rng.random(n_ev) < 0.1] = 0
    n = int(ev["N"].sum())
    print(f"[max-buffer] {n_ev} events {n} photons", flush=True)
    with fd2lib.Transport(p) as tr:
        tr.submit(ev)
        a = (tr.prf(), tr.hist_counts(), tr.counters())
        tr.reset()
        half = n_ev // 2 + 7
        tr.submit(ev[:half])
        tr.submit(ev[half:])
        b = (tr.prf(), tr.hist_counts(), tr.counters())
    c = a[2]
    assert c["generated"] == n
    assert n == sum(c[k] for k in ("detected", "abs_top", "abs_bulk", "lost",
                                   "outofcol", "theta1"))
    assert int(a[1].sum()) == c["detected"]
    assert int(a[0].sum()) <= c["detected"]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
