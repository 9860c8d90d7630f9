"""The gfortran-ABI entry points on the GPU.

The tally's call sequences (main/hybridMANTIS_tallyEnergyDepositionEvents.f:
198-304, main/hybridMANTIS_penEasy.f:434-443) are driven through the COMMON
blocks, and the accumulated outputimage_/outputdetprim_/optstats_ must equal
what the native API gives and what the oracle (its device arithmetic on the
production streams, the reference's PHS rule) gives for the same events,
photon counts and seeds (bit-exact: PRF, PHS and counters are integers).  The PHS of gpuoptical_
covers the whole run; that of cpuoptical_ only the events of its call, as in
the reference's drivers (hybridMANTIS_cuda_ver1_0.cu:376-399 after call 101,
hybridMANTIS_c_ver1_0.c:343-366 per call).
"""
import os

import numpy as np
import pytest

from hybridmantis_amd import events as E
from hybridmantis_amd import fd2
from hybridmantis_amd import fortran as F

pytestmark = pytest.mark.gpu

SEED = 4242
COUNTERS = [("glgen", "generated"), ("gldetect", "detected"), ("glabstop", "abs_top"),
            ("glabsbulk", "abs_bulk"), ("gllost", "lost"), ("gloutofcol", "outofcol"),
            ("gltheta1", "theta1")]


@pytest.fixture(scope="module")
def shim():
    os.environ["FD2_SEED"] = str(SEED)
    C = F.commons()
    yield C
    os.environ.pop("FD2_SEED", None)


def shim_params(p):
    """The COMMON /outputdetprim/ block holds 1000 PHS bins (the entry
    points abort on more): 1000 bins of 2 photons."""
    q = dict(p, num_bins=1000, max_ph=2000, min_ph=0)
    assert q["num_bins"] <= 1000 and (q["ubx"] - q["lbx"]) / q["pixel"] <= 501
    return q


def as_buffer_events(ev, p, seed, first):
    """Events as the entry points see them: positions through the cm buffer
    (x_cm * 1e4 rounded to float) and Poisson photon counts."""
    out = ev.copy()
    for k in ("x", "y", "z"):
        cm = np.asarray(ev[k], np.float64) / 1e4
        out[k] = (cm * 1e4).astype(np.float32)
    out["N"] = F.photon_counts(ev, p["yield_per_ev"], seed, first)
    return out


def native(p, buffers, seed):
    with fd2.Transport([p], seed=seed) as tr:
        for b in buffers:
            tr.submit(b)
        tr.sync()
        return tr.prf(0), tr.phs(0), tr.counters(0)


def oracle_tallies(oracle, p, buffers, seed):
    """The oracle on the same converted events, numbered on across the
    buffers as one context numbers them: PRF, the reference's PHS rule over
    the per-history counts (hybridMANTIS_cuda_ver1_0.cu:376-399) and the
    counters."""
    ev = np.concatenate(buffers)
    o = oracle.run(p, ev, p["num_hist"], rng=oracle.RNG_COUNTER, seed=seed,
                   math=oracle.MATH_DEVICE)
    phs = oracle.phs(o.hist_count, p["min_ph"], p["max_ph"], p["num_bins"])
    return o.prf, phs, dict(zip(oracle.COUNTER_NAMES, (int(v) for v in o.counters)))


def check_against_oracle(C, want_prf, want_phs, want_cnt):
    assert np.array_equal(image_part(C, want_prf), want_prf)
    assert C.image.sum() == want_prf.sum()
    assert np.array_equal(C.phs[:len(want_phs)], want_phs)
    st = C.stats()
    for a, b in COUNTERS:
        assert st[a] == want_cnt[b], a


def image_part(C, prf):
    xd, yd = prf.shape
    return C.image[:xd, :yd]


def test_gpuoptical_protocol(shim, oracle):
    C = shim
    p, ev = E.workload("mono25", scale=0.25)     # 500 histories
    p = shim_params(p)
    C.set_params(p)
    C.clear_outputs()
    cuts = [0, len(ev) // 3, 2 * len(ev) // 3, len(ev)]
    bufs = [ev[a:b] for a, b in zip(cuts[:-1], cuts[1:])]
    F.gpuoptical(99, len(bufs[0]))
    for i, b in enumerate(bufs):
        C.fill(b)
        F.gpuoptical(101 if i == len(bufs) - 1 else 100, len(b))
    conv = [as_buffer_events(b, p, SEED, a) for a, b in zip(cuts[:-1], bufs)]
    prf, phs, cnt = native(p, conv, SEED)
    assert cnt["generated"] == sum(int(b["N"].sum()) for b in conv) > 0
    assert np.array_equal(image_part(C, prf), prf)
    assert C.image.sum() == prf.sum()
    assert np.array_equal(C.phs[:len(phs)], phs)
    st = C.stats()
    for a, b in COUNTERS:
        assert st[a] == cnt[b], a
    assert st["glgputime"] > 0
    # and the oracle, bit for bit, after the 99/100/100/101 sequence
    check_against_oracle(C, *oracle_tallies(oracle, p, conv, SEED))
    # a second run accumulates on top (COMMON blocks are only ever added to)
    F.gpuoptical(99, len(bufs[0]))
    C.fill(ev)
    F.gpuoptical(101, len(ev))
    conv2 = as_buffer_events(ev, p, SEED, 0)
    prf2, phs2, cnt2 = native(p, [conv2], SEED)
    assert np.array_equal(image_part(C, prf), prf + prf2)
    assert C.stats()["glgen"] == cnt["generated"] + cnt2["generated"]


def test_cpuoptical_per_call_phs_and_range(shim, oracle):
    """cpuoptical_(gpusize) runs events [gpusize, myctropt); its PHS covers
    that call's events only."""
    C = shim
    p, ev = E.workload("mono25", scale=0.1)
    p = shim_params(p)
    C.set_params(p)
    C.clear_outputs()
    C.fill(ev)
    C.optical.myctropt = len(ev)
    split = len(ev) // 2
    F.cpuoptical(split)
    F.cpuoptical(0)
    # the k-th cpuoptical_ call of the process uses seed SEED + 1 + k; no
    # other test in this process calls it
    a = as_buffer_events(ev[split:], p, SEED + 1, 0)
    b = as_buffer_events(ev, p, SEED + 2, 0)
    prf_a, phs_a, cnt_a = native(p, [a], SEED + 1)
    prf_b, phs_b, cnt_b = native(p, [b], SEED + 2)
    assert np.array_equal(image_part(C, prf_a), prf_a + prf_b)
    assert np.array_equal(C.phs[:len(phs_a)], phs_a + phs_b)
    st = C.stats()
    for a_, b_ in COUNTERS:
        assert st[a_] == cnt_a[b_] + cnt_b[b_], a_
    # the oracle: each call's events on that call's seed, each call's own
    # PHS (hybridMANTIS_c_ver1_0.c:343-366) added into outputdetprim_
    oa, ob = oracle_tallies(oracle, p, [a], SEED + 1), oracle_tallies(oracle, p, [b], SEED + 2)
    check_against_oracle(C, oa[0] + ob[0], oa[1] + ob[1],
                         {k: oa[2][k] + ob[2][k] for k in oa[2]})


def test_load_balancing_probe(shim):
    """gpuoptlb_ transports events [0, lbbuf) and adds its tallies;
    cpuoptlb_ leaves every output untouched and reports +inf."""
    C = shim
    p, ev = E.workload("mono25", scale=0.05)
    p = shim_params(p)
    C.set_params(p)
    C.clear_outputs()
    C.fill(ev)
    before = C.image.copy()
    t = F.cpuoptlb(len(ev))
    assert np.isinf(t) and np.array_equal(C.image, before) and C.stats()["glgen"] == 0
    gt = F.gpuoptlb(len(ev))
    assert 0 < gt < 60
    # first gpuoptlb_ call of the process: seed SEED - 1
    b = as_buffer_events(ev, p, SEED - 1, 0)
    prf, phs, cnt = native(p, [b], SEED - 1)
    assert np.array_equal(image_part(C, prf), prf)
    assert np.array_equal(C.phs[:len(phs)], phs)
    st = C.stats()
    for a_, b_ in COUNTERS:
        assert st[a_] == cnt[b_], a_
