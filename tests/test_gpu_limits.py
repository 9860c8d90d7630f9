"""Limits of the numbering and of one launch (GPU, through the C ABI).

* Event numbers are 64-bit: a context whose numbering passes 2^32 gives
  every photon the oracle's stream for its 64-bit event number (photon by
  photon), and events 2^32 apart get different photons.
* One launch numbers at most tune.max_chunks work items with its 32-bit
  queue offsets: a submit from host memory above the limit runs as several
  launches with the same photons as one; a device submit above it fails with
  FD2_E_INVAL at fd2_sync, tallies nothing, and leaves the context usable.
* The reference's RANECU seeding takes the event index as an int: the
  trajectory kernel rejects a range past 2^31 - 1.
"""
import numpy as np
import pytest

from hybridmantis_amd import events as E

pytestmark = pytest.mark.gpu


def small_case(n_hist=40, seed=5):
    p = E.params(num_hist=n_hist)
    ev = E.generate(n_hist, p["H"], 25.0, 150.0, seed)
    return p, ev


def test_event_numbers_past_2_32_match_oracle(fd2lib, oracle):
    p, ev = small_case()
    base = 2**32 - len(ev) // 2  # the events straddle 2^32
    print(f"[limits] {len(ev)} events from event {base}", flush=True)
    with fd2lib.Transport(p) as tr:
        tr.set_next_event(base)
        rec = tr.submit_record(ev)
        prf = tr.prf()
    o = oracle.run(p, ev, p["num_hist"], rng=oracle.RNG_COUNTER, seed=271828182,
                   event_base=base, record=True, math=oracle.MATH_DEVICE)
    assert np.array_equal(rec["fate"], o.photons["fate"])
    assert np.array_equal(rec["pix"], o.photons["pix"])
    assert np.array_equal(prf, o.prf)
    # the same events numbered 2^32 lower take other streams
    with fd2lib.Transport(p) as tr:
        tr.set_next_event(base - 2**32 + 2**31)  # any other numbering
        low = tr.submit_record(ev)
    assert not np.array_equal(rec["fate"], low["fate"])


def test_events_2_32_apart_have_distinct_streams(fd2lib):
    p, ev = small_case(20, 7)
    with fd2lib.Transport(p) as tr:
        tr.set_next_event(11)
        a = tr.submit_record(ev)
    with fd2lib.Transport(p) as tr:
        tr.set_next_event(11 + 2**32)
        b = tr.submit_record(ev)
    same = (a["fate"] == b["fate"]) & (a["pix"] == b["pix"])
    # independent streams: about half the photons share a fate by chance
    # (detected/lost dominate), never all of them
    assert same.mean() < 0.9, same.mean()


def test_host_submit_split_into_launches_keeps_photons(fd2lib):
    p, ev = small_case(60, 11)
    chunks = int(np.sum((ev["N"] + 127) // 128))
    print(f"[limits] {len(ev)} events, {chunks} work items, limit 37", flush=True)
    with fd2lib.Transport(p) as tr:
        tr.submit(ev)
        ref = tr.prf(), tr.hist_counts(), tr.counters()
    with fd2lib.Transport(p) as tr:
        tr.set_tuning(max_chunks=37)
        assert tr.tuning()["max_chunks"] == 37
        tr.submit(ev)
        got = tr.prf(), tr.hist_counts(), tr.counters()
        rec = tr.submit_record(ev)  # record submits split too
    assert np.array_equal(got[0], ref[0])
    assert np.array_equal(got[1], ref[1])
    assert got[2] == ref[2]
    with fd2lib.Transport(p) as tr:
        tr.submit(ev)  # the record submit above numbered its events from len(ev)
        want = tr.submit_record(ev)
    assert np.array_equal(rec, want)


def test_device_submit_over_limit_fails_cleanly(fd2lib):
    p, ev = small_case(30, 13)
    chunks = int(np.sum((ev["N"] + 127) // 128))
    with fd2lib.Transport(p) as tr:
        tr.set_tuning(max_chunks=max(1, chunks - 1))
        buf = tr.upload_events(ev)
        with pytest.raises(fd2lib.Fd2Error):
            tr.submit_device(buf, len(ev))
            tr.sync()
        assert tr.counters()["generated"] == 0
        assert int(tr.prf().sum()) == 0
        tr.set_tuning(max_chunks=chunks)  # exactly at the limit: runs
        tr.submit_device(buf, len(ev))
        tr.sync()
        assert tr.counters()["generated"] == int(ev["N"].sum())


def test_trace_rejects_event_index_past_int32(fd2lib):
    p, ev = small_case(5, 3)
    assert len(ev) >= 2
    with fd2lib.Transport(p) as tr:
        with pytest.raises(fd2lib.Fd2Error):
            tr.trace_ranecu(ev, event_base=2**31 - 2)
        tr.trace_ranecu(ev[:1], event_base=2**31 - 1)  # the last index is fine
