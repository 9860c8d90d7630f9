"""fd2_submit_device_noalloc (the allocation-free submit, work-queue scratch
in the library's static device array) against fd2_submit_device on the
same device events, through the C ABI: integer tallies bit for bit after
each of two calls in a row, the second with fewer events (the scratch keeps
the first call's contents, which must not leak into the second)."""
import numpy as np
import pytest

from hybridmantis_amd import events as E

pytestmark = pytest.mark.gpu


def tallies(tr):
    return tr.prf(), tr.hist_counts(), tr.counters()


def assert_same(a, b):
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2]


def test_noalloc_matches_submit_device_twice(fd2lib):
    p = E.params(num_hist=400)
    ev1 = E.generate(400, p["H"], 25.0, 150.0, 2024)
    ev2 = ev1[: len(ev1) // 3].copy()
    ev2["hist"] = np.minimum(ev2["hist"], 400)
    print(f"[noalloc] call 1: {len(ev1)} events, call 2: {len(ev2)} events", flush=True)
    with fd2lib.Transport(p) as ref, fd2lib.Transport(p) as na:
        r1, r2 = ref.upload_events(ev1), ref.upload_events(ev2)
        n1, n2 = na.upload_events(ev1), na.upload_events(ev2)
        ref.submit_device(r1, len(ev1))
        ref.sync()
        na.submit_device_noalloc(n1, len(ev1))
        assert_same(tallies(na), tallies(ref))
        ref.submit_device(r2, len(ev2))
        ref.sync()
        na.submit_device_noalloc(n2, len(ev2))
        assert_same(tallies(na), tallies(ref))
        assert na.counters()["generated"] == int(ev1["N"].sum() + ev2["N"].sum())


def test_noalloc_limits(fd2lib):
    p = E.params(num_hist=10)
    ev = E.generate(10, p["H"], 25.0, 150.0, 7)
    with fd2lib.Transport(p) as na:
        buf = na.upload_events(ev)
        with pytest.raises(fd2lib.Fd2Error):  # above the scratch's size
            na.submit_device_noalloc(buf, 1048577)
        na.submit_device_noalloc(buf, 0)
        assert na.counters()["generated"] == 0
        bad = ev.copy()
        bad["geom"][0] = 7  # an invalid event: FD2_E_INVAL from the queue build
        b2 = na.upload_events(bad)
        with pytest.raises(fd2lib.Fd2Error):
            na.submit_device_noalloc(b2, len(bad))
        na.reset()
        na.submit_device_noalloc(buf, len(ev))  # the error flag was cleared
        assert na.counters()["generated"] == int(ev["N"].sum())
