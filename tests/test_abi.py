"""libfd2.so loads without a GPU and exports every symbol include/fd2.h
declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b((?:fd2|gpuoptical|cpuoptical|gpuoptlb|cpuoptlb)\w*)\s*\(", txt)))


def test_library_exports_all_declared_symbols():
    from hybridmantis_amd import fd2

    lib = fd2.lib()
    names = declared("fd2.h")
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_strerror_and_host_phs_without_gpu():
    import numpy as np

    from hybridmantis_amd import fd2

    assert fd2.lib().fd2_strerror(-4).decode().startswith("event has more")
    h = fd2.phs_from_counts([0, 1, 10, 1399, 1400], 0, 1400, 140)
    assert h[0] == 1 and h[1] == 1 and h[139] == 2 and h.sum() == 4


def test_create_without_gpu_reports_no_device():
    from hybridmantis_amd import events as E, fd2

    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present")
    with pytest.raises(fd2.Fd2Error):
        fd2.Transport(E.params())
