"""MI355X-native fastDETECT2 optical-photon transport (hybridMANTIS).

    from hybridmantis_amd import events, fd2
    p, ev = events.workload("mono25")
    with fd2.Transport(p) as tr:
        tr.submit(ev)
        prf, phs = tr.prf(), tr.phs()

The transport runs in libfd2.so (HIP kernels for gfx950 + C host library);
build it with `python -m hybridmantis_amd.build`.
"""
from . import events  # noqa: F401

__all__ = ["events", "fd2", "build"]
