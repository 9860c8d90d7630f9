"""The gfortran-ABI shim (libfd2_fortran.so): exports, COMMON layouts and
the host photon-count sampler.  No GPU calls."""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest
from scipy import stats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol():
    from hybridmantis_amd import fortran as F

    txt = open(os.path.join(ROOT, "include", "fd2_fortran.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = sorted(set(re.findall(r"\b(\w+_|fd2f_\w+)\s*\(", txt)))
    assert {"gpuoptical_", "cpuoptical_", "gpuoptlb_", "cpuoptlb_", "fd2f_poisson"} <= set(names)
    L = F.lib()
    assert all(hasattr(L, n) for n in names)


def test_common_block_layouts():
    """Sizes of the COMMON blocks as the reference declares them
    (kernel_cuda_c_ver1_0.cu:86-234) and as the library defines them."""
    from hybridmantis_amd import fortran as F

    B = F.BUFSIZE
    want = {"optical_": 4 * 8 * B + 4 * B + 8, "optstats_": 64, "outputdetprim_": 4000,
            "outputimage_": 2 * 501 * 501 * 8, "inputargs_": 17 * 8 + 7 * 4}
    assert ctypes.sizeof(F.Optical) == want["optical_"]
    assert ctypes.sizeof(F.OptStats) == want["optstats_"]
    assert ctypes.sizeof(F.OutputImage) == want["outputimage_"]
    out = subprocess.run(["nm", "-D", "-S", "--defined-only", F.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    sizes = {}
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 4 and parts[3] in want:
            sizes[parts[3]] = int(parts[1], 16)
    for name, n in want.items():
        # the C struct of /inputargs/ carries 4 bytes of tail padding
        assert sizes[name] in (n, n + 4), (name, sizes.get(name), n)
    # member offsets of /optical/ follow Fortran storage order
    assert F.Optical.nbufopt.offset == 4 * 8 * B
    assert F.Optical.myctropt.offset == 4 * 8 * B + 4 * B


def test_cpu_probe_reports_infinite_time():
    """cpuoptlb_ returns +inf so the tally's balancer gives the GPU the whole
    buffer: factorLB = ((0/794) + 1) / ((0/gpuspeed) + 1) = 1 (tally
    .f:209-221)."""
    from hybridmantis_amd import fortran as F

    t = F.cpuoptlb(100000)
    assert math.isinf(t) and t > 0
    cpuspeed = 1000 / t
    factor = ((cpuspeed / 794) + 1) / ((cpuspeed / 12345.0) + 1)
    assert factor == 1.0


@pytest.mark.parametrize("mu", [0.3, 2.5, 9.99, 10.0, 37.0, 1375.0, 55000.0])
def test_poisson_distribution(mu):
    """Photon counts follow Poisson(mu): chi-square over the central bins of
    the scipy pmf, plus mean and variance within 5 standard errors."""
    from hybridmantis_amd import fortran as F

    n = 200000
    k = F.poisson(np.full(n, mu), seed=12345, first=7)
    assert abs(k.mean() - mu) < 5 * math.sqrt(mu / n)
    var_se = math.sqrt((mu + 2 * mu * mu) / n)  # var of the sample variance ~ (mu + 2mu^2)/n
    assert abs(k.var() - mu) < 5 * var_se
    lo, hi = stats.poisson.ppf(1e-4, mu), stats.poisson.ppf(1 - 1e-4, mu)
    edges = np.unique(np.linspace(lo, hi + 1, 41).astype(int))
    obs = np.histogram(k, bins=np.concatenate([[-1], edges[1:-1], [10 ** 9]]))[0]
    cdf = stats.poisson.cdf(np.concatenate([edges[1:-1] - 1, [10 ** 9]]), mu)
    exp = np.diff(np.concatenate([[0.0], cdf])) * n
    ok = exp > 5
    chi2 = ((obs[ok] - exp[ok]) ** 2 / exp[ok]).sum()
    p = stats.chi2.sf(chi2, ok.sum() - 1)
    assert p > 1e-4, (mu, chi2, ok.sum())


def test_poisson_streams_are_keyed_by_ordinal():
    """Counts depend on (seed, ordinal) only: splitting the run into buffers
    does not change them; another seed does."""
    from hybridmantis_amd import fortran as F

    mu = np.random.default_rng(0).uniform(0, 3000, 5000)
    a = F.poisson(mu, 99)
    b = np.concatenate([F.poisson(mu[:1234], 99, 0), F.poisson(mu[1234:], 99, 1234)])
    assert np.array_equal(a, b)
    assert not np.array_equal(a, F.poisson(mu, 100))
    assert np.all(F.poisson(np.zeros(10), 5) == 0)
    with pytest.raises(Exception):
        F.poisson(np.array([-1.0]), 5)
    with pytest.raises(Exception):
        F.poisson(np.array([np.nan]), 5)


def test_entry_point_errors_abort_with_a_message():
    """Errors in the entry points print a message and abort (the reference's
    cutilSafeCall behaviour); no silent return.  Here: a PHS of 2000 bins
    does not fit /outputdetprim/, or (without a GPU) no device."""
    import signal
    import sys

    code = (
        "from hybridmantis_amd import events as E, fortran as F\n"
        "C = F.commons()\n"
        "C.set_params(E.params(num_bins=2000))\n"
        "F.gpuoptical(99, 10)\n"
        "print('returned')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == -signal.SIGABRT, (r.returncode, r.stdout, r.stderr)
    assert "returned" not in r.stdout
    assert "fastDETECT2 (gpuoptical)" in r.stderr
