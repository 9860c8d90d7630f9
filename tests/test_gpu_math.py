"""The transport's device arithmetic on the GPU (fd2_debug_math) against the
same functions on the CPU (the oracle's fdo_math_eval), bit for bit, and the
GPU square root against the correctly rounded float32 square root.  The
faithful arithmetic's pieces (restated C-library acosf/asinf, float-pair
sin/cos/log) likewise: the CPU side of those is pinned to the C library and
to the double functions by oracle/check_faithful.c."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FNS = {  # argument ranges the transport passes
    "sin": (-3.2, 6.4), "cos": (-3.2, 6.4), "log": (6e-8, 1.0), "asin": (-1.0, 1.0),
    "acos": (-1.0, 1.0), "sin_turns": (0.0, 1.0), "cos_turns": (0.0, 1.0),
    "rsqrt": (1e-6, 1e6), "sqrt_unit": (0.0, 4.0), "angle_over_157": (-1.0, 1.0),
    "acosf_libc": (-1.0, 1.0), "asinf_libc": (-1.0, 1.0), "ff_sin": (-8.0, 8.0),
    "ff_cos": (-8.0, 8.0), "ff_log": (4.6e-10, 1.0),
}


def _args(fn, rng):
    lo, hi = FNS[fn]
    x = rng.uniform(lo, hi, 1 << 20).astype(np.float32)
    edges = np.array([lo, hi, 0.0, 0.5, -0.5, 1.0, -1.0, 7.90e-4, 8.03e-4, 7.9632671e-4],
                     np.float32)
    return np.concatenate([x, edges[(edges >= lo) & (edges <= hi)]])


@pytest.mark.parametrize("fn", sorted(FNS))
def test_device_math_matches_oracle(fd2lib, oracle, fn):
    from hybridmantis_amd import fd2

    x = _args(fn, np.random.default_rng(sum(map(ord, fn))))
    print(f"[gpu-math] {fn} n={len(x)}", flush=True)
    g, c = fd2.debug_math(fn, x), oracle.math_eval(fn, x)
    same = (g.view(np.uint32) == c.view(np.uint32)) | (np.isnan(g) & np.isnan(c))
    assert same.all(), f"{fn}: {(~same).sum()} differ, e.g. x={x[~same][:4]}"


def test_angle_band_exhaustive(fd2lib, oracle):
    """Every float between 7.0e-4 and 9.0e-4 (around cos 1.57) and NaN/|x|>1."""
    from hybridmantis_amd import fd2

    lo = np.float32(7.0e-4).view(np.uint32)
    hi = np.float32(9.0e-4).view(np.uint32)
    x = np.arange(lo, hi + 1, dtype=np.uint32).view(np.float32)
    x = np.concatenate([x, np.array([np.nan, 1.0000001, -1.0000001, 2.0, -2.0], np.float32)])
    print(f"[gpu-math] angle band n={len(x)}", flush=True)
    assert np.array_equal(fd2.debug_math("angle_over_157", x), oracle.math_eval("angle_over_157", x))


def test_sqrt_unit_correctly_rounded(fd2lib):
    """Every float in [0.5, 2) (two binades: every mantissa, both exponent
    parities) and a geometric sweep from 2^-96 to 2^100, zeros and inf."""
    from hybridmantis_amd import fd2

    a = np.float32(0.5).view(np.uint32)
    b = np.float32(2.0).view(np.uint32)
    x = np.concatenate([
        np.arange(a, b, dtype=np.uint32).view(np.float32),
        np.geomspace(2.0 ** -96, 2.0 ** 100, 1 << 20).astype(np.float32),
        np.array([0.0, -0.0, np.inf, 2.0 ** -96, 1.0 - 2.0 ** -24], np.float32)])
    print(f"[gpu-math] sqrt_unit n={len(x)}", flush=True)
    y = fd2.debug_math("sqrt_unit", x)
    assert np.array_equal(y.view(np.uint32), np.sqrt(x).view(np.uint32))
    assert np.isnan(fd2.debug_math("sqrt_unit", np.array([-1.0, np.nan], np.float32))).all()


def test_libc_arc_functions_every_float_in_two_binades(fd2lib, oracle):
    """The restated acosf/asinf on the GPU against the CPU restatement for
    every float in [0.5, 1] and [-1, -0.5] (their half-angle branches, with
    the |x| > 0.975 branch of asinf) and every 64th float of (-0.5, 0.5)."""
    from hybridmantis_amd import fd2

    h = np.float32(0.5).view(np.uint32)
    one = np.float32(1.0).view(np.uint32)
    top = np.arange(h, one + 1, dtype=np.uint32)
    mid = np.arange(0, h, 64, dtype=np.uint32)
    bits = np.concatenate([top, top | np.uint32(0x80000000), mid, mid | np.uint32(0x80000000)])
    x = bits.view(np.float32)
    for fn in ("acosf_libc", "asinf_libc"):
        print(f"[gpu-math] {fn} n={len(x)}", flush=True)
        g, c = fd2.debug_math(fn, x), oracle.math_eval(fn, x)
        assert np.array_equal(g.view(np.uint32), c.view(np.uint32)), fn
