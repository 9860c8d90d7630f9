"""fd2_math.h (the kernel's elementary functions) against double precision."""
import numpy as np
import pytest


def ulp_err(got, ref):
    ref32 = ref.astype(np.float32)
    u = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / u


@pytest.mark.parametrize("fn,lo,hi,ref,tol", [
    ("sin", -3.2, 6.4, np.sin, 2.0),
    ("cos", -3.2, 6.4, np.cos, 2.0),
    ("log", 6e-8, 1.0, np.log, 2.0),
    ("asin", -1.0, 1.0, np.arcsin, 3.0),
    ("acos", -1.0, 1.0, np.arccos, 2.0),
    ("sin_turns", 0.0, 1.0, lambda u: np.sin(2 * np.pi * u), 2.0),
    ("cos_turns", 0.0, 1.0, lambda u: np.cos(2 * np.pi * u), 2.0),
    ("rsqrt", 1e-6, 1e6, lambda x: 1.0 / np.sqrt(x), 2.0),
])
def test_accuracy(oracle, fn, lo, hi, ref, tol):
    x = np.linspace(lo, hi, 400001).astype(np.float32)
    y = oracle.math_eval(fn, x)
    r = ref(x.astype(np.float64))
    keep = np.abs(r) > 1e-3  # relative error near the zeros is not meaningful
    assert ulp_err(y[keep], r[keep]).max() <= tol


def test_domains(oracle):
    assert np.isnan(oracle.math_eval("acos", [1.0000001, -1.5, np.nan])).all()
    assert np.isnan(oracle.math_eval("asin", [1.0000001, 2.0])).all()
    assert oracle.math_eval("log", [0.0])[0] == -np.inf
    assert oracle.math_eval("acos", [1.0])[0] == 0.0
    y = oracle.math_eval("acos", [-1.0])[0]
    assert abs(y - np.pi) < 3e-7



def test_turns_and_rsqrt_special_values(oracle):
    """Exact quarter turns give exact zeros and ones; the uniforms the
    kernel draws (24-bit, in (0, 1]) stay within 2 ulp; rsqrt over many
    binades including values just below 1 (normalisations)."""
    u = np.array([0.25, 0.5, 0.75, 1.0], np.float32)
    assert np.array_equal(oracle.math_eval("sin_turns", u), [1.0, 0.0, -1.0, 0.0])
    assert np.array_equal(oracle.math_eval("cos_turns", u), [0.0, -1.0, 0.0, 1.0])
    k = np.arange(1, 1 << 16, dtype=np.float64)
    uu = (k * 256.0 * 2.0 ** -24).astype(np.float32)  # a stride of the draws' lattice
    for fn, ref in (("sin_turns", np.sin), ("cos_turns", np.cos)):
        r = ref(2 * np.pi * uu.astype(np.float64))
        keep = np.abs(r) > 1e-3
        assert ulp_err(oracle.math_eval(fn, uu)[keep], r[keep]).max() <= 2.0
    x = np.concatenate([np.geomspace(1e-30, 1e30, 20001),
                        np.linspace(0.999, 1.001, 20001)]).astype(np.float32)
    assert ulp_err(oracle.math_eval("rsqrt", x), 1.0 / np.sqrt(x.astype(np.float64))).max() <= 2.0


def test_device_shortcut_facts_exhaustive():
    """Every float in [-1, 1]: the arc-cosine facts the GPU's branch-free
    1.57 angle test relies on; every uniform the kernel draws: the bound of
    its bulk-absorption shortcut (oracle/check_device.c)."""
    import os
    import subprocess
    here = os.path.join(os.path.dirname(__file__), "..", "oracle")
    subprocess.run(["make", "-s", "-C", here, "check_device"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))
    r = subprocess.run([os.path.join(here, "check_device")], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(" 0 violations") == 2, r.stdout


def test_angle_and_sqrt_unit_oracle(oracle):
    """angle_over_157 on fd2_acos and fd2_sqrt_unit (the CPU side of the
    device functions) at the band edges and special values."""
    x = np.array([-1.0, 0.0, 7.90e-4, 7.9632671e-4, 8.03e-4, 1.0, 1.0000001, -1.0000001,
                  np.nan], np.float32)
    a = oracle.math_eval("angle_over_157", x)
    assert list(a[:2]) == [1.0, 1.0] and a[2] == 1.0 and a[4] == 0.0 and a[5] == 0.0
    assert list(a[6:]) == [0.0, 0.0, 0.0]  # NaN angles pass the test
    s = np.array([0.0, 2.0 ** -24, 0.25, 1.0, np.inf], np.float32)
    assert np.array_equal(oracle.math_eval("sqrt_unit", s), np.sqrt(s))


def test_faithful_arithmetic_against_c_library():
    """oracle/check_faithful.c: the restated acosf/asinf equal the C
    library's for every float in [-1, 1]; the float-pair sin/cos/log are
    within 2^-45 of the double functions; each double expression of the
    reference, evaluated with float pairs, rounds to the same float as its
    double evaluation (at most 16 of 2^22 random arguments may differ; the
    expected rate is about 1e-6); the crosstalk pieces equal float sums."""
    import os
    import subprocess
    here = os.path.join(os.path.dirname(__file__), "..", "oracle")
    subprocess.run(["make", "-s", "-C", here, "check_faithful"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))
    r = subprocess.run([os.path.join(here, "check_faithful")], capture_output=True, text=True,
                       env=env, timeout=900)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("violations 0") == 5, r.stdout
