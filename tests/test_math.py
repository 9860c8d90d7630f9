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

