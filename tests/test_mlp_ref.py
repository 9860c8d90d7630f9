"""The surrogate MLP's CPU reference against finite differences (float64,
no rounding): pins the reference's analytical backward pass, which in turn
checks the GPU kernels' gradients (tests/test_gpu_surrogate.py)."""
import numpy as np

from oracle import mlp_ref as M


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(3).astype(np.float64)
    p[M.B_OFF[0]:M.B_OFF[0] + 256] = rng.normal(0, 0.1, 256)
    B = 8
    x = rng.uniform(0, 1, (B, 5))
    y = np.zeros((B, 65))
    y[:, 0] = rng.uniform(0.1, 0.9, B)
    q = rng.dirichlet(np.ones(64), B)
    q[0] = 0.0  # a sample without detections: no radial term
    y[:, 1:] = q
    _, g = M.loss_and_grads(p, x, y, rnd=False)
    idx = np.concatenate([rng.choice(np.arange(M.W_OFF[l], M.W_OFF[l] + 5 * (256 if l < 3 else 80)), 6)
                          for l in range(4)] + [np.arange(M.B_OFF[3], M.B_OFF[3] + 3),
                                                np.arange(M.B_OFF[1], M.B_OFF[1] + 3)])
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grads(pp, x, y, rnd=False)[0] - M.loss_and_grads(pm, x, y, rnd=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_probs_are_distributions():
    p = M.init_params(1)
    x = np.random.default_rng(1).uniform(0, 1, (32, 5))
    out = M.probs(p, x)
    assert np.all((out[:, 0] > 0) & (out[:, 0] < 1))
    assert np.allclose(out[:, 1:].sum(1), 1.0)


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.0e-5, 65504.0], np.float32)
    r = M.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # ties to even
    assert r[2] == 1.0078125
    assert abs(r[3] - v[3]) / abs(v[3]) < 2 ** -8
