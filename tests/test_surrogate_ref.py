"""Pins of the surrogate checker (tests/surrogate_ref.py) by what the
mathematics fixes: every derivative against central finite differences,
gauge invariance of the action under random unitary transformations, and
bfloat16 rounding on hand-picked values."""
import numpy as np
import pytest

import surrogate_ref as sr

DIMS = (2, 2, 2, 4)


def rand_links(V, N, seed, eps=0.3):
    rng = np.random.default_rng(seed)
    return np.eye(N)[None, None] + eps * (rng.standard_normal((V, 5, N, N))
                                          + 1j * rng.standard_normal((V, 5, N, N)))


def rand_unitary(V, N, seed):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((V, N, N)) + 1j * rng.standard_normal((V, N, N))
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=-2, axis2=-1)
    return q * (d / np.abs(d))[:, None, :]


@pytest.fixture(scope="module")
def geo():
    return sr.Geometry(DIMS)


def test_bf16_rounding():
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, -2.5, 1 + 2 ** -7 + 2 ** -8], np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -7, -2.5, 1 + 2 ** -6], np.float32)
    assert np.array_equal(sr.bf16(vals), want)


@pytest.mark.parametrize("N", [2, 3])
def test_force_is_the_action_derivative(geo, N):
    th = sr.init_params(1, scale=2.0)
    U = rand_links(geo.V, N, 2)
    S, F = sr.action_and_force(geo, th, U)
    rng = np.random.default_rng(3)
    for _ in range(3):
        d = rng.standard_normal(U.shape) + 1j * rng.standard_normal(U.shape)
        h = 1e-5
        Sp, _ = sr.forward(geo, th, sr.features(geo, U + h * d))
        Sm, _ = sr.forward(geo, th, sr.features(geo, U - h * d))
        fd = (Sp - Sm) / (2 * h)
        an = 2 * np.vdot(F, d).real
        assert abs(fd - an) < 1e-6 * max(1.0, abs(an)), (fd, an)


@pytest.mark.parametrize("N", [2, 4])
def test_feature_tangent(geo, N):
    U = rand_links(geo.V, N, 5)
    R = rand_links(geo.V, N, 6) - np.eye(N)
    h = 1e-6
    fd = (sr.features(geo, U + h * R) - sr.features(geo, U - h * R)) / (2 * h)
    assert np.max(np.abs(fd - sr.features_jvp(geo, U, R))) < 1e-6


def test_gauge_invariance(geo):
    N = 3
    th = sr.init_params(7, scale=2.0)
    U = rand_links(geo.V, N, 8)
    g = rand_unitary(geo.V, N, 9)
    Ug = np.empty_like(U)
    for a in range(5):
        Ug[:, a] = g @ U[:, a] @ sr.dag(g[geo.shift(sr.EV[a])])
    assert np.allclose(sr.features(geo, Ug), sr.features(geo, U), atol=1e-12)
    S0, _ = sr.forward(geo, th, sr.features(geo, U))
    S1, _ = sr.forward(geo, th, sr.features(geo, Ug))
    assert abs(S0 - S1) < 1e-10 * abs(S0)


def test_conv_adjoint(geo):
    rng = np.random.default_rng(11)
    W = rng.standard_normal((sr.NTAP, 4, 3))
    x = rng.standard_normal((geo.V, 3))
    y = rng.standard_normal((geo.V, 4))
    assert abs(np.sum(y * sr.conv(geo, W, x)) - np.sum(x * sr.conv_t(geo, W, y))) < 1e-10


def test_force_loss_gradient(geo):
    N = 2
    th = sr.init_params(12, scale=2.0)
    th["b1"] = np.random.default_rng(13).standard_normal(sr.NCH) * 0.3
    th["b2"] = np.random.default_rng(14).standard_normal(sr.NCH) * 0.3
    U = rand_links(geo.V, N, 15)
    Ft = rand_links(geo.V, N, 16) - np.eye(N)
    L, g = sr.force_loss_grad(geo, th, U, Ft)
    assert abs(L - sr.force_loss(geo, th, U, Ft)) < 1e-12
    rng = np.random.default_rng(17)
    h = 1e-6
    for name in ("W1", "b1", "W2", "b2", "w3"):
        for _ in range(3):
            idx = tuple(rng.integers(0, s) for s in th[name].shape)
            tp = {k: v.copy() for k, v in th.items()}
            tm = {k: v.copy() for k, v in th.items()}
            tp[name][idx] += h
            tm[name][idx] -= h
            fd = (sr.force_loss(geo, tp, U, Ft) - sr.force_loss(geo, tm, U, Ft)) / (2 * h)
            assert abs(fd - g[name][idx]) < 1e-6 * max(1.0, abs(fd)), (name, idx, fd,
                                                                        g[name][idx])
