"""The ODE driver ``tools.evolution.evolve`` (tools/evolution.py:144-460) on
host right-hand sides, against closed-form solutions of linear ODEs
(dense ``scipy.linalg.expm``)."""

import numpy as np
import pytest
from scipy.linalg import expm

from quspin_amd.tools.evolution import evolve


def _system(n=6, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, n))
    H = (X + X.T) / 2
    v0 = rng.normal(size=n) + 1j * rng.normal(size=n)
    return H, v0 / np.linalg.norm(v0)


def test_complex_schroedinger_matches_expm():
    H, v0 = _system()
    times = np.array([0.0, 0.3, 1.1])
    out = evolve(v0, 0.0, times, lambda t, y: -1j * (H @ y))
    assert out.shape == (6, 3) and out.dtype == np.complex128
    for k, t in enumerate(times):
        np.testing.assert_allclose(out[:, k], expm(-1j * t * H) @ v0, atol=1e-7)


def test_stacked_real_form_matches_complex():
    """u' = H v, v' = -H u for y = u + i v (the stacked form real H allows)."""
    H, v0 = _system(seed=1)
    n = H.shape[0]

    def f(t, y):
        return np.concatenate([H @ y[n:], -(H @ y[:n])])

    times = [0.0, 0.7]
    out = evolve(v0, 0.0, times, f, stack_state=True)
    np.testing.assert_allclose(out[:, 1], expm(-0.7j * H) @ v0, atol=1e-7)
    np.testing.assert_allclose(out[:, 0], v0, atol=0)


def test_imaginary_time_keeps_norm():
    H, v0 = _system(seed=2)
    v0 = 2.0 * v0.real
    out = evolve(v0, 0.0, [0.5, 2.0], lambda t, y: -(H @ y), real=True, imag_time=True)
    for k, t in enumerate([0.5, 2.0]):
        w = expm(-t * H) @ v0
        np.testing.assert_allclose(out[:, k], w / np.linalg.norm(w) * np.linalg.norm(v0), atol=1e-7)


def test_iterate_scalar_and_2d_states():
    H, v0 = _system(seed=3)
    V0 = np.stack([v0, v0.conj()], axis=1)
    f = lambda t, y: -1j * (H @ y.reshape(V0.shape)).ravel()  # noqa: E731
    times = [0.0, 0.4, 0.9]
    full = evolve(V0, 0.0, times, f)
    assert full.shape == (6, 2, 3)
    it = list(evolve(V0, 0.0, times, f, iterate=True))
    for k in range(3):
        np.testing.assert_allclose(it[k], full[..., k], atol=1e-12)
    one = evolve(V0, 0.0, 0.9, f)
    np.testing.assert_allclose(one, full[..., 2], atol=1e-12)
    np.testing.assert_allclose(full[:, 0, 2], expm(-0.9j * H) @ v0, atol=1e-7)


def test_errors():
    H, v0 = _system()
    f = lambda t, y: y  # noqa: E731
    with pytest.raises(ValueError):
        evolve(np.zeros((2, 2, 2)), 0.0, [1.0], f)
    with pytest.raises(ValueError):
        evolve(v0, 0.0, np.array([1j]), f)
    with pytest.raises(ValueError):
        evolve(v0, 0.0, [1.0], f, stack_state=True, imag_time=True)
