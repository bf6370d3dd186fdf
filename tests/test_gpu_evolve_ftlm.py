"""``hamiltonian.evolve`` with device right-hand sides (hamiltonian_core.py:1390-1685)
against closed-form propagators, and the FTLM / LTLM iterations
(tools/lanczos/_FTLM.py, _LTLM.py) on device Lanczos data against the oracle."""

import numpy as np
import pytest
from scipy.linalg import expm
from scipy.sparse.linalg import expm_multiply

from oracle import qsx_oracle as O
from tests.models import xxz_static

pytestmark = pytest.mark.gpu


def _chain(L, dtype=np.float64, dynamic=()):
    from quspin_amd.basis import spin_basis_1d
    from quspin_amd.operators import hamiltonian
    rng = np.random.default_rng(L)
    basis = spin_basis_1d(L, m=0, pauli=0)
    h = rng.uniform(-1, 1, size=L)
    static = xxz_static(L, J=1.0, Delta=0.7, h=h)
    dyn = [[op, bonds, f, args] for (op, bonds) in static for (f, args) in dynamic]
    return hamiltonian(static, dyn, basis=basis, dtype=dtype), basis


def _state(n, seed, complex_=True):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=n) + (1j * rng.normal(size=n) if complex_ else 0)
    return v / np.linalg.norm(v)


def test_evolve_static_schroedinger():
    print("evolve SE static L=10", flush=True)
    H, basis = _chain(10)
    v0 = _state(basis.Ns, 1)
    times = np.linspace(0, 2.0, 5)
    out = H.evolve(v0, 0.0, times)
    assert out.shape == (basis.Ns, 5)
    Hc = H.tocsr()
    for k, t in enumerate(times):
        np.testing.assert_allclose(out[:, k], expm_multiply(-1j * t * Hc, v0), atol=1e-7)
    # the stacked real form of a real Hamiltonian gives the same states
    outs = H.evolve(v0, 0.0, times, stack_state=True)
    np.testing.assert_allclose(outs, out, atol=1e-7)
    # a device start vector, several columns, and the generator form
    import torch
    V0 = np.stack([v0, _state(basis.Ns, 2)], axis=1)
    outm = H.evolve(torch.from_numpy(V0).cuda(), 0.0, times)
    assert outm.shape == (basis.Ns, 2, 5)
    np.testing.assert_allclose(outm[:, 0, :], out, atol=1e-7)
    it = list(H.evolve(v0, 0.0, times, iterate=True))
    np.testing.assert_allclose(np.stack(it, axis=1), out, atol=1e-12)


def test_evolve_driven_schroedinger():
    """H(t) = (1 + cos(w t)) H0 commutes with itself at all times:
    U(t) = exp(-i H0 (t + sin(w t) / w))."""
    print("evolve SE driven L=10", flush=True)
    w = 3.0
    drive = (lambda t, w: np.cos(w * t), (w,))
    H, basis = _chain(10, dynamic=[drive])
    H0, _ = _chain(10)
    v0 = _state(basis.Ns, 3)
    times = [0.0, 0.4, 1.5]
    out = H.evolve(v0, 0.0, times)
    H0c = H0.tocsr()
    for k, t in enumerate(times):
        np.testing.assert_allclose(out[:, k], expm_multiply(-1j * (t + np.sin(w * t) / w) * H0c, v0), atol=1e-7)


def test_evolve_imaginary_time():
    print("evolve SE imag_time L=10", flush=True)
    H, basis = _chain(10)
    v0 = 3.0 * _state(basis.Ns, 4, complex_=False)
    taus = [0.5, 2.0]
    out = H.evolve(v0, 0.0, taus, imag_time=True)
    assert out.dtype == np.float64
    Hc = H.tocsr()
    for k, t in enumerate(taus):
        w = expm_multiply(-t * Hc, v0)
        np.testing.assert_allclose(out[:, k], w / np.linalg.norm(w) * 3.0, atol=1e-7)


def test_evolve_liouville_von_neumann():
    print("evolve LvNE L=8", flush=True)
    H, basis = _chain(8, dtype=np.complex128)
    rng = np.random.default_rng(5)
    X = rng.normal(size=(basis.Ns, basis.Ns)) + 1j * rng.normal(size=(basis.Ns, basis.Ns))
    rho0 = X @ X.conj().T
    rho0 /= np.trace(rho0).real
    times = [0.0, 0.6]
    out = H.evolve(rho0, 0.0, times, eom="LvNE")
    assert out.shape == (basis.Ns, basis.Ns, 2)
    U = expm(-0.6j * H.toarray())
    np.testing.assert_allclose(out[..., 1], U @ rho0 @ U.conj().T, atol=1e-7)
    np.testing.assert_allclose(out[..., 0], rho0, atol=0)


def test_evolve_argument_errors():
    H, basis = _chain(6)
    v0 = _state(basis.Ns, 6)
    with pytest.raises(ValueError):
        H.evolve(v0[:-1], 0.0, [1.0])
    with pytest.raises(ValueError):
        H.evolve(v0, 0.0, [1.0], eom="XX")
    with pytest.raises(NotImplementedError):
        H.evolve(v0, 0.0, [1.0], imag_time=True, stack_state=True)
    with pytest.raises(ValueError):
        H.evolve(v0, 0.0, [1.0], eom="LvNE")
    Hc, _ = _chain(6, dtype=np.complex128)
    with pytest.raises(ValueError):
        Hc.evolve(v0, 0.0, [1.0], stack_state=True)


def _ftlm_setup(L=12, m=30):
    import torch
    from quspin_amd.operators import hamiltonian
    from quspin_amd.tools.lanczos import lanczos_full
    H, basis = _chain(L)
    # a non-Hermitian, complex observable, so that overlap orientation and
    # conjugation both matter
    Op = hamiltonian([["+-", [[1.0 + 0.5j, 0, 3]]], ["z", [[0.3, 1]]]], [], basis=basis, dtype=np.complex128,
                     check_herm=False)
    r = _state(basis.Ns, 7, complex_=False)
    E, V, Q = lanczos_full(H, torch.from_numpy(r).cuda(), m)
    return H, Op, E, V, Q


def test_ftlm_device_matches_oracle():
    print("FTLM L=12 m=30", flush=True)
    from quspin_amd.tools.lanczos import FTLM_static_iteration, lanczos_iter
    H, Op, E, V, Q = _ftlm_setup()
    betas = np.array([0.0, 0.3, 1.0, 5.0])
    got, c0 = FTLM_static_iteration({"O": Op, "H": H}, E, V, Q, beta=betas)
    Qh = Q.cpu().numpy()
    want, c0o = O.ftlm_static({"O": Op.tocsr(), "H": H.tocsr()}, E, V, Qh, beta=betas)
    for k in ("O", "H"):
        np.testing.assert_allclose(got[k], want[k], rtol=1e-10, atol=1e-12)
    assert np.iscomplexobj(got["O"]) and not np.iscomplexobj(got["H"])
    np.testing.assert_allclose(c0, c0o, rtol=1e-12)
    # beta = 0 gives <r|O|r> (V orthogonal)
    r = Qh[0]
    np.testing.assert_allclose(got["O"][0], np.vdot(r, Op.tocsr() @ r), rtol=1e-10)
    # host Krylov vectors and a scalar beta
    got_h, _ = FTLM_static_iteration({"O": Op}, E, V, Qh, beta=0.3)
    np.testing.assert_allclose(got_h["O"], want["O"][1], rtol=1e-10)
    # the vector iterator of lanczos_iter
    E2, V2, it = lanczos_iter(H, r, 30)
    got_i, _ = FTLM_static_iteration({"O": Op}, E2, V2, it, beta=betas)
    np.testing.assert_allclose(got_i["O"], want["O"], rtol=1e-8, atol=1e-10)


def test_ltlm_device_matches_oracle():
    print("LTLM L=12 m=30", flush=True)
    from quspin_amd.tools.lanczos import LTLM_static_iteration
    H, Op, E, V, Q = _ftlm_setup()
    Qh = Q.cpu().numpy()
    for betas in (np.array([0.0, 0.5, 2.0]), np.array([60.0, 80.0])):
        got, Id = LTLM_static_iteration({"O": Op, "H": H}, E, V, Q, beta=betas)
        want, Ido = O.ltlm_static({"O": Op.tocsr(), "H": H.tocsr()}, E, V, Qh, beta=betas)
        for k in ("O", "H"):
            np.testing.assert_allclose(got[k], want[k], rtol=1e-10, atol=1e-300)
        np.testing.assert_allclose(Id, Ido, rtol=1e-12)
