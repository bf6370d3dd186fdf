"""The split-chain schedule of the fused H.v (blocks of the upper half-chain
state; sectors without symmetry maps) against the oracle: the element
formula is the reference's (base_general.py:354-410 with no maps), only the
order in which rows and terms are visited differs."""

import numpy as np
import pytest

from oracle import qsx_oracle as O
from tests.models import flat_terms, xxz_static

pytestmark = pytest.mark.gpu

KW = dict(check_symm=False, check_herm=False, check_pcon=False)


def _oracle(B, static):
    b = B.states.astype(np.uint64)
    g = O.Group(B.N, [], [], [])
    return O.make_csr(b, np.ones(len(b), np.int64), g, flat_terms(static), np.complex128)


def _model(L, cplx, seed):
    rng = np.random.default_rng(seed)
    st = xxz_static(L, Delta=0.8, h=rng.uniform(-2, 2, L))
    # next-nearest hops: across the halves and inside them; complex couplings
    # (with their conjugates, Hermitian) for the complex dtypes
    J2 = 0.35 + (0.25j if cplx else 0.0)
    st += [["+-", [[J2, i, (i + 2) % L] for i in range(L)]],
           ["-+", [[np.conj(J2), i, (i + 2) % L] for i in range(L)]],
           ["zz", [[0.2, i, (i + 3) % L] for i in range(L)]]]
    # a hop beside a z: where the z sits in the other half-chain the term
    # flips bits of one half and reads the other (the "low" and "up" term
    # classes of the split layout)
    J3 = 0.15 - (0.1j if cplx else 0.0)
    st += [["z+-", [[J3, i, (i + 1) % L, (i + 2) % L] for i in range(L)]],
           ["z-+", [[np.conj(J3), i, (i + 1) % L, (i + 2) % L] for i in range(L)]]]
    return st


@pytest.mark.parametrize("sched", ["split", "split2"])
@pytest.mark.parametrize("dt", [np.float64, np.complex128, np.float32, np.complex64])
def test_split_schedule_vs_oracle(dt, sched, monkeypatch):
    from quspin_amd import _lib as Lb
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    monkeypatch.setenv("QSX_SCHEDULE", sched)
    cplx = np.dtype(dt).kind == "c"
    tol = 1e-12 if np.dtype(dt) in (np.float64, np.complex128) else 3e-6
    for L, Nup in ((12, 6), (17, 6), (20, 10), (20, 3)):
        B = spin_basis_general(L, Nup=Nup, pauli=0)
        st = _model(L, cplx, L + Nup)
        A = _oracle(B, st)
        if not cplx:
            A = A.real
        Q = quantum_LinearOperator(st, basis=B, dtype=dt, **KW)
        rng = np.random.default_rng(Nup)
        v = rng.normal(size=B.Ns) + (1j * rng.normal(size=B.Ns) if cplx else 0)
        v = v.astype(dt)
        got = Lb.to_host(Q.dot(Lb.to_device(v)))
        cfg = next(iter(Q._qsx_operator().config.values()))
        print(f"case L={L} Nup={Nup} Ns={B.Ns} dtype={np.dtype(dt).name} cfg={cfg}", flush=True)
        assert cfg["schedule"] == sched
        ref = A @ v.astype(np.complex128 if cplx else np.float64)
        assert np.abs(got - ref).max() <= tol * max(1.0, np.abs(ref).max())
        # transposed and conjugated products (reference quirk B.1 included)
        if cplx:
            gotT = Lb.to_host(Q.T.dot(Lb.to_device(v)))
            np.testing.assert_allclose(gotT, A.T @ v.astype(np.complex128), rtol=0,
                                       atol=tol * max(1.0, np.abs(ref).max()))
            gotH = Lb.to_host(Q.H.dot(Lb.to_device(v)))
            np.testing.assert_allclose(gotH, A.conj().T @ v.astype(np.complex128), rtol=0,
                                       atol=tol * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("sched", ["split", "split2"])
@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_split_schedule_expm_and_lanczos(dt, sched, monkeypatch):
    """The Taylor loop (expm_multiply_parallel on the matrix-free operator;
    with split2 each term's epilogue runs inside the strip pass) and a
    captured Lanczos recurrence on the split schedule, against the oracle's
    expm and dense ground energy (complex64: FP32 vectors, FP64 sums)."""
    import torch
    from scipy.linalg import expm
    from quspin_amd import _lib as Lb
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.tools.evolution import expm_multiply_parallel
    from quspin_amd.tools.lanczos import lanczos_full
    monkeypatch.setenv("QSX_SCHEDULE", sched)
    L = 14
    B = spin_basis_general(L, Nup=7, pauli=0)
    st = _model(L, True, 3)
    A = _oracle(B, st).toarray()
    Q = quantum_LinearOperator(st, basis=B, dtype=dt, **KW)
    rng = np.random.default_rng(0)
    v = rng.normal(size=B.Ns) + 1j * rng.normal(size=B.Ns)
    v = (v / np.linalg.norm(v)).astype(dt)
    print(f"case sched={sched} dtype={np.dtype(dt).name} Ns={B.Ns}", flush=True)
    got = expm_multiply_parallel(Q, a=-0.4j, dtype=dt).dot(v)
    tol_x, tol_e = (1e-11, 1e-9) if dt == np.complex128 else (2e-6, 2e-4)
    np.testing.assert_allclose(got, expm(-0.4j * A) @ v.astype(np.complex128), atol=tol_x, rtol=0)
    E, _, _ = lanczos_full(Q, torch.from_numpy(v).cuda(), 80)
    E = np.asarray(E.cpu() if hasattr(E, "cpu") else E)
    assert abs(E[0] - np.linalg.eigvalsh(A)[0]) < tol_e
