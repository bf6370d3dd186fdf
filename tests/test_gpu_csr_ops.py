"""Device CSR algebra (qsx_csr_sum_*, qsx_csr_transpose) against scipy, and
the hamiltonian operations built on it: H(t) (tocsr, _qsx_operator,
trace), transpose / conj / getH, +, scalar *, -, astype
(hamiltonian_core.py:1686-1720 and the hamiltonian algebra)."""

import numpy as np
import pytest
import scipy.sparse as sp

from tests.models import chain_maps, xxz_static

pytestmark = pytest.mark.gpu


def _rand(n, m, density, dtype, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, m, density=density, random_state=seed, format="csr")
    if np.dtype(dtype).kind == "c":
        A = A + 1j * sp.random(n, m, density=density, random_state=seed + 100, format="csr")
    A = sp.csr_matrix(A, dtype=dtype)
    A.data[:] = (rng.normal(size=A.nnz) + (1j * rng.normal(size=A.nnz) if A.dtype.kind == "c" else 0)).astype(dtype)
    A.sum_duplicates()
    return A


def _check_equal(M, ref, rtol):
    got = M.to_scipy()
    ref = sp.csr_matrix(ref)
    ref.eliminate_zeros()
    ref.sort_indices()
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got.indptr, ref.indptr)
    np.testing.assert_array_equal(got.indices, ref.indices)
    np.testing.assert_allclose(got.data, ref.data, rtol=rtol, atol=0)


@pytest.mark.parametrize("dts", [(np.float64, np.float64), (np.float32, np.float64), (np.complex128, np.float64),
                                 (np.complex64, np.complex64)])
def test_csr_combine_vs_scipy(dts):
    from quspin_amd.operators._devop import DeviceCSR, csr_combine
    print(f"case dtypes={[np.dtype(d).name for d in dts]}", flush=True)
    A, B = _rand(300, 300, 0.03, dts[0], 1), _rand(300, 300, 0.05, dts[1], 2)
    C = _rand(300, 300, 0.01, dts[1], 3)
    Ad, Bd, Cd = (DeviceCSR.from_scipy(X) for X in (A, B, C))
    tol = 1e-6 if np.float32 in dts or np.complex64 in dts else 1e-15
    _check_equal(csr_combine([Ad, Bd, Cd], [1.0, -0.5, 2.0]), A - 0.5 * B + 2.0 * C, tol)
    _check_equal(csr_combine([Ad, Bd], [0.3 + 0.2j, 1.0]), (0.3 + 0.2j) * A + B, tol)
    _check_equal(csr_combine([Ad, Bd], [1.0, 2.0], conj=[True, False]), A.conj() + 2.0 * B, tol)
    # exact cancellation leaves no entry
    Z = csr_combine([Ad, Ad], [1.0, -1.0])
    assert Z.nnz == 0 and Z.to_scipy().shape == (300, 300)
    # more matrices than one launch takes, and 64-bit indices
    mats = [DeviceCSR.from_scipy(_rand(64, 64, 0.1, np.float64, 10 + k)) for k in range(40)]
    ref = sum(0.5 ** k * m.to_scipy() for k, m in enumerate(mats))
    _check_equal(csr_combine(mats, [0.5 ** k for k in range(40)]), ref, 1e-14)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.complex64])
def test_csr_transpose_vs_scipy(dtype):
    from quspin_amd.operators._devop import DeviceCSR, csr_transpose
    A = _rand(257, 131, 0.04, dtype, 4)
    Ad = DeviceCSR.from_scipy(A)
    _check_equal(csr_transpose(Ad), A.T, 0)
    _check_equal(csr_transpose(Ad, conj=True), A.conj().T, 0)
    E = DeviceCSR.from_scipy(sp.csr_matrix((5, 7), dtype=dtype))
    T = csr_transpose(E)
    assert T.nnz == 0 and T.to_scipy().shape == (7, 5)


def test_hamiltonian_algebra_on_device():
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian
    L = 10
    m = chain_maps(L)
    B = spin_basis_general(L, Nup=5, pauli=0, kblock=(m["T"], 2))
    kw = dict(basis=B, check_symm=False, check_herm=False, check_pcon=False)
    drive = lambda t, w: np.cos(w * t) + 0.3j * np.sin(t)  # noqa: E731
    static = xxz_static(L, Delta=0.4)
    dyn = [["zz", [[0.7, i, (i + 2) % L] for i in range(L)], drive, (2.0,)],
           ["+-", [[0.2, i, (i + 3) % L] for i in range(L)], np.sin, ()]]
    H = hamiltonian(static, dyn, dtype=np.complex128, **kw)
    S = H._static.to_scipy()
    parts = {f: M.to_scipy() for f, M in H._dynamic.items()}
    for t in (0.0, 0.37, 1.9):
        ref = S + sum(f(t) * M for f, M in parts.items())
        _check_equal(H._qsx_operator(t), ref, 1e-14)
        np.testing.assert_allclose(H.tocsr(t).toarray(), ref.toarray(), rtol=0, atol=1e-14)
        np.testing.assert_allclose(H.trace(t), ref.diagonal().sum(), rtol=1e-13)
    t = 0.8
    ref = H.tocsr(t).toarray()
    np.testing.assert_allclose(H.transpose().tocsr(t).toarray(), ref.T, atol=1e-14)
    # conj() conjugates the matrices and the drives, so H.conj()(t) = conj(H(t))
    np.testing.assert_allclose(H.conj().tocsr(t).toarray(), ref.conj(), atol=1e-14)
    np.testing.assert_allclose(H.getH().tocsr(t).toarray(), ref.conj().T, atol=1e-14)
    np.testing.assert_allclose(H.conj().conj().tocsr(t).toarray(), ref, atol=1e-14)
    np.testing.assert_allclose((H + H.transpose()).tocsr(t).toarray(), ref + ref.T, atol=1e-14)
    np.testing.assert_allclose((2.5 * H).tocsr(t).toarray(), 2.5 * ref, atol=1e-14)
    np.testing.assert_allclose((-H).tocsr(t).toarray(), -ref, atol=1e-14)
    H64 = H.astype(np.complex64)
    assert H64.dtype == np.complex64 and H64._static.dtype == np.complex64
    np.testing.assert_allclose(H64.tocsr(t).toarray(), ref, atol=1e-6)
