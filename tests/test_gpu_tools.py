"""expm_multiply_parallel and Lanczos on the device vs the oracle / scipy
(test_expm_multiply_parallel*.py, test_Lanczos.py)."""

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import expm_multiply

from oracle import qsx_oracle as O
from tests.models import chain_maps, xxz_static

pytestmark = pytest.mark.gpu


def test_expm_random_complex_matrix():
    """test_expm_multiply_parallel.py:53-76 (N=3500, density log N / N), atol 1e-15
    relaxed to 1e-13 for the parallel reduction order."""
    from quspin_amd.tools.evolution import expm_multiply_parallel
    rng = np.random.RandomState(0)
    N = 3500
    for i in range(3):
        A = sp.random(N, N, density=np.log(N) / N, random_state=rng) + 1j * sp.random(
            N, N, density=np.log(N) / N, random_state=rng)
        A = A.tocsr()
        v = rng.normal(0, 1, size=N) + 1j * rng.normal(0, 1, size=N)
        v /= np.linalg.norm(v)
        v1 = expm_multiply(A, v)
        v2 = expm_multiply_parallel(A).dot(v)
        np.testing.assert_allclose(v1, v2, rtol=0, atol=1e-13)
        v3, _ = O.expm_multiply(A, v)
        np.testing.assert_allclose(v3, v2, rtol=0, atol=1e-13)


def test_expm_int8_matrix_complex_a():
    from quspin_amd.tools.evolution import expm_multiply_parallel
    rng = np.random.RandomState(0)
    N = 2000
    A = sp.random(N, N, density=np.log(N) / N, random_state=rng,
                  data_rvs=lambda n: rng.randint(-100, 100, size=n).astype(np.int8), dtype=np.int8).tocsr()
    v = rng.normal(0, 1, size=N) + 1j * rng.normal(0, 1, size=N)
    v /= np.linalg.norm(v)
    v1 = expm_multiply(-0.01j * A, v)
    v2 = expm_multiply_parallel(A, a=-0.01j, dtype=np.complex128).dot(v)
    np.testing.assert_allclose(v1, v2, rtol=0, atol=1e-13)


def test_expm_batch_and_dense_check():
    from quspin_amd.tools.evolution import expm_multiply_parallel
    from scipy.linalg import expm
    rng = np.random.default_rng(2)
    N = 200
    A = sp.random(N, N, density=0.05, random_state=3).tocsr()
    A = A + A.T
    V = rng.normal(size=(N, 10)) + 1j * rng.normal(size=(N, 10))
    U = expm_multiply_parallel(A, a=-0.3j)
    got = U.dot(V)
    ref = expm(-0.3j * A.toarray()) @ V
    np.testing.assert_allclose(got, ref, atol=1e-12)


def test_expm_imag_time_chain():
    """test_expm_multiply_parallel.py:9-50 (L reduced to 14): repeated
    imaginary-time steps agree with scipy."""
    from quspin_amd.basis import spin_basis_1d
    from quspin_amd.operators import hamiltonian
    from quspin_amd.tools.evolution import expm_multiply_parallel
    L = 14
    basis = spin_basis_1d(L, m=0, kblock=0, pblock=1, zblock=1)
    J = [[1.0, i, (i + 1) % L] for i in range(L)]
    static = [["xx", J], ["yy", J], ["zz", J]]
    H = hamiltonian(static, [], basis=basis, dtype=np.float64)
    E = np.linalg.eigvalsh(H.toarray())[0]
    A = -(H.tocsr() - E * sp.eye(H.Ns, format="csr", dtype=np.float64))
    U = expm_multiply_parallel(A)
    rng = np.random.RandomState(0)
    v1 = rng.uniform(-1, 1, size=H.Ns)
    v1 /= np.linalg.norm(v1)
    v2 = v1.copy()
    for i in range(30):
        v2 = U.dot(v2)
        v2 /= np.linalg.norm(v2)
        v1 = expm_multiply(A, v1)
        v1 /= np.linalg.norm(v1)
    np.testing.assert_allclose(v1, v2, rtol=0, atol=1e-13)


def test_expm_matrix_free_equals_csr():
    import torch
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian, quantum_LinearOperator
    from quspin_amd.tools.evolution import expm_multiply_parallel
    L = 16
    m = chain_maps(L)
    B = spin_basis_general(L, Nup=8, pauli=0, kblock=(m["T"], 2))
    static = xxz_static(L, Delta=0.5)
    Q = quantum_LinearOperator(static, basis=B, dtype=np.complex128, check_symm=False, check_herm=False,
                               check_pcon=False)
    H = hamiltonian(static, [], basis=B, dtype=np.complex128, check_symm=False, check_herm=False, check_pcon=False)
    rng = np.random.default_rng(0)
    v = rng.normal(size=B.Ns) + 1j * rng.normal(size=B.Ns)
    v /= np.linalg.norm(v)
    Um = expm_multiply_parallel(Q, a=-0.5j)
    Uc = expm_multiply_parallel(H.tocsr(), a=-0.5j)
    assert Um.partition == Uc.partition
    np.testing.assert_allclose(Um.dot(v), Uc.dot(v), atol=1e-13)
    ref = expm_multiply(-0.5j * H.tocsr(), v)
    np.testing.assert_allclose(Um.dot(v), ref, atol=1e-12)


def test_lanczos_ground_state_and_expm():
    """test_Lanczos.py with L reduced: E_lanczos vs exact to 1e-10,
    expm_lanczos vs scipy expm_multiply."""
    from quspin_amd.basis import spin_basis_1d
    from quspin_amd.operators import hamiltonian
    from quspin_amd.tools.lanczos import lanczos_full, lanczos_iter, lin_comb_Q_T, expm_lanczos
    L = 16
    basis = spin_basis_1d(L, m=0, kblock=0, pblock=1, zblock=1)
    J_list = [[1.0, i, (i + 1) % L] for i in range(L)]
    static = [[op, J_list] for op in ["xx", "yy", "zz"]]
    H = -hamiltonian(static, [], basis=basis, dtype=np.float64)
    np.random.seed(0)
    v0 = np.random.normal(0, 1, size=basis.Ns)
    v0 /= np.linalg.norm(v0)
    v1 = v0.copy()
    v2 = v0.copy()
    for i in range(10):
        E1, V1, lv1 = lanczos_full(H, v1, 20)
        E2, V2, lv2 = lanczos_iter(H, v2, 20)
        v0 = expm_multiply(-0.1j * H.static, v0)
        v1 = expm_lanczos(E1, V1, lv1, a=-0.1j)
        v2 = expm_lanczos(E2, V2, lv2, a=-0.1j)
        np.testing.assert_allclose(v1, v0, atol=1e-10, rtol=0)
        np.testing.assert_allclose(v2, v0, atol=1e-10, rtol=0)
    E_GS = np.linalg.eigvalsh(H.toarray())[-1]
    v0 = np.random.normal(0, 1, size=basis.Ns)
    v0 /= np.linalg.norm(v0)
    E, V, Q = lanczos_full(H, v0, 50, full_ortho=False)
    assert abs(E[-1] - E_GS) < 1e-10
    Eo, Vo, Qo, ao, bo = O.lanczos_full(lambda x: H.tocsr() @ x, v0, 50)
    # extreme (converged) Ritz values agree; interior ones depend on rounding
    # once orthogonality is lost, so only the first recurrence steps are compared
    assert abs(E[0] - Eo[0]) < 1e-9 and abs(E[-1] - Eo[-1]) < 1e-9
    np.testing.assert_allclose(Q[:8], Qo[:8], atol=1e-10)
    v1 = lin_comb_Q_T(V[:, -1], Q)
    E, V, Q_iter = lanczos_iter(H, v0, 50, return_vec_iter=True)
    v2 = lin_comb_Q_T(V[:, -1], Q_iter)
    np.testing.assert_allclose(v2, v1, atol=1e-10, rtol=0)


def test_lanczos_device_L20_kpz_ground_energy():
    """100-step Lanczos on the L=20 k=0, p=+1, z=+1 sector from device vectors: ground energy
    agrees with the oracle recurrence on the same device matvec to 1e-9."""
    import torch
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.tools.lanczos import lanczos_full
    L = 20
    m = chain_maps(L)
    B = spin_basis_general(L, Nup=10, pauli=0, kblock=(m["T"], 0), pblock=(m["P"], 0), zblock=(m["Z"], 0))
    Q = quantum_LinearOperator(xxz_static(L), basis=B, dtype=np.float64, check_symm=False, check_herm=False,
                               check_pcon=False)
    rng = np.random.default_rng(0)
    v0 = rng.normal(size=B.Ns)
    E, V, Qm = lanczos_full(Q, torch.from_numpy(v0).cuda(), 100)
    Eo, _, _, _, _ = O.lanczos_full(lambda x: Q.dot(x), v0 / np.linalg.norm(v0), 100)
    assert abs(E[0] - Eo[0]) < 1e-9
    # exact ground energy of the sector by dense diagonalisation of the oracle matrix
    from tests.test_gpu_hv import _oracle_H
    E_exact = np.linalg.eigvalsh(_oracle_H(B, xxz_static(L)).toarray())[0]
    assert abs(E[0] - E_exact) < 1e-9


def test_degenerate_operators_and_steps():
    """An operator with no terms (products are zero, its exponential is the
    identity), a zero time step (e^{0 H} v = v), and a Krylov space as large
    as the sector allows (m = Ns - 1): every Ritz value equals the oracle
    recurrence's on the same matrix and start vector."""
    import torch
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian, quantum_LinearOperator
    from quspin_amd.tools.evolution import expm_multiply_parallel
    from quspin_amd.tools.lanczos import lanczos_full
    from tests.test_gpu_hv import _oracle_H
    kw = dict(check_symm=False, check_herm=False, check_pcon=False)
    L = 8
    m = chain_maps(L)
    B = spin_basis_general(L, Nup=4, pauli=0, kblock=(m["T"], 0), zblock=(m["Z"], 0))
    rng = np.random.default_rng(5)
    v = (rng.normal(size=B.Ns) + 1j * rng.normal(size=B.Ns)).astype(np.complex128)
    print(f"case L={L} Ns={B.Ns} no terms", flush=True)
    Q0 = quantum_LinearOperator([], basis=B, dtype=np.complex128, **kw)
    np.testing.assert_array_equal(Q0.dot(v), np.zeros_like(v))
    # (hamiltonian([], [], basis=...) has no shape to take, as in the
    # reference: hamiltonian_core.py:359-391)
    with pytest.raises(ValueError):
        hamiltonian([], [], basis=B, dtype=np.complex128, **kw)
    np.testing.assert_allclose(expm_multiply_parallel(Q0, a=-1.0j, dtype=np.complex128).dot(v.copy()), v,
                               atol=1e-14, rtol=0)
    print("case zero step", flush=True)
    Q = quantum_LinearOperator(xxz_static(L), basis=B, dtype=np.complex128, **kw)
    np.testing.assert_allclose(expm_multiply_parallel(Q, a=0.0, dtype=np.complex128).dot(v.copy()), v,
                               atol=1e-14, rtol=0)
    print("case Lanczos m = Ns - 1", flush=True)
    Qr = quantum_LinearOperator(xxz_static(L), basis=B, dtype=np.float64, **kw)
    Hd = _oracle_H(B, xxz_static(L)).toarray().real
    v0 = rng.normal(size=B.Ns)
    E, _, _ = lanczos_full(Qr, torch.from_numpy(v0).cuda(), B.Ns - 1)
    E = E.cpu().numpy() if hasattr(E, "cpu") else np.asarray(E)
    Eo = O.lanczos_full(lambda x: Hd @ x, v0 / np.linalg.norm(v0), B.Ns - 1)[0]
    np.testing.assert_allclose(np.sort(E), np.sort(Eo), atol=1e-9, rtol=0)


def test_scipy_eigsh_on_linear_operator():
    """quantum_LinearOperator is a scipy LinearOperator (the reference's usage
    in eigsh(H, k, which="SA")): scipy's Lanczos driven by the device product
    gives the lowest eigenvalues of the oracle's matrix, for a real k=0/p/z
    sector and a complex k-sector."""
    from scipy.sparse.linalg import eigsh
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from tests.test_gpu_hv import _oracle_H
    kw = dict(check_symm=False, check_herm=False, check_pcon=False)
    L = 14
    m = chain_maps(L)
    for blocks, dt in (({"kblock": (m["T"], 0), "pblock": (m["P"], 0), "zblock": (m["Z"], 0)}, np.float64),
                       ({"kblock": (m["T"], 3)}, np.complex128)):
        B = spin_basis_general(L, Nup=L // 2, pauli=0, **blocks)
        print(f"case L={L} Ns={B.Ns} dtype={np.dtype(dt).name}", flush=True)
        Q = quantum_LinearOperator(xxz_static(L, Delta=0.9), basis=B, dtype=dt, **kw)
        ev = np.linalg.eigvalsh(_oracle_H(B, xxz_static(L, Delta=0.9)).toarray())
        w = np.sort(eigsh(Q, k=3, which="SA", return_eigenvectors=False))
        np.testing.assert_allclose(w, ev[:3], atol=1e-9, rtol=0)


def test_expm_set_a_work_array_and_overwrite():
    """ExpmMultiplyParallel options (EMP.html:117-374): set_a re-times the
    same operator (a new partition), several vectors at once with a caller's
    work_array, overwrite_v writing into the caller's array; against dense
    exponentials of the oracle's matrix."""
    from scipy.linalg import expm
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.tools.evolution import expm_multiply_parallel
    from tests.test_gpu_hv import _oracle_H
    L = 12
    m = chain_maps(L)
    B = spin_basis_general(L, Nup=6, pauli=0, kblock=(m["T"], 2))
    static = xxz_static(L, Delta=0.5)
    Q = quantum_LinearOperator(static, basis=B, dtype=np.complex128, check_symm=False, check_herm=False,
                               check_pcon=False)
    Hd = _oracle_H(B, static).toarray()
    rng = np.random.default_rng(8)
    V = (rng.normal(size=(B.Ns, 3)) + 1j * rng.normal(size=(B.Ns, 3))).astype(np.complex128)
    E = expm_multiply_parallel(Q, a=-0.5j, dtype=np.complex128)
    print(f"case Ns={B.Ns} a=-0.5i", flush=True)
    np.testing.assert_allclose(E.dot(V[:, 0]), expm(-0.5j * Hd) @ V[:, 0], atol=1e-11, rtol=0)
    for a in (-1.3j, 0.2):
        print(f"case set_a({a})", flush=True)
        E.set_a(a)
        work = np.zeros(2 * V.size, np.complex128)
        np.testing.assert_allclose(E.dot(V, work_array=work), expm(a * Hd) @ V, atol=1e-10, rtol=0)
    W = V.copy()
    out = E.dot(W, overwrite_v=True)
    assert out is W
    np.testing.assert_allclose(W, expm(0.2 * Hd) @ V, atol=1e-10, rtol=0)
