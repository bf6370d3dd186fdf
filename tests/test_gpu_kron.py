"""Species-factorised H.v (H_up (x) 1 + 1 (x) H_down + static diagonal) on
spinful fermion sectors: product, 1-norm and device time stepping against the
oracle's matrix of the same static list, and against the CSR form."""

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.linalg import expm

from oracle import qsx_oracle as O
from tests.models import flat_terms, hubbard_static

pytestmark = pytest.mark.gpu

_KW = dict(check_symm=False, check_herm=False, check_pcon=False)


def _adv(term, L):
    opstr, indx, J = term
    i = opstr.index("|")
    return (opstr.replace("|", ""), tuple(indx[:i]) + tuple(j + L for j in indx[i:]), J)


def _oracle_H(B, static, L):
    """Oracle CSR of the static list on the basis states (no symmetry maps)."""
    g = O.Group(B.N, [], [], [], fermion=True)
    b = B.states.astype(np.uint64)
    return O.make_csr(b, np.ones(len(b), np.int64), g, [_adv(t, L) for t in flat_terms(static)], np.complex128,
                      fermion=True)


def _model(L, complex_hop=False, V=0.7, mu=0.3, U=2.5):
    """Hubbard chain with on-site U, chemical potential and a
    nearest-neighbour density coupling between the species (diagonal terms on
    both species); complex_hop gives the hopping a Peierls phase."""
    ph = np.exp(0.37j) if complex_hop else 1.0
    hop_r = [[-1.0 * ph, i, (i + 1) % L] for i in range(L)]
    hop_l = [[+1.0 * np.conj(ph), i, (i + 1) % L] for i in range(L)]
    st = [["+-|", hop_r], ["-+|", hop_l], ["|+-", hop_r], ["|-+", hop_l]]
    st += [["n|n", [[U, i, i] for i in range(L)]], ["n|", [[-mu, i] for i in range(L)]],
           ["|n", [[-mu, i] for i in range(L)]], ["n|n", [[V, i, (i + 1) % L] for i in range(L)]],
           ["|zz", [[0.2, i, (i + 2) % L] for i in range(L)]]]
    return st


def _vec(n, dtype, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=n)
    if np.dtype(dtype).kind == "c":
        v = v + 1j * rng.normal(size=n)
    return v.astype(dtype)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("dtype,cplx,tol", [(np.float64, False, 1e-12), (np.complex128, True, 1e-12),
                                            (np.float32, False, 3e-6), (np.complex64, True, 3e-6)])
def test_kron_product_matches_oracle(dtype, cplx, tol, mode):
    """Both kernel forms (1: one pass, 2: row pass + column-strip pass)."""
    from quspin_amd import _lib as L
    from quspin_amd.basis import spinful_fermion_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.operators._kron import KronOp
    Lh = 6
    B = spinful_fermion_basis_general(Lh, Nf=(3, 2))
    st = _model(Lh, complex_hop=cplx)
    ref = _oracle_H(B, st, Lh)
    Q = quantum_LinearOperator(st, basis=B, dtype=np.complex128 if cplx else np.float64, **_KW)
    K = KronOp(Q, dtype)
    K.set_mode(mode)
    print(f"case: L={Lh} Nf=(3,2) dtype={np.dtype(dtype).name} mode={mode} Nu={K.Nu} Nd={K.Nd}", flush=True)
    assert (K.Nu, K.Nd) == (20, 15) and K.sdiag is not None
    np.testing.assert_allclose(K.to_scipy().toarray(), ref.toarray(), atol=1e-6 if tol > 1e-9 else 1e-13)
    x = _vec(B.Ns, dtype, 1)
    xd = _vec(B.Ns, np.float64, 2)
    y0 = _vec(B.Ns, dtype, 3)
    a = (0.3 - 0.8j) if cplx else -1.7
    xs = 0.45
    X, XD = L.to_device(x), L.to_device(xd)
    Y = L.to_device(y0.copy())
    K.apply(X, Y, a=a, beta=1, xdiag=XD, xs=xs)
    expect = y0.astype(np.complex128) + a * (ref @ x.astype(np.complex128) + xs * xd * x)
    got = L.to_host(Y).astype(np.complex128)
    scale = np.abs(expect).max()
    np.testing.assert_allclose(got, expect if cplx else expect.real, rtol=0, atol=tol * scale)
    Y2 = L.empty(B.Ns, dtype)
    K.apply(X, Y2)
    np.testing.assert_allclose(L.to_host(Y2), ref @ x if cplx else (ref @ x).real, rtol=0,
                               atol=tol * np.abs(ref @ x).max())


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 3e-6)])
def test_kron_two_pass_odd_up_rows(dtype, tol):
    """The row pass stages two up rows per block: an odd number of up rows
    (Nf=(3,2) on 7 sites: 35 x 21) leaves the last group with one row."""
    from quspin_amd import _lib as L
    from quspin_amd.basis import spinful_fermion_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.operators._kron import KronOp
    Lh = 7
    B = spinful_fermion_basis_general(Lh, Nf=(3, 2))
    st = _model(Lh, complex_hop=True)
    K = KronOp(quantum_LinearOperator(st, basis=B, dtype=np.complex128, **_KW), dtype)
    K.set_mode(2)
    print(f"case: L={Lh} Nf=(3,2) dtype={np.dtype(dtype).name} Nu={K.Nu} Nd={K.Nd}", flush=True)
    assert (K.Nu, K.Nd) == (35, 21)
    ref = _oracle_H(B, st, Lh)
    x = _vec(B.Ns, dtype, 5)
    Y = L.empty(B.Ns, dtype)
    K.apply(L.to_device(x), Y)
    e = ref @ x.astype(np.complex128)
    np.testing.assert_allclose(L.to_host(Y), e, rtol=0, atol=tol * np.abs(e).max())


@pytest.mark.parametrize("mode", [1, 2])
def test_kron_unstaged_rows_match_oracle(mode):
    """A down sector wider than the LDS row stage (C(15,7) complex128 rows:
    103 KB) takes the direct-read kernels."""
    from quspin_amd import _lib as L
    from quspin_amd.basis import spinful_fermion_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.operators._kron import KronOp
    Lh = 15
    B = spinful_fermion_basis_general(Lh, Nf=(1, 7))
    st = _model(Lh, complex_hop=True)
    K = KronOp(quantum_LinearOperator(st, basis=B, dtype=np.complex128, **_KW), np.complex128)
    K.set_mode(mode)
    print(f"case: L={Lh} Nf=(1,7) Ns={B.Ns} mode={mode}", flush=True)
    assert K.Nd * 16 > 64 * 1024
    ref = _oracle_H(B, st, Lh)
    x = _vec(B.Ns, np.complex128, 4)
    Y = L.empty(B.Ns, np.complex128)
    K.apply(L.to_device(x), Y)
    e = ref @ x
    np.testing.assert_allclose(L.to_host(Y), e, rtol=0, atol=1e-12 * np.abs(e).max())


def test_kron_one_norm_matches_oracle():
    from quspin_amd import _lib as L
    from quspin_amd.basis import spinful_fermion_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.operators._kron import KronOp
    Lh = 6
    B = spinful_fermion_basis_general(Lh, Nf=(3, 3))
    st = _model(Lh, complex_hop=True)
    ref = _oracle_H(B, st, Lh)
    K = KronOp(quantum_LinearOperator(st, basis=B, dtype=np.complex128, **_KW), np.complex128)
    xd = _vec(B.Ns, np.float64, 5)
    for shift, xs in ((0.0, 0.0), (-1.3 + 0.2j, 0.0), (0.4, 2.5), (-2.0, -1.5)):
        got = K.one_norm(shift, L.to_device(xd) if xs else None, xs)
        expect = O.csr_1_norm(ref + xs * sp.diags(xd), shift)
        assert abs(got - expect) <= 1e-12 * expect, (shift, xs, got, expect)
    assert abs(K.trace() - O.csr_trace(ref)) < 1e-10


def test_kron_rejects_operators_that_do_not_factorise():
    from quspin_amd.basis import spinful_fermion_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.operators._kron import KronOp
    Lh = 4
    B = spinful_fermion_basis_general(Lh, Nf=(2, 2))
    spin_flip = [["+-|-+", [[0.5, i, (i + 1) % Lh, i, (i + 1) % Lh] for i in range(Lh)]]]
    with pytest.raises(TypeError, match="both species"):
        KronOp(quantum_LinearOperator(spin_flip, basis=B, dtype=np.float64, **_KW))
    T = (np.arange(Lh) + 1) % Lh
    Bk = spinful_fermion_basis_general(Lh, Nf=(2, 2), kblock=(T, 0))
    with pytest.raises(TypeError, match="symmetry"):
        KronOp(quantum_LinearOperator(hubbard_static(Lh), basis=Bk, dtype=np.float64, **_KW))
    Bs = spinful_fermion_basis_general(Lh, Nf=[(2, 2), (1, 3)])
    with pytest.raises(TypeError, match="sector"):
        KronOp(quantum_LinearOperator(hubbard_static(Lh), basis=Bs, dtype=np.float64, **_KW))


def _driven(L_, dtype=np.complex128):
    from quspin_amd.basis import spinful_fermion_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    B = spinful_fermion_basis_general(L_, Nf=(L_ // 2, L_ // 2))
    hop = quantum_LinearOperator(hubbard_static(L_, U=0.0), basis=B, dtype=dtype, **_KW)
    D = quantum_LinearOperator([["n|n", [[1.0, i, i] for i in range(L_)]]], basis=B, dtype=np.float64, **_KW)
    f = lambda t: 4.0 + 2.0 * np.cos(3.0 * t)  # noqa: E731
    phase = lambda t: (3.0 * t) % (2 * np.pi)  # noqa: E731
    return B, hop, D, f, phase


def test_kron_stepping_matches_dense_midpoint_exponentials():
    from quspin_amd import _lib as L
    from quspin_amd.tools.learned_step import DrivenEvolution
    B, hop, D, f, phase = _driven(6)
    dt = 0.05
    ev = DrivenEvolution(hop, D, f, dt, dtype=np.complex128, phase=phase, form="kron")
    assert ev.kron is not None and ev.kron.dtype == np.float64  # real hopping: real factors
    Hd = _oracle_H(B, hubbard_static(6, U=0.0), 6).toarray()
    Dd = np.diag(L.to_host(ev.D))
    rng = np.random.default_rng(5)
    psi0 = rng.normal(size=B.Ns) + 1j * rng.normal(size=B.Ns)
    psi0 /= np.linalg.norm(psi0)
    psi = L.to_device(psi0.copy())
    ev.run(psi, 0.3, 12, learned=False)
    ref = psi0.copy()
    for k in range(12):
        t = 0.3 + (k + 0.5) * dt
        ref = expm(-1j * dt * (Hd + f(t) * Dd)) @ ref
    np.testing.assert_allclose(L.to_host(psi), ref, rtol=0, atol=1e-12)
    assert ev.stats["exact_steps"] == 12 and ev.stats["norm_passes"] == 12


def test_kron_graph_replay_equals_csr_form():
    """Fixed partitions through captured step sequences: every replay of the
    factorised form equals the CSR form on the same records (FP64), and the
    FP32-term / FP64-state form stays within FP32 accuracy of it."""
    from quspin_amd import _lib as L
    from quspin_amd.tools.learned_step import DrivenEvolution, PartitionMLP
    B, hop, D, f, phase = _driven(8)
    net = PartitionMLP(seed=0)
    # near-constant prediction m ~ 40, s ~ 1: the device truncation test ends
    # each step early, the tail check passes
    net.set_normalisation(np.zeros((2, 4)), np.array([[40.0, 1.0], [40.0, 1.0]]))
    rng = np.random.default_rng(7)
    psi0 = rng.normal(size=B.Ns) + 1j * rng.normal(size=B.Ns)
    psi0 /= np.linalg.norm(psi0)
    out = {}
    for form, dt_, sdt, mode in (("csr", np.complex128, None, 0), ("kron", np.complex128, None, 1),
                                 ("kron", np.complex128, None, 2), ("kron", np.complex64, np.complex128, 1),
                                 ("kron", np.complex64, np.complex128, 2)):
        print(f"case: form={form} dtype={np.dtype(dt_).name} mode={mode}", flush=True)
        ev = DrivenEvolution(hop, D, f, 0.02, dtype=dt_, phase=phase, predictor=net, m_margin=0, state_dtype=sdt,
                             form=form)
        if mode:
            ev.kron.set_mode(mode)
        reps = []
        for rep in range(3):
            p = L.to_device(psi0.astype(sdt or dt_))
            ev.run(p, 0.0, 20, learned=True, block=10)
            L.torch().cuda.synchronize()
            reps.append(L.to_host(p).astype(np.complex128))
            assert ev.stats["fallback_blocks"] == 0
        for r in reps[1:]:
            np.testing.assert_array_equal(r, reps[0])
        out[(form, np.dtype(dt_).name, mode)] = reps[0]
    for mode in (1, 2):
        np.testing.assert_allclose(out[("kron", "complex128", mode)], out[("csr", "complex128", 0)], rtol=0,
                                   atol=1e-13)
        err = np.linalg.norm(out[("kron", "complex64", mode)] - out[("csr", "complex128", 0)])
        print(f"kron mode {mode} complex64 terms + complex128 state vs FP64: {err:.3e}", flush=True)
        assert err < 1e-5
