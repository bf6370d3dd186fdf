"""The benchmark workloads' code paths at sizes the oracle finishes in
seconds, the partition choice of expm_multiply_parallel, the
reference-shaped single-term export, and graph captures across operators
that are freed and rebuilt."""

import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import qsx_oracle as O
from tests.models import chain_maps, flat_terms, xxz_static

pytestmark = pytest.mark.gpu

KW = dict(check_symm=False, check_herm=False, check_pcon=False)


def _group(basis, fermion=False):
    return O.Group(basis.N, list(basis._maps) if len(basis._pers) else [], list(basis._pers), list(basis._qs),
                   fermion=fermion)


def _oracle_csr(basis, static, dtype=np.complex128):
    b = basis.states.astype(np.uint64)
    n = basis._n.astype(np.int64) if len(basis._pers) else np.ones(len(b), np.int64)
    return O.make_csr(b, n, _group(basis), flat_terms(static), dtype)


def test_xxz_L12_sector_dot_lanczos_expm():
    """XXZ L=12, J=1, Delta=0.5, Sz=0 (924 states): hamiltonian.dot (CSR)
    on a seeded complex128 unit vector, the 50-step lanczos_full ground
    energy and expm_multiply_parallel to t=1 (a = -i), each against the
    oracle's matrix (hamiltonian_core.py:723-772; _lanczos_utils.py:77-227;
    EMP.html:267-374)."""
    from scipy.linalg import expm
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian
    from quspin_amd.tools.evolution import expm_multiply_parallel
    from quspin_amd.tools.lanczos import lanczos_full
    L = 12
    B = spin_basis_general(L, Nup=L // 2, pauli=0)
    assert B.Ns == 924
    static = xxz_static(L, J=1.0, Delta=0.5)
    H = hamiltonian(static, [], basis=B, dtype=np.float64, **KW)
    A = _oracle_csr(B, static)
    rng = np.random.default_rng(0)
    v = rng.normal(size=B.Ns) + 1j * rng.normal(size=B.Ns)
    v /= np.linalg.norm(v)
    print(f"case dot Ns={B.Ns}", flush=True)
    np.testing.assert_allclose(H.dot(v), A @ v, rtol=0, atol=1e-13)
    Hd = A.toarray()
    E_exact = np.linalg.eigvalsh(Hd)[0]
    print("case lanczos m=50", flush=True)
    E, V, Q = lanczos_full(H, v.real / np.linalg.norm(v.real), 50)
    E = np.asarray(E.cpu() if hasattr(E, "cpu") else E)
    assert abs(E[0] - E_exact) < 1e-10
    print("case expm t=1", flush=True)
    U = expm_multiply_parallel(H.tocsr(), a=-1j)
    np.testing.assert_allclose(U.dot(v), expm(-1j * Hd) @ v, rtol=0, atol=1e-12)


def test_xxz_fields_complex64_matrix_free_expm_vs_fp64_oracle():
    """The precision workload's path at L=18: XXZ Delta=1, h_i ~ U[-3,3]
    (seed 0), Sz=0 without spatial symmetry (48,620 states), Neel start;
    20 steps of expm_multiply_parallel(quantum_LinearOperator(float32),
    a=-0.1i, dtype=complex64) -- the matrix-free Taylor loop with the
    direct-rank lookup -- against the oracle's FP64 Taylor loop on the CSR of
    the same operator.  Bounds: state error <= 1e-6 and max |<Sz_i>| error
    <= 1e-6 after every step (FP32 vectors, tolerance eps32/2 per step).
    Also one complex64 H.v on that sector against the oracle, 2e-6
    relative."""
    import torch
    from quspin_amd import _lib as Lb
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.tools.evolution import expm_multiply_parallel
    L, dt, steps = 18, 0.1, 20
    h = np.random.default_rng(0).uniform(-3, 3, L)
    static = xxz_static(L, Delta=1.0, h=h)
    B = spin_basis_general(L, Nup=L // 2, pauli=0)
    Q = quantum_LinearOperator(static, basis=B, dtype=np.float32, **KW)
    A = _oracle_csr(B, static).real.tocsr()
    rng = np.random.default_rng(1)
    x = (rng.normal(size=B.Ns) + 1j * rng.normal(size=B.Ns)).astype(np.complex64)
    print(f"case H.v complex64 Ns={B.Ns}", flush=True)
    y = Lb.to_host(Q.dot(Lb.to_device(x)))
    ref = A @ x.astype(np.complex128)
    assert np.abs(y - ref).max() <= 2e-6 * np.abs(ref).max()
    E = expm_multiply_parallel(Q, a=-1j * dt, dtype=np.complex64)
    i0 = B.index("10" * (L // 2))
    psi = torch.zeros(B.Ns, dtype=torch.complex64, device="cuda")
    psi[i0] = 1
    work = torch.empty(2 * B.Ns, dtype=torch.complex64, device="cuda")
    ref = np.zeros(B.Ns, np.complex128)
    ref[i0] = 1
    bits = ((B.states.astype(np.int64)[:, None] >> (L - 1 - np.arange(L))[None, :]) & 1) - 0.5
    for k in range(steps):
        psi = E.dot(psi, work_array=work, overwrite_v=True)
        ref, _ = O.expm_multiply(A, ref, a=-1j * dt, dtype=np.complex128)
        p = Lb.to_host(psi).astype(np.complex128)
        err = np.linalg.norm(p - ref)
        sz = np.abs((np.abs(p) ** 2) @ bits - (np.abs(ref) ** 2) @ bits).max()
        print(f"step {k + 1}: state err {err:.2e}, <Sz_i> err {sz:.2e}", flush=True)
        assert err <= 1e-6 and sz <= 1e-6


def _partition_case(E, A, a):
    """mu, ||A - mu I||_1 and (m*, s) of a device expm_multiply_parallel
    against the oracle's reference selection on the matrix A."""
    N = A.shape[0]
    mu = O.csr_trace(A) / N
    nrm = O.csr_1_norm(A, -mu)
    assert abs(complex(E._mu) - mu) <= 1e-12 * max(1.0, abs(mu))
    assert abs(E._A_1_norm - nrm) <= 1e-12 * nrm
    tol = np.finfo(E._dtype).eps / 2
    part = O.fragment_3_1_partition(A, a, mu, nrm, tol)
    return part, abs(a) * nrm


def test_expm_partition_mu_and_norm_equal_reference_selection():
    """ExpmMultiplyParallel's mu = tr(A)/N, ||A - mu I||_1 and (m*, s)
    (EMP.html:152-203, 377-385; scipy _expm_multiply.py:503-558) equal the
    oracle's reference selection.  theta-branch inputs (|a| ||A - mu I||_1
    <= 63.36, the deterministic table branch) on a CSR operator, a
    matrix-free operator of a real sector and one of a complex k-sector; and
    one onenormest-branch input on a random CSR matrix, with the global
    numpy seed set before each selection (onenormest draws from it)."""
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.tools.evolution import expm_multiply_parallel
    L = 14
    m = chain_maps(L)
    static = xxz_static(L, Delta=0.7)
    cases = [("real k=0 p z", {"kblock": (m["T"], 0), "pblock": (m["P"], 0), "zblock": (m["Z"], 0)}, np.float64),
             ("complex k=3", {"kblock": (m["T"], 3)}, np.complex128)]
    for name, blocks, dt in cases:
        B = spin_basis_general(L, Nup=L // 2, pauli=0, **blocks)
        A = _oracle_csr(B, static)
        if dt == np.float64:
            A = A.real.tocsr()
        for a in (-0.5j, -2.0j, 1.5):
            for form in ("matrix-free", "csr"):
                print(f"case {name} Ns={B.Ns} a={a} {form}", flush=True)
                op = quantum_LinearOperator(static, basis=B, dtype=dt, **KW) if form == "matrix-free" else A
                E = expm_multiply_parallel(op, a=a)
                part, an = _partition_case(E, A, a)
                assert an <= 63.36, an
                assert E.partition == part, (E.partition, part)
    rng = np.random.RandomState(4)
    N = 400
    A = (sp.random(N, N, density=0.03, random_state=rng) - 0.5 * sp.eye(N)).tocsr()
    a = -20.0
    np.random.seed(11)
    E = expm_multiply_parallel(A, a=a)
    np.random.seed(11)
    part, an = _partition_case(E, A, a)
    print(f"case onenormest branch |a| ||A - mu||_1 = {an:.1f}: {E.partition} vs {part}", flush=True)
    assert an > 63.36
    assert E.partition == part


def test_inplace_op_export_all_flag_pairs_vs_oracle():
    """qsx_inplace_op, the export that keeps _core.inplace_op's argument list
    (base_general.py:395-408: conjugated before transposed), called through
    ctypes once per term for all four (conjugated, transposed) pairs on a
    complex k-sector, against the oracle's per-term scatter (base.py:14-23,
    125-162)."""
    from quspin_amd import _lib as Lb
    from quspin_amd.basis import spin_basis_general
    L = 10
    m = chain_maps(L)
    B = spin_basis_general(L, Nup=L // 2, pauli=0, kblock=(m["T"], 2), zblock=(m["Z"], 1))
    b = B.states.astype(np.uint64)
    n = B._n.astype(np.int64)
    g = _group(B)
    terms = [("+-", (0, 1), 0.3 + 0.7j), ("-+", (4, 2), -0.45 + 0.1j), ("zz", (1, 6), 0.8), ("z", (3,), -1.1 + 0.2j),
             ("+-", (2, 7), 1.25)]
    rng = np.random.default_rng(3)
    nv = 2
    vin = (rng.normal(size=(B.Ns, nv)) + 1j * rng.normal(size=(B.Ns, nv))).astype(np.complex128)
    lib = Lb.load()
    for conj in (0, 1):
        for trans in (0, 1):
            print(f"case conjugated={conj} transposed={trans}", flush=True)
            vin_d = Lb.to_device(vin)
            vout = (rng.normal(size=(B.Ns, nv)) + 1j * rng.normal(size=(B.Ns, nv))).astype(np.complex128)
            vout_d = Lb.to_device(vout)
            ref = vout.copy()
            for opstr, indx, J in terms:
                idx = np.asarray(indx, np.int32)
                rc = lib.qsx_inplace_op(B._core.h, Lb.ptr(vin_d), Lb.ptr(vout_d), Lb.C128, nv, conj, trans,
                                        opstr.encode(), idx.ctypes.data_as(C.c_void_p), len(opstr), J.real,
                                        complex(J).imag, *B._dev_args(), Lb.stream_handle())
                Lb.check(rc, "qsx_inplace_op")
                O.inplace_op(vin, b, n, g, [(opstr, indx, J)], transposed=bool(trans), conjugated=bool(conj),
                             v_out=ref)
            np.testing.assert_allclose(Lb.to_host(vout_d), ref, rtol=0, atol=1e-12)


def test_captures_follow_rebuilt_and_changed_operators():
    """Captured launch sequences never outlive the operator state they were
    recorded with: same-shape operators built, freed and rebuilt in one
    process (so the allocator hands back recycled addresses), a diagonal
    replaced by set_diagonal between two captured Lanczos runs, and captured
    driven-evolution blocks on rebuilt operators; every result against the
    oracle (dense eigenvalues, dense midpoint exponentials)."""
    import gc
    import torch
    from scipy.linalg import expm
    from quspin_amd import _lib as Lb
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.tools.lanczos import lanczos_full
    L = 12
    m = chain_maps(L)
    B = spin_basis_general(L, Nup=L // 2, pauli=0, kblock=(m["T"], 0), pblock=(m["P"], 0), zblock=(m["Z"], 0))
    rng = np.random.default_rng(6)
    v0 = rng.normal(size=B.Ns)
    v0d = torch.from_numpy(v0).cuda()
    for rep, Delta in enumerate((0.5, 1.7, 0.5, -0.3)):
        static = xxz_static(L, Delta=Delta)
        Q = quantum_LinearOperator(static, basis=B, dtype=np.float64, **KW)
        Hd = _oracle_csr(B, static).toarray().real
        E, _, _ = lanczos_full(Q, v0d, min(40, B.Ns - 1))
        E = np.asarray(E.cpu() if hasattr(E, "cpu") else E)
        print(f"rebuild {rep}: Delta={Delta} E0 {E[0]:.12f}", flush=True)
        assert abs(E[0] - np.linalg.eigvalsh(Hd)[0]) < 1e-9
        diag = rng.uniform(-2, 2, size=B.Ns)
        Q.set_diagonal(diag)
        E, _, _ = lanczos_full(Q, v0d, min(40, B.Ns - 1))
        E = np.asarray(E.cpu() if hasattr(E, "cpu") else E)
        # set_diagonal replaces the diagonal terms' part only; off-diagonal
        # terms that fold onto the diagonal of a symmetric sector remain
        # (quantum_LinearOperator_core.py:215-234, 520-540)
        Hd2 = _oracle_csr(B, [st for st in static if st[0] != "zz"]).toarray().real + np.diag(diag)
        print(f"rebuild {rep}: after set_diagonal E0 {E[0]:.12f}", flush=True)
        assert abs(E[0] - np.linalg.eigvalsh(Hd2)[0]) < 1e-9
        del Q
        gc.collect()
    # captured driven steps (learned-step block path) on rebuilt operators
    from quspin_amd.basis import spinful_fermion_basis_general
    from quspin_amd.tools.learned_step import DrivenEvolution
    Lh = 6
    f = lambda t: 4.0 + 2.0 * np.cos(3.0 * t)  # noqa: E731
    phase = lambda t: (3.0 * t) % (2 * np.pi)  # noqa: E731
    Bh = spinful_fermion_basis_general(Lh, Nf=(3, 3))
    psi0 = rng.normal(size=Bh.Ns) + 1j * rng.normal(size=Bh.Ns)
    psi0 /= np.linalg.norm(psi0)
    for rep, thop in enumerate((1.0, 0.6, 1.0)):
        hr = [[-thop, i, (i + 1) % Lh] for i in range(Lh)]
        hl = [[+thop, i, (i + 1) % Lh] for i in range(Lh)]
        hop_st = [["+-|", hr], ["-+|", hl], ["|+-", hr], ["|-+", hl]]
        hop = quantum_LinearOperator(hop_st, basis=Bh, dtype=np.complex128, **KW)
        D = quantum_LinearOperator([["n|n", [[1.0, i, i] for i in range(Lh)]]], basis=Bh, dtype=np.float64, **KW)
        Hh = _dense(hop, Bh.Ns)
        Dm = _dense(D, Bh.Ns)
        ref = psi0.copy()
        for s_ in range(8):
            ref = expm(-1j * 0.05 * (Hh + f((s_ + 0.5) * 0.05) * Dm)) @ ref
        for form in ("csr", "kron"):
            ev = DrivenEvolution(hop, D, f, 0.05, dtype=np.complex128, phase=phase, form=form,
                                 predictor=_FixedPartition(24, 2))
            psi = Lb.to_device(psi0.copy())
            ev.run(psi, 0.0, 8, learned=True, block=8, graph=True)
            got = Lb.to_host(psi)
            print(f"driven rebuild {rep} {form}: {ev.stats}, max err {np.abs(got - ref).max():.2e}", flush=True)
            assert ev.stats["learned_steps"] == 8 and ev.stats["fallback_blocks"] == 0
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10)
            del ev
        del hop, D
        gc.collect()


class _FixedPartition:
    """Stand-in predictor: writes one fixed (m, s) into every step record, so
    that the captured block path runs without a trained network."""

    def __init__(self, m, s):
        self.m, self.s = m, s

    def write_partitions(self, X, records, *args):
        from quspin_amd import _lib as Lb
        from quspin_amd.tools.learned_step import RECORD
        r = Lb.to_host(records).view(RECORD).copy()
        r["m"], r["s"] = self.m, self.s
        records.copy_(Lb.torch().from_numpy(r.view(np.uint8)))


def _dense(op, n):
    """Dense matrix of a device operator, column by column through dot."""
    eye = np.eye(n, dtype=np.complex128)
    return np.stack([np.asarray(op.dot(eye[:, j])) for j in range(n)], axis=1)
