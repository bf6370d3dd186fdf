"""Fused matrix-free H.v (quantum_LinearOperator), device CSR
(hamiltonian.dot), Op and inplace_Op against the oracle."""

import numpy as np
import pytest

from oracle import qsx_oracle as O
from tests.models import chain_maps, flat_terms, hubbard_static, ladder_maps, xxz_static

pytestmark = pytest.mark.gpu


def _oracle_H(basis, static, fermion=False, spinful_L=None):
    items = [(k, basis._maps_dict[k], None, None) for k in basis._maps_dict]
    g = O.Group(basis.N, list(basis._maps) if len(basis._pers) else [], list(basis._pers), list(basis._qs),
                fermion=fermion)
    terms = flat_terms(static)
    if spinful_L is not None:
        terms = [_adv(t, spinful_L) for t in terms]
    b = basis.states.astype(np.uint64)
    n = basis._n.astype(np.int64) if len(basis._pers) else np.ones(len(b), np.int64)
    return O.make_csr(b, n, g, terms, np.complex128, fermion=fermion)


def _adv(term, L):
    opstr, indx, J = term
    i = opstr.index("|")
    return (opstr.replace("|", ""), tuple(indx[:i]) + tuple(j + L for j in indx[i:]), J)


def _vec(n, dtype, seed=0, nvec=None):
    rng = np.random.default_rng(seed)
    shape = (n,) if nvec is None else (n, nvec)
    v = rng.normal(size=shape)
    if np.dtype(dtype).kind == "c":
        v = v + 1j * rng.normal(size=shape)
    return v.astype(dtype)


def _tol(dtype):
    return 1e-12 if np.dtype(dtype) in (np.float64, np.complex128) else 2e-5


CASES = []
for L, k, p, z in [(10, 0, 0, 0), (10, 0, 1, None), (12, 3, None, None), (12, 6, 1, 1), (11, 2, None, None)]:
    CASES.append((L, k, p, z))


@pytest.mark.parametrize("L,k,p,z", CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.complex64])
def test_qlo_dot_vs_oracle(L, k, p, z, dtype):
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    m = chain_maps(L)
    blocks = {"kblock": (m["T"], k)}
    if p is not None:
        blocks["pblock"] = (m["P"], p)
    if z is not None:
        blocks["zblock"] = (m["Z"], z)
    print(f"case L={L} k={k} p={p} z={z} dtype={np.dtype(dtype)}", flush=True)
    B = spin_basis_general(L, Nup=L // 2, pauli=0, **blocks)
    static = xxz_static(L, J=1.0, Delta=0.7, h=None)
    complex_chars = (2 * k) % L != 0
    H = _oracle_H(B, static)
    if np.dtype(dtype).kind != "c" and complex_chars:
        with pytest.raises(TypeError):
            quantum_LinearOperator(static, basis=B, dtype=dtype, check_symm=False, check_herm=False,
                                   check_pcon=False).dot(_vec(B.Ns, dtype))
        return
    Q = quantum_LinearOperator(static, basis=B, dtype=dtype, check_symm=False, check_herm=False, check_pcon=False)
    v = _vec(B.Ns, dtype)
    ref = H @ v.astype(np.complex128)
    got = Q.dot(v)
    np.testing.assert_allclose(got, ref if np.dtype(dtype).kind == "c" else ref.real,
                               atol=_tol(dtype) * max(1, np.abs(ref).max()), rtol=0)
    # transposed / conjugated / Hermitian conjugate
    for op, R in [(Q.T, H.T), (Q.conj(copy=True), H.conj()), (Q.H, H.conj().T)]:
        got = op.dot(v)
        ref = R @ v.astype(np.complex128)
        np.testing.assert_allclose(got, ref if np.dtype(dtype).kind == "c" else ref.real,
                                   atol=_tol(dtype) * max(1, np.abs(ref).max()), rtol=0)


def test_qlo_batch_and_out_semantics():
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    L = 12
    m = chain_maps(L)
    B = spin_basis_general(L, Nup=6, pauli=0, kblock=(m["T"], 1))
    static = xxz_static(L, Delta=0.5)
    H = _oracle_H(B, static)
    Q = quantum_LinearOperator(static, basis=B, dtype=np.complex128, check_symm=False, check_herm=False,
                               check_pcon=False)
    V = _vec(B.Ns, np.complex128, nvec=5)
    np.testing.assert_allclose(Q.dot(V), H @ V, atol=1e-12)
    # dot(out=, a=): diagonal not scaled by a (quantum_LinearOperator_core.py:332-342)
    v = V[:, 0].copy()
    out = np.ones_like(v)
    Q.dot(v, out=out, a=2.0)
    Hd = H.toarray()
    D = np.diag(np.diag(Hd))
    ref = D @ v + 2.0 * ((Hd - D) @ v)
    np.testing.assert_allclose(out, ref, atol=1e-12)
    # no diagonal: accumulates into out
    Q2 = quantum_LinearOperator(static[:2], basis=B, dtype=np.complex128, check_symm=False, check_herm=False,
                                check_pcon=False)
    H2 = _oracle_H(B, static[:2])
    out = np.ones_like(v)
    Q2.dot(v, out=out, a=1.5)
    np.testing.assert_allclose(out, 1.0 + 1.5 * (H2 @ v), atol=1e-12)


def test_hamiltonian_vs_qlo_ladder():
    """test_inplace_op.py: 3x2 ladder with tx,ty,px,py,z; H.dot == H_op.dot,
    also transposed; full, pcon and symmetric bases."""
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian, quantum_LinearOperator
    Lx, Ly = 3, 2
    N = Lx * Ly
    m = ladder_maps(Lx, Ly)
    bases = [spin_basis_general(N, pauli=False), spin_basis_general(N, pauli=False, m=0.0),
             spin_basis_general(N, pauli=False, m=0.0, tx=(m["tx"], 0), ty=(m["ty"], 0), px=(m["px"], 0),
                                py=(m["py"], 0), z=(m["z"], 0))]
    tx, ty = m["tx"], m["ty"]
    Jzz = [[0.0, i, tx[i]] for i in range(N)] + [[0.0, i, ty[i]] for i in range(N)]
    Jxy = [[0.5, i, tx[i]] for i in range(N)] + [[0.5, i, ty[i]] for i in range(N)]
    static = [["+-", Jxy], ["-+", Jxy], ["zz", Jzz]]
    rng = np.random.default_rng(0)
    for b in bases:
        for dt in (np.float64, np.complex128):
            H = hamiltonian(static, [], basis=b, dtype=dt, check_symm=False, check_herm=False, check_pcon=False)
            Hop = quantum_LinearOperator(static, basis=b, dtype=dt, check_symm=False, check_herm=False,
                                         check_pcon=False)
            ref = _oracle_H(b, static)
            for _ in range(3):
                v = rng.uniform(-1, 1, b.Ns) + 1j * rng.uniform(-1, 1, b.Ns)
                v /= np.linalg.norm(v)
                np.testing.assert_allclose(H.dot(v), Hop.dot(v), atol=100 * np.finfo(dt).eps)
                np.testing.assert_allclose(H.dot(v), ref @ v, atol=100 * np.finfo(dt).eps)
                np.testing.assert_allclose(H.T.dot(v), Hop.T.dot(v), atol=100 * np.finfo(dt).eps)
            np.testing.assert_allclose(H.toarray(), ref.toarray(), atol=1e-13)


def test_op_coo_vs_oracle():
    from quspin_amd.basis import spin_basis_general
    L = 10
    m = chain_maps(L)
    B = spin_basis_general(L, Nup=5, pauli=0, kblock=(m["T"], 2), zblock=(m["Z"], 1))
    g = O.Group(L, list(B._maps), list(B._pers), list(B._qs))
    for opstr, indx, J in [("+-", (0, 1), 0.5), ("zz", (2, 5), 1.3), ("xy", (3, 7), 0.25), ("-+", (9, 0), 1.0)]:
        ME, row, col = B.Op(opstr, indx, J, np.complex128)
        me_o, r_o, c_o = O.op_coo(B.states.astype(np.uint64), B._n.astype(np.int64), g, opstr, indx, J)
        np.testing.assert_array_equal(col, c_o)
        np.testing.assert_array_equal(row, r_o)
        np.testing.assert_allclose(ME, me_o, atol=1e-14)


def test_inplace_op_vs_oracle_flags():
    from quspin_amd.basis import spin_basis_general
    L = 10
    m = chain_maps(L)
    B = spin_basis_general(L, Nup=5, pauli=1, kblock=(m["T"], 3))
    g = O.Group(L, list(B._maps), list(B._pers), list(B._qs))
    ops = [["+-", (0, 1), 0.5 + 0.25j], ["zz", (2, 3), 1.0], ["y-", (4, 6), 0.3]]
    v = _vec(B.Ns, np.complex128, nvec=3)
    b = B.states.astype(np.uint64)
    n = B._n.astype(np.int64)
    for T in (False, True):
        for Cj in (False, True):
            got = B.inplace_Op(v, ops, np.complex128, transposed=T, conjugated=Cj, a=0.7)
            scaled = [(o, i, J * (1 << len(o.replace("I", "")))) for o, i, J in ops]
            ref = O.inplace_op(v, b, n, g, scaled, transposed=T, conjugated=Cj, a=0.7)
            np.testing.assert_allclose(got, ref, atol=1e-12)


def test_pauli_scaling():
    """pauli in {0, 1, -1} (spin.py:216-276, test_pauli.py)."""
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian
    L = 6
    st = [["x", [[1.0, 0]]], ["+-", [[1.0, 1, 2]]], ["zz", [[1.0, 3, 4]]]]
    mats = {}
    for p in (0, 1, -1):
        B = spin_basis_general(L, pauli=p)
        mats[p] = [hamiltonian([s], [], basis=B, dtype=np.float64, check_herm=False, check_pcon=False,
                               check_symm=False).toarray() for s in st]
    np.testing.assert_allclose(mats[1][0], 2 * mats[0][0])
    np.testing.assert_allclose(mats[1][1], 4 * mats[0][1])
    np.testing.assert_allclose(mats[-1][1], mats[0][1])
    np.testing.assert_allclose(mats[-1][2], 4 * mats[0][2])


def test_fermion_hubbard_vs_oracle():
    from quspin_amd.basis import spinful_fermion_basis_general
    from quspin_amd.operators import hamiltonian, quantum_LinearOperator
    L = 5
    B = spinful_fermion_basis_general(L, Nf=(2, 3))
    static = hubbard_static(L, t=1.0, U=2.5)
    ref = _oracle_H(B, static, fermion=True, spinful_L=L)
    H = hamiltonian(static, [], basis=B, dtype=np.float64, check_symm=False, check_herm=False, check_pcon=False)
    np.testing.assert_allclose(H.toarray(), ref.toarray().real, atol=1e-13)
    Q = quantum_LinearOperator(static, basis=B, dtype=np.float64, check_symm=False, check_herm=False,
                               check_pcon=False)
    v = _vec(B.Ns, np.float64)
    np.testing.assert_allclose(Q.dot(v), ref.real @ v, atol=1e-12)


def test_fermion_momentum_sectors_spectrum():
    """Union of k-sector spectra == spectrum without translation symmetry."""
    from quspin_amd.basis import spinful_fermion_basis_general
    from quspin_amd.operators import hamiltonian
    L = 6
    T = (np.arange(L) + 1) % L
    static = hubbard_static(L, U=3.0)
    full = hamiltonian(static, [], basis=spinful_fermion_basis_general(L, Nf=(2, 2)), dtype=np.float64,
                       check_symm=False, check_herm=False, check_pcon=False)
    E_full = np.linalg.eigvalsh(full.toarray())
    Es = []
    for k in range(L):
        b = spinful_fermion_basis_general(L, Nf=(2, 2), kblock=(T, k))
        Hk = hamiltonian(static, [], basis=b, dtype=np.complex128, check_symm=False, check_herm=False,
                         check_pcon=False)
        Es.extend(np.linalg.eigvalsh(Hk.toarray()))
    np.testing.assert_allclose(np.sort(Es), E_full, atol=1e-11)


def test_large_C2_hermiticity_and_linearity():
    """Full C2 size: <u, H v> == <H u, v>, linearity, and CSR == matrix-free."""
    import torch
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian, quantum_LinearOperator
    L = 28
    m = chain_maps(L)
    B = spin_basis_general(L, Nup=14, pauli=0, kblock=(m["T"], 0), pblock=(m["P"], 0), zblock=(m["Z"], 0))
    static = xxz_static(L, J=1.0, Delta=1.0)
    Q = quantum_LinearOperator(static, basis=B, dtype=np.float64, check_symm=False, check_herm=False,
                               check_pcon=False)
    g = torch.Generator(device="cuda").manual_seed(0)
    u = torch.randn(B.Ns, dtype=torch.float64, device="cuda", generator=g)
    v = torch.randn(B.Ns, dtype=torch.float64, device="cuda", generator=g)
    Hu, Hv = Q.dot(u), Q.dot(v)
    lhs = torch.dot(u, Hv).item()
    rhs = torch.dot(Hu, v).item()
    assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(lhs))
    Huv = Q.dot(u + 2.5 * v)
    assert torch.max(torch.abs(Huv - (Hu + 2.5 * Hv))).item() < 1e-10
    H = hamiltonian(static, [], basis=B, dtype=np.float64, check_symm=False, check_herm=False, check_pcon=False)
    Hv2 = H.dot(v)
    assert torch.max(torch.abs(Hv2 - Hv)).item() < 1e-11


@pytest.mark.parametrize("schedule", ["rows", "batch4", "batch8", "batch14", "batch14x"])
def test_large_C2_forced_schedules_vs_csr(schedule, monkeypatch):
    """Full C2 size (1,412 blocks of 256 rows, so the per-XCD block orders
    renumber most of the grid): every table schedule of the batched kernel
    against the device CSR, which is assembled through the orbit search."""
    import torch
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian, quantum_LinearOperator
    monkeypatch.setenv("QSX_LOOKUP", "table")
    monkeypatch.setenv("QSX_SCHEDULE", schedule)
    L = 28
    m = chain_maps(L)
    print(f"case L={L} schedule={schedule}", flush=True)
    B = spin_basis_general(L, Nup=14, pauli=0, kblock=(m["T"], 0), pblock=(m["P"], 0), zblock=(m["Z"], 0))
    static = xxz_static(L, J=1.0, Delta=1.0)
    kw = dict(check_symm=False, check_herm=False, check_pcon=False)
    Q = quantum_LinearOperator(static, basis=B, dtype=np.float64, **kw)
    v = torch.randn(B.Ns, dtype=torch.float64, device="cuda", generator=torch.Generator(device="cuda").manual_seed(3))
    Hv = Q.dot(v)
    assert all(c["schedule"] == schedule for c in Q._qsx_operator().config.values())
    Hv2 = hamiltonian(static, [], basis=B, dtype=np.float64, **kw).dot(v)
    assert torch.max(torch.abs(Hv2 - Hv)).item() <= 1e-12 * torch.max(torch.abs(Hv2)).item()


@pytest.mark.parametrize("lookup", ["search", "table"])
@pytest.mark.parametrize("schedule", ["rows", "rows2", "rows4", "rows8", "queue", "batch4", "batch8", "batch14",
                                      "batch14x"])
def test_forced_configs_vs_oracle(lookup, schedule, monkeypatch):
    """Every (lookup, schedule) pair of the fused kernel against the oracle:
    complex k-sector, real k=0/p/z sector, fermion k-sector."""
    from quspin_amd.basis import spin_basis_general, spinful_fermion_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    monkeypatch.setenv("QSX_LOOKUP", lookup)
    monkeypatch.setenv("QSX_SCHEDULE", schedule)
    print(f"case lookup={lookup} schedule={schedule}", flush=True)
    L = 14
    m = chain_maps(L)
    cases = [(spin_basis_general(L, Nup=7, pauli=0, kblock=(m["T"], 3)), xxz_static(L, Delta=0.3), np.complex128,
              False, None),
             (spin_basis_general(L, Nup=7, pauli=0, kblock=(m["T"], 0), pblock=(m["P"], 1), zblock=(m["Z"], 0)),
              xxz_static(L, Delta=1.0), np.float64, False, None)]
    Lf = 6
    T = (np.arange(Lf) + 1) % Lf
    cases.append((spinful_fermion_basis_general(Lf, Nf=(3, 2), kblock=(T, 1)), hubbard_static(Lf, U=2.0),
                  np.complex128, True, Lf))
    for B, static, dt, ferm, sl in cases:
        Q = quantum_LinearOperator(static, basis=B, dtype=dt, check_symm=False, check_herm=False, check_pcon=False)
        H = _oracle_H(B, static, fermion=ferm, spinful_L=sl)
        for nvec in (None, 3):
            v = _vec(B.Ns, dt, seed=5, nvec=nvec)
            got = Q.dot(v)
            ref = H @ v.astype(np.complex128)
            np.testing.assert_allclose(got, ref if np.dtype(dt).kind == "c" else ref.real, atol=1e-11, rtol=0)
        cfg = Q._qsx_operator().config
        if len(B._pers):
            assert all(c["lookup"] == lookup for c in cfg.values()), cfg


@pytest.mark.parametrize("lookup", ["table", "search"])
def test_merged_term_pairs_vs_oracle(lookup, monkeypatch):
    """Op lists whose '+-' / '-+' terms share a bond with equal couplings (one
    merged term with two accepted bit patterns), unequal couplings (kept
    apart), complex couplings and extra one-site flips, on a symmetric sector
    (table or orbit search) and on a sector without symmetry (direct rank):
    every product against the oracle's matrix."""
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    monkeypatch.setenv("QSX_LOOKUP", lookup)
    monkeypatch.setenv("QSX_SCHEDULE", "rows")
    L = 12
    eq = [[0.5, i, (i + 1) % L] for i in range(L)]
    # translation-invariant couplings (the k-sector product is the gather form)
    ne_p = [[0.6, i, (i + 2) % L] for i in range(L)]
    ne_m = [[0.8, i, (i + 2) % L] for i in range(L)]
    cj = [[0.2 + 0.4j, i, (i + 3) % L] for i in range(L)]
    static_c = [["+-", eq], ["-+", eq], ["+-", ne_p], ["-+", ne_m], ["+-", cj],
                ["-+", [[np.conj(b[0]), b[1], b[2]] for b in cj]], ["zz", eq]]
    static_r = [["+-", eq], ["-+", eq], ["+-", ne_p], ["-+", ne_m], ["zz", eq], ["z", [[0.3, 2]]]]
    m = chain_maps(L)
    cases = [(spin_basis_general(L, Nup=6, pauli=0, kblock=(m["T"], 2)), static_c, np.complex128),
             (spin_basis_general(L, Nup=6, pauli=0), static_r, np.float64),
             (spin_basis_general(L, Nup=5, pauli=0), static_c, np.complex128)]
    for B, static, dt in cases:
        print(f"case lookup={lookup} Ns={B.Ns} maps={len(B._pers)} dtype={np.dtype(dt).name}", flush=True)
        Q = quantum_LinearOperator(static, basis=B, dtype=dt, check_symm=False, check_herm=False, check_pcon=False)
        H = _oracle_H(B, static)
        v = _vec(B.Ns, dt, seed=7)
        got = Q.dot(v)
        ref = H @ v.astype(np.complex128)
        np.testing.assert_allclose(got, ref if np.dtype(dt).kind == "c" else ref.real, atol=1e-11, rtol=0)
        # the adjoint applies the transposed term list; Q.T applies conj(H)
        # (quantum_LinearOperator passes (conjugated, transposed) into the
        # (transposed, conjugated) slots, quantum_LinearOperator_core.py:335-343)
        for op, M in ((Q.H, H.conj().T), (Q.T, H.conj())):
            e = M @ v.astype(np.complex128)
            np.testing.assert_allclose(op.dot(v), e if np.dtype(dt).kind == "c" else e.real, atol=1e-11, rtol=0)


@pytest.mark.parametrize("lookup", ["table", "search"])
@pytest.mark.parametrize("schedule", ["rows", "batch4", "batch8", "batch14", "batch14x", "chunk4", "chunk8"])
def test_fast_rows_pauli_strings_vs_oracle(lookup, schedule, monkeypatch):
    """The fast-row kernels (rows: one term at a time with the direct rank,
    batches sized from the term count with the table; batches of 4, 8 and 14;
    chunks of 4 and 8 terms with predicated gathers for the direct rank)
    on x / y / z strings: odd and even numbers of y (factors i^(3 ny)), z
    factors on flipped and unflipped sites, complex couplings, a term count
    that is not a multiple of the batch, on a k-sector (sector table) and on
    a sector without symmetry (direct rank); products, adjoints and
    transposes against the oracle's matrix."""
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    monkeypatch.setenv("QSX_LOOKUP", lookup)
    monkeypatch.setenv("QSX_SCHEDULE", schedule)
    L = 12
    m = chain_maps(L)
    b1 = [[0.7, i, (i + 1) % L] for i in range(L)]
    b2 = [[0.3 - 0.2j, i, (i + 2) % L] for i in range(L)]
    tri = [[0.45, i, (i + 1) % L, (i + 2) % L] for i in range(L)]
    static = [["xy", b1], ["yx", [[np.conj(b[0]), b[1], b[2]] for b in b1]], ["xx", b2],
              ["yy", [[np.conj(b[0]), b[1], b[2]] for b in b2]], ["+z-", tri], ["-z+", tri], ["zz", b1],
              ["yzy", tri]]
    cases = [(spin_basis_general(L, Nup=6, pauli=0, kblock=(m["T"], 1)), np.complex128),
             (spin_basis_general(L, Nup=5, pauli=0), np.complex128)]
    for B, dt in cases:
        print(f"case lookup={lookup} schedule={schedule} Ns={B.Ns} maps={len(B._pers)}", flush=True)
        Q = quantum_LinearOperator(static, basis=B, dtype=dt, check_symm=False, check_herm=False, check_pcon=False)
        H = _oracle_H(B, static)
        v = _vec(B.Ns, dt, seed=11)
        np.testing.assert_allclose(Q.dot(v), H @ v, atol=1e-11, rtol=0)
        for op, M in ((Q.H, H.conj().T), (Q.T, H.conj())):
            np.testing.assert_allclose(op.dot(v), M @ v, atol=1e-11, rtol=0)


@pytest.mark.parametrize("schedule", ["rows", "batch4", "batch8", "batch14", "batch14x"])
def test_fast_rows_u64_table_vs_oracle(schedule, monkeypatch):
    """Sector tables with 8-byte entries (index | coefficient id << 40, the
    layout large bases need), forced on small bases: the fast-row kernels
    read them as the 4-byte ones, complex k-sector and real k=0/p/z sector,
    against the oracle."""
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    monkeypatch.setenv("QSX_LOOKUP", "table")
    monkeypatch.setenv("QSX_SCHEDULE", schedule)
    L = 14
    m = chain_maps(L)
    cases = [(spin_basis_general(L, Nup=7, pauli=0, kblock=(m["T"], 3)), xxz_static(L, Delta=0.3), np.complex128),
             (spin_basis_general(L, Nup=7, pauli=0, kblock=(m["T"], 0), pblock=(m["P"], 1), zblock=(m["Z"], 0)),
              xxz_static(L, Delta=1.0), np.float64)]
    for B, static, dt in cases:
        tab = B.lookup_table(entry_bytes=8)
        print(f"case schedule={schedule} Ns={B.Ns} entry_bytes={tab[1]} lut={tab[3] is not None}", flush=True)
        assert tab[1] == 8 and tab[0].element_size() == 8
        Q = quantum_LinearOperator(static, basis=B, dtype=dt, check_symm=False, check_herm=False, check_pcon=False)
        H = _oracle_H(B, static)
        v = _vec(B.Ns, dt, seed=13)
        ref = H @ v.astype(np.complex128)
        np.testing.assert_allclose(Q.dot(v), ref if np.dtype(dt).kind == "c" else ref.real, atol=1e-11, rtol=0)
        assert all(c["lookup"] == "table" for c in Q._qsx_operator().config.values())


def test_empty_and_single_state_sectors():
    """Degenerate sectors: no state at all (L=4 half filling with k=1 and
    z=+1, an eigenvalue pair no state carries; L=6 all down with k=1) and a
    single state (all down, k=0; all up without symmetry): the basis sizes
    match the oracle's, products return empty or the one diagonal element,
    and an expm over a single state is the scalar exponential."""
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian, quantum_LinearOperator
    from quspin_amd.tools.evolution import expm_multiply_parallel
    kw = dict(check_symm=False, check_herm=False, check_pcon=False)
    for L, Nup, blocks in [(4, 2, ("T", 1, "Z", 0)), (6, 0, ("T", 1)), (6, 0, ("T", 0)), (5, 5, ())]:
        m = chain_maps(L)
        bk = {}
        names = {"T": "kblock", "Z": "zblock", "P": "pblock"}
        for i in range(0, len(blocks), 2):
            bk[names[blocks[i]]] = (m[blocks[i]], blocks[i + 1])
        print(f"case L={L} Nup={Nup} blocks={blocks}", flush=True)
        B = spin_basis_general(L, Nup=Nup, pauli=0, **bk)
        maps = [m[blocks[i]] for i in range(0, len(blocks), 2)]
        qs = [blocks[i + 1] for i in range(0, len(blocks), 2)]
        g = O.Group(L, maps, [O.process_map(x)[2] for x in maps], qs)
        res = O.make_basis(L, g, Np=Nup)
        ref_basis = res[0] if isinstance(res, tuple) else res
        assert B.Ns == len(ref_basis)
        static = xxz_static(L, Delta=0.7)
        dt = np.complex128
        Q = quantum_LinearOperator(static, basis=B, dtype=dt, **kw)
        H = hamiltonian(static, [], basis=B, dtype=dt, **kw)
        v = _vec(B.Ns, dt, seed=3)
        if B.Ns == 0:
            assert Q.dot(v).shape == (0,) and H.dot(v).shape == (0,)
            continue
        Hm = _oracle_H(B, static)
        ref = Hm @ v
        np.testing.assert_allclose(Q.dot(v), ref, atol=1e-12, rtol=0)
        np.testing.assert_allclose(H.dot(v), ref, atol=1e-12, rtol=0)
        if B.Ns == 1:
            e = complex(Hm.toarray()[0, 0]) if hasattr(Hm, "toarray") else complex(np.asarray(Hm)[0, 0])
            out = expm_multiply_parallel(Q, a=-0.3j, dtype=dt).dot(v.copy())
            np.testing.assert_allclose(out, np.exp(-0.3j * e) * v, atol=1e-12, rtol=0)


def test_degenerate_fermion_and_1d_sectors():
    """Fermion sectors with a single state (the spinful vacuum, a filled
    spinless chain) and an empty spin_basis_1d sector: basis sizes, products
    against the oracle's matrix, and empty products."""
    from quspin_amd.basis import spin_basis_1d, spinful_fermion_basis_general, spinless_fermion_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    kw = dict(check_symm=False, check_herm=False, check_pcon=False)
    L = 4
    T = (np.arange(L) + 1) % L
    cases = [("spinful vacuum", spinful_fermion_basis_general(L, Nf=(0, 0), kblock=(T, 0)),
              hubbard_static(L, U=2.0, mu=0.3), True, L),
             ("spinless filled", spinless_fermion_basis_general(L, Nf=L),
              [["+-", [[-1.0, i, (i + 1) % L] for i in range(L)]], ["-+", [[1.0, i, (i + 1) % L] for i in range(L)]],
               ["n", [[0.5, i] for i in range(L)]]], True, None)]
    for name, B, static, ferm, sl in cases:
        print(f"case {name} Ns={B.Ns}", flush=True)
        assert B.Ns == 1
        Q = quantum_LinearOperator(static, basis=B, dtype=np.complex128, **kw)
        H = _oracle_H(B, static, fermion=ferm, spinful_L=sl)
        v = _vec(B.Ns, np.complex128, seed=2)
        np.testing.assert_allclose(Q.dot(v), H @ v, atol=1e-12, rtol=0)
    print("case spin_basis_1d empty k-sector", flush=True)
    B1 = spin_basis_1d(6, Nup=0, kblock=1)
    assert B1.Ns == 0
    Q1 = quantum_LinearOperator(xxz_static(6), basis=B1, dtype=np.complex128, **kw)
    assert Q1.dot(np.zeros(0, np.complex128)).shape == (0,)


@pytest.mark.parametrize("lookup", ["table", "search"])
def test_64bit_states_vs_oracle(lookup, monkeypatch):
    """Bases wider than 32 bits (64-bit state words): an L=36 chain with two
    up spins, in the k=1 sector (sector table or orbit search) and without
    symmetry (direct rank); the basis is bit-exact against the oracle's
    representatives of the sector (built from its 630 states) and products
    equal the oracle's matrix."""
    import itertools
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    monkeypatch.setenv("QSX_LOOKUP", lookup)
    L, Nup = 36, 2
    m = chain_maps(L)
    static = xxz_static(L, Delta=0.6)
    cand = np.array(sorted((sum(1 << b for b in c) for c in itertools.combinations(range(L), Nup)), reverse=True),
                    dtype=np.uint64)
    for blocks in ({"kblock": (m["T"], 1)}, {}):
        print(f"case L={L} Nup={Nup} lookup={lookup} blocks={list(blocks)}", flush=True)
        B = spin_basis_general(L, Nup=Nup, pauli=0, **blocks)
        if blocks:
            items = O.order_maps(blocks)
            g = O.Group(L, [it[1] for it in items], [it[2] for it in items], [it[3] for it in items])
            rep, _, _ = g.representative(cand)
            c = cand[rep == cand]
            n = g.normalization(c)
            b, n = c[n > 0], n[n > 0]
        else:
            g = O.Group(L, [], [], [])
            b, n = cand, np.ones(len(cand), np.int64)
        assert B.states.dtype == np.uint64
        np.testing.assert_array_equal(B.states, b)
        H = O.make_csr(b, n, g, flat_terms(static), np.complex128)
        Q = quantum_LinearOperator(static, basis=B, dtype=np.complex128, check_symm=False, check_herm=False,
                                   check_pcon=False)
        for nvec in (None, 2):
            v = _vec(B.Ns, np.complex128, seed=9, nvec=nvec)
            np.testing.assert_allclose(Q.dot(v), H @ v, atol=1e-12, rtol=0)


def test_hamiltonian_dynamic_terms_and_out_semantics():
    """hamiltonian with a time-dependent term (hamiltonian_core.py dot):
    H(t) v = H_s v + f(t, *args) H_d v at several times; out= with
    overwrite_out=False accumulates a H(t) v into out; a vector of times
    applies H(t_i) to column i."""
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian
    L = 10
    m = chain_maps(L)
    B = spin_basis_general(L, Nup=5, pauli=0, kblock=(m["T"], 0))
    static = xxz_static(L, Delta=0.8)
    nnn = [[0.4, i, (i + 2) % L] for i in range(L)]

    def f(t, w):
        return np.cos(w * t)

    H = hamiltonian(static, [["zz", nnn, f, [2.0]]], basis=B, dtype=np.complex128, check_symm=False,
                    check_herm=False, check_pcon=False)
    Hs = _oracle_H(B, static)
    Hd = _oracle_H(B, [["zz", nnn]])
    v = _vec(B.Ns, np.complex128, seed=4)
    for t in (0.0, 0.7, 1.9):
        print(f"case t={t}", flush=True)
        np.testing.assert_allclose(H.dot(v, time=t), Hs @ v + f(t, 2.0) * (Hd @ v), atol=1e-12, rtol=0)
    out = np.ones_like(v)
    H.dot(v, time=0.7, out=out, overwrite_out=False, a=0.5)
    np.testing.assert_allclose(out, 1.0 + 0.5 * (Hs @ v + f(0.7, 2.0) * (Hd @ v)), atol=1e-12, rtol=0)
    V2 = np.stack([v, 2.0 * v], axis=1)
    got = H.dot(V2, time=[0.3, 1.1])
    for i, t in enumerate((0.3, 1.1)):
        np.testing.assert_allclose(got[:, i], Hs @ V2[:, i] + f(t, 2.0) * (Hd @ V2[:, i]), atol=1e-12, rtol=0)


def test_spinful_without_double_occupancy_vs_oracle():
    """Spinful fermions with double_occupancy=False (basis_general drops the
    doubly occupied states): hops that would create a doubly occupied site
    leave the basis and give no element; products with and without
    translation symmetry against the oracle's matrix on the same basis."""
    from quspin_amd.basis import spinful_fermion_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    L = 5
    T = (np.arange(L) + 1) % L
    static = hubbard_static(L, U=0.0, mu=0.2)
    for blocks in ({}, {"kblock": (T, 1)}):
        B = spinful_fermion_basis_general(L, Nf=(2, 2), double_occupancy=False, **blocks)
        print(f"case L={L} Nf=(2,2) no double occupancy blocks={list(blocks)} Ns={B.Ns}", flush=True)
        assert not np.any((B.states >> L) & B.states & ((1 << L) - 1))
        Q = quantum_LinearOperator(static, basis=B, dtype=np.complex128, check_symm=False, check_herm=False,
                                   check_pcon=False)
        H = _oracle_H(B, static, fermion=True, spinful_L=L)
        v = _vec(B.Ns, np.complex128, seed=6)
        np.testing.assert_allclose(Q.dot(v), H @ v, atol=1e-12, rtol=0)
