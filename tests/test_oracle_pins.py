"""Pins of the oracle (CPU only): known answers printed by the reference's
notebooks and tests, closed forms, and brute force on tiny systems.  Each
test is chosen so that a plausible mistake in the oracle (a dropped term, a
wrong sign or index, a transposed operand, a wrong phase convention) fails."""

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.special import comb

from oracle import qsx_oracle as O
from tests.models import chain_maps, flat_terms, xxz_static


def test_ssh_momentum_block_printout():
    """SSH.ipynb cell 5: L=6, one fermion, translation by a=2, kblock=1 prints
    [[0.5, -0.65-0.78j], [-0.65+0.78j, -0.5]] (pull action, chi = e^{+2pi i q l/m})."""
    L, J, dJ, D = 6, 1.0, 0.1, 0.5
    T2 = [(i + 2) % L for i in range(L)]
    g = O.Group(L, [T2], [3], [1], fermion=True)
    b, n = O.make_basis(L, g, Np=1)
    static = ([("+-", (i, (i + 1) % L), -J - dJ * (-1) ** i) for i in range(L)]
              + [("-+", (i, (i + 1) % L), J + dJ * (-1) ** i) for i in range(L)]
              + [("n", (i,), D * (-1) ** i) for i in range(L)])
    H = O.reduced_op(b, n, g, static, fermion=True)
    np.testing.assert_allclose(np.round(H, 2), [[0.5, -0.65 - 0.78j], [-0.65 + 0.78j, -0.5]], atol=1e-12)
    # k = 2 block is the complex conjugate (SSH.ipynb cell 5, last block)
    g2 = O.Group(L, [T2], [3], [2], fermion=True)
    H2 = O.reduced_op(b, n, g2, static, fermion=True)
    np.testing.assert_allclose(H2, H.conj(), atol=1e-12)


@pytest.mark.parametrize("L", [8, 10])
def test_reduced_equals_projected_full(L):
    """A.5 element formula == P^H H_full P with the get_proj coefficient
    1/sqrt(n |G|) (base_general.py:571-575), P^H P = 1, and the union of all
    sector spectra is the full spectrum."""
    m = chain_maps(L)
    st = flat_terms(xxz_static(L, Delta=0.6))
    full = O.full_basis_states(L, L // 2)
    Hf = O.full_hamiltonian(full, L, st)
    Es = []
    for k in range(L):
        for z in (0, 1):
            items = O.order_maps({"k": (m["T"], k), "z": (m["Z"], z)})
            g = O.Group(L, [it[1] for it in items], [it[2] for it in items], [it[3] for it in items])
            b, n = O.make_basis(L, g, Np=L // 2)
            if len(b) == 0:
                continue
            P = O.projector(b, n, g, full)
            np.testing.assert_allclose(P.conj().T @ P, np.eye(len(b)), atol=1e-12)
            np.testing.assert_allclose(O.reduced_op(b, n, g, st), P.conj().T @ Hf @ P, atol=1e-12)
            Es.extend(np.linalg.eigvalsh(O.reduced_op(b, n, g, st)))
    np.testing.assert_allclose(np.sort(Es), np.linalg.eigvalsh(Hf), atol=1e-10)


def test_sector_sizes_closed_form():
    """Sum over k of the k-sector sizes is C(L, L/2); L=12 k=0,p=0,z=0 size
    from Burnside's count of symmetric orbits."""
    L = 12
    m = chain_maps(L)
    tot = 0
    for k in range(L):
        g = O.Group(L, [m["T"]], [L], [k])
        tot += len(O.make_basis(L, g, Np=L // 2)[0])
    assert tot == comb(L, L // 2, exact=True)


def test_spin_operator_table_vs_kron():
    """Local ops (pauli=0) and the site<->bit convention against explicit
    2x2 spin matrices and Kronecker products (site 0 = leftmost factor =
    highest bit)."""
    L = 3
    Sp = np.array([[0, 1], [0, 0]], complex)  # basis (up, down) = (bit 1, bit 0)
    Sm = Sp.T
    Sz = np.diag([0.5, -0.5]).astype(complex)
    Sx = 0.5 * np.array([[0, 1], [1, 0]], complex)
    Sy = 0.5 * np.array([[0, -1j], [1j, 0]])
    ops = {"+": Sp, "-": Sm, "z": Sz, "x": Sx, "y": Sy, "n": np.diag([1.0, 0.0]).astype(complex)}
    full = O.full_basis_states(L)  # descending: index 0 = all up = kron index 0
    rng = np.random.default_rng(1)
    for _ in range(20):
        chars = "".join(rng.choice(list(ops), size=2))
        sites = tuple(rng.choice(L, size=2, replace=False))
        M = O.full_hamiltonian(full, L, [(chars, sites, 1.0)])
        mats = [np.eye(2, dtype=complex)] * L
        mats = list(mats)
        for ch, i in zip(chars, sites):
            mats[i] = ops[ch] @ mats[i]
        K = mats[0]
        for A in mats[1:]:
            K = np.kron(K, A)
        np.testing.assert_allclose(M, K, atol=1e-14)


def test_fermion_anticommutation():
    """{c_i, c_j^+} = delta_ij and {c_i, c_j} = 0 on the full Fock space."""
    L = 4
    full = O.full_basis_states(L)

    def mat(ch, i):
        M = np.zeros((full.size, full.size))
        me, s2 = O.apply_term(full, ch, (i,), L, fermion=True)
        rows = O.index_of(full, s2)
        ok = me != 0
        M[rows[ok], np.nonzero(ok)[0]] = me[ok].real
        return M

    for i in range(L):
        for j in range(L):
            ci, cjd = mat("-", i), mat("+", j)
            np.testing.assert_allclose(ci @ cjd + cjd @ ci, np.eye(full.size) * (i == j), atol=1e-14)
            cj = mat("-", j)
            np.testing.assert_allclose(ci @ cj + cj @ ci, 0, atol=1e-14)


@pytest.mark.parametrize("U,E", [(0.0, -12.0), (1.0, -11.8364690235642218), (2.0, -11.7104580743242632),
                                 (4.0, -11.5302924026297138)])
def test_hubbard_4x4_literature_energies(U, E):
    """test_FHM_energies.py:21-33 (arXiv:cond-mat/0604319): 4x4 torus,
    Nf=(2,2), hopping +-| with J=+1, -+| with J=-1, U n|n at U = linspace(0,4,9)[j];
    the reference asserts atol 1e-13, eigsh here reaches ~1e-12."""
    from scipy.sparse.linalg import eigsh
    Lx = Ly = 4
    N = Lx * Ly
    i = np.arange(N)
    x, y = i % Lx, i // Lx
    tx, ty = (x + 1) % Lx + y * Lx, x + ((y + 1) % Ly) * Lx
    terms = []
    for j in range(N):
        for t in (tx[j], ty[j]):
            terms += [("+-", (j, int(t)), 1.0), ("-+", (j, int(t)), -1.0),
                      ("+-", (j + N, int(t) + N), 1.0), ("-+", (j + N, int(t) + N), -1.0)]
        if U:
            terms.append(("nn", (j, j + N), U))
    g = O.Group(2 * N, [], [], [], fermion=True)
    # the (2,2) sector directly from 16-bit halves (up in the high half)
    import itertools
    half = [sum(1 << p for p in c) for c in itertools.combinations(range(N), 2)]
    states = np.array(sorted(((u << N) | d for u in half for d in half), reverse=True), dtype=np.uint64)
    H = O.make_csr(states, np.ones(states.size, np.int64), g, terms, np.float64, fermion=True)
    e = eigsh(H, k=1, which="SA", tol=1e-14)[0][0]
    assert abs(e - E) < 1e-11


def test_expm_against_dense_and_scipy_partition():
    """Taylor loop vs scipy.linalg.expm on a small matrix; (m*, s) equals
    scipy's _fragment_3_1 choice for the same norm."""
    from scipy.linalg import expm
    rng = np.random.default_rng(0)
    A = sp.random(60, 60, density=0.2, random_state=3).tocsr()
    A = A + A.T
    v = rng.normal(size=60) + 1j * rng.normal(size=60)
    for a in (-0.5j, -3.0j, 0.7):
        got, part = O.expm_multiply(A, v, a=a)
        np.testing.assert_allclose(got, expm(a * A.toarray()) @ v, atol=1e-12)
        assert part[0] >= 1 and part[1] >= 1


def test_partition_inputs_and_theta_branch_choice():
    """The oracle's mu, ||A - mu I||_1 and theta-branch (m*, s) against what
    fixes them independently: the dense trace and column sums; and, for
    |a| ||A - mu I||_1 <= 63.36, the branch's defining property (Al-Mohy &
    Higham 2011, eq. 3.10 and Table 3.1): s = ceil(|a| ||.||_1 / theta_m) for
    the chosen m, and no other m of the theta table has a smaller cost m * s
    (ties to the smaller m, which the table's ascending scan keeps)."""
    from scipy.sparse.linalg._expm_multiply import _theta
    rng = np.random.default_rng(1)
    A = sp.random(80, 80, density=0.1, random_state=5, dtype=np.float64).tocsr()
    A = (A + 1j * sp.random(80, 80, density=0.05, random_state=6)).tocsr()
    Ad = A.toarray()
    mu = O.csr_trace(A) / 80
    assert abs(mu - np.trace(Ad) / 80) < 1e-14
    nrm = O.csr_1_norm(A, -mu)
    assert abs(nrm - np.abs(Ad - mu * np.eye(80)).sum(axis=0).max()) < 1e-12
    for a in (-0.01j, -0.3j, -2.0j, 1.7, -5.5j):
        an = abs(a) * nrm
        assert an <= 63.36
        m, s = O.fragment_3_1_partition(A, a, mu, nrm, 2.0 ** -53)
        assert s == int(np.ceil(an / _theta[m]))
        costs = {mm: mm * int(np.ceil(an / th)) for mm, th in _theta.items() if mm <= 55}
        best = min(costs.values())
        assert m * s == best and m == min(mm for mm, c in costs.items() if c == best)
    assert O.fragment_3_1_partition(A, 0.0, mu, nrm, 2.0 ** -53) == (0, 1)
    v = rng.normal(size=80) + 0j
    got, _ = O.expm_multiply(A, v, a=0.0)
    np.testing.assert_array_equal(got, v)


def test_lanczos_ground_energy():
    L = 10
    m = chain_maps(L)
    g = O.Group(L, [m["T"]], [L], [0])
    b, n = O.make_basis(L, g, Np=L // 2)
    H = O.make_csr(b, n, g, flat_terms(xxz_static(L)), np.float64)
    v0 = np.random.default_rng(0).normal(size=len(b))
    E, V, Q, a, bb = O.lanczos_full(lambda x: H @ x, v0, 60)
    assert abs(E[0] - np.linalg.eigvalsh(H.toarray())[0]) < 1e-10
    assert np.allclose(Q[:10] @ Q[:10].T, np.eye(10), atol=1e-8)


def test_c_oracle_matches_numpy_oracle():
    """The C/OpenMP restatement (used at full sizes and as the CPU baseline)
    equals the numpy oracle on small sectors: basis, n and H.v."""
    from oracle import c_oracle as C
    for L, k, z in [(12, 0, 1), (12, 3, None), (11, 2, None)]:
        m = chain_maps(L)
        blocks = {"k": (m["T"], k)}
        if z is not None:
            blocks["z"] = (m["Z"], z)
        items = O.order_maps(blocks)
        maps, pers, qs = [it[1] for it in items], [it[2] for it in items], [it[3] for it in items]
        g = O.Group(L, maps, pers, qs)
        b, n = O.make_basis(L, g, Np=L // 2)
        cg = C.Group(L, maps, pers, qs)
        cb, cn = cg.make_basis(L // 2)
        np.testing.assert_array_equal(cb, b.astype(np.uint64))
        np.testing.assert_array_equal(cn, n)
        if (2 * k) % L == 0:
            terms = flat_terms(xxz_static(L, Delta=0.4))
            H = O.make_csr(b, n, g, terms, np.float64)
            v = np.random.default_rng(0).normal(size=len(b))
            for fused in (True, False):
                np.testing.assert_allclose(C.hv_real(cg, cb, cn, terms, v, fused=fused), H @ v, atol=1e-12)


def test_mlp_backward_finite_differences():
    """Gradients of the oracle MLP against central differences of
    sum(dy * y) (every parameter block)."""
    rng = np.random.default_rng(4)
    shapes = [(64, 4), (64,), (64, 64), (64,), (2, 64), (2,)]
    params = [rng.normal(0, 0.3, size=s) for s in shapes]
    X = rng.normal(size=(7, 4))
    dy = rng.normal(size=(7, 2))
    y, cache = O.mlp_forward(params, X)
    grads = O.mlp_backward(params, cache, dy)
    h = 1e-6
    for bi, (P, G) in enumerate(zip(params, grads)):
        flat = P.reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            yp = O.mlp_forward(params, X)[0]
            flat[idx] = old - h
            ym = O.mlp_forward(params, X)[0]
            flat[idx] = old
            fd = np.sum(dy * (yp - ym)) / (2 * h)
            assert abs(fd - G.reshape(-1)[idx]) < 1e-7 * max(1.0, abs(fd)), (bi, idx)


def test_shift_sector_and_get_amp_identities():
    """op_shift_sector: for an operator that commutes with the group it
    equals the reduced operator of the sector (P^H H P, the A.5 element
    formula); for the momentum-q operator sum_j e^{-2 pi i q j / L} z_j the
    shifted parts over all target sectors k2 rebuild H_full P_k1 v
    (sum_k2 P_k2 P_k2^H = 1).  get_amp: psi_full[s] = C_s psi_rep(s) for every
    state of the sector, and |C_s|^2 sums to 1 over each orbit."""
    L = 8
    m = chain_maps(L)
    full = O.full_basis_states(L, L // 2)
    rng = np.random.default_rng(2)
    groups = {k: O.Group(L, [m["T"]], [L], [k]) for k in range(L)}
    bases = {k: O.make_basis(L, groups[k], Np=L // 2) for k in range(L)}
    st = flat_terms(xxz_static(L, Delta=0.4))
    b, n = bases[3]
    v = rng.normal(size=len(b)) + 1j * rng.normal(size=len(b))
    got = O.op_shift_sector(b, n, groups[3], b, n, groups[3], full, L, st, v)
    np.testing.assert_allclose(got, O.reduced_op(b, n, groups[3], st) @ v, atol=1e-12)
    q = 2
    zq = [("z", (j,), np.exp(-2j * np.pi * q * j / L)) for j in range(L)]
    Hf = O.full_hamiltonian(full, L, zq)
    P1 = O.projector(b, n, groups[3], full)
    acc = np.zeros(full.size, complex)
    for k2 in range(L):
        b2, n2 = bases[k2]
        if len(b2) == 0:
            continue
        part = O.op_shift_sector(b2, n2, groups[k2], b, n, groups[3], full, L, zq, v)
        if k2 != (3 + q) % L:
            assert np.abs(part).max() < 1e-12
        acc += O.projector(b2, n2, groups[k2], full) @ part
    np.testing.assert_allclose(acc, Hf @ (P1 @ v), atol=1e-12)
    # get_amp
    for k in (0, 3):
        b, n = bases[k]
        psi = rng.normal(size=len(b)) + 1j * rng.normal(size=len(b))
        psi_full = O.projector(b, n, groups[k], full) @ psi
        C, rep = O.get_amp(b, n, groups[k], full, full)
        cols = O.index_of(b, rep)
        ok = cols >= 0
        np.testing.assert_allclose(psi_full[ok], C[ok] * psi[cols[ok]], atol=1e-12)
        assert np.abs(psi_full[~ok]).max(initial=0) < 1e-12
        for c in range(len(b)):
            assert abs(np.sum(np.abs(C[cols == c]) ** 2) - 1.0) < 1e-12


def _full_krylov_data(n, seed):
    """Real symmetric H, real start vector r and an orthonormal basis Q (rows)
    of the whole space with q_0 = r; T = Q H Q^T and its eigenpairs."""
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, n))
    H = (X + X.T) / 2
    r = rng.normal(size=n)
    r /= np.linalg.norm(r)
    B, _ = np.linalg.qr(np.column_stack([r, rng.normal(size=(n, n - 1))]))
    B[:, 0] *= np.sign(B[:, 0] @ r)
    Q = B.T
    E, V = np.linalg.eigh(Q @ H @ Q.T)
    return H, r, Q, E, V


def test_ftlm_ltlm_closed_forms():
    """Whole-space Krylov data: FTLM = <r|e^{-bH} O|r>, LTLM = <r|e^{-bH/2} O
    e^{-bH/2}|r>, c_0 = <I>_r = <r|e^{-bH}|r>, with a non-Hermitian complex O
    so that a transposed or conjugated overlap fails."""
    from scipy.linalg import expm
    n = 9
    H, r, Q, E, V = _full_krylov_data(n, 3)
    rng = np.random.default_rng(4)
    Op = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    betas = np.array([0.0, 0.4, 1.3])
    f_res, c0 = O.ftlm_static({"O": Op, "Os": sp.csr_matrix(Op)}, E, V, Q, beta=betas)
    l_res, Id = O.ltlm_static({"O": Op}, E, V, Q, beta=betas)
    for k, b in enumerate(betas):
        eb = expm(-b * H)
        eh = expm(-b * H / 2)
        np.testing.assert_allclose(f_res["O"][k], r @ eb @ Op @ r, rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(f_res["Os"][k], r @ eb @ Op @ r, rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(l_res["O"][k], r @ eh @ Op @ eh @ r, rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(c0[k], r @ eb @ r, rtol=1e-12)
        np.testing.assert_allclose(Id[k], r @ eb @ r, rtol=1e-12)
    # the two differ for a non-commuting O, so neither formula passes for the other
    assert abs(f_res["O"][2] - l_res["O"][2]) > 1e-3


def test_ftlm_ltlm_beta_zero_any_krylov_size():
    """At b = 0 both reduce to <r|O|r> whatever the Krylov size (V is
    orthogonal); complex Hermitian H and complex r, so a missing conjugate fails."""
    n, m = 40, 12
    rng = np.random.default_rng(5)
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    H = (X + X.conj().T) / 2
    r = rng.normal(size=n) + 1j * rng.normal(size=n)
    r /= np.linalg.norm(r)
    Op = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    E, V, Q, _, _ = O.lanczos_full(lambda v: H @ v, r, m)
    f_res, c0 = O.ftlm_static({"O": Op}, E, V, Q, beta=0.0)
    l_res, Id = O.ltlm_static({"O": Op}, E, V, Q, beta=0.0)
    want = np.vdot(r, Op @ r)
    np.testing.assert_allclose(f_res["O"], want, rtol=1e-10)
    np.testing.assert_allclose(l_res["O"], want, rtol=1e-10)
    np.testing.assert_allclose(c0, 1.0, rtol=1e-12)
    np.testing.assert_allclose(Id, 1.0, rtol=1e-12)


def test_c_oracle_table_form_and_real_csr_taylor():
    """The C restatement's timed forms: the sector-table H.v equals its
    orbit-search H.v (same matrix elements through another lookup); the real
    CSR (+ diagonal) 1-norm equals the dense column sum and its Taylor loop
    equals dense expm."""
    from scipy.linalg import expm
    from oracle import c_oracle as C
    L = 14
    m = chain_maps(L)
    g = C.Group(L, [m["P"], m["Z"], m["T"]], [2, 2, L], [0, 0, 0])
    b, n = g.make_basis(L // 2)
    tab = C.table_build(g, L // 2, b)
    terms = flat_terms(xxz_static(L, Delta=0.6))
    v = np.random.default_rng(1).normal(size=len(b))
    np.testing.assert_allclose(C.hv_table_real(g, b, n, tab, terms, v), C.hv_real(g, b, n, terms, v), atol=1e-13)
    rng = np.random.default_rng(2)
    N = 50
    A = sp.random(N, N, density=0.15, random_state=3)
    A = (A + A.T).tocsr()
    D = rng.normal(size=N)
    M = (A + 0.4 * sp.diags(D)).toarray()
    mu = np.trace(M) / N
    for idx in (np.int32, np.int64):
        R = C.RealCSR(A.indptr.astype(idx), A.indices.astype(idx), A.data, D)
        np.testing.assert_allclose(R.one_norm(0.4, mu), np.abs(M - mu * np.eye(N)).sum(axis=0).max(), rtol=1e-13)
        w = rng.normal(size=N) + 1j * rng.normal(size=N)
        np.testing.assert_allclose(R.expm(w, 2, 40, 2.0 ** -53, mu, -0.6j, 0.4), expm(-0.6j * M) @ w, atol=1e-13)


@pytest.mark.parametrize("Nf", [3, 4])
def test_fermion_momentum_sectors_block_diagonalise(Nf):
    """Spinless fermions on a ring (hopping c+_i c_j - c_i c+_j and nn), all
    momentum sectors: P^H P = 1, the reduced operator equals P^H H_full P,
    and the union of the sector spectra is the full spectrum.  With an even
    number of fermions a translation permutes them with sign -1 across the
    boundary, so dropping the permutation sign of the maps breaks the last
    identity (checked: error 1.9 at Nf = 4); an odd number exercises the
    sign-free case."""
    L = 8
    T = [(i + 1) % L for i in range(L)]
    static = ([("+-", (i, (i + 1) % L), -1.0) for i in range(L)] + [("-+", (i, (i + 1) % L), 1.0) for i in range(L)]
              + [("nn", (i, (i + 1) % L), 0.7) for i in range(L)])
    full = O.full_basis_states(L, Nf)
    Hf = O.full_hamiltonian(full, L, static, fermion=True)
    np.testing.assert_allclose(Hf, Hf.conj().T, atol=0)
    Es = []
    for k in range(L):
        g = O.Group(L, [T], [L], [k], fermion=True)
        b, n = O.make_basis(L, g, Np=Nf)
        if len(b) == 0:
            continue
        P = O.projector(b, n, g, full)
        np.testing.assert_allclose(P.conj().T @ P, np.eye(len(b)), atol=1e-12)
        Hr = O.reduced_op(b, n, g, static, fermion=True)
        np.testing.assert_allclose(Hr, P.conj().T @ Hf @ P, atol=1e-12)
        Es.extend(np.linalg.eigvalsh(Hr))
    np.testing.assert_allclose(np.sort(Es), np.linalg.eigvalsh(Hf), atol=1e-10)
