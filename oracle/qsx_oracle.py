"""numpy restatement of the reference's symmetry-reduced basis, matrix elements,
Taylor-series expm and Lanczos — TEST INFRASTRUCTURE ONLY (the oracle).

Never imported by the product package.  Every function cites the reference
file:line whose behaviour it restates (paths relative to the reference root).
Pins (checked in ``tests/test_oracle_*.py``):

* ``process_map``/``Group``      — periods of translation/parity/inversion maps
  (closed form), orbit sizes x stabiliser = |G| (group theory).
* ``make_basis``                  — brute force over all Fock states; sector
  sizes by the character formula (C2 = 361,270 at L=28 k=0,p=0,z=0);
  sum over all k sectors == C(L, L/2); FHM printout order (FHM.ipynb cell 3).
* ``reduced_op`` (A.5 formula)   — equals P^H H_full P built from ``projector``
  (get_proj convention, base_general.py:571-575) for every sector of small
  chains/ladders; SSH k-block printout (SSH.ipynb cell 5: -0.65-0.78j);
  2x3 ladder inplace test (test_inplace_op.py) structure.
* ``apply_term``                  — explicit 2x2 spin matrices (spin.py:66-110
  docstring) and fermion anticommutation {c_i, c_j^+} = delta_ij on full Fock
  space; Hubbard 4x4 (2,2) literature ground energies (test_FHM_energies.py:21-33).
* ``expm_multiply``               — dense ``scipy.linalg.expm`` on small
  matrices to 1e-13; scipy's own ``expm_multiply`` (test_expm_multiply_parallel.py).
* ``lanczos_full``                — dense ``eigh`` ground energy to 1e-10
  (test_Lanczos.py:56-70); Krylov orthonormality.
* ``mlp_*``                       — finite differences of the loss.
* ``op_shift_sector``/``get_amp`` — built on ``projector`` (pinned above);
  P^H P = 1 and the shifted product of a symmetric operator equals the
  reduced operator (tests/test_oracle_pins.py).
* ``ftlm_static``/``ltlm_static`` — with Krylov data spanning the whole space
  they equal <r|e^{-bH} O|r> and <r|e^{-bH/2} O e^{-bH/2}|r> for a
  non-Hermitian O (dense ``expm``), and <r|O|r> at b = 0 for any Krylov size
  (tests/test_oracle_pins.py).  LTLM's truncation of Krylov vectors whose
  weight falls below machine epsilon restates _LTLM.py:95-101 and is pinned
  only where it removes nothing.
"""

import itertools
import math
from fractions import Fraction

import numpy as np
import scipy.sparse as sp

__all__ = [
    "process_map", "order_maps", "Group", "make_basis", "basis_dtype",
    "apply_term", "reduced_op", "projector", "full_basis_states",
    "op_coo", "make_csr", "inplace_op", "index_of",
    "expm_multiply", "fragment_3_1_partition", "lanczos_full",
    "csr_trace", "csr_1_norm", "mlp_forward", "mlp_backward", "op_shift_sector", "get_amp",
    "ftlm_static", "ltlm_static",
]


# ----------------------------------------------------------------------------
# symmetry maps
# ----------------------------------------------------------------------------
def process_map(m):
    """Site map -> (perm, inv, period).  Restates ``process_map``
    (src/quspin/basis/basis_general/base_general.py:22-51): entry ``-(j+1)``
    means site ``j`` with inversion; the period is the number of applications
    that return every site label (with its inversion flag) to itself."""
    m = np.asarray(m, dtype=np.int64)
    inv = m < 0
    perm = np.where(inv, -(m + 1), m)
    if np.any(np.sort(perm) - np.arange(len(m))):
        raise ValueError("map must be a one-to-one site mapping.")
    sites = np.arange(len(m), dtype=np.int64)
    order = sites.copy()
    per = 0
    while True:
        sites = sites.copy()
        sites[inv] = -(sites[inv] + 1)
        sites = sites[perm]
        per += 1
        if np.array_equal(order, sites):
            break
    return perm, inv, per


def order_maps(blocks, block_order=None):
    """Order symmetry blocks and drop identity maps the way
    ``basis_general.__init__`` does (base_general.py:135-203): ascending period
    (stable), or ``reversed(block_order)``; period-1 maps are looked for among
    all but the last item and popped in ascending index order (quirk kept).

    ``blocks``: dict name -> (map, q).  Returns list of (name, map, per, q)."""
    items = {k: (np.asarray(v[0]), process_map(v[0])[2], int(v[1]))
             for k, v in blocks.items() if v is not None}
    if block_order is None:
        sorted_items = sorted(items.items(), key=lambda kv: kv[1][1])
    else:
        bo = list(block_order)
        bo.reverse()
        sorted_items = [(k, items[k]) for k in bo]
    remove = [i for i, it in enumerate(sorted_items[:-1]) if it[1][1] == 1]
    lst = list(sorted_items)
    for i in sorted(remove):
        lst.pop(i)
    return [(k, v[0], v[1], v[2]) for k, v in lst]


def apply_map(states, N, perm, inv):
    """Pull action (g.s)[i] = s[perm[i]] xor inv[i]; site i <-> bit N-1-i
    (sphinx/source/tutorials/user_basis.rst:56-79).  ``states``: uint64 array."""
    s = np.asarray(states, dtype=np.uint64)
    out = np.zeros_like(s)
    one = np.uint64(1)
    for i in range(N):
        b = (s >> np.uint64(N - 1 - int(perm[i]))) & one
        if inv[i]:
            b ^= one
        out |= b << np.uint64(N - 1 - i)
    return out


def fermion_map_sign(states, N, perm):
    """Sign of re-ordering the creation operators after the pull map (mode j
    precedes mode k iff j < k).  A fermion on old site perm[i] lands on site i;
    the sign is the parity of inversions of the new positions listed in old
    order.  (Fermionic symmetry signs of the general cores, fermion.py:1-6;
    only spectra pin the convention.)"""
    s = np.asarray(states, dtype=np.uint64)
    newpos = np.empty(N, dtype=np.int64)
    newpos[np.asarray(perm)] = np.arange(N)
    occ = [((s >> np.uint64(N - 1 - j)) & np.uint64(1)).astype(bool) for j in range(N)]
    parity = np.zeros(s.shape, dtype=np.int64)
    for a in range(N):
        for b in range(a + 1, N):
            if newpos[a] > newpos[b]:
                parity += (occ[a] & occ[b]).astype(np.int64)
    return np.where(parity % 2 == 1, -1, 1).astype(np.int64)


class Group:
    """Abelian group generated by the ordered maps.  Elements enumerated with
    the first map outermost: g = Q_{M-1}^{l_{M-1}} ... Q_0^{l_0}.

    Character chi(g) = exp(+2 pi i sum_k q_k l_k / m_k) (spin.py:20 eigenvalue
    exp(-2 pi i q/m) of Q, projector weight on g.r, App. A.4 of the survey)."""

    def __init__(self, N, maps, pers, qs, fermion=False):
        self.N = N
        self.maps = [np.asarray(m) for m in maps]
        self.pers = [int(p) for p in pers]
        self.qs = [int(q) for q in qs]
        self.fermion = fermion
        self.pi = [process_map(m)[:2] for m in self.maps]
        if fermion and any(inv.any() for _, inv in self.pi):
            raise NotImplementedError("fermion maps with inversion")
        self.elements = list(itertools.product(*[range(p) for p in self.pers]))
        self.G = len(self.elements)
        ch = []
        for ls in self.elements:
            fr = sum((Fraction(q * l, p) for q, l, p in zip(self.qs, ls, self.pers)),
                     Fraction(0)) % 1
            ch.append(_exact_phase(fr))
        self.chars = np.array(ch, dtype=np.complex128)

    def orbit(self, states):
        """(|G|, n) images and (|G|, n) fermion signs of ``states``."""
        s = np.asarray(states, dtype=np.uint64)
        if self.G == 0 or len(self.maps) == 0:
            return s[None, :].copy(), np.ones((1, s.size), dtype=np.int64)
        imgs = np.empty((self.G, s.size), dtype=np.uint64)
        sg = np.ones((self.G, s.size), dtype=np.int64)
        for e, ls in enumerate(self.elements):
            t = s.copy()
            sign = np.ones(s.size, dtype=np.int64)
            for k, l in enumerate(ls):
                perm, inv = self.pi[k]
                for _ in range(l):
                    if self.fermion:
                        sign *= fermion_map_sign(t, self.N, perm)
                    t = apply_map(t, self.N, perm, inv)
            imgs[e] = t
            sg[e] = sign
        return imgs, sg

    def representative(self, states):
        """Orbit maximum, exponents of the first element reaching it, and its
        fermion sign (``representative`` base_general.py:1035-1117; max not
        min, survey App. A.2)."""
        imgs, sg = self.orbit(states)
        arg = np.argmax(imgs, axis=0)
        cols = np.arange(imgs.shape[1])
        rep = imgs[arg, cols]
        if len(self.maps) == 0:
            g = np.zeros((imgs.shape[1], 0), dtype=np.int32)
        else:
            g = np.array([self.elements[a] for a in arg], dtype=np.int32)
        return rep, g, sg[arg, cols].astype(np.int8)

    def normalization(self, states):
        """n = sum over stabiliser of chi(h) sign(h) (``normalization``
        base_general.py:1119-1181; 0 outside the sector)."""
        s = np.asarray(states, dtype=np.uint64)
        imgs, sg = self.orbit(s)
        if len(self.maps) == 0:
            return np.ones(s.size, dtype=np.int64)
        stab = imgs == s[None, :]
        tot = (stab * sg * self.chars[:, None]).sum(axis=0)
        return np.rint(tot.real).astype(np.int64)

    def phase_of(self, g):
        idx = 0
        for k, p in enumerate(self.pers):
            idx = idx * p + int(g[k])
        return self.chars[idx] if len(self.pers) else 1.0


def _exact_phase(fr):
    if fr == 0:
        return 1.0 + 0j
    if fr == Fraction(1, 2):
        return -1.0 + 0j
    if fr == Fraction(1, 4):
        return 1j
    if fr == Fraction(3, 4):
        return -1j
    a = 2 * math.pi * float(fr)
    return complex(math.cos(a), math.sin(a))


# ----------------------------------------------------------------------------
# basis
# ----------------------------------------------------------------------------
def basis_dtype(N):
    """``get_basis_type(N, None, 2)``: uint32 up to 32 bits, else uint64."""
    return np.uint32 if N <= 32 else np.uint64


def full_basis_states(N, Np=None, spinful_L=None, double_occupancy=True):
    """All Fock states of the particle sector(s), descending.
    ``Np``: int, list of ints, or for spinful (``spinful_L`` set) a list of
    (Nup, Ndown) tuples."""
    if spinful_L is not None:
        L = spinful_L
        out = []
        sectors = Np if isinstance(Np, list) else [Np]
        for s in range(1 << (2 * L)):
            up, dn = s >> L, s & ((1 << L) - 1)
            if Np is not None and (bin(up).count("1"), bin(dn).count("1")) not in [tuple(x) for x in sectors]:
                continue
            if not double_occupancy and (up & dn):
                continue
            out.append(s)
        return np.array(sorted(out, reverse=True), dtype=np.uint64)
    allowed = None if Np is None else (set(Np) if isinstance(Np, (list, tuple)) else {Np})
    out = [s for s in range(1 << N) if allowed is None or bin(s).count("1") in allowed]
    return np.array(sorted(out, reverse=True), dtype=np.uint64)


def make_basis(N, group, Np=None, spinful_L=None, double_occupancy=True):
    """Representatives (orbit max with n > 0), descending, and n.
    Restates ``basis_general.make`` (base_general.py:817-899) +
    ``_core.make_basis``: enumerate the particle sector, keep s == rep(s) with
    normalization > 0, sort descending."""
    cand = full_basis_states(N, Np, spinful_L, double_occupancy)
    if cand.size == 0:
        return cand.astype(basis_dtype(N)), np.zeros(0, dtype=np.int64)
    rep, _, _ = group.representative(cand)
    keep = rep == cand
    cand = cand[keep]
    n = group.normalization(cand)
    keep = n > 0
    return cand[keep].astype(basis_dtype(N)), n[keep]


def index_of(basis, states):
    """Position in a descending basis, -1 if absent (``_get_basis_index``)."""
    b = np.asarray(basis, dtype=np.uint64)[::-1]
    s = np.asarray(states, dtype=np.uint64)
    pos = np.searchsorted(b, s)
    pos_c = np.minimum(pos, b.size - 1)
    ok = (b.size > 0) & (b[pos_c] == s) if b.size else np.zeros(s.shape, bool)
    return np.where(ok, b.size - 1 - pos_c, -1)


# ----------------------------------------------------------------------------
# local operator strings
# ----------------------------------------------------------------------------
# (factor on bit 0, factor on bit 1, flips) per character.
_HCB = {
    "I": (1.0, 1.0, False),
    "z": (-0.5, 0.5, False),
    "n": (0.0, 1.0, False),
    "+": (1.0, 0.0, True),
    "-": (0.0, 1.0, True),
    "x": (0.5, 0.5, True),
    "y": (-0.5j, 0.5j, True),
}
_FERMION = {
    "I": (1.0, 1.0, False),
    "z": (-0.5, 0.5, False),
    "n": (0.0, 1.0, False),
    "+": (1.0, 0.0, True),
    "-": (0.0, 1.0, True),
    "x": (1.0, 1.0, True),
    "y": (1.0j, -1.0j, True),
}


def popcount64(x):
    """Bit count of each uint64 (SWAR)."""
    x = np.asarray(x, dtype=np.uint64).copy()
    m1, m2, m4 = np.uint64(0x5555555555555555), np.uint64(0x3333333333333333), np.uint64(0x0F0F0F0F0F0F0F0F)
    x = x - ((x >> np.uint64(1)) & m1)
    x = (x & m2) + ((x >> np.uint64(2)) & m2)
    x = (x + (x >> np.uint64(4))) & m4
    return ((x * np.uint64(0x0101010101010101)) >> np.uint64(56)).astype(np.int64)


def apply_term(states, opstr, indx, N, fermion=False):
    """Apply an operator string right-to-left to Fock states.
    hcb/spin-1/2 (``pauli=0``): z=+-1/2, n, +, -, x=sigma^x/2, y=sigma^y/2
    (spin.py:66-110 docstring; base_hcb.py:81 allowed ops).
    Fermions: +=c^dag, -=c, n, z=n-1/2, x=c+c^dag, y=-i(c-c^dag)
    (fermion.py:31-40 table) with Jordan-Wigner sign (-1)^(#occupied modes
    of lower index) per c/c^dag (convention: modes ordered by site index).
    Returns (ME complex, new states)."""
    s = np.asarray(states, dtype=np.uint64).copy()
    me = np.ones(s.shape, dtype=np.complex128)
    tab = _FERMION if fermion else _HCB
    one = np.uint64(1)
    for ch, site in reversed(list(zip(opstr, indx))):
        pos = N - 1 - int(site)
        b = ((s >> np.uint64(pos)) & one).astype(bool)
        f0, f1, flip = tab[ch]
        me *= np.where(b, f1, f0)
        if fermion and flip:
            cnt = popcount64(s >> np.uint64(pos + 1))
            me *= np.where(cnt % 2 == 1, -1.0, 1.0)
        if flip:
            s ^= one << np.uint64(pos)
    return me, s


# ----------------------------------------------------------------------------
# reduced matrix elements, projector, COO / CSR
# ----------------------------------------------------------------------------
def op_coo(basis, n, group, opstr, indx, J, fermion=False):
    """(ME, row, col) of J*opstr in the reduced basis, one element per basis
    state (column), zeros dropped.  Restates ``_Op`` (base_general.py:292-352)
    with the reduced element (survey App. A.5):
      H[r', r] += J * m * sign * chi(g) * sqrt(n_r' / n_r),   g.s' = r'."""
    b = np.asarray(basis, dtype=np.uint64)
    me, s2 = apply_term(b, opstr, indx, group.N, fermion)
    rep, g, sgn = group.representative(s2)
    row = index_of(b, rep)
    ok = (row >= 0) & (me != 0)
    col = np.arange(b.size)
    phase = np.array([group.phase_of(gi) for gi in g], dtype=np.complex128) \
        if len(group.pers) else np.ones(b.size, dtype=np.complex128)
    nn = np.asarray(n, dtype=np.float64)
    n_row = np.where(row >= 0, nn[np.maximum(row, 0)], 1.0)
    val = J * me * sgn * phase * np.sqrt(n_row / nn)
    val = np.where(ok, val, 0)
    keep = val != 0
    return val[keep], row[keep], col[keep]


def make_csr(basis, n, group, static, dtype=np.complex128, fermion=False):
    """CSR of sum_terms op_coo, duplicates summed (``_make_matrix``
    base.py:265-304; scipy COO->CSR).  ``static`` = [(opstr, indx, J), ...]."""
    Ns = len(basis)
    H = sp.csr_matrix((Ns, Ns), dtype=np.complex128)
    for opstr, indx, J in static:
        ME, row, col = op_coo(basis, n, group, opstr, indx, J, fermion)
        H = H + sp.csr_matrix((ME, (row, col)), shape=(Ns, Ns))
    H.sum_duplicates()
    H.sort_indices()
    if np.dtype(dtype).kind != "c":
        assert np.abs(H.data.imag).max(initial=0) < 1e-13
        H = H.real
    return H.astype(dtype)


def inplace_op(v_in, basis, n, group, op_list, transposed=False,
               conjugated=False, a=1.0, v_out=None, fermion=False):
    """v_out += a * Op v_in, one scatter pass per term (base.py:125-162 +
    ``_coo_dot`` base.py:14-23; per-term loop base_general.py:392-408)."""
    v_in = np.asarray(v_in)
    vi = v_in.reshape(len(basis), -1)
    if v_out is None:
        v_out = np.zeros(vi.shape, dtype=np.result_type(v_in.dtype, np.complex128))
    vo = v_out.reshape(len(basis), -1)
    for opstr, indx, J in op_list:
        ME, row, col = op_coo(basis, n, group, opstr, indx, a * J, fermion)
        if transposed:
            row, col = col, row
        if conjugated:
            ME = ME.conj()
        for r, c, m in zip(row, col, ME):
            vo[r] += m * vi[c]
    return v_out


def projector(basis, n, group, full_states):
    """P (full x reduced): column r = c_r sum_g chi(g) sign(g) |g.r>,
    c_r = 1/sqrt(n_r |G|) (``get_proj`` base_general.py:521-582, coefficient
    :571-575)."""
    b = np.asarray(basis, dtype=np.uint64)
    imgs, sg = group.orbit(b)
    Gsz = max(group.G, 1) if len(group.pers) else 1
    c = 1.0 / np.sqrt(np.asarray(n, dtype=np.float64) * Gsz)
    fs = np.asarray(full_states, dtype=np.uint64)
    P = np.zeros((fs.size, b.size), dtype=np.complex128)
    chars = group.chars if len(group.pers) else np.ones(1)
    for e in range(imgs.shape[0]):
        rows = index_of(fs, imgs[e])
        assert np.all(rows >= 0)
        np.add.at(P, (rows, np.arange(b.size)), c * chars[e] * sg[e])
    return P


def op_shift_sector(basis2, n2, group2, basis1, n1, group1, full_states, N, op_list, v_in, fermion=False):
    """P_2^H Op P_1 v_in: the "formally equivalent" computation the
    reference's ``Op_shift_sector`` docstring states
    (base_general.py:419-427; checked that way by
    test/test_Op_shift_sector.py:31-46), with dense projectors and the brute
    force operator on ``full_states``."""
    P1 = projector(basis1, n1, group1, full_states)
    P2 = projector(basis2, n2, group2, full_states)
    H = full_hamiltonian(full_states, N, op_list, fermion)
    return P2.conj().T @ (H @ (P1 @ np.asarray(v_in)))


def get_amp(basis, n, group, full_states, states):
    """psi_s / psi_r^sym = P[s, rep(s)] read off the dense projector
    (``get_amp`` base_general.py:1183-1269, which test/test_get_amp.py:127-150
    checks against ``get_vec``); 0 for states whose representative is not in
    the basis.  Returns (C, representatives)."""
    P = projector(basis, n, group, full_states)
    st = np.asarray(states, dtype=np.uint64)
    rep = group.representative(st)[0] if len(group.pers) else st
    rows = index_of(full_states, st)
    cols = index_of(basis, rep)
    C = np.where(cols >= 0, P[rows, np.maximum(cols, 0)], 0.0)
    return C, rep


def full_hamiltonian(full_states, N, static, fermion=False):
    """Dense H on an unreduced Fock space (brute force)."""
    fs = np.asarray(full_states, dtype=np.uint64)
    H = np.zeros((fs.size, fs.size), dtype=np.complex128)
    for opstr, indx, J in static:
        me, s2 = apply_term(fs, opstr, indx, N, fermion)
        rows = index_of(fs, s2)
        ok = (me != 0) & (rows >= 0)
        np.add.at(H, (rows[ok], np.nonzero(ok)[0]), J * me[ok])
    return H


def reduced_op(basis, n, group, static, fermion=False):
    """Dense reduced H from the A.5 element formula."""
    return make_csr(basis, n, group, static, np.complex128, fermion).toarray()


# ----------------------------------------------------------------------------
# CSR helpers and Taylor-series expm (Al-Mohy & Higham, Alg. 3.2)
# ----------------------------------------------------------------------------
def csr_trace(A):
    """``_wrapper_csr_trace`` (EMP.html:196): sum of diagonal entries."""
    return A.diagonal().sum()


def csr_1_norm(A, shift):
    """``_wrapper_csr_1_norm(..., shift)`` (EMP.html:200-202): max column
    abs-sum of A + shift*I (scipy ``_exact_1_norm`` SCIPY:27-35)."""
    B = A + shift * sp.eye(A.shape[0], format="csr", dtype=np.result_type(A.dtype, np.asarray(shift).dtype))
    return float(np.abs(B).sum(axis=0).max()) if A.shape[0] else 0.0


def fragment_3_1_partition(A, a, mu, A_1_norm, tol, ell=2):
    """(m_star, s) exactly as ``ExpmMultiplyParallel._calculate_partition``
    (EMP.html:377-385): scipy ``_fragment_3_1`` with n0=1 and the lazy norm
    info of a*(A - mu I)."""
    from scipy.sparse.linalg._expm_multiply import _fragment_3_1
    from scipy.sparse.linalg import LinearOperator, onenormest

    if abs(a) * A_1_norm == 0:
        return 0, 1

    class _Info:
        def __init__(self):
            self._d = {}

        def onenorm(self):
            return abs(a) * A_1_norm

        def d(self, p):
            if p not in self._d:
                dt = np.result_type(A.dtype, np.asarray(a).dtype, np.float64)
                op = LinearOperator(A.shape, dtype=dt,
                                    matvec=lambda v: a * (A.dot(v) - mu * v),
                                    rmatvec=lambda v: np.conj(a) * (A.conj().T.dot(v) - np.conj(mu) * v))
                self._d[p] = onenormest(op ** p) ** (1.0 / p)
            return self._d[p]

        def alpha(self, p):
            return max(self.d(p), self.d(p + 1))

    return _fragment_3_1(_Info(), 1, tol, ell=ell)


def _inf_norm(B):
    B = np.asarray(B)
    if B.ndim == 1:
        return np.abs(B).max(initial=0.0)
    return np.abs(B).sum(axis=1).max(initial=0.0)


def expm_multiply(A, v, a=1.0, dtype=None, tol=None, partition=None):
    """e^{aA} v by the Taylor loop of ``_expm_multiply_simple_core``
    (SCIPY:280-303) applied to (A - mu I) with mu = tr(A)/N, as
    ``ExpmMultiplyParallel`` does (EMP.html:152-203, :347-372)."""
    N = A.shape[0]
    if dtype is None:
        dtype = np.result_type(A.dtype, np.min_scalar_type(a), np.float64)
    if tol is None:
        tol = np.finfo(dtype).eps / 2
    mu = csr_trace(A) / N
    if partition is None:
        nrm = csr_1_norm(A, -mu)
        m_star, s = fragment_3_1_partition(A, a, mu, nrm, tol)
    else:
        m_star, s = partition
    vdt = np.result_type(dtype, np.asarray(v).dtype)
    F = np.array(v, dtype=vdt, copy=True)
    a = np.asarray(a, dtype=vdt)
    mu = np.asarray(mu, dtype=vdt)
    eta = np.exp(a * mu / float(s))
    B = F.copy()
    for _ in range(s):
        c1 = _inf_norm(B)
        for j in range(m_star):
            coeff = a / float(s * (j + 1))
            B = coeff * (A.dot(B) - mu * B)
            c2 = _inf_norm(B)
            F = F + B
            if c1 + c2 <= tol * _inf_norm(F):
                break
            c1 = c2
        F = eta * F
        B = F
    return F, (m_star, s)


# ----------------------------------------------------------------------------
# Lanczos (src/quspin/tools/lanczos/_lanczos_utils.py:77-227)
# ----------------------------------------------------------------------------
def lanczos_full(matvec, v0, m, eps=None):
    """3-term Lanczos: a_i = Re<q_i, r>, b_i = ||r||, break if b_i < eps.
    Returns (E, V, Q[:m]) with E, V from the tridiagonal eigenproblem."""
    from scipy.linalg import eigh_tridiagonal
    v0 = np.asarray(v0)
    n = v0.size
    dtype = v0.dtype
    Q = np.zeros((m, n), dtype=dtype)
    Q[0] = v0
    if eps is None:
        eps = np.finfo(dtype).eps
    qn = np.linalg.norm(Q[0])
    if abs(qn - 1.0) > eps:
        Q[0] /= qn
    a = np.zeros(m)
    b = np.zeros(m)
    r = matvec(Q[0])
    a[0] = np.vdot(Q[0], r).real
    r = r - a[0] * Q[0]
    b[0] = np.linalg.norm(r)
    mm = m
    for i in range(1, m):
        v = Q[i - 1].copy()
        Q[i] = r / b[i - 1]
        r = matvec(Q[i])
        r = r - b[i - 1] * v
        a[i] = np.vdot(Q[i], r).real
        r = r - a[i] * Q[i]
        b[i] = np.linalg.norm(r)
        if b[i] < eps:
            mm = i
            break
    E, V = eigh_tridiagonal(a[:mm], b[:mm - 1])
    return E, V, Q[:mm], a[:mm], b[:mm]


# ----------------------------------------------------------------------------
# FTLM / LTLM (src/quspin/tools/lanczos/_FTLM.py:10-105, _LTLM.py:9-139)
# ----------------------------------------------------------------------------
def ftlm_static(O_mats, E, V, Q, beta=0):
    """One FTLM sample (_FTLM.py:76-103): with w_j(b) = exp(-b E_j) and
    c_i(b) = sum_j V[i,j] V[0,j] w_j(b), <O>_r(b) = sum_i c_i(b) <q_i|O q_0>.
    ``O_mats`` maps names to dense or scipy matrices; ``Q`` holds the Krylov
    vectors as rows.  Returns ({name: <O>_r}, c_0)."""
    beta = np.atleast_1d(np.asarray(beta, dtype=np.float64))
    E = np.asarray(E)
    V = np.asarray(V)
    Q = np.asarray(Q)
    w = np.exp(-E[:, None] * beta[None, :])
    c = (V * V[0][None, :]) @ w
    out = {}
    for name, O in O_mats.items():
        Or = O @ Q[0]
        ov = np.conj(Q) @ Or
        out[name] = np.squeeze((c * ov[:, None]).sum(axis=0))
    return out, np.squeeze(c[0])


def ltlm_static(O_mats, E, V, Q, beta=0):
    """One LTLM sample (_LTLM.py:86-137): keep the first nv Krylov vectors,
    nv = first j with exp(-min(b) E_j) max|V[0]| < eps (all when none);
    M[i,j] = <q_j|O q_i>, A = V^T M V on the kept block, c_j = V[0,j]
    exp(-b E_j / 2), <O>_r = sum_{jl} c_j c_l A_{jl}; also returns
    <I>_r = sum_j V[0,j]^2 exp(-b E_j)."""
    beta = np.atleast_1d(np.asarray(beta, dtype=np.float64))
    E = np.asarray(E)
    V = np.asarray(V)
    Q = np.asarray(Q)
    nv = E.size
    weight = np.exp(-E * beta.min()) * np.abs(V[0]).max()
    small = weight < np.finfo(weight.dtype).eps
    if small.any():
        nv = int(np.nonzero(small)[0][0])
    Qk = Q[:nv]
    Vk = V[:nv, :nv]
    c = Vk[0][:, None] * np.exp(-E[:nv, None] * beta[None, :] / 2.0)
    out = {}
    for name, O in O_mats.items():
        OQ = np.asarray(O @ Qk.T)
        M = (np.conj(Qk) @ OQ).T
        A = Vk.T @ M @ Vk
        out[name] = np.squeeze(np.einsum("jb,lb,jl->b", c, c, A))
    Id = (Vk[0] ** 2) @ np.exp(-E[:nv, None] * beta[None, :])
    return out, np.squeeze(Id)


# ----------------------------------------------------------------------------
# MLP (m, s) predictor reference: 4 -> 64 -> 64 -> 2, tanh hidden, linear out
# ----------------------------------------------------------------------------
def mlp_forward(params, X):
    W1, b1, W2, b2, W3, b3 = params
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    y = h2 @ W3.T + b3
    return y, (X, h1, h2)


def mlp_backward(params, cache, dy):
    """Gradients of sum(dy * y) w.r.t. params."""
    W1, b1, W2, b2, W3, b3 = params
    X, h1, h2 = cache
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    dh2 = (dy @ W3) * (1 - h2 ** 2)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (1 - h1 ** 2)
    gW1 = dh1.T @ X
    gb1 = dh1.sum(0)
    return [gW1, gb1, gW2, gb2, gW3, gb3]
