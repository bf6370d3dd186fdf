"""Species-factorised device operator for spinful fermions.

In a sector (N_up, N_down) of ``spinful_fermion_basis_general`` without
symmetry maps the descending basis is the grid  index = u * N_d + d  of the
up and down occupations (the up bits are the high bits; basis_general/
fermion.py:466-474 sizes the sector as C(N, N_up) C(N, N_down)).  A term of
the static list whose operators all sit on one species acts as
``H_up (x) 1`` or ``1 (x) H_down`` with the same term evaluated on the
spinless sector of that species: particle-conserving terms carry an even
number of '+'/'-', so the Jordan-Wigner count of the other species cancels.
Terms on both species are admitted when they are diagonal ('n', 'z', 'I')
and go into a static diagonal.

``KronOp`` holds the two factor CSR matrices and the static diagonal in HBM
and exposes the product, the 1-norm of H - mu I + xs diag(D) and the device
time stepping of ``qsx_evolve_kron_steps``; every operation runs in
``libquspin_hip.so``.
"""

import ctypes as C

import numpy as np
import scipy.sparse as sp
from scipy.special import comb

from .. import _lib as L
from .._lib import ptr, stream_handle

_DIAG_OPS = set("nzI")
_FACTOR_OPS = set("+-nzI")


def _vec_dtype_code(t):
    return L.dtype_code(L.np_dtype(t.dtype))


def factorisation_terms(op):
    """Split the static list of ``op`` (a ``quantum_LinearOperator`` or a
    static ``hamiltonian`` on a spinful basis) into up-only, down-only and
    diagonal two-species terms, as lists in the static-list format of the
    spinless bases (up) / (down) and of the spinful basis (mixed).  Raises
    TypeError with the reason when the operator does not factorise."""
    from ..basis.fermion import spinful_fermion_basis_general
    basis = getattr(op, "basis", None)
    if not isinstance(basis, spinful_fermion_basis_general):
        raise TypeError("the factorised form needs a spinful_fermion_basis_general")
    if len(basis._pers):
        raise TypeError("the factorised form needs a basis without symmetry maps")
    Np = basis._Np
    if Np is None or len(Np) != 1 or not basis._double_occupancy:
        raise TypeError("the factorised form needs one (N_up, N_down) sector with double occupancy")
    if getattr(op, "_dynamic", None):
        raise TypeError("the factorised form covers static operators")
    stat = getattr(op, "_static_list", None)
    if stat is None:
        raise TypeError("operator has no static operator list")
    N = basis.N // 2
    up, dn, mixed = [], [], []
    for opstr, indx, J in stat:
        a_opstr, a_indx, a_J = basis._opstr_prepare(opstr, indx, J)
        a_indx = [int(i) for i in a_indx]
        ops = set(a_opstr)
        if ops - _FACTOR_OPS:
            raise TypeError("operator string {!r} is outside the pcon alphabet +-nzI".format(opstr))
        if all(i < N for i in a_indx):
            up.append([a_opstr, [[a_J] + a_indx]])
        elif all(i >= N for i in a_indx):
            dn.append([a_opstr, [[a_J] + [i - N for i in a_indx]]])
        elif ops <= _DIAG_OPS:
            mixed.append([opstr, [[J] + list(indx)]])
        else:
            raise TypeError("off-diagonal term {!r} acts on both species".format(opstr))
    return up, dn, mixed


def _spinless_csr(N, Nf, terms, dtype, transposed, conjugated):
    """Host CSR of the terms on the spinless sector (N, Nf) in ``dtype``
    (zero matrix for no terms); assembled in complex128, so that a real
    dtype for a complex factor is refused rather than truncated."""
    from ..basis.fermion import spinless_fermion_basis_general
    from .quantum_LinearOperator_core import quantum_LinearOperator
    B = spinless_fermion_basis_general(N, Nf=Nf)
    if not terms:
        return sp.csr_matrix((B.Ns, B.Ns), dtype=dtype)
    Q = quantum_LinearOperator(terms, basis=B, dtype=np.complex128, check_symm=False, check_herm=False,
                               check_pcon=False)
    M = Q._qsx_operator().to_csr(np.complex128).to_scipy().tocsr()
    if conjugated:
        M = M.conj()
    if transposed:
        M = M.T.tocsr()
    if np.dtype(dtype).kind == "f":
        if np.any(M.data.imag != 0):
            raise TypeError("complex factor in a real dtype")
        M = M.real
    M = M.astype(dtype)
    M.sort_indices()
    return M


class KronOp:
    """H = H_up (x) 1 + 1 (x) H_down + diag(sdiag) on a spinful sector.

    ``op``: a ``quantum_LinearOperator`` (or a static ``hamiltonian``) on a
    spinful basis without symmetry maps; ``dtype``: the factors' dtype (a
    real dtype for a real operator halves the bytes of the factors)."""

    is_matrix_free = False

    def __init__(self, op, dtype=None):
        up, dn, mixed = factorisation_terms(op)
        basis = op.basis
        dtype = np.dtype(dtype if dtype is not None else op.dtype)
        if dtype not in (np.float32, np.float64, np.complex64, np.complex128):
            raise TypeError("unsupported dtype {}".format(dtype))
        self.dtype = dtype
        tr = bool(getattr(op, "_transposed", False))
        cj = bool(getattr(op, "_conjugated", False))
        N = basis.N // 2
        Nup, Ndn = basis._Np[0]
        self.Nu, self.Nd = comb(N, Nup, exact=True), comb(N, Ndn, exact=True)
        self.N = self.Nu * self.Nd
        if basis.Ns != self.N:
            raise TypeError("basis size {} is not C(N, N_up) C(N, N_down) = {}".format(basis.Ns, self.N))
        Hu = _spinless_csr(N, Nup, up, dtype, tr, cj)
        Hd = _spinless_csr(N, Ndn, dn, dtype, tr, cj)
        self.host = (Hu, Hd)
        # the static diagonal: mixed diagonal terms and a diagonal set on the operator
        sd = None
        if mixed:
            from .quantum_LinearOperator_core import quantum_LinearOperator
            Q = quantum_LinearOperator(mixed, basis=basis, dtype=np.complex128, check_symm=False, check_herm=False,
                                       check_pcon=False)
            sd = L.to_host(Q._qsx_operator().diagonal(np.complex128))
        ud = getattr(op, "_user_diag", None)
        if ud is not None:
            u = L.to_host(ud) if L.is_device_tensor(ud) else np.asarray(ud)
            sd = u.astype(np.complex128) if sd is None else sd + u
        if sd is not None:
            if np.any(sd.imag != 0):
                raise TypeError("the factorised form takes a real static diagonal")
            self.sdiag = L.to_device(np.ascontiguousarray(sd.real, np.float64))
        else:
            self.sdiag = None
        self._dev = []
        for M in (Hu, Hd):
            self._dev.append((L.to_device(M.indptr.astype(np.int32)),
                              L.to_device(M.indices.astype(np.int32)) if M.nnz else L.empty(1, np.int32),
                              L.to_device(M.data.astype(dtype)) if M.nnz else L.empty(1, dtype), M.nnz))
        # factor columns: off-diagonal 1-norms and the diagonal (re, im) for the norm kernel
        self._norm_parts = []
        for M in (Hu, Hd):
            C_ = M.tocsc()
            dg = M.diagonal().astype(np.complex128)
            off = np.asarray(abs(C_).sum(axis=0)).ravel() - np.abs(dg)
            self._norm_parts.append((L.to_device(np.maximum(off, 0.0).astype(np.float64)),
                                     L.to_device(np.ascontiguousarray(np.stack([dg.real, dg.imag], 1).ravel()))))
        self._nwork = L.empty(1024, np.float64)
        self._nout = L.empty(1, np.float64)
        (ipu, ixu, dau, nnu), (ipd, ixd, dad, nnd) = self._dev
        h = L._vp()
        L.check(L.load().qsx_kron_create(ptr(ipu), ptr(ixu), ptr(dau), 0, self.Nu, nnu, ptr(ipd), ptr(ixd), ptr(dad),
                                         0, self.Nd, nnd, ptr(self.sdiag) if self.sdiag is not None else None,
                                         L.dtype_code(dtype), C.byref(h)), "qsx_kron_create")
        self.h = h
        self._trace = complex(Hu.diagonal().sum()) * self.Nd + complex(Hd.diagonal().sum()) * self.Nu + (
            complex(sd.sum()) if sd is not None else 0.0)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                L.load().qsx_kron_destroy(h)
            except Exception:
                pass
            self.h = None

    def set_mode(self, mode):
        """Kernel form: 0 by size, 1 one pass, 2 two passes (row pass over
        H_down, L2-resident column-strip pass over H_up)."""
        L.check(L.load().qsx_kron_set_mode(self.h, int(mode)), "qsx_kron_set_mode")

    @property
    def shape(self):
        return (self.N, self.N)

    @property
    def Ns(self):
        return self.N

    def apply(self, x, y, a=1.0, beta=0, xdiag=None, xs=0.0):
        """y = a (H + xs diag(xdiag)) x (+ y when beta); one vector."""
        if x.numel() != self.N or y.numel() != self.N:
            raise ValueError("the factorised product takes one vector of {} elements".format(self.N))
        ac, xc = complex(a), complex(xs)
        L.check(L.load().qsx_kron_matvec(self.h, ptr(xdiag) if xdiag is not None else None, xc.real, xc.imag, ptr(x),
                                         ptr(y), _vec_dtype_code(x), ac.real, ac.imag, int(bool(beta)),
                                         stream_handle()), "qsx_kron_matvec")

    def trace(self):
        return self._trace

    def one_norm(self, shift=0.0, xdiag=None, xs=0.0):
        """||H + shift I + xs diag(xdiag)||_1 (one device pass)."""
        (ou, gu), (od, gd) = self._norm_parts
        sc, xc = complex(shift), complex(xs)
        L.check(L.load().qsx_kron_1_norm(self.h, ptr(ou), ptr(gu), ptr(od), ptr(gd),
                                         ptr(xdiag) if xdiag is not None else None, xc.real, xc.imag, -sc.real,
                                         -sc.imag, ptr(self._nwork), ptr(self._nout), stream_handle()),
                "qsx_kron_1_norm")
        return float(L.to_host(self._nout)[0])

    def to_scipy(self):
        """The assembled operator (host, for checks)."""
        Hu, Hd = self.host
        M = sp.kron(Hu, sp.identity(self.Nd, format="csr")) + sp.kron(sp.identity(self.Nu, format="csr"), Hd)
        if self.sdiag is not None:
            M = M + sp.diags(L.to_host(self.sdiag))
        return M.tocsr()

    def evolve_steps(self, xdiag, rec, nsteps, s_max, m_max, psi, work, vdt, sdt, fail, ctl, graph):
        L.check(L.load().qsx_evolve_kron_steps(self.h, ptr(xdiag) if xdiag is not None else None, int(nsteps),
                                               ptr(rec), int(s_max), int(m_max), ptr(psi), ptr(work),
                                               L.dtype_code(vdt), L.dtype_code(sdt), ptr(fail), ptr(ctl),
                                               int(ctl.numel() * 8), int(bool(graph)), stream_handle()),
                "qsx_evolve_kron_steps")
