"""``quantum_LinearOperator``: matrix-free operator applied by the fused
device H.v kernel (src/quspin/operators/quantum_LinearOperator_core.py).

Construction consolidates the term list (_make_hamiltonian.py:19-41) and
splits off the terms that do not flip any site (their sum is the diagonal,
evaluated on the fly by the kernel).  ``dot``/``_matvec`` keep the
reference's argument conventions, including its quirks:
  * (conjugated, transposed) are passed into the (transposed, conjugated)
    slots of ``inplace_Op`` (quantum_LinearOperator_core.py:335-343, 531-538);
  * ``dot(out=...)`` overwrites ``out`` only through the diagonal product and
    accumulates into it when there is no diagonal (:332-343);
  * with ``out`` given the diagonal is not scaled by ``a``; the diagonal is
    never conjugated (:332-333, :529);
  * ``transpose()``/``conjugate()`` with ``copy=False`` modify the operator.
"""

import numpy as np
from scipy.sparse.linalg import LinearOperator

from .. import _lib as L
from ._devop import MatrixFreeOp
from ._make_hamiltonian import _consolidate_static
from .hamiltonian_core import supported_dtypes, _default_basis


def _check_static(sub_list):
    if isinstance(sub_list, (list, tuple)) and len(sub_list) == 2:
        return isinstance(sub_list[0], str) and isinstance(sub_list[1], (list, tuple, np.ndarray))
    return False


def isquantum_LinearOperator(obj):
    return isinstance(obj, quantum_LinearOperator)


class quantum_LinearOperator(LinearOperator):
    """Matrix-free operator (see module docstring)."""

    def __init__(self, static_list, N=None, basis=None, diagonal=None, check_symm=True, check_herm=True,
                 check_pcon=True, dtype=np.complex128, copy=False, **basis_args):
        if type(static_list) in [list, tuple]:
            for ele in static_list:
                if not _check_static(ele):
                    raise ValueError("quantum_LinearOperator only supports operator string representations.")
        else:
            raise TypeError("expecting list/tuple of lists/tuples containing opstr and list of indx")
        if np.dtype(dtype) not in [np.dtype(d) for d in supported_dtypes]:
            raise TypeError("hamiltonian does not support type: " + str(dtype))
        self._dtype = np.dtype(dtype)
        if N == []:
            raise ValueError("second argument of `quantum_LinearOperator()` canNOT be an empty list.")
        elif type(N) is int and basis is None:
            self._basis = _default_basis(N, **basis_args)
        elif N is None and basis is not None:
            self._basis = basis
        else:
            raise ValueError("expecting integer for N or basis object for basis.")
        self._unique_me = True
        self._transposed = False
        self._conjugated = False
        self._scale = np.array(1.0, dtype=dtype)
        self._ndim = 2
        self._shape = (self._basis.Ns, self._basis.Ns)
        if check_herm:
            self.basis.check_hermitian(static_list, [])
        if check_symm:
            self.basis.check_symm(static_list, [])
        if check_pcon:
            self.basis.check_pcon(static_list, [])
        self._user_diag = None
        if diagonal is not None:
            self.set_diagonal(diagonal, copy=copy)
        self._public_static_list = list(static_list)
        self._static_list = _consolidate_static(static_list)
        self._oplist = self._basis._compile_oplist(self._static_list)
        self._ops = {}

    # -------------------------------------------------------------- props
    @property
    def shape(self):
        return self._shape

    @property
    def basis(self):
        return self._basis

    @property
    def ndim(self):
        return self._ndim

    @property
    def dtype(self):
        return self._dtype

    @property
    def Ns(self):
        return self._shape[0]

    @property
    def static_list(self):
        return self._public_static_list

    @property
    def T(self):
        return self.transpose(copy=True)

    @property
    def H(self):
        return self.getH(copy=True)

    def _has_diagonal(self):
        return self._user_diag is not None or self._oplist.n_diag > 0

    @property
    def diagonal(self):
        """Diagonal matrix elements (user diagonal plus the non-flipping
        terms), or None when there is none."""
        if not self._has_diagonal():
            return None
        d = self._op_for(False, False).diagonal(np.complex128)
        d = L.to_host(d)
        if self._dtype.kind != "c":
            d = d.real
        return d.astype(self._dtype)

    def set_diagonal(self, diagonal, copy=True):
        diagonal = np.asanyarray(diagonal)
        if diagonal.ndim != 1:
            diagonal = diagonal.squeeze()
        if diagonal.shape != (self.Ns,):
            raise ValueError("diagonal must be 1-d array with size Ns.")
        cplx = np.iscomplexobj(diagonal)
        self._user_diag = L.to_device(diagonal.astype(np.complex128 if cplx else np.float64))
        if hasattr(self, "_oplist"):
            # the new array replaces the whole diagonal, the diagonal terms'
            # part included (the reference overwrites ``_diagonal``,
            # quantum_LinearOperator_core.py:231-234)
            off = [t for t in self._static_list if not self._diagonal_term(*t)]
            self._static_list = off
            self._oplist = self._basis._compile_oplist(off)
        self._ops = {}

    def _diagonal_term(self, opstr, indx, J):
        """True for a term that changes no state (its flips cancel), which
        the reference moves into ``_diagonal`` (quantum_LinearOperator_core.py:146-156)."""
        opstr, indx, _ = self._basis._opstr_prepare(opstr, indx, J)
        flip = 0
        for ch, i in zip(opstr, indx):
            if ch in "+-xy":
                flip ^= 1 << int(i)
        return flip == 0

    # ------------------------------------------------------------ kernels
    def _op_for(self, transposed, conjugated):
        """Device operator with effective (transposed, conjugated) flags on
        the off-diagonal terms; the diagonal is never conjugated."""
        key = (bool(transposed), bool(conjugated))
        if key not in self._ops:
            self._ops[key] = MatrixFreeOp(self._basis, self._oplist, self._dtype, transposed=key[0],
                                          conjugated=key[1], diag_conj=False, diag=self._user_diag)
        return self._ops[key]

    def _qsx_operator(self):
        # the swap of B.1: the operator's (conjugated, transposed) land in the
        # (transposed, conjugated) slots
        return self._op_for(self._conjugated, self._transposed)

    def _apply(self, other, out, a, a_diag, beta):
        on_dev = L.is_device_tensor(other)
        odt = L.np_dtype(other.dtype) if on_dev else np.asanyarray(other).dtype
        result_dtype = np.result_type(self._dtype, odt)
        x = L.to_device(other, result_dtype)
        if out is None:
            y = L.empty(tuple(x.shape), result_dtype)
        elif L.is_device_tensor(out):
            y = out
        else:
            y = L.to_device(out, result_dtype)
        op = self._qsx_operator()
        op.apply(x, y, a=a, beta=beta, a_diag=a_diag)
        if on_dev or L.is_device_tensor(out):
            return y
        res = L.to_host(y)
        if out is not None:
            out[...] = res
            return out
        return res

    def dot(self, other, out=None, a=1.0):
        """a*H|other> (quantum_LinearOperator_core.py:288-347)."""
        if out is not None:
            on_dev = L.is_device_tensor(other)
            odt = L.np_dtype(other.dtype) if on_dev else np.asanyarray(other).dtype
            result_dtype = np.result_type(self._dtype, odt)
            out_dt = L.np_dtype(out.dtype) if L.is_device_tensor(out) else out.dtype
            contiguous = True if L.is_device_tensor(out) else out.flags["CARRAY"]
            if not contiguous or (out_dt, tuple(out.shape)) != (result_dtype, tuple(other.shape)):
                raise ValueError("out must be C-congituous writable array with dtype {} and shape {}.".format(
                    result_dtype, out.shape))
            if self._has_diagonal():
                return self._apply(other, out, a=a, a_diag=1.0, beta=0)
            return self._apply(other, out, a=a, a_diag=0.0, beta=1)
        return self._apply(other, None, a=a, a_diag=a, beta=0)

    def _matvec(self, other):
        return self._apply(other, None, a=1.0, a_diag=1.0, beta=0)

    def matvec(self, other):
        return self._matvec(other)

    def _rmatvec(self, other):
        return self.T.conj()._matvec(other)

    def _matmat(self, other):
        return self._matvec(other)

    def __mul__(self, other):
        if np.isscalar(other):
            return NotImplemented
        return self._matvec(other)

    def __rmul__(self, other):
        return NotImplemented

    # ------------------------------------------------------- transforms
    def _copy_flags(self):
        new = object.__new__(quantum_LinearOperator)
        new.__dict__.update(self.__dict__)
        new._ops = {}
        return new

    def transpose(self, copy=False):
        target = self._copy_flags() if copy else self
        target._transposed = not target._transposed
        target._ops = {} if copy else target._ops
        return target

    def conjugate(self, copy=False):
        target = self._copy_flags() if copy else self
        target._conjugated = not target._conjugated
        return target

    def conj(self, copy=False):
        return self.conjugate(copy=copy)

    def getH(self, copy=False):
        return self.conj(copy=copy).transpose(copy=copy)

    def copy(self):
        return self._copy_flags()

    # ------------------------------------------------------ measurements
    def expt_value(self, V, time=0, check=True, enforce_pure=False):
        HV = self._matvec(V)
        if L.is_device_tensor(HV):
            HV = L.to_host(HV)
            V = L.to_host(V)
        V = np.asarray(V)
        if V.ndim == 1:
            return np.vdot(V, HV)
        return np.einsum("ij,ij->j", V.conj(), HV)

    def quant_fluct(self, V, enforce_pure=False, check=True, time=0):
        e = self.expt_value(V)
        HV = self._matvec(V)
        HV = L.to_host(HV) if L.is_device_tensor(HV) else HV
        e2 = np.vdot(HV, HV) if np.ndim(HV) == 1 else np.einsum("ij,ij->j", HV.conj(), HV)
        return e2 - e ** 2

    def eigsh(self, **eigsh_args):
        """ARPACK with device matvecs (Hermitian operators)."""
        from scipy.sparse.linalg import eigsh
        op = LinearOperator(self._shape, matvec=self._matvec, dtype=self._dtype)
        return eigsh(op, **eigsh_args)

    def __repr__(self):
        return "<{0}x{1} quspin quantum_LinearOperator of type '{2}'>".format(*(self._shape + (self._dtype,)))
