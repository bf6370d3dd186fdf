"""``hamiltonian``: static + time-dependent sparse operator with the CSR held
in HBM (src/quspin/operators/hamiltonian_core.py:221-1958, the parts on the
H.v path: construction, ``dot``, ``tocsr``, format/dtype handling).

The static opstr terms are assembled on the device (``qsx_csr_builder_*``),
the dynamic terms likewise, one CSR per drive function
(``make_static``/``make_dynamic`` _make_hamiltonian.py:83-133).  ``dot`` runs
the CSR SpMV kernel with the reference's ``overwrite_out`` / ``a``
semantics (hamiltonian_core.py:615-788).
"""

import numpy as np
import scipy.sparse as sp

from .. import _lib as L
from ._devop import DeviceCSR, csr_combine, csr_transpose
from ._make_hamiltonian import _consolidate_static, _consolidate_dynamic, function

supported_dtypes = (np.float32, np.float64, np.complex64, np.complex128)


def ishamiltonian(obj):
    return isinstance(obj, hamiltonian)


def _default_basis(N, **kw):
    from ..basis import spin_basis_general
    return spin_basis_general(N, **kw)


def _is_opstr_entry(ele):
    return isinstance(ele, (list, tuple)) and len(ele) >= 2 and isinstance(ele[0], str)


class hamiltonian:
    """Sparse static + dynamic operator (see module docstring)."""

    def __init__(self, static_list, dynamic_list, N=None, basis=None, shape=None, dtype=np.complex128,
                 static_fmt=None, dynamic_fmt=None, copy=True, check_symm=True, check_herm=True, check_pcon=True,
                 **basis_kwargs):
        if np.dtype(dtype) not in [np.dtype(d) for d in supported_dtypes]:
            raise TypeError("hamiltonian does not support type: " + str(dtype))
        self._dtype = np.dtype(dtype)
        static_opstr = [e for e in static_list if _is_opstr_entry(e)]
        static_other = [e for e in static_list if not _is_opstr_entry(e)]
        dyn_opstr = [e for e in dynamic_list if _is_opstr_entry(e)]
        dyn_other = [e for e in dynamic_list if not _is_opstr_entry(e)]
        self._basis = None
        if static_opstr or dyn_opstr:
            if basis is None:
                if N is None:
                    raise ValueError("expecting integer for N or basis object for basis.")
                basis = _default_basis(N, **basis_kwargs)
            self._basis = basis
            Ns = basis.Ns
            self._shape = (Ns, Ns)
            if check_herm:
                basis.check_hermitian(static_opstr, dyn_opstr)
            if check_symm:
                basis.check_symm(static_opstr, dyn_opstr)
            if check_pcon:
                basis.check_pcon(static_opstr, dyn_opstr)
        else:
            if shape is None:
                mats = [e for e in static_other] + [e[0] for e in dyn_other]
                if not mats:
                    raise ValueError("missing argument shape")
                shape = mats[0].shape
            self._shape = tuple(shape)
        self._basis_for_build = basis
        # static part
        host_static = None
        for M in static_other:
            M = sp.csr_matrix(M, dtype=self._dtype)
            host_static = M if host_static is None else host_static + M
        self._static = self._assemble(_consolidate_static(static_opstr), host_static)
        # the term list of an operator-string-only static part (the species-factorised form reads it)
        self._static_list = _consolidate_static(static_opstr) if not static_other else None
        # dynamic parts: one CSR per (f, f_args)
        self._dynamic = {}
        groups = _consolidate_dynamic(dyn_opstr)
        for (f, f_args), ops in groups.items():
            func = function(f, f_args)
            self._dynamic[func] = self._assemble(ops, None)
        for item in dyn_other:
            M, f, f_args = item[0], item[1], tuple(item[2])
            func = function(f, f_args)
            Md = DeviceCSR.from_scipy(sp.csr_matrix(M, dtype=self._dtype), self._dtype)
            if func in self._dynamic:
                self._dynamic[func] = csr_combine([self._dynamic[func], Md], [1.0, 1.0], self._dtype)
            else:
                self._dynamic[func] = Md

    # ------------------------------------------------------------ assembly
    def _assemble(self, op_list, host_extra):
        Ns = self._shape[0]
        if op_list:
            from ._devop import build_csr
            ol = self._basis._compile_oplist(op_list)
            M = build_csr(self._basis, ol, self._dtype)
        else:
            M = DeviceCSR.from_scipy(sp.csr_matrix((Ns, Ns), dtype=self._dtype), self._dtype)
        if host_extra is not None:
            M = csr_combine([M, DeviceCSR.from_scipy(host_extra, self._dtype)], [1.0, 1.0], self._dtype)
        return M

    @classmethod
    def _from_parts(cls, static, dynamic, dtype, shape, basis):
        H = cls.__new__(cls)
        H._dtype = np.dtype(dtype)
        H._static = static
        H._dynamic = dict(dynamic)
        H._shape = tuple(shape)
        H._basis = basis
        H._basis_for_build = basis
        H._static_list = None
        return H

    # ------------------------------------------------------------ properties
    @property
    def basis(self):
        if self._basis is None:
            raise AttributeError("object has no attribute 'basis'")
        return self._basis

    @property
    def Ns(self):
        return self._shape[0]

    @property
    def shape(self):
        return self._shape

    @property
    def ndim(self):
        return 2

    @property
    def dtype(self):
        return self._dtype

    @property
    def static(self):
        """scipy.sparse.csr: static part (downloaded)."""
        return self._static.to_scipy()

    @property
    def dynamic(self):
        return {f: M.to_scipy() for f, M in self._dynamic.items()}

    @property
    def is_dense(self):
        return False

    @property
    def nbytes(self):
        n = self._static.data.numel() * self._static.data.element_size()
        return n

    # ------------------------------------------------------------ algebra
    # every part stays a device CSR: transposes and linear combinations run
    # on the device (csr_transpose, csr_combine)
    def _map(self, fpart, shape=None, conj_funcs=False):
        st = fpart(self._static)
        dy = {(f.conj() if conj_funcs else f): fpart(M) for f, M in self._dynamic.items()}
        return hamiltonian._from_parts(st, dy, st.dtype, shape or self._shape, self._basis)

    def transpose(self, copy=False):
        return self._map(csr_transpose, self._shape[::-1])

    def conjugate(self):
        """Conjugated matrices and drives (hamiltonian_core.py:2058-2079)."""
        return self._map(lambda M: csr_combine([M], [1.0], conj=[True]) if M.dtype.kind == "c" else M,
                         conj_funcs=True)

    conj = conjugate

    def getH(self, copy=False):
        return self._map(lambda M: csr_transpose(M, conj=True), self._shape[::-1], conj_funcs=True)

    @property
    def T(self):
        return self.transpose()

    @property
    def H(self):
        return self.getH()

    def __neg__(self):
        return self._map(lambda M: csr_combine([M], [-1.0]))

    def __mul__(self, other):
        if np.isscalar(other):
            dt = np.result_type(self._dtype, np.min_scalar_type(other))
            return self._map(lambda M: csr_combine([M], [other], dtype=np.result_type(M.dtype, dt)))
        return self.dot(other)

    def __rmul__(self, other):
        if np.isscalar(other):
            return self.__mul__(other)
        return NotImplemented

    def __truediv__(self, other):
        if np.isscalar(other):
            return self.__mul__(1.0 / other)
        return NotImplemented

    def __add__(self, other):
        if ishamiltonian(other):
            st = csr_combine([self._static, other._static], [1.0, 1.0])
            dy = {f: M for f, M in self._dynamic.items()}
            for f, M in other._dynamic.items():
                dy[f] = csr_combine([dy[f], M], [1.0, 1.0]) if f in dy else M
            return hamiltonian._from_parts(st, dy, st.dtype, self._shape, self._basis)
        if sp.issparse(other) or isinstance(other, np.ndarray):
            st = csr_combine([self._static, DeviceCSR.from_scipy(sp.csr_matrix(other))], [1.0, 1.0])
            return hamiltonian._from_parts(st, self._dynamic, st.dtype, self._shape, self._basis)
        return NotImplemented

    __radd__ = __add__

    def __sub__(self, other):
        return self.__add__(-other if ishamiltonian(other) else -other)

    def astype(self, dtype, copy=False):
        H = self._map(lambda M: csr_combine([M], [1.0], dtype=dtype))
        H._dtype = np.dtype(dtype)
        return H

    # ------------------------------------------------------------- formats
    def _at(self, time):
        """Device CSR of H(time) = H_static + sum_f f(time) H_f, combined on
        the device (hamiltonian_core.py:1686-1720)."""
        if not self._dynamic:
            return self._static
        parts = [self._static] + list(self._dynamic.values())
        coefs = [1.0] + [f(time) for f in self._dynamic]
        return csr_combine(parts, coefs)

    def tocsr(self, time=0):
        """Host CSR of H(time) (hamiltonian_core.py:1686-1720): combined on the
        device, copied once."""
        return sp.csr_matrix(self._at(time).to_scipy(), copy=True)

    def tocsc(self, time=0):
        return self.tocsr(time).tocsc()

    def toarray(self, time=0, order=None, out=None):
        return self.tocsr(time).toarray()

    def todense(self, time=0, order=None, out=None):
        return np.asmatrix(self.toarray(time))

    def aslinearoperator(self, time=0.0):
        from scipy.sparse.linalg import LinearOperator
        return LinearOperator(self._shape, matvec=lambda v: self.dot(v, time=time),
                              rmatvec=lambda v: self.H.dot(v, time=time), dtype=self._dtype)

    def _qsx_operator(self, time=0.0):
        """Device CSR of H(time) for the tools."""
        return self._at(time)

    # ------------------------------------------------------------------ dot
    def dot(self, V, time=0, check=True, out=None, overwrite_out=True, a=1.0):
        """a*H(time)|V> with the reference's out/overwrite_out semantics."""
        if ishamiltonian(V):
            return self * V if a == 1.0 else a * (self * V)
        on_dev = L.is_device_tensor(V)
        if not on_dev:
            if sp.issparse(V):
                if out is not None:
                    raise TypeError("'out' option does not apply for sparse inputs.")
                res = self._static.to_scipy() * V
                for f, M in self._dynamic.items():
                    res = res + f(time) * (M.to_scipy().dot(V))
                return a * res
            V = np.asanyarray(V)
        shape = tuple(V.shape)
        if check:
            if shape[0] != self._shape[1]:
                raise ValueError("matrix dimension mismatch with shapes: {0} and {1}.".format(shape, self._shape))
            if len(shape) > 3:
                raise ValueError("Expecting V.ndim < 4.")
        vdt = L.np_dtype(V.dtype) if on_dev else V.dtype
        result_dtype = np.result_type(vdt, self._dtype)
        if result_dtype not in [np.dtype(d) for d in supported_dtypes]:
            raise TypeError("resulting dtype is not supported.")
        times = np.array(time)
        if len(shape) == 3 and times.ndim == 0:
            times = np.broadcast_to(times, (shape[-1],))
        if times.ndim > 0:
            if times.ndim > 1:
                raise ValueError("Expecting time to be one dimensional array-like.")
            if shape[-1] != times.shape[0]:
                raise ValueError("For non-scalar times V.shape[-1] must be equal to len(time).")
            cols = [self.dot(V[..., i], time=t, check=False, a=a) for i, t in enumerate(times)]
            if on_dev:
                return L.torch().stack(cols, dim=-1)
            return np.stack(cols, axis=-1)
        x = L.to_device(V, result_dtype)
        if out is None:
            y = L.empty(shape, result_dtype)
            beta = 0
        else:
            odt = L.np_dtype(out.dtype) if L.is_device_tensor(out) else out.dtype
            if odt != result_dtype:
                raise TypeError("'out' must be array with correct dtype and dimensions for output array.")
            if tuple(out.shape) != shape:
                raise ValueError("'out' must be array with correct dtype and dimensions for output array.")
            y = out if L.is_device_tensor(out) else L.to_device(out, result_dtype)
            beta = 0 if overwrite_out else 1
        ac = complex(a) if np.iscomplexobj(a) else float(a)
        self._static.apply(x, y, a=ac, beta=beta)
        for f, M in self._dynamic.items():
            M.apply(x, y, a=ac * f(time), beta=1)
        if on_dev:
            return y
        res = L.to_host(y)
        if out is not None and not L.is_device_tensor(out):
            out[...] = res
            return out
        return res

    # ------------------------------------------------------------ evolution
    def evolve(self, v0, t0, times, eom="SE", solver_name="dop853", stack_state=False, verbose=False,
               iterate=False, imag_time=False, **solver_args):
        """Solve the Schroedinger (``eom="SE"``, real or imaginary time) or
        Liouville-von Neumann (``"LvNE"``) equation with scipy's ``ode``
        integrators (hamiltonian_core.py:1494-1685); every right-hand side
        evaluation applies H(t) on the device (``_DeviceEOM``)."""
        from ..tools.evolution import evolve
        if L.is_device_tensor(v0):
            v0 = L.to_host(v0)
        v0 = np.asanyarray(v0)
        if np.iscomplexobj(times):
            raise ValueError("times must be real number(s).")
        kw = dict(solver_args, solver_name=solver_name, stack_state=stack_state, verbose=verbose,
                  iterate=iterate, imag_time=imag_time)
        if eom == "SE":
            if v0.ndim > 2:
                raise ValueError("v0 must have ndim <= 2")
            if v0.shape[0] != self.Ns:
                raise ValueError("v0 must have {0} elements".format(self.Ns))
            if imag_time:
                if stack_state:
                    raise NotImplementedError("stack state is not compatible with imaginary time evolution.")
                dt = np.result_type(v0.dtype, self._dtype, np.float64)
                f = _DeviceEOM(self, "ISO", v0.shape, dt)
                kw["real"] = dt.kind != "c"
            elif stack_state:
                if self._dtype.kind == "c":
                    raise ValueError("stack_state option cannot be used with complex-valued Hamiltonians")
                f = _DeviceEOM(self, "SO_real", (2 * v0.shape[0],) + v0.shape[1:], np.float64)
                kw["real"] = False
            else:
                f = _DeviceEOM(self, "SO", v0.shape, np.complex128)
                kw["real"] = False
        elif eom == "LvNE":
            if v0.ndim != 2:
                raise ValueError("v0 must have ndim = 2")
            if v0.shape != self._shape:
                raise ValueError("v0 must be same shape as Hamiltonian")
            if imag_time:
                raise NotImplementedError("imaginary time not implemented for Liouville-von Neumann dynamics")
            if stack_state:
                raise NotImplementedError("stack_state not implemented for Liouville-von Neumann dynamics")
            f = _DeviceEOM(self, "LvNE", v0.shape, np.complex128)
            kw["real"] = False
        else:
            raise ValueError("'{} equation' not recognized, must be 'SE' or 'LvNE'".format(eom))
        return evolve(v0, t0, times, f, **kw)

    def rdot(self, V, time=0, check=True, out=None, overwrite_out=True, a=1.0):
        return self.transpose().dot(V.T if not L.is_device_tensor(V) else V.T, time=time, check=check, a=a).T

    # ----------------------------------------------------- measurements etc.
    def expt_value(self, V, time=0, check=True, enforce_pure=False):
        """<V|H(time)|V> for pure states (columns of V)."""
        on_dev = L.is_device_tensor(V)
        HV = self.dot(V, time=time)
        if on_dev:
            from ._devop import vdot_dev
            if V.dim() == 1:
                return vdot_dev(V.to(HV.dtype).contiguous(), HV.contiguous())
            V = L.to_host(V)
            HV = L.to_host(HV)
        V = np.asarray(V)
        if V.ndim == 1:
            return np.vdot(V, HV)
        return np.einsum("ij,ij->j", V.conj(), HV)

    def quant_fluct(self, V, time=0, check=True, enforce_pure=False):
        e = self.expt_value(V, time)
        HV = self.dot(V, time=time)
        HV = L.to_host(HV) if L.is_device_tensor(HV) else HV
        e2 = np.vdot(HV, HV) if np.ndim(HV) == 1 else np.einsum("ij,ij->j", HV.conj(), HV)
        return e2 - e ** 2

    def eigsh(self, time=0, **eigsh_args):
        """Host ARPACK on ``tocsr(time)`` (hamiltonian_core.py:1269-1310)."""
        from scipy.sparse.linalg import eigsh
        return eigsh(self.tocsr(time), **eigsh_args)

    def eigh(self, time=0, **eigh_args):
        return np.linalg.eigh(self.toarray(time))

    def eigvalsh(self, time=0, **eigvalsh_args):
        return np.linalg.eigvalsh(self.toarray(time))

    def trace(self, time=0):
        return self._at(time).trace()

    def __repr__(self):
        return "<{0}x{1} qspin dtype:{2}>".format(*(self._shape + (self._dtype,)))


class _DeviceEOM:
    """Right-hand sides of the equations of motion (hamiltonian_core.py:1390-1492)
    as scipy ``ode`` callables: the state arrives as a host array, H(t) acts
    on it on the device through the operator's CSR kernels (static part with
    overwrite, each drive term f(t)*H_f accumulated), and the derivative
    returns as a host array.

    kind "SO":      -i H(t) V                         (complex V)
    kind "ISO":     -H(t) V                           (imaginary time)
    kind "SO_real": [H(t) v; -H(t) u] for V = [u; v]  (real H, stacked state)
    kind "LvNE":    -i (H(t) rho - rho H(t)), rho H = (H^T rho^T)^T
    """

    def __init__(self, H, kind, shape, dtype):
        self.H = H
        self.kind = kind
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.x = L.empty(self.shape, self.dtype)
        self.y = L.empty(self.shape, self.dtype)
        if kind == "LvNE":
            self.HT = H.transpose()
            self.yt = L.empty(self.shape, self.dtype)

    @staticmethod
    def _apply(Hop, x, y, time, a, beta):
        Hop._static.apply(x, y, a=a, beta=beta)
        for f, M in Hop._dynamic.items():
            M.apply(x, y, a=a * f(time), beta=1)

    def __call__(self, time, V):
        torch = L.torch()
        self.x.copy_(torch.from_numpy(np.ascontiguousarray(np.asarray(V).reshape(self.shape), dtype=self.dtype)))
        x, y = self.x, self.y
        if self.kind == "SO":
            self._apply(self.H, x, y, time, -1j, 0)
        elif self.kind == "ISO":
            self._apply(self.H, x, y, time, -1.0, 0)
        elif self.kind == "SO_real":
            Ns = self.shape[0] // 2
            self._apply(self.H, x[Ns:], y[:Ns], time, 1.0, 0)
            self._apply(self.H, x[:Ns], y[Ns:], time, -1.0, 0)
        else:
            self._apply(self.H, x, y, time, 1.0, 0)
            xt = x.transpose(0, 1).contiguous()
            self._apply(self.HT, xt, self.yt, time, 1.0, 0)
            y.sub_(self.yt.transpose(0, 1))
            y.mul_(-1j)
        return L.to_host(y).ravel()
