"""``expm_multiply_parallel``: e^{aA} v with the Taylor loop on the device.

Mirrors ``ExpmMultiplyParallel`` (parallel_sparse_tools, rendered source
docs/modules/parallel_sparse_tools/expm_multiply_parallel_core/
expm_multiply_parallel_core.html:117-466):
  * dtype = result_type(A.dtype, min_scalar_type(a), float64) unless given;
    tol = eps(dtype)/2; mu = tr(A)/N; ||A - mu I||_1 (device kernels);
  * (m_star, s) from scipy's ``_fragment_3_1`` with n0=1, ell=2 and the lazy
    1-norm estimates of a(A - mu I) (EMP.html:377-466);
  * ``dot`` runs Al-Mohy & Higham Alg. 3.2 (scipy _expm_multiply.py:280-303)
    in place with a work array of 2*v.size: one fused SpMV + F update + inf-norm
    kernel per Taylor term, the truncation test evaluated on the device.
``A`` may be a scipy matrix, a ``hamiltonian`` (its static part), a device CSR
or a ``quantum_LinearOperator`` (matrix-free Taylor loop).
"""

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import LinearOperator, onenormest
from scipy.sparse.linalg._expm_multiply import _fragment_3_1

from .. import _lib as L
from ..operators._devop import DeviceCSR, MatrixFreeOp, SectorBatch, as_device_operator, csr_transpose


class LazyOperatorNormInfo:
    """1-norm information of a(A - mu I), computed lazily
    (EMP.html:399-466).  Products run on the device."""

    def __init__(self, op, A_1_norm, a, mu, dtype, ell=2, adjoint=None):
        self._op = op
        self._adj = adjoint
        self._a = a
        self._mu = mu
        self._dtype = np.dtype(dtype)
        self._A_1_norm = A_1_norm
        self._ell = ell
        self._d = {}

    def onenorm(self):
        return np.abs(self._a) * self._A_1_norm

    def _mv(self, op, v, a, mu):
        v = np.asarray(v)
        dt = np.result_type(self._dtype, v.dtype)
        x = L.to_device(np.ascontiguousarray(v), dt)
        y = L.empty(tuple(x.shape), dt)
        op.apply(x, y, a=1.0, beta=0)
        r = L.to_host(y)
        return a * (r - mu * v)

    @property
    def linear_operator(self):
        a, mu = self._a, self._mu
        adj = self._adj() if self._adj is not None else None
        return LinearOperator(
            self._op.shape, dtype=self._dtype,
            matvec=lambda v: self._mv(self._op, v, a, mu),
            rmatvec=lambda v: self._mv(adj, v, np.conj(a), np.conj(mu)),
            matmat=lambda V: self._mv(self._op, V, a, mu),
            rmatmat=lambda V: self._mv(adj, V, np.conj(a), np.conj(mu)))

    def d(self, p):
        if p not in self._d:
            est = onenormest(self.linear_operator ** p)
            self._d[p] = est ** (1.0 / p)
        return self._d[p]

    def alpha(self, p):
        return max(self.d(p), self.d(p + 1))


class expm_multiply_parallel:
    """Action of e^{aA} on vectors, computed on the device."""

    def __init__(self, A, a=1.0, dtype=None, copy=False):
        if np.array(a).ndim == 0:
            self._a = a
        else:
            raise ValueError("a must be scalar value.")
        if A.shape[0] != A.shape[1]:
            raise ValueError("A must be a square matrix.")
        op = as_device_operator(A)
        if op is None:
            op = DeviceCSR.from_scipy(sp.csr_matrix(A))
        self._op = op
        self._A = A
        A_dtype = np.dtype(op.dtype)
        a_dtype_min = np.min_scalar_type(self._a)
        if dtype is None:
            self._dtype = np.result_type(A_dtype, a_dtype_min, np.float64)
        else:
            min_dtype = np.result_type(A_dtype, a_dtype_min, np.float32)
            if not np.can_cast(min_dtype, dtype):
                raise ValueError("dtype not sufficient to represent a*A to at least float32 precision.")
            self._dtype = np.dtype(dtype)
        tol = np.finfo(self._dtype).eps / 2
        self._tol = np.array(tol, dtype=np.finfo(self._dtype).eps.dtype)
        # (an operator on an empty sector has no diagonal to average: mu = 0)
        mu = op.trace() / op.shape[0] if op.shape[0] else 0.0
        self._mu = np.array(mu, dtype=self._dtype)
        self._A_1_norm = op.one_norm(-complex(self._mu))
        self._partition_override = None
        self._calculate_partition()

    @property
    def a(self):
        return self._a

    @property
    def A(self):
        return self._A

    @property
    def partition(self):
        """(m_star, s) used by ``dot``."""
        return self._m_star, self._s

    def set_a(self, a, dtype=None):
        if np.array(a).ndim == 0:
            self._a = a
            a_dtype_min = np.min_scalar_type(self._a)
            if dtype is None:
                self._dtype = np.result_type(self._op.dtype, a_dtype_min, np.float64)
            else:
                min_dtype = np.result_type(self._op.dtype, a_dtype_min, np.float32)
                if not np.can_cast(min_dtype, dtype):
                    raise ValueError("dtype not sufficient to represent a*A to at least float32 precision.")
                self._dtype = np.dtype(dtype)
            tol = np.finfo(self._dtype).eps / 2
            self._tol = np.array(tol, dtype=np.finfo(self._dtype).eps.dtype)
            self._mu = np.array(self._mu, dtype=self._dtype)
            self._calculate_partition()
        else:
            raise ValueError("expecting 'a' to be scalar.")

    def set_partition(self, m_star, s):
        """Use (m_star, s) chosen elsewhere (e.g. by the learned predictor)."""
        self._m_star, self._s = int(m_star), int(s)

    def _adjoint(self):
        op = self._op
        if isinstance(op, MatrixFreeOp):
            return MatrixFreeOp(op.basis, op.oplist, op.dtype, transposed=not op.transposed,
                                conjugated=not op.conjugated, diag_conj=not op.diag_conj, diag=op.diag)
        if not hasattr(self, "_adj_csr"):
            self._adj_csr = csr_transpose(op, conj=True)
        return self._adj_csr

    def _calculate_partition(self):
        if np.abs(self._a) * self._A_1_norm == 0:
            self._m_star, self._s = 0, 1
        else:
            ell = 2
            info = LazyOperatorNormInfo(self._op, self._A_1_norm, self._a, complex(self._mu), self._dtype, ell=ell,
                                        adjoint=self._adjoint)
            self._m_star, self._s = _fragment_3_1(info, 1, self._tol, ell=ell)

    def dot(self, v, work_array=None, overwrite_v=False, tol=None):
        """e^{aA} v (EMP.html:267-374)."""
        on_dev = L.is_device_tensor(v)
        if not on_dev:
            v = np.asarray(v)
        if v.ndim > 2:
            raise ValueError("array must have ndim <= 2.")
        if v.shape[0] != self._op.shape[1]:
            raise ValueError("dimension mismatch {}, {}".format(self._op.shape, v.shape))
        vdt = L.np_dtype(v.dtype) if on_dev else v.dtype
        v_dtype = np.result_type(self._dtype, vdt)
        if overwrite_v:
            if v_dtype != vdt:
                raise ValueError("if overwrite_v is True, the input array must match correct output dtype for "
                                 "matrix multiplication.")
            if not on_dev and not v.flags["CARRAY"]:
                raise TypeError("input array must a contiguous and writable.")
        if on_dev:
            F = v if (overwrite_v and v.is_contiguous()) else v.to(L.torch_dtype(v_dtype)).clone()
        else:
            F = L.to_device(np.array(v, dtype=v_dtype, order="C", copy=True))
        if work_array is None:
            work = L.empty((2 * v.shape[0],) + tuple(v.shape[1:]), v_dtype)
        else:
            wsize = work_array.numel() if L.is_device_tensor(work_array) else np.asarray(work_array).size
            if wsize != 2 * int(np.prod(v.shape)):
                raise ValueError("work_array must have twice the number of elements as in v.")
            wdt = L.np_dtype(work_array.dtype) if L.is_device_tensor(work_array) else work_array.dtype
            if wdt != v_dtype:
                raise ValueError("work_array must be array of dtype which matches the result of the matrix-vector "
                                 "multiplication.")
            work = work_array if L.is_device_tensor(work_array) else L.empty(2 * int(np.prod(v.shape)), v_dtype)
        tol_use = float(self._tol) if tol is None else float(tol)
        a = np.array(self._a, dtype=v_dtype)
        mu = np.array(self._mu, dtype=v_dtype)
        self._op.expm_inplace(F, work, self._s, self._m_star, tol_use, complex(mu), complex(a))
        if on_dev:
            return F
        res = L.to_host(F)
        if overwrite_v:
            v[...] = res
            return v
        return res


ExpmMultiplyParallel = expm_multiply_parallel


class expm_multiply_sectors:
    """e^{a A_k} V_k for many symmetry sectors k and ``nvec`` vectors each,
    all sectors in one launch per Taylor term (optionally one hipGraph replay
    for the whole loop).

    No counterpart of its own in the reference: sector k gives what
    ``expm_multiply_parallel(A_k, a, dtype).dot(V_k)`` gives (EMP.html:117-374),
    with the same per-sector mu, tolerance and (m*, s).  ``ops`` are
    ``quantum_LinearOperator`` (or other matrix-free) operators of equal
    dtype and site count.  ``form``: "mf" (matrix elements on the fly), "csr"
    (per-sector CSR built once on the device) or None (autotuned)."""

    def __init__(self, ops, a=1.0, dtype=None, nvec=1, chunk_rows=None, graph=True, form=None):
        self._single = [expm_multiply_parallel(op, a=a, dtype=dtype) for op in ops]
        self._dtype = self._single[0]._dtype
        for e in self._single:
            if e._dtype != self._dtype:
                raise ValueError("all sectors must share one dtype")
        mfs = [e._op for e in self._single]
        self._batch = SectorBatch(mfs, self._dtype, nvec, form=form, chunk_rows=chunk_rows)
        self._a = a
        self._graph = bool(graph)
        self._work = None

    @property
    def batch(self):
        return self._batch

    @property
    def partitions(self):
        """Per-sector (m_star, s)."""
        return [e.partition for e in self._single]

    def set_partitions(self, parts):
        for e, (m, s) in zip(self._single, parts):
            e.set_partition(m, s)

    def dot(self, V, overwrite_v=False):
        """``V``: stacked (rows, nvec) device array, or a list of per-sector
        arrays.  Returns the stacked result (in place when overwrite_v)."""
        B = self._batch
        if isinstance(V, (list, tuple)):
            F = B.stack(V)
        else:
            if L.np_dtype(V.dtype) != B.vdt:
                raise ValueError("stacked input must have dtype {}".format(B.vdt))
            F = V if overwrite_v else V.clone()
            F = F.reshape(B.rows, B.nvec)
        if self._work is None:
            self._work = L.empty(2 * B.rows * B.nvec, B.vdt)
        parts = [e.partition for e in self._single]
        tols = [float(e._tol) for e in self._single]
        mus = [complex(e._mu) for e in self._single]
        B.expm_inplace(F, self._work, parts, tols, mus, complex(self._a), graph=self._graph)
        return F
