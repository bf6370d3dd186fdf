"""Lanczos tools (src/quspin/tools/lanczos/_lanczos_utils.py:77-425,
_expm.py:7-52) with the recurrence on the device.

For operators the library can run (``hamiltonian``, ``quantum_LinearOperator``,
scipy matrices, device CSR) each step is one H.v launch plus two fused
update kernels (r -= b v, a = Re<q,r>; r -= a q, b = ||r||) with FP64
accumulation; a and b stay in device memory until the loop ends, and the
tridiagonal problem is solved in FP64 with ``scipy.linalg.eigh_tridiagonal``
as in the reference.  Other operators go through their own ``dot``.
"""

import numpy as np
from scipy.linalg import eigh_tridiagonal

from .. import _lib as L
from .._lib import ptr, stream_handle
from ..operators._devop import as_device_operator, nrm2_dev
from ._ftlm import FTLM_static_iteration, LTLM_static_iteration  # noqa: F401

__all__ = ["lanczos_full", "lanczos_iter", "lin_comb_Q_T", "expm_lanczos", "FTLM_static_iteration",
           "LTLM_static_iteration"]


def _dtype_of(A, v0):
    adt = np.dtype(getattr(A, "dtype", np.float64))
    vdt = L.np_dtype(v0.dtype) if L.is_device_tensor(v0) else np.asarray(v0).dtype
    return np.result_type(adt, vdt)


def _scale_div(x, s, dtype):
    """x /= s in place on the device."""
    sd = L.to_device(np.array([s, 0.0], dtype=np.float64))
    rc = L.load().qsx_lanczos_div(L.dtype_code(dtype), ptr(x), ptr(sd), ptr(x), int(x.numel()), stream_handle())
    L.check(rc, "qsx_lanczos_div")


def _device_loop(op, v0, m, dtype, eps, store_Q, Q_out=None):
    """Run the recurrence; returns (a, b, m_eff, Q or None)."""
    lib = L.load()
    n = op.shape[0]
    code = L.dtype_code(dtype)
    st = stream_handle()
    q0 = L.to_device(v0, dtype).reshape(n).clone()
    qn = nrm2_dev(q0)
    if abs(qn - 1.0) > eps:
        _scale_div(q0, qn, dtype)
    if store_Q:
        Q = Q_out if Q_out is not None else L.empty((m, n), dtype)
        Q[0].copy_(q0)
        qbuf = [Q[i] for i in range(m)]
    else:
        Q = None
        ring = [q0, L.empty(n, dtype), L.empty(n, dtype)]
        qbuf = None
    r = L.empty(n, dtype)
    ab = L.zeros((m, 2, 2), np.float64)  # [i, 0] = a_i, [i, 1] = b_i (2 doubles each)
    kind = _run_kind(op, dtype)
    if kind is not None and m >= 8:
        # the whole recurrence as one device launch sequence, captured once per
        # operator and shape and replayed (the workspace keeps its addresses)
        ws = _workspace(op, dtype, n, m, store_Q)
        Qbuf, r, ab = ws["Q"], ws["r"], ws["ab"]
        Qbuf[0].copy_(q0)
        ctx, csr = kind
        if csr is None:
            args = (0, ctx.h, None, None, None, 0, 0, 0, 0)
        else:
            d = csr._data_for(dtype)
            args = (1, None, ptr(csr.indptr), ptr(csr.indices), ptr(d), csr.idx64,
                    L.dtype_code(L.np_dtype(d.dtype)), csr.nnz, 0)
        rc = lib.qsx_lanczos_run(*args, code, n, m, ptr(Qbuf), int(bool(store_Q)), ptr(r), ptr(ab), 1, st)
        L.check(rc, "qsx_lanczos_run")
        if store_Q:
            Q.copy_(Qbuf)
        return _finish(ab, m, eps, Q)

    def qi(i):
        return qbuf[i] if store_Q else ring[i % 3]

    op.apply(qi(0), r, a=1.0, beta=0)
    rc = lib.qsx_lanczos_update(code, None, None, ptr(qi(0)), ptr(r), n, ptr(ab[0, 0]), ptr(ab[0, 1]), st)
    L.check(rc, "qsx_lanczos_update")
    for i in range(1, m):
        rc = lib.qsx_lanczos_div(code, ptr(r), ptr(ab[i - 1, 1]), ptr(qi(i)), n, st)
        L.check(rc, "qsx_lanczos_div")
        op.apply(qi(i), r, a=1.0, beta=0)
        rc = lib.qsx_lanczos_update(code, ptr(qi(i - 1)), ptr(ab[i - 1, 1]), ptr(qi(i)), ptr(r), n,
                                    ptr(ab[i, 0]), ptr(ab[i, 1]), st)
        L.check(rc, "qsx_lanczos_update")
    return _finish(ab, m, eps, Q)


def _finish(ab, m, eps, Q):
    """(a, b, effective m, Q) from the device alpha/beta array: the
    recurrence stops at the first b_i < eps (_lanczos_utils.py:191-201)."""
    h = L.to_host(ab)
    a = h[:, 0, 0].copy()
    b = h[:, 1, 0].copy()
    m_eff = m
    for i in range(1, m):
        if b[i] < eps:
            m_eff = i
            break
    return a, b, m_eff, Q


def _workspace(op, dtype, n, m, store_Q):
    """Per-operator device buffers of the captured recurrence."""
    cache = op.__dict__.setdefault("_lanczos_ws", {})
    key = (np.dtype(dtype).str, n, m, bool(store_Q))
    if key not in cache:
        if len(cache) >= 4:
            cache.clear()
        cache[key] = {"Q": L.empty((m if store_Q else 3, n), dtype), "r": L.empty(n, dtype),
                      "ab": L.zeros((m, 2, 2), np.float64)}
    return cache[key]


def _run_kind(op, dtype):
    """(H.v context, None) or (None, csr) for the one-call device recurrence,
    or None."""
    from ..operators._devop import DeviceCSR, MatrixFreeOp
    if isinstance(op, MatrixFreeOp):
        return op._context(np.dtype(dtype), 1), None
    if isinstance(op, DeviceCSR):
        return None, op
    return None


def _host_loop(A, v0, m, dtype, eps):
    """Recurrence through A.dot for operators without a device form."""
    v0 = np.asarray(v0)
    n = v0.size
    Q = np.zeros((m, n), dtype=dtype)
    Q[0, :] = v0
    a = np.zeros(m)
    b = np.zeros(m)
    qn = np.linalg.norm(Q[0])
    if abs(qn - 1.0) > eps:
        Q[0] /= qn
    r = np.asarray(A.dot(Q[0]), dtype=dtype)
    a[0] = np.vdot(Q[0], r).real
    r = r - a[0] * Q[0]
    b[0] = np.linalg.norm(r)
    m_eff = m
    for i in range(1, m):
        v = Q[i - 1].copy()
        Q[i] = r / b[i - 1]
        r = np.asarray(A.dot(Q[i]), dtype=dtype)
        r = r - b[i - 1] * v
        a[i] = np.vdot(Q[i], r).real
        r = r - a[i] * Q[i]
        b[i] = np.linalg.norm(r)
        if b[i] < eps:
            m_eff = i
            break
    return a, b, m_eff, Q


def lanczos_full(A, v0, m, full_ortho=False, out=None, eps=None):
    """Lanczos basis Q (m x n) and eigenpairs of the Krylov tridiagonal
    (_lanczos_utils.py:77-227)."""
    on_dev = L.is_device_tensor(v0)
    n = int(v0.numel()) if on_dev else np.asanyarray(v0).size
    ndim = v0.dim() if on_dev else np.asanyarray(v0).ndim
    dtype = _dtype_of(A, v0)
    if ndim != 1:
        raise ValueError("expecting array with ndim=1 for initial Lanczos vector.")
    if m >= n:
        raise ValueError("Requested size of Lanczos basis must be smaller then size of original space (e.g. m < n).")
    if out is not None:
        oshape = tuple(out.shape)
        odt = L.np_dtype(out.dtype) if L.is_device_tensor(out) else out.dtype
        if oshape != (m, n):
            raise ValueError("argument 'out' must have shape (m,n), see documentation.")
        if odt != dtype:
            raise ValueError("argument 'out' has dtype {}, expecting dtype {}".format(odt, dtype))
        if not L.is_device_tensor(out) and not out.flags["CARRAY"]:
            raise ValueError("argument 'out' must be C-contiguous and writable.")
    if eps is None:
        eps = np.finfo(dtype).eps
    op = as_device_operator(A)
    if op is not None:
        Qd = out if L.is_device_tensor(out) else None
        a, b, mm, Q = _device_loop(op, v0, m, dtype, eps, True, Qd)
    else:
        a, b, mm, Q = _host_loop(A, v0, m, dtype, eps)
    if full_ortho:
        Qh = L.to_host(Q[:mm]) if L.is_device_tensor(Q) else Q[:mm]
        q, _ = np.linalg.qr(Qh.T)
        Qh = np.ascontiguousarray(q.T)
        h = np.zeros((mm, mm), dtype=a.dtype)
        for i in range(mm):
            r = A.dot(Qh[i])
            r = L.to_host(r) if L.is_device_tensor(r) else np.asarray(r)
            h[i, i:] = np.dot(Qh[i:mm], np.conj(r)).real
        E, V = np.linalg.eigh(h, UPLO="U")
        if L.is_device_tensor(Q):
            Q[:mm].copy_(L.to_device(Qh, dtype))
        else:
            Q[:mm] = Qh
    else:
        E, V = eigh_tridiagonal(a[:mm], b[:mm - 1])
    if on_dev:
        return E, V, Q[:mm]
    Qn = L.to_host(Q) if L.is_device_tensor(Q) else Q
    if out is not None and not L.is_device_tensor(out):
        out[...] = Qn
        return E, V, out[:mm]
    return E, V, Qn[:mm]


class _lanczos_vec_iter:
    """Regenerates the Lanczos vectors from (a, b) on demand
    (_lanczos_utils.py:15-74)."""

    def __init__(self, A, v0, a, b, dtype):
        self._A, self._v0, self._a, self._b, self._dtype = A, v0, a, b, dtype

    def __iter__(self):
        A, a, b, dtype = self._A, self._a, self._b, self._dtype
        on_dev = L.is_device_tensor(self._v0)
        op = as_device_operator(A)
        if op is None:
            q = np.array(self._v0, dtype=dtype, copy=True)
            qn = np.linalg.norm(q)
            if abs(qn - 1.0) > np.finfo(dtype).eps:
                q /= qn
            r = np.asarray(A.dot(q), dtype=dtype) - a[0] * q
            yield q.copy()
            for i in range(1, a.size):
                v = q.copy()
                q = r / b[i - 1]
                r = np.asarray(A.dot(q), dtype=dtype) - b[i - 1] * v - a[i] * q
                yield q.copy()
            return
        n = op.shape[0]
        q = L.to_device(self._v0, dtype).reshape(n).clone()
        qn = nrm2_dev(q)
        if abs(qn - 1.0) > np.finfo(dtype).eps:
            _scale_div(q, qn, dtype)
        lib = L.load()
        code = L.dtype_code(dtype)
        st = stream_handle()
        r = L.empty(n, dtype)
        op.apply(q, r, a=1.0, beta=0)
        lib.qsx_axpy(code, ptr(q), ptr(r), n, -float(a[0]), 0.0, None, 1.0, 0, st)
        yield q.clone() if on_dev else L.to_host(q)
        for i in range(1, a.size):
            v = q.clone()
            q = L.empty(n, dtype)
            bi = L.to_device(np.array([b[i - 1], 0.0]))
            lib.qsx_lanczos_div(code, ptr(r), ptr(bi), ptr(q), n, st)
            op.apply(q, r, a=1.0, beta=0)
            lib.qsx_axpy(code, ptr(v), ptr(r), n, -float(b[i - 1]), 0.0, None, 1.0, 0, st)
            lib.qsx_axpy(code, ptr(q), ptr(r), n, -float(a[i]), 0.0, None, 1.0, 0, st)
            yield q.clone() if on_dev else L.to_host(q)


def lanczos_iter(A, v0, m, return_vec_iter=True, copy_v0=True, copy_A=False, eps=None):
    """Memory-light Lanczos: (E, V[, vector iterator]) (_lanczos_utils.py:230-351)."""
    on_dev = L.is_device_tensor(v0)
    n = int(v0.numel()) if on_dev else np.asanyarray(v0).size
    ndim = v0.dim() if on_dev else np.asanyarray(v0).ndim
    dtype = _dtype_of(A, v0)
    if ndim != 1:
        raise ValueError("expecting array with ndim=1 for initial Lanczos vector.")
    if m >= n:
        raise ValueError("Requested size of Lanczos basis must be smaller then size of original space (e.g. m < n).")
    if eps is None:
        eps = np.finfo(dtype).eps
    if copy_v0 and return_vec_iter:
        v0 = v0.clone() if on_dev else np.array(v0, copy=True)
    op = as_device_operator(A)
    if op is not None:
        a, b, mm, _ = _device_loop(op, v0, m, dtype, eps, False)
    else:
        a, b, mm, _ = _host_loop(A, v0, m, dtype, eps)
    E, V = eigh_tridiagonal(a[:mm], b[:mm - 1])
    if return_vec_iter:
        return E, V, _lanczos_vec_iter(A, v0, a[:mm], b[:mm - 1], dtype)
    return E, V


def lin_comb_Q_T(coeff, Q_T, out=None):
    """v = sum_i coeff_i Q_T[i] (_lanczos_utils.py:366-425)."""
    coeff = np.asanyarray(coeff)
    if L.is_device_tensor(Q_T):
        m, n = Q_T.shape
        dtype = np.result_type(L.np_dtype(Q_T.dtype), coeff.dtype)
        Q = Q_T if L.np_dtype(Q_T.dtype) == dtype else Q_T.to(L.torch_dtype(dtype))
        c = L.to_device(np.ascontiguousarray(coeff[:m], dtype=np.complex128))
        res = out if (out is not None and L.is_device_tensor(out)) else L.empty(n, dtype)
        rc = L.load().qsx_lincomb(L.dtype_code(dtype), ptr(Q.contiguous()), m, n, ptr(c), ptr(res), stream_handle())
        L.check(rc, "qsx_lincomb")
        return res
    if isinstance(Q_T, np.ndarray):
        Q_iter = iter(Q_T[:])
    else:
        Q_iter = iter(Q_T)
    q = next(Q_iter)
    on_dev = L.is_device_tensor(q)
    if on_dev:
        q = L.to_host(q)
    dtype = np.result_type(q.dtype, coeff.dtype)
    if out is not None:
        if out.shape != q.shape:
            raise ValueError("'out' must have same shape as a Lanczos vector.")
        if out.dtype != dtype:
            raise ValueError("argument 'out' has dtype {}, expecting dtype {}".format(out.dtype, dtype))
        if not out.flags["CARRAY"]:
            raise ValueError("argument 'out' must be C-contiguous and writable.")
    else:
        out = np.zeros(q.shape, dtype=dtype)
    np.multiply(q, coeff[0], out=out)
    for weight, q in zip(coeff[1:], Q_iter):
        if L.is_device_tensor(q):
            q = L.to_host(q)
        out += weight * q
    return out


def expm_lanczos(E, V, Q_T, a=1.0, out=None):
    """e^{aA} v0 from Lanczos data: V (e^{aE} * V[0, :]) combined with the
    Lanczos vectors (_expm.py:7-52)."""
    c = V.dot(np.exp(a * E) * V[0, :])
    return lin_comb_Q_T(c, Q_T, out=out)
