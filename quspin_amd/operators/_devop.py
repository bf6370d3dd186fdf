"""Device operators used by ``hamiltonian``, ``quantum_LinearOperator`` and
the tools: a CSR matrix in HBM and the matrix-free fused H.v, both exposing

    apply(x, y, a=1, beta=0)   y <- beta*y + a*A x   (device tensors)
    trace(), one_norm(shift), expm_inplace(...)

All arithmetic runs in ``libquspin_hip.so``; torch only holds the buffers.
"""

import ctypes as C

import numpy as np
import scipy.sparse as sp

from .. import _lib as L
from .. import tuning
from .._lib import ptr, stream_handle


def _vec_dtype_code(t):
    return L.dtype_code(L.np_dtype(t.dtype))


class DeviceCSR:
    """CSR matrix resident in HBM (int32 or int64 indices)."""

    is_matrix_free = False

    def __init__(self, indptr, indices, data, shape, dtype):
        self.nnz = int(indices.shape[0])
        if self.nnz == 0:
            # the kernels take non-null entry pointers: a matrix without
            # entries keeps one element that no row's indptr range reaches
            t = L.torch()
            indices = t.zeros(1, dtype=indices.dtype, device=indices.device)
            data = t.zeros(1, dtype=data.dtype, device=data.device)
        self.indptr, self.indices, self.data = indptr, indices, data
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.idx64 = 1 if indptr.dtype == L.torch().int64 else 0
        self._host = None

    @property
    def Ns(self):
        return self.shape[0]

    @classmethod
    def from_scipy(cls, A, dtype=None):
        A = sp.csr_matrix(A)
        if dtype is not None:
            A = A.astype(dtype)
        A.sum_duplicates()
        dt = A.dtype
        if dt not in (np.float32, np.float64, np.complex64, np.complex128):
            dt = np.result_type(dt, np.float64)
            A = A.astype(dt)
        idx = np.int64 if (A.nnz >= 2 ** 31 - 1 or A.shape[0] >= 2 ** 31 - 1) else np.int32
        ip = L.to_device(A.indptr.astype(idx))
        ix = L.to_device(A.indices.astype(idx)) if A.nnz else L.empty(1, idx)[:0]
        da = L.to_device(A.data) if A.nnz else L.empty(1, dt)[:0]
        M = cls(ip, ix, da, A.shape, dt)
        return M

    def to_scipy(self):
        if self._host is None:
            self._host = sp.csr_matrix((L.to_host(self.data)[:self.nnz], L.to_host(self.indices)[:self.nnz],
                                        L.to_host(self.indptr)), shape=self.shape)
        return self._host

    def apply(self, x, y, a=1.0, beta=0):
        """y = a*A*x (+ y when beta)."""
        vdt = _vec_dtype_code(x)
        mdt = L.dtype_code(self.dtype)
        nvec = int(np.prod(x.shape[1:])) if x.dim() > 1 else 1
        ac = complex(a)
        data = self._data_for(L.np_dtype(x.dtype))
        mdt_use = L.dtype_code(L.np_dtype(data.dtype))
        rc = L.load().qsx_csr_matvec(ptr(self.indptr), ptr(self.indices), ptr(data), self.idx64, mdt_use,
                                     self.shape[0], self.nnz, ptr(x), ptr(y), vdt, nvec, ac.real, ac.imag,
                                     int(bool(beta)), 0, stream_handle())
        L.check(rc, "qsx_csr_matvec")

    def _data_for(self, vdt):
        """Values in the matrix dtype the kernel pairs with vector dtype
        ``vdt``: the vector dtype itself, or its real part for a real matrix."""
        vdt = np.dtype(vdt)
        if self.dtype.kind == "c":
            tgt = vdt if vdt.kind == "c" else np.dtype(np.complex128 if vdt == np.float64 else np.complex64)
        else:
            tgt = np.dtype(np.float64) if vdt in (np.float64, np.complex128) else np.dtype(np.float32)
        if tgt == self.dtype:
            return self.data
        if not hasattr(self, "_casts"):
            self._casts = {}
        if tgt not in self._casts:
            self._casts[tgt] = self.data.to(L.torch_dtype(tgt))
        return self._casts[tgt]

    def trace(self):
        out = L.zeros(2, np.float64)
        rc = L.load().qsx_csr_trace(ptr(self.indptr), ptr(self.indices), ptr(self.data), self.idx64,
                                    L.dtype_code(self.dtype), self.shape[0], ptr(out), stream_handle())
        L.check(rc, "qsx_csr_trace")
        o = L.to_host(out)
        return complex(o[0], o[1]) if self.dtype.kind == "c" else float(o[0])

    def one_norm(self, shift=0.0):
        colsum = L.empty(max(self.shape[0], 1), np.float64)
        out = L.zeros(2, np.float64)
        sc = complex(shift)
        rc = L.load().qsx_csr_1_norm(ptr(self.indptr), ptr(self.indices), ptr(self.data), self.idx64,
                                     L.dtype_code(self.dtype), self.shape[0], sc.real, sc.imag, ptr(colsum),
                                     ptr(out), stream_handle())
        L.check(rc, "qsx_csr_1_norm")
        return float(L.to_host(out)[0])

    def expm_inplace(self, v, work, s, m_star, tol, mu, a):
        """Taylor loop of ``_wrapper_expm_multiply[_batch]`` on device."""
        vdt = _vec_dtype_code(v)
        nvec = int(np.prod(v.shape[1:])) if v.dim() > 1 else 1
        data = self._data_for(L.np_dtype(v.dtype))
        mdt = L.dtype_code(L.np_dtype(data.dtype))
        mu, a = complex(mu), complex(a)
        rc = L.load().qsx_expm_multiply_csr(ptr(self.indptr), ptr(self.indices), ptr(data), self.idx64, mdt,
                                            self.shape[0], self.nnz, int(s), int(m_star), float(tol), mu.real,
                                            mu.imag, a.real, a.imag, ptr(v), ptr(work), vdt, nvec, 0,
                                            stream_handle())
        L.check(rc, "qsx_expm_multiply_csr")


class MatrixFreeOp:
    """Fused matrix-free H.v over a device basis and operator list.

    ``transposed``/``conjugated`` act on the off-diagonal terms, ``diag_conj``
    on the diagonal terms; ``diag`` is an optional extra device diagonal."""

    is_matrix_free = True

    def __init__(self, basis, oplist, dtype, transposed=False, conjugated=False, diag_conj=False, diag=None):
        self.basis = basis
        self.oplist = oplist
        self.dtype = np.dtype(dtype)
        self.transposed, self.conjugated, self.diag_conj = bool(transposed), bool(conjugated), bool(diag_conj)
        self.diag = diag
        self._ctx = {}
        self.shape = (basis.Ns, basis.Ns)

    @property
    def Ns(self):
        return self.basis.Ns

    def _raw_ctx(self, vdt, nvec):
        h = L._vp()
        dcplx = 0
        if self.diag is not None:
            dcplx = 1 if L.np_dtype(self.diag.dtype).kind == "c" else 0
        rc = L.load().qsx_hv_ctx_create(self.basis._core.h, self.oplist.h, *self.basis._dev_args(),
                                        L.dtype_code(vdt), nvec, int(self.transposed), int(self.conjugated),
                                        int(self.diag_conj), ptr(self.diag), dcplx, C.byref(h))
        L.check(rc, "qsx_hv_ctx_create")
        return _Ctx(h)

    def _folded_diagonal(self):
        """The operator's whole diagonal (diagonal terms and extra diagonal)
        as one device vector, computed once: float64 when real, else
        complex128.  None when the operator has no diagonal part."""
        if not hasattr(self, "_dfold"):
            self._dfold = None
            if self.oplist.n_diag > 0 or self.diag is not None:
                ctx = self._raw_ctx(np.dtype(np.complex128), 1)
                t = L.torch()
                ones = t.ones(self.Ns, dtype=t.complex128, device="cuda")
                d = L.zeros(self.Ns, np.complex128)
                L.check(L.load().qsx_hv_ctx_apply(ctx.h, ptr(ones), ptr(d), 0.0, 0.0, 1.0, 0.0, 0,
                                                  stream_handle()), "qsx_hv_ctx_apply")
                del ctx
                if bool((d.imag != 0).any().item()):
                    self._dfold = d
                else:
                    self._dfold = d.real.contiguous()
        return self._dfold

    def _new_ctx(self, vdt, nvec):
        """H.v context whose diagonal is read from the folded per-row array
        (one 8- or 16-byte read per row instead of evaluating every diagonal
        term per product), as quantum_LinearOperator keeps ``_diagonal``."""
        ctx = self._raw_ctx(vdt, nvec)
        d = self._folded_diagonal()
        if d is not None:
            cplx = 1 if L.np_dtype(d.dtype).kind == "c" else 0
            L.check(L.load().qsx_hv_ctx_set_diag(ctx.h, ptr(d), cplx), "qsx_hv_ctx_set_diag")
        return ctx

    def _context(self, vdt, nvec):
        key = (vdt, nvec)
        if key not in self._ctx:
            ctx = self._new_ctx(vdt, nvec)
            cfg = tuning.choose(self, ctx, vdt, nvec)
            ctx.configure(cfg, self.basis)
            self._ctx[key] = ctx
        return self._ctx[key]

    @property
    def config(self):
        """Chosen (lookup, schedule) per (dtype, nvec)."""
        return {k: v.cfg for k, v in self._ctx.items()}

    def apply(self, x, y, a=1.0, beta=0, a_diag=None):
        vdt = L.np_dtype(x.dtype)
        nvec = int(np.prod(x.shape[1:])) if x.dim() > 1 else 1
        ctx = self._context(vdt, nvec)
        ac = complex(a)
        ad = ac if a_diag is None else complex(a_diag)
        rc = L.load().qsx_hv_ctx_apply(ctx.h, ptr(x), ptr(y), ac.real, ac.imag, ad.real, ad.imag,
                                       int(bool(beta)), stream_handle())
        L.check(rc, "qsx_hv_ctx_apply")

    def diagonal(self, dtype=np.complex128):
        """Diagonal of the operator as a device vector."""
        ones = L.torch().ones(self.Ns, dtype=L.torch_dtype(dtype), device="cuda")
        out = L.zeros(self.Ns, dtype)
        self.apply(ones, out, a=0.0, a_diag=1.0)
        return out

    def _column_norms(self, shift):
        """(column 1-norms of A + shift I, diagonal of A as (re, im) pairs),
        device arrays from one pass over the operator's columns
        (``qsx_hv_ctx_column_norms``), or None when the operator needs its
        CSR for them (a column with more than 64 distinct rows, or a
        transposed list that does not commute with the group)."""
        ctx = self._raw_ctx(np.dtype(np.complex128), 1)
        colabs = L.empty(self.Ns, np.float64)
        dg = L.empty((self.Ns, 2), np.float64)
        sc = complex(shift)
        rc = L.load().qsx_hv_ctx_column_norms(ctx.h, sc.real, sc.imag, ptr(colabs), ptr(dg), stream_handle())
        if rc == L.QSX_ERR_UNSUPPORTED:
            return None
        L.check(rc, "qsx_hv_ctx_column_norms")
        return colabs, dg

    def trace(self):
        """tr(A) over the whole reduced matrix: the diagonal terms plus the
        off-diagonal terms that map a representative onto itself (what
        _wrapper_csr_trace sums on the reference's CSR, EMP.html:196)."""
        if self.Ns == 0:
            return 0.0
        r = self._column_norms(0.0)
        if r is None:
            return self.to_csr(self.dtype if self.dtype.kind == "c" else np.complex128).trace()
        out = L.zeros(2, np.float64)
        L.check(L.load().qsx_reduce(0, ptr(r[1]), self.Ns, ptr(out), stream_handle()), "qsx_reduce")
        o = L.to_host(out)
        return complex(o[0], o[1]) if (self.dtype.kind == "c" or o[1] != 0) else float(o[0])

    def one_norm(self, shift=0.0):
        """max_c sum_r |A_rc + shift delta_rc| (_wrapper_csr_1_norm,
        EMP.html:200-202) from one pass over the operator's columns; the CSR
        is assembled only when that pass cannot hold a column."""
        if self.Ns == 0:
            return 0.0
        r = self._column_norms(shift)
        if r is None:
            return self.to_csr().one_norm(shift)
        out = L.zeros(2, np.float64)
        L.check(L.load().qsx_reduce(1, ptr(r[0]), self.Ns, ptr(out), stream_handle()), "qsx_reduce")
        return float(L.to_host(out)[0])

    def to_csr(self, dtype=np.complex128):
        return build_csr(self.basis, self.oplist, dtype, self.transposed, self.conjugated, diag=self.diag)

    def expm_inplace(self, v, work, s, m_star, tol, mu, a):
        vdt = L.np_dtype(v.dtype)
        nvec = int(np.prod(v.shape[1:])) if v.dim() > 1 else 1
        ctx = self._context(vdt, nvec)
        mu, a = complex(mu), complex(a)
        rc = L.load().qsx_expm_multiply_mf(ctx.h, self.Ns, int(s), int(m_star), float(tol), mu.real, mu.imag,
                                           a.real, a.imag, ptr(v), ptr(work), L.dtype_code(vdt), nvec,
                                           stream_handle())
        L.check(rc, "qsx_expm_multiply_mf")


class SectorBatch:
    """Many sectors x ``nvec`` vectors in one launch (``qsx_hv_batch_*``).

    ``ops`` are matrix-free operators of the same dtype and site count (one
    per symmetry sector).  Their vectors are stacked row-wise into one
    ``(rows, nvec)`` device array, sector k starting at ``row_off[k]``.
    ``form``: "mf" evaluates the matrix elements on the fly (wavefront per
    row); "csr" builds every sector's CSR matrix once on the device and
    streams it (``qsx_csr_batch_create``); None lets the autotuner time both
    and keep the faster."""

    def __new__(cls, ops, vdt, nvec, form="mf", chunk_rows=None):
        if form is None:
            form = tuning.choose_batch_form(cls, ops, vdt, nvec)
        self = super().__new__(cls)
        self._form = form
        return self

    def __init__(self, ops, vdt, nvec, form="mf", chunk_rows=None):
        if not ops:
            raise ValueError("at least one sector is needed")
        for op in ops:
            if not isinstance(op, MatrixFreeOp):
                raise TypeError("batched sectors need matrix-free operators")
        self.ops = list(ops)
        self.vdt = np.dtype(vdt)
        self.nvec = int(nvec)
        self.form = self._form
        lib = L.load()
        h = L._vp()
        n = len(self.ops)
        if self.form == "mf":
            self._ctxs = []
            for op in self.ops:
                ctx = op._new_ctx(self.vdt, self.nvec)
                ctx.configure({"lookup": "table" if len(op.basis._pers) else "search", "schedule": "rows"},
                              op.basis)
                self._ctxs.append(ctx)
            arr = (C.c_void_p * n)(*[c.h.value for c in self._ctxs])
            L.check(lib.qsx_hv_batch_create(n, C.cast(arr, C.c_void_p), self.nvec, int(chunk_rows or 16),
                                            C.byref(h)), "qsx_hv_batch_create")
        elif self.form == "csr":
            self._csr = [op.to_csr(self.vdt) for op in self.ops]
            for M in self._csr:
                if L.np_dtype(M.indices.dtype) != np.int32:
                    raise NotImplementedError("CSR batches need int32 column indices")
                if L.np_dtype(M.data.dtype) != self.vdt:
                    raise TypeError("CSR batch values must have the vector dtype")
            Ns = np.array([M.shape[0] for M in self._csr], np.int64)
            ip = (C.c_void_p * n)(*[M.indptr.data_ptr() for M in self._csr])
            i64 = np.array([int(L.np_dtype(M.indptr.dtype) == np.int64) for M in self._csr], np.int32)
            ix = (C.c_void_p * n)(*[M.indices.data_ptr() for M in self._csr])
            dt = (C.c_void_p * n)(*[M.data.data_ptr() for M in self._csr])
            L.check(lib.qsx_csr_batch_create(n, Ns.ctypes.data_as(L._vp), C.cast(ip, C.c_void_p),
                                             i64.ctypes.data_as(L._vp), C.cast(ix, C.c_void_p),
                                             C.cast(dt, C.c_void_p), L.dtype_code(self.vdt), self.nvec,
                                             int(min(chunk_rows or 16, 63)), C.byref(h)), "qsx_csr_batch_create")
        else:
            raise ValueError("form must be 'mf', 'csr' or None")
        self.h = h
        ro = np.zeros(n + 1, np.int64)
        lib.qsx_hv_batch_rows(self.h, ro.ctypes.data_as(L._vp))
        self.row_off = ro
        self.rows = int(ro[-1])
        self.chunk_rows, self.xcd_slices = int(chunk_rows or 16), 0
        if chunk_rows is None:
            self.set_schedule(*tuning.choose_batch_schedule(self))

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and L._lib is not None:
                L._lib.qsx_hv_batch_destroy(self.h)
        except Exception:
            pass

    def set_schedule(self, chunk_rows, xcd_slices=0):
        """Rows per wavefront chunk and chunk order (XCD slices or dispatch order)."""
        L.check(L.load().qsx_hv_batch_set_schedule(self.h, int(chunk_rows), int(bool(xcd_slices))),
                "qsx_hv_batch_set_schedule")
        self.chunk_rows, self.xcd_slices = int(chunk_rows), int(bool(xcd_slices))

    def stack(self, vs):
        """One (rows, nvec) device array from per-sector (Ns_k, nvec) arrays."""
        t = L.torch()
        parts = [L.to_device(np.asarray(v), self.vdt) if not L.is_device_tensor(v) else v.to(L.torch_dtype(self.vdt))
                 for v in vs]
        return t.cat([p.reshape(p.shape[0], self.nvec) for p in parts], 0).contiguous()

    def unstack(self, X):
        return [X[self.row_off[k]:self.row_off[k + 1]] for k in range(len(self.ops))]

    def empty(self):
        return L.empty((self.rows, self.nvec), self.vdt)

    def apply(self, x, y, a=1.0, beta=0):
        """y = a H x (+ y), every sector in one launch."""
        self._check(x)
        self._check(y)
        ac = complex(a)
        L.check(L.load().qsx_hv_batch_apply(self.h, ptr(x), ptr(y), ac.real, ac.imag, int(bool(beta)),
                                            stream_handle()), "qsx_hv_batch_apply")

    def expm_inplace(self, v, work, parts, tols, mus, a, graph=True):
        """v <- e^{a H_k} v_k for every sector (per-sector (s, m), tol, mu)."""
        self._check(v)
        if work.numel() != 2 * self.rows * self.nvec or L.np_dtype(work.dtype) != self.vdt:
            raise ValueError("work must hold 2 * rows * nvec elements of the vector dtype")
        s = np.ascontiguousarray([p[1] for p in parts], dtype=np.int32)
        m = np.ascontiguousarray([p[0] for p in parts], dtype=np.int32)
        tol = np.ascontiguousarray(tols, dtype=np.float64)
        mu = np.ascontiguousarray(np.column_stack([np.real(mus), np.imag(mus)]).ravel(), dtype=np.float64)
        ac = complex(a)
        L.check(L.load().qsx_expm_multiply_batch(self.h, s.ctypes.data_as(L._vp), m.ctypes.data_as(L._vp),
                                                 tol.ctypes.data_as(L._vp), mu.ctypes.data_as(L._vp), ac.real,
                                                 ac.imag, ptr(v), ptr(work), int(bool(graph)), stream_handle()),
                "qsx_expm_multiply_batch")

    def _check(self, x):
        if tuple(x.shape) not in ((self.rows, self.nvec),) + (((self.rows,),) if self.nvec == 1 else ()):
            raise ValueError("expected a stacked ({}, {}) array, got {}".format(self.rows, self.nvec, tuple(x.shape)))
        if L.np_dtype(x.dtype) != self.vdt or not x.is_contiguous():
            raise ValueError("stacked arrays must be contiguous {}".format(self.vdt))


class _Ctx:
    def __init__(self, h):
        self.h = h
        self.cfg = {"lookup": "search", "schedule": "rows"}

    def configure(self, cfg, basis):
        lib = L.load()
        self.cfg = dict(cfg)
        if cfg.get("lookup") == "table":
            tab = basis.lookup_table()
            if tab is None:
                self.cfg["lookup"] = "search"
                L.check(lib.qsx_hv_ctx_set_table(self.h, None, 0, 0, None, 0), "qsx_hv_ctx_set_table")
            else:
                t, eb, ib, lut = tab
                L.check(lib.qsx_hv_ctx_set_table(self.h, ptr(t), eb, ib, ptr(lut) if lut is not None else None,
                                                  int(lut.numel()) if lut is not None else 0),
                        "qsx_hv_ctx_set_table")
        else:
            L.check(lib.qsx_hv_ctx_set_table(self.h, None, 0, 0, None, 0), "qsx_hv_ctx_set_table")
        code = tuning.SCHEDULES.get(cfg.get("schedule"), 0)
        rc = lib.qsx_hv_ctx_set_schedule(self.h, code)
        if rc == L.QSX_ERR_UNSUPPORTED and code == tuning.SCHEDULES["split"]:
            # the split-chain layout needs one hcb sector without symmetry
            # maps and bit-parallel terms: other operators keep one lane per row
            self.cfg["schedule"] = "rows"
            rc = lib.qsx_hv_ctx_set_schedule(self.h, 0)
        L.check(rc, "qsx_hv_ctx_set_schedule")

    def __del__(self):
        try:
            if L._lib is not None:
                L._lib.qsx_hv_ctx_destroy(self.h)
        except Exception:
            pass


def _idx_dtype(nnz, rows):
    return np.int64 if (nnz >= 2 ** 31 - 1 or rows >= 2 ** 31 - 1) else np.int32


def _combine_dtype(mats, coefs, conj):
    dt = np.result_type(*[M.dtype for M in mats])
    if dt.kind != "c" and any(complex(c).imag != 0 for c in coefs):
        dt = np.result_type(dt, np.complex64)
    return np.dtype(dt)


def csr_combine(mats, coefs, dtype=None, conj=None):
    """sum_k coefs[k] * op_k(mats[k]) of device CSR matrices of one shape,
    op_k the complex conjugate where conj[k], on the device
    (``qsx_csr_sum_*``); entries summing to exactly zero are dropped.  The
    result dtype defaults to that of the matrices, complex when a
    coefficient is."""
    K = len(mats)
    if K == 0:
        raise ValueError("nothing to combine")
    conj = [bool(c) for c in (conj if conj is not None else [False] * K)]
    coefs = [complex(c) for c in coefs]
    dt = np.dtype(dtype) if dtype is not None else _combine_dtype(mats, coefs, conj)
    if K > 32:
        head = csr_combine(mats[:32], coefs[:32], dt, conj[:32])
        return csr_combine([head] + list(mats[32:]), [1.0] + coefs[32:], dt, [False] + conj[32:])
    shape = mats[0].shape
    if any(M.shape != shape for M in mats):
        raise ValueError("matrices of different shapes")
    lib = L.load()
    vp = C.c_void_p * K
    ia = C.c_int * K
    coef_arr = (C.c_double * (2 * K))(*[x for c in coefs for x in (c.real, c.imag)])
    h = L._vp()
    nnz = L._i64()
    rc = lib.qsx_csr_sum_create(K, vp(*[ptr(M.indptr) for M in mats]), vp(*[ptr(M.indices) for M in mats]),
                                vp(*[ptr(M.data) for M in mats]), ia(*[M.idx64 for M in mats]),
                                ia(*[L.dtype_code(M.dtype) for M in mats]), ia(*[int(c) for c in conj]), coef_arr,
                                shape[0], C.byref(h), C.byref(nnz), stream_handle())
    try:
        L.check(rc, "qsx_csr_sum_create")
        n = nnz.value
        idx = _idx_dtype(n, shape[0])
        indptr = L.empty(shape[0] + 1, idx)
        indices = L.empty(max(n, 1), idx)
        data = L.empty(max(n, 1), dt)
        L.check(lib.qsx_csr_sum_finish(h, L.dtype_code(dt), np.dtype(idx).itemsize, ptr(indptr), ptr(indices),
                                       ptr(data), stream_handle()), "qsx_csr_sum_finish")
    finally:
        lib.qsx_csr_sum_destroy(h)
    return DeviceCSR(indptr, indices[:n], data[:n], shape, dt)


def csr_transpose(M, conj=False, dtype=None):
    """The transpose (conj: the conjugate transpose) of a device CSR, on the
    device (``qsx_csr_transpose``)."""
    dt = np.dtype(dtype) if dtype is not None else M.dtype
    Nr, Nc = M.shape
    idx = _idx_dtype(M.nnz, Nc)
    indptr = L.empty(Nc + 1, idx)
    indices = L.empty(max(M.nnz, 1), idx)
    data = L.empty(max(M.nnz, 1), dt)
    L.check(L.load().qsx_csr_transpose(ptr(M.indptr), ptr(M.indices), ptr(M.data), M.idx64, L.dtype_code(M.dtype),
                                       Nr, Nc, M.nnz, int(bool(conj)), L.dtype_code(dt), np.dtype(idx).itemsize,
                                       ptr(indptr), ptr(indices), ptr(data), stream_handle()), "qsx_csr_transpose")
    return DeviceCSR(indptr, indices[:M.nnz], data[:M.nnz], (Nc, Nr), dt)


def csr_diagonal(d, dtype):
    """Diagonal device CSR with the entries of the device vector d."""
    t = L.torch()
    n = int(d.numel())
    idx = _idx_dtype(n, n)
    ar = t.arange(n + 1, dtype=L.torch_dtype(idx), device="cuda")
    return DeviceCSR(ar, ar[:n], d.to(L.torch_dtype(np.dtype(dtype))).contiguous(), (n, n), dtype)


def build_csr(basis, oplist, dtype, transposed=False, conjugated=False, a=1.0, diag=None):
    """Assemble a*Op as a canonical device CSR (``qsx_csr_builder_*``).
    Lists that do not commute with the symmetry group are assembled through
    their transpose (rows gathered with the original terms), transposed
    back on the device."""
    if len(basis._pers) and not oplist.symmetric and not transposed:
        M = csr_transpose(build_csr(basis, oplist, dtype, True, conjugated, a, None))
        if diag is not None:
            M = csr_combine([M, csr_diagonal(diag, dtype)], [1.0, 1.0], dtype)
        return M
    lib = L.load()
    dtype = np.dtype(dtype)
    h = L._vp()
    nnz = L._i64()
    ac = complex(a)
    rc = lib.qsx_csr_builder_create(basis._core.h, oplist.h, *basis._dev_args(), int(transposed), int(conjugated),
                                    ac.real, ac.imag, C.byref(h), C.byref(nnz), stream_handle())
    L.check(rc, "qsx_csr_builder_create")
    try:
        n = nnz.value
        Ns = basis.Ns
        idx = np.int64 if (n >= 2 ** 31 - 1 or Ns >= 2 ** 31 - 1) else np.int32
        indptr = L.empty(Ns + 1, idx)
        indices = L.empty(max(n, 1), idx)
        data = L.empty(max(n, 1), dtype)
        rc = lib.qsx_csr_builder_finish(h, L.dtype_code(dtype), np.dtype(idx).itemsize, ptr(indptr), ptr(indices),
                                        ptr(data), stream_handle())
        L.check(rc, "qsx_csr_builder_finish")
    finally:
        lib.qsx_csr_builder_destroy(h)
    M = DeviceCSR(indptr, indices[:n], data[:n], (basis.Ns, basis.Ns), dtype)
    if diag is not None:
        M = csr_combine([M, csr_diagonal(diag, dtype)], [1.0, 1.0], dtype)
    return M


def as_device_operator(A, dtype=None):
    """Device operator for the tools: our operators, scipy sparse or dense."""
    if hasattr(A, "_qsx_operator"):
        return A._qsx_operator()
    if isinstance(A, (DeviceCSR, MatrixFreeOp)):
        return A
    if sp.issparse(A) or isinstance(A, np.ndarray):
        return DeviceCSR.from_scipy(sp.csr_matrix(A), dtype)
    return None


def vdot_dev(x, y):
    """sum conj(x) y on the device (FP64 accumulation, deterministic)."""
    out = L.zeros(2, np.float64)
    rc = L.load().qsx_vdot(_vec_dtype_code(x), ptr(x), ptr(y), int(x.numel()), ptr(out), stream_handle())
    L.check(rc, "qsx_vdot")
    o = L.to_host(out)
    return complex(o[0], o[1])


def nrm2_dev(x):
    out = L.zeros(2, np.float64)
    rc = L.load().qsx_nrm2(_vec_dtype_code(x), ptr(x), int(x.numel()), ptr(out), stream_handle())
    L.check(rc, "qsx_nrm2")
    return float(L.to_host(out)[0])
