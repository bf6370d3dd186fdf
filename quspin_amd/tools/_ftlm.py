"""Finite- and low-temperature Lanczos iterations (FTLM, LTLM) on Lanczos
data held on the device.

Restates ``FTLM_static_iteration`` (src/quspin/tools/lanczos/_FTLM.py:10-105)
and ``LTLM_static_iteration`` (_LTLM.py:9-139).  The Krylov quantities
(E, V) are small host arrays; the operator products A.dot(q) and the
overlaps <q_i | A q> run on the device when ``Q_T`` is a device array
(``lanczos_full`` on a device start vector returns one), with FP64
accumulation.  Host arrays and iterators of vectors take the same path
vector by vector.
"""

import numpy as np

from .. import _lib as L
from .._lib import ptr, stream_handle

__all__ = ["FTLM_static_iteration", "LTLM_static_iteration"]


def _rows(Q_T, limit=None):
    """The Lanczos vectors as a list of device vectors (at most ``limit``)."""
    out = []
    it = Q_T if L.is_device_tensor(Q_T) else (Q_T[:] if isinstance(Q_T, np.ndarray) else Q_T)
    for i, q in enumerate(it):
        if limit is not None and i >= limit:
            break
        out.append(q if L.is_device_tensor(q) else L.to_device(np.ascontiguousarray(q)))
    return out


def _dot_dev(A, q):
    """A.dot(q) as a device vector of the result dtype."""
    r = A.dot(q)
    if not L.is_device_tensor(r):
        r = L.to_device(np.ascontiguousarray(r))
    return r


def _overlaps(rows, x):
    """<rows[i] | x> for every row (np.vdot order: the rows are conjugated),
    one FP64 device reduction each."""
    lib = L.load()
    st = stream_handle()
    n = len(rows)
    out = L.zeros((max(n, 1), 2), np.float64)
    for i, q in enumerate(rows):
        dt = np.result_type(L.np_dtype(q.dtype), L.np_dtype(x.dtype))
        qa = q if L.np_dtype(q.dtype) == dt else q.to(L.torch_dtype(dt))
        xa = x if L.np_dtype(x.dtype) == dt else x.to(L.torch_dtype(dt))
        L.check(lib.qsx_vdot(L.dtype_code(dt), ptr(qa.contiguous()), ptr(xa.contiguous()), int(qa.numel()),
                             ptr(out[i]), st), "qsx_vdot")
    o = L.to_host(out)[:n]
    return o[:, 0] + 1j * o[:, 1]


def FTLM_static_iteration(O_dict, E, V, Q_T, beta=0):
    """One FTLM sample (_FTLM.py:10-105): for the start vector r = Q_T[0],
    <O>_r = sum_i c_i <q_i | O r> with c_i = sum_j V[0, j] V[i, j] e^{-beta E_j},
    and the identity's c_0.  Returns (dict of results shaped like beta, c_0)."""
    E = np.asarray(E)
    V = np.asarray(V)
    p = np.exp(-np.outer(np.atleast_1d(beta), E))
    c = np.einsum("j,aj,...j->a...", V[0, :], V, p)
    rows = _rows(Q_T)
    r = rows[0]
    results = {}
    for key, A in O_dict.items():
        Ar = _dot_dev(A, r)
        ov = _overlaps(rows, Ar)
        if L.np_dtype(r.dtype).kind != "c" and L.np_dtype(Ar.dtype).kind != "c":
            ov = ov.real
        res = sum(c[i, ...] * ov[i] for i in range(len(rows)))
        results[key] = np.squeeze(res)
    return results, np.squeeze(c[0, ...])


def LTLM_static_iteration(O_dict, E, V, Q_T, beta=0):
    """One LTLM sample (_LTLM.py:9-139): the Krylov vectors whose weight
    e^{-beta_min E} |V[0]| stays above machine epsilon, the operator in that
    space <q_j | O q_i>, and <O>_r = sum_{jl} c_j c_l (V^T O V)_{jl} with
    c_j = V[0, j] e^{-beta E_j / 2}.  Returns (dict, <I>_r)."""
    E = np.asarray(E)
    V = np.asarray(V)
    nv = E.size
    beta = np.atleast_1d(beta)
    p = np.exp(-E * beta.min()) * np.abs(V[0, :]).max()
    mask = p < np.finfo(p.dtype).eps
    if np.any(mask):
        nv = int(np.argmax(mask))
    rows = _rows(Q_T, limit=nv)
    Ome = {}
    for i, q in enumerate(rows):
        for key, A in O_dict.items():
            Ar = _dot_dev(A, q)
            ov = _overlaps(rows, Ar)
            if key not in Ome:
                dtype = np.result_type(L.np_dtype(q.dtype), np.dtype(getattr(A, "dtype", np.float64)))
                Ome[key] = np.zeros((nv, nv), dtype=dtype)
            Ome[key][i, :] = ov if Ome[key].dtype.kind == "c" else ov.real
    p = np.exp(-np.outer(E[:nv], beta / 2.0))
    Vn = V[:nv, :nv].copy()
    c = np.einsum("j,j...->j...", Vn[0, :], p)
    results = {}
    for key, Ame in Ome.items():
        A_diag = Vn.T.dot(Ame.dot(Vn))
        results[key] = np.squeeze(np.einsum("j...,l...,jl->...", c, c, A_diag))
    p = np.exp(-np.outer(E[:nv], beta))
    Id = np.einsum("j,j...->...", Vn[0, :] ** 2, p)
    return results, np.squeeze(Id)
