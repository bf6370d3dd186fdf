"""``matvec`` / ``get_matvec_function`` (parallel_sparse_tools matvec_core,
rendered at docs/modules/parallel_sparse_tools/matvec/matvec_core.html) on
the device CSR kernel: x = a A v, or x += a A v when ``overwrite_out`` is
False and ``out`` is given."""

import numpy as np
import scipy.sparse as sp

from .. import _lib as L
from ..operators._devop import DeviceCSR, as_device_operator


def _matvec(array, other, overwrite_out=False, out=None, a=1.0):
    op = as_device_operator(array)
    if op is None:
        op = DeviceCSR.from_scipy(sp.csr_matrix(array))
    on_dev = L.is_device_tensor(other)
    odt = L.np_dtype(other.dtype) if on_dev else np.asarray(other).dtype
    result_dtype = np.result_type(op.dtype, odt, np.min_scalar_type(a))
    x = L.to_device(other, result_dtype)
    if out is None:
        y = L.empty(tuple(x.shape), result_dtype)
        beta = 0
    else:
        y = out if L.is_device_tensor(out) else L.to_device(out, result_dtype)
        beta = 0 if overwrite_out else 1
    op.apply(x, y, a=a, beta=beta)
    if on_dev or L.is_device_tensor(out):
        return y
    res = L.to_host(y)
    if out is not None:
        out[...] = res
        return out
    return res


def matvec(array, other, overwrite_out=False, out=None, a=1.0):
    """x = a A v (new array) or x (+)= a A v into ``out``."""
    return _matvec(array, other, overwrite_out=overwrite_out, out=out, a=a)


def get_matvec_function(array):
    """Bind the device CSR of ``array`` once; returns f(array, other, ...)."""
    op = as_device_operator(array)
    if op is None:
        op = DeviceCSR.from_scipy(sp.csr_matrix(array))

    def f(_array, other, overwrite_out=False, out=None, a=1.0):
        return _matvec(op, other, overwrite_out=overwrite_out, out=out, a=a)

    return f


_get_matvec_function = get_matvec_function
