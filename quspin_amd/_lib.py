"""ctypes binding of ``libquspin_hip.so`` (the C-ABI declared in
``include/quspin_hip.h``) and device-buffer helpers.

The library is the only compute path: nothing here falls back to the CPU.
Loading fails loudly when the shared library is missing, and every device
entry point raises when no ROCm device is visible.
"""

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libquspin_hip.so")

F32, F64, C64, C128 = 0, 1, 2, 3
KIND_HCB, KIND_SPINLESS, KIND_SPINFUL = 0, 1, 2

QSX_OK = 0
QSX_ERR_ARG = -1
QSX_ERR_HIP = -2
QSX_ERR_OVERFLOW = -3
QSX_ERR_COMPLEX = -4
QSX_ERR_UNSUPPORTED = -5

_vp = C.c_void_p
_i32 = C.c_int
_i64 = C.c_int64
_dbl = C.c_double
_pd = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol declared in include/quspin_hip.h
SIGNATURES = {
    "qsx_version": (_i32, []),
    "qsx_last_error": (C.c_char_p, []),
    "qsx_core_create": (_i32, [_i32, _i32, _vp, _vp, _vp, _i32, _i32, C.POINTER(_vp)]),
    "qsx_core_destroy": (None, [_vp]),
    "qsx_core_group_order": (_i32, [_vp]),
    "qsx_core_map_info": (_i32, [_vp, _i32, C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_i32)]),
    "qsx_make_basis": (_i32, [_vp, _vp, _i32, _vp, _vp, _i32, _i64, C.POINTER(_i64), _vp]),
    "qsx_make_basis_blocks": (_i32, [_vp, _i32, _vp, _i64, _vp, _vp, _vp]),
    "qsx_representative": (_i32, [_vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "qsx_normalization": (_i32, [_vp, _vp, _vp, _i32, _i64, _vp]),
    "qsx_oplist_create": (_i32, [_vp, _i32, C.c_char_p, _vp, _vp, _vp, _i32, C.POINTER(_vp)]),
    "qsx_oplist_destroy": (None, [_vp]),
    "qsx_oplist_info": (_i32, [_vp, C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_i32)]),
    "qsx_hv": (_i32, [_vp, _vp, _vp, _vp, _i32, _i64, _vp, _vp, _i32, _vp, _vp, _i32, _i64, _i32, _i32, _i32,
                      _dbl, _dbl, _dbl, _dbl, _i32, _vp, _i32, _vp]),
    "qsx_inplace_op": (_i32, [_vp, _vp, _vp, _i32, _i64, _i32, _i32, C.c_char_p, _vp, _i32, _dbl, _dbl, _vp, _vp,
                              _i32, _i64, _vp, _vp, _i32, _vp]),
    "qsx_op": (_i32, [_vp, _vp, _vp, _i32, _vp, C.c_char_p, _vp, _i32, _dbl, _dbl, _vp, _vp, _i32, _i64, _vp, _vp,
                      _i32, _vp]),
    "qsx_hv_ctx_create": (_i32, [_vp, _vp, _vp, _vp, _i32, _i64, _vp, _vp, _i32, _i32, _i64, _i32, _i32, _i32, _vp,
                                 _i32, C.POINTER(_vp)]),
    "qsx_hv_ctx_destroy": (None, [_vp]),
    "qsx_hv_ctx_set_table": (_i32, [_vp, _vp, _i32, _i32, _vp, _i32]),
    "qsx_hv_ctx_set_diag": (_i32, [_vp, _vp, _i32]),
    "qsx_hv_ctx_set_schedule": (_i32, [_vp, _i32]),
    "qsx_table_info": (_i32, [_vp, _i64, _vp, _i32, C.POINTER(_i64), C.POINTER(_i32), C.POINTER(_i32),
                              C.POINTER(_i64)]),
    "qsx_table_build": (_i32, [_vp, _vp, _i64, _vp, _vp, _i32, _vp, _i32, _vp, _i32, _i32, _vp, _i64, _vp]),
    "qsx_hv_ctx_apply": (_i32, [_vp, _vp, _vp, _dbl, _dbl, _dbl, _dbl, _i32, _vp]),
    "qsx_hv_ctx_column_norms": (_i32, [_vp, _dbl, _dbl, _vp, _vp, _vp]),
    "qsx_reduce": (_i32, [_i32, _vp, _i64, _vp, _vp]),
    "qsx_op_shift_sector": (_i32, [_vp, _vp, _vp, _vp, _i32, _i64, C.c_char_p, _vp, _i32, _dbl, _dbl, _vp, _vp, _i32,
                                   _i64, _vp, _vp, _i32, _vp, _vp, _i32, _i64, _vp]),
    "qsx_get_amp": (_i32, [_vp, _vp, _vp, _i32, _i64, _i32, _vp]),
    "qsx_project_from_dense": (_i32, [_vp, _vp, _vp, _i32, _i64, _vp, _vp, _i32, _i64, _i32, _i64, _vp]),
    "qsx_project_to_dense": (_i32, [_vp, _vp, _vp, _i32, _i64, _vp, _vp, _i32, _i64, _i32, _i64, _vp]),
    "qsx_get_proj": (_i32, [_vp, _vp, _vp, _i32, _i64, _i32, _i64, _i32, _i32, _vp, _vp, _vp, C.POINTER(_i64),
                            _vp]),
    "qsx_csr_builder_create": (_i32, [_vp, _vp, _vp, _vp, _i32, _i64, _vp, _vp, _i32, _i32, _i32, _dbl, _dbl,
                                      C.POINTER(_vp), C.POINTER(_i64), _vp]),
    "qsx_csr_builder_complex": (_i32, [_vp]),
    "qsx_csr_builder_finish": (_i32, [_vp, _i32, _i32, _vp, _vp, _vp, _vp]),
    "qsx_csr_builder_destroy": (None, [_vp]),
    "qsx_csr_matvec": (_i32, [_vp, _vp, _vp, _i32, _i32, _i64, _i64, _vp, _vp, _i32, _i64, _dbl, _dbl, _i32, _i32,
                              _vp]),
    "qsx_csr_trace": (_i32, [_vp, _vp, _vp, _i32, _i32, _i64, _vp, _vp]),
    "qsx_csr_1_norm": (_i32, [_vp, _vp, _vp, _i32, _i32, _i64, _dbl, _dbl, _vp, _vp, _vp]),
    "qsx_csr_1_norm_diag": (_i32, [_vp, _vp, _vp, _i32, _i32, _i64, _dbl, _dbl, _vp, _dbl, _dbl, _vp, _vp, _vp]),
    "qsx_csr_col_split": (_i32, [_vp, _vp, _vp, _i32, _i32, _i64, _vp, _vp, _vp]),
    "qsx_csr_sum_create": (_i32, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "qsx_csr_sum_finish": (_i32, [_vp, _i32, _i32, _vp, _vp, _vp, _vp]),
    "qsx_csr_sum_destroy": (None, [_vp]),
    "qsx_csr_transpose": (_i32, [_vp, _vp, _vp, _i32, _i32, _i64, _i64, _i64, _i32, _i32, _i32, _vp, _vp, _vp, _vp]),
    "qsx_partition_exact": (_i32, [_vp, _vp, _vp, _i64, _vp, _i32, _vp, _vp, _i32, _dbl, _vp, _vp, _vp]),
    "qsx_expm_multiply_csr": (_i32, [_vp, _vp, _vp, _i32, _i32, _i64, _i64, _i32, _i32, _dbl, _dbl, _dbl, _dbl,
                                     _dbl, _vp, _vp, _i32, _i64, _i32, _vp]),
    "qsx_expm_multiply_csr_noalloc": (_i32, [_vp, _vp, _vp, _i32, _i32, _i64, _i64, _i32, _i32, _dbl, _dbl, _dbl,
                                             _dbl, _dbl, _vp, _vp, _i32, _i64, _i32, _vp]),
    "qsx_expm_multiply_csr_noalloc_max_bytes": (_i64, []),
    "qsx_expm_multiply_mf": (_i32, [_vp, _i64, _i32, _i32, _dbl, _dbl, _dbl, _dbl, _dbl, _vp, _vp, _i32, _i64,
                                    _vp]),
    "qsx_vdot": (_i32, [_i32, _vp, _vp, _i64, _vp, _vp]),
    "qsx_nrm2": (_i32, [_i32, _vp, _i64, _vp, _vp]),
    "qsx_axpy": (_i32, [_i32, _vp, _vp, _i64, _dbl, _dbl, _vp, _dbl, _i32, _vp]),
    "qsx_lanczos_update": (_i32, [_i32, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "qsx_lanczos_div": (_i32, [_i32, _vp, _vp, _vp, _i64, _vp]),
    "qsx_lanczos_run": (_i32, [_i32, _vp, _vp, _vp, _vp, _i32, _i32, _i64, _i32, _i32, _i64, _i32, _vp, _i32, _vp, _vp,
                               _i32, _vp]),
    "qsx_lincomb": (_i32, [_i32, _vp, _i64, _i64, _vp, _vp, _vp]),
    "qsx_copy": (_i32, [_vp, _vp, _i64, _vp]),
    "qsx_sector_rank": (_i32, [_vp, _vp, _i64, _vp, _vp, _vp]),
    "qsx_hv_batch_create": (_i32, [_i32, _vp, _i64, _i32, C.POINTER(_vp)]),
    "qsx_csr_batch_create": (_i32, [_i32, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _i32, C.POINTER(_vp)]),
    "qsx_hv_batch_destroy": (None, [_vp]),
    "qsx_hv_batch_rows": (_i32, [_vp, _vp]),
    "qsx_hv_batch_set_schedule": (_i32, [_vp, _i32, _i32]),
    "qsx_hv_batch_apply": (_i32, [_vp, _vp, _vp, _dbl, _dbl, _i32, _vp]),
    "qsx_expm_multiply_batch": (_i32, [_vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp, _i32, _vp]),
    "qsx_expm_record_bytes": (_i32, []),
    "qsx_evolve_csr_ctl_bytes": (_i64, [_i32, _i32]),
    "qsx_evolve_csr_steps": (_i32, [_vp, _vp, _vp, _i32, _i32, _i64, _i64, _vp, _i32, _vp, _i32, _i32, _vp, _vp, _i32,
                                    _i32, _i64, _vp, _vp, _i64, _i32, _i32, _vp]),
    "qsx_graph_cache_clear": (None, []),
    "qsx_kron_create": (_i32, [_vp, _vp, _vp, _i32, _i64, _i64, _vp, _vp, _vp, _i32, _i64, _i64, _vp, _i32,
                               C.POINTER(_vp)]),
    "qsx_kron_destroy": (None, [_vp]),
    "qsx_kron_set_mode": (_i32, [_vp, _i32]),
    "qsx_kron_matvec": (_i32, [_vp, _vp, _dbl, _dbl, _vp, _vp, _i32, _dbl, _dbl, _i32, _vp]),
    "qsx_kron_1_norm": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _dbl, _dbl, _vp, _vp, _vp]),
    "qsx_evolve_kron_steps": (_i32, [_vp, _vp, _i32, _vp, _i32, _i32, _vp, _vp, _i32, _i32, _vp, _vp, _i64, _i32,
                                     _vp]),
    "qsx_mlp_param_count": (_i32, []),
    "qsx_mlp_forward": (_i32, [_vp, _vp, _vp, _i64, _vp, _vp]),
    "qsx_mlp_grad_work": (_i64, []),
    "qsx_mlp_grad": (_i32, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp]),
    "qsx_mlp_adam": (_i32, [_vp, _vp, _vp, _vp, _i64, _dbl, _dbl, _dbl, _dbl, _vp]),
    "qsx_mlp_partition": (_i32, [_vp, _vp, _vp, _i64, _vp, _i32, _i32, _i32, _vp]),
}

_lib = None


class QsxError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} not found: build it with quspin_amd._build.build() "
            "(there is no CPU fallback)")
    # torch ships its own libamdhip64 with SONAME libamdhip64.so.7; loading it
    # first makes this library bind to the same HIP runtime (one runtime per
    # process, so torch's streams and allocations are valid here)
    torch()
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def last_error():
    return load().qsx_last_error().decode(errors="replace")


def check(rc, what=""):
    """Map a negative status to the exception the reference raises."""
    if rc >= 0:
        return rc
    msg = last_error()
    if rc == QSX_ERR_COMPLEX:
        raise TypeError(msg)
    if rc in (QSX_ERR_ARG, QSX_ERR_OVERFLOW):
        raise ValueError(msg)
    if rc == QSX_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise QsxError(f"{what}: {msg}")


# ----------------------------------------------------------------------------
# device helpers (torch is used for allocation and streams only)
# ----------------------------------------------------------------------------
_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as _t
        _torch = _t
    return _torch


def require_device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("quspin_amd needs a ROCm GPU (no CPU fallback exists)")
    load()
    return t


def stream_handle():
    t = torch()
    return _vp(t.cuda.current_stream().cuda_stream)


def ptr(x):
    if x is None:
        return None
    return _vp(x.data_ptr())


_NP2CODE = {np.dtype(np.float32): F32, np.dtype(np.float64): F64,
            np.dtype(np.complex64): C64, np.dtype(np.complex128): C128}


def dtype_code(dt):
    dt = np.dtype(dt)
    if dt not in _NP2CODE:
        raise TypeError(f"unsupported dtype {dt}")
    return _NP2CODE[dt]


def torch_dtype(dt):
    t = torch()
    return {np.dtype(np.float32): t.float32, np.dtype(np.float64): t.float64,
            np.dtype(np.complex64): t.complex64, np.dtype(np.complex128): t.complex128,
            np.dtype(np.int32): t.int32, np.dtype(np.int64): t.int64,
            np.dtype(np.uint8): t.uint8, np.dtype(np.int8): t.int8,
            np.dtype(np.int16): t.int16}[np.dtype(dt)]


def np_dtype(tdt):
    t = torch()
    return np.dtype({t.float32: np.float32, t.float64: np.float64, t.complex64: np.complex64,
                     t.complex128: np.complex128, t.int32: np.int32, t.int64: np.int64,
                     t.uint8: np.uint8, t.int8: np.int8, t.int16: np.int16}[tdt])


def is_device_tensor(x):
    if _torch is None:
        try:
            import torch as _t  # noqa: F401
        except Exception:
            return False
    t = torch()
    return isinstance(x, t.Tensor) and x.is_cuda


def to_device(a, dtype=None):
    """numpy (or tensor) -> contiguous device tensor of ``dtype``."""
    t = require_device()
    if isinstance(a, t.Tensor):
        x = a
        if dtype is not None and np_dtype(x.dtype) != np.dtype(dtype):
            x = x.to(torch_dtype(dtype))
        if not x.is_cuda:
            x = x.cuda()
        return x.contiguous()
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    elif a.dtype == np.uint64:
        a = a.view(np.int64)
    elif a.dtype == np.uint16:
        a = a.view(np.int16)
    return t.from_numpy(a.copy() if not a.flags.writeable else a).cuda()


def empty(shape, dtype):
    t = require_device()
    return t.empty(shape, dtype=torch_dtype(dtype), device="cuda")


def zeros(shape, dtype):
    t = require_device()
    return t.zeros(shape, dtype=torch_dtype(dtype), device="cuda")


def to_host(x, dtype=None):
    a = x.detach().cpu().numpy()
    if dtype is not None:
        a = a.view(dtype) if np.dtype(dtype).itemsize == a.dtype.itemsize else a.astype(dtype)
    return a
