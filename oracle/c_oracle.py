"""Build and load the C/OpenMP restatement (``oracle/csrc/qsx_oracle.c``) —
TEST INFRASTRUCTURE ONLY (checker and CPU baseline).

Built with gcc -O3 -fopenmp into ``oracle/_build/libqsx_oracle.so``.
"""

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "qsx_oracle.c")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libqsx_oracle.so")

_lib = None


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-o", LIB + ".tmp", SRC, "-lm"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(p.stderr)
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    lib.qo_group_create.restype = vp
    lib.qo_group_create.argtypes = [i32, i32, vp, vp, vp]
    lib.qo_group_destroy.argtypes = [vp]
    lib.qo_make_basis.restype = i64
    lib.qo_make_basis.argtypes = [vp, i32, vp, vp, i64]
    lib.qo_representative.restype = C.c_uint64
    lib.qo_representative.argtypes = [vp, C.c_uint64, C.POINTER(i32), C.POINTER(i32)]
    lib.qo_normalization.restype = i32
    lib.qo_normalization.argtypes = [vp, C.c_uint64]
    lib.qo_hv_real.argtypes = [vp, vp, vp, i64, i32, C.c_char_p, vp, vp, vp, vp, vp, i32]
    lib.qo_csr_spmv_c128.argtypes = [i64, vp, vp, vp, vp, vp]
    lib.qo_expm_c128.argtypes = [i64, vp, vp, vp, i32, i32, dbl, dbl, dbl, dbl, dbl, vp, vp]
    lib.qo_table_build.restype = i64
    lib.qo_table_build.argtypes = [vp, i32, vp, i64, vp]
    lib.qo_hv_table_real.argtypes = [vp, vp, vp, i64, vp, i32, C.c_char_p, vp, vp, vp, vp, vp]
    lib.qo_rcsr_1_norm.restype = dbl
    lib.qo_rcsr_1_norm.argtypes = [i64, vp, vp, i32, vp, vp, dbl, dbl, dbl, vp]
    lib.qo_expm_rcsr_c128.argtypes = [i64, vp, vp, i32, vp, vp, dbl, i32, i32, dbl, dbl, dbl, dbl, dbl, vp, vp]
    lib.qo_max_threads.restype = i32
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Group:
    def __init__(self, N, maps, pers, qs):
        lib = load()
        self.N = N
        self.maps = np.ascontiguousarray(np.asarray(maps, dtype=np.int32).reshape(-1, N) if len(pers) else
                                         np.zeros((0, N), np.int32))
        self.pers = np.ascontiguousarray(pers, dtype=np.int32)
        self.qs = np.ascontiguousarray(qs, dtype=np.int32)
        self.h = lib.qo_group_create(N, len(self.pers), _p(self.maps), _p(self.pers), _p(self.qs))
        if not self.h:
            raise ValueError("too many maps or sites for the C oracle")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qo_group_destroy(self.h)

    def make_basis(self, Np):
        lib = load()
        Ns = lib.qo_make_basis(self.h, Np, None, None, 0)
        b = np.empty(max(Ns, 1), np.uint64)
        n = np.empty(max(Ns, 1), np.int32)
        lib.qo_make_basis(self.h, Np, _p(b), _p(n), Ns)
        return b[:Ns], n[:Ns]

    def representative(self, s):
        code, cnt = C.c_int(), C.c_int()
        r = load().qo_representative(self.h, int(s), C.byref(code), C.byref(cnt))
        return r, code.value, cnt.value


def hv_real(group, basis, n, terms, vin, fused=True):
    """terms: [(opstr, sites, J)] with real J (spin operators, hcb)."""
    lib = load()
    ops = "".join(t[0] for t in terms).encode()
    lens = np.array([len(t[0]) for t in terms], np.int32)
    sites = np.array([s for t in terms for s in t[1]], np.int32)
    J = np.array([float(np.real(t[2])) for t in terms], np.float64)
    basis = np.ascontiguousarray(basis, np.uint64)
    n = np.ascontiguousarray(n, np.int32)
    vin = np.ascontiguousarray(vin, np.float64)
    vout = np.empty_like(vin)
    lib.qo_hv_real(group.h, _p(basis), _p(n), basis.size, len(terms), ops, _p(lens), _p(sites), _p(J), _p(vin),
                   _p(vout), 1 if fused else 0)
    return vout


def table_build(group, Np, basis):
    """Sector table of the Np-particle sector (8-byte entries: basis index of
    the representative, element code, element count), as a uint64 array."""
    lib = load()
    basis = np.ascontiguousarray(basis, np.uint64)
    from math import comb
    tab = np.empty(comb(group.N, Np), np.uint64)
    n = lib.qo_table_build(group.h, Np, _p(basis), basis.size, _p(tab))
    if n != tab.size:
        raise ValueError("qo_table_build: unsupported sizes")
    return tab


def hv_table_real(group, basis, n, tab, terms, vin):
    """Fused real H.v through the sector table of ``table_build``."""
    lib = load()
    ops = "".join(t[0] for t in terms).encode()
    lens = np.array([len(t[0]) for t in terms], np.int32)
    sites = np.array([s for t in terms for s in t[1]], np.int32)
    J = np.array([float(np.real(t[2])) for t in terms], np.float64)
    basis = np.ascontiguousarray(basis, np.uint64)
    n = np.ascontiguousarray(n, np.int32)
    vin = np.ascontiguousarray(vin, np.float64)
    vout = np.empty_like(vin)
    lib.qo_hv_table_real(group.h, _p(basis), _p(n), basis.size, _p(tab), len(terms), ops, _p(lens), _p(sites), _p(J),
                         _p(vin), _p(vout))
    return vout


def expm_c128(A, v, s, m_star, tol, mu, a):
    """Taylor loop on a complex128 CSR; returns e^{aA} v."""
    lib = load()
    A = A.tocsr().astype(np.complex128)
    ip = np.ascontiguousarray(A.indptr, np.int64)
    ix = np.ascontiguousarray(A.indices, np.int64)
    d = np.ascontiguousarray(A.data)
    F = np.array(v, dtype=np.complex128, copy=True)
    work = np.empty(2 * F.size, np.complex128)
    mu, a = complex(mu), complex(a)
    lib.qo_expm_c128(A.shape[0], _p(ip), _p(ix), _p(d), s, m_star, float(tol), mu.real, mu.imag, a.real, a.imag,
                     _p(F), _p(work))
    return F


class RealCSR:
    """A real CSR (numpy index arrays of int32 or int64, float64 values) plus
    an optional diagonal D, for the real-matrix / complex-state Taylor loop."""

    def __init__(self, indptr, indices, data, D=None):
        self.indptr = np.ascontiguousarray(indptr)
        self.indices = np.ascontiguousarray(indices)
        if self.indptr.dtype != self.indices.dtype or self.indptr.dtype not in (np.int32, np.int64):
            raise TypeError("indptr and indices must both be int32 or both int64")
        self.data = np.ascontiguousarray(data, np.float64)
        self.D = None if D is None else np.ascontiguousarray(D, np.float64)
        self.Ns = self.indptr.size - 1
        self.idx64 = 1 if self.indptr.dtype == np.int64 else 0
        self._colsum = None

    def _args(self):
        return (self.Ns, _p(self.indptr), _p(self.indices), self.idx64, _p(self.data),
                None if self.D is None else _p(self.D))

    def one_norm(self, dscale=0.0, mu=0.0):
        """||A + dscale D - mu I||_1 (max column sum of absolute values)."""
        if self._colsum is None:
            self._colsum = np.empty(max(self.Ns, 1), np.float64)
        mu = complex(mu)
        return load().qo_rcsr_1_norm(*self._args(), float(dscale), mu.real, mu.imag, _p(self._colsum))

    def expm(self, v, s, m_star, tol, mu, a, dscale=0.0):
        """e^{a (A + dscale D)} v by the Taylor loop with shift mu."""
        F = np.array(v, dtype=np.complex128, copy=True)
        work = np.empty(2 * F.size, np.complex128)
        mu, a = complex(mu), complex(a)
        load().qo_expm_rcsr_c128(*self._args(), float(dscale), int(s), int(m_star), float(tol), mu.real, mu.imag,
                                 a.real, a.imag, _p(F), _p(work))
        return F


def max_threads():
    return load().qo_max_threads()
