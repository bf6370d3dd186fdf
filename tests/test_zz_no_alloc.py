"""The allocation-free Taylor loop (qsx_expm_multiply_csr_noalloc) against
the allocating one (qsx_expm_multiply_csr, the _wrapper_expm_multiply
replacement) through the C-ABI: two calls in a row, the second on a smaller
operator, each bit-identical to the allocating entry point (same kernels,
fixed reduction order); a partition whose control block exceeds the scratch
is rejected."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _case(N, seed, nvec):
    rng = np.random.RandomState(seed)
    A = sp.random(N, N, density=8.0 / N, random_state=rng, format="csr")
    A = (A + A.T).tocsr()
    A.sort_indices()
    V = rng.normal(size=(N, nvec)) + 1j * rng.normal(size=(N, nvec))
    return A, V


def _run(fn, A, V, s, m, a, mu, tol):
    from quspin_amd import _lib as L
    t = L.torch()
    N, nvec = V.shape
    ip = L.to_device(A.indptr.astype(np.int32))
    ix = L.to_device(A.indices.astype(np.int32))
    dt = L.to_device(A.data.astype(np.float64))
    v = L.to_device(V.astype(np.complex128))
    work = L.empty((2 * N, nvec), np.complex128)
    rc = fn(L.ptr(ip), L.ptr(ix), L.ptr(dt), 0, L.F64, N, A.nnz, s, m, tol, mu.real, mu.imag, a.real, a.imag,
            L.ptr(v), L.ptr(work), L.C128, nvec, 0, L.stream_handle())
    L.check(rc, "expm")
    t.cuda.synchronize()
    return L.to_host(v)


def test_noalloc_expm_equals_allocating_entry_point():
    from scipy.linalg import expm
    from quspin_amd import _lib as L
    lib = L.load()
    tol = 2.0 ** -53
    for N, seed, nvec, s, m, a in ((3000, 0, 3, 2, 24, -0.7j), (700, 1, 1, 1, 30, -0.3j)):
        A, V = _case(N, seed, nvec)
        mu = complex(A.diagonal().sum() / N)
        print(f"case N={N} nvec={nvec} s={s} m={m}", flush=True)
        ref = _run(lib.qsx_expm_multiply_csr, A, V, s, m, a, mu, tol)
        got = _run(lib.qsx_expm_multiply_csr_noalloc, A, V, s, m, a, mu, tol)
        np.testing.assert_array_equal(got, ref)
        if N <= 1000:
            np.testing.assert_allclose(got, expm(a * A.toarray()) @ V, atol=1e-12, rtol=0)
    cap = int(lib.qsx_expm_multiply_csr_noalloc_max_bytes())
    assert cap == 4 << 20
    A, V = _case(200, 2, 1)
    with pytest.raises(ValueError):
        _run(lib.qsx_expm_multiply_csr_noalloc, A, V, 4000, 55, -0.1j, 0j, tol)
