"""CPU-side checks (no GPU needed): the C-ABI library exports exactly what
include/quspin_hip.h declares, the ctypes signature table covers it, and the
host logic that mirrors the reference (map processing and ordering, term
consolidation, tuning rules, error mapping) behaves like the reference.
No compute call is made here."""

import ctypes
import os
import re
import subprocess
import sys
import warnings

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "quspin_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return set(re.findall(r"\b(qsx_[a-z0-9_]+)\s*\(", text))


@pytest.fixture(scope="module")
def libpath():
    from quspin_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from quspin_amd import _build
        _build.build()
    return _lib.LIB_PATH


def test_header_symbols_exported(libpath):
    out = subprocess.run(["nm", "-D", "--defined-only", libpath], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    declared = _declared()
    assert len(declared) >= 30
    assert declared - exported == set()
    # every extern "C" entry point of the library is declared in the header
    c_exports = {s for s in exported if s.startswith("qsx_")}
    assert c_exports - declared == set()


def test_library_loads_and_signatures_cover_header(libpath):
    from quspin_amd import _lib
    lib = _lib.load()  # dlopen + ctypes prototypes; no HIP call is made
    assert set(_lib.SIGNATURES) == _declared()
    for name in _lib.SIGNATURES:
        assert isinstance(getattr(lib, name), ctypes._CFuncPtr)
    assert isinstance(_lib.last_error(), str)


def test_product_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from quspin_amd.basis import spin_basis_general
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        spin_basis_general(8, Nup=4)


def test_status_codes_map_to_reference_exceptions(libpath):
    from quspin_amd import _lib
    _lib.load()
    assert _lib.check(3) == 3
    for rc, exc in [(_lib.QSX_ERR_COMPLEX, TypeError), (_lib.QSX_ERR_ARG, ValueError),
                    (_lib.QSX_ERR_OVERFLOW, ValueError), (_lib.QSX_ERR_UNSUPPORTED, NotImplementedError)]:
        with pytest.raises(exc):
            _lib.check(rc)


def test_bad_handles_and_arguments_rejected_before_any_device_call(libpath):
    """Entry points check their arguments on the host and return the
    argument error (ValueError in Python) with a message before any HIP call;
    here: null contexts."""
    from quspin_amd import _lib
    lib = _lib.load()
    calls = [lambda: lib.qsx_hv_ctx_set_table(None, None, 0, 0, None, 0),
             lambda: lib.qsx_hv_ctx_set_schedule(None, 0),
             lambda: lib.qsx_hv_ctx_set_diag(None, None, 0),
             lambda: lib.qsx_hv_ctx_apply(None, None, None, 1.0, 0.0, 1.0, 0.0, 0, None)]
    for call in calls:
        rc = call()
        assert rc == _lib.QSX_ERR_ARG
        with pytest.raises(ValueError):
            _lib.check(rc)
        assert _lib.last_error()


def test_process_map_periods_and_errors():
    """Periods of translation, parity, inversion (closed forms); the product's
    map processing agrees with the oracle's independent one."""
    from oracle import qsx_oracle as O
    from quspin_amd.basis.base_general import GeneralBasisWarning, process_map
    L = 12
    i = np.arange(L)
    cases = {"T": ((i + 1) % L, L), "T3": ((i + 3) % L, 4), "P": (i[::-1], 2), "Z": (-(i + 1), 2),
             "PZ": (-(i[::-1] + 1), 2), "TZ": (-((i + 1) % L + 1), L)}
    for name, (m, per) in cases.items():
        assert process_map(m, 0)[1] == per, name
        assert O.process_map(m)[2] == per, name
    # half-chain inversion with shift: period 2 x 12 -> lcm
    m = np.where(i < 6, -((i + 1) % L + 1), (i + 1) % L)
    assert process_map(m, 0)[1] == O.process_map(m)[2]
    with pytest.raises(ValueError):
        process_map([0, 0, 1], 0)
    with pytest.warns(GeneralBasisWarning):
        process_map(i, 0)


class _HostOnly:
    """basis_general's constructor up to (not including) the device core."""

    @staticmethod
    def make(N, block_order=None, **blocks):
        from quspin_amd.basis.base_general import basis_general

        class B(basis_general):
            pass
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            return B(N, block_order=block_order, **blocks)


def test_map_ordering_matches_oracle():
    """Ascending period (stable), block_order reversed, identity maps removed
    with the reference's ascending pops (base_general.py:135-203)."""
    from oracle import qsx_oracle as O
    L = 8
    i = np.arange(L)
    T, P, Z, E = (i + 1) % L, i[::-1], -(i + 1), i.copy()
    for blocks, order in [({"k": (T, 1), "p": (P, 0), "z": (Z, 1)}, None),
                          ({"k": (T, 3), "z": (Z, 0)}, ["k", "z"]),
                          ({"e1": (E, 0), "p": (P, 1), "e2": (E, 0), "k": (T, 2)}, None)]:
        B = _HostOnly.make(L, block_order=order, **blocks)
        ref = O.order_maps(blocks, block_order=order)
        assert B._pers.tolist() == [r[2] for r in ref]
        assert B._qs.tolist() == [r[3] for r in ref]
        np.testing.assert_array_equal(B._maps, np.vstack([r[1] for r in ref]))
    with pytest.raises(ValueError):
        _HostOnly.make(L, block_order=["k"], k=(T, 0), p=(P, 0))
    with pytest.raises(ValueError):
        _HostOnly.make(L + 1, k=(T, 0))


def test_consolidate_static_sums_and_drops():
    """_make_hamiltonian.py:19-41: repeated (opstr, indx) summed, |J| <= 10 eps dropped."""
    from quspin_amd.operators._make_hamiltonian import _consolidate_static
    out = _consolidate_static([["zz", [[1.0, 0, 1], [0.5, 0, 1], [2.0, 1, 2]]],
                               ["x", [[1e-17, 3], [1.0, 2]]], ["zz", [[-2.0, 1, 2]]]])
    d = {(o, i): J for o, i, J in out}
    assert d == {("zz", (0, 1)): 1.5, ("x", (2,)): 1.0}


def test_tuning_static_rule_and_candidates():
    from quspin_amd import tuning

    class _B:
        def __init__(self, pers, Np):
            self._pers = np.asarray(pers, np.int32)
            self._qs = np.zeros(len(pers), np.int32)
            self._Np, self.N, self.Ns = Np, 16, 1000

    class _OL:
        n_off, n_diag = 32, 16

    class _Op:
        def __init__(self, b):
            self.basis, self.oplist, self.transposed, self.conjugated = b, _OL(), False, False

    from quspin_amd.basis.base_general import hcb_basis_general

    class _HB(hcb_basis_general):
        def __init__(self, Np):
            self._pers = np.zeros(0, np.int32)
            self._qs = np.zeros(0, np.int32)
            self._Np, self._N, self._Ns = Np, 16, 1000

    sym = _Op(_B([16, 2], 8))
    free = _Op(_B([], 8))
    assert tuning.static_rule(sym, np.float64, 1) == {"lookup": "table", "schedule": "batch8"}
    assert tuning.static_rule(free, np.float64, 1) == {"lookup": "search", "schedule": "batch8"}
    # a spin / hard-core-boson sector without symmetry: the one-term-at-a-time kernel
    assert tuning.static_rule(_Op(_HB(8)), np.float64, 1) == {"lookup": "search", "schedule": "rows"}
    assert tuning.static_rule(sym, np.float64, 4)["schedule"] == "rows"
    names = {(c["lookup"], c["schedule"]) for c in tuning.candidates(sym, np.float64, 1)}
    assert ("table", "batch8") in names and ("search", "queue") in names and ("table", "batch14") in names
    assert ("table", "batch14x") in names
    assert all(c["schedule"] in tuning.SCHEDULES for c in tuning.candidates(sym, np.float64, 1))


def test_basis_word_types():
    from quspin_amd.basis.base_general import _get_index_type, get_basis_type
    assert get_basis_type(32, None, 2) is np.uint32
    assert get_basis_type(33, None, 2) is np.uint64
    assert get_basis_type(64, None, 2) is np.uint64
    with pytest.raises(NotImplementedError):
        get_basis_type(65, None, 2)
    assert _get_index_type(10) is np.int32
    assert _get_index_type(2 ** 31) is np.int64


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_spin_basis_1d_argument_checks():
    """spin_basis_1d refuses what the reference refuses (unexpected keys,
    spin inversion outside the zero-magnetisation sector, sublattice
    inversion with a magnetisation sector, parity with sublattice inversion)
    and the momentum-parity sectors at 0 < k < pi that the 1D core builds
    from sigma = +-1 real states; all before any device work."""
    from quspin_amd.basis import spin_basis_1d
    with pytest.raises(ValueError):
        spin_basis_1d(6, Nup=3, qblock=1)
    with pytest.raises(ValueError):
        spin_basis_1d(6, Nup=2, zblock=1)
    with pytest.raises(ValueError):
        spin_basis_1d(6, Nup=3, zAblock=1)
    with pytest.raises(ValueError):
        spin_basis_1d(6, pblock=1, zAblock=1)
    with pytest.raises(ValueError):
        spin_basis_1d(6, Nup=3, pblock=2)
    with pytest.raises(TypeError):
        spin_basis_1d(6, Nup=3, kblock=1.0)
    for blocks in ({"kblock": 1, "pblock": 1}, {"kblock": 2, "pzblock": -1}, {"kblock": 1, "a": 2, "pblock": 1},
                   {"kblock": -1, "pblock": 1}):
        with pytest.raises(NotImplementedError):
            spin_basis_1d(8, **blocks)
