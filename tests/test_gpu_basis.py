"""Device basis construction, representatives and normalisations vs the
oracle (bit-exact integers)."""

import itertools

import numpy as np
import pytest

from oracle import qsx_oracle as O
from tests.models import chain_maps, ladder_maps

pytestmark = pytest.mark.gpu


def _oracle_basis(N, blocks, Np, fermion=False, spinful_L=None, double_occupancy=True):
    items = O.order_maps(blocks)
    g = O.Group(N, [it[1] for it in items], [it[2] for it in items], [it[3] for it in items], fermion=fermion)
    b, n = O.make_basis(N, g, Np=Np, spinful_L=spinful_L, double_occupancy=double_occupancy)
    return g, b, n


CHAIN_CASES = []
for L in (8, 10, 12):
    m = chain_maps(L)
    for k in (0, 1, L // 2):
        for p in (None, 0, 1):
            for z in (None, 0, 1):
                if p is not None and k not in (0, L // 2):
                    continue
                blocks = {"kblock": (m["T"], k)}
                if p is not None:
                    blocks["pblock"] = (m["P"], p)
                if z is not None:
                    blocks["zblock"] = (m["Z"], z)
                CHAIN_CASES.append((L, blocks))


@pytest.mark.parametrize("L,blocks", CHAIN_CASES[::3])
def test_chain_basis_bit_exact(L, blocks):
    from quspin_amd.basis import spin_basis_general
    print(f"case L={L} blocks={ {k: v[1] for k, v in blocks.items()} }", flush=True)
    g, b, n = _oracle_basis(L, blocks, L // 2)
    B = spin_basis_general(L, Nup=L // 2, pauli=0, **blocks)
    assert B.Ns == len(b)
    np.testing.assert_array_equal(B.states, b.astype(np.uint32))
    np.testing.assert_array_equal(B._n.astype(np.int64), n)


def test_ladder_basis_and_representatives():
    from quspin_amd.basis import spin_basis_general
    m = ladder_maps(3, 2)
    blocks = {"tx": (m["tx"], 0), "ty": (m["ty"], 0), "px": (m["px"], 0), "py": (m["py"], 0), "z": (m["z"], 0)}
    g, b, n = _oracle_basis(6, blocks, 3)
    B = spin_basis_general(6, m=0.0, pauli=False, **blocks)
    np.testing.assert_array_equal(B.states, b.astype(np.uint32))
    np.testing.assert_array_equal(B._n.astype(np.int64), n)
    states = np.arange(64, dtype=np.uint32)
    r, gg, sg = B.representative(states, return_g=True, return_sign=True)
    items = O.order_maps(blocks)
    go = O.Group(6, [it[1] for it in items], [it[2] for it in items], [it[3] for it in items])
    r_o, g_o, s_o = go.representative(states)
    np.testing.assert_array_equal(r, r_o.astype(np.uint32))
    np.testing.assert_array_equal(gg, g_o)
    np.testing.assert_array_equal(B.normalization(states).astype(np.int64), go.normalization(states))


def test_no_symmetry_and_multi_sector():
    from quspin_amd.basis import spin_basis_general
    B = spin_basis_general(10, Nup=[3, 5, 7], pauli=0)
    full = O.full_basis_states(10, [3, 5, 7])
    np.testing.assert_array_equal(B.states, full.astype(np.uint32))
    B2 = spin_basis_general(9, pauli=0)
    np.testing.assert_array_equal(B2.states, np.arange(511, -1, -1, dtype=np.uint32))


def test_spinful_order_and_printout():
    """FHM.ipynb cell 3: L=4, Nf=(2,2) basis order |up>|down>, descending."""
    from quspin_amd.basis import spinful_fermion_basis_general
    B = spinful_fermion_basis_general(4, Nf=(2, 2))
    assert B.Ns == 36
    assert B.int_to_state(B.states[0]) == "|1 1 0 0>|1 1 0 0>"
    assert B.int_to_state(B.states[1]) == "|1 1 0 0>|1 0 1 0>"
    assert B.int_to_state(B.states[6]) == "|1 0 1 0>|1 1 0 0>"
    assert B.int_to_state(B.states[35]) == "|0 0 1 1>|0 0 1 1>"
    full = O.full_basis_states(8, [(2, 2)], spinful_L=4)
    np.testing.assert_array_equal(B.states, full.astype(np.uint32))


def test_spinful_symmetric_and_no_double_occupancy():
    from quspin_amd.basis import spinful_fermion_basis_general
    L = 6
    T = (np.arange(L) + 1) % L
    for k in range(L):
        B = spinful_fermion_basis_general(L, Nf=(3, 2), kblock=(T, k))
        adv = np.hstack([T, T + L])
        g, b, n = _oracle_basis(2 * L, {"kblock": (adv, k)}, [(3, 2)], fermion=True, spinful_L=L)
        np.testing.assert_array_equal(B.states, b.astype(np.uint32))
        np.testing.assert_array_equal(B._n.astype(np.int64), n)
    B = spinful_fermion_basis_general(5, Nf=(2, 2), double_occupancy=False)
    full = O.full_basis_states(10, [(2, 2)], spinful_L=5, double_occupancy=False)
    np.testing.assert_array_equal(B.states, full.astype(np.uint32))


def test_sector_sizes_sum_to_pcon_size():
    """test_sq_lat_Heis.py:28-32: k-sector sizes add up to C(L, L/2)."""
    from quspin_amd.basis import spin_basis_general
    from scipy.special import comb
    L = 16
    T = (np.arange(L) + 1) % L
    tot = sum(spin_basis_general(L, Nup=L // 2, kblock=(T, k)).Ns for k in range(L))
    assert tot == comb(L, L // 2, exact=True)


def test_C2_sector_size_closed_form():
    """Heisenberg L=28, k=0, p=0, z=0: 361,270 states (character formula)."""
    from quspin_amd.basis import spin_basis_general
    m = chain_maps(28)
    B = spin_basis_general(28, Nup=14, pauli=0, kblock=(m["T"], 0), pblock=(m["P"], 0), zblock=(m["Z"], 0))
    assert B.Ns == 361270
    st = B.states
    assert np.all(np.diff(st.astype(np.int64)) < 0)
    # every stored state is its own representative with the stored n
    idx = np.random.default_rng(0).choice(B.Ns, 2000, replace=False)
    r = B.representative(st[idx])
    np.testing.assert_array_equal(r, st[idx])
    np.testing.assert_array_equal(B.normalization(st[idx]).astype(np.int64), B._n[idx].astype(np.int64))


def test_blocks_bracket_prefixes():
    from quspin_amd.basis import spin_basis_general
    m = chain_maps(14)
    B = spin_basis_general(14, Nup=7, kblock=(m["T"], 0))
    st = B.states.astype(np.int64)
    beg, end = B._basis_begin, B._basis_end
    sh = 14 - B._N_p
    for p in range(1 << B._N_p):
        sel = st[beg[p]:end[p]]
        assert np.all((sel >> sh) == p)
        assert np.sum((st >> sh) == p) == end[p] - beg[p]


def test_Ns_block_est_overflow():
    """make() raises when the size estimate is too low (base_general.py:849-872)."""
    from quspin_amd.basis import spin_basis_general
    m = chain_maps(20)
    with pytest.raises(ValueError):
        spin_basis_general(20, Nup=10, kblock=(m["T"], 0), Ns_block_est=1001)


def test_sector_rank_all_modes():
    """Device combinadic rank in every sector mode (hcb sector, spinful pair,
    full space) against direct enumeration; regression test for the
    full-space branch."""
    from quspin_amd import _lib as L
    from quspin_amd.basis import spin_basis_general, spinful_fermion_basis_general
    cases = [(spin_basis_general(8, Nup=3), 8, lambda s: bin(s).count("1") == 3),
             (spin_basis_general(7), 7, lambda s: True),
             (spinful_fermion_basis_general(4, Nf=(2, 1)), 8,
              lambda s: bin(s >> 4).count("1") == 2 and bin(s & 15).count("1") == 1)]
    for B, N, member in cases:
        allst = np.arange(1 << N, dtype=np.uint32)
        sd = L.to_device(allst)
        out = L.empty(allst.size, np.int64)
        info = np.zeros(8, np.int64)
        L.check(L.load().qsx_sector_rank(B._core.h, L.ptr(sd), allst.size, L.ptr(out),
                                         info.ctypes.data_as(L._vp), L.stream_handle()))
        got = L.to_host(out)
        members = [s for s in range(1 << N) if member(s)]
        expect = np.full(1 << N, -1, np.int64)
        expect[members] = np.arange(len(members))
        np.testing.assert_array_equal(got, expect)
