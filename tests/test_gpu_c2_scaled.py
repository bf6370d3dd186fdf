"""The 64-bit-state sector-table path that the Heisenberg chain at L=34
(Sz=0, k=0, p=+1, z=+1; 17,179,435 states) runs on: bit-exact bases and
products against the oracle on 34-site sectors small enough for it, then the
full sector, where the table-based matrix-free product must agree with the
device CSR assembled through the orbit search (an independent lookup)."""

import itertools

import numpy as np
import pytest

from oracle import qsx_oracle as O
from tests.models import chain_maps, flat_terms, xxz_static

pytestmark = pytest.mark.gpu

KW = dict(check_symm=False, check_herm=False, check_pcon=False)


def _oracle_sector(L, Nup, blocks):
    cand = np.array(sorted((sum(1 << b for b in c) for c in itertools.combinations(range(L), Nup)), reverse=True),
                    dtype=np.uint64)
    items = O.order_maps(blocks)
    g = O.Group(L, [it[1] for it in items], [it[2] for it in items], [it[3] for it in items])
    rep, _, _ = g.representative(cand)
    c = cand[rep == cand]
    n = g.normalization(c)
    return g, c[n > 0], n[n > 0]


@pytest.mark.parametrize("schedule", ["rows", "batch4"])
@pytest.mark.parametrize("entry_bytes", [None, 8])
@pytest.mark.parametrize("k,p,dtype", [(0, 0, np.float64), (5, None, np.complex128)])
def test_u64_table_sectors_vs_oracle(k, p, dtype, entry_bytes, schedule, monkeypatch):
    """L=34 with four up spins: 64-bit state words, sector table with 4-byte
    (default) or 8-byte entries; the coefficient table has one row per
    distinct character (1 at k=0, p=+1; 34 at k=5) and distinct n."""
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    monkeypatch.setenv("QSX_LOOKUP", "table")
    monkeypatch.setenv("QSX_SCHEDULE", schedule)
    L, Nup = 34, 4
    m = chain_maps(L)
    blocks = {"kblock": (m["T"], k)}
    if p is not None:
        blocks["pblock"] = (m["P"], p)
    print(f"case L={L} Nup={Nup} k={k} p={p} entry_bytes={entry_bytes} schedule={schedule}", flush=True)
    B = spin_basis_general(L, Nup=Nup, pauli=0, **blocks)
    g, b, n = _oracle_sector(L, Nup, blocks)
    assert B.states.dtype == np.uint64
    np.testing.assert_array_equal(B.states, b)
    tab = B.lookup_table(entry_bytes=entry_bytes)
    assert tab[1] == (entry_bytes or 4)
    nchi = 1 if k == 0 else L
    assert tab[3] is not None and tab[3].numel() == nchi * len(np.unique(n))
    static = xxz_static(L, Delta=0.8)
    Q = quantum_LinearOperator(static, basis=B, dtype=dtype, **KW)
    H = O.make_csr(b, n.astype(np.int64), g, flat_terms(static), np.complex128)
    rng = np.random.default_rng(21)
    v = rng.normal(size=B.Ns) + (1j * rng.normal(size=B.Ns) if np.dtype(dtype).kind == "c" else 0)
    ref = H @ v.astype(np.complex128)
    np.testing.assert_allclose(Q.dot(v), ref if np.dtype(dtype).kind == "c" else ref.real, atol=1e-12, rtol=0)
    assert all(c["lookup"] == "table" for c in Q._qsx_operator().config.values())


def test_heisenberg_L34_kpz_full_sector():
    """The full sector: 17,179,435 states (character formula), 4-byte table
    entries with a one-row-per-n coefficient table, and the table-based
    matrix-free product equal to the device CSR product (orbit search) and
    symmetric (<u|Hv> = <Hu|v>)."""
    import torch
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian, quantum_LinearOperator
    L = 34
    m = chain_maps(L)
    print(f"case L={L} Nup=17 k=0 p=+1 z=+1", flush=True)
    B = spin_basis_general(L, Nup=L // 2, pauli=0, kblock=(m["T"], 0), pblock=(m["P"], 0), zblock=(m["Z"], 0))
    assert B.Ns == 17179435
    tab = B.lookup_table()
    assert tab[1] == 4 and tab[3] is not None and tab[3].numel() == len(torch.unique(B._dev["n"].to(torch.int32)))
    static = xxz_static(L, Delta=1.0)
    Q = quantum_LinearOperator(static, basis=B, dtype=np.float64, **KW)
    gen = torch.Generator(device="cuda").manual_seed(34)
    u = torch.randn(B.Ns, dtype=torch.float64, device="cuda", generator=gen)
    v = torch.randn(B.Ns, dtype=torch.float64, device="cuda", generator=gen)
    Hv = Q.dot(v)
    Hu = Q.dot(u)
    assert all(c["lookup"] == "table" for c in Q._qsx_operator().config.values())
    lhs, rhs = float(torch.dot(u, Hv)), float(torch.dot(Hu, v))
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)
    del Hu, u
    H = hamiltonian(static, [], basis=B, dtype=np.float64, **KW)
    Hv2 = H.dot(v)
    err = float((Hv - Hv2).abs().max())
    print(f"nnz={H._static.nnz} max|table-csr|={err:.3e}", flush=True)
    assert err <= 1e-12 * float(Hv.abs().max())
