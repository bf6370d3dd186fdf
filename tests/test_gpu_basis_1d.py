"""spin_basis_1d on the device core against the oracle's general basis.

The reference's own test (test/test_general_spin_opstr.py:58-135,166-250)
asserts that ``spin_basis_1d`` and ``spin_basis_general`` with the maps
t = (i+1) % L, p = L-1-i, z = -(i+1) agree in states, normalisations and
every ``Op`` matrix element at L = 6, for every momentum, parity at k = 0 and
k = L/2, and spin inversion.  These tests run the same grid through this
build's ``spin_basis_1d`` and compare with the oracle's general basis built
from those maps (1D block b = +-1 <-> quantum number q with (-1)^q = b)."""

import itertools

import numpy as np
import pytest

from oracle import qsx_oracle as O

pytestmark = pytest.mark.gpu

L6 = 6
_T = np.array([(i + 1) % L6 for i in range(L6)])
_P = np.array([L6 - i - 1 for i in range(L6)])
_Z = np.array([-(i + 1) for i in range(L6)])


def _grid():
    for Nup, k, p, z in itertools.product([None, 3], [None] + list(range(L6)), [None, 0, 1], [None, 0, 1]):
        if p is not None and k not in (0, L6 // 2):
            continue
        yield Nup, k, p, z


@pytest.mark.parametrize("Nup", [None, 3])
def test_spin_basis_1d_equals_general_convention(Nup):
    from quspin_amd.basis import spin_basis_1d
    ops = ["x", "y", "z", "+", "-", "I"]
    for Nup_, k, p, z in _grid():
        if Nup_ != Nup:
            continue
        blocks, maps, pers, qs = {}, [], [], []
        if p is not None:
            blocks["pblock"] = (-1) ** p
            maps.append(_P); pers.append(2); qs.append(p)
        if z is not None:
            blocks["zblock"] = (-1) ** z
            maps.append(_Z); pers.append(2); qs.append(z)
        if k is not None:
            blocks["kblock"] = k
            maps.append(_T); pers.append(L6); qs.append(k)
        print(f"case Nup={Nup} blocks={blocks}", flush=True)
        B = spin_basis_1d(L6, Nup=Nup, pauli=0, **blocks)
        g = O.Group(L6, maps, pers, qs)
        ob, on = O.make_basis(L6, g, Np=Nup)
        assert B.Ns == ob.size
        np.testing.assert_array_equal(B.states.astype(np.uint64), ob.astype(np.uint64))
        if len(pers):
            np.testing.assert_array_equal(B._n.astype(np.int64), on.astype(np.int64))
        if B.Ns == 0:
            continue
        for length in (1, 2):
            for i0 in range(L6 - length + 1):
                indx = list(range(i0, i0 + length))
                for opstr in itertools.product(ops, repeat=length):
                    opstr = "".join(opstr)
                    me, row, col = B.Op(opstr, indx, 1.0, np.complex128)
                    ome, orow, ocol = O.op_coo(ob, on, g, opstr, indx, 1.0)
                    a = np.lexsort((row, col))
                    b = np.lexsort((orow, ocol))
                    np.testing.assert_array_equal(col[a], ocol[b])
                    np.testing.assert_array_equal(row[a], orow[b])
                    np.testing.assert_allclose(me[a], ome[b], atol=1e-12, rtol=0, err_msg=f"{opstr} {indx}")


def test_spin_basis_1d_sublattice_inversion():
    """zAblock / zBblock: spin inversion of the even / odd sites as general
    maps (no magnetisation sector, as the reference requires), with
    translation by a = 2: states against the oracle's general basis, and the
    sectors' spectra together are the full spectrum (the 1D core's own gauge
    for these blocks is not pinned by any reference output)."""
    from quspin_amd.basis import spin_basis_1d
    from quspin_amd.operators import hamiltonian
    L = 8
    # invariant under each sublattice inversion: x on any site, z z and y y
    # only within one sublattice (next-nearest neighbours)
    static = [["xx", [[1.0, i, (i + 1) % L] for i in range(L)]], ["zz", [[0.7, i, (i + 2) % L] for i in range(L)]],
              ["yy", [[0.4, i, (i + 2) % L] for i in range(L)]], ["x", [[0.3, i] for i in range(L)]]]
    kw = dict(check_symm=False, check_herm=False, check_pcon=False)
    full = np.linalg.eigvalsh(hamiltonian(static, [], basis=spin_basis_1d(L, pauli=0), dtype=np.float64,
                                          **kw).toarray())
    es = []
    for zA, zB in itertools.product((1, -1), (1, -1)):
        for k in range(L // 2):
            print(f"case zA={zA} zB={zB} k={k}", flush=True)
            B = spin_basis_1d(L, pauli=0, zAblock=zA, zBblock=zB, kblock=k, a=2)
            maps = [np.where(np.arange(L) % 2 == 0, -(np.arange(L) + 1), np.arange(L)),
                    np.where(np.arange(L) % 2 == 1, -(np.arange(L) + 1), np.arange(L)),
                    (np.arange(L) + 2) % L]
            g = O.Group(L, maps, [2, 2, L // 2], [(1 - zA) // 2, (1 - zB) // 2, k])
            ob, on = O.make_basis(L, g)
            np.testing.assert_array_equal(B.states.astype(np.uint64), ob.astype(np.uint64))
            if B.Ns == 0:
                continue
            H = hamiltonian(static, [], basis=B, dtype=np.complex128, **kw).toarray()
            es.append(np.linalg.eigvalsh(H))
    np.testing.assert_allclose(np.sort(np.concatenate(es)), full, atol=1e-10)
