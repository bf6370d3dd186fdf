"""Operators between symmetry sectors and projections to and from the full
space on the device, against the oracle's dense projector
(test_Op_shift_sector.py, test_get_amp.py, test_project_to.py,
test_general_spin_get_vec.py of the reference)."""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import qsx_oracle as O
from tests.models import chain_maps, flat_terms

pytestmark = pytest.mark.gpu


def _group(basis, fermion=False):
    return O.Group(basis.N, list(basis._maps) if len(basis._pers) else [], list(basis._pers), list(basis._qs),
                   fermion=fermion)


def _bn(basis):
    b = basis.states.astype(np.uint64)
    n = basis._n.astype(np.int64) if len(basis._pers) else np.ones(len(b), np.int64)
    return b, n


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=n) + 1j * rng.normal(size=n)
    return v / np.linalg.norm(v)


def test_op_shift_sector_momentum_parity_and_particle_number():
    """test_Op_shift_sector.py:10-121 at L=10: z_q = sum_j e^{-2 pi i q j/L}
    z_j from every k-sector to k+q; parity-odd z pairs between parity sectors;
    sum_j S^+_j (Nup -> Nup+1) and sum_j S^-_j (Nup -> Nup-1); every result
    against P_2^H Op P_1 v of the oracle, atol 1e-13 (as the reference)."""
    from quspin_amd.basis import spin_basis_general
    L = 10
    t = (np.arange(L) + 1) % L
    p = np.arange(L)[::-1].copy()
    full = O.full_basis_states(L)
    for k in range(L):
        for q in (1, 4):
            op_list = [["z", [i], np.exp(-2j * np.pi * q * i / L)] for i in range(L)]
            b1 = spin_basis_general(L, kblock=(t, k), pauli=0)
            b2 = spin_basis_general(L, kblock=(t, (k + q) % L), pauli=0)
            v = _rand(b1.Ns, 10 * k + q)
            got = b2.Op_shift_sector(b1, op_list, v)
            B1, n1 = _bn(b1)
            B2, n2 = _bn(b2)
            ref = O.op_shift_sector(B2, n2, _group(b2), B1, n1, _group(b1), full, L,
                                    [(o, tuple(i), J) for o, i, J in op_list], v)
            print(f"case k={k} q={q} Ns {b1.Ns}->{b2.Ns} err {np.abs(got - ref).max():.1e}", flush=True)
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13)
    for q1 in (0, 1):
        for q2 in (0, 1):
            i = 1
            op_list = [["z", [i], 1], ["z", [L - i - 1], (-1) ** q2]]
            b1 = spin_basis_general(L, block=(p, q1), pauli=0)
            b2 = spin_basis_general(L, block=(p, (q1 + q2) % 2), pauli=0)
            v = _rand(b1.Ns, q1 + 2 * q2)
            got = b2.Op_shift_sector(b1, op_list, v)
            B1, n1 = _bn(b1)
            B2, n2 = _bn(b2)
            ref = O.op_shift_sector(B2, n2, _group(b2), B1, n1, _group(b1), full, L,
                                    [(o, tuple(i), J) for o, i, J in op_list], v)
            print(f"case parity q1={q1} q2={q2}", flush=True)
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13)
    for Nup, ch, dN in [(n, "+", 1) for n in range(0, L)] + [(n, "-", -1) for n in range(1, L + 1)]:
        op_list = [[ch, [i], 1.0] for i in range(L)]
        b1 = spin_basis_general(L, Nup=Nup, pauli=0)
        b2 = spin_basis_general(L, Nup=Nup + dN, pauli=0)
        v = _rand(b1.Ns, Nup)
        got = b2.Op_shift_sector(b1, op_list, v)
        B1, n1 = _bn(b1)
        B2, n2 = _bn(b2)
        ref = O.op_shift_sector(B2, n2, _group(b2), B1, n1, _group(b1), full, L,
                                [(o, tuple(i), J) for o, i, J in op_list], v)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13)


def test_op_shift_sector_fermions_and_several_vectors():
    """A spinless-fermion annihilator with momentum q, c_q = sum_j
    e^{-2 pi i q j/L} c_j, from (Nf, k) to (Nf - 1, k - q) -- fermion signs of
    the translations on both sides -- for three vectors at once and with a
    device v_out accumulated into, against the oracle."""
    import torch
    from quspin_amd import _lib as Lb
    from quspin_amd.basis import spinless_fermion_basis_general
    L, Nf, k, q = 8, 3, 2, 3
    t = (np.arange(L) + 1) % L
    full = O.full_basis_states(L)
    op_list = [["-", [j], np.exp(-2j * np.pi * q * j / L)] for j in range(L)]
    b1 = spinless_fermion_basis_general(L, Nf=Nf, kblock=(t, k))
    hit = []
    for k2 in range(L):
        b2 = spinless_fermion_basis_general(L, Nf=Nf - 1, kblock=(t, k2))
        rng = np.random.default_rng(k2)
        V = rng.normal(size=(b1.Ns, 3)) + 1j * rng.normal(size=(b1.Ns, 3))
        W0 = rng.normal(size=(b2.Ns, 3)) + 1j * rng.normal(size=(b2.Ns, 3))
        W = Lb.to_device(W0.copy())
        b2.Op_shift_sector(b1, op_list, Lb.to_device(V), v_out=W)
        B1, n1 = _bn(b1)
        B2, n2 = _bn(b2)
        ref = O.op_shift_sector(B2, n2, _group(b2, True), B1, n1, _group(b1, True), full, L,
                                [(o, tuple(i), J) for o, i, J in op_list], V, fermion=True)
        print(f"case fermion k={k} -> k2={k2}: |ref| {np.abs(ref).max():.2e}", flush=True)
        np.testing.assert_allclose(Lb.to_host(W), W0 + ref, rtol=0, atol=1e-13)
        if np.abs(ref).max() > 1e-12:
            hit.append(k2)
    # the operator carries momentum: exactly one target sector receives it
    assert len(hit) == 1


def _chain_sectors(L):
    m = chain_maps(L)
    return [dict(), dict(Nup=L // 2), dict(zblock=(m["Z"], 1)), dict(pblock=(m["P"], 1), Nup=L // 2),
            dict(Nup=L // 2, kblock=(m["T"], 0), pblock=(m["P"], 0), zblock=(m["Z"], 0)),
            dict(Nup=3, kblock=(m["T"], L - 1)), dict(kblock=(m["T"], 2), zblock=(m["Z"], 0)),
            dict(Nup=1, kblock=(m["T"], L // 2), pblock=(m["P"], 0))]


def test_get_proj_and_projections_vs_oracle():
    """get_proj (full space and particle sector), project_from / get_vec and
    project_to, dense and sparse, host and device vectors, on the sectors of
    test_project_to.py:38-67 (L=10 here), against the oracle's dense
    projector: P, P v, P^H v_full (dense project_to), P^T v_full (the
    reference's sparse project_to, base_general.py:671)."""
    import torch
    from quspin_amd import _lib as Lb
    from quspin_amd.basis import spin_basis_general
    L = 10
    for kw in _chain_sectors(L):
        B = spin_basis_general(L, pauli=0, **kw)
        b, n = _bn(B)
        g = _group(B)
        pcons = (False, True) if "Nup" in kw else (False,)
        for pcon in pcons:
            full = O.full_basis_states(L, kw["Nup"] if pcon else None)
            Pref = O.projector(b, n, g, full)
            dt = np.complex128 if np.abs(Pref.imag).max(initial=0) > 0 else np.float64
            P = B.get_proj(dt, pcon=pcon)
            print(f"case {sorted(kw)} pcon={pcon} Ns={B.Ns} full={full.size} nnz={P.nnz}", flush=True)
            assert sp.isspmatrix_csc(P) and P.shape == Pref.shape
            np.testing.assert_allclose(P.toarray(), Pref, rtol=0, atol=1e-14)
            np.testing.assert_allclose((Pref.conj().T @ Pref), np.eye(B.Ns), atol=1e-12)
            rng = np.random.default_rng(B.Ns)
            if B.Ns == 0:
                # an empty sector gives (n_out, 0) arrays (base_general.py:649-656, 759-766)
                assert B.project_from(np.zeros((0, 2)), sparse=False, pcon=pcon).shape == (full.size, 0)
                assert B.project_to(np.zeros((full.size, 2)), sparse=False, pcon=pcon).shape == (0, 0)
                continue
            v = (rng.normal(size=(B.Ns, 2)) + 1j * rng.normal(size=(B.Ns, 2))).astype(np.complex128)
            vf = (rng.normal(size=(full.size, 2)) + 1j * rng.normal(size=(full.size, 2))).astype(np.complex128)
            np.testing.assert_allclose(B.project_from(v, sparse=False, pcon=pcon), Pref @ v, atol=1e-13)
            np.testing.assert_allclose(B.get_vec(v[:, 0], sparse=False, pcon=pcon), Pref @ v[:, 0], atol=1e-13)
            np.testing.assert_allclose(B.project_to(vf, sparse=False, pcon=pcon), Pref.conj().T @ vf, atol=1e-13)
            np.testing.assert_allclose(B.project_from(v, sparse=True, pcon=pcon).toarray(), Pref @ v, atol=1e-13)
            np.testing.assert_allclose(B.project_to(vf, sparse=True, pcon=pcon).toarray(), Pref.T @ vf, atol=1e-13)
            got = B.project_from(torch.from_numpy(v).cuda(), sparse=False, pcon=pcon)
            np.testing.assert_allclose(Lb.to_host(got), Pref @ v, atol=1e-13)
            if dt == np.float64:
                vr = v.real.copy()
                np.testing.assert_allclose(B.project_from(vr, sparse=False, pcon=pcon), (Pref @ vr).real, atol=1e-13)


def test_get_amp_modes_vs_oracle():
    """get_amp (test_get_amp.py:127-150 on chains): 'representative' mode on
    the basis states rescales reduced amplitudes to full-basis ones equal to
    get_vec's rows; 'full_basis' mode on every state of the particle sector
    gives P[s, rep(s)], replaces the states by their representatives, and
    divides full amplitudes back to the reduced ones."""
    from quspin_amd.basis import spin_basis_general
    L = 10
    m = chain_maps(L)
    for kw in (dict(kblock=(m["T"], 3)), dict(kblock=(m["T"], 0), pblock=(m["P"], 1), zblock=(m["Z"], 0)),
               dict(kblock=(m["T"], 5), zblock=(m["Z"], 1))):
        B = spin_basis_general(L, Nup=L // 2, pauli=0, **kw)
        Bf = spin_basis_general(L, Nup=L // 2, pauli=0)
        b, n = _bn(B)
        g = _group(B)
        full = O.full_basis_states(L, L // 2)
        print(f"case {sorted(kw)} Ns={B.Ns}", flush=True)
        psi = _rand(B.Ns, 1)
        states = B.states.copy()
        amps = psi.copy()
        B.get_amp(states, amps=amps, mode="representative")
        inds = O.index_of(full, states.astype(np.uint64))
        psi_full = B.get_vec(psi, pcon=True, sparse=False)
        np.testing.assert_allclose(amps, psi_full[inds], rtol=0, atol=1e-14)
        sf = Bf.states.copy()
        C = B.get_amp(sf, mode="full_basis")
        Cref, rep = O.get_amp(b, n, g, full, Bf.states.astype(np.uint64))
        np.testing.assert_allclose(C, Cref, rtol=0, atol=1e-14)
        np.testing.assert_array_equal(sf.astype(np.uint64), rep)
        sf2 = Bf.states.copy()
        amps_full = psi_full.copy()
        ok = np.abs(Cref) > 0
        B.get_amp(sf2, amps=amps_full, mode="full_basis")
        cols = O.index_of(b, rep)
        np.testing.assert_allclose(amps_full[ok], psi[cols[ok]], rtol=0, atol=1e-13)
