"""Batched sectors x vectors (one launch for all momentum sectors): H.V and
the lockstep Taylor loop, each sector against the oracle's reduced matrix
and against the single-sector path."""

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import qsx_oracle as O
from tests.models import chain_maps, flat_terms, hubbard_static, xxz_static

pytestmark = pytest.mark.gpu


def _sectors(L, ks, Delta=0.7, extra=None, dtype=np.complex128):
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    m = chain_maps(L)
    static = xxz_static(L, Delta=Delta)
    bases, ops, dense = [], [], []
    for k in ks:
        B = spin_basis_general(L, Nup=L // 2, pauli=0, kblock=(m["T"], k), **(extra or {}))
        Q = quantum_LinearOperator(static, basis=B, dtype=dtype, check_symm=False, check_herm=False,
                                   check_pcon=False)
        items = O.order_maps(dict({"kblock": (m["T"], k)}, **(extra or {})))
        g = O.Group(L, [it[1] for it in items], [it[2] for it in items], [it[3] for it in items])
        b, n = O.make_basis(L, g, Np=L // 2)
        assert np.array_equal(b.astype(np.uint32), B.states)
        dense.append(O.reduced_op(b, n, g, flat_terms(static)))
        bases.append(B)
        ops.append(Q)
    return bases, ops, dense


def _rand(rng, shape, dt):
    x = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    return x.astype(dt)


@pytest.mark.parametrize("form", ["mf", "csr"])
@pytest.mark.parametrize("nvec,dt,chunk", [(1, np.complex128, 16), (5, np.complex128, 3), (64, np.complex64, 16),
                                           (70, np.complex128, 32)])
def test_batched_apply_all_k_sectors(nvec, dt, chunk, form):
    from quspin_amd import _lib as L
    from quspin_amd.operators._devop import SectorBatch
    L_ = 12
    print(f"case L={L_} nvec={nvec} dtype={np.dtype(dt).name} chunk={chunk} form={form}", flush=True)
    bases, ops, dense = _sectors(L_, range(L_))
    mfs = [op._qsx_operator() for op in ops]
    Bt = SectorBatch(mfs, dt, nvec, form=form, chunk_rows=chunk)
    assert Bt.rows == sum(b.Ns for b in bases)
    rng = np.random.default_rng(5)
    xs = [_rand(rng, (b.Ns, nvec), dt) for b in bases]
    X = Bt.stack(xs)
    Y = Bt.empty()
    Bt.apply(X, Y, a=0.5 - 0.25j)
    got = Bt.unstack(Y)
    tol = 2e-5 if dt == np.complex64 else 1e-12
    for k in range(L_):
        ref = (0.5 - 0.25j) * dense[k] @ xs[k].astype(np.complex128)
        np.testing.assert_allclose(L.to_host(got[k]), ref, rtol=0, atol=tol * max(1.0, np.abs(ref).max()))
    # accumulate form
    Y2 = Y.clone()
    Bt.apply(X, Y2, a=1.0, beta=1)
    for k in range(L_):
        ref = (1.5 - 0.25j) * dense[k] @ xs[k].astype(np.complex128)
        np.testing.assert_allclose(L.to_host(Bt.unstack(Y2)[k]), ref, rtol=0, atol=tol * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("form", ["mf", "csr"])
@pytest.mark.parametrize("xcd_slices", [0, 1])
def test_batched_apply_large_grid_block_orders(form, xcd_slices):
    """All 20 momentum sectors of an L=20 chain (184,756 rows, 4-row chunks:
    about 11,500 blocks, so the per-XCD block orders renumber most of the
    grid) against each sector's own single-sector product."""
    import torch
    from quspin_amd import _lib as L
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.operators._devop import SectorBatch
    L_, nvec = 20, 2
    print(f"case L={L_} nvec={nvec} form={form} xcd_slices={xcd_slices}", flush=True)
    m = chain_maps(L_)
    static = xxz_static(L_, Delta=0.7)
    ops = [quantum_LinearOperator(static, basis=spin_basis_general(L_, Nup=L_ // 2, pauli=0, kblock=(m["T"], k)),
                                  dtype=np.complex128, check_symm=False, check_herm=False, check_pcon=False)
           for k in range(L_)]
    Bt = SectorBatch([op._qsx_operator() for op in ops], np.complex128, nvec, form=form, chunk_rows=4)
    Bt.set_schedule(4, xcd_slices)
    assert Bt.rows == 184756
    rng = np.random.default_rng(11)
    xs = [_rand(rng, (op.shape[0], nvec), np.complex128) for op in ops]
    X = Bt.stack(xs)
    Y = Bt.empty()
    Bt.apply(X, Y)
    got = Bt.unstack(Y)
    for k, op in enumerate(ops):
        ref = op.dot(torch.from_numpy(xs[k]).cuda())
        np.testing.assert_allclose(L.to_host(got[k]), L.to_host(ref), rtol=0, atol=1e-12 * max(1.0, float(ref.abs().max())))


def test_batched_matches_single_sector_kernel_with_parity_and_inversion():
    """k=0 and k=L/2 sectors with parity and spin inversion (real characters),
    float64: batched product == the single-operator launch, bit for bit up to
    summation order."""
    from quspin_amd import _lib as L
    from quspin_amd.operators._devop import SectorBatch
    L_ = 14
    m = chain_maps(L_)
    bases, ops, dense = _sectors(L_, [0, L_ // 2], extra={"pblock": (m["P"], 0), "zblock": (m["Z"], 1)})
    mfs = [op._qsx_operator() for op in ops]
    for dt, form in ((np.float64, "mf"), (np.complex128, "mf"), (np.float64, "csr"), (np.complex128, "csr")):
        Bt = SectorBatch(mfs, dt, 3, form=form)
        rng = np.random.default_rng(1)
        xs = [rng.normal(size=(b.Ns, 3)).astype(dt) for b in bases]
        Y = Bt.empty()
        Bt.apply(Bt.stack(xs), Y)
        for k, (mf, x) in enumerate(zip(mfs, xs)):
            xd = L.to_device(x)
            yd = L.empty(x.shape, dt)
            mf.apply(xd, yd)
            np.testing.assert_allclose(L.to_host(Bt.unstack(Y)[k]), L.to_host(yd), rtol=0, atol=1e-12)
            np.testing.assert_allclose(L.to_host(yd), dense[k] @ x, rtol=0, atol=1e-11)


def test_batched_rank_lookup_and_fermions():
    """No-symmetry sector (rank lookup) and spinful-fermion momentum sectors
    (u64 table entries with the permutation sign)."""
    from quspin_amd import _lib as L
    from quspin_amd.basis import spin_basis_general, spinful_fermion_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.operators._devop import SectorBatch
    B = spin_basis_general(12, Nup=5, pauli=0)
    Q = quantum_LinearOperator(xxz_static(12, h=np.linspace(-1, 1, 12)), basis=B, dtype=np.float64,
                               check_symm=False, check_herm=False, check_pcon=False)
    Bt = SectorBatch([Q._qsx_operator()], np.complex128, 4)
    x = _rand(np.random.default_rng(2), (B.Ns, 4), np.complex128)
    Y = Bt.empty()
    Bt.apply(Bt.stack([x]), Y)
    H = O.full_hamiltonian(O.full_basis_states(12, 5), 12, flat_terms(xxz_static(12, h=np.linspace(-1, 1, 12))))
    np.testing.assert_allclose(L.to_host(Y), H @ x, rtol=0, atol=1e-11)

    Lf = 6
    T = (np.arange(Lf) + 1) % Lf
    st = hubbard_static(Lf, U=2.0)
    ops, dense = [], []
    for k in range(Lf):
        Bf = spinful_fermion_basis_general(Lf, Nf=(3, 2), kblock=(T, k))
        Qf = quantum_LinearOperator(st, basis=Bf, dtype=np.complex128, check_symm=False, check_herm=False,
                                    check_pcon=False)
        ops.append(Qf._qsx_operator())
        dense.append(Qf)
    Bt = SectorBatch(ops, np.complex128, 2, form="mf")
    Bc = SectorBatch(ops, np.complex128, 2, form="csr")
    rng = np.random.default_rng(3)
    xs = [_rand(rng, (op.Ns, 2), np.complex128) for op in ops]
    Y = Bt.empty()
    Bt.apply(Bt.stack(xs), Y)
    Yc = Bc.empty()
    Bc.apply(Bc.stack(xs), Yc)
    for k in range(Lf):
        ref = dense[k].dot(xs[k])
        np.testing.assert_allclose(L.to_host(Bt.unstack(Y)[k]), ref, rtol=0, atol=1e-11)
        np.testing.assert_allclose(L.to_host(Bc.unstack(Yc)[k]), ref, rtol=0, atol=1e-11)


@pytest.mark.parametrize("graph,form", [(False, "mf"), (True, "mf"), (True, "csr")])
def test_batched_expm_matches_dense_per_sector(graph, form):
    """e^{-2i H_k} V_k for all k-sectors of an L=12 chain, 8 vectors each:
    every sector equals the dense exponential of its reduced matrix and the
    single-sector expm_multiply_parallel; graph replays repeat exactly."""
    from quspin_amd import _lib as L
    from quspin_amd.tools.evolution import expm_multiply_parallel, expm_multiply_sectors
    L_ = 12
    bases, ops, dense = _sectors(L_, range(L_))
    nv = 8
    E = expm_multiply_sectors(ops, a=-2.0j, dtype=np.complex128, nvec=nv, graph=graph, form=form)
    rng = np.random.default_rng(7)
    xs = [_rand(rng, (b.Ns, nv), np.complex128) for b in bases]
    X = E.batch.stack(xs)
    outs = []
    for rep in range(3 if graph else 1):
        F = E.dot(X)
        outs.append(L.to_host(F))
        L.torch().cuda.synchronize()
        _ = (X * 2).sum().item()  # unrelated eager work between replays
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])
    got = E.batch.unstack(L.to_device(outs[0]))
    for k in range(L_):
        ref = expm(-2.0j * dense[k]) @ xs[k]
        np.testing.assert_allclose(L.to_host(got[k]), ref, rtol=0, atol=1e-11)
        single = expm_multiply_parallel(ops[k], a=-2.0j, dtype=np.complex128).dot(xs[k])
        np.testing.assert_allclose(L.to_host(got[k]), single, rtol=0, atol=1e-12)


def test_batched_expm_complex64_fp64_reductions():
    """complex64 vectors: error against the FP64 batched result within the
    single-precision tolerance, uneven partitions across sectors."""
    from quspin_amd import _lib as L
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.tools.evolution import expm_multiply_sectors
    L_ = 14
    bases, ops, dense = _sectors(L_, range(0, L_, 3))
    nv = 16
    rng = np.random.default_rng(11)
    xs = [_rand(rng, (b.Ns, nv), np.complex128) for b in bases]
    xs = [x / np.linalg.norm(x, axis=0) for x in xs]
    ops32 = [quantum_LinearOperator(xxz_static(L_, Delta=0.7), basis=b, dtype=np.complex64, check_symm=False,
                                    check_herm=False, check_pcon=False) for b in bases]
    E64 = expm_multiply_sectors(ops, a=-2.0j, dtype=np.complex128, nvec=nv, form="mf")
    E32 = expm_multiply_sectors(ops32, a=-2.0j, dtype=np.complex64, nvec=nv, form="csr")
    F64 = L.to_host(E64.dot(E64.batch.stack(xs)))
    F32 = L.to_host(E32.dot(E32.batch.stack([x.astype(np.complex64) for x in xs])))
    err = np.linalg.norm(F32.astype(np.complex128) - F64, axis=0).max()
    assert err < 2e-6, err
    # uneven: force a different (m, s) per sector; results stay within tolerance
    parts = E64.partitions
    E64.set_partitions([(m + 4 * (k % 3), s) for k, (m, s) in enumerate(parts)])
    F64b = L.to_host(E64.dot(E64.batch.stack(xs)))
    np.testing.assert_allclose(F64b, F64, rtol=0, atol=1e-12)


@pytest.mark.parametrize("form", ["mf", "csr"])
def test_batched_expm_with_empty_sectors(form):
    """Sector lists that contain empty sectors (L=4 half filling with z=+1:
    k=1 and k=3 hold no state): the batch stacks, evolves and unstacks them
    as zero-row blocks, and the other sectors equal their dense exponential."""
    from quspin_amd import _lib as L
    from quspin_amd.tools.evolution import expm_multiply_sectors
    m = chain_maps(4)
    print(f"case L=4 z=+1 k=0..3 form={form}", flush=True)
    bases, ops, dense = _sectors(4, range(4), extra={"zblock": (m["Z"], 0)})
    assert [b.Ns for b in bases].count(0) == 2
    nv = 3
    E = expm_multiply_sectors(ops, a=-2.0j, dtype=np.complex128, nvec=nv, graph=False, form=form)
    rng = np.random.default_rng(11)
    xs = [_rand(rng, (b.Ns, nv), np.complex128) for b in bases]
    got = E.batch.unstack(E.dot(E.batch.stack(xs)))
    for k in range(4):
        g = L.to_host(got[k])
        assert g.shape == (bases[k].Ns, nv)
        if bases[k].Ns:
            np.testing.assert_allclose(g, expm(-2.0j * dense[k]) @ xs[k], rtol=0, atol=1e-11)
