"""Benchmark of the symmetry-reduced matrix-free H.v on one MI355X.

Workload (N=1): spin-1/2 Heisenberg chain L=28, periodic, Sz=0 + momentum
k=0 + parity + spin inversion (361,270 reduced states), float64, seeded
Gaussian vector; one step = one fused H.v launch (quantum_LinearOperator.dot).
Also timed in the same run: 200-step Lanczos on the same operator (a warm
call; the first call, which captures the recurrence, is reported beside it),
the same chain at L=34 (17,179,435 states, 64-bit state words: the
HBM-bound size), and the expm workloads (precision, batched sectors,
learned step).

Prints ONE JSON line.  ``roofline`` uses the algorithmic bytes of one launch
B = Ns*(s_i + s_n) + 2*Ns*s_v (state word, normalisation byte, v_in, v_out:
21 B per state at L=28; the off-diagonal gathers of v_in count as cache
hits) over the average launch time measured with HIP events on the launch
stream; the figure with the folded diagonal (8 B per state more) is given
beside it.  ``cpu_baseline`` is the C/OpenMP restatement in oracle/ running
the same algorithm (fused rows through a sector lookup table) on the host
cores, timed on a bounded number of matvecs; the reference's own per-element
orbit search is timed beside it.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the dominant kernel, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes scaled by the copy-kernel calibration (fetch x2.0, write
# x1.0); keyed by (L, lookup, schedule, dtype).  Source and commands:
# profiles/k_hv_rows_heisenberg_L28_kpz.md, profiles/k_hv_rows_heisenberg_L34_kpz.md
# For these operators the "rows", "batch4", "batch8", "batch14" and "batch14x"
# schedules all launch k_hv_rows_fast_batch (batches of 7 or 6, 4, 8 and 14
# terms; each XCD takes runs of 64 or 256 blocks, batch14x one run per XCD).
MEASURED_TRAFFIC = {(28, "table", "rows", "float64"): 68.5e6, (28, "table", "batch4", "float64"): 67.9e6,
                    (28, "table", "batch8", "float64"): 67.8e6, (28, "table", "batch14", "float64"): 71.3e6,
                    (28, "table", "batch14x", "float64"): 45.8e6, (34, "table", "batch14x", "float64"): 7.05e9,
                    (34, "table", "rows", "float64"): 7.42e9, (34, "table", "batch4", "float64"): 7.35e9,
                    (34, "table", "batch8", "float64"): 7.32e9, (34, "table", "batch14", "float64"): 7.30e9,
                    ("csr", 34, "float64"): 11.99e9}
# The same for the precision workload's term kernels (full launches, per
# launch): profiles/k_csr_spmv_expm_hubbard_L14_xxz_L28.md
C3_TRAFFIC = {("mf", "complex64", "rows"): 3.76e9, ("mf", "complex128", "rows"): 7.44e9,
              ("mf", "complex64", "split2"): 2.95e9, ("mf", "complex128", "split2"): 5.17e9,
              ("csr", "complex64"): 8.22e9, ("csr", "complex128"): 15.03e9}
# ... and of one batched H.V of the 26-sector workload (CSR form, all sectors):
# profiles/k_csr_wave_batched_heisenberg_L26_64vec.md
C4_TRAFFIC = {("csr", "fp32"): 59.1e9, ("csr", "fp64"): 161.6e9}


def chain_static(L, J=1.0, Delta=1.0):
    hop = [[0.5 * J, i, (i + 1) % L] for i in range(L)]
    zz = [[J * Delta, i, (i + 1) % L] for i in range(L)]
    return [["+-", hop], ["-+", hop], ["zz", zz]]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--L", type=int, default=28)
    ap.add_argument("--lanczos", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="only run the plain copy kernel on this many MiB (HBM counter calibration)")
    ap.add_argument("--no-expm", action="store_true", help="skip the expm workloads (precision, batched, learned)")
    ap.add_argument("--no-scaled", action="store_true", help="skip the L=34 form of the headline workload")
    ap.add_argument("--scaled-steps", type=int, default=20)
    ap.add_argument("--expm-only", default="", help="comma list of c3,c4,c5 to run (default all)")
    ap.add_argument("--c3-steps", type=int, default=100)
    ap.add_argument("--c5-steps", type=int, default=500)
    ap.add_argument("--c5-runs", default="", help="comma list of the driven-Hubbard runs to time (default all)")
    ap.add_argument("--no-graph", action="store_true",
                    help="driven-Hubbard steps launched directly, without capturing step sequences")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: it measures the single-GPU kernels; the multi-GPU "
              "sharding of these workloads is a design, not a run", file=sys.stderr)
        sys.exit(2)

    import torch
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.tools.lanczos import lanczos_full

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    if args.copy_mb:
        from quspin_amd import _lib
        nb = args.copy_mb << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(args.warmup + args.steps):
            _lib.check(_lib.load().qsx_copy(_lib.ptr(src), _lib.ptr(dst), nb, _lib.stream_handle()), "qsx_copy")
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": nb, "launches": args.warmup + args.steps}))
        return
    L = args.L
    i = np.arange(L)
    T, P, Z = (i + 1) % L, i[::-1].copy(), -(i + 1)
    t0 = time.time()
    B = spin_basis_general(L, Nup=L // 2, pauli=0, kblock=(T, 0), pblock=(P, 0), zblock=(Z, 0))
    torch.cuda.synchronize()
    t_basis = time.time() - t0
    static = chain_static(L)
    Q = quantum_LinearOperator(static, basis=B, dtype=np.float64, check_symm=False, check_herm=False,
                               check_pcon=False)
    Ns = B.Ns
    gen = torch.Generator(device="cuda").manual_seed(1234)
    v = torch.randn(Ns, dtype=torch.float64, device="cuda", generator=gen)
    out = torch.empty_like(v)
    op = Q._qsx_operator()
    for _ in range(args.warmup):
        op.apply(v, out)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall0 = time.perf_counter()
    e0.record(st)
    for _ in range(args.steps):
        op.apply(v, out)
    e1.record(st)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall0
    ms_total = e0.elapsed_time(e1)
    ms_per = ms_total / args.steps
    value = 1000.0 / ms_per

    # per reduced state: state word (4 B up to 32 sites, else 8 B), 1 B n, 8 B x read, 8 B y written;
    # the kernel also reads the folded diagonal (8 B), reported beside it
    s_i = 4 if L <= 32 else 8
    nbytes = Ns * (s_i + 1) + 2 * Ns * 8
    achieved = nbytes / (ms_per * 1e-3) / 1e9
    achieved_fd = (nbytes + 8 * Ns) / (ms_per * 1e-3) / 1e9

    # 200-step Lanczos on the same operator (device recurrence)
    lz = {}
    if args.lanczos > 0:
        # first call: captures the recurrence as one graph; the timed call replays it
        torch.cuda.synchronize()
        tl0 = time.perf_counter()
        E, V, Qm = lanczos_full(Q, v, args.lanczos)
        torch.cuda.synchronize()
        t_first = time.perf_counter() - tl0
        del Qm
        tl0 = time.perf_counter()
        E, V, Qm = lanczos_full(Q, v, args.lanczos)
        torch.cuda.synchronize()
        tl = time.perf_counter() - tl0
        lz = {"steps": int(Qm.shape[0]), "seconds": round(tl, 4), "steps_per_s": round(Qm.shape[0] / tl, 1),
              "first_call_seconds": round(t_first, 4), "E0": float(E[0])}
        del Qm

    cpu = None
    if not args.no_cpu and L <= 30:
        from oracle import c_oracle
        g = c_oracle.Group(L, [P, Z, T], [2, 2, L], [0, 0, 0])
        cb, cn = g.make_basis(L // 2)
        if not np.array_equal(cb, B.states.astype(np.uint64)):
            raise SystemExit("CPU restatement basis differs from the device basis")
        tc0 = time.perf_counter()
        ctab = c_oracle.table_build(g, L // 2, cb)
        t_tab = time.perf_counter() - tc0
        terms = [(op_, tuple(b[1:]), b[0]) for op_, bonds in static for b in bonds]
        vh = v.cpu().numpy()
        ref = c_oracle.hv_table_real(g, cb, cn, ctab, terms, vh)  # warm-up + parity
        err = float(np.abs(ref - out.cpu().numpy()).max())

        def timed_calls(fn, budget, cap):
            nrep, tsum = 0, 0.0
            while tsum < budget and nrep < cap:
                tc = time.perf_counter()
                fn()
                tsum += time.perf_counter() - tc
                nrep += 1
            return nrep, tsum

        nrep, tsum = timed_calls(lambda: c_oracle.hv_table_real(g, cb, cn, ctab, terms, vh), args.cpu_seconds, 2000)
        nref, tref = timed_calls(lambda: c_oracle.hv_real(g, cb, cn, terms, vh, fused=True), args.cpu_seconds / 3, 50)
        cpu = {"value": round(nrep / tsum, 3), "unit": "matvecs/s", "cores": c_oracle.max_threads(),
               "kind": "port",
               "sample": f"{nrep} fused H.v calls on the same L={L} k=0,p=+1,z=+1 sector ({Ns} states), "
                         f"C/OpenMP restatement of the same algorithm (rows through a sector lookup table, "
                         f"oracle/csrc/qsx_oracle.c qo_hv_table_real; table built once in {t_tab:.1f} s, not timed); "
                         f"max|GPU-CPU| = {err:.2e}",
               "reference_algorithm": {"value": round(nref / tref, 3), "unit": "matvecs/s",
                                       "sample": f"{nref} calls of the per-element orbit search + binary search "
                                                 "form (the reference's loop, fused over terms; qo_hv_real)"}}

    cfg = next(iter(op.config.values()), {})
    traffic = MEASURED_TRAFFIC.get((L, cfg.get("lookup"), cfg.get("schedule"), "float64"))
    scaled = None
    if not args.no_scaled and L == 28:
        scaled = run_c2_scaled(args.scaled_steps)
    expm = {}
    if not args.no_expm:
        which = set(args.expm_only.split(",")) if args.expm_only else {"c3", "c4", "c5"}
        if "c3" in which:
            expm["precision_L28"] = run_precision(args.c3_steps, cpu=not args.no_cpu)
        if "c4" in which:
            expm["batched_L26"] = run_batched(cpu=not args.no_cpu)
        if "c5" in which:
            expm["learned_hubbard_L14"] = run_learned(args.c5_steps, runs=args.c5_runs, graph=not args.no_graph,
                                                      cpu=not args.no_cpu)

    line = {
        "metric": "H.v matvecs/s (symmetry-reduced matrix-free, fp64)",
        "value": round(value, 2),
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Heisenberg chain L={L} PBC, Sz=0, k=0, p=+1, z=+1 ({Ns} states), "
                               "quantum_LinearOperator.dot on a seeded Gaussian vector",
                   "Ns": Ns, "basis_build_s": round(t_basis, 3), "wall_s": round(t_wall, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "alg_bytes_per_state": s_i + 1 + 16,
                     "achieved_with_folded_diag": round(achieved_fd, 2),
                     "limit": ("cache-resident working set: divergent gathers (about one cache line per lane "
                               "for each applied term's table entry and x element) through each CU's vector "
                               "memory path; halving the L2 misses (schedule batch14x) leaves the time unchanged "
                               "(profiles/k_hv_rows_heisenberg_L28_kpz.md); the HBM-bound size is "
                               "c2_scaled_L34") if L <= 30 else "sector-table reads from HBM"},
        "cpu_baseline": cpu,
        "lanczos": lz,
        "c2_scaled_L34": scaled,
        "expm": expm,
        "kernel_config": {str(k): v for k, v in op.config.items()},
    }
    print(json.dumps(line))


def run_c2_scaled(steps, L=34):
    """The headline chain at L=34: Sz=0, k=0, p=+1, z=+1 (17,179,435 states,
    64-bit state words; 429 MB of algorithmic bytes per product, above the
    256 MB MALL).  Matrix-free quantum_LinearOperator.dot (sector table,
    4-byte entries) and hamiltonian.dot on the device CSR of the same
    operator, each timed over `steps` launches with HIP events; the two
    results are compared."""
    import time
    import torch
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian, quantum_LinearOperator
    i = np.arange(L)
    T, P, Z = (i + 1) % L, i[::-1].copy(), -(i + 1)
    t0 = time.perf_counter()
    B = spin_basis_general(L, Nup=L // 2, pauli=0, kblock=(T, 0), pblock=(P, 0), zblock=(Z, 0))
    torch.cuda.synchronize()
    t_basis = time.perf_counter() - t0
    t0 = time.perf_counter()
    tab = B.lookup_table()
    torch.cuda.synchronize()
    t_tab = time.perf_counter() - t0
    static = chain_static(L)
    kw = dict(check_symm=False, check_herm=False, check_pcon=False)
    Q = quantum_LinearOperator(static, basis=B, dtype=np.float64, **kw)
    Ns = B.Ns
    gen = torch.Generator(device="cuda").manual_seed(1234)
    v = torch.randn(Ns, dtype=torch.float64, device="cuda", generator=gen)
    y = torch.empty_like(v)
    op = Q._qsx_operator()
    op.apply(v, y)
    torch.cuda.synchronize()
    st, e0, e1 = _ev_timer()
    e0.record(st)
    for _ in range(steps):
        op.apply(v, y)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    alg = Ns * (8 + 1) + 2 * Ns * 8
    cfg = next(iter(op.config.values()), {})
    traffic = MEASURED_TRAFFIC.get((L, cfg.get("lookup"), cfg.get("schedule"), "float64"))
    t0 = time.perf_counter()
    H = hamiltonian(static, [], basis=B, dtype=np.float64, **kw)
    torch.cuda.synchronize()
    t_csr = time.perf_counter() - t0
    M = H._static
    y2 = torch.empty_like(v)
    M.apply(v, y2)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(steps):
        M.apply(v, y2)
    e1.record(st)
    torch.cuda.synchronize()
    ms_csr = e0.elapsed_time(e1) / steps
    alg_csr = 4 * (Ns + 1) + M.nnz * (4 + 8) + 2 * Ns * 8
    err = float((y - y2).abs().max())
    out = {"workload": f"Heisenberg chain L={L} PBC, Sz=0, k=0, p=+1, z=+1 ({Ns} states, uint64 words), "
                       "seeded Gaussian vector",
           "Ns": Ns, "basis_build_s": round(t_basis, 3), "table_build_s": round(t_tab, 3),
           "table_bytes": int(tab[0].numel() * tab[0].element_size()),
           "matrix_free": {"matvecs_per_s": round(1000.0 / ms, 2), "ms_per_step": round(ms, 5),
                           "kernel_config": cfg,
                           "roofline": {"bound": "hbm", "achieved": round(alg / ms / 1e6, 2), "peak": HBM_PEAK_GBS,
                                        "unit": "GB/s", "frac": round(alg / ms / 1e6 / HBM_PEAK_GBS, 5),
                                        "traffic": traffic, "alg_bytes_per_state": 25,
                                        "achieved_with_folded_diag": round((alg + 8 * Ns) / ms / 1e6, 2),
                                        "limit": "L2-miss traffic of the sector-table reads (one 32 B sector per "
                                                 "applied term; profiles/k_hv_rows_heisenberg_L34_kpz.md)"}},
           "csr": {"matvecs_per_s": round(1000.0 / ms_csr, 2), "ms_per_step": round(ms_csr, 5), "nnz": int(M.nnz),
                   "assembly_s": round(t_csr, 3),
                   "roofline": {"bound": "hbm", "achieved": round(alg_csr / ms_csr / 1e6, 2), "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": round(alg_csr / ms_csr / 1e6 / HBM_PEAK_GBS, 5),
                                "alg_bytes": int(alg_csr), "traffic": MEASURED_TRAFFIC.get(("csr", L, "float64"))}},
           "max_abs_diff_matrix_free_vs_csr": err}
    del H, M, Q, op, tab, B, v, y, y2
    torch.cuda.empty_cache()
    return out


def _ev_timer():
    import torch
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    return st, e0, e1


def run_precision(steps, L=28, dt=0.1, cpu=True):
    """Both forms of the L=28 precision workload: matrix-free
    (quantum_LinearOperator) and CSR (hamiltonian); with ``cpu`` the CSR form
    also times one FP64 step of the C/OpenMP Taylor loop on the host."""
    out = {"matrix_free": _run_precision(steps, "mf", L, dt)}
    out["csr"] = _run_precision(steps, "csr", L, dt, cpu=cpu)
    return out


def _kernel_roofline(op, x, form, Ns, traffic):
    """One H.v launch of the Taylor loop's operator on x, timed over 5
    launches with HIP events, against its algorithmic bytes."""
    import torch
    y = torch.empty_like(x)
    op.apply(x, y)
    torch.cuda.synchronize()
    st, e0, e1 = _ev_timer()
    e0.record(st)
    for _ in range(5):
        op.apply(x, y)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    esz = x.element_size()
    if form == "mf":
        alg = Ns * 4 + 2 * Ns * esz
    else:
        alg = op.indptr.element_size() * (Ns + 1) + op.nnz * (op.indices.element_size() + op.data.element_size()) \
            + 2 * Ns * esz
    gbs = alg / ms / 1e6
    out = {"ms": round(ms, 4), "alg_bytes": int(alg), "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}
    if form == "mf":
        cfg = next((c for (dt_, nv), c in op.config.items() if np.dtype(dt_) == np.dtype(_np_dtype(x)) and nv == 1),
                   {})
        out["schedule"] = cfg.get("schedule")
        traffic = C3_TRAFFIC.get(("mf", np.dtype(_np_dtype(x)).name, cfg.get("schedule")))
    out["traffic"] = traffic
    return out


def _np_dtype(t):
    import torch
    return {torch.complex64: np.complex64, torch.complex128: np.complex128, torch.float32: np.float32,
            torch.float64: np.float64}[t.dtype]


def _run_precision(steps, form, L=28, dt=0.1, cpu=False):
    """XXZ chain L=28, Delta=1, h_i ~ U[-3, 3] (seed 0), Sz=0 without spatial
    symmetry (40,116,600 states); Neel start; `steps` matrix-free expm steps
    of dt in complex64 (FP64 accumulation inside every kernel) and in
    complex128; state error and max |<Sz_i>| error after every step."""
    import torch
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import hamiltonian, quantum_LinearOperator
    from quspin_amd.operators._devop import vdot_dev
    from quspin_amd.tools.evolution import expm_multiply_parallel
    rng = np.random.default_rng(0)
    h = rng.uniform(-3, 3, L)
    static = chain_static(L) + [["z", [[float(h[i]), i] for i in range(L)]]]
    B = spin_basis_general(L, Nup=L // 2, pauli=0)
    kw = dict(basis=B, check_symm=False, check_herm=False, check_pcon=False)
    Op = quantum_LinearOperator if form == "mf" else (lambda st, **k: hamiltonian(st, [], **k))
    E = {np.complex64: expm_multiply_parallel(Op(static, dtype=np.float32, **kw), a=-1j * dt, dtype=np.complex64),
         np.complex128: expm_multiply_parallel(Op(static, dtype=np.float64, **kw), a=-1j * dt, dtype=np.complex128)}
    zops = [quantum_LinearOperator([["z", [[1.0, i]]]], dtype=np.float64, **kw)._qsx_operator() for i in range(L)]
    i0 = B.index("10" * (L // 2))
    psi = {}
    for dt_ in E:
        psi[dt_] = torch.zeros(B.Ns, dtype=torch.complex64 if dt_ == np.complex64 else torch.complex128, device="cuda")
        psi[dt_][i0] = 1
    work = {dt_: torch.empty(2 * B.Ns, dtype=psi[dt_].dtype, device="cuda") for dt_ in E}
    tmp = torch.empty(B.Ns, dtype=torch.complex128, device="cuda")

    def sz(p):
        p = p if p.dtype == torch.complex128 else p.to(torch.complex128)
        out = []
        for zo in zops:
            zo.apply(p, tmp)
            out.append(vdot_dev(p, tmp).real)
        return np.array(out)

    ms = {dt_: 0.0 for dt_ in E}
    errs, szerr = [], []
    st, e0, e1 = _ev_timer()
    for k in range(steps):
        for dt_ in E:
            e0.record(st)
            psi[dt_] = E[dt_].dot(psi[dt_], work_array=work[dt_], overwrite_v=True)
            e1.record(st)
            torch.cuda.synchronize()
            ms[dt_] += e0.elapsed_time(e1)
        p32 = psi[np.complex64].to(torch.complex128)
        p64 = psi[np.complex128]
        d2 = vdot_dev(p32, p32).real + vdot_dev(p64, p64).real - 2 * vdot_dev(p32, p64).real
        errs.append(float(np.sqrt(max(d2, 0.0))))
        szerr.append(float(np.abs(sz(p32) - sz(p64)).max()))
    sp32 = 1000.0 * steps / ms[np.complex64]
    sp64 = 1000.0 * steps / ms[np.complex128]
    kern = {}
    for dt_, name in ((np.complex64, "fp32"), (np.complex128, "fp64")):
        kern[name] = _kernel_roofline(E[dt_]._op, psi[dt_], form, B.Ns,
                                      C3_TRAFFIC.get((form, np.dtype(dt_).name)))
    res = {"workload": f"XXZ L={L} Delta=1 h_i~U[-3,3] seed 0, Sz=0 ({B.Ns} states), Neel start, {steps} steps dt={dt}",
           "path": ("matrix-free Taylor loop (quantum_LinearOperator)" if form == "mf" else
                    "CSR Taylor loop (hamiltonian, real FP32/FP64 values)"),
           "partition": [int(x) for x in E[np.complex64].partition],
           "steps_per_s_fp32": round(sp32, 3), "steps_per_s_fp64": round(sp64, 3),
           "speedup_fp32_vs_fp64": round(sp32 / sp64, 3),
           "state_err_max": max(errs), "state_err_final": errs[-1], "state_err_every_10": errs[9::10],
           "sz_err_max": max(szerr), "sz_err_every_10": szerr[9::10],
           "hv_kernel": kern}
    if cpu and form == "csr":
        res["cpu_baseline"] = _cpu_taylor_step(E[np.complex128], B.Ns, i0)
    return res


def _cpu_taylor_step(Ed, Ns, i0):
    """One FP64 step of the C/OpenMP Taylor loop on the host (real CSR
    values, complex128 state; the same partition, shift and tolerance as the
    device), from the Neel state; compared with the device's FP64 step."""
    import time
    import torch
    from quspin_amd import _lib
    from oracle import c_oracle
    M = Ed._op
    ip, ix = _lib.to_host(M.indptr), _lib.to_host(M.indices)
    if ip.dtype != ix.dtype:
        ip, ix = ip.astype(np.int64), ix.astype(np.int64)
    R = c_oracle.RealCSR(ip, ix, _lib.to_host(M.data).astype(np.float64))
    m_star, s = (int(x) for x in Ed.partition)
    psi0 = np.zeros(Ns, np.complex128)
    psi0[i0] = 1
    t0 = time.perf_counter()
    F = R.expm(psi0, s, m_star, float(Ed._tol), complex(Ed._mu), complex(Ed._a))
    t_cpu = time.perf_counter() - t0
    p = torch.zeros(Ns, dtype=torch.complex128, device="cuda")
    p[i0] = 1
    G = Ed.dot(p)
    err = float(np.abs(_lib.to_host(G) - F).max())
    return {"value": round(1.0 / t_cpu, 4), "unit": "steps/s", "cores": c_oracle.max_threads(), "kind": "port",
            "sample": f"1 FP64 step (partition m*={m_star}, s={s}) of the C/OpenMP Taylor loop "
                      "(oracle/csrc/qsx_oracle.c qo_expm_rcsr_c128) on the host copy of the CSR; "
                      f"max|GPU-CPU| after the step = {err:.2e}"}


def _cpu_sector_expm(op, V0, a, n_sectors, nv, n_cpu=2):
    """The C/OpenMP complex128 CSR Taylor loop on the host for `n_cpu` of the
    first sector's vectors (same partition, shift and tolerance as the
    device), scaled to all sectors and vectors; compared with the device."""
    import time
    from quspin_amd import _lib
    from quspin_amd.tools.evolution import expm_multiply_parallel
    from oracle import c_oracle
    E = expm_multiply_parallel(op, a=a, dtype=np.complex128)
    A = E._op.to_csr(np.complex128).to_scipy()
    m_star, s = (int(x) for x in E.partition)
    Vh = _lib.to_host(V0[:, :n_cpu]).astype(np.complex128)
    t0 = time.perf_counter()
    F = np.stack([c_oracle.expm_c128(A, Vh[:, j], s, m_star, float(E._tol), complex(E._mu), a)
                  for j in range(n_cpu)], axis=1)
    per_vec = (time.perf_counter() - t0) / n_cpu
    G = _lib.to_host(E.dot(V0[:, :n_cpu].contiguous().to(_lib.torch_dtype(np.complex128))))
    err = float(np.abs(G - F).max())
    total = per_vec * nv * n_sectors
    return {"value": round(1.0 / total, 6), "unit": "batched expm steps/s", "cores": c_oracle.max_threads(),
            "kind": "port",
            "sample": f"{n_cpu} vectors of sector k=0 (partition m*={m_star}, s={s}) through the C/OpenMP complex128 "
                      f"CSR Taylor loop (oracle/csrc/qsx_oracle.c qo_expm_c128), {per_vec:.2f} s each, scaled by "
                      f"{nv} vectors x {n_sectors} sectors (estimate: {total:.0f} s per step); "
                      f"max|GPU-CPU| = {err:.2e}"}


def run_batched(L=26, nv=64, t=2.0, plain_sectors=2, cpu=True):
    """Heisenberg L=26, Sz=0, all 26 momentum sectors x 64 seeded states:
    batched expm to t=2 (one launch per Taylor term for every sector, hipGraph
    replay, autotuned form and schedule) in complex64 and complex128; the
    plain FP64 path (one expm_multiply_parallel per sector) timed on
    `plain_sectors` sectors and scaled to 26."""
    import torch
    from quspin_amd import _lib
    from quspin_amd.basis import spin_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.tools.evolution import expm_multiply_parallel, expm_multiply_sectors
    i = np.arange(L)
    T = (i + 1) % L
    static = chain_static(L)
    bases = [spin_basis_general(L, Nup=L // 2, pauli=0, kblock=(T, k)) for k in range(L)]
    kw = dict(check_symm=False, check_herm=False, check_pcon=False)
    ops = {dt_: [quantum_LinearOperator(static, basis=b, dtype=dt_, **kw) for b in bases]
           for dt_ in (np.complex64, np.complex128)}
    rows = sum(b.Ns for b in bases)
    g = torch.Generator(device="cuda").manual_seed(7)
    V = torch.randn(rows, nv, dtype=torch.complex128, device="cuda", generator=g)
    out = {"workload": f"Heisenberg L={L}, Sz=0, k=0..{L - 1} ({rows} states, sectors "
                       f"{min(b.Ns for b in bases)}..{max(b.Ns for b in bases)}) x {nv} vectors, expm to t={t}"}
    F = {}
    st, e0, e1 = _ev_timer()
    for dt_, name in ((np.complex64, "fp32"), (np.complex128, "fp64")):
        E = expm_multiply_sectors(ops[dt_], a=-1j * t, dtype=dt_, nvec=nv)
        X = V.to(_lib.torch_dtype(dt_))
        E.dot(X)
        torch.cuda.synchronize()
        e0.record(st)
        F[name] = E.dot(X)
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        # one batched H.V (the dominant kernel) for the roofline
        Y = torch.empty_like(X)
        E.batch.apply(X, Y)
        torch.cuda.synchronize()
        e0.record(st)
        for _ in range(5):
            E.batch.apply(X, Y)
        e1.record(st)
        torch.cuda.synchronize()
        mv = e0.elapsed_time(e1) / 5
        esz = np.dtype(dt_).itemsize
        alg = rows * (4 + 1) + 2 * rows * nv * esz
        traffic = C4_TRAFFIC.get((E.batch.form, name))
        out[name] = {"expm_s": round(ms / 1000.0, 4), "form": E.batch.form,
                     "schedule": [E.batch.chunk_rows, E.batch.xcd_slices],
                     "partitions": sorted({tuple(int(x) for x in p) for p in E.partitions}),
                     "batched_hv_ms": round(mv, 4),
                     "batched_hv_roofline": {"bound": "hbm", "achieved": round(alg / mv / 1e6, 1),
                                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                             "frac": round(alg / mv / 1e6 / HBM_PEAK_GBS, 4), "alg_bytes": int(alg),
                                             "traffic": traffic,
                                             "traffic_GBps": round(traffic / mv / 1e6, 1) if traffic else None}}
        del E, X, Y
    f32 = F["fp32"].to(torch.complex128)
    f64 = F["fp64"]
    out["state_err_max_column"] = float(((f32 - f64).norm(dim=0) / V.norm(dim=0)).max().item())
    del F, f32, f64
    # plain FP64 GPU path: one expm_multiply_parallel per sector (64 vectors each)
    tp = 0.0
    off = 0
    for k in range(plain_sectors):
        Ek = expm_multiply_parallel(ops[np.complex128][k], a=-1j * t, dtype=np.complex128)
        Vk = V[off:off + bases[k].Ns]
        off += bases[k].Ns
        torch.cuda.synchronize()
        e0.record(st)
        Ek.dot(Vk)
        e1.record(st)
        torch.cuda.synchronize()
        tp += e0.elapsed_time(e1) / 1000.0
    plain = tp * L / plain_sectors
    if cpu:
        out["cpu_baseline"] = _cpu_sector_expm(ops[np.complex128][0], V[:bases[0].Ns], -1j * t, L, nv)
    out["plain_fp64_s_est"] = round(plain, 3)
    out["plain_fp64_sample"] = (f"estimate: {plain_sectors} of {L} sectors timed ({tp:.3f} s), scaled by "
                                f"{L}/{plain_sectors} (the sectors hold 400,023 or 400,024 states each)")
    out["speedup_fp32_batched_vs_plain_fp64_est"] = round(plain / out["fp32"]["expm_s"], 2)
    out["speedup_fp32_batched_vs_batched_fp64"] = round(out["fp64"]["expm_s"] / out["fp32"]["expm_s"], 2)
    return out


def _cpu_driven_steps(hop, D, f, dt, i0, nsteps):
    """`nsteps` FP64 steps of the driven workload with the C/OpenMP Taylor
    loop on the host, each with the reference's per-step work: the 1-norm of
    H(t) - mu I (one pass over the CSR of H0 plus f(t) D), the partition by
    scipy's _fragment_3_1, then the Taylor loop.  Returns the state and the
    seconds per step."""
    import time
    from quspin_amd import _lib
    from quspin_amd.operators._devop import as_device_operator
    from quspin_amd.tools.learned_step import exact_partition
    from oracle import c_oracle
    M = as_device_operator(hop).to_csr(np.float64)
    ip, ix = _lib.to_host(M.indptr), _lib.to_host(M.indices)
    if ip.dtype != ix.dtype:
        ip, ix = ip.astype(np.int64), ix.astype(np.int64)
    Dh = _lib.to_host(as_device_operator(D).diagonal(np.float64)).astype(np.float64)
    R = c_oracle.RealCSR(ip, ix, _lib.to_host(M.data).astype(np.float64), Dh)
    N = Dh.size
    rows = np.repeat(np.arange(N), np.diff(ip.astype(np.int64)))
    trH0 = float(R.data[ix == rows].sum())
    trD = float(Dh.sum())
    tol = 2.0 ** -53
    F = np.zeros(N, np.complex128)
    F[i0] = 1
    t0 = time.perf_counter()
    for k in range(nsteps):
        t = (k + 0.5) * dt
        mu = (trH0 + f(t) * trD) / N
        m_star, s = exact_partition(R.one_norm(f(t), mu), -1j * dt, tol)
        F = R.expm(F, s, m_star, tol, mu, -1j * dt, dscale=f(t))
    return F, (time.perf_counter() - t0) / nsteps


def run_learned(steps, L=14, dt=0.02, runs="", graph=True, cpu=True):
    """1D spinful Fermi-Hubbard L=14 PBC, t_hop=1, U(t)=4+2cos(3t), (7,7)
    (11,778,624 states): `steps` expm steps of dt in complex64 with the
    partition predicted by the device MLP (trained here on reference
    partitions of L=10 and L=12), against the exact per-step selection in
    complex64 and the plain FP64 path (complex128, exact selection); each
    with H0 as one CSR and as the species-factorised operator.  `runs`
    selects run names (the FP64 CSR reference always runs); `graph=False`
    launches every step directly instead of replaying captured blocks."""
    runs_sel = set(r for r in runs.split(",") if r)
    import time
    import torch
    from quspin_amd.basis import spinful_fermion_basis_general
    from quspin_amd.operators import quantum_LinearOperator
    from quspin_amd.operators._devop import vdot_dev
    from quspin_amd.tools.learned_step import DrivenEvolution, PartitionMLP, training_set
    f = lambda t: 4.0 + 2.0 * np.cos(3.0 * t)  # noqa: E731
    phase = lambda t: (3.0 * t) % (2 * np.pi)  # noqa: E731
    kw = dict(check_symm=False, check_herm=False, check_pcon=False)

    def hubbard(Lh, dtype):
        Bh = spinful_fermion_basis_general(Lh, Nf=(Lh // 2, Lh // 2))
        hop_r = [[-1.0, i, (i + 1) % Lh] for i in range(Lh)]
        hop_l = [[+1.0, i, (i + 1) % Lh] for i in range(Lh)]
        hop = quantum_LinearOperator([["+-|", hop_r], ["-+|", hop_l], ["|+-", hop_r], ["|-+", hop_l]], basis=Bh,
                                     dtype=dtype, **kw)
        D = quantum_LinearOperator([["n|n", [[1.0, i, i] for i in range(Lh)]]], basis=Bh, dtype=np.float64, **kw)
        return Bh, hop, D

    t0 = time.perf_counter()
    evs = []
    for Lh in (10, 12):
        _, hop, D = hubbard(Lh, np.complex128)
        evs.append(DrivenEvolution(hop, D, f, dt, dtype=np.complex128, phase=phase))
    X, Y = training_set(evs, t_max=2 * np.pi / 3 * 2, dts=(0.005, 0.01, 0.02, 0.04, 0.08), n_t=40, seed=0)
    t_data = time.perf_counter() - t0
    net = PartitionMLP(seed=0)
    t0 = time.perf_counter()
    hist = net.fit(X, Y, epochs=3000, lr=3e-3)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    del evs
    B, hop32, D = hubbard(L, np.complex64)
    _, hop64, _ = hubbard(L, np.complex128)
    rng = np.random.default_rng(3)
    i0 = int(rng.integers(B.Ns))
    res = {}
    # fp32: complex64 matrix, Taylor terms and SpMV with the state accumulated in complex128
    # "csr": H0 streamed as one CSR per Taylor term; "kron": H0 = H_up (x) 1 + 1 (x) H_down, psi read once per term
    # exact selection: "host" = per step on the host, each step launched on
    # its own (the reference's structure); "device" = per block of 100 steps
    # on the device, the block replayed like the learned blocks (the fair
    # baseline of the network)
    runs = (("learned_fp32", np.complex64, hop32, net, True, "csr", "host"),
            ("exact_dev_fp32", np.complex64, hop32, None, False, "csr", "device"),
            ("exact_fp32", np.complex64, hop32, None, False, "csr", "host"),
            ("exact_fp64", np.complex128, hop64, None, False, "csr", "host"),
            ("exact_dev_fp64", np.complex128, hop64, None, False, "csr", "device"),
            ("learned_fp32_kron", np.complex64, hop32, net, True, "kron", "host"),
            ("exact_dev_fp32_kron", np.complex64, hop32, None, False, "kron", "device"),
            ("learned_fp64_kron", np.complex128, hop64, net, True, "kron", "host"),
            ("exact_dev_fp64_kron", np.complex128, hop64, None, False, "kron", "device"))
    if runs_sel:
        runs = tuple(r for r in runs if r[0] in runs_sel or r[0] == "exact_fp64")
    psis = {}
    for name, dt_, hop, pred, learned, form, exact in runs:
        ev = DrivenEvolution(hop, D, f, dt, dtype=dt_, phase=phase, predictor=pred, state_dtype=np.complex128,
                             form=form)
        psi = torch.zeros(B.Ns, dtype=torch.complex128, device="cuda")
        psi[i0] = 1
        # capture / first launches
        ev.run(psi, 0.0, min(steps, 100), learned=learned, block=100, graph=graph, exact=exact)
        psi.zero_()
        psi[i0] = 1
        ev.stats = {k: 0 for k in ev.stats}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev.run(psi, 0.0, steps, learned=learned, block=100, graph=graph, exact=exact)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        res[name] = {"steps_per_s": round(steps / wall, 2), "ms_per_step": round(1000.0 * wall / steps, 4),
                     "stats": dict(ev.stats)}
        psis[name] = psi.to(torch.complex128)
        del ev
    p64 = psis["exact_fp64"]
    for name in [r[0] for r in runs if r[0] != "exact_fp64"]:
        p = psis[name]
        d2 = vdot_dev(p, p).real + vdot_dev(p64, p64).real - 2 * vdot_dev(p, p64).real
        res[name]["state_err_vs_fp64"] = float(np.sqrt(max(d2, 0.0)))
    res["workload"] = (f"spinful Fermi-Hubbard L={L} PBC, t=1, U(t)=4+2cos(3t), Nf=({L // 2},{L // 2}) ({B.Ns} "
                       f"states), {steps} steps dt={dt} from a seeded basis state")
    res["training"] = {"samples": int(X.shape[0]), "sizes": [10, 12], "epochs": 3000,
                       "loss_first_last": [hist[0][1], hist[-1][1]], "data_s": round(t_data, 2),
                       "fit_s": round(t_fit, 2)}
    def ratio(a, b):
        if a in res and b in res:
            return round(res[a]["steps_per_s"] / res[b]["steps_per_s"], 3)
        return None
    # same operator form, same structure of launches: what FP32 and the network each gain
    res["speedup_learned_fp32_vs_exact_dev_fp64"] = ratio("learned_fp32", "exact_dev_fp64")
    res["gain_learned_vs_exact_dev_fp32"] = ratio("learned_fp32", "exact_dev_fp32")
    res["speedup_exact_dev_fp32_vs_exact_dev_fp64"] = ratio("exact_dev_fp32", "exact_dev_fp64")
    res["speedup_learned_fp32_kron_vs_exact_dev_fp64_kron"] = ratio("learned_fp32_kron", "exact_dev_fp64_kron")
    res["gain_learned_vs_exact_dev_fp32_kron"] = ratio("learned_fp32_kron", "exact_dev_fp32_kron")
    # against the plain FP64 path with the reference's per-step host selection
    res["speedup_learned_fp32_vs_exact_fp64"] = ratio("learned_fp32", "exact_fp64")
    res["gain_learned_vs_exact_fp32"] = ratio("learned_fp32", "exact_fp32")
    res["speedup_learned_fp32_kron_vs_exact_fp64_csr"] = ratio("learned_fp32_kron", "exact_fp64")
    if cpu:
        nc = 2
        F, sec = _cpu_driven_steps(hop64, D, f, dt, i0, nc)
        ev = DrivenEvolution(hop64, D, f, dt, dtype=np.complex128, phase=phase)
        p = torch.zeros(B.Ns, dtype=torch.complex128, device="cuda")
        p[i0] = 1
        ev.run(p, 0.0, nc, learned=False)
        from quspin_amd import _lib
        from oracle import c_oracle
        err = float(np.abs(_lib.to_host(p) - F).max())
        res["cpu_baseline"] = {"value": round(1.0 / sec, 4), "unit": "steps/s", "cores": c_oracle.max_threads(),
                               "kind": "port",
                               "sample": f"{nc} FP64 steps on the host: 1-norm pass, _fragment_3_1 and the C/OpenMP "
                                         "Taylor loop on the real CSR of H0 plus f(t) D (oracle/csrc/qsx_oracle.c "
                                         f"qo_rcsr_1_norm, qo_expm_rcsr_c128); max|GPU-CPU| after them = {err:.2e}"}
    return res


if __name__ == "__main__":
    main()
