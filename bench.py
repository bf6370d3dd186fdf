"""Benchmark of the symmetry-reduced matrix-free H.v on one MI355X.

Workload (N=1): spin-1/2 Heisenberg chain L=28, periodic, Sz=0 + momentum
k=0 + parity + spin inversion (361,270 reduced states), float64, seeded
Gaussian vector; one step = one fused H.v launch (quantum_LinearOperator.dot).
Also timed in the same run: 200-step Lanczos on the same operator.

Prints ONE JSON line.  ``roofline`` uses the algorithmic bytes of one launch
B = Ns*(4 + 1) + 2*Ns*8 (basis word, normalisation byte, v_in, v_out; the
off-diagonal gathers of v_in count as cache hits) over the average launch
time measured with HIP events on the launch stream.  ``cpu_baseline`` is
the C/OpenMP restatement in oracle/ (fused all-terms gather) timed on the
host cores on a bounded number of matvecs.
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
# profiles/k_hv_rows_heisenberg_L28_kpz.md
MEASURED_TRAFFIC = {(28, "table", "rows", "float64"): 93.5e6}


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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build measures the single-GPU kernel "
              "(multi-GPU sharding is described in DESIGN.md, not run)", file=sys.stderr)
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

    nbytes = Ns * (4 + 1) + 2 * Ns * 8
    achieved = nbytes / (ms_per * 1e-3) / 1e9

    # 200-step Lanczos on the same operator (device recurrence)
    lz = {}
    if args.lanczos > 0:
        torch.cuda.synchronize()
        tl0 = time.perf_counter()
        E, V, Qm = lanczos_full(Q, v, args.lanczos)
        torch.cuda.synchronize()
        tl = time.perf_counter() - tl0
        lz = {"steps": int(Qm.shape[0]), "seconds": round(tl, 4), "steps_per_s": round(Qm.shape[0] / tl, 1),
              "E0": float(E[0])}
        del Qm

    cpu = None
    if not args.no_cpu:
        from oracle import c_oracle
        g = c_oracle.Group(L, [P, Z, T], [2, 2, L], [0, 0, 0])
        tc0 = time.perf_counter()
        cb, cn = g.make_basis(L // 2)
        t_cb = time.perf_counter() - tc0
        if not np.array_equal(cb, B.states.astype(np.uint64)):
            raise SystemExit("CPU restatement basis differs from the device basis")
        terms = [(op_, tuple(b[1:]), b[0]) for op_, bonds in static for b in bonds]
        vh = v.cpu().numpy()
        ref = c_oracle.hv_real(g, cb, cn, terms, vh, fused=True)  # warm-up + parity
        err = float(np.abs(ref - out.cpu().numpy()).max())
        nrep, tsum = 0, 0.0
        while tsum < args.cpu_seconds and nrep < 1000:
            tc = time.perf_counter()
            c_oracle.hv_real(g, cb, cn, terms, vh, fused=True)
            tsum += time.perf_counter() - tc
            nrep += 1
        cpu = {"value": round(nrep / tsum, 3), "unit": "matvecs/s", "cores": c_oracle.max_threads(),
               "kind": "port",
               "sample": f"{nrep} fused H.v calls on the same L={L} k=0,p=0,z=0 sector ({Ns} states), "
                         f"C/OpenMP restatement (oracle/csrc/qsx_oracle.c); max|GPU-CPU| = {err:.2e}"}

    cfg = next(iter(op.config.values()), {})
    traffic = MEASURED_TRAFFIC.get((L, cfg.get("lookup"), cfg.get("schedule"), "float64"))
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
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic},
        "cpu_baseline": cpu,
        "lanczos": lz,
        "kernel_config": {str(k): v for k, v in op.config.items()},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
