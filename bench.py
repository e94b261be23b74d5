#!/usr/bin/env python3
"""Benchmark of the HDG Newton-GMRES path on one MI355X.

Workload (default): steady convection-diffusion on the unit cube, 32^3 hex mesh,
p = 4, two components, velocity (1, 0.5, 0.25), diffusivity 1e-2, seeded
uniform[-1, 1] nodal source, Dirichlet u = 0 on x = 0 and homogeneous Neumann
elsewhere, GMRES(100) to 1e-10 (FP64) preconditioned by block-ILU(0) over
2x2x2-element superelements.  One "step" is one full solve from the
input state: hdgGetQ, residual, assembly + static condensation, GMRES,
hdgGetDUDG, residual (one Newton step, the problem is linear), with every input
already resident in HBM.

Prints one JSON line (rank 0, N = 1).  ``roofline`` is the dominant kernel of
the GMRES loop, the element matvec: algorithmic bytes per launch
(8*m^2*ne for H, 8*m*ne for the gathered vector, 4*(m/ncu)*ne for the
connectivity, 8*m*ne for the result) over its average launch time measured
with HIP events on the solver's stream during the timed solves.
``cpu_baseline`` is the numpy oracle on one core, timed on a bounded sample of
the same workload (see ``cpu_sample``) and scaled to one full solve.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s + p50/p99 latency, mixed vs FP64 GPU at rel err ≤1e-8; HBM/MFMA %"
HBM_PEAK_GBS = 8000.0


def build_case(path, n, porder, precond, restart, tol, maxit, se):
    from exasim_amd import cases
    cart = [3, n, n, n, se, se, se, 0] if precond == "bilu0" else None
    return cases.convdiff3d(path, n=n, porder=porder, ncu=2, kappa=1e-2, vel=(1.0, 0.5, 0.25),
                            tau=1.0, precond=precond, restart=restart, gmres_tol=tol,
                            gmres_maxit=maxit, newton_iter=1, newton_tol=1e-8, cartgridpart=cart)


def cpu_sample(porder, restart, n_sample, precond, se):
    """Child process: time the oracle on one core on a reduced mesh of the workload."""
    from oracle import hdg as ora
    with tempfile.TemporaryDirectory() as d:
        build_case(d, n_sample, porder, precond, restart, 1e-10, 100000, min(se, n_sample))
        P = ora.HDGProblem(d, "convdiff")
        P.get_q(P.udg, P.uh)
        t0 = time.perf_counter()
        E, S, b = P.assemble()
        t_asm = time.perf_counter() - t0
        t0 = time.perf_counter()
        prec = P.make_preconditioner(S["H"])
        t_pre = time.perf_counter() - t0
        # GMRES iteration cost: matvec + preconditioner + MGS against restart/2 vectors
        N = b.size
        V = np.random.default_rng(0).standard_normal((restart // 2, N))
        w = b.copy()
        reps = 10
        t0 = time.perf_counter()
        for _ in range(reps):
            w = prec(P.matvec(S["H"], w / np.linalg.norm(w)))
            for k in range(V.shape[0]):
                h = np.dot(w, V[k])
                w = w - h * V[k]
        t_it = (time.perf_counter() - t0) / reps
    return dict(ne=int(P.ne), N=int(N), t_asm=float(t_asm), t_pre=float(t_pre), t_iter=float(t_it))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32, help="elements per direction")
    ap.add_argument("--porder", type=int, default=4)
    ap.add_argument("--restart", type=int, default=100)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--maxit", type=int, default=20000)
    ap.add_argument("--ortho", type=int, default=0)
    ap.add_argument("--precond", default="bilu0", choices=["bilu0", "bjacobi"])
    ap.add_argument("--se", type=int, default=2, help="superelement edge (elements) of block-ILU(0)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy a MIB-MiB device buffer 3 times with an elementwise kernel "
                         "(the scale for HBM byte counters)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-n", type=int, default=4)
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_child:
        print(json.dumps(cpu_sample(args.porder, args.restart, args.cpu_sample_n, args.precond, args.se)))
        return 0
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (the domain "
              "decomposition design is described in DESIGN.md); use --gpus 1", file=sys.stderr)
        return 2

    from exasim_amd.hdg import HDGSolver

    if args.copy_calibration:
        import torch
        a = torch.empty(args.copy_calibration * (1 << 20) // 8, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)     # elementwise kernel: reads and writes MIB MiB
        torch.cuda.synchronize()
        print("# copy calibration: 3 copies of %d MiB" % args.copy_calibration, file=sys.stderr, flush=True)
        del a, b

    work = tempfile.mkdtemp(prefix="exa_bench_")
    t0 = time.perf_counter()
    info = build_case(work, args.n, args.porder, args.precond, args.restart, args.tol, args.maxit, args.se)
    t_write = time.perf_counter() - t0
    print("# datain written in %.1f s: %s" % (t_write, info), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    S = HDGSolver(work, "convdiff")
    S.set_option("ortho", args.ortho)
    t_setup = time.perf_counter() - t0
    z = S.sizes
    print("# setup (read, connectivity, geometry, q-equation) %.2f s; sizes %s" % (t_setup, z),
          file=sys.stderr, flush=True)

    for w in range(args.warmup):
        S.reset()
        it = S.solve()
        print("# warmup %d: newton %d gmres %d final residual %.3e" % (w, *it), file=sys.stderr, flush=True)

    S.sync()
    S.set_option("timing", 1)
    S.reset_timings()
    lat = []
    gm = []
    res = []
    t_all = time.perf_counter()
    for k in range(args.steps):
        S.reset()
        t0 = time.perf_counter()
        ni, gi, nrm = S.solve()     # synchronises at the end
        lat.append(time.perf_counter() - t0)
        gm.append(gi)
        res.append(nrm)
        print("# step %d: %.3f s  newton %d gmres %d residual %.3e" % (k, lat[-1], ni, gi, nrm),
              file=sys.stderr, flush=True)
    t_all = time.perf_counter() - t_all
    tm = S.timings()
    S.close()

    ne, m, n, ncu, N = z["ne"], z["m"], z["n"], z["ncu"], z["N"]
    mv_ms, mv_cnt = tm["matvec"]
    mv_avg_s = mv_ms / max(mv_cnt, 1) / 1e3
    bytes_mv = 8.0 * m * m * ne + 8.0 * m * ne + 4.0 * (m // ncu) * ne + 8.0 * m * ne
    achieved = bytes_mv / mv_avg_s / 1e9 if mv_cnt else None
    solves_s = args.steps / t_all
    lat_ms = np.array(lat) * 1e3

    cpu = None
    if not args.no_cpu_baseline:
        env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", "--porder",
                            str(args.porder), "--restart", str(args.restart), "--cpu-sample-n",
                            str(args.cpu_sample_n), "--precond", args.precond, "--se", str(args.se)],
                           env=env, capture_output=True, text=True, timeout=900)
        if r.returncode == 0:
            c = json.loads(r.stdout.strip().splitlines()[-1])
            iters = float(np.mean(gm))
            t_solve = (c["t_asm"] + c["t_pre"]) / c["ne"] * ne + iters * c["t_iter"] / c["N"] * N
            cpu = {"value": 1.0 / t_solve, "unit": "solves/s", "cores": 1, "kind": "port",
                   "sample": "numpy oracle, 1 thread: assembly+condensation (%.2f s) and %s "
                             "set-up (%.2f s) of the %d^3 mesh (%d elements), and one GMRES "
                             "iteration (matvec, preconditioner, MGS over %d vectors; %.4f s at %d "
                             "unknowns), scaled linearly to %d elements and %d unknowns at the GPU's "
                             "%.0f GMRES iterations"
                             % (c["t_asm"], args.precond, c["t_pre"], args.cpu_sample_n, c["ne"],
                                args.restart // 2, c["t_iter"], c["N"], ne, N, iters)}
        else:
            print("# cpu baseline failed: " + r.stderr[-2000:], file=sys.stderr)

    out = {
        "metric": METRIC,
        "value": solves_s,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(lat_ms)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "C2 steady convection-diffusion 3D, %d^3 hex, p=%d, 2 components, "
                               "kappa=1e-2, velocity (1,0.5,0.25), GMRES(%d) tol %g, FP64, %s"
                               % (args.n, args.porder, args.restart, args.tol,
                                  "block-ILU(0) over %d^3-element superelements" % args.se
                                  if args.precond == "bilu0" else "block-Jacobi face preconditioner"),
                   "ne": ne, "trace_dofs": N, "n": n, "m": m, "precision": "fp64"},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "dof_per_s": N * solves_s,
        "newton_steps_per_s": solves_s,
        "gmres_iterations": gm,
        "final_residual": res,
        "kernel_ms": {k: {"total_ms": v[0], "launches": v[1]} for k, v in tm.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "kernel": "k_matvec_elem", "bytes_per_launch": bytes_mv,
                     "avg_launch_ms": mv_avg_s * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
