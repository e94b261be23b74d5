#!/usr/bin/env python3
"""Benchmark of the HDG Newton-GMRES path on one MI355X.

Workload (default): steady convection-diffusion on the unit cube, 32^3 hex mesh,
p = 4, two components, velocity (1, 0.5, 0.25), diffusivity 1e-2, seeded
uniform[-1, 1] nodal source, Dirichlet u = 0 on x = 0 and homogeneous Neumann
elsewhere, GMRES(100) to 1e-10 with classical Gram-Schmidt (GMRESortho = 1)
preconditioned by block-ILU(0) over 2x2x2-element superelements.  One "step" is one full solve from the input
state: hdgGetQ, residual, assembly + static condensation, GMRES, hdgGetDUDG,
residual (one Newton step, the problem is linear), with every input already
resident in HBM.

Two timed legs run on the same solver: the plain FP64 form and the mixed form
(FP32 storage of the condensed operator and the local-solve factors, BF16 storage
of the block-ILU(0) factors -- ``--precond-fp32`` keeps them in FP32; FP64
residuals, LU pivots and GMRES vectors; iterative refinement to 1e-10 of the
first residual).  ``value`` is the mixed form's
solves/s; the FP64 form, the ratio and the mixed solution's relative error
against the FP64 one are reported beside it.

The same line carries ``batch``: 256 parametrized 2D convection-diffusion instances of uneven
size and order (see ``run_batch``) solved as four batches (one per polynomial order, shared
launches, per-instance Newton/GMRES), in the mixed and FP64 forms, next to the plain FP64 form
solving them one at a time; and ``batch3d``: the same comparison on 48 parametrized 3D
convection-diffusion instances (n^3 hexes, n in {4..12}, p in {2, 3, 4}).  ``--workload batch``
and ``--workload batch3d`` print one of them alone, and ``--no-batch`` leaves both out.

Prints one JSON line (rank 0, N = 1).  ``roofline`` is the dominant kernel of
the GMRES loop, the element matvec of the leg reported in ``value``:
algorithmic bytes per launch (s_H*m^2*ne for H, with s_H = 4 or 8, 8*m*ne for
the gathered vector, 4*(m/ncu)*ne for the connectivity, 8*m*ne for the result)
over its average launch time measured with HIP events on the solver's stream
during the timed solves.  ``cpu_baseline`` is the numpy oracle on one core,
timed on a bounded sample of the same workload (see ``cpu_sample``) and scaled
to one full FP64 solve.
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
MFMA_PEAK_FP32 = 157.3   # TFLOP/s, dense
MFMA_PEAK_FP64 = 78.6
PREC_NAME = {0: "fp64", 1: "mixed"}
PREC_ID = {"fp64": 0, "mixed": 1}


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "convdiff3d_32cube_p4_hbm_traffic.md")


def measured_traffic():
    """HBM bytes per launch of the matvec and the block-ILU(0) apply on the 32^3 p=4 workload, from the
    two counter passes (FETCH_SIZE, WRITE_SIZE) summarised in the profile file, each scaled by what the
    counter reads for an elementwise copy of known size. Empty when the file is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return {k: float(v) for k, v in json.loads(line.split(":", 1)[1]).items()}
    except OSError:
        pass
    return {}


def build_case(path, n, porder, precond, restart, tol, maxit, se):
    from exasim_amd import cases
    cart = [3, n, n, n, se, se, se, 0] if precond == "bilu0" else None
    return cases.convdiff3d(path, n=n, porder=porder, ncu=2, kappa=1e-2, vel=(1.0, 0.5, 0.25),
                            tau=1.0, precond=precond, restart=restart, gmres_tol=tol,
                            gmres_maxit=maxit, newton_iter=1, newton_tol=1e-8, cartgridpart=cart)


def cpu_sample(porder, restart, n_sample, precond, se):
    """Child process: time the oracle's full solve (assembly and condensation, preconditioner
    set-up, preconditioned GMRES to the workload's tolerance) on a reduced mesh of the workload."""
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
        t0 = time.perf_counter()
        _, its = ora.gmres(lambda q: P.matvec(S["H"], q), prec, b, P.restart, P.gmres_maxit, P.gmres_tol,
                           P.newton_tol, P.ortho)
        t_gmres = time.perf_counter() - t0
    return dict(ne=int(P.ne), N=int(b.size), t_asm=float(t_asm), t_pre=float(t_pre), t_gmres=float(t_gmres),
                its=int(its), t_iter=float(t_gmres / max(its, 1)))


def cpu_threads():
    """Host cores this process may use (at most 16, the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def batch_cpu_sample(budget_s=20.0, nmax=32):
    """Child process: the oracle (1 thread) on the batch workload's instances with n <= nmax, in
    instance order, until about budget_s seconds of solves; returns per-instance seconds and
    trace unknowns."""
    from exasim_amd import cases
    from oracle import hdg as ora
    out = []
    with tempfile.TemporaryDirectory() as d:
        inst = cases.convdiff2d_batch(d, count=256, select=lambda i, n, p: n <= nmax and i < 96, **BATCH_KW)
        spent = 0.0
        for info in inst:
            P = ora.HDGProblem(info["path"], "cd2d")
            t0 = time.perf_counter()
            ora.newton_solve(P)
            t = time.perf_counter() - t0
            spent += t
            out.append(dict(t=t, N=int(P.nf * P.npf * P.ncu), p=info["porder"], n=info["n"]))
            if spent > budget_s:
                break
    return out


# solver settings of the batch workload's instances (GMRESortho = 1: classical Gram-Schmidt)
BATCH_KW = dict(precond="eas", restart=100, gmres_tol=1e-10, newton_tol=1e-8, ortho=1)


# mixed-form refinement settings of the batch legs.  3D: the C2 line's (first solve and corrections
# to 1e-5).  2D: its convection-dominated instances spend hundreds of iterations in GMRES(100)'s
# plateau, which a correction solved from scratch pays again, so each solve runs one restart cycle
# (refine_cycles = 1) and the next correction restarts from the FP64 residual (measured: the 2D
# batch's four groups one after the other 0.83 s -> 0.68 s; the 3D batch, whose solves end within a
# cycle, 0.39 s -> 0.44 s, so it keeps the 1e-5 settings)
# the mixed batch legs' inverted Schwarz blocks in BF16 (option precond_bf16) or FP32; measured: 3D
# batch 140.6 -> 147.5 solves/s with BF16 (same GMRES counts, error 2.2e-9), 2D batch 489 -> 476 (its
# slowest instance needs 700 instead of 600 GMRES iterations), so the 2D leg keeps FP32 blocks
BATCH_PRECOND_BF16 = {"2d": 0, "3d": 1}

BATCH_REFINE = {"2d": dict(refine_first_tol=1e-7, refine_tol=1e-7, refine_cycles=1, refine_maxit=40),
                "3d": dict(refine_first_tol=1e-5, refine_tol=1e-5)}


def run_batch(steps, warmup, cpu_baseline=True, count=256, verbose=True, concurrent=True, kind="2d",
              precond_bf16=None):
    """The batch workload: ``count`` parametrized 2D convection-diffusion instances (n x n quads,
    n in {8..64}, p in {2..5}, kappa log-uniform [1e-3, 1], velocity angle uniform, source
    sin(k pi x) cos(k pi y)), one batch handle per polynomial order, the four batches solved
    concurrently (one host thread and one HIP stream each).  A step solves every instance from
    its input state.  Reported: batched mixed and FP64 solves/s, the plain FP64 form (one
    instance at a time through exa_create) on the same instances, per-instance time to solution,
    and the mixed results' largest relative error against the plain FP64 ones."""
    import collections
    import threading
    from exasim_amd import cases
    from exasim_amd.hdg import HDGSolver

    tag = "batch" if kind == "2d" else "batch3d"

    def log(msg):
        if verbose:
            print("# %s: %s" % (tag, msg), file=sys.stderr, flush=True)

    work = tempfile.mkdtemp(prefix="exa_batch_")
    t0 = time.perf_counter()
    if kind == "2d":
        inst = cases.convdiff2d_batch(work, count=count, **BATCH_KW)
        model = "cd2d"
    else:
        inst = cases.convdiff3d_batch(work, count=count)
        model = "convdiff"
    log("%d instances written in %.1f s" % (count, time.perf_counter() - t0))
    groups = collections.defaultdict(list)
    for i, d in enumerate(inst):
        groups[d["porder"]].append(i)
    # plain FP64 form: each instance alone, inputs resident (set-up outside the timed solve; an
    # untimed solve first, so that the GEMM tile tuner has timed the instance's product shapes)
    plain_t, uh_ref = [], {}
    for i, d in enumerate(inst):
        with HDGSolver(d["path"], model) as S:
            S.solve()
            S.reset()
            S.sync()
            t1 = time.perf_counter()
            S.solve()
            plain_t.append(time.perf_counter() - t1)
            uh_ref[i] = S.get("uh")
    plain_t = np.array(plain_t)
    log("plain FP64: %.1f solves/s (sum of %d solve times %.2f s)" % (count / plain_t.sum(), count, plain_t.sum()))
    legs = {}
    for prec in (0, 1):
        hs = {}
        for p, ids in sorted(groups.items()):
            hs[p] = HDGSolver.batch([inst[i]["path"] for i in ids], model)
            hs[p].set_option("precision", prec)
            if prec:
                for k, v in BATCH_REFINE[kind].items():
                    hs[p].set_option(k, v)
                bf16 = BATCH_PRECOND_BF16[kind] if precond_bf16 is None else precond_bf16
                hs[p].set_option("precond_bf16", int(bf16))
        lat = collections.defaultdict(list)
        walls = []
        for k in range(warmup + steps):
            for h in hs.values():
                h.reset()
                h.sync()
            done = {}
            t1 = time.perf_counter()

            def go(p, h):
                h.solve()
                done[p] = time.perf_counter() - t1
            if concurrent:
                th = [threading.Thread(target=go, args=(p, h)) for p, h in hs.items()]
                for t in th:
                    t.start()
                for t in th:
                    t.join()
            else:  # one batch after the other on one host thread
                for p, h in hs.items():
                    go(p, h)
            wall = time.perf_counter() - t1
            if k >= warmup:
                walls.append(wall)
                for p, ids in groups.items():
                    lat[p].append(done[p])
            log("%s step %d: %.3f s (%s)" % (PREC_NAME[prec], k, wall,
                                              ", ".join("p=%d %.3f s" % (p, done[p]) for p in sorted(done))))
        err, gm = 0.0, []
        for p, ids in groups.items():
            a, g, nrm = hs[p].instance_stats()
            gm += list(g)
            for j, i in enumerate(ids):
                u = hs[p].get_instance(j, "uh")
                err = max(err, float(np.abs(u - uh_ref[i]).max() / np.abs(uh_ref[i]).max()))
        # per-instance time to solution: its batch's completion time
        tts = np.concatenate([np.repeat(np.mean(lat[p]), len(ids)) for p, ids in groups.items()])
        tm = {p: h.timings() for p, h in hs.items()}
        for h in hs.values():
            h.close()
        legs[prec] = dict(solves_s=count / float(np.mean(walls)), wall_s=float(np.mean(walls)), err=err,
                          gmres_mean=float(np.mean(gm)), gmres_max=int(max(gm)),
                          p50_ms=float(np.percentile(tts, 50) * 1e3), p99_ms=float(np.percentile(tts, 99) * 1e3),
                          group_s={int(p): float(np.mean(v)) for p, v in lat.items()})
    cpu = None
    if cpu_baseline and kind == "2d":
        env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--batch-cpu-child"], env=env,
                           capture_output=True, text=True, timeout=900)
        if r.returncode == 0:
            c = json.loads(r.stdout.strip().splitlines()[-1])
            rate = sum(x["t"] for x in c) / sum(x["N"] for x in c)      # oracle seconds per trace unknown
            total_N = sum(int(np.sum(uh_ref[i].size)) for i in range(count))
            cpu = {"value": count / (rate * total_N), "unit": "solves/s", "cores": 1, "kind": "port",
                   "sample": "numpy oracle (FP64, 1 thread), full Newton-GMRES solves of the first %d "
                             "instances with n <= 32 (%.1f s), scaled by trace unknowns to the %d instances"
                             % (len(c), sum(x["t"] for x in c), count)}
        else:
            log("cpu baseline failed: " + r.stderr[-1500:])
    plain = dict(solves_s=count / float(plain_t.sum()), p50_ms=float(np.percentile(plain_t, 50) * 1e3),
                 p99_ms=float(np.percentile(plain_t, 99) * 1e3))
    return dict(instances=count, groups={int(p): len(v) for p, v in groups.items()}, plain_fp64=plain,
                batched_fp64=legs[0], batched_mixed=legs[1],
                mixed_batched_over_plain_fp64=legs[1]["solves_s"] / plain["solves_s"],
                fp64_batched_over_plain_fp64=legs[0]["solves_s"] / plain["solves_s"],
                mixed_batched_over_fp64_batched=legs[1]["solves_s"] / legs[0]["solves_s"], cpu_baseline=cpu,
                mixed_refinement=dict(BATCH_REFINE[kind], precond_bf16=int(
                    BATCH_PRECOND_BF16[kind] if precond_bf16 is None else precond_bf16)))


def run_serve(seconds, rate, check_alone=6, refill=True, lu_inverses=1):
    """The served workload: a Poisson stream of Navier-Stokes (70 %) and convection-diffusion
    solve requests (n x n quads, n in {32, 64, 128}, p in {2..5}; see cases.serve_stream)
    replayed in real time into the continuous-batching server (HBM budget 200 GB, mixed
    precision; a finished request's slot takes a waiting request of the same mesh between Newton
    iterations or time steps).  Reports completed solves/s over the stream, p50/p99 time to
    solution (arrival to result), the number of requests admitted into running batches, and
    whether a sample of requests gets bitwise the same result alone."""
    from exasim_amd import cases, serve
    from exasim_amd.hdg import HDGSolver
    work = tempfile.mkdtemp(prefix="exa_serve_")
    t0 = time.perf_counter()
    reqs = cases.serve_stream(work, seconds=seconds, rate=rate)
    print("# serve: %d requests written in %.1f s" % (len(reqs), time.perf_counter() - t0), file=sys.stderr,
          flush=True)
    srv = serve.Server(hbm_budget_bytes=200e9, max_batch=64, max_concurrent=8, precision=1, refill=refill,
                       options={} if lu_inverses else {"lu_inverses": 0},
                       keep_results=check_alone > 0, window=0.02,
                       log=lambda msg: print("# serve: " + msg, file=sys.stderr, flush=True))
    summ = srv.run_stream(reqs)
    print("# serve: %s" % json.dumps({k: summ[k] for k in ("completed", "solves_per_s", "p50_ms", "p99_ms", "refills")}),
          file=sys.stderr, flush=True)
    same, err64, by_model = [], 0.0, {}
    for rid, path, model, t in reqs[:check_alone]:
        if rid not in srv.results:
            continue
        with HDGSolver.batch([path], model) as S:
            S.set_option("precision", srv.records[rid].get("precision", 1))
            S.solve()
            same.append(bool(np.array_equal(S.get_instance(0, "uh"), srv.results[rid])))
        with HDGSolver.batch([path], model) as S:     # the same request in the plain FP64 form
            S.set_option("precision", 0)
            S.solve()
            u64 = S.get_instance(0, "uh")
        e = float(np.abs(srv.results[rid] - u64).max() / np.abs(u64).max())
        err64 = max(err64, e)
        by_model[model] = max(by_model.get(model, 0.0), e)
    kinds = {}
    for rid, path, model, t in reqs:
        kinds[model] = kinds.get(model, 0) + 1
    summ.update(stream_seconds=seconds, rate_per_s=rate, mix=kinds, alone_equal_sample=same,
                rel_err_vs_fp64_sample=err64 if same else None, rel_err_vs_fp64_by_model=by_model)
    return summ


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
    ap.add_argument("--ortho", type=int, default=1,
                    help="GMRESortho: 0 modified, 1 classical Gram-Schmidt (default; fewer passes over the basis)")
    ap.add_argument("--refine-first-tol", type=float, default=1e-5,
                    help="mixed form: relative tolerance of the first FP32 solve (measured best at 32^3 p=4)")
    ap.add_argument("--refine-tol", type=float, default=1e-5,
                    help="mixed form: relative tolerance of each refinement correction")
    ap.add_argument("--precond-fp32", action="store_true",
                    help="mixed leg: block-ILU(0) factors in FP32 (default: BF16, same iteration counts)")
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"],
                    help="timed legs: FP64, mixed FP32/FP64, or both (value = the mixed leg)")
    ap.add_argument("--precond", default="bilu0", choices=["bilu0", "bjacobi"])
    ap.add_argument("--se", type=int, default=2, help="superelement edge (elements) of block-ILU(0)")
    ap.add_argument("--workspace-gb", type=float, default=None,
                    help="assembly workspace budget in GB (the library's default: 16)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy a MIB-MiB device buffer 3 times with an elementwise kernel "
                         "(the scale for HBM byte counters)")
    ap.add_argument("--batch-serial", action="store_true",
                    help="solve the batch workload's per-order batches one after the other (default: concurrently)")
    ap.add_argument("--gmres-graph", type=int, default=0, choices=[0, 1],
                    help="1: replay the GMRES steps from captured HIP graphs; 0 (default): launch them one by one")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-n", type=int, default=4)
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--batch-cpu-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--serve-seconds", type=float, default=1.0, help="length of the served request stream")
    ap.add_argument("--no-serve", action="store_true", help="skip the served stream in the c2 line")
    ap.add_argument("--serve-rate", type=float, default=50.0, help="request arrivals per second")
    ap.add_argument("--gemm-autotune", type=int, default=1,
                    help="1: time the GEMM workgroup tiles per product shape (first call) and keep the fastest; "
                         "0: the default 64x64 tile")
    ap.add_argument("--serve-lu-inverses", type=int, default=1, choices=[0, 1],
                    help="served stream, mixed form: 1 FP32 LU with triangle inverses, 0 by substitution")
    ap.add_argument("--serve-refill", type=int, default=1,
                    help="1: running batches admit waiting requests between steps; 0: each batch solves to the end")
    ap.add_argument("--workload", default="c2", choices=["c2", "batch", "batch3d", "serve"],
                    help="c2: the steady 3D convection-diffusion solve (default, with the batch "
                         "workloads' results as sub-objects); batch: the 2D batch workload alone; "
                         "batch3d: the 3D convection-diffusion batch alone")
    ap.add_argument("--no-batch", action="store_true", help="skip the batch workload in the c2 line")
    ap.add_argument("--batch-precond-bf16", type=int, default=None, choices=[0, 1],
                    help="mixed batch legs: inverted Schwarz blocks in BF16 (1) or FP32 (0); default per workload")
    args = ap.parse_args()
    os.environ["EXA_GEMM_AUTOTUNE"] = str(args.gemm_autotune)   # read by the library at its first GEMM

    if args.cpu_child:
        print(json.dumps(cpu_sample(args.porder, args.restart, args.cpu_sample_n, args.precond, args.se)))
        return 0
    if args.batch_cpu_child:
        print(json.dumps(batch_cpu_sample()))
        return 0
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (the domain "
              "decomposition design is described in DESIGN.md); use --gpus 1", file=sys.stderr)
        return 2
    from exasim_amd import hdg as _hdg
    _hdg.DEFAULT_OPTIONS["gmres_graph"] = args.gmres_graph

    from exasim_amd.hdg import HDGSolver

    if args.workload == "serve":
        sv = run_serve(args.serve_seconds, args.serve_rate, refill=bool(args.serve_refill),
                       lu_inverses=args.serve_lu_inverses)
        print(json.dumps({
            "metric": METRIC, "value": sv["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": 1,
            "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": "served stream: Poisson arrivals %.0f/s for %.0f s, 70%% Navier-Stokes "
                                   "(2 DIRK(2,2) steps of 5e-3) / 30%% convection-diffusion, n in {32,64,128}, p in "
                                   "{2..5}, HBM budget 200 GB, mixed precision" % (args.serve_rate,
                                                                                   args.serve_seconds)},
            "p50_ms": sv["p50_ms"], "p99_ms": sv["p99_ms"], "serve": sv, "roofline": None, "cpu_baseline": None}))
        return 0
    if args.workload == "batch3d":
        b = run_batch(args.steps, args.warmup, cpu_baseline=False, count=48, kind="3d",
                      concurrent=not args.batch_serial, precond_bf16=args.batch_precond_bf16)
        m = b["batched_mixed"]
        print(json.dumps({
            "metric": METRIC, "value": m["solves_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": m["wall_s"] * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": "batch of 48 parametrized 3D convection-diffusion solves (n^3 hexes, n in "
                                   "{4,6,8,12}, p in {2,3,4}, 2 components, kappa log-uniform [0.01, 0.1]), "
                                   "one batch per p, mixed FP32/FP64 with refinement, elemental Schwarz, "
                                   "GMRES(100) CGS tol 1e-10"},
            "p50_ms": m["p50_ms"], "p99_ms": m["p99_ms"], "batch3d": b, "roofline": None, "cpu_baseline": None,
            "gemm_autotune": gemm_summary()}))
        return 0
    if args.workload == "batch":
        b = run_batch(args.steps, args.warmup, cpu_baseline=not args.no_cpu_baseline,
                      concurrent=not args.batch_serial, precond_bf16=args.batch_precond_bf16)
        m = b["batched_mixed"]
        print(json.dumps({
            "metric": METRIC, "value": m["solves_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": m["wall_s"] * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": "batch of 256 parametrized 2D convection-diffusion solves (n in {8..64}, "
                                   "p in {2..5}), one batch per p, mixed FP32/FP64 with refinement, "
                                   "elemental Schwarz, GMRES(100) CGS tol 1e-10"},
            "p50_ms": m["p50_ms"], "p99_ms": m["p99_ms"], "batch": b, "roofline": None,
            "cpu_baseline": b["cpu_baseline"], "gemm_autotune": gemm_summary()}))
        return 0

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
    if args.refine_first_tol is not None:
        S.set_option("refine_first_tol", args.refine_first_tol)
    if args.refine_tol is not None:
        S.set_option("refine_tol", args.refine_tol)
    if not args.precond_fp32:
        S.set_option("precond_bf16", 1)
    if args.workspace_gb is not None:
        S.set_option("workspace_bytes", args.workspace_gb * 1e9)
    t_setup = time.perf_counter() - t0
    z = S.sizes
    print("# setup (read, connectivity, geometry, q-equation) %.2f s; sizes %s" % (t_setup, z),
          file=sys.stderr, flush=True)
    ne, m, n, ncu, N = z["ne"], z["m"], z["n"], z["ncu"], z["N"]
    traffic = measured_traffic() if (args.n, args.porder, args.se, args.precond) == (32, 4, 2, "bilu0") else {}

    def leg(precision):
        """W untimed + K timed solves from the input state at one precision."""
        S.set_option("precision", precision)
        S.set_option("timing", 0)
        for w in range(args.warmup):
            S.reset()
            it = S.solve()
            print("# %s warmup %d: newton %d gmres %d final residual %.3e"
                  % (PREC_NAME[precision], w, *it), file=sys.stderr, flush=True)
        S.sync()
        S.set_option("timing", 1)
        S.reset_timings()
        lat, gm, res, refs = [], [], [], []
        t_all = time.perf_counter()
        for k in range(args.steps):
            S.reset()
            t0 = time.perf_counter()
            ni, gi, nrm = S.solve()     # synchronises at the end
            lat.append(time.perf_counter() - t0)
            gm.append(gi)
            res.append(nrm)
            nref = S.state()["refinements"]
            refs.append(nref)
            print("# %s step %d: %.3f s  newton %d gmres %d refinements %d residual %.3e"
                  % (PREC_NAME[precision], k, lat[-1], ni, gi, nref, nrm), file=sys.stderr, flush=True)
        t_all = time.perf_counter() - t_all
        tm = S.timings()
        uh = S.get("uh")
        mv_ms, mv_cnt = tm["matvec"]
        mv_avg_s = mv_ms / max(mv_cnt, 1) / 1e3
        sH = 4.0 if precision else 8.0
        bytes_mv = sH * m * m * ne + 8.0 * m * ne + 4.0 * (m // ncu) * ne + 8.0 * m * ne
        achieved = bytes_mv / mv_avg_s / 1e9 if mv_cnt else None
        lat_ms = np.array(lat) * 1e3
        kmv = "k_matvec_elem<%s>" % ("float" if precision else "double")
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic.get(kmv),
                "kernel": kmv, "bytes_per_launch": bytes_mv, "avg_launch_ms": mv_avg_s * 1e3}
        st = S.state()
        pc_roof = None
        if st["preconditioner"] == 2 and st["ilu_nnz"] > 0:
            # the block-ILU(0) apply: its factored blocks are read once per application
            pc_ms, pc_cnt = tm["precond"]
            ncf = z["npf"] * ncu
            s_f = (2.0 if not args.precond_fp32 else 4.0) if precision else 8.0
            bytes_pc = s_f * ncf * ncf * st["ilu_nnz"] * st["ilu_nse"]
            pc_avg_s = pc_ms / max(pc_cnt, 1) / 1e3
            bf16 = precision and not args.precond_fp32
            wave = bf16 and ncf == 50 and os.environ.get("EXA_BILU_WAVE", "1") != "0"  # the library's choice
            kpc = "k_bilu_apply%s<%s>" % ("_wave" if wave else "",
                                          ("bf16" if bf16 else "float") if precision else "double")
            ach = bytes_pc / pc_avg_s / 1e9 if pc_cnt else None
            pc_roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": (ach / HBM_PEAK_GBS) if ach else None, "traffic": traffic.get(kpc),
                       "kernel": kpc, "bytes_per_launch": bytes_pc, "avg_launch_ms": pc_avg_s * 1e3}
        return dict(solves_s=args.steps / t_all, lat_ms=lat_ms, gm=gm, res=res, tm=tm, uh=uh, refs=refs,
                    roofline=roof, precond_roofline=pc_roof)

    legs = {}
    for p in ([0, 1] if args.precision == "both" else [PREC_ID[args.precision]]):
        legs[p] = leg(p)
    S.close()
    # the C2 legs' products only (before the batch workloads); FP32 ones when the mixed leg is `value`
    mfma = mfma_roofline(gemm_summary(top=40), fp32=1 if 1 in legs else 0)
    head = legs[1] if 1 in legs else legs[0]
    err = None
    if 0 in legs and 1 in legs:
        u64, u32 = legs[0]["uh"], legs[1]["uh"]
        err = float(np.abs(u32 - u64).max() / np.abs(u64).max())
    gm = head["gm"]
    lat_ms = head["lat_ms"]
    tm = head["tm"]
    solves_s = head["solves_s"]

    cpu = None
    if not args.no_cpu_baseline:
        # the oracle is one-threaded numpy: the host's cores each run one copy of the sample at once,
        # and the aggregate rate is the sum of the copies' rates (contention for memory included)
        nth = cpu_threads()
        env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", "--porder", str(args.porder),
               "--restart", str(args.restart), "--cpu-sample-n", str(args.cpu_sample_n), "--precond",
               args.precond, "--se", str(args.se)]
        t_wall = time.perf_counter()
        procs = [subprocess.Popen(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
                 for _ in range(nth)]
        outs = [pr.communicate(timeout=900) for pr in procs]
        t_wall = time.perf_counter() - t_wall
        ok = [json.loads(o.strip().splitlines()[-1]) for (o, e), pr in zip(outs, procs) if pr.returncode == 0]
        r = subprocess.CompletedProcess(cmd, 0 if ok else 1, "", "" if ok else outs[0][1])
        if r.returncode == 0:
            iters = float(np.mean(legs[0]["gm"] if 0 in legs else gm))
            rates = [1.0 / ((c["t_asm"] + c["t_pre"]) / c["ne"] * ne + iters * c["t_iter"] / c["N"] * N) for c in ok]
            c = ok[0]
            cpu = {"value": float(np.sum(rates)), "unit": "solves/s", "cores": len(ok), "kind": "port",
                   "sample": "numpy oracle (FP64), %d one-threaded copies at once, one per host core, each a "
                             "full solve of the %d^3 mesh (%d elements, %d unknowns): assembly+condensation, "
                             "%s set-up and preconditioned GMRES(%d) to 1e-10 (%d iterations); copy 0: "
                             "%.2f s + %.2f s + %.2f s (%.4f s per iteration); each copy's time scaled "
                             "linearly to %d elements and %d unknowns at the GPU FP64 form's %.0f GMRES "
                             "iterations, the rates summed; %.1f s of wall time"
                             % (len(ok), args.cpu_sample_n, c["ne"], c["N"], args.precond, args.restart,
                                c["its"], c["t_asm"], c["t_pre"], c["t_gmres"], c["t_iter"], ne, N, iters, t_wall),
                   "per_core_value": float(np.mean(rates))}
        else:
            print("# cpu baseline failed: " + r.stderr[-2000:], file=sys.stderr)

    mixed = 1 in legs
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
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": "steady convection-diffusion 3D, %d^3 hex, p=%d, 2 components, "
                               "kappa=1e-2, velocity (1,0.5,0.25), GMRES(%d) tol %g %s, %s, %s"
                               % (args.n, args.porder, args.restart, args.tol,
                                  "CGS" if args.ortho else "MGS",
                                  "block-ILU(0) over %d^3-element superelements" % args.se
                                  if args.precond == "bilu0" else "block-Jacobi face preconditioner",
                                  ("mixed FP32 storage (%s block-ILU(0) factors) / FP64 refinement to 1e-10 "
                                   "(first solve %g, corrections %g)"
                                   % ("FP32" if args.precond_fp32 else "BF16", args.refine_first_tol,
                                      args.refine_tol)) if mixed else "FP64"),
                   "ne": ne, "trace_dofs": N, "n": n, "m": m,
                   "precision": "mixed" if mixed else "fp64"},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "dof_per_s": N * solves_s,
        "newton_steps_per_s": solves_s,
        "gmres_iterations": gm,
        "refinements": head["refs"],
        "final_residual": head["res"],
        "kernel_ms": {k: {"total_ms": v[0], "launches": v[1]} for k, v in tm.items()},
        "roofline": head["roofline"],
        "precond_roofline": head["precond_roofline"],
        "mfma_roofline": mfma,
        "cpu_baseline": cpu,
    }
    if 0 in legs:
        L = legs[0]
        out["fp64"] = {"solves_per_s": L["solves_s"], "p50_ms": float(np.percentile(L["lat_ms"], 50)),
                       "p99_ms": float(np.percentile(L["lat_ms"], 99)), "gmres_iterations": L["gm"],
                       "final_residual": L["res"], "roofline": L["roofline"],
                       "precond_roofline": L["precond_roofline"],
                       "kernel_ms": {k: {"total_ms": v[0], "launches": v[1]} for k, v in L["tm"].items()}}
    if mixed and 0 in legs:
        out["mixed_over_fp64"] = solves_s / legs[0]["solves_s"]
        out["mixed_rel_err_vs_fp64"] = err
    if not args.no_batch:
        try:
            out["batch"] = run_batch(2, 1, cpu_baseline=not args.no_cpu_baseline)
        except Exception as e:          # the c2 line is still printed
            out["batch"] = {"error": repr(e)}
        try:
            out["batch3d"] = run_batch(2, 1, cpu_baseline=False, count=48, kind="3d")
        except Exception as e:
            out["batch3d"] = {"error": repr(e)}
    if not args.no_serve:
        try:
            out["serve"] = run_serve(args.serve_seconds, args.serve_rate, refill=bool(args.serve_refill),
                                     lu_inverses=args.serve_lu_inverses)
        except Exception as e:
            out["serve"] = {"error": repr(e)}
    out["gemm_autotune"] = gemm_summary()
    print(json.dumps(out))
    return 0


def gemm_summary(top=6):
    """The GEMM tile autotuner's record for this run (exa_gemm_tuning): shapes timed, the tiles
    chosen, and the time of the chosen tile against the default 64x64 tile, summed over the shapes
    (each shape once); the largest shapes listed."""
    from exasim_amd.hdg import gemm_tuning
    rows = gemm_tuning()
    if not rows:
        return {"shapes": 0, "mode": os.environ.get("EXA_GEMM_AUTOTUNE", "1")}
    hist = {}
    for r in rows:
        hist[r["tile"]] = hist.get(r["tile"], 0) + 1
    def best(r):      # (-1: a tile the product cannot take)
        return min(x for x in r["us"] if x > 0.0)
    t_def = sum(r["us"][0] for r in rows)
    t_best = sum(best(r) for r in rows)
    big = sorted(rows, key=lambda r: -r["us"][0])[:top]

    def tflops(r):    # 2 M N K per product over the chosen tile's time
        return 2.0 * r["M"] * r["N"] * r["K"] * r["batch"] / (best(r) * 1e-6) / 1e12
    return {"shapes": len(rows), "tiles_chosen": hist, "default_us_sum": t_def, "chosen_us_sum": t_best,
            "default_over_chosen": t_def / t_best if t_best > 0 else None,
            "largest": [{"M": r["M"], "N": r["N"], "K": r["K"], "batch": r["batch"], "fp32": r["fp32"],
                         "tile": r["tile"], "us_default": r["us"][0], "us_chosen": best(r),
                         "tflops": tflops(r),
                         "mfma_frac": tflops(r) / (MFMA_PEAK_FP32 if r["fp32"] else MFMA_PEAK_FP64)}
                        for r in big]}


def mfma_roofline(summary, fp32=None):
    """The MFMA roofline of the largest product shape of the line's GEMMs (of FP32 or FP64 products
    when fp32 is given): TFLOP/s of the chosen tile in the tuner's timing against the dense MFMA peak
    of the product's type."""
    big = [r for r in (summary.get("largest") or []) if fp32 is None or r["fp32"] == fp32]
    if not big:
        return None
    r = big[0]
    return {"bound": "mfma", "achieved": r["tflops"], "peak": MFMA_PEAK_FP32 if r["fp32"] else MFMA_PEAK_FP64,
            "unit": "TFLOP/s", "frac": r["mfma_frac"], "traffic": None,
            "kernel": "k_gemm_mfma<%s> %dx%dx%d, %d products, tile %s" % (
                "float" if r["fp32"] else "double", r["M"], r["N"], r["K"], r["batch"], r["tile"]),
            "flop_per_launch": 2.0 * r["M"] * r["N"] * r["K"] * r["batch"], "avg_launch_ms": r["us_chosen"] / 1e3}


if __name__ == "__main__":
    sys.exit(main())
