"""Serving of a stream of HDG solve requests on one GPU (continuous batching).

Requests (a datain directory and a model name) arrive over time.  They are grouped by
element shape -- (model, nd, p, quadrature, ncu), the condition for sharing launches --
into lanes.  A dispatcher starts a batch as soon as requests of a lane wait (after the
batching window) and the HBM budget admits them, while fewer than ``max_concurrent`` batches
run: it takes every waiting request of the lane that fits (up to ``max_batch``), builds one
batch handle (``exa_batch_create``: the instances' meshes, Jacobians and Krylov vectors
resident on the device for the whole solve, one HIP stream) and runs it on a thread of its own
(shared launches, per-instance Newton/GMRES, see exasim_amd/csrc/batch.cpp).

The batch advances one Newton iteration (steady requests) or one time step (DIRK requests) at a
time (``exa_batch_step``).  After each step every request that has finished gets its result at
once, and its slot is admitted the oldest waiting request of the lane with the same mesh
(``exa_batch_refill``: same element shape and mesh.bin, any coordinates, parameters and initial
state): the new request's Newton iterations run in the same launches as the requests already
under way, in the slot and memory the finished one leaves.  A request with no free slot of its
mesh is admitted into a new batch as soon as memory and a batch slot free up, and the batches run
concurrently on their own streams, so their work overlaps on the device.  An instance's result
does not depend on the batch it was solved in, nor on when it joined (the batch keeps every
Krylov/Newton scalar and the model time per instance), so a request gets the same answer alone
and in the mix.

No counterpart in the reference (one problem per process there); each solve follows
exa_solve's algorithm.
"""
import collections
import hashlib
import os
import threading
import time

import numpy as np

from .datain import read_sizes
from .hdg import ExaError, HDGSolver


def shape_key(path, model):
    """Requests with equal keys can share a batch: the model and the master element (nd, ncu,
    element type, p, quadrature)."""
    z = read_sizes(path)
    return (model, z["nd"], z["ncu"], z["elemtype"], z["porder"], z["pgauss"], z["npe"], z["npf"], z["nge"],
            z["ngf"])


def mesh_key(path):
    """Requests with equal keys (and equal shape keys) can take each other's slot in a running
    batch: the digest of mesh.bin (connectivity, boundary faces and face numbering; the node
    coordinates are in sol.bin)."""
    with open(os.path.join(path, "mesh.bin"), "rb") as f:
        return hashlib.sha1(f.read()).hexdigest()


def estimate_bytes(ne, npe, npf, nfe, ncu, nd, restart, precision=1):
    """Device bytes of one instance in a batch: condensed operator H and DinvF, q-equation
    matrices, element Schwarz blocks, FP32 copies in the mixed form, Krylov basis and the
    geometry, plus 20 % for vectors and workspace."""
    n, m = npe * ncu, npf * nfe * ncu
    nf = ne * nfe / 2.0 + 2.0 * np.sqrt(ne) * nd
    N = ncu * npf * nf
    b = 8.0 * ne * (m * m + n * m + npe * nd * npe + npe * nd * npf * nfe + m * m + n * n)
    if precision:
        b += 4.0 * ne * (m * m + n * m + n * n + m * n + m * m)
    b += 8.0 * (restart + 1) * N
    return 1.2 * b


Request = collections.namedtuple("Request", "rid datain model arrival")


class Server:
    """Continuous-batching request loop over batch handles."""

    def __init__(self, hbm_budget_bytes=200e9, max_batch=64, max_concurrent=8, precision=1, options=None,
                 keep_results=False, log=None, window=0.0, refill=True, accept_norm=1e-8):
        self.budget = float(hbm_budget_bytes)
        self.max_batch = max_batch
        self.max_concurrent = int(max_concurrent)  # batches solving at once (a host thread and a HIP stream each)
        self.running = 0
        self.threads = []
        self.dispatcher = None
        self.precision = precision
        # a request is solved when its final Newton residual norm is finite and at most accept_norm;
        # otherwise the mixed form's result is solved again in FP64, and an FP64 solve that ends
        # above it (Newton stagnated or diverged) is reported as failed
        self.accept_norm = float(accept_norm)
        self.options = dict(options or {})
        # each batch handle's assembly workspace, counted against the budget per batch
        self.options.setdefault("workspace_bytes", 1e9)
        self.keep_results = keep_results
        self.log = log           # callable(str) told about every finished batch, or None
        # batching window: a worker starts a batch once its lane's oldest waiting request is
        # this old (seconds), or max_batch requests wait, or the stream has ended
        self.window = float(window)
        self.lock = threading.Condition()
        self.queues = collections.defaultdict(collections.deque)   # lane key -> waiting requests
        self.in_flight = 0.0
        self.peak_bytes = 0.0
        self.closed = False
        self.records = {}        # rid -> dict(arrival, admit, done, batch, newton, gmres, norm)
        self.results = {}        # rid -> uh (keep_results)
        self.errors = []
        self.batches = []        # (lane, size, seconds)
        self.fallbacks = []      # (lane, size) of batches solved again in FP64
        self.t0 = None
        self._est = {}
        # iteration-level admission: batches step (exa_batch_step) and finished slots take waiting
        # requests of the same mesh (exa_batch_refill); False: each batch solves to the end
        self.refill = bool(refill)
        self.refills = 0
        self._mesh = {}          # datain -> mesh_key
        self._no_refill = set()  # requests a running batch refused (they wait for a batch of their own)

    # -- requests --------------------------------------------------------
    def submit(self, rid, datain, model, arrival=None):
        key = shape_key(datain, model)
        mk = mesh_key(datain) if self.refill else None
        with self.lock:
            self._mesh[datain] = mk
            if self.t0 is None:
                self.t0 = time.perf_counter()
            t = time.perf_counter() - self.t0 if arrival is None else arrival
            self.queues[key].append(Request(rid, datain, model, t))
            self.records[rid] = dict(arrival=t)
            if self.dispatcher is None:
                self.dispatcher = threading.Thread(target=self._dispatch, daemon=True)
                self.dispatcher.start()
            self.lock.notify_all()

    def close(self):
        """No more requests: the dispatcher returns once every waiting request has been solved."""
        with self.lock:
            self.closed = True
            self.lock.notify_all()

    def join(self):
        if self.dispatcher is not None:
            self.dispatcher.join()
        for th in list(self.threads):
            th.join()

    # -- dispatch ----------------------------------------------------------
    def _take(self, key):
        """Requests of the lane that fit the budget now (at least one when nothing is in flight)."""
        q = self.queues[key]
        take, need = [], float(self.options["workspace_bytes"])
        while q and len(take) < self.max_batch:
            req = q[0]
            est = self._estimate(req)
            if self.in_flight + need + est > self.budget and (take or self.in_flight > 0):
                break
            take.append(q.popleft())
            need += est
        return take, need

    def _estimate(self, req):
        if req.datain not in self._est:
            z = read_sizes(req.datain)
            self._est[req.datain] = estimate_bytes(z["ne"], z["npe"], z["npf"], z["nfe"], z["ncu"], z["nd"],
                                                   int(self.options.get("gmres_restart", z["restart"])),
                                                   self.precision)
        return self._est[req.datain]

    def _ready(self, now):
        """The lane whose batch should start now (oldest waiting request first), or None."""
        best = None
        for key, q in self.queues.items():
            if not q:
                continue
            age = now - q[0].arrival
            if age < self.window and len(q) < self.max_batch and not self.closed:
                continue
            if best is None or q[0].arrival < self.queues[best][0].arrival:
                best = key
        return best

    def _dispatch(self):
        """Start batches while requests wait, slots are free and the budget admits them."""
        with self.lock:
            while True:
                now = time.perf_counter() - self.t0
                key = self._ready(now) if self.running < self.max_concurrent else None
                take = []
                if key is not None:
                    take, need = self._take(key)
                if take:
                    self.in_flight += need
                    self.peak_bytes = max(self.peak_bytes, self.in_flight)
                    self.running += 1
                    for r in take:
                        self.records[r.rid]["admit"] = now
                    th = threading.Thread(target=self._run_batch, args=(key, take, need), daemon=True)
                    self.threads.append(th)
                    th.start()
                    continue
                if self.closed and self.running == 0 and not any(self.queues.values()):
                    return
                self.lock.wait(0.005)

    def _solved(self, norm):
        return bool(np.isfinite(norm)) and norm <= self.accept_norm

    def _take_same_mesh(self, key, mk):
        """The oldest waiting request of the lane whose mesh is mk (removed from the queue), or None."""
        with self.lock:
            q = self.queues[key]
            for idx, req in enumerate(q):
                if self._mesh.get(req.datain) == mk and req.rid not in self._no_refill:
                    del q[idx]
                    self.records[req.rid]["admit"] = time.perf_counter() - self.t0
                    return req
        return None

    def _run_batch(self, key, take, need):
        try:
            try:
                failed = self._solve(key, take, self.precision)
            except Exception as e:
                # a numerical failure of the whole batch (EXA_ERR_NUMERIC) in the mixed form:
                # solve it again in FP64
                if self.precision == 0 or getattr(e, "code", None) != 4:
                    raise
                failed = take
            if failed and self.precision != 0:
                # instances whose Newton iteration broke down with the FP32 factors: again, in FP64
                with self.lock:
                    self.fallbacks.append((key, len(failed)))
                failed = self._solve(key, failed, 0, admit=False)
            with self.lock:
                for r in failed:
                    self.records[r.rid]["failed"] = True
        except Exception as e:  # the requests fail, the server goes on
            with self.lock:
                self.errors.append((key, [r.rid for r in getattr(e, "requests", take)], repr(e)))
        finally:
            with self.lock:
                self.in_flight -= need
                self.running -= 1
                self.lock.notify_all()

    def _solve(self, key, take, precision, admit=True):
        if self.refill:
            return self._solve_steps(key, take, precision, admit)
        t1 = time.perf_counter()
        S = HDGSolver.batch([r.datain for r in take], take[0].model)
        try:
            S.set_option("precision", precision)
            for k, v in self.options.items():
                S.set_option(k, v)
            S.solve()
            newton, gmres, norm = S.instance_stats()
            done = time.perf_counter() - self.t0
            uhs = [S.get_instance(i, "uh") for i in range(len(take))] if self.keep_results else None
        finally:
            S.close()
        if self.log:
            self.log("batch of %d %s p=%d requests solved in %.2f s (%s), t = %.1f s"
                     % (len(take), key[0], key[4], time.perf_counter() - t1, "mixed" if precision else "fp64", done))
        failed = []
        with self.lock:
            self.batches.append((key, len(take), time.perf_counter() - t1))
            for i, r in enumerate(take):
                if not self._solved(norm[i]):       # the instance's Newton iteration broke down
                    self.records[r.rid].update(norm=float(norm[i]), precision=precision)
                    failed.append(r)
                    continue
                self.records[r.rid].update(done=done, batch=len(take), newton=int(newton[i]), gmres=int(gmres[i]),
                                           norm=float(norm[i]), precision=precision)
                self.records[r.rid].pop("failed", None)
                if uhs is not None:
                    self.results[r.rid] = uhs[i]
        return failed

    def _solve_steps(self, key, take, precision, admit):
        """Step one batch to the end; after each step hand out the finished requests' results and
        refill their slots from the lane's queue (admit) -- requests of the same mesh."""
        t1 = time.perf_counter()
        S = HDGSolver.batch([r.datain for r in take], take[0].model)
        failed, slots, served = [], list(take), len(take)
        try:
            S.set_option("precision", precision)
            for k, v in self.options.items():
                S.set_option(k, v)
            while any(r is not None for r in slots):
                try:
                    st = S.step()
                except ExaError as e:
                    # a numerical failure of the whole batch (EXA_ERR_NUMERIC) in the mixed form:
                    # the requests under way are solved again in FP64
                    if precision == 0 or e.code != 4:
                        e.requests = [r for r in slots if r is not None]
                        raise
                    failed += [r for r in slots if r is not None]
                    break
                fin = [i for i, r in enumerate(slots) if r is not None and st[i] != 1]
                if not fin:
                    continue
                newton, gmres, norm = S.instance_stats()
                done = time.perf_counter() - self.t0
                for i in fin:
                    r = slots[i]
                    uh = S.get_instance(i, "uh") if self.keep_results else None
                    with self.lock:
                        if st[i] < 0 or not self._solved(norm[i]):   # its Newton iteration broke down
                            self.records[r.rid].update(norm=float(norm[i]), precision=precision)
                            failed.append(r)
                        else:
                            self.records[r.rid].update(done=done, batch=len(take), newton=int(newton[i]),
                                                       gmres=int(gmres[i]), norm=float(norm[i]),
                                                       precision=precision, refilled=r not in take)
                            self.records[r.rid].pop("failed", None)
                            if uh is not None:
                                self.results[r.rid] = uh
                    slots[i] = None
                    nxt = self._take_same_mesh(key, self._mesh.get(r.datain)) if admit else None
                    if nxt is None:
                        continue
                    try:
                        S.refill(i, nxt.datain)
                    except ExaError:  # other options or time steps: it waits for a batch of its own
                        with self.lock:
                            self._no_refill.add(nxt.rid)
                            self.queues[key].appendleft(nxt)
                            self.lock.notify_all()
                        continue
                    slots[i] = nxt
                    served += 1
                    with self.lock:
                        self.refills += 1
        finally:
            S.close()
        if self.log:
            self.log("batch of %d (%d admitted while it ran) %s p=%d requests solved in %.2f s (%s), t = %.1f s"
                     % (served, served - len(take), key[0], key[4], time.perf_counter() - t1,
                        "mixed" if precision else "fp64", time.perf_counter() - self.t0))
        with self.lock:
            self.batches.append((key, served, time.perf_counter() - t1))
        return failed

    # -- replay of a timed stream ------------------------------------------
    def run_stream(self, requests):
        """requests: [(rid, datain, model, arrival_seconds)], submitted at their arrival times
        (relative to the call); returns when every request is done."""
        reqs = sorted(requests, key=lambda r: r[3])
        with self.lock:
            self.t0 = time.perf_counter()
        for rid, din, model, t in reqs:
            wait = t - (time.perf_counter() - self.t0)
            if wait > 0:
                time.sleep(wait)
            self.submit(rid, din, model, arrival=t)
        self.close()
        self.join()
        return self.summary()

    def summary(self):
        done = [r for r in self.records.values() if "done" in r and not r.get("failed")]
        tts = np.array([r["done"] - r["arrival"] for r in done]) if done else np.zeros(0)
        span = (max(r["done"] for r in done) - min(r["arrival"] for r in self.records.values())) if done else 0.0
        nfail = sum(1 for r in self.records.values() if r.get("failed")) + sum(len(e[1]) for e in self.errors)
        return dict(requests=len(self.records), completed=len(done), failed=nfail,
                    solves_per_s=len(done) / span if span > 0 else None,
                    p50_ms=float(np.percentile(tts, 50) * 1e3) if len(tts) else None,
                    p99_ms=float(np.percentile(tts, 99) * 1e3) if len(tts) else None,
                    batches=len(self.batches),
                    mean_batch=float(np.mean([b[1] for b in self.batches])) if self.batches else None,
                    peak_in_flight_gb=self.peak_bytes / 1e9, errors=[e[2] for e in self.errors[:3]],
                    refills=self.refills,
                    fp64_fallback_batches=len(self.fallbacks),
                    fp64_fallback_requests=sum(f[1] for f in self.fallbacks),
                    lanes={"%s p=%d" % (k[0], k[4]): dict(batches=len(v), requests=int(sum(b[1] for b in v)),
                                                         mean_batch_s=float(np.mean([b[2] for b in v])))
                           for k, v in self._by_lane().items()})

    def _by_lane(self):
        out = collections.defaultdict(list)
        for b in self.batches:
            out[b[0]].append(b)
        return out
