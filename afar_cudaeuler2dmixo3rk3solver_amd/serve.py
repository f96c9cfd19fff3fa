"""Served solves with continuous batching on one GPU.

Requests (a corner flow run until its density residual has fallen by a
given factor, or a four-quadrant Riemann problem run to a fixed time) arrive
over time.  Each is
prepared on the host (mesh, RCM, stencils, RBF precompute), then admitted to
the GPU when its device footprint fits in the HBM budget and a slot is free.
Admitted requests run concurrently, each on its own HIP stream with its state
resident on the device, advanced in graph-launched chunks of device-clocked
RK3 steps; requests that arrive while others run join as soon as memory and a
slot free up, so their kernels overlap on the device.  A request's result
does not depend on what else is running: every kernel of a solver reads only
that solver's buffers, and the chunking that decides when a residual is
checked is fixed per request.

    python -m afar_cudaeuler2dmixo3rk3solver_amd.serve [--rate 20] [--duration 120] [--nc-min 1e4] [--nc-max 2e6]

prints one JSON line per request and a summary line with p50/p99
time-to-solution (arrival to result, precompute included), how many requests
met their stop criterion (converged) or failed, the aggregate cell-steps/s
and the L1 density error of a sample of requests against the all-FP64 solve.
A request whose preparation or solve raises is recorded as failed (its error
message kept) and the loop goes on.
"""
import argparse
import faulthandler
import json
import math
import os
import queue
import sys
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import cases
from .solver import BasicCUDASolver, device_info, schemes


@dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the run
    kind: str                 # "corner" or "riemann"
    nc_target: int
    seed: int
    t_end: float = 0.0        # riemann
    res_tol: float = 0.0      # corner: stop when residual / first residual < res_tol
    max_steps: int = 0        # corner: step cap
    # filled while serving
    times: dict = field(default_factory=dict)
    steps: int = 0
    cells: int = 0
    result: object = None
    residual: float = float("nan")   # last density residual (corner)
    residual0: float = float("nan")  # density residual after the first chunk (corner)
    converged: bool = False          # the stop criterion was met (not the step cap)
    error: str = ""                  # set when preparation or solving raised


def request_stream(rate=20.0, duration=120.0, seed=3, nc_min=1e4, nc_max=2e6, t_end=0.02, res_tol=1e-8,
                   max_steps=400):
    """Poisson arrivals at `rate` per second for `duration` seconds; half
    corner flows, half Riemann problems; cells log-uniform in [nc_min, nc_max]."""
    rng = np.random.default_rng(seed)
    out, t, k = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        kind = "corner" if rng.uniform() < 0.5 else "riemann"
        nc = int(round(math.exp(rng.uniform(math.log(nc_min), math.log(nc_max)))))
        out.append(Request(k, t, kind, nc, int(rng.integers(0, 2 ** 31)), t_end=t_end, res_tol=res_tol,
                           max_steps=max_steps))
        k += 1
    return out


def build_case(req, sc=None, gpu=False):
    """Mesh and initial state of a request; precomputed (the RBF part on the
    GPU when gpu is set) when `sc` is given."""
    if req.kind == "corner":
        ny = max(4, int(round(math.sqrt(req.nc_target / 2.0))))
        nx = max(8, int(round(req.nc_target / ny)))
        mesh, W0 = cases.corner_case(nx=nx, ny=ny, seed=req.seed % 1000)
    else:
        # a quad split in two with probability 0.5 gives 1.5 cells per quad
        n = max(8, int(round(math.sqrt(req.nc_target / 1.5))))
        mesh, W0 = cases.riemann_case(n=n, seed=req.seed % 1000)
    if sc is not None:
        mesh.precompute(sc, gpu=gpu)
    return mesh, W0


def footprint_bytes(mesh, precision):
    """Device memory of a solver for this mesh (the buffers it allocates)."""
    s = mesh.sizes()
    nc, nf, nb, maxS, maxC = s["nc"], s["nf"], s["nb"], s["maxS"], s["maxC"]
    w = 8 if precision == "fp64" else 4
    nslot = (nc + 63) // 64 * 64 * maxC
    cell = nc * (4 * w * 4 + 32 * 3 + 8 + 4 * 2 * w + 2 + 2) + (nc + nb) * (4 * w + 1)
    slots = nslot * (2 * 1 + 4 + 16 + 16 + 2 * w) + (nc + 63) // 64 * 64 * maxS * 2
    rbf = nslot * (maxS + 1) * w
    faces = nf * (8 + 64 + 2 * 4 * w + 2 * w) + (nf + nb) * 32 + nb * (4 + 1 + 64)
    # cell and face centres (the FP64 reconstruction, the face passes), the
    # periodic shifts (counted always: an upper bound), Sf and magSf
    centres = 16 * (nc + nb + nf + nb) + 16 * nf + 24 * nf
    return int(cell + slots + rbf + faces + centres + 2 * nc * 4 + (1 << 20))


class Server:
    """Continuous-batching request loop (see the module docstring).

    Host threads prepare requests, the earliest arrival first or, with
    prep_order="smallest", the pending request with the fewest cells first:
    mesh generation, precompute (its RBF part on the GPU with gpu_precompute)
    and the solver with its state on the device and its step graph captured
    (the library's calls are safe from several threads).  One scheduler thread steps the solvers: it admits
    prepared requests while their footprint fits in the HBM budget and a slot
    is free, enqueues chunks of graph-launched steps on each active solver's
    stream without waiting, and polls the streams; a request whose stream is
    idle either gets its next chunk, or is finished (result copied back,
    memory and slot released, the solver handed to a reaper thread that
    destroys it, so that the scheduler never waits on the library's lock)."""

    def __init__(self, precision="mixed", hbm_budget=64 << 30, max_active=8, chunk=50, prep_workers=2,
                 log=None, poll_s=2e-4, gpu_precompute=False, max_prepared=None, prep_order="arrival"):
        if prep_order not in ("arrival", "smallest"):
            raise ValueError("prep_order must be 'arrival' or 'smallest'")
        self.precision = precision
        # which pending request a free preparation thread takes: the earliest
        # arrival, or the one with the fewest cells (shortest job first)
        self.prep_order = prep_order
        self.gpu_precompute = gpu_precompute
        self.budget = hbm_budget
        self.max_active = max_active
        self.chunk = chunk
        self.prep_workers = prep_workers
        # prepared requests hold their mesh and coefficients in host memory
        # until admitted: preparation waits while this many are ready
        self.max_prepared = max_prepared if max_prepared is not None else max(2 * max_active, prep_workers)
        self.poll_s = poll_s
        self.log = log or (lambda m: None)
        self.sc = schemes(cOrder=7, fOrder=2, maxCo=0.5, precision=precision)
        self.sc_riemann = schemes(cOrder=7, fOrder=2, maxCo=0.5, precision=precision, boundaryFirstOrder=True)
        self.peak_active = 0

    def schemes_of(self, req):
        return self.sc if req.kind == "corner" else self.sc_riemann

    def _create(self, req, mesh, W0):
        """The request's solver, state on the device, step graph captured
        (preparation threads: the scheduler then only launches)."""
        g = BasicCUDASolver(mesh, self.schemes_of(req), W0)
        g.mesh = None  # the solver holds its own copy: the host mesh can go
        g.evolve_auto(0)  # captures the step graph; no step is taken
        req.cells = mesh.nc
        return g

    # -- per-request steps, all on the scheduler thread
    def _start(self, req, g, t0):
        req.times["start"] = time.perf_counter() - t0
        if req.kind == "riemann":
            g.set_time(0.0, req.t_end)
        self._enqueue(req, g)

    def _enqueue(self, req, g):
        k = self.chunk if req.kind == "riemann" else min(self.chunk, req.max_steps - req.steps)
        g.evolve_auto(k)
        req.steps += k

    def _advance(self, req, g):
        """Called when the solver's stream is idle: True when the request is done.
        Corner requests stop when the density residual has fallen below res_tol
        times its value after the first chunk (the residual's absolute size
        depends on the mesh and the inflow), or at the step cap."""
        if req.kind == "riemann":
            if g.device_time() >= req.t_end:
                req.steps = g.device_steps()  # steps that moved the clock
                req.converged = True
                return True
        else:
            req.residual = g.residual_norm()
            if math.isnan(req.residual0):
                req.residual0 = req.residual
            if req.residual0 > 0 and req.residual < req.res_tol * req.residual0:
                req.converged = True
                return True
            if req.steps >= req.max_steps or not math.isfinite(req.residual):
                return True
        self._enqueue(req, g)
        return False

    def _finish(self, req, g, t0, reap=None):
        """Result to the host; the solver is destroyed by `reap` (a thread
        that frees device memory off the scheduler) or here."""
        try:
            req.result = g.conserved()
        finally:
            req.times["finish"] = time.perf_counter() - t0
            if reap is None:
                g.close()
            else:
                reap(g)

    @staticmethod
    def _fail(req, exc, t0):
        req.error = "%s: %s" % (type(exc).__name__, exc)
        req.times.setdefault("finish", time.perf_counter() - t0)

    def solve(self, req, mesh, W0, t0):
        """One request alone, with the same chunking as in the loop."""
        if not mesh.precomputed:
            mesh.precompute(self.schemes_of(req), gpu=self.gpu_precompute)
        g = self._create(req, mesh, W0)
        self._start(req, g, t0)
        while True:
            g.synchronize()
            if self._advance(req, g):
                break
        self._finish(req, g, t0)

    def run(self, requests, realtime=True):
        t0 = time.perf_counter()
        prep_q = queue.PriorityQueue()  # (key, request); the end markers sort after every request
        ready_q = queue.Queue()
        failed_q = queue.Queue()
        slots = threading.Semaphore(self.max_prepared)  # prepared, not yet admitted
        by_size = self.prep_order == "smallest"

        def arrivals():
            for r in requests:
                if realtime:
                    delay = r.arrival - (time.perf_counter() - t0)
                    if delay > 0:
                        time.sleep(delay)
                prep_q.put(((0, r.nc_target if by_size else r.arrival, r.rid), r))
            for k in range(self.prep_workers):
                prep_q.put(((1, 0, k), None))

        def preparer():
            try:
                while True:
                    r = prep_q.get()[1]
                    if r is None:
                        return
                    slots.acquire()
                    r.times["prep_start"] = time.perf_counter() - t0
                    try:
                        mesh, W0 = build_case(r, self.schemes_of(r), gpu=self.gpu_precompute)
                        need = footprint_bytes(mesh, self.precision)
                        g = self._create(r, mesh, W0)
                        del mesh, W0
                    except Exception as e:  # a request that cannot be prepared fails alone
                        r.times["prepared"] = time.perf_counter() - t0
                        self._fail(r, e, t0)
                        slots.release()
                        failed_q.put(r)
                        continue
                    r.times["prepared"] = time.perf_counter() - t0
                    ready_q.put([r, g, need])
            finally:
                ready_q.put(None)  # always posted, so the scheduler's producer count reaches zero

        reap_q = queue.Queue()

        def reaper():  # destroys finished solvers: a destroy waits for the solver's stream and takes the
            while True:  # library's lock, which the scheduler thread should not
                g = reap_q.get()
                if g is None:
                    return
                g.close()

        feeder = threading.Thread(target=arrivals, daemon=True)
        preps = [threading.Thread(target=preparer, daemon=True) for _ in range(self.prep_workers)]
        reap_t = threading.Thread(target=reaper, daemon=True)
        feeder.start()
        reap_t.start()
        for p in preps:
            p.start()
        busy = self.sched_s = dict(start=0.0, advance=0.0, finish=0.0, idle=0.0)

        waiting = []   # prepared, not admitted (arrival order)
        active = []    # (request, solver, footprint)
        done = []
        used = 0
        producers = self.prep_workers
        while producers > 0 or waiting or active:
            progressed = False
            while True:  # requests whose preparation failed
                try:
                    done.append(failed_q.get_nowait())
                except queue.Empty:
                    break
            while True:  # collect prepared requests
                try:
                    item = ready_q.get_nowait()
                except queue.Empty:
                    break
                if item is None:
                    producers -= 1
                else:
                    waiting.append(item)
            while waiting and len(active) < self.max_active:
                if used + waiting[0][2] > self.budget and active:
                    break
                r, g, need = waiting.pop(0)
                r.times["admitted"] = time.perf_counter() - t0
                progressed = True
                slots.release()
                c0 = time.perf_counter()
                try:
                    self._start(r, g, t0)
                except Exception as e:
                    reap_q.put(g)
                    self._fail(r, e, t0)
                    done.append(r)
                    continue
                finally:
                    busy["start"] += time.perf_counter() - c0
                active.append((r, g, need))
                used += need
                self.peak_active = max(self.peak_active, len(active))
            still = []
            for r, g, need in active:
                c0 = time.perf_counter()
                try:
                    finished = g.idle() and self._advance(r, g)
                    c1 = time.perf_counter()
                    busy["advance"] += c1 - c0
                    if finished:
                        self._finish(r, g, t0, reap_q.put)
                        busy["finish"] += time.perf_counter() - c1
                except Exception as e:
                    reap_q.put(g)
                    self._fail(r, e, t0)
                    finished = True
                if finished:
                    used -= need
                    done.append(r)
                    self.log(json.dumps(summary_row(r)))
                    progressed = True
                else:
                    still.append((r, g, need))
            active = still
            if not progressed:
                if producers > 0 and not active and not waiting:
                    try:
                        item = ready_q.get(timeout=0.05)
                        if item is None:
                            producers -= 1
                        else:
                            waiting.append(item)
                    except queue.Empty:
                        pass
                else:
                    c0 = time.perf_counter()
                    time.sleep(self.poll_s)
                    busy["idle"] += time.perf_counter() - c0
        feeder.join()
        for p in preps:
            p.join()
        reap_q.put(None)
        reap_t.join()
        while not failed_q.empty():
            done.append(failed_q.get_nowait())
        return sorted(done, key=lambda r: r.rid), time.perf_counter() - t0


def summary_row(r):
    tts = r.times["finish"] - r.arrival
    out = dict(rid=r.rid, kind=r.kind, cells=r.cells, steps=r.steps, arrival=round(r.arrival, 3),
               tts_s=round(tts, 3), converged=r.converged)
    if "prepared" in r.times and "prep_start" in r.times:
        out["prep_s"] = round(r.times["prepared"] - r.times["prep_start"], 3)
    if "admitted" in r.times and "prepared" in r.times:
        out["queue_s"] = round(r.times["admitted"] - r.times["prepared"], 3)
    if "start" in r.times:
        out["gpu_s"] = round(r.times["finish"] - r.times["start"], 3)
    if not math.isnan(r.residual):
        out["residual"] = r.residual
        out["residual_rel"] = r.residual / r.residual0 if r.residual0 > 0 else None
    if r.error:
        out["error"] = r.error
    return out


def fp64_reference(req, chunk=50):
    """The request solved alone on the all-FP64 path (the FP64 GPU path is
    pinned to the CPU oracle elsewhere), with the same stepping."""
    srv = Server(precision="fp64", max_active=1, chunk=chunk, gpu_precompute=True)
    mesh, W0 = build_case(req)
    r = Request(req.rid, 0.0, req.kind, req.nc_target, req.seed, req.t_end, req.res_tol, req.max_steps)
    srv.solve(r, mesh, W0, time.perf_counter())
    return r


def main(argv=None):
    faulthandler.enable()
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--duration", type=float, default=120.0)
    ap.add_argument("--nc-min", type=float, default=1e4)
    ap.add_argument("--nc-max", type=float, default=2e6)
    ap.add_argument("--t-end", type=float, default=0.02, help="end time of the Riemann requests")
    ap.add_argument("--res-tol", type=float, default=1e-8,
                    help="corner requests stop when the density residual falls below this times its first value")
    ap.add_argument("--max-steps", type=int, default=400, help="step cap of the corner requests")
    ap.add_argument("--max-active", type=int, default=8)
    ap.add_argument("--chunk", type=int, default=50)
    ap.add_argument("--sample", type=float, default=0.05, help="fraction checked against FP64")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--gpu-precompute", action="store_true",
                    help="RBF precompute on the GPU, from the preparation threads (default: on the host)")
    ap.add_argument("--prep-workers", type=int, default=2, help="host threads preparing requests")
    ap.add_argument("--prep-order", default="arrival", choices=["arrival", "smallest"],
                    help="which pending request a free preparation thread takes: the earliest arrival, or the "
                         "one with the fewest cells")
    ap.add_argument("--host-threads", type=int, default=0,
                    help="threads of each request's parallel host loops (0: the CPU share / prep workers)")
    args = ap.parse_args(argv)
    # the preparation threads already run requests side by side: each request's own parallel loops
    # get a share of the CPUs, not all of them (AFAR_HOST_THREADS, read by the library per loop)
    cpus = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cpus = min(cpus, max(1, int(os.environ["OMP_NUM_THREADS"])))
    os.environ["AFAR_HOST_THREADS"] = str(args.host_threads or max(1, cpus // max(1, args.prep_workers)))
    reqs = request_stream(args.rate, args.duration, args.seed, args.nc_min, args.nc_max, args.t_end, args.res_tol,
                          args.max_steps)
    srv = Server(max_active=args.max_active, chunk=args.chunk, log=lambda m: print(m, flush=True),
                 gpu_precompute=args.gpu_precompute, prep_workers=args.prep_workers, prep_order=args.prep_order)
    done, wall = srv.run(reqs)
    ok = [r for r in done if not r.error]
    tts = np.array([r.times["finish"] - r.arrival for r in ok])
    cell_steps = float(sum(r.cells * r.steps for r in ok))
    # a seeded sample of ceil(fraction x requests) solved requests, at least one
    rng = np.random.default_rng(args.seed + 1)
    k = min(len(ok), max(1, int(np.ceil(args.sample * len(ok)))))
    order = sorted(ok, key=lambda r: r.rid)
    sample = [order[i] for i in sorted(rng.choice(len(order), size=k, replace=False))] if ok else []
    import concurrent.futures

    def err_of(r):
        ref = fp64_reference(r, args.chunk)
        return float(np.abs(r.result[:, 0] - ref.result[:, 0]).sum() / np.abs(ref.result[:, 0]).sum())

    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        errs = list(ex.map(err_of, sample))
    def pct(vals):
        vals = np.asarray([v for v in vals if v is not None], dtype=float)
        return [round(float(np.percentile(vals, 50)), 3), round(float(np.percentile(vals, 99)), 3)] if len(vals) else None

    rows = [summary_row(r) for r in ok]
    corner_rel = [r.residual / r.residual0 for r in ok if r.kind == "corner" and r.residual0 > 0]
    print(json.dumps(dict(
        device=device_info()[0].strip(), requests=len(done), failed=len(done) - len(ok),
        p50_p99_s=dict(prep=pct(x.get("prep_s") for x in rows), queue=pct(x.get("queue_s") for x in rows),
                       gpu=pct(x.get("gpu_s") for x in rows)),
        corner_residual_rel_median=float(np.median(corner_rel)) if corner_rel else None,
        cells_total=int(sum(r.cells for r in ok)), gpu_precompute=args.gpu_precompute,
        prep_workers=args.prep_workers, prep_order=args.prep_order, max_active=args.max_active, chunk=args.chunk,
        converged=dict(corner=sum(r.converged for r in ok if r.kind == "corner"),
                       corner_of=sum(r.kind == "corner" for r in ok),
                       riemann=sum(r.converged for r in ok if r.kind == "riemann"),
                       riemann_of=sum(r.kind == "riemann" for r in ok)),
        rate=args.rate, duration=args.duration, res_tol_rel=args.res_tol, corner_step_cap=args.max_steps,
        riemann_t_end=args.t_end,
        nc_range=[args.nc_min, args.nc_max], wall_s=round(wall, 2),
        host_threads_per_request=int(os.environ["AFAR_HOST_THREADS"]),
        p50_tts_s=round(float(np.percentile(tts, 50)), 3) if len(tts) else None,
        p99_tts_s=round(float(np.percentile(tts, 99)), 3) if len(tts) else None,
        cell_steps_per_s=round(cell_steps / wall, 1),
        Mcell_steps_per_s=round(cell_steps / wall / 1e6, 2), peak_active=srv.peak_active,
        scheduler_s={k: round(v, 2) for k, v in srv.sched_s.items()},
        fp64_sample=len(errs), rho_L1_rel_vs_fp64_max=max(errs) if errs else None,
        rho_L1_rel_vs_fp64_median=float(np.median(errs)) if errs else None)), flush=True)


if __name__ == "__main__":
    sys.exit(main())
