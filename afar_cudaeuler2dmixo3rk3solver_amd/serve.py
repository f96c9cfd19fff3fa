"""Served solves with continuous batching on one GPU.

Requests (a corner flow run to a density-residual tolerance, or a
four-quadrant Riemann problem run to a fixed time) arrive over time.  Each is
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
time-to-solution (arrival to result, precompute included), the aggregate
cell-steps/s and the L1 density error of a sample of requests against the
all-FP64 solve.
"""
import argparse
import json
import math
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
    res_tol: float = 0.0      # corner
    max_steps: int = 0        # corner
    # filled while serving
    times: dict = field(default_factory=dict)
    steps: int = 0
    cells: int = 0
    result: object = None
    residual: float = float("nan")


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


def build_case(req, sc):
    """Mesh (precomputed) and initial state of a request."""
    if req.kind == "corner":
        ny = max(4, int(round(math.sqrt(req.nc_target / 2.0))))
        nx = max(8, int(round(req.nc_target / ny)))
        mesh, W0 = cases.corner_case(nx=nx, ny=ny, seed=req.seed % 1000)
    else:
        # a quad split in two with probability 0.5 gives 1.5 cells per quad
        n = max(8, int(round(math.sqrt(req.nc_target / 1.5))))
        mesh, W0 = cases.riemann_case(n=n, seed=req.seed % 1000)
    mesh.precompute(sc)
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
    return int(cell + slots + rbf + faces + 2 * nc * 4 + (1 << 20))


class Server:
    """Continuous-batching request loop (see the module docstring)."""

    def __init__(self, precision="mixed", hbm_budget=64 << 30, max_active=8, chunk=50, prep_workers=2,
                 log=None):
        self.precision = precision
        self.budget = hbm_budget
        self.max_active = max_active
        self.chunk = chunk
        self.prep_workers = prep_workers
        self.log = log or (lambda m: None)
        self.sc = schemes(cOrder=7, fOrder=2, maxCo=0.5, precision=precision)
        self.sc_riemann = schemes(cOrder=7, fOrder=2, maxCo=0.5, precision=precision, boundaryFirstOrder=True)
        self._cv = threading.Condition()
        self._used = 0
        self._active = 0
        self.peak_active = 0

    def _admit(self, need):
        with self._cv:
            while self._active >= self.max_active or (self._used + need > self.budget and self._active > 0):
                self._cv.wait()
            self._used += need
            self._active += 1
            self.peak_active = max(self.peak_active, self._active)

    def _release(self, need):
        with self._cv:
            self._used -= need
            self._active -= 1
            self._cv.notify_all()

    def solve(self, req, mesh, W0, t0):
        """Runs one admitted request on the GPU (own solver, own stream)."""
        sc = self.sc if req.kind == "corner" else self.sc_riemann
        g = BasicCUDASolver(mesh, sc, W0)
        req.times["start"] = time.perf_counter() - t0
        if req.kind == "riemann":
            req.steps = g.run_until(req.t_end, chunk=self.chunk)
        else:
            n = 0
            while n < req.max_steps:
                k = min(self.chunk, req.max_steps - n)
                g.evolve_auto(k)
                n += k
                req.residual = g.residual_norm()
                if req.residual < req.res_tol:
                    break
            req.steps = n
        req.result = g.conserved()
        req.cells = mesh.nc
        g.close()

    def run(self, requests, realtime=True):
        t0 = time.perf_counter()
        prep_q = queue.Queue()
        threads = []
        results_lock = threading.Lock()
        done = []

        def arrivals():
            for r in requests:
                if realtime:
                    delay = r.arrival - (time.perf_counter() - t0)
                    if delay > 0:
                        time.sleep(delay)
                prep_q.put(r)
            for _ in range(self.prep_workers):
                prep_q.put(None)

        def runner(r, mesh, W0, need):
            try:
                self.solve(r, mesh, W0, t0)
            finally:
                self._release(need)
            r.times["finish"] = time.perf_counter() - t0
            with results_lock:
                done.append(r)
            self.log(json.dumps(summary_row(r)))

        def preparer():
            while True:
                r = prep_q.get()
                if r is None:
                    return
                r.times["prep_start"] = time.perf_counter() - t0
                mesh, W0 = build_case(r, self.sc if r.kind == "corner" else self.sc_riemann)
                r.times["prepared"] = time.perf_counter() - t0
                need = footprint_bytes(mesh, self.precision)
                self._admit(need)
                r.times["admitted"] = time.perf_counter() - t0
                th = threading.Thread(target=runner, args=(r, mesh, W0, need))
                th.start()
                threads.append(th)

        feeder = threading.Thread(target=arrivals)
        feeder.start()
        preps = [threading.Thread(target=preparer) for _ in range(self.prep_workers)]
        for p in preps:
            p.start()
        feeder.join()
        for p in preps:
            p.join()
        for th in list(threads):
            th.join()
        return sorted(done, key=lambda r: r.rid), time.perf_counter() - t0


def summary_row(r):
    tts = r.times["finish"] - r.arrival
    return dict(rid=r.rid, kind=r.kind, cells=r.cells, steps=r.steps, arrival=round(r.arrival, 3),
                tts_s=round(tts, 3), prep_s=round(r.times["prepared"] - r.times["prep_start"], 3),
                queue_s=round(r.times["admitted"] - r.times["prepared"], 3),
                gpu_s=round(r.times["finish"] - r.times["start"], 3),
                residual=None if math.isnan(r.residual) else r.residual)


def fp64_reference(req):
    """The request solved alone on the all-FP64 path (the FP64 GPU path is
    pinned to the CPU oracle elsewhere), with the same stepping."""
    srv = Server(precision="fp64", max_active=1)
    sc = srv.sc if req.kind == "corner" else srv.sc_riemann
    mesh, W0 = build_case(req, sc)
    r = Request(req.rid, 0.0, req.kind, req.nc_target, req.seed, req.t_end, req.res_tol, req.max_steps)
    srv.solve(r, mesh, W0, time.perf_counter())
    return r


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--duration", type=float, default=120.0)
    ap.add_argument("--nc-min", type=float, default=1e4)
    ap.add_argument("--nc-max", type=float, default=2e6)
    ap.add_argument("--t-end", type=float, default=0.02, help="end time of the Riemann requests")
    ap.add_argument("--res-tol", type=float, default=1e-8)
    ap.add_argument("--max-steps", type=int, default=400, help="step cap of the corner requests")
    ap.add_argument("--max-active", type=int, default=8)
    ap.add_argument("--chunk", type=int, default=50)
    ap.add_argument("--sample", type=float, default=0.05, help="fraction checked against FP64")
    ap.add_argument("--seed", type=int, default=3)
    args = ap.parse_args(argv)
    reqs = request_stream(args.rate, args.duration, args.seed, args.nc_min, args.nc_max, args.t_end, args.res_tol,
                          args.max_steps)
    srv = Server(max_active=args.max_active, chunk=args.chunk, log=lambda m: print(m, flush=True))
    done, wall = srv.run(reqs)
    tts = np.array([r.times["finish"] - r.arrival for r in done])
    cell_steps = float(sum(r.cells * r.steps for r in done))
    rng = np.random.default_rng(args.seed + 1)
    sample = [r for r in done if rng.uniform() < args.sample]
    errs = []
    for r in sample:
        ref = fp64_reference(r)
        d = np.abs(r.result[:, 0] - ref.result[:, 0]).sum() / np.abs(ref.result[:, 0]).sum()
        errs.append(float(d))
    print(json.dumps(dict(
        device=device_info()[0].strip(), requests=len(done), rate=args.rate, duration=args.duration,
        nc_range=[args.nc_min, args.nc_max], wall_s=round(wall, 2), p50_tts_s=round(float(np.percentile(tts, 50)), 3),
        p99_tts_s=round(float(np.percentile(tts, 99)), 3), cell_steps_per_s=round(cell_steps / wall, 1),
        Mcell_steps_per_s=round(cell_steps / wall / 1e6, 2), peak_active=srv.peak_active,
        fp64_sample=len(errs), rho_L1_rel_vs_fp64_max=max(errs) if errs else None,
        rho_L1_rel_vs_fp64_median=float(np.median(errs)) if errs else None)), flush=True)


if __name__ == "__main__":
    sys.exit(main())
