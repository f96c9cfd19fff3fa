"""Batched solves: many independent meshes advanced together on one GPU.

A batch (``afar_batch_*`` of the C-ABI) stores its meshes one after another in
one device layout, so each kernel launch of an RK3 step covers every mesh and
one graph launch advances the whole batch by one step; each mesh keeps its own
CFL time step and gets exactly the trajectory it gets alone.

The autotuner groups meshes into size buckets and, per bucket, times each LDS
tile (the cell kernels' workgroup size: 64, 128 or 256 cells) on a sample of
the bucket's meshes, keeping the fastest.  ``BucketedBatch`` then runs one
batch per bucket with its tile.
"""
import ctypes as C
import json
import math
import time

import numpy as np

from ._lib import check, dptr, lib, ptr
from .solver import schemes as make_schemes

TILES = (64, 128, 256)


class BatchSolver:
    """One afar_batch over `meshes` (precomputed Mesh objects) with initial
    states `W0s` (caller order of each mesh) and one schemes `sc`."""

    def __init__(self, meshes, sc, W0s):
        self.meshes = list(meshes)
        n = len(self.meshes)
        if n == 0 or n != len(W0s):
            raise ValueError("need one initial state per mesh")
        self._w0 = [np.ascontiguousarray(w, dtype=np.float64).reshape(-1, 4) for w in W0s]
        hs = (C.c_void_p * n)(*[m.handle.value for m in self.meshes])
        ws = (C.POINTER(C.c_double) * n)(*[dptr(w) for w in self._w0])
        h = C.c_void_p()
        check(lib().afar_batch_create(n, C.cast(hs, C.POINTER(C.c_void_p)), C.byref(sc), ws, C.byref(h)))
        self._h = h
        self.schemes = sc
        self.ncs = np.array([m.nc for m in self.meshes], dtype=np.int64)
        self.steps = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().afar_batch_destroy(h)
            self._h = None

    def close(self):
        self.__del__()

    @property
    def cells(self):
        return int(self.ncs.sum())

    def evolve(self, nsteps):
        check(lib().afar_batch_evolve(self._h, nsteps))
        self.steps += nsteps

    def time_steps(self, nsteps):
        ms = C.c_double()
        check(lib().afar_batch_time_steps(self._h, nsteps, C.byref(ms)))
        self.steps += nsteps
        return ms.value

    def synchronize(self):
        check(lib().afar_batch_synchronize(self._h))

    def reset(self):
        check(lib().afar_batch_reset(self._h))
        self.steps = 0

    def sizes(self):
        o = np.zeros(8, dtype=np.int64)
        check(lib().afar_batch_sizes(self._h, ptr(o, C.c_int64)))
        return dict(zip(["ninst", "nc", "nf", "nb", "tile", "maxS", "maxC", "maxHalo"], (int(v) for v in o)))

    def last_dt(self):
        dt = np.zeros(len(self.meshes))
        check(lib().afar_batch_last_dt(self._h, dptr(dt)))
        return dt

    def set_time(self, t, t_end=float("inf"), inst=-1):
        check(lib().afar_batch_set_time(self._h, inst, t, t_end))

    def device_time(self):
        t = np.zeros(len(self.meshes))
        check(lib().afar_batch_get_time(self._h, dptr(t), None))
        return t

    def device_steps(self):
        n = np.zeros(len(self.meshes))
        check(lib().afar_batch_get_time(self._h, None, dptr(n)))
        return n.astype(np.int64)

    def conserved(self, inst):
        U = np.zeros((self.meshes[inst].nc, 4))
        check(lib().afar_batch_get_conserved(self._h, inst, dptr(U)))
        return U

    def write_field(self, inst):
        W = np.zeros((self.meshes[inst].nc, 4))
        sh = np.zeros(self.meshes[inst].nc, dtype=np.int8)
        check(lib().afar_batch_write_field(self._h, inst, dptr(W), ptr(sh, C.c_int8)))
        return W, sh

    def profile(self, reps=1):
        names = ["cfl", "reconstruct", "bvd", "flux", "update", "flux_dt", "dt_reduce"]
        ms = np.zeros(len(names))
        cnt = np.zeros(len(names), dtype=np.int32)
        check(lib().afar_batch_profile(self._h, reps, dptr(ms), ptr(cnt, C.c_int32)))
        return dict(zip(names, zip(ms.tolist(), cnt.tolist())))


def size_buckets(ncs, per_decade=2):
    """Mesh indices grouped by log10(cells), `per_decade` buckets per decade."""
    keys = np.floor(np.log10(np.asarray(ncs, dtype=float)) * per_decade).astype(int)
    return {int(k): [int(i) for i in np.where(keys == k)[0]] for k in np.unique(keys)}


def autotune(meshes, W0s, sc_kwargs, tiles=TILES, steps=10, warmup=3, sample=16, seed=0):
    """Fastest tile for these meshes: each tile timed on the same random
    sample (at most `sample` meshes) of `steps` graph-launched RK3 steps.
    Returns (best tile, {tile: ns per cell-step})."""
    rng = np.random.default_rng(seed)
    idx = np.arange(len(meshes))
    if len(idx) > sample:
        idx = np.sort(rng.choice(idx, sample, replace=False))
    ms_ = [meshes[i] for i in idx]
    ws_ = [W0s[i] for i in idx]
    cells = sum(m.nc for m in ms_)
    timing = {}
    for t in tiles:
        b = BatchSolver(ms_, make_schemes(**dict(sc_kwargs, tile=t)), ws_)
        b.evolve(warmup)
        b.synchronize()
        ms = b.time_steps(steps)
        b.close()
        timing[t] = ms * 1e6 / (cells * steps)
    best = min(timing, key=timing.get)
    return best, timing


class BucketedBatch:
    """Meshes split into size buckets, one autotuned batch per bucket.  All
    batches step together: evolve(n) enqueues n steps on every bucket's
    stream before waiting, so the buckets overlap on the device."""

    def __init__(self, meshes, W0s, sc_kwargs, per_decade=2, tune=True, tune_steps=10, log=None):
        self.meshes = list(meshes)
        self.buckets = size_buckets([m.nc for m in self.meshes], per_decade)
        self.batches = []
        self.tuning = {}
        for key, ids in self.buckets.items():
            ms_ = [self.meshes[i] for i in ids]
            ws_ = [W0s[i] for i in ids]
            if tune:
                tile, timing = autotune(ms_, ws_, sc_kwargs, steps=tune_steps, seed=key)
            else:
                tile, timing = sc_kwargs.get("tile", 256), {}
            self.tuning[key] = dict(meshes=len(ids), cells=int(sum(m.nc for m in ms_)), tile=tile,
                                    ns_per_cell_step={str(k): round(v, 4) for k, v in timing.items()})
            if log:
                log("bucket %s: %s" % (key, json.dumps(self.tuning[key])))
            self.batches.append((ids, BatchSolver(ms_, make_schemes(**dict(sc_kwargs, tile=tile)), ws_)))

    @property
    def cells(self):
        return sum(b.cells for _, b in self.batches)

    def evolve(self, nsteps):
        for _, b in self.batches:
            b.evolve(nsteps)

    def synchronize(self):
        for _, b in self.batches:
            b.synchronize()

    def reset(self):
        for _, b in self.batches:
            b.reset()

    def conserved(self, i):
        for ids, b in self.batches:
            if i in ids:
                return b.conserved(ids.index(i))
        raise IndexError(i)

    def close(self):
        for _, b in self.batches:
            b.close()
        self.batches = []


def corner_instances(n=256, seed=2, nc_min=5e3, nc_max=2e5, mach=(1.5, 4.0), angle=(5.0, 25.0)):
    """Parameters of `n` corner-flow instances: cells log-uniform in
    [nc_min, nc_max] with nx:ny about 2:1, Mach and corner angle uniform."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nc = math.exp(rng.uniform(math.log(nc_min), math.log(nc_max)))
        ny = max(4, int(round(math.sqrt(nc / 2.0))))
        nx = max(8, int(round(nc / ny)))
        out.append(dict(nx=nx, ny=ny, mach_u=float(rng.uniform(*mach)), angle_deg=float(rng.uniform(*angle)),
                        seed=int(rng.integers(0, 2 ** 31))))
    return out


def build_corner_instances(params, sc, log=None):
    """Meshes (precomputed with `sc`) and initial states of corner instances."""
    from .cases import corner_case
    meshes, W0s = [], []
    t0 = time.perf_counter()
    for k, p in enumerate(params):
        m, w = corner_case(nx=p["nx"], ny=p["ny"], seed=p["seed"], mach_u=p["mach_u"], angle_deg=p["angle_deg"])
        m.precompute(sc)
        meshes.append(m)
        W0s.append(w)
        if log and (k + 1) % 32 == 0:
            log("precomputed %d/%d meshes, %.1f s" % (k + 1, len(params), time.perf_counter() - t0))
    return meshes, W0s
