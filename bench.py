#!/usr/bin/env python3
"""Throughput benchmark of the AFAR SSP-RK3 pipeline on one MI355X.

Workload (default): the 1000x1000 jittered unit square (+-20%, seed 1), each
quad split into two triangles with probability 0.5 -> 1,500,605 cells;
four-quadrant Riemann initial data; transmissive boundaries; AFAR switch
with KFVS/Roe; all-FP64; CFL 0.5.  A step = CFL time-step reduction + one
SSP-RK3 step (3 stages), time step kept on the device, replayed as a
hipGraph; inputs resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--case riemann|corner] [--precision fp64|mixed]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0


def mesh_stats(mesh):
    S, Cc, _, cf = mesh.stencils()
    s = mesh.sizes()
    nf = s["nf"]
    internal = (np.abs(cf) - 1 < nf) & (cf != 0)
    n_int = internal.sum(axis=1)
    return dict(nc=s["nc"], nf=nf, nb=s["nb"], sumS=int(S.sum()), sumC=int(Cc.sum()),
                sum_int_coef=int((n_int * (S.astype(np.int64) + 1)).sum()), meanS=float(S.mean()),
                meanC=float(Cc.mean()))


def kernel_bytes(st, w=8):
    """Compulsory bytes per launch of each kernel (every array element touched
    once; neighbour re-reads counted as cache hits).  w = width of the
    reconstruction/flux type (8 FP64, 4 FP32-mixed); state stays FP64."""
    nc, nf, nb, sS, sC, sR = st["nc"], st["nf"], st["nb"], st["sumS"], st["sumC"], st["sum_int_coef"]
    ns = 2 * nf  # internal face slots
    recon = (4 * w * (nc + nb) + 2 * nc + 2 * sS + 4 * sC + 16 * ns + 16 * ns + 2 * w * sC + w * sR
             + 8 * w * nc + 4 * w * nc + nc + 4 * w * ns + w * ns)
    bvd = ((nc + nb) + 2 * nc + 2 * sS + 4 * sC + 16 * ns + w * ns + 4 * w * nc + 2 * w * nc + w * nc
           + nc + 4 * w * nc)
    flux = 8 * nf + 64 * nf + 8 * w * nf + nc * (4 * w + 4 * w + 8 * w + 1) + 32 * nb + 4 * w * nb + 32 * (nf + nb)
    update = nc + 4 * sC + 32 * (nf + nb) + 32 * nc * 2 + 8 * nc + 65 * nb + 32 * nc + 4 * w * (nc + nb) + 32 * nc / 3
    cfl = 8 * nf + 48 * nf + 4 * w * nc
    nblk = 8 * (((nf + nb + 255) // 256 + 7) // 8)  # flux grid: one CFL partial per block
    return dict(cfl=cfl, reconstruct=recon, bvd=bvd, flux=flux, update=update, flux_dt=flux + 8 * nblk,
                dt_reduce=8 * nblk)


def survey_balg_per_cell_step(st, w=8):
    """Compulsory traffic per cell-step of the whole pipeline with no
    intermediate arrays (static data streamed once per stage)."""
    nc, nf, nb, sS, sC, sR = st["nc"], st["nf"], st["nb"], st["sumS"], st["sumC"], st["sum_int_coef"]
    per_stage = (nc * (2 * 4 * w + 3 * 32 + 8 + 16 + 2 + 4) + 2 * w * sC + 4 * sS + 4 * sC
                 + nf * (40 + 8) + w * sR + nb * (40 + 2 * 4 * w + 5))
    per_step = 3 * per_stage + 48 * nf + (16 + 4 * w) * nc
    return per_step / nc


def measured_traffic(kernel, case, precision):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE scaled by a copy
    kernel's reading), or None when there is none."""
    path = os.path.join(ROOT, "profiles", "riemann1000_%s_hbm_traffic.txt" % precision)
    if case != "riemann" or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if f and not line.startswith("#") and f[0] == kernel and len(f) == 5:
            return int(f[4])
    return None


def cpu_baseline(mesh, sc, W0, budget_s=12.0, max_steps=40):
    """The FP64 oracle (OpenMP) on the same mesh: steps until budget_s."""
    from oracle import pyoracle as O
    ref = mesh.reference_arrays(sc, W0)
    orc = O.OracleState(ref, sc.R, sc.Cv, boundary_first_order=bool(sc.boundaryFirstOrder))
    del ref
    orc.run(1, sc.maxCo)  # warm-up step
    t0 = time.perf_counter()
    n = 0
    while n < max_steps:
        orc.run(1, sc.maxCo)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    return dict(value=mesh.nc * n / el / 1e6, unit="Mcell-RK3-steps/s", cores=cores, kind="port",
                sample="%d RK3 steps (with CFL) of the same %d-cell mesh, FP64 oracle, %d OpenMP threads, %.1f s"
                       % (n, mesh.nc, cores, el))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--case", default="riemann", choices=["riemann", "corner", "batch"],
                    help="batch: 256 corner flows (5k-200k cells, seed 2), FP32-mixed, autotuned per size bucket")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--tile", type=int, default=64)
    ap.add_argument("--recon-lanes", type=int, default=0, choices=[0, 1, 4],
                    help="reconstruction threads per cell (0: 4 when every cell has at most four faces)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--faithful-boundary", action="store_true",
                    help="riemann case without the boundaryFirstOrder extension (loses positivity)")
    ap.add_argument("--batch-size", type=int, default=256, help="meshes of the batch case")
    ap.add_argument("--batch-fp64", action="store_true", help="batch case in FP64 (default FP32-mixed)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md on sharding); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 8, "bytes_written": n * 8, "copies": 3}))
        return

    import afar_cudaeuler2dmixo3rk3solver_amd as A
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    if args.case == "batch":
        return run_batch(args)

    t0 = time.perf_counter()
    if args.case == "riemann":
        mesh, W0 = cases.riemann_case(n=1000, seed=1)
        workload = "four-quadrant Riemann, 1000x1000 jittered (+-20%, seed 1) quads split p=0.5 -> triangles/quads"
    else:
        mesh, W0 = cases.corner_case()
        workload = "Mach-2 15-degree corner, 200x100 jittered (+-15%, seed 0) quads"
    # The four-quadrant problem loses positivity in boundary triangles with the
    # unmodified reference algorithm within a few steps; the bench runs it with
    # the boundaryFirstOrder extension so the timed steps act on a physical state.
    bfo = args.case == "riemann" and not args.faithful_boundary
    sc = A.schemes(cOrder=7, fOrder=2, maxCo=0.5, precision=args.precision, tile=args.tile, boundaryFirstOrder=bfo,
                   reconLanes=args.recon_lanes)
    mesh.precompute(sc)
    t_pre = time.perf_counter() - t0
    print("bench: %s, nc=%d, precompute %.1f s" % (args.case, mesh.nc, t_pre), file=sys.stderr, flush=True)
    st = mesh_stats(mesh)
    gpu = A.BasicCUDASolver(mesh, sc, W0)
    dev_name = A.device_info()[0]

    gpu.evolve_auto(args.warmup)
    gpu.synchronize()
    wall0 = time.perf_counter()
    ms_events = gpu.time_steps(args.steps, graph=True)
    gpu.synchronize()
    wall = time.perf_counter() - wall0
    seconds = ms_events / 1e3
    value = mesh.nc * args.steps / seconds / 1e6
    print("bench: %d steps in %.3f s (events) / %.3f s (wall)" % (args.steps, seconds, wall), file=sys.stderr,
          flush=True)

    W_end, _, _ = gpu.write_field()
    finite = bool(np.isfinite(W_end).all() and (W_end[:, 0] > 0).all() and (W_end[:, 3] > 0).all())
    prof = gpu.profile(reps=20)  # back-to-back launches per kernel; after the timed run
    w = 8 if args.precision == "fp64" else 4
    kb = kernel_bytes(st, w)
    kernels = {}
    for k, (avg, cnt) in prof.items():
        kernels[k] = dict(avg_ms=avg, count_per_step=cnt, alg_bytes=kb[k], GBps=kb[k] / (avg * 1e-3) / 1e9)
    # cfl: the standalone pass of host-requested time steps; a device-time-step
    # RK3 step gets its time step from flux_dt + dt_reduce instead (count 0)
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["count_per_step"])
    d = kernels[dom]
    balg = survey_balg_per_cell_step(st, w)
    roofline = dict(bound="hbm", achieved=round(d["GBps"], 1), peak=PEAK_HBM_GBPS, unit="GB/s",
                    frac=round(d["GBps"] / PEAK_HBM_GBPS, 4),
                    traffic=measured_traffic(dom, args.case, args.precision), kernel=dom,
                    alg_bytes_per_launch=int(d["alg_bytes"]), avg_launch_ms=round(d["avg_ms"], 5))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mesh, sc, W0)
    out = {
        "metric": "Mcell-RK3-steps/s", "value": round(value, 2), "unit": "Mcell-RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_events / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-reconstruction/f64-state", "data": "synthetic",
        "config": {"workload": workload, "case": args.case, "cells": mesh.nc, "internal_faces": st["nf"],
                   "boundary_faces": st["nb"], "mean_stencil": round(st["meanS"], 3),
                   "mean_compact": round(st["meanC"], 3), "precision": args.precision, "cfl": 0.5,
                   "tile": args.tile, "recon_lanes": args.recon_lanes or "auto", "device": dev_name.strip(),
                   "precompute_s": round(t_pre, 1),
                   "boundary_first_order": bfo, "state_positive_after_timed_steps": finite},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
        "whole_step": {"balg_bytes_per_cell_step": round(balg, 1),
                       "balg_GBps": round(balg * value * 1e6 / 1e9, 1),
                       "kernel_alg_bytes_per_cell_step": round(
                           sum(kb[k] * kernels[k]["count_per_step"] for k in kb) / mesh.nc, 1)},
    }
    print(json.dumps(out))


def run_batch(args):
    """256 independent corner flows advanced together (one batch per size
    bucket, tile autotuned per bucket); a step = one RK3 step of every mesh."""
    import afar_cudaeuler2dmixo3rk3solver_amd as A
    from afar_cudaeuler2dmixo3rk3solver_amd import batch as Bt
    log = lambda m: print("bench: " + m, file=sys.stderr, flush=True)
    precision = "mixed" if args.precision == "fp64" and not args.batch_fp64 else args.precision
    sc_kwargs = dict(cOrder=7, fOrder=2, maxCo=0.5, precision=precision)
    t0 = time.perf_counter()
    params = Bt.corner_instances(n=args.batch_size, seed=2)
    meshes, W0s = Bt.build_corner_instances(params, A.schemes(**sc_kwargs), log=log)
    t_pre = time.perf_counter() - t0
    t1 = time.perf_counter()
    bb = Bt.BucketedBatch(meshes, W0s, sc_kwargs, log=log)
    t_tune = time.perf_counter() - t1
    bb.evolve(args.warmup)
    bb.synchronize()
    bb.reset()
    for _, b in bb.batches:
        b.set_time(0.0)
    bb.synchronize()
    w0 = time.perf_counter()
    bb.evolve(args.steps)
    bb.synchronize()
    wall = time.perf_counter() - w0
    cells = bb.cells
    value = cells * args.steps / wall / 1e6
    finite = all(np.isfinite(b.conserved(k)).all() for _, b in bb.batches for k in (0, len(b.meshes) - 1))
    out = {
        "metric": "Mcell-RK3-steps/s", "value": round(value, 2), "unit": "Mcell-RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f32-reconstruction/f64-state", "data": "synthetic",
        "config": {"workload": "%d corner flows, cells log-uniform 5e3-2e5, Mach U[1.5,4], angle U[5,25] deg, seed 2"
                               % len(meshes), "case": "batch", "meshes": len(meshes), "cells": cells,
                   "precision": precision, "cfl": 0.5, "device": A.device_info()[0].strip(),
                   "precompute_s": round(t_pre, 1), "autotune_s": round(t_tune, 1),
                   "buckets": {str(k): v for k, v in bb.tuning.items()}, "state_finite": bool(finite),
                   "timing": "host wall clock around the enqueue of all buckets' steps and their completion"},
        "roofline": None, "cpu_baseline": None,
    }
    bb.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
