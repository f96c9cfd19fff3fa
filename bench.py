#!/usr/bin/env python3
"""Throughput benchmark of the AFAR SSP-RK3 pipeline on one MI355X.

Workload (default): the 1000x1000 jittered unit square (+-20%, seed 1), each
quad split into two triangles with probability 0.5 -> 1,500,605 cells;
four-quadrant Riemann initial data; transmissive boundaries; AFAR switch
with KFVS/Roe; all-FP64; CFL 0.5.  The run is first advanced (untimed) to
t = 0.15, the middle of the t = 0.3 run, where the waves cover much of the
domain; then W warm-up steps and K timed steps.  A step = CFL time-step
reduction + one SSP-RK3 step (3 stages), time step kept on the device,
replayed as a hipGraph; inputs resident in HBM.  Boundary cells take
first-order face states (the boundaryFirstOrder extension; the unmodified
algorithm loses positivity at the boundary of this mesh), stated in config.
Prints one JSON line:

  value          cell-RK3-steps per second over the K timed steps (Mcell)
  roofline       per RK3 step: the pipeline's compulsory bytes (B_alg per cell
                 x cells) over the measured step time against 8 TB/s; traffic =
                 HBM bytes per step from the committed rocprofv3 counter
                 summaries (profiles/riemann1000_*_hbm_traffic.txt); the
                 dominant kernel's own figures beside it
  cpu_baseline   the FP64 oracle (OpenMP, the host's CPUs) on the same mesh
                 from the same state for about 15 s
  accuracy       a fresh GPU solve from the oracle's starting state against
                 the FP64 oracle after the same number of steps (FP64:
                 bitwise); FP32-mixed also against the FP32-mixed oracle
                 (bitwise)

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--case riemann|corner|batch] [--precision fp64|mixed]
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


def mesh_stats(mesh, field=None, precision="fp64"):
    """Mesh sizes for the byte model; with field (write_field(): primitive
    state in caller order, ghosts, types) also the cells that take the
    reconstruction's full path (T not constant over the compact stencil,
    cudaLimiter.h:207-233): only those read their RBF coefficients and
    neighbour-centre offsets.  The test runs in the reconstruction's type:
    FP32-mixed rounds the primitive state to float first, as the device
    state does."""
    S, Cc, sten, cf = mesh.stencils()
    s = mesh.sizes()
    nf, nc = s["nf"], s["nc"]
    internal = (np.abs(cf) - 1 < nf) & (cf != 0)
    n_int = internal.sum(axis=1)
    coef = n_int * (S.astype(np.int64) + 1)
    full = np.ones(nc, dtype=bool)
    if field is not None:
        W, ghost, _ = field
        Tm = np.empty(nc + s["nb"])
        Tm[:nc] = W[mesh.permutation(), 3]  # mesh order
        Tm[nc:] = ghost[:, 3]
        cols = np.arange(sten.shape[1])[None, :] < Cc[:, None]
        Ts = np.where(cols, Tm[np.where(cols, sten, 0)], Tm[:nc, None])
        ft = np.float64 if precision == "fp64" else np.float32
        Tm, Ts = Tm.astype(ft), Ts.astype(ft)
        rng = np.maximum(Ts.max(axis=1), Tm[:nc]) - np.minimum(Ts.min(axis=1), Tm[:nc])
        full = ~(np.abs(rng) < ft(1e-10))
    return dict(nc=nc, nf=nf, nb=s["nb"], maxC=s["maxC"], sumS=int(S.sum()), sumC=int(Cc.sum()),
                sum_int_coef=int(coef.sum()), sum_int_coef_full=int(coef[full].sum()),
                int_slots_full=int(n_int[full].sum()), full_cells=int(full.sum()),
                meanS=float(S.mean()), meanC=float(Cc.mean()))


def kernel_bytes(st, w=8, ngp=0, geo=False):
    """Compulsory bytes per launch of each kernel (every array element touched
    once; neighbour re-reads counted as cache hits).  w = width of the
    reconstruction/flux type (8 FP64, 4 FP32-mixed); state stays FP64.  The
    reconstruction reads the RBF coefficients and the neighbour-centre
    distances (slot_dr) only for cells on the full path; face offsets
    (slot_d) and distances are stored in the kernels' precision.  geo (the
    FP64 four-lane kernel at the face centre): instead of the face offsets,
    distances and least-squares weights it reads the cell and pseudo-cell
    centres and, per compact slot, the face centre (16 B each).  ngp > 0
    (flux quadrature points): the reconstruction also reads each point's
    coefficients (full path) and offsets and writes each point's states; the
    flux reads those states and offsets instead of the face-centre states."""
    nc, nf, nb, sS, sC = st["nc"], st["nf"], st["nb"], st["sumS"], st["sumC"]
    sR, sfull = st["sum_int_coef_full"], st["int_slots_full"]
    ns = 2 * nf  # internal face slots
    slot_geo = 16 * (nc + nb) + 16 * sC if geo else 2 * w * ns + w * sfull + 2 * w * sC
    # writes: gradients 8w, limiters 4w, smoothness flag and limiter condition per cell; O3 face state 4w
    # (FP64: one array per variable, the temperatures being what the BVD test reads; FP32: records and the
    # temperatures again, w) and the BVD test's limited linear temperature w per internal face side
    recon = (4 * w * (nc + nb) + 2 * nc + 2 * sS + 4 * sC + slot_geo + w * sR
             + 8 * w * nc + 4 * w * nc + 2 * nc + 4 * w * ns + w * ns + (w * ns if w == 4 else 0))
    # bvd reads flags, counts, stencil ids, face slots, both sides' O3 face temperatures and this side's
    # limited linear one, the limiter condition; writes the types (the face pass applies the limiter transform)
    bvd = ((nc + nb) + 2 * nc + 2 * sS + 4 * sC + w * ns + w * ns + nc + nc)
    # FP32-mixed: internal fluxes at the face centre are stored as floats, and
    # the flux pass without the CFL bound reads the face geometry in float.
    # FP64 at the face centre: both passes read Sf, magSf and the face centre
    # (40 B per face) and the cells' centres (16 B per cell; a periodic mesh
    # adds its 16-B shift per face), and form the offsets and the distance;
    # so does the FP32-mixed pass with the CFL bound
    fw = 4 * w if (w == 4 and not ngp) else 32
    cg = w == 8 and not ngp
    if cg:
        dt_extra = 0
    elif not ngp:
        dt_extra = (40 - 8 * w) * nf + 16 * nc
    else:  # flux points: the 64-B FP64 record
        dt_extra = (64 - 8 * w) * nf
    geo = 40 if cg else 8 * w
    flux = (8 * nf + geo * nf + 8 * w * nf + nc * (4 * w + 4 * w + 8 * w + 1 + (16 if cg else 0)) + 32 * nb
            + 4 * w * nb + fw * nf + 32 * nb)
    # U and Uold are read in the second and third stages, Uold alone in the first; the first two stages
    # write U, the last writes Uold only
    update = nc + 4 * sC + fw * nf + 32 * nb + 32 * nc * 5 / 3 + 8 * nc + 65 * nb + 32 * nc + 4 * w * (nc + nb)
    cfl = 8 * nf + 48 * nf + 4 * w * nc
    if ngp:
        recon += ngp * (w * sR + 4 * w * ns + 4 * w * ns)
        flux += ngp * (8 * w + 4 * w) * nf - 8 * w * nf
    nblk = 8 * (((nf + nb + 255) // 256 + 7) // 8)  # flux grid: one CFL partial per block
    return dict(cfl=cfl, reconstruct=recon, bvd=bvd, flux=flux, update=update,
                flux_dt=flux + dt_extra + 8 * nblk, dt_reduce=8 * nblk)


def compulsory_bytes_per_cell_step(st, w=8):
    """Compulsory traffic per cell-step of the whole pipeline with no
    intermediate arrays (static data streamed once per stage)."""
    nc, nf, nb, sS, sC, sR = st["nc"], st["nf"], st["nb"], st["sumS"], st["sumC"], st["sum_int_coef"]
    per_stage = (nc * (2 * 4 * w + 3 * 32 + 8 + 16 + 2 + 4) + 2 * w * sC + 4 * sS + 4 * sC
                 + nf * (40 + 8) + w * sR + nb * (40 + 2 * 4 * w + 5))
    per_step = 3 * per_stage + 48 * nf + (16 + 4 * w) * nc
    return per_step / nc


def measured_traffic(kernel, case, precision):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (the rows hbm_table prints: read requests by
    size and WRITE_SIZE, each scaled by a copy kernel's reading), or None
    when there is none."""
    path = os.path.join(ROOT, "profiles", "riemann1000_%s_hbm_traffic.txt" % precision)
    if case != "riemann" or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if f and not line.startswith("#") and f[0] == kernel and len(f) == 5:
            return int(f[4])
    return None


KERNEL_KEYS = [("k_reconstruct", "reconstruct"), ("k_bvd", "bvd"), ("k_flux<double, true", "flux_dt"),
               ("k_flux<float, true", "flux_dt"), ("k_flux<", "flux"), ("k_dt_reduce", "dt_reduce"),
               ("k_update", "update"), ("k_cfl", "cfl")]


def _pmc_per_dispatch(csv_path):
    """{dispatch id: (kernel name, {counter: value})} from a rocprofv3
    counter_collection.csv."""
    import csv
    out = {}
    for r in csv.DictReader(open(csv_path)):
        d = out.setdefault(r["Dispatch_Id"], (r["Kernel_Name"], {}))
        d[1][r["Counter_Name"]] = d[1].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def _read_bytes(c):
    # TCC read requests by size; FETCH_SIZE counts a 128 B request as 64 B
    return (32 * c.get("TCC_EA0_RDREQ_32B_sum", 0.0) + 64 * c.get("TCC_EA0_RDREQ_64B_sum", 0.0)
            + 128 * c.get("TCC_EA0_RDREQ_128B_sum", 0.0))


def hbm_table(reads_csv, writes_csv, copy_reads_csv, copy_writes_csv, copy_mib=1024):
    """HBM bytes per launch of each solver kernel from two rocprofv3 --pmc
    runs (TCC_EA0_RDREQ_{32B,64B,128B}_sum, then WRITE_SIZE), each scaled by
    what the same counters read for a device-to-device copy of copy_mib MiB
    (bench.py --copy-calibration).  Prints one row per kernel:
    kernel dispatches read_bytes write_bytes traffic_bytes."""
    size = copy_mib * (1 << 20)

    def copies(path, fn):
        v = [fn(c) for name, c in _pmc_per_dispatch(path).values() if "copyBuffer" in name]
        return max(v) if v else None

    rd = copies(copy_reads_csv, _read_bytes)
    wr = copies(copy_writes_csv, lambda c: 1024.0 * c.get("WRITE_SIZE", 0.0))
    fr, fw = size / rd, size / wr
    print("# calibration: a %d MiB device-to-device copy reads %.0f B (32/64/128 B requests) and writes %.0f B "
          "(WRITE_SIZE) -> factors read x%.4f, write x%.4f" % (copy_mib, rd, wr, fr, fw))
    print("# kernel  dispatches  read_bytes  write_bytes  traffic_bytes")
    acc = {}
    for path, fn, k in [(reads_csv, _read_bytes, 0), (writes_csv, lambda c: 1024.0 * c.get("WRITE_SIZE", 0.0), 1)]:
        for name, c in _pmc_per_dispatch(path).values():
            key = next((v for pat, v in KERNEL_KEYS if pat in name), None)
            if key is None:
                continue
            a = acc.setdefault(key, [[], []])
            a[k].append(fn(c))
    for key in dict.fromkeys(v for _, v in KERNEL_KEYS):
        if key not in acc or not acc[key][0] or not acc[key][1]:
            continue
        r = float(np.mean(acc[key][0])) * fr
        w = float(np.mean(acc[key][1])) * fw
        print("%s %d %.0f %.0f %d" % (key, len(acc[key][0]), r, w, int(round(r + w))))


def bandwidth_probe(mib=1024, reps=20):
    """Practical HBM rates on this GPU: torch's own copy (read + write),
    read-only sum and fill of `mib` MiB of float64, timed with events on the
    current stream after three untimed runs each."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    nb = n * 8
    ms = dict(copy=timed(lambda: y.copy_(x)), sum=timed(lambda: x.sum()), fill=timed(lambda: y.fill_(1.0)))
    print(json.dumps({"bytes": nb, "reps": reps, "device": torch.cuda.get_device_name(0),
                      "copy_GBps": round(2 * nb / ms["copy"] / 1e6, 1), "sum_read_GBps": round(nb / ms["sum"] / 1e6, 1),
                      "fill_write_GBps": round(nb / ms["fill"] / 1e6, 1), "ms": {k: round(v, 4) for k, v in ms.items()}}))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh, sc, W0, budget_s=15.0, check_steps=100):
    """The FP64 oracle (OpenMP) on the same mesh from the same state: at
    least check_steps steps, then on until budget_s.  Returns the baseline
    entry and, for the accuracy check, the oracle's state after check_steps
    steps."""
    from oracle import pyoracle as O
    ref = mesh.reference_arrays(sc, W0)
    orc = O.OracleState(ref, sc.R, sc.Cv, boundary_first_order=bool(sc.boundaryFirstOrder))
    del ref
    t0 = time.perf_counter()
    n, check = 0, None
    while True:
        orc.run(1, sc.maxCo)
        n += 1
        if n == check_steps:
            check = dict(U=orc.U.copy(), steps=n)
        if n >= check_steps and time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    entry = dict(value=mesh.nc * n / el / 1e6, unit="Mcell-RK3-steps/s", cores=cores, kind="port",
                 cpu=cpu_model(),
                 sample="%d RK3 steps (with CFL) of the same %d-cell mesh from the same state, FP64 oracle, "
                        "%d OpenMP threads, %.1f s" % (n, mesh.nc, cores, el))
    return entry, check


def mixed_oracle_state(mesh, sc, W0, steps):
    """The FP32-mixed oracle's conserved state after `steps` steps (CFL time
    step of each) from W0: the checker of the mixed path."""
    from oracle import pyoracle as O
    ref = mesh.reference_arrays(sc, W0)
    orc = O.OracleState(ref, sc.R, sc.Cv, boundary_first_order=bool(sc.boundaryFirstOrder), mixed=True)
    del ref
    orc.run(steps, sc.maxCo)
    return orc.U.copy()


def accuracy_vs_oracle(mesh, sc, W0, check, U_mixed=None):
    """A fresh GPU solver from the oracle's starting state W0 (both built by
    the same export), as many device-time-step steps as the oracle ran:
    relative L1 of each conserved variable (density first) against the FP64
    oracle's state, and whether they are equal bit for bit; with U_mixed (the
    FP32-mixed oracle's state after as many steps), the same against it."""
    import afar_cudaeuler2dmixo3rk3solver_amd as A
    g = A.BasicCUDASolver(mesh, sc, W0)
    g.evolve_auto(check["steps"])
    g.synchronize()
    U = g.conserved()[mesh.permutation()]
    g.close()
    Uo = check["U"]
    rel = np.abs(U - Uo).sum(axis=0) / np.abs(Uo).sum(axis=0)
    out = dict(reference="FP64 oracle (CPU), same mesh and starting state", steps=check["steps"],
               rho_L1_rel=float(rel[0]), L1_rel=[float(v) for v in rel], bitwise_equal=bool(np.array_equal(U, Uo)))
    if U_mixed is not None:
        relm = np.abs(U - U_mixed).sum(axis=0) / np.abs(U_mixed).sum(axis=0)
        out["mixed_oracle"] = dict(reference="FP32-mixed oracle (CPU), same mesh, starting state and steps",
                                   rho_L1_rel=float(relm[0]), bitwise_equal=bool(np.array_equal(U, U_mixed)))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--case", default="riemann", choices=["riemann", "corner", "batch"],
                    help="batch: 256 corner flows (5k-200k cells, seed 2), FP32-mixed, autotuned per size bucket")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--tile", type=int, default=64)
    ap.add_argument("--t-start", type=float, default=None,
                    help="simulated time reached (untimed) before the warm-up; default 0.15 for the Riemann case "
                         "(mid-run of t = 0.3, when the waves cover much of the domain), 0 otherwise")
    ap.add_argument("--recon-lanes", type=int, default=0, choices=[0, 1, 4],
                    help="reconstruction kernel: 1 one thread per cell, 4 four per cell; 0 automatic")
    ap.add_argument("--cell-order", type=int, default=0, choices=[0, 1, 2],
                    help="device cell order: 0 compact tiles along a Hilbert curve with 32 (FP64) or 64 (mixed) "
                         "cell blocks ordered by face count and stencil size, 1 the mesh's (RCM) order, "
                         "2 compact tiles in curve order throughout")
    ap.add_argument("--face-quadrature", type=int, default=0, choices=[0, 1],
                    help="flux at the face centre (0, the reference) or at the fOrder 2 rule's two points per face "
                         "(1, extension)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--faithful-boundary", action="store_true",
                    help="riemann case without the boundaryFirstOrder extension (loses positivity)")
    ap.add_argument("--batch-size", type=int, default=256, help="meshes of the batch case")
    ap.add_argument("--batch-fp64", action="store_true", help="batch case in FP64 (default FP32-mixed)")
    ap.add_argument("--save-state", metavar="NPZ", default=None,
                    help="after reaching t-start (untimed), save the conserved state and time to NPZ and exit")
    ap.add_argument("--restart", metavar="NPZ", default=None,
                    help="start from a state saved by --save-state instead of advancing to t-start "
                         "(keeps profiled runs short)")
    ap.add_argument("--precompute-timing", action="store_true",
                    help="only time the RBF precompute, host vs GPU, on the Riemann meshes (250^2 and 1000^2 nodes)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (calibrates HBM byte counters) and exit")
    ap.add_argument("--hbm-table", nargs=4, default=None,
                    metavar=("READS_CSV", "WRITES_CSV", "COPY_READS_CSV", "COPY_WRITES_CSV"),
                    help="only print HBM bytes per launch per kernel from rocprofv3 --pmc counter_collection.csv "
                         "files (no GPU needed)")
    ap.add_argument("--bandwidth-probe", action="store_true",
                    help="only time torch's 1 GiB copy, read-only sum and fill on the GPU (practical HBM rates)")
    args = ap.parse_args()
    if args.hbm_table:
        return hbm_table(*args.hbm_table, copy_mib=args.copy_calibration or 1024)
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (meshes would shard by cell ranges "
              "with halo exchanges, not implemented); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.bandwidth_probe:
        return bandwidth_probe()
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
    if args.precompute_timing:
        return run_precompute_timing()

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
                   reconLanes=args.recon_lanes, cellOrder=args.cell_order, faceQuadrature=args.face_quadrature)
    mesh.precompute(sc)
    t_pre = time.perf_counter() - t0
    print("bench: %s, nc=%d, precompute %.1f s" % (args.case, mesh.nc, t_pre), file=sys.stderr, flush=True)
    gpu = A.BasicCUDASolver(mesh, sc, W0)
    dev_name = A.device_info()[0]
    t_start = args.t_start if args.t_start is not None else (0.15 if args.case == "riemann" else 0.0)
    pre_steps = 0
    if args.restart:
        z = np.load(args.restart, allow_pickle=False)
        if int(z["cells"]) != mesh.nc or str(z["precision"]) != args.precision:
            raise SystemExit("bench: %s holds another workload" % args.restart)
        gpu.set_conserved(z["U"])
        gpu.set_time(float(z["t"]))
        t_start, pre_steps = float(z["t"]), int(z["steps"])
        print("bench: restarted from %s at t = %.4f" % (args.restart, t_start), file=sys.stderr, flush=True)
    elif t_start > 0:
        gpu.run_until(t_start, chunk=100)
        pre_steps = gpu.device_steps()
        gpu.set_time(gpu.device_time())  # clock keeps running, no end time
        print("bench: advanced to t = %.4f in %d steps (untimed)" % (gpu.device_time(), pre_steps),
              file=sys.stderr, flush=True)
    if args.save_state:
        np.savez(args.save_state, U=gpu.conserved(), t=gpu.device_time(), steps=pre_steps, cells=mesh.nc,
                 precision=args.precision)
        print("bench: state at t = %.4f saved to %s" % (gpu.device_time(), args.save_state), file=sys.stderr)
        return
    W_start = gpu.write_field()[0]  # the CPU baseline starts from the same state

    gpu.evolve_auto(args.warmup)
    gpu.synchronize()
    U_t0, t_t0 = gpu.conserved(), gpu.device_time()
    wall0 = time.perf_counter()
    ms_events = gpu.time_steps(args.steps, graph=True)
    gpu.synchronize()
    wall = time.perf_counter() - wall0
    U_t1 = gpu.conserved()
    seconds = ms_events / 1e3
    value = mesh.nc * args.steps / seconds / 1e6
    print("bench: %d steps in %.3f s (events) / %.3f s (wall)" % (args.steps, seconds, wall), file=sys.stderr,
          flush=True)

    field = gpu.write_field()
    W_end = field[0]
    finite = bool(np.isfinite(W_end).all() and (W_end[:, 0] > 0).all() and (W_end[:, 3] > 0).all())
    st = mesh_stats(mesh, field, args.precision)
    t_end = gpu.device_time()
    # Per-kernel durations: the same number of further steps, graph-replayed
    # with an event after every kernel (each kernel's time in the running
    # pipeline, inter-kernel gap included); then back-to-back launches of each
    # kernel alone, reported beside it (and the only figure for the
    # standalone CFL pass, which a device-time-step RK3 step does not run).
    live_ms, live = gpu.time_steps_kernels(args.steps)
    b2b = gpu.profile(reps=20)
    # stability: the timed run again five times from the same state (after
    # everything above, so they change neither value nor the byte model's
    # state); each must end in the same state bit for bit
    reps, same = [], True
    for _ in range(5):
        gpu.set_conserved(U_t0)
        gpu.set_time(t_t0)
        reps.append(gpu.time_steps(args.steps, graph=True) / args.steps)
        same = same and bool(np.array_equal(gpu.conserved(), U_t1))
    w = 8 if args.precision == "fp64" else 4
    ngp = 2 if args.face_quadrature else 0
    lanes = args.recon_lanes or (4 if st["maxC"] <= 4 else 1)
    kb = kernel_bytes(st, w, ngp, geo=w == 8 and ngp == 0 and lanes == 4)
    kernels = {}
    for k, (avg_b2b, cnt) in b2b.items():
        avg = live[k][0] if live[k][1] else avg_b2b
        kernels[k] = dict(avg_ms=avg, count_per_step=cnt, alg_bytes=kb[k], GBps=kb[k] / (avg * 1e-3) / 1e9,
                          back_to_back_ms=avg_b2b)
    # cfl: the standalone pass of host-requested time steps; a device-time-step
    # RK3 step gets its time step from flux_dt + dt_reduce instead (count 0)
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["count_per_step"])
    # the committed counter summaries cover the default configuration only
    default_layout = (not args.face_quadrature and args.cell_order == 0 and args.tile == 64
                      and args.recon_lanes in (0, 4))
    d = kernels[dom]
    # The roofline unit is one RK3 step (one graph launch: every kernel of the
    # step, CFL reduction included).  Its algorithmic bytes are the whole
    # pipeline's compulsory traffic with no intermediate arrays (static data
    # streamed once per stage, state read and written), evaluated on this
    # mesh's stencil and compact sizes, times the cells.
    balg = compulsory_bytes_per_cell_step(st, w)
    step_s = seconds / args.steps
    achieved = balg * mesh.nc / step_s / 1e9
    step_bytes = None
    if default_layout:
        # measured HBM bytes of one RK3 step: each kernel's committed counter
        # summary x its launches per step
        tr = {k: measured_traffic(k, args.case, args.precision) for k in kernels if kernels[k]["count_per_step"]}
        if all(v is not None for v in tr.values()):
            step_bytes = int(sum(tr[k] * kernels[k]["count_per_step"] for k in tr))
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBPS, unit="GB/s",
                    frac=round(achieved / PEAK_HBM_GBPS, 4), traffic=step_bytes,
                    unit_of_work="one RK3 step of %d cells (graph launch, CFL reduction included)" % mesh.nc,
                    alg_bytes_per_cell_step=round(balg, 1), alg_bytes_per_step=int(balg * mesh.nc),
                    avg_step_ms=round(step_s * 1e3, 5),
                    traffic_note="HBM bytes per step from the rocprofv3 counter summary "
                                 "profiles/riemann1000_%s_hbm_traffic.txt (per kernel x launches per step)"
                                 % args.precision if step_bytes else None,
                    traffic_GBps=round(step_bytes / step_s / 1e9, 1) if step_bytes else None,
                    dominant_kernel=dict(kernel=dom, alg_bytes_per_launch=int(d["alg_bytes"]),
                                         avg_launch_ms=round(d["avg_ms"], 5), GBps=round(d["GBps"], 1),
                                         frac=round(d["GBps"] / PEAK_HBM_GBPS, 4),
                                         traffic=measured_traffic(dom, args.case, args.precision)
                                         if default_layout else None,
                                         timing="hipEvents after every kernel of %d graph-replayed steps following "
                                                "the timed ones (%.4f ms per step with the events)"
                                                % (args.steps, live_ms / max(args.steps, 1))))
    whole_step = {"kernel_alg_bytes_per_cell_step": round(
        sum(kb[k] * kernels[k]["count_per_step"] for k in kb) / mesh.nc, 1)}
    cpu, accuracy = None, None
    if not args.no_cpu:
        cpu, check = cpu_baseline(mesh, sc, W_start)
        U_mixed = mixed_oracle_state(mesh, sc, W_start, check["steps"]) if args.precision == "mixed" else None
        gpu.close()  # the accuracy check's solver takes its place in memory
        accuracy = accuracy_vs_oracle(mesh, sc, W_start, check, U_mixed)
    out = {
        "metric": "Mcell-RK3-steps/s", "value": round(value, 2), "unit": "Mcell-RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_events / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-reconstruction/f64-state", "data": "synthetic",
        "config": {"workload": workload, "case": args.case, "cells": mesh.nc, "internal_faces": st["nf"],
                   "boundary_faces": st["nb"], "mean_stencil": round(st["meanS"], 3),
                   "mean_compact": round(st["meanC"], 3), "precision": args.precision, "cfl": 0.5,
                   "tile": args.tile, "recon_lanes": args.recon_lanes or "auto",
                   "cell_order": ["compact tiles (Hilbert), %d-cell blocks by face count and stencil size"
                                  % (32 if args.precision == "fp64" else 64), "mesh (RCM)",
                                  "compact tiles (Hilbert)"][args.cell_order],
                   "flux_points_per_face": ngp or 1, "device": dev_name.strip(),
                   "precompute_s": round(t_pre, 1),
                   "boundary_first_order": bfo, "state_positive_after_timed_steps": finite,
                   "t_start": t_start, "untimed_steps_to_t_start": pre_steps, "t_after_timed_steps": round(t_end, 5),
                   "full_path_cells": st["full_cells"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy": accuracy,
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
        "whole_step": whole_step,
        "repeat_runs": {"ms_per_step": [round(r, 4) for r in reps], "median_ms_per_step": round(float(np.median(reps)), 4),
                        "median_value": round(mesh.nc / (float(np.median(reps)) * 1e-3) / 1e6, 2),
                        "same_final_state": same},
    }
    out["command"] = " ".join(["python", "bench.py"] + sys.argv[1:])
    print(json.dumps(out))


def run_precompute_timing():
    """Host precompute (all host threads) against the GPU precompute, and
    whether the GPU coefficients equal the host's bit for bit."""
    import afar_cudaeuler2dmixo3rk3solver_amd as A
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    sc = A.schemes(cOrder=7, fOrder=2)
    for n in (250, 1000):
        a, _ = cases.riemann_case(n=n, seed=1)
        b, _ = cases.riemann_case(n=n, seed=1)
        t0 = time.perf_counter()
        a.precompute(sc)
        t1 = time.perf_counter()
        b.precompute(sc, gpu=True)
        t2 = time.perf_counter()
        print(json.dumps(dict(n=n, cells=a.nc, host_s=round(t1 - t0, 3), gpu_s=round(t2 - t1, 3),
                              equal=bool(np.array_equal(a.rbf_basis()[0], b.rbf_basis()[0])))), flush=True)


def run_batch(args):
    """256 independent corner flows advanced together (one batch per size
    bucket, tile autotuned per bucket); a step = one RK3 step of every mesh."""
    import afar_cudaeuler2dmixo3rk3solver_amd as A
    from afar_cudaeuler2dmixo3rk3solver_amd import batch as Bt
    log = lambda m: print("bench: " + m, file=sys.stderr, flush=True)
    precision = "mixed" if args.precision == "fp64" and not args.batch_fp64 else args.precision
    # With the reference's boundary treatment 18 of the 256 flows (inflow Mach
    # above about 3 with the steeper corners) lose positivity within 100
    # steps; the boundaryFirstOrder extension keeps every one physical.
    sc_kwargs = dict(cOrder=7, fOrder=2, maxCo=0.5, precision=precision, boundaryFirstOrder=True)
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
    def physical(U):
        return bool(np.isfinite(U).all() and (U[:, 0] > 0).all())
    n_physical = sum(physical(b.conserved(k)) for _, b in bb.batches for k in range(len(b.meshes)))
    k_small = int(np.argmin([m.nc for m in meshes]))  # its state is checked against the oracle below
    U_small = bb.conserved(k_small)[meshes[k_small].permutation()]
    bb.close()
    # The roofline unit is one RK3 step of every mesh: the pipeline's
    # compulsory bytes (no intermediate arrays) of each mesh, summed.
    w = 8 if precision == "fp64" else 4
    step_bytes = sum(compulsory_bytes_per_cell_step(mesh_stats(m), w) * m.nc for m in meshes)
    step_s = wall / args.steps
    achieved = step_bytes / step_s / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBPS, unit="GB/s",
                    frac=round(achieved / PEAK_HBM_GBPS, 4), traffic=None,
                    unit_of_work="one RK3 step of all %d meshes (%d cells)" % (len(meshes), cells),
                    alg_bytes_per_step=int(step_bytes), alg_bytes_per_cell_step=round(step_bytes / cells, 1),
                    avg_step_ms=round(step_s * 1e3, 5),
                    traffic_note="no counter summary for the batch: HBM bytes per step not measured")
    cpu, accuracy = None, None
    if not args.no_cpu:
        # the oracle on the median-sized mesh of the batch, from its starting state
        k = int(np.argsort([m.nc for m in meshes])[len(meshes) // 2])
        cpu, _ = cpu_baseline(meshes[k], A.schemes(**dict(sc_kwargs, precision="fp64")), W0s[k])
        cpu["sample"] = "mesh %d of the batch (median size): " % k + cpu["sample"]
        # the smallest mesh's state after the timed steps against the oracle of the batch's precision
        from oracle import pyoracle as O
        sc = A.schemes(**sc_kwargs)
        orc = O.OracleState(meshes[k_small].reference_arrays(sc, W0s[k_small]), sc.R, sc.Cv,
                            boundary_first_order=True, mixed=precision == "mixed")
        orc.run(args.steps, sc.maxCo)
        rel = np.abs(U_small - orc.U).sum(axis=0) / np.abs(orc.U).sum(axis=0)
        accuracy = dict(reference="%s oracle (CPU), mesh %d of the batch (the smallest, %d cells), same starting "
                                  "state and steps" % ("FP32-mixed" if precision == "mixed" else "FP64", k_small,
                                                       meshes[k_small].nc),
                        steps=args.steps, rho_L1_rel=float(rel[0]), bitwise_equal=bool(np.array_equal(U_small, orc.U)))
    out = {
        "metric": "Mcell-RK3-steps/s", "value": round(value, 2), "unit": "Mcell-RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f32-reconstruction/f64-state", "data": "synthetic",
        "config": {"workload": "%d corner flows, cells log-uniform 5e3-2e5, Mach U[1.5,4], angle U[5,25] deg, seed 2"
                               % len(meshes), "case": "batch", "meshes": len(meshes), "cells": cells,
                   "precision": precision, "cfl": 0.5, "device": A.device_info()[0].strip(),
                   "precompute_s": round(t_pre, 1), "autotune_s": round(t_tune, 1),
                   "buckets": {str(k): v for k, v in bb.tuning.items()}, "boundary_first_order": True,
                   "meshes_finite_positive_after_timed_steps": n_physical,
                   "timing": "host wall clock around the enqueue of all buckets' steps and their completion"},
        "roofline": roofline, "cpu_baseline": cpu, "accuracy": accuracy,
    }
    out["command"] = " ".join(["python", "bench.py"] + sys.argv[1:])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
