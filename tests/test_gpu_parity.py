"""GPU parity: the HIP path (through the C-ABI) against the FP64 oracle.

Bar (FP64 path): bit-exact reconstruction-switch flags and time steps, and the
conserved state within 1e-12 relative L1 (in practice the two are bitwise
equal, since both sides run the same IEEE operations in the same order).
"""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def rel_l1(a, b):
    return np.abs(a - b).sum(axis=0) / np.maximum(np.abs(b).sum(axis=0), 1e-300)


def make_pair(afar, mesh, W0, sc):
    from afar_cudaeuler2dmixo3rk3solver_amd.solver import BasicCUDASolver
    mesh.precompute(sc)
    ref = mesh.reference_arrays(sc, W0)
    orc = O.OracleState(ref, sc.R, sc.Cv)
    gpu = BasicCUDASolver(mesh, sc, W0)
    return gpu, orc, ref


def gpu_state_mesh_order(gpu, nc):
    """W (mesh order) and conserved U (mesh order) from the device."""
    W = gpu.fetch("W", np.float64, (nc + gpu.nb, 4))
    U = gpu.fetch("U", np.float64, (nc, 4))
    return W, U


def test_math_bit_exact(afar):
    say("case: device exp/erfc/sqrt/div vs host, 40000 points (FP64) and 50000 (float)")
    import ctypes as C
    from afar_cudaeuler2dmixo3rk3solver_amd._lib import check, dptr, lib
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-30, 30, 10000), rng.uniform(-1, 1, 10000), rng.uniform(-745, 709, 10000),
                        rng.normal(0, 3, 10000)])
    dev = np.zeros((len(x), 4))
    host = np.zeros((len(x), 4))
    check(lib().afar_selftest_math(dptr(x), dptr(dev), len(x)))
    check(lib().afar_host_math(dptr(x), dptr(host), len(x)))
    assert np.array_equal(dev.view(np.uint64), host.view(np.uint64))
    for i, v in enumerate(x[:50]):
        assert host[i, 0] == O.lib().oracle_exp(v)
        assert host[i, 1] == O.lib().oracle_erfc(v)
    # the float functions of the FP32-mixed flux
    from afar_cudaeuler2dmixo3rk3solver_amd._lib import ptr
    xf = np.concatenate([x, rng.uniform(-104, 90, 10000)]).astype(np.float32)
    devf = np.zeros((len(xf), 4), np.float32)
    hostf = np.zeros((len(xf), 4), np.float32)
    check(lib().afar_selftest_math_f32(ptr(xf, C.c_float), ptr(devf, C.c_float), len(xf)))
    check(lib().afar_host_math_f32(ptr(xf, C.c_float), ptr(hostf, C.c_float), len(xf)))
    assert np.array_equal(devf.view(np.uint32), hostf.view(np.uint32))


def test_stage_kernels_corner(afar):
    """Each kernel of one RK stage against the oracle, after 5 steps of evolution."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    say("case: corner 200x100 seed 0, stage kernels after 5 steps")
    mesh, W0 = cases.corner_case()
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    gpu, orc, ref = make_pair(afar, mesh, W0, sc)
    nc = mesh.nc
    for _ in range(5):
        dt_o = orc.adjust_time_step(sc.maxCo)
        dt_g = gpu.set_delta_t()
        assert dt_o == dt_g
        orc.evolve(dt_o)
        gpu.evolve(dt_g)
    gpu.synchronize()
    W, U = gpu_state_mesh_order(gpu, nc)
    assert np.array_equal(U, orc.U)
    assert np.array_equal(W, orc.a["FD"])
    # K5
    orc.reconstruct()
    gpu.run_stage(0)
    g = np.concatenate([gpu.fetch(n, np.float64, (nc, 2)) for n in ("gp", "gu", "gv", "gT")], axis=1)
    go = orc.a["gradFD"][:, [0, 1, 2, 4, 3, 5, 6, 7]]  # (px,py),(Uxx,Uyx),(Uxy,Uyy),(Tx,Ty)
    assert np.array_equal(g, go)
    assert np.array_equal(gpu.fetch("lim", np.float64, (nc, 4)), orc.a["limiter"])
    assert np.array_equal(gpu.fetch("smooth", np.int8, nc + gpu.nb)[:nc], orc.smooth_flags)
    assert np.array_equal(gpu.fetch("fv", np.float64, (ref["sizes"]["nf"], 2, 4)), orc.a["faceFD"])
    # K6
    orc.bvd()
    gpu.run_stage(1)
    ftype = gpu.fetch("ftype", np.int8, nc)
    assert np.array_equal(ftype, orc.final_types)
    # the device keeps the raw K5 limiter; the face pass applies K6's
    # transform (cudaLimiter.h:453-461) as it reads it: the same values
    raw = gpu.fetch("lim", np.float64, (nc, 4))
    assert np.all(raw >= 0)
    y, sh = raw, (ftype < 0).astype(np.float64)[:, None]
    bj = np.fmin(2.0 * y, 1.0)
    venkat = np.fmin((y * y + y + 1e-10) / (y * y + 0.5 + 0.5 * y + 1e-10), 1.0)
    assert np.array_equal(sh * venkat + (1.0 - sh) * bj, orc.a["limiter"])
    # K7
    orc.flux()
    gpu.run_stage(2)
    assert np.array_equal(gpu.fetch("flux", np.float64, orc.a["Flux"].shape), orc.a["Flux"])


def run_parity(afar, mesh, W0, sc, nsteps, label):
    gpu, orc, ref = make_pair(afar, mesh, W0, sc)
    nc = mesh.nc
    say("case: %s, %d steps, nc=%d" % (label, nsteps, nc))
    for s in range(nsteps):
        dt_o = orc.adjust_time_step(sc.maxCo)
        dt_g = gpu.set_delta_t()
        assert dt_o == dt_g, (s, dt_o, dt_g)
        orc.evolve(dt_o)
        gpu.evolve(dt_g)
    gpu.synchronize()
    W, U = gpu_state_mesh_order(gpu, nc)
    err = rel_l1(U, orc.U)
    say("  rel L1 (rho, rhoU, rhoV, rhoE) = %s" % err)
    assert np.all(err <= 1e-12)
    _, _, shock = gpu.write_field()
    perm = mesh.permutation()
    types_caller = np.zeros(nc, np.int8)
    types_caller[perm] = orc.final_types
    assert np.array_equal(shock, types_caller)
    assert np.array_equal(gpu.fetch("smooth", np.int8, nc + gpu.nb)[:nc], orc.smooth_flags)
    return gpu, orc


def test_corner_parity_500_steps(afar):
    """Config 0: 20,000-cell jittered corner, FP64, 500 steps."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    mesh, W0 = cases.corner_case()
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    gpu, orc = run_parity(afar, mesh, W0, sc, 500, "corner 200x100 seed 0")
    assert np.array_equal(gpu.conserved()[mesh.permutation()], orc.U)


def test_riemann_small_parity(afar):
    """Mixed triangle/quad mesh with the four-quadrant Riemann problem (KFVS
    active).  40 steps: on this 96x96 mesh the reference algorithm loses
    positivity at step 44 (the oracle does too)."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    mesh, W0 = cases.riemann_case(n=96, seed=1)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    gpu, orc = run_parity(afar, mesh, W0, sc, 40, "riemann 96x96 seed 1")
    assert (orc.final_types < 0).sum() > 0  # the KFVS branch ran


def fan_case(afar, k=20):
    """k triangles around one vertex inside a slip wall: every cell's
    edgeNeighbour stencil holds the other k - 1 triangles and its ghost, so
    k = 20 gives the largest stencil the reference accepts (20,
    basicCUDAsolver.cpp:393-397) in every cell, all in one partial tile."""
    ang = np.arange(k) * 2 * np.pi / k
    pts = np.vstack([[0.0, 0.0], np.stack([np.cos(ang), np.sin(ang)], 1)])
    cells = []
    for i in range(k):
        cells += [0, 1 + i, 1 + (i + 1) % k]
    be = [[1 + i, 1 + (i + 1) % k] for i in range(k)]
    mesh = afar.Mesh.from_polygons(pts, np.arange(0, 3 * k + 1, 3), cells, be, [0] * k, [2])
    th = (np.arange(k) + 0.5) * 2 * np.pi / k
    W0 = np.stack([1.0 + 0.3 * np.cos(th) + 0.1 * np.sin(3 * th), 0.1 * np.sin(th), -0.2 * np.cos(th),
                   1.0 + 0.2 * np.sin(2 * th)], 1)
    return mesh, W0


@pytest.mark.parametrize("lanes", [1, 4])
def test_max_stencil_fan_parity(afar, lanes):
    mesh, W0 = fan_case(afar)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5, reconLanes=lanes)
    gpu, orc = run_parity(afar, mesh, W0, sc, 200, "fan of 20 triangles (stencils of 20), lanes %d" % lanes)
    assert mesh.sizes()["maxS"] == 20
    assert np.array_equal(gpu.conserved()[mesh.permutation()], orc.U)
    assert (orc.final_types != 0).any() and (orc.final_types == 0).any()  # both branches ran


@pytest.mark.parametrize("split_p", [0.0, 1.0])
def test_periodic_vortex_parity(afar, split_p):
    """Isentropic vortex on a fully periodic jittered mesh (no boundary faces:
    nb = 0, every stencil wraps), quads or triangles, 60 steps."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    mesh, W0 = cases.vortex_case(n=24, seed=4, split_p=split_p)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    gpu, orc = run_parity(afar, mesh, W0, sc, 60, "periodic vortex 24x24, split_p %.1f" % split_p)
    assert mesh.sizes()["nb"] == 0
    assert np.array_equal(gpu.conserved()[mesh.permutation()], orc.U)


def test_periodic_vortex_device_time_step(afar):
    """Periodic vortex with the device time step (graph replays): the faces
    across the wrap take the shifted neighbour centre in their offsets and in
    the CFL bound's centre distance, which the face pass forms on the device;
    time steps and state equal the oracle's host loop."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    say("case: periodic vortex 24x24 split_p 0.5, 40 device-dt graph steps")
    mesh, W0 = cases.vortex_case(n=24, seed=4, split_p=0.5)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    gpu, orc, ref = make_pair(afar, mesh, W0, sc)
    assert np.any(ref["faceShift"] != 0.0)
    dts = orc.run(40, sc.maxCo)
    gpu.evolve_auto(40)
    gpu.synchronize()
    assert gpu.last_dt() == dts[-1]
    assert np.array_equal(gpu.conserved()[mesh.permutation()], orc.U)


def test_riemann_quads_to_t03_parity(afar):
    """Jittered quad mesh, four-quadrant Riemann problem run to t = 0.3."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    mesh, W0 = cases.riemann_case(n=64, seed=1, split_p=0.0)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    gpu, orc, ref = make_pair(afar, mesh, W0, sc)
    say("case: riemann quads 64x64 seed 1 to t=0.3")
    t = 0.0
    while t < 0.3:
        dt_o = orc.adjust_time_step(sc.maxCo)
        dt_g = gpu.set_delta_t()
        assert dt_o == dt_g
        dt = min(dt_o, 0.3 - t)
        orc.evolve(dt)
        gpu.evolve(dt)
        t += dt
    gpu.synchronize()
    W, U = gpu_state_mesh_order(gpu, mesh.nc)
    assert np.array_equal(U, orc.U)
    assert np.array_equal(gpu.fetch("ftype", np.int8, mesh.nc), orc.final_types)
    assert np.all(orc.W[:, 0] > 0) and np.all(orc.W[:, 3] > 0)


def test_evolve_auto_equals_host_dt(afar):
    """Device-resident time step with graph replay gives the same trajectory."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    from afar_cudaeuler2dmixo3rk3solver_amd.solver import BasicCUDASolver
    say("case: corner 100x50, evolve_auto vs set_delta_t+evolve, 40 steps")
    mesh, W0 = cases.corner_case(nx=100, ny=50)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    mesh.precompute(sc)
    a = BasicCUDASolver(mesh, sc, W0)
    b = BasicCUDASolver(mesh, sc, W0)
    dt_a = 0.0
    for _ in range(40):
        dt_a = a.set_delta_t()
        a.evolve(dt_a)
    b.evolve_auto(40)
    dt_b = b.last_dt()  # time step of the 40th step
    a.synchronize()
    assert dt_b == dt_a
    assert np.array_equal(a.conserved(), b.conserved())


def test_graph_replays_each_equal_eager(afar):
    """Every replay of the captured step graph equals the eager step (host
    time step) bit for bit, with eager work of another solver and host
    synchronisations between the replays."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    from afar_cudaeuler2dmixo3rk3solver_amd.solver import BasicCUDASolver
    say("case: corner 80x40, 8 single-step graph replays vs eager steps, eager work in between")
    mesh, W0 = cases.corner_case(nx=80, ny=40)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    mesh.precompute(sc)
    a = BasicCUDASolver(mesh, sc, W0)  # eager
    b = BasicCUDASolver(mesh, sc, W0)  # graph
    c = BasicCUDASolver(mesh, sc, W0)  # unrelated eager work
    for i in range(8):
        dt_a = a.set_delta_t()
        a.evolve(dt_a)
        a.synchronize()
        b.evolve_auto(1)
        c.evolve(c.set_delta_t())
        b.synchronize()
        assert b.last_dt() == dt_a, i
        assert np.array_equal(a.conserved(), b.conserved()), i
        c.synchronize()
    for g in (a, b, c):
        g.close()


def test_riemann_boundary_first_order_to_t03(afar):
    """Mixed mesh, four-quadrant problem with the boundaryFirstOrder extension,
    run to t = 0.3 on the device and on the oracle: bitwise equal, positive."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    from afar_cudaeuler2dmixo3rk3solver_amd.solver import BasicCUDASolver
    mesh, W0 = cases.riemann_case(n=96, seed=1)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5, boundaryFirstOrder=True)
    mesh.precompute(sc)
    orc = O.OracleState(mesh.reference_arrays(sc, W0), sc.R, sc.Cv, boundary_first_order=True)
    gpu = BasicCUDASolver(mesh, sc, W0)
    say("case: riemann 96x96 seed 1, boundaryFirstOrder, to t=0.3")
    t, n = 0.0, 0
    while t < 0.3:
        dt_o = orc.adjust_time_step(sc.maxCo)
        assert dt_o == gpu.set_delta_t()
        dt = min(dt_o, 0.3 - t)
        orc.evolve(dt)
        gpu.evolve(dt)
        t += dt
        n += 1
    gpu.synchronize()
    W, U = gpu_state_mesh_order(gpu, mesh.nc)
    say("  %d steps" % n)
    assert np.array_equal(U, orc.U)
    assert np.array_equal(gpu.fetch("ftype", np.int8, mesh.nc), orc.final_types)
    assert np.all(orc.W[:, 0] > 0) and np.all(orc.W[:, 3] > 0)
    assert (orc.final_types < 0).sum() > 0


def test_device_clock_matches_host_loop(afar):
    """run_until (device time steps clipped to the end time on the device)
    follows the host loop dt = min(dt, t_end - t) exactly, and steps past the
    end time change nothing."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    from afar_cudaeuler2dmixo3rk3solver_amd.solver import BasicCUDASolver
    say("case: riemann quads 64x64 to t=0.05, device clock vs host loop")
    mesh, W0 = cases.riemann_case(n=64, seed=1, split_p=0.0)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    mesh.precompute(sc)
    a = BasicCUDASolver(mesh, sc, W0)
    b = BasicCUDASolver(mesh, sc, W0)
    t, n = 0.0, 0
    while t < 0.05:
        dt = min(a.set_delta_t(), 0.05 - t)
        a.evolve(dt)
        t += dt
        n += 1
    a.synchronize()
    launches = b.run_until(0.05, chunk=16)
    assert launches >= n
    assert b.device_time() == t
    assert b.device_steps() == n
    assert np.array_equal(a.conserved(), b.conserved())
    b.evolve_auto(5)  # past the end: no change
    assert np.array_equal(a.conserved(), b.conserved())
    assert b.device_time() == t


def test_residual_norm_matches_oracle(afar):
    import ctypes as C
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    from afar_cudaeuler2dmixo3rk3solver_amd.solver import BasicCUDASolver
    say("case: corner 100x50, density residual norm vs oracle after 20 steps")
    mesh, W0 = cases.corner_case(nx=100, ny=50)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    gpu, orc, ref = make_pair(afar, mesh, W0, sc)
    for _ in range(20):
        dt = orc.adjust_time_step(sc.maxCo)
        assert dt == gpu.set_delta_t()
        orc.evolve(dt)
        gpu.evolve(dt)
    r_gpu = gpu.residual_norm()
    orc.reconstruct()
    orc.bvd()
    orc.flux()
    O.lib().oracle_evaluate_residual(C.byref(orc.f))
    q = orc.a["Res"][:, 0] / ref["cellVolume"]
    r_orc = float(np.sqrt(np.mean(q * q)))
    say("  residual %.6e (gpu) %.6e (oracle)" % (r_gpu, r_orc))
    assert abs(r_gpu - r_orc) <= 1e-12 * r_orc
    assert np.array_equal(gpu.fetch("ftype", np.int8, mesh.nc), orc.final_types)


def tiny_case(afar, kind):
    """The smallest meshes: one unit square whose four faces are all
    boundary faces (no internal face: the CFL bound has none to take, so the
    time step is the reference's 1e5 cap, cudaFieldOperation.h:124-180), two
    squares sharing one face, one triangle, a strip of three quads."""
    if kind == "one square":
        pts = [[0, 0], [1, 0], [1, 1], [0, 1]]
        ptr, cells, be = [0, 4], [0, 1, 2, 3], [[0, 1], [1, 2], [2, 3], [3, 0]]
        bp, labels = [0, 0, 0, 0], [4]
    elif kind == "two squares":
        pts = [[0, 0], [1, 0], [2, 0], [2, 1], [1, 1], [0, 1]]
        ptr, cells = [0, 4, 8], [0, 1, 4, 5, 1, 2, 3, 4]
        be, bp, labels = [[0, 1], [1, 2], [2, 3], [3, 4], [4, 5], [5, 0]], [0, 0, 1, 0, 0, 1], [2, 4]
    elif kind == "one triangle":
        pts = [[0, 0], [1, 0.2], [0.3, 1]]
        ptr, cells, be, bp, labels = [0, 3], [0, 1, 2], [[0, 1], [1, 2], [2, 0]], [0, 0, 0], [0]
    else:  # strip of three quads
        pts = [[0, 0], [1, 0], [2, 0.1], [3, 0], [3, 1], [2, 1.1], [1, 1], [0, 1]]
        ptr, cells = [0, 4, 8, 12], [0, 1, 6, 7, 1, 2, 5, 6, 2, 3, 4, 5]
        be = [[0, 1], [1, 2], [2, 3], [3, 4], [4, 5], [5, 6], [6, 7], [7, 0]]
        bp, labels = [0, 0, 0, 1, 0, 0, 0, 2], [2, 4, 1]
    mesh = afar.Mesh.from_polygons(np.array(pts, float), np.array(ptr), np.array(cells), np.array(be),
                                   np.array(bp), labels)
    nc = mesh.nc
    k = np.arange(nc)
    W0 = np.stack([1.0 + 0.2 * k, 0.3 - 0.1 * k, 0.05 * k, 1.4 + 0.1 * (k % 2)], 1)
    return mesh, W0


@pytest.mark.parametrize("lanes", [1, 4])
@pytest.mark.parametrize("kind", ["one square", "two squares", "one triangle", "strip of three quads"])
def test_tiny_meshes_parity(afar, kind, lanes):
    mesh, W0 = tiny_case(afar, kind)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5, reconLanes=lanes)
    mesh.precompute(sc)
    gpu, orc, ref = make_pair(afar, mesh, W0, sc)
    say("case: %s (nc=%d, nf=%d, nb=%d), lanes %d, 20 steps" % (kind, mesh.nc, mesh.sizes()["nf"],
                                                                 mesh.sizes()["nb"], lanes))
    for s in range(20):
        dt_o = orc.adjust_time_step(sc.maxCo)
        dt_g = gpu.set_delta_t()
        assert dt_o == dt_g, (s, dt_o, dt_g)
        orc.evolve(dt_o)
        gpu.evolve(dt_g)
    if mesh.sizes()["nf"] == 0:
        assert dt_g == 1e5
    U = gpu.conserved()[mesh.permutation()]
    assert np.all(np.isfinite(U))
    assert np.array_equal(U, orc.U)
    g = afar.BasicCUDASolver(mesh, sc, W0)  # device time steps: the same trajectory
    g.evolve_auto(20)
    assert np.array_equal(g.conserved()[mesh.permutation()], orc.U)
