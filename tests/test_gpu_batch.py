"""GPU: batched solves and the four-lane reconstruction kernel.

- every mesh of a batch follows exactly (bitwise) the trajectory it follows
  alone, time steps included, in FP64 and in FP32-mixed;
- the four-lanes-per-cell reconstruction gives the one-lane kernel's values
  bit for bit (flags, gradients, limiters, face values, state).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def corner_meshes(afar, sc):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    specs = [dict(nx=40, ny=20, mach_u=2.0, angle_deg=15.0, seed=0),
             dict(nx=24, ny=12, mach_u=3.5, angle_deg=8.0, seed=5),
             dict(nx=64, ny=30, mach_u=1.6, angle_deg=22.0, seed=9),
             dict(nx=30, ny=16, mach_u=2.7, angle_deg=12.0, seed=3)]
    meshes, W0s = [], []
    for s in specs:
        m, w = cases.corner_case(**s)
        m.precompute(sc)
        meshes.append(m)
        W0s.append(w)
    return meshes, W0s


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_alone(afar, precision):
    from afar_cudaeuler2dmixo3rk3solver_amd.batch import BatchSolver
    from afar_cudaeuler2dmixo3rk3solver_amd.solver import BasicCUDASolver
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5, precision=precision, tile=128)
    meshes, W0s = corner_meshes(afar, sc)
    say("case: batch of 4 corner meshes (%s), 60 steps, vs each alone" % precision)
    b = BatchSolver(meshes, sc, W0s)
    assert b.sizes()["ninst"] == 4
    b.evolve(60)
    b.synchronize()
    dts = b.last_dt()
    for k, (m, w) in enumerate(zip(meshes, W0s)):
        g = BasicCUDASolver(m, sc, w)
        g.evolve_auto(60)
        assert g.last_dt() == dts[k]
        assert np.array_equal(g.conserved(), b.conserved(k)), k
        Wg, _, sg = g.write_field()
        Wb, sb = b.write_field(k)
        assert np.array_equal(Wg, Wb)
        assert np.array_equal(sg, sb)
        g.close()
    # reset returns to the initial state: the same 60 steps again
    U0 = [b.conserved(k) for k in range(4)]
    b.reset()
    b.evolve(60)
    b.synchronize()
    for k in range(4):
        assert np.array_equal(b.conserved(k), U0[k])


def test_batch_rejects_tiny_meshes(afar):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    from afar_cudaeuler2dmixo3rk3solver_amd._lib import AfarError
    from afar_cudaeuler2dmixo3rk3solver_amd.batch import BatchSolver
    say("case: batch with a 6x3 mesh (too few faces) is refused")
    sc = afar.schemes(cOrder=7, fOrder=2)
    m, w = cases.corner_case(nx=6, ny=3)
    m.precompute(sc)
    with pytest.raises(AfarError):
        BatchSolver([m], sc, [w])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("lanes", [4])
def test_reconstruct_lanes_equal(afar, precision, lanes):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    from afar_cudaeuler2dmixo3rk3solver_amd.solver import BasicCUDASolver
    say("case: riemann 96x96, reconstruction kernel 1 vs %d (%s), stage kernels and 30 steps" % (lanes, precision))
    mesh, W0 = cases.riemann_case(n=96, seed=1)
    sc1 = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5, precision=precision, reconLanes=1, boundaryFirstOrder=True,
                       tile=64)
    scl = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5, precision=precision, reconLanes=lanes,
                       boundaryFirstOrder=True, tile=64)
    mesh.precompute(sc1)
    a = BasicCUDASolver(mesh, sc1, W0)
    b = BasicCUDASolver(mesh, scl, W0)
    a.evolve_auto(30)
    b.evolve_auto(30)
    a.run_stage(0)
    b.run_stage(0)
    nc, nf = mesh.nc, mesh.sizes()["nf"]
    dt = np.float64 if precision == "fp64" else np.float32
    for name, shape in [("gp", (nc, 2)), ("gu", (nc, 2)), ("gv", (nc, 2)), ("gT", (nc, 2)), ("lim", (nc, 4)),
                        ("fv", (nf, 2, 4))]:
        assert np.array_equal(a.fetch(name, dt, shape), b.fetch(name, dt, shape)), name
    assert np.array_equal(a.fetch("smooth", np.int8, nc + a.nb), b.fetch("smooth", np.int8, nc + b.nb))
    assert np.array_equal(a.conserved(), b.conserved())
    assert a.last_dt() == b.last_dt()
