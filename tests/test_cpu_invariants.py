"""Invariants of the oracle's full RK3 step, checked on the CPU.

- free-stream preservation: a uniform flow with transmissive boundaries on a
  jittered mixed mesh stays uniform to round-off;
- discrete conservation: the face-to-cell gather sums to the boundary fluxes
  (internal fluxes cancel exactly, so the residual sum is the boundary sum up
  to summation rounding);
- mirror symmetry: the four-quadrant Riemann problem is symmetric about
  x = y; on a mesh symmetric about the diagonal the solution stays symmetric;
- the seam-1 export reproduces the case's inflow patch value (label 6).
"""
import numpy as np
import pytest

from oracle import pyoracle as O


def test_free_stream_preservation(afar):
    m = afar.Mesh.riemann(n=30, jitter=0.2, seed=4, split_p=0.5)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    m.precompute(sc)
    W0 = np.tile([0.9, 0.7, -0.4, 0.9 / (1.2 * sc.R)], (m.nc, 1))
    orc = O.OracleState(m.reference_arrays(sc, W0), sc.R, sc.Cv)
    U0 = orc.U.copy()
    orc.run(10, sc.maxCo)
    assert np.abs(orc.U - U0).max() < 1e-12 * np.abs(U0).max()


def test_discrete_conservation(afar):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    m, W0 = cases.riemann_case(n=32, seed=1)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    m.precompute(sc)
    orc = O.OracleState(m.reference_arrays(sc, W0), sc.R, sc.Cv)
    orc.run(3, sc.maxCo)
    orc.reconstruct()
    orc.bvd()
    orc.flux()
    import ctypes as C
    O.lib().oracle_evaluate_residual(C.byref(orc.f))
    nf = orc.s["nf"]
    total_res = orc.a["Res"].sum(axis=0)
    total_bnd = orc.a["Flux"][nf:].sum(axis=0)
    scale = np.abs(orc.a["Flux"]).sum(axis=0)
    assert np.all(np.abs(total_res - total_bnd) <= 1e-13 * scale)


def test_mirror_symmetry_riemann(afar):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    # unjittered, unsplit square grid: symmetric about x = y
    m = afar.Mesh.riemann(n=24, jitter=0.0, seed=1, split_p=0.0)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    m.precompute(sc)
    xy_all, _ = m.cells()
    nc = m.nc
    perm = m.permutation()
    xy = np.zeros((nc, 2))
    xy[perm] = xy_all[:nc]
    W0 = cases.riemann_state(xy, sc.R)
    orc = O.OracleState(m.reference_arrays(sc, W0), sc.R, sc.Cv)
    orc.run(20, sc.maxCo)
    W = orc.W  # mesh order
    xm = xy_all[:nc]
    key = {(round(x, 9), round(y, 9)): i for i, (x, y) in enumerate(xm)}
    mirror = np.array([key[(round(y, 9), round(x, 9))] for x, y in xm])
    Wm = W[mirror]
    assert np.allclose(W[:, 0], Wm[:, 0], rtol=0, atol=1e-11)
    assert np.allclose(W[:, 1], Wm[:, 2], rtol=0, atol=1e-11)
    assert np.allclose(W[:, 3], Wm[:, 3], rtol=0, atol=1e-11)
    # the flags are symmetric too
    t = orc.final_types
    assert np.array_equal(t, t[mirror])


def test_inflow_patch_value(afar):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    m, W0 = cases.corner_case(nx=20, ny=10)
    sc = afar.schemes(cOrder=7, fOrder=2)
    m.precompute(sc)
    a = m.reference_arrays(sc, W0)
    nc = m.nc
    inlet = a["boundaryFacesType"] == 6
    assert inlet.sum() == 10
    assert np.allclose(a["FD"][nc:][inlet], [1 / 1.4, 2.0, 0.0, (1 / 1.4) / sc.R])
    # T = p/(rho R) with rho = 1
    slip = a["boundaryFacesType"] == 2
    assert slip.sum() == 40


def test_boundary_first_order_keeps_riemann_positive(afar):
    """The reference algorithm loses positivity in boundary triangles of the
    four-quadrant problem (the ghost-state boundary flux against the
    reconstructed interior faces); the boundaryFirstOrder extension keeps the
    run positive to t = 0.3, and leaves cells without boundary faces alone."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    mesh, W0 = cases.riemann_case(n=48, seed=1)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    mesh.precompute(sc)
    ref = mesh.reference_arrays(sc, W0)
    plain = O.OracleState(ref, sc.R, sc.Cv)
    fixed = O.OracleState(ref, sc.R, sc.Cv, boundary_first_order=True)
    t = 0.0
    lost = False
    while t < 0.3:
        dt = fixed.adjust_time_step(sc.maxCo)
        dt = min(dt, 0.3 - t)
        fixed.evolve(dt)
        if not lost:
            plain.evolve(plain.adjust_time_step(sc.maxCo))
            Wp = plain.W
            lost = bool(np.isnan(Wp).any() or (Wp[:, 0] <= 0).any() or (Wp[:, 3] <= 0).any())
        t += dt
    W = fixed.W
    assert np.all(W[:, 0] > 0) and np.all(W[:, 3] > 0)
    assert lost, "the unmodified algorithm was expected to lose positivity on this mesh"
