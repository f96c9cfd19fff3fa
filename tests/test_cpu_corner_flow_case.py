"""The reference's example case (simulation/cornerFlow/cornerFlow_cuda) at
reduced resolution, written by cases.write_corner_flow_case without
OpenFOAM: its polyMesh, 0/ fields and dictionaries read back through the
library's case I/O with the case's patches, labels and values, and the
oracle runs it.

Finding: with the case's own maxCo 0.3 the algorithm loses positivity in its
first step.  The first time step is set by the quiescent interior (sound
speed 1), while the inflow enters at |u| + c = 4.08 + 2.45 = 6.5, so the
inflow cells run at a Courant number of about 2 whatever the resolution
(an explicit scheme is positive only up to 1).  maxCo 0.15 keeps the inflow
Courant number below 1 and runs to the case's end time."""
import os

import numpy as np

from oracle import pyoracle as O


def read_case(afar, d, co):
    m = afar.Mesh.read_foam(d)
    W0, pv, pf = m.read_foam_fields(d, "0")
    for p, lab in enumerate([5, 2, 4]):  # fvSchemes CUDA.boundary.patch0..2
        m.set_patch(p, lab)
    sc = afar.schemes(cOrder=7, fOrder=5, maxCo=co)
    m.precompute(sc)
    return m, sc, W0, pv, pf


def test_case_files_read_back(afar, tmp_path):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    d = str(tmp_path)
    cases.write_corner_flow_case(d, nx=40, ny=20)
    m, sc, W0, pv, pf = read_case(afar, d, 0.3)
    s = m.sizes()
    # two blocks of 40 x 20: 1600 quads; inlet = upper half of x = 0 (20 faces), obstacle = lower
    # half of x = 0 plus the walls y = 0 and y = 1 (20 + 40 + 40), outlet = x = 1 (40)
    assert (s["nc"], s["nf"], s["nb"]) == (1600, 2 * 40 * 39, 160)
    assert [n for _, n in m.patches()] == [20, 100, 40, 0]
    assert m.patches()[-1][0] == "frontAndBack"
    assert np.all(W0 == [1.0, 0.0, 0.0, 1.0])
    assert pf.tolist() == [True, False, False, False]
    assert np.array_equal(pv[0], cases.CORNER_FLOW_INLET)
    fv = open(os.path.join(d, "system", "fvSchemes")).read()
    for key in ("cOrder 7;", "fOrder 5;", "stencilType  edgeNeighbour;", "patch0  5;", "patch1  2;", "patch2  4;"):
        assert key in fv, key
    cd = open(os.path.join(d, "system", "controlDict")).read()
    assert "maxCo           0.3;" in cd and "adjustTimeStep  yes;" in cd
    # the labels reach the boundary table: inlet ghosts carry label 5's state
    ref = m.reference_arrays(sc, W0)
    types = ref["boundaryFacesType"]
    assert sorted(np.unique(types).tolist()) == [2, 4, 5]
    assert np.all(ref["FD"][s["nc"]:][types == 5] == cases.CORNER_FLOW_INLET)


def run_case(orc, co, end=0.15, write=0.01):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    n = 0
    for t, dt, _ in cases.foam_time_loop(end, write, lambda t: orc.adjust_time_step(co)):
        orc.evolve(dt)
        n += 1
        W = orc.W
        if not (np.isfinite(W).all() and (W[:, 0] > 0).all() and (W[:, 3] > 0).all()):
            return n, t, False
    return n, t, True


def test_case_settings_lose_positivity_in_first_step(afar, tmp_path):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    d = str(tmp_path)
    cases.write_corner_flow_case(d, nx=40, ny=20)
    m, sc, W0, _, _ = read_case(afar, d, 0.3)
    orc = O.OracleState(m.reference_arrays(sc, W0), sc.R, sc.Cv)
    dt = orc.adjust_time_step(0.3)
    # the first time step is the quiescent interior's: 0.3 h / 1 with h = 1/40
    assert abs(dt - 0.3 / 40) < 1e-9
    n, t, ok = run_case(orc, 0.3)
    assert not ok and n == 1


def test_case_runs_to_end_time_at_half_courant(afar, tmp_path):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    d = str(tmp_path)
    cases.write_corner_flow_case(d, nx=40, ny=20, max_co=0.15)
    m, sc, W0, _, _ = read_case(afar, d, 0.15)
    orc = O.OracleState(m.reference_arrays(sc, W0), sc.R, sc.Cv)
    n, t, ok = run_case(orc, 0.15)
    assert ok and abs(t - 0.15) < 1e-12
    # the inflow has entered: compressed gas behind the shock in the upper half, a flow
    # around the corner into the lower half, a quiescent far corner
    xy, _ = m.cells()
    nc = m.nc
    W = orc.W
    rho = W[:, 0] / (sc.R * W[:, 3])
    x, y = xy[:nc, 0], xy[:nc, 1]
    assert rho[(x < 0.2) & (y > 0.6)].min() > 5.0
    assert W[(x < 0.3) & (y < 0.45) & (y > 0.3), 2].min() < -0.5  # turning down around the corner
    far = (x > 0.9) & (y < 0.1)
    assert np.allclose(W[far], [1.0, 0.0, 0.0, 1.0], atol=1e-12)
    assert n > 200  # time steps to t = 0.15 at maxCo 0.15 on h = 1/40
