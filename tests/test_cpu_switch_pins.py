"""Hand-computed pins of the oracle's switch logic: the limiter, the
smoothness flag, the BVD switch, the CFL bound and the boundary ghosts.

The cases run on an unjittered n x n grid of squares of side h = 0.1 (cell
centres at odd multiples of h/2), where an interior cell's compact stencil is
its four face neighbours at (+-h, 0), (0, +-h), its least-squares gradient is
the central difference ((E - W)/2h, (N - S)/2h) and its face offsets are
(+-h/2, 0), (0, +-h/2).  Each expected value below follows from that
arithmetic and the reference's formulas, worked out in the comments:

  limiterProcess        src/CUDA/cudaEuler2DmixO3RK3/cudaLimiter.h:35-50
  smoothness flag       cudaLimiter.h:286-313
  BVD switch, transform cudaLimiter.h:385-461 (and 9-32)
  CFL bound             cudaFieldOperation.h:145-157
  boundary ghosts       src/CUDA/BCs/2D/*.h, table cudaEulerBCs.h:34-39

The pin_* functions raise AssertionError when the oracle disagrees;
tests/test_cpu_oracle_mutations.py runs them against deliberately broken
builds of the oracle to show that each catches its slip.  Each pin also runs
on the FP32-mixed oracle (mixed=True): the same expected values, with
tolerances of float round-off where a value is not exact.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as O

H = 0.1
N = 10


def _grid(labels=None, values=None, n=N, jitter=0.0, seed=1):
    import afar_cudaeuler2dmixo3rk3solver_amd as A
    m = A.Mesh.riemann(n=n, jitter=jitter, seed=seed, split_p=0.0)
    for patch, label in enumerate(labels or []):
        m.set_patch(patch, label, None if values is None else values[patch])
    sc = A.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    m.precompute(sc)
    return m, sc


def _oracle(m, sc, field, mixed=False):
    """Oracle state (FP64, or FP32-mixed with ``mixed``) whose cell values are
    field(x, y) -> (nc, 4) at the cell centres (mesh order); ghosts from the
    boundary table."""
    xy, _ = m.cells()
    nc = m.nc
    Wm = field(xy[:nc, 0], xy[:nc, 1])
    W0 = np.empty_like(Wm)
    W0[m.permutation()] = Wm
    orc = O.OracleState(m.reference_arrays(sc, W0), sc.R, sc.Cv, mixed=mixed)
    return orc, xy


def _cell_at(xy, nc, x, y):
    d = np.hypot(xy[:nc, 0] - x, xy[:nc, 1] - y)
    c = int(np.argmin(d))
    assert d[c] < 1e-9
    return c


def _col(x, a, b):
    """a where x is left of b's column centre ... (step functions by column)."""
    return np.where(x < a, 0, np.where(x < b, 1, 2))


def pin_limiter_branches(mixed=False):
    """Raw K5 limiters of the interior cell c at (0.45, 0.55).

    p: 0.9 left of c's column, 1.0 in it, 1.3 right of it.  gx = 0.4/(2h),
    gy = 0.  East face: phif - pc = gx h/2 = 0.1 >= 0, d1 = pMax - pc = 0.3,
    y = 0.5*0.3/0.1 = 1.5.  West face: phif - pc = -0.1 < 0, d1 = pMin - pc =
    -0.1, y = 0.5*(-0.1)/(-0.1) = 0.5.  North/south: d2 = 0 < SMALL, y = 1.
    Raw limiter min(2, 1.5, 0.5, 1, 1) = 0.5 (the phif < phic branch binds).

    Ux: -0.3 | 0 | 0.1.  gx = 0.4/(2h); east: d2 = 0.1, d1 = UxMax - 0 = 0.1,
    y = 0.5; west: d2 = -0.1, d1 = -0.3, y = 1.5.  Raw 0.5 (the phif >= phic
    branch binds).

    Uy: 0 | 0 | 1e-6.  Both x faces have |d2| = 2.5e-7 < SMALL = 1e-5, so
    y = 1 (without the guard the west face would give d1 = 0, y = 0): raw 1
    exactly, and the range 1e-6 < SMALL makes the face values the unlimited
    linear ones, Uy_c + gx*dx = +-2.5e-7.

    T = 1 + 0.5 y, linear: north/south y = 0.5*(0.5h)/(0.5*0.5h/1)... = 1,
    east/west guarded: raw 1 (to round-off)."""
    m, sc = _grid()
    nc = m.nc

    def field(x, y):
        col = _col(x, 0.4, 0.5)
        p = np.choose(col, [0.9, 1.0, 1.3])
        ux = np.choose(col, [-0.3, 0.0, 0.1])
        uy = np.choose(col, [0.0, 0.0, 1e-6])
        return np.stack([p, ux, uy, 1.0 + 0.5 * y], axis=1)

    orc, xy = _oracle(m, sc, field, mixed)
    c = _cell_at(xy, nc, 0.45, 0.55)
    orc.reconstruct()
    lim = orc.a["limiter"][c]
    tol = 1e-6 if mixed else 1e-8  # round-off of the float or double arithmetic
    assert abs(lim[0] - 0.5) < tol, lim
    assert abs(lim[1] - 0.5) < tol, lim
    assert lim[2] == 1.0, lim
    assert abs(lim[3] - 1.0) < tol, lim
    # the SMALL range rule: Uy face values are the linear extrapolation
    fv = orc.a["faceFD"]
    cf = orc.a["cellFaces"][c][: orc.a["compactStencilSize"][c]]
    for sf in cf:
        f, side = abs(sf) - 1, int(sf < 0)
        xf = orc.a["FACE"][f, 0]
        expect = 2.5e-7 * np.sign(xf - 0.45) if abs(xf - 0.45) > 1e-9 else 0.0
        assert abs(fv[f, side, 2] - expect) < (1e-13 if mixed else 1e-15), (sf, fv[f, side, 2], expect)


def pin_limiter_cap(mixed=False):
    """The cap y <= 2 (cudaLimiter.h:47).  Corner cell c at (0.05, 0.05) with
    ghosts set by hand: pseudo-cells L at (0, 0.05) and B at (0.05, 0) (h/2
    away) carry the extremes, and boundary faces are not limited.  p: c = 20,
    L = 29, B = 9, E = 24.6, N = 14.6.  Sxx = Syy = h^2/4 + h^2 = 1.25 h^2,
    Sxy = 0; gx = (-h/2*9 + h*4.6)/(1.25 h^2) = 0.08/h, gy = (-h/2*(-11) +
    h*(-5.4))/(1.25 h^2) = 0.08/h.  East face: d2 = 0.04, d1 = 29 - 20 = 9,
    0.5*9/0.04 = 112.5 -> 2; north face the same.  Raw limiter exactly 2."""
    m, sc = _grid()
    nc = m.nc

    def field(x, y):
        p = np.full_like(x, 20.0)
        p[(np.abs(x - 0.15) < 1e-9) & (np.abs(y - 0.05) < 1e-9)] = 24.6
        p[(np.abs(x - 0.05) < 1e-9) & (np.abs(y - 0.15) < 1e-9)] = 14.6
        return np.stack([p, np.zeros_like(x), np.zeros_like(x), 1.0 + 0.5 * x + 0.25 * y], axis=1)

    orc, xy = _oracle(m, sc, field, mixed)
    c = _cell_at(xy, nc, 0.05, 0.05)
    FD = orc.a["FD"]
    for g in range(nc, len(FD)):
        if abs(xy[g, 0]) < 1e-12 and abs(xy[g, 1] - 0.05) < 1e-9:
            FD[g, 0] = 29.0
        if abs(xy[g, 1]) < 1e-12 and abs(xy[g, 0] - 0.05) < 1e-9:
            FD[g, 0] = 9.0
    orc.reconstruct()
    assert orc.a["limiter"][c, 0] == 2.0, orc.a["limiter"][c]


def pin_smoothness_kink(mixed=False):
    """Smoothness flag on p = 1 + s |x - 0.45| (kink at the centre of column
    4), T = p / R (rho = 1).  At the kink cell gx = 0, so phif = pc at every
    face; the east face gives |(1/2) pc + (1/2) pE - pc| / h = (s h/2)/h =
    s/2, the north/south faces 0.  Normalised by (pMax + pMin)/2 = (2 + s h)/2
    the value is s/(2 + s h): > 1 exactly when s > 2/(1 - h) = 2.22.  Every
    other interior column sees a linear field, smoothness 0.  So s = 2.5
    (value 1.11) flags the kink column only, s = 2.0 (0.91) flags nothing."""
    m, sc = _grid()
    nc = m.nc
    for s, expect in ((2.5, 1), (2.0, 0)):
        def field(x, y, s=s):
            p = 1.0 + s * np.abs(x - 0.45)
            return np.stack([p, np.zeros_like(x), np.zeros_like(x), p / sc.R], axis=1)

        orc, xy = _oracle(m, sc, field, mixed)
        orc.reconstruct()
        flags = orc.smooth_flags
        x = xy[:nc, 0]
        inner = (x > 0.2) & (x < 0.8)  # columns away from the pseudo-cells of the side walls
        kink = np.abs(x - 0.45) < 1e-9
        assert np.all(flags[inner & kink] == expect), (s, flags[inner & kink])
        assert np.all(flags[inner & ~kink] == 0), s


def _bvd_case(Tc, Tcur, Tnei, raw_T, flags_on=(), raw_other=0.3, mixed=False):
    """BVD of the interior cell c at (0.45, 0.55) with hand-set inputs:
    zero gradients (so Tf_cur_TVD = Tc), the raw limiter (raw_other for p,
    Ux, Uy; raw_T for T), smoothness flags 1 on the cells at the positions
    in flags_on, and on c's east face the O3 temperatures Tcur (c's side) and
    Tnei (the neighbour's side); c's other faces are flat (both sides equal).
    Returns (final type, transformed limiter of c)."""
    m, sc = _grid()
    nc = m.nc
    orc, xy = _oracle(m, sc, lambda x, y: np.stack([np.ones_like(x), np.zeros_like(x), np.zeros_like(x),
                                                    np.ones_like(x)], axis=1), mixed)
    c = _cell_at(xy, nc, 0.45, 0.55)
    a = orc.a
    a["gradFD"][:] = 0.0
    a["limiter"][:] = raw_other
    a["limiter"][c, 3] = raw_T
    a["FD"][c, 3] = Tc
    si = a["shockIndicator"]
    si[:] = 0
    for (x, y) in flags_on:
        si[2 * _cell_at(xy, nc, x, y)] = 1
    a["faceFD"][:, :, 3] = 1.0
    for sf in a["cellFaces"][c][: a["compactStencilSize"][c]]:
        f, side = abs(sf) - 1, int(sf < 0)
        if abs(a["FACE"][f, 0] - 0.5) < 1e-9:  # east face
            a["faceFD"][f, side, 3] = Tcur
            a["faceFD"][f, 1 - side, 3] = Tnei
    orc.bvd()
    return int(orc.final_types[c]), a["limiter"][c].copy()


def _bj(y):
    return min(2.0 * y, 1.0)


def _venkat(y):
    return min((y * y + y + 1e-10) / (y * y + 0.5 + 0.5 * y + 1e-10), 1.0)


def pin_bvd_truth_table(mixed=False):
    """Each factor of the BVD switch (cudaLimiter.h:431-451) on c's east face,
    with T_TVD = Tc (zero gradient), l = min(2 raw_T, 1):
      typeval   = |Tc - Tnei| < 0.5 |Tcur - Tnei|
      isNotFlat = |Tcur - Tnei| > 1e-10
      isLimited = (l < 0.9) + (extended shock count > 0)
      final     = -2 [compact count > 0][extended count > 1] + [BVD > 0]
    c's compact stencil is its face neighbours, the extended-only members
    its diagonal neighbours (edgeNeighbour)."""
    D, E = (0.35, 0.45), (0.55, 0.55)  # a diagonal and the east neighbour of c
    rows = [
        # Tc,   Tcur, Tnei, raw_T, flags,    final   why
        (1.9, 1.0, 2.0, 0.40, (), 1),           # |-0.1| < 0.5: typeval; l = 0.8 < 0.9: limited
        (1.9, 1.0, 2.0, 0.45, (), 0),           # l = 0.9 is not < 0.9: not limited
        (1.45, 1.0, 2.0, 0.40, (), 0),          # |-0.55| < 0.5 fails: typeval 0
        (1.9, 2.0, 2.0, 0.40, (), 0),           # Tcur = Tnei: flat face
        (1.9, 1.0, 2.0, 0.45, (D,), 1),         # limited by the extended count (1 > 0)
        (1.45, 1.0, 2.0, 0.40, ((0.45, 0.55), D), -2),  # compact count 1 > 0, extended 2 > 1: shock
        (1.9, 1.0, 2.0, 0.40, ((0.45, 0.55), D), -1),   # shock and BVD: -2 + 1
        (1.9, 1.0, 2.0, 0.45, ((0.45, 0.55),), 1),      # extended count 1 is not > 1: no shock; limited by it
        (1.45, 1.0, 2.0, 0.40, (E,), 0),        # compact count 1, extended 1: no shock
        (1.45, 1.0, 2.0, 0.40, (E, D), -2),     # compact 1, extended 2: shock
    ]
    for Tc, Tcur, Tnei, raw_T, flags, expect in rows:
        got, lim = _bvd_case(Tc, Tcur, Tnei, raw_T, flags, mixed=mixed)
        assert got == expect, (Tc, Tcur, Tnei, raw_T, flags, got, expect)
        # the transform of cudaLimiter.h:453-461: Venkatakrishnan for shock types
        tr = _venkat if expect < 0 else _bj
        tol = 1e-7 if mixed else 1e-15
        assert abs(lim[0] - tr(0.3)) < tol and abs(lim[3] - tr(raw_T)) < tol, (expect, lim)
    # the two transforms at the raw value 0.3: BJ 0.6, Venkatakrishnan 0.39/0.74
    assert abs(_venkat(0.3) - 0.39 / 0.74) < 1e-9 and _bj(0.3) == 0.6


def pin_bvd_corner_cells(mixed=False):
    """A cell with two boundary faces takes type 1 (cudaLimiter.h:419-429)
    whatever its faces hold; a cell with one does not."""
    m, sc = _grid()
    nc = m.nc
    orc, xy = _oracle(m, sc, lambda x, y: np.stack([np.ones_like(x), np.zeros_like(x), np.zeros_like(x),
                                                    np.ones_like(x)], axis=1), mixed)
    orc.reconstruct()
    orc.bvd()
    t = orc.final_types
    for (x, y) in ((0.05, 0.05), (0.95, 0.05), (0.05, 0.95), (0.95, 0.95)):
        assert t[_cell_at(xy, nc, x, y)] == 1
    assert t[_cell_at(xy, nc, 0.05, 0.45)] == 0
    assert t[_cell_at(xy, nc, 0.45, 0.45)] == 0


def pin_cfl_brute_force(mixed=False):
    """adjustTimeStep against a numpy evaluation of cudaFieldOperation.h:
    145-157 over every internal face of jittered 3 x 3 and 6 x 6 meshes with
    random states: dt = min(1e5, min_f CFL dr |S| / ((1 - a)|U_o.S| +
    a |U_n.S| + sqrt(gamma R ((1 - a) T_o + a T_n)) |S|)), a = |x_f - x_o|/dr;
    boundary faces do not take part.  Also per face on a 2-cell strip."""
    rng = np.random.default_rng(7)
    for n in (3, 6):
        m, sc = _grid(n=n, jitter=0.2, seed=3)
        nc = m.nc

        def field(x, y):
            k = len(x)
            return np.stack([rng.uniform(0.5, 2.0, k), rng.uniform(-3, 3, k), rng.uniform(-3, 3, k),
                             rng.uniform(0.5, 4.0, k)], axis=1)

        orc, xy = _oracle(m, sc, field, mixed)
        a = orc.a
        nf = orc.s["nf"]
        own, nei = a["neighbour"][:, 0], a["neighbour"][:, 1]
        F = a["FACE"][:nf]
        W = a["FD"].astype(np.float32).astype(np.float64) if mixed else a["FD"]  # the states the bound reads
        xo, xn = xy[own], xy[nei] + a["faceShift"]
        dr = np.sqrt(((xn - xo) ** 2).sum(axis=1))
        alpha = np.sqrt(((F[:, :2] - xo) ** 2).sum(axis=1)) / dr
        S = F[:, 3:5]
        un = np.abs((W[own, 1:3] * S).sum(axis=1))
        unn = np.abs((W[nei, 1:3] * S).sum(axis=1))
        gam = sc.R / sc.Cv + 1.0
        T = (1 - alpha) * W[own, 3] + alpha * W[nei, 3]
        dtf = 0.5 * dr * F[:, 2] / ((1 - alpha) * un + alpha * unn + np.sqrt(gam * sc.R * T) * F[:, 2])
        expect = min(1e5, dtf.min())
        got = orc.adjust_time_step(0.5)
        assert abs(got - expect) <= 1e-13 * expect, (n, got, expect)
        # the weights matter here: the face at the minimum is off centre
        swapped = 0.5 * dr * F[:, 2] / (alpha * un + (1 - alpha) * unn + np.sqrt(gam * sc.R * (
            alpha * W[own, 3] + (1 - alpha) * W[nei, 3])) * F[:, 2])
        assert abs(swapped.min() - expect) > 1e-6 * expect


def pin_boundary_ghosts(mixed=False):
    """Ghost states of every boundary label from the adjacent cell's state
    w = (p, Ux, Uy, T) (src/CUDA/BCs/2D/*.h): 0 (p, 0, 0, 1); 1 (p, 0, 0, T);
    2 (p, U - (U.n) n, T), so U_g.n = 0 and U_g.t = U.t; 3 (1, 3, 0, 1);
    4 w; 5 (30.05945, 4.07795, 0, 5.97677); 6 the patch's value."""
    rng = np.random.default_rng(11)
    v6 = [(0.7, 1.5, -0.5, 2.0)] * 4
    for labels in ((0, 1, 3, 5), (2, 6, 4, 2)):
        m, sc = _grid(labels=labels, values=v6, n=6, jitter=0.2, seed=5)
        nc = m.nc

        def field(x, y):
            k = len(x)
            return np.stack([rng.uniform(0.5, 2.0, k), rng.uniform(-3, 3, k), rng.uniform(-3, 3, k),
                             rng.uniform(0.5, 4.0, k)], axis=1)

        orc, xy = _oracle(m, sc, field, mixed)
        a = orc.a
        nf = orc.s["nf"]
        a["FD"][nc:] = -99.0
        O.lib().oracle_update_boundary(C.byref(orc.f), sc.R, sc.Cv)
        for b in range(orc.s["nb"]):
            w = a["FD"][a["boundaryFaceNeiLabel"][b]]
            g = a["FD"][nc + b]
            lab = a["boundaryFacesType"][b]
            n = a["FACE"][nf + b, 3:5] / a["FACE"][nf + b, 2]
            if lab == 0:
                assert np.array_equal(g, [w[0], 0.0, 0.0, 1.0])
            elif lab == 1:
                assert np.array_equal(g, [w[0], 0.0, 0.0, w[3]])
            elif lab == 2:
                assert g[0] == w[0] and g[3] == w[3]
                assert abs(g[1] * n[0] + g[2] * n[1]) < 1e-14
                t = (-n[1], n[0])
                assert abs((g[1] - w[1]) * t[0] + (g[2] - w[2]) * t[1]) < 1e-14
            elif lab == 3:
                assert np.array_equal(g, [1.0, 3.0, 0.0, 1.0])
            elif lab == 4:
                assert np.array_equal(g, w)
            elif lab == 5:
                assert np.array_equal(g, [30.05945, 4.07795, 0.0, 5.97677])
            else:
                assert np.array_equal(g, v6[0])
        assert set(np.unique(a["boundaryFacesType"])) == set(labels)


def pin_transformed_limiters_bounded(mixed=False):
    """After K6 every limiter lies in [0, 1] (both transforms are min(., 1) of
    a non-negative value), on a jittered triangle/quad Riemann run."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    import afar_cudaeuler2dmixo3rk3solver_amd as A
    m, W0 = cases.riemann_case(n=24, seed=1)
    sc = A.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    m.precompute(sc)
    orc = O.OracleState(m.reference_arrays(sc, W0), sc.R, sc.Cv, mixed=mixed)
    orc.run(5, sc.maxCo)
    orc.reconstruct()
    raw = orc.a["limiter"].copy()
    assert raw.min() >= 0.0 and raw.max() <= 2.0
    orc.bvd()
    lim = orc.a["limiter"]
    assert lim.min() >= 0.0 and lim.max() <= 1.0
    assert (orc.final_types < 0).any() and (orc.final_types > 0).any()


PINS = [pin_limiter_branches, pin_limiter_cap, pin_smoothness_kink, pin_bvd_truth_table, pin_bvd_corner_cells,
        pin_cfl_brute_force, pin_boundary_ghosts, pin_transformed_limiters_bounded]


@pytest.fixture(params=[False, True], ids=["fp64", "mixed"])
def mixed(request):
    return request.param


def test_limiter_branches(afar, mixed):
    pin_limiter_branches(mixed)


def test_limiter_cap(afar, mixed):
    pin_limiter_cap(mixed)


def test_smoothness_kink(afar, mixed):
    pin_smoothness_kink(mixed)


def test_bvd_truth_table(afar, mixed):
    pin_bvd_truth_table(mixed)


def test_bvd_corner_cells(afar, mixed):
    pin_bvd_corner_cells(mixed)


def test_cfl_brute_force(afar, mixed):
    pin_cfl_brute_force(mixed)


def test_boundary_ghosts(afar, mixed):
    pin_boundary_ghosts(mixed)


def test_transformed_limiters_bounded(afar, mixed):
    pin_transformed_limiters_bounded(mixed)
