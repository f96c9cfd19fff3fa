"""Benchmark and test cases: meshes with their initial primitive states.

All states are (p, Ux, Uy, T) per cell in the caller's cell order, with the
reference's gas constants R = 0.7142857, Cv = 1.7857142857 (gamma = 1.4;
simulation/cornerFlow/cornerFlow_cuda/system/fvSchemes:25-26) unless given.
"""
import math

import numpy as np

from .solver import Mesh

R_REF = 0.7142857
CV_REF = 1.7857142857


def _centroids(mesh):
    """Cell centroids in the caller's order."""
    xy, _ = mesh.cells()
    nc = mesh.nc
    out = np.zeros((nc, 2))
    out[mesh.permutation()] = xy[:nc]
    return out


def uniform_state(nc, p, ux, uy, T):
    W = np.empty((nc, 4))
    W[:] = (p, ux, uy, T)
    return W


def corner_case(nx=200, ny=100, seed=0, jitter=0.15, mach_u=2.0, angle_deg=15.0, corner_x=0.5, top_label=2,
                R=R_REF):
    """Supersonic flow over an expansion corner: freestream rho = 1,
    p = 1/1.4, u = mach_u, v = 0, T = p/(rho R); inflow patch label 6."""
    p0, rho0 = 1.0 / 1.4, 1.0
    T0 = p0 / (rho0 * R)
    mesh = Mesh.corner(nx=nx, ny=ny, jitter=jitter, seed=seed, corner_x=corner_x, angle_deg=angle_deg,
                       inflow=(p0, mach_u, 0.0, T0), top_label=top_label)
    return mesh, uniform_state(nx * ny, p0, mach_u, 0.0, T0)


# four-quadrant Riemann problem: (rho, u, v, p) in quadrants 1..4 counter-clockwise from x>.5, y>.5
RIEMANN_STATES = {1: (1.5, 0.0, 0.0, 1.5), 2: (0.5323, 1.206, 0.0, 0.3), 3: (0.138, 1.206, 1.206, 0.029),
                  4: (0.5323, 0.0, 1.206, 0.3)}


def riemann_state(xy, R=R_REF):
    x, y = xy[:, 0], xy[:, 1]
    q = np.where(x > 0.5, np.where(y > 0.5, 1, 4), np.where(y > 0.5, 2, 3))
    W = np.empty((len(xy), 4))
    for k, (rho, u, v, p) in RIEMANN_STATES.items():
        m = q == k
        W[m] = (p, u, v, p / (rho * R))
    return W


def riemann_case(n=1000, seed=1, jitter=0.2, split_p=0.5, R=R_REF):
    mesh = Mesh.riemann(n=n, jitter=jitter, seed=seed, split_p=split_p)
    return mesh, riemann_state(_centroids(mesh), R)


def vortex_exact(xy, t, beta=5.0, gamma=1.4, L=10.0, x0=5.0, y0=5.0, u_inf=1.0, v_inf=1.0, R=R_REF):
    """Isentropic vortex on (rho, u, v, p) = (1, 1, 1, 1) at point positions xy,
    advected periodically; returns (p, Ux, Uy, T) with T = p/(rho R)."""
    xc = x0 + u_inf * t
    yc = y0 + v_inf * t
    dx = (xy[:, 0] - xc + 0.5 * L) % L - 0.5 * L
    dy = (xy[:, 1] - yc + 0.5 * L) % L - 0.5 * L
    r2 = dx * dx + dy * dy
    e = np.exp(0.5 * (1.0 - r2))
    du = -beta / (2 * math.pi) * e * dy
    dv = beta / (2 * math.pi) * e * dx
    Tstd = 1.0 - (gamma - 1.0) * beta * beta / (8 * gamma * math.pi ** 2) * np.exp(1.0 - r2)
    rho = Tstd ** (1.0 / (gamma - 1.0))
    p = rho * Tstd
    W = np.stack([p, u_inf + du, v_inf + dv, p / (rho * R)], axis=1)
    return W, rho


def triangle_rule(n):
    """Conical product rule on the triangle (0,0), (1,0), (0,1): Gauss-Legendre
    in u and v mapped by (xi, eta) = (u, v (1 - u)) with Jacobian (1 - u);
    n*n points, exact for polynomials of degree 2n - 2 (weights sum to 1/2)."""
    x, w = np.polynomial.legendre.leggauss(n)
    u, wu = 0.5 * (x + 1.0), 0.5 * w
    U, V = np.meshgrid(u, u, indexing="ij")
    WU, WV = np.meshgrid(wu, wu, indexing="ij")
    return U.ravel(), (V * (1.0 - U)).ravel(), (WU * WV * (1.0 - U)).ravel()


def cell_average(mesh, fn, n=5):
    """Cell averages of fn(x, y) -> (npts, m) over each polygonal cell (fan
    triangulation, conical product rule per triangle), in the caller's order."""
    vptr, vxy = mesh.polygons()
    nc = mesh.nc
    nv = np.diff(vptr)
    xi, eta, wq = triangle_rule(n)
    acc = None
    area = np.zeros(nc)
    v0 = vxy[vptr[:-1]]
    for k in range(1, int(nv.max()) - 1):
        live = nv > k + 1
        idx = np.where(live, vptr[:-1] + k, vptr[:-1])
        a = vxy[idx] - v0
        b = vxy[np.where(live, vptr[:-1] + k + 1, vptr[:-1])] - v0
        det = np.where(live, a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0], 0.0)
        pts = v0[:, None, :] + xi[None, :, None] * a[:, None, :] + eta[None, :, None] * b[:, None, :]
        vals = fn(pts[..., 0].ravel(), pts[..., 1].ravel())
        vals = vals.reshape(nc, len(wq), -1)
        part = np.einsum("q,cqm->cm", wq, vals) * det[:, None]
        acc = part if acc is None else acc + part
        area += 0.5 * det
    avg = acc / area[:, None]
    out = np.empty_like(avg)
    out[mesh.permutation()] = avg
    return out


def vortex_conserved(x, y, t=0.0, beta=5.0, gamma=1.4, L=10.0, x0=5.0, y0=5.0, u_inf=1.0, v_inf=1.0):
    """Conserved (rho, rho u, rho v, rho E) of the isentropic vortex at points
    (x, y); E = p/(rho (gamma - 1)) + |u|^2/2."""
    W, rho = vortex_exact(np.stack([x, y], axis=1), t, beta, gamma, L, x0, y0, u_inf, v_inf, R=1.0)
    p, u, v = W[:, 0], W[:, 1], W[:, 2]
    return np.stack([rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)], axis=1)


def primitive_from_conserved(U, R=R_REF, Cv=CV_REF):
    """(p, Ux, Uy, T) whose conserved state, as the solver forms it, is U."""
    gamma = R / Cv + 1.0
    rho = U[:, 0]
    ux, uy = U[:, 1] / rho, U[:, 2] / rho
    p = (gamma - 1.0) * (U[:, 3] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / rho)
    return np.stack([p, ux, uy, p / (R * rho)], axis=1)


def vortex_case(n=64, seed=0, jitter=0.15, L=10.0, R=R_REF, averaged=False, quad=5, split_p=1.0):
    """Isentropic vortex (strength 5) on the doubly periodic [0, L]^2 mesh of
    2 n^2 jittered triangles.  averaged: the initial state is the cell average
    of the conserved variables (quad x quad conical rule per triangle); else
    point values at the centroids.  Returns mesh, W0 (and with averaged, the
    exact cell-average conserved state, which is also the exact solution at
    t = L, one period of the (1, 1) advection)."""
    mesh = Mesh.periodic(n=n, L=L, jitter=jitter, seed=seed, split_p=split_p)
    if not averaged:
        W, _ = vortex_exact(_centroids(mesh), 0.0, L=L, R=R)
        return mesh, W
    Ubar = cell_average(mesh, lambda x, y: vortex_conserved(x, y, L=L), n=quad)
    return mesh, primitive_from_conserved(Ubar, R), Ubar


# ---------------------------------------------------------------- the reference's example case

# inflow state of the cornerFlow case's inlet patch (p, Ux, Uy, T): the fixedValue of its 0/ fields,
# which is also the value of boundary label 5 (src/CUDA/BCs/2D/fixedValue2Boundary.h)
CORNER_FLOW_INLET = (30.05945, 4.07795, 0.0, 5.97677)


def corner_flow_mesh(nx=500, ny=250, renumber=True):
    """The geometry of the reference's example case
    (simulation/cornerFlow/cornerFlow_cuda/system/blockMeshDict:18-75): the
    unit square as two blocks of nx x ny quads, y in [0, 0.5] and [0.5, 1]
    (the reference has nx = 500, ny = 250).  Patches in the case's order:
    0 inlet (x = 0, y > 0.5; label 5), 1 obstacle (x = 0, y < 0.5, and the
    walls y = 0 and y = 1; label 2), 2 outlet (x = 1; label 4).  Cells are
    RCM-renumbered, as the case's `renumberMesh -overwrite` does."""
    nyt = 2 * ny
    xs = np.linspace(0.0, 1.0, nx + 1)
    ys = np.concatenate([np.linspace(0.0, 0.5, ny + 1), np.linspace(0.5, 1.0, ny + 1)[1:]])
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel()], axis=1)

    def P(i, j):
        return i * (nyt + 1) + j

    I, J = np.meshgrid(np.arange(nx), np.arange(nyt), indexing="ij")
    I, J = I.T.ravel(), J.T.ravel()  # blockMesh order: x fastest, block 0 (lower) first
    quads = np.stack([P(I, J), P(I + 1, J), P(I + 1, J + 1), P(I, J + 1)], axis=1)
    cell_ptr = np.arange(0, 4 * len(quads) + 1, 4)
    bedges, bpatch = [], []
    for j in range(nyt):
        bedges.append((P(0, j), P(0, j + 1)))
        bpatch.append(0 if j >= ny else 1)
        bedges.append((P(nx, j), P(nx, j + 1)))
        bpatch.append(2)
    for i in range(nx):
        bedges.append((P(i, 0), P(i + 1, 0)))
        bpatch.append(1)
        bedges.append((P(i, nyt), P(i + 1, nyt)))
        bpatch.append(1)
    return Mesh.from_polygons(pts, cell_ptr, quads.ravel(), bedges, bpatch, [5, 2, 4], renumber=renumber)


def _foam_header(cls, obj):
    return ("FoamFile\n{\n    version     2.0;\n    format      ascii;\n    class       %s;\n"
            "    object      %s;\n}\n\n" % (cls, obj))


def write_corner_flow_case(case_dir, nx=500, ny=250, end_time=0.15, write_interval=0.01, max_co=0.3,
                           c_order=7, f_order=5, R=R_REF, Cv=CV_REF):
    """A complete case directory with the settings of the reference's
    cornerFlow case: constant/polyMesh of corner_flow_mesh(nx, ny) (what
    blockMesh and renumberMesh give), 0/{p,U,T} (quiescent p = 1, U = 0,
    T = 1; inlet fixedValue CORNER_FLOW_INLET, obstacle zeroGradient / slip,
    outlet zeroGradient; 0/p:19-35, 0/U:19-35, 0/T:19-35 of the case),
    system/controlDict (adjustable time step with maxCo, endTime, writes
    every write_interval; controlDict:25-51) and system/fvSchemes
    (gaussQuadrature cOrder/fOrder; CUDA R, Cv, edgeNeighbour, recalcBasis,
    boundary patch0 5, patch1 2, patch2 4; fvSchemes:17-36).  Returns the
    mesh (mesh order = file order)."""
    import os
    mesh = corner_flow_mesh(nx, ny)
    mesh.write_foam(case_dir)
    names = [n for n, _ in mesh.patches()]
    os.makedirs(os.path.join(case_dir, "0"), exist_ok=True)
    os.makedirs(os.path.join(case_dir, "system"), exist_ok=True)
    inlet, obstacle, outlet = names[:3]
    p_in, ux_in, uy_in, T_in = CORNER_FLOW_INLET

    def field(cls, obj, dims, internal, inlet_value, obstacle_type):
        return (_foam_header(cls, obj) + "dimensions      %s;\n\ninternalField   %s;\n\nboundaryField\n{\n"
                "    %s\n    {\n        type            fixedValue;\n        value           %s;\n    }\n"
                "    %s\n    {\n        type            %s;\n    }\n"
                "    %s\n    {\n        type            zeroGradient;\n    }\n"
                "    frontAndBack\n    {\n        type            empty;\n    }\n}\n"
                % (dims, internal, inlet, inlet_value, obstacle, obstacle_type, outlet))

    with open(os.path.join(case_dir, "0", "p"), "w") as f:
        f.write(field("volScalarField", "p", "[1 -1 -2 0 0 0 0]", "uniform 1", "uniform %.17g" % p_in,
                      "zeroGradient"))
    with open(os.path.join(case_dir, "0", "U"), "w") as f:
        f.write(field("volVectorField", "U", "[0 1 -1 0 0 0 0]", "uniform (0 0 0)",
                      "uniform (%.17g %.17g 0)" % (ux_in, uy_in), "slip"))
    with open(os.path.join(case_dir, "0", "T"), "w") as f:
        f.write(field("volScalarField", "T", "[0 0 0 1 0 0 0]", "uniform 1", "uniform %.17g" % T_in,
                      "zeroGradient"))
    with open(os.path.join(case_dir, "system", "controlDict"), "w") as f:
        f.write(_foam_header("dictionary", "controlDict"))
        f.write("application     cudaEuler2DmixO3RK3solver;\nstartFrom       startTime;\nstartTime       0;\n"
                "stopAt          endTime;\nendTime         %r;\ndeltaT          1e-05;\n"
                "writeControl    adjustableRunTime;\nwriteInterval   %r;\nadjustTimeStep  yes;\n"
                "maxCo           %r;\n" % (float(end_time), float(write_interval), float(max_co)))
    with open(os.path.join(case_dir, "system", "fvSchemes"), "w") as f:
        f.write(_foam_header("dictionary", "fvSchemes"))
        f.write("gaussQuadrature\n{\n    cOrder %d;\n    fOrder %d;\n}\n\nCUDA\n{\n    R  %r;\n    Cv  %r;\n"
                "    stencilType  edgeNeighbour;\n    recalcBasis  true;\n\n    boundary\n    {\n"
                "        patch0  5;\n        patch1  2;\n        patch2  4;\n    }\n}\n" % (c_order, f_order, R, Cv))
    return mesh


def foam_time_loop(end_time, write_interval, step):
    """The application's time loop (cudaEuler2DmixO3RK3solver.cpp:53-79 with
    Foam::Time's adjustableRunTime writes): step(t) -> CFL time step of the
    current state, clipped to the next write time and to the end time; then
    the time advances and `yield t, dt, write_now` (the caller evolves by dt)."""
    t, nxt = 0.0, write_interval
    tol = 1e-12 * write_interval
    while t < end_time - 1e-12 * (end_time if end_time > 0 else 1.0):
        d = step(t)
        to_write, to_end = nxt - t, end_time - t
        if to_write > 0 and d > to_write:
            d = to_write
        if to_end > 0 and d > to_end:
            d = to_end
        t += d
        write_now = False
        if t >= nxt - tol:
            write_now = True
            while nxt <= t + tol:
                nxt += write_interval
        if t >= end_time - tol:
            write_now = True
        yield t, d, write_now
