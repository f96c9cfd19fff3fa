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
