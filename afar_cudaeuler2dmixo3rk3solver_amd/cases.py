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


def vortex_case(n=64, seed=0, jitter=0.15, L=10.0, R=R_REF):
    mesh = Mesh.periodic(n=n, L=L, jitter=jitter, seed=seed)
    W, _ = vortex_exact(_centroids(mesh), 0.0, L=L, R=R)
    return mesh, W
