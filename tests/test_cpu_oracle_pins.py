"""Pins of the oracle's point functions against what mathematics fixes.

- exp / erfc against mpmath (40 digits): <= 1 ulp and <= 8 ulp; their float
  versions (the FP32-mixed flux) <= 1.5 ulp and <= 5 ulp of float.
- Roe and KFVS flux consistency: F(U, U) equals the physical Euler flux.
- Roe upwinding: for a supersonic face with all wave speeds positive and no
  entropy fix active, Roe(UL, UR) = F(UL) when the true jumps are used.
- KFVS at high Mach number tends to the upwind flux.
- The Roe flux's entropy fix: a first-order solve of a transonic rarefaction
  stays free of an expansion shock.
- The SSP-RK3 update kernel with exact coefficients reproduces the RK3
  amplification polynomial 1 + z + z^2/2 + z^3/6 on y' = lambda y, and the
  reference's truncated literals deviate from it by O(1e-6 z^3).
"""
import ctypes as C

import mpmath as mp
import numpy as np
import pytest

from oracle import pyoracle as O

R_GAS = 0.7142857
CV = 1.7857142857
GAMMA = (CV + R_GAS) / CV


def ulp_err(a, b):
    if a == b:
        return 0.0
    return abs(a - b) / np.spacing(abs(b))


def test_exp_against_mpmath():
    mp.mp.dps = 40
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-700, 700, 2000), rng.uniform(-2, 2, 2000), [0.0, 1.0, -1.0, 709.7, -700.0]])
    worst = max(ulp_err(O.lib().oracle_exp(float(x)), float(mp.exp(mp.mpf(float(x))))) for x in xs)
    assert worst <= 1.0
    assert O.lib().oracle_exp(0.0) == 1.0
    assert O.lib().oracle_exp(800.0) == float("inf")
    assert O.lib().oracle_exp(-800.0) == 0.0


def test_erfc_against_mpmath():
    mp.mp.dps = 40
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(-6, 6, 3000), rng.uniform(0, 26, 1000), [0.0, 0.5, -0.5, 3.75]])
    worst = 0.0
    for x in xs:
        ref = float(mp.erfc(mp.mpf(float(x))))
        if ref > 1e-300:
            worst = max(worst, ulp_err(O.lib().oracle_erfc(float(x)), ref))
    assert worst <= 8.0
    assert ulp_err(O.lib().oracle_erfc(0.0), 1.0) <= 1.0
    assert O.lib().oracle_erfc(30.0) == 0.0
    assert O.lib().oracle_erfc(-30.0) == 2.0


def ulp_err_f32(a, ref):
    """|a - ref| in units of the float spacing at ref (ref exact, in double)."""
    if float(a) == ref:
        return 0.0
    return abs(float(a) - ref) / float(np.spacing(np.abs(np.float32(ref))))


def test_expf_against_mpmath():
    """The float exp of the FP32-mixed flux: <= 1.5 ulp of float (measured
    1.06), overflow, underflow and the exact point."""
    mp.mp.dps = 40
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-87, 88.7, 4000), rng.uniform(-1, 1, 2000)]).astype(np.float32)
    worst = max(ulp_err_f32(O.lib().oracle_expf(float(x)), float(mp.exp(mp.mpf(float(x))))) for x in xs)
    assert worst <= 1.5, worst
    assert O.lib().oracle_expf(0.0) == 1.0
    assert O.lib().oracle_expf(90.0) == float("inf")
    assert O.lib().oracle_expf(-110.0) == 0.0
    # subnormal results: within one subnormal spacing
    for x in (-90.0, -95.5, -100.0, -103.0):
        ref = float(mp.exp(mp.mpf(x)))
        assert abs(O.lib().oracle_expf(x) - ref) <= 2.0 ** -149, x


def test_erfcf_against_mpmath():
    """The float erfc of the FP32-mixed flux: <= 5 ulp of float where the
    result is a normal float (measured 3.9), the limits and the symmetry."""
    mp.mp.dps = 40
    rng = np.random.default_rng(4)
    xs = np.concatenate([rng.uniform(-10, 9.2, 4000), rng.uniform(-1, 1, 1000),
                         rng.normal(0, 3, 1000)]).astype(np.float32)
    worst = 0.0
    for x in xs:
        ref = float(mp.erfc(mp.mpf(float(x))))
        if ref >= 1.2e-38:
            worst = max(worst, ulp_err_f32(O.lib().oracle_erfcf(float(x)), ref))
    assert worst <= 5.0, worst
    assert O.lib().oracle_erfcf(0.0) == 1.0
    assert O.lib().oracle_erfcf(11.0) == 0.0 and O.lib().oracle_erfcf(-11.0) == 2.0
    assert O.lib().oracle_erfcf(1.5) + O.lib().oracle_erfcf(-1.5) == pytest.approx(2.0, abs=2.5e-7)


def test_library_float_math_equals_oracle(afar):
    """The library's host evaluation of the float functions (the code the
    device compiles) equals the oracle's, bit for bit."""
    from afar_cudaeuler2dmixo3rk3solver_amd._lib import check, lib, ptr
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-100, 90, 3000), rng.uniform(-12, 12, 3000)]).astype(np.float32)
    out = np.zeros((len(x), 4), np.float32)
    check(lib().afar_host_math_f32(ptr(x, C.c_float), ptr(out, C.c_float), len(x)))
    for i in range(0, len(x), 7):
        assert out[i, 0] == np.float32(O.lib().oracle_expf(float(x[i])))
        assert out[i, 1] == np.float32(O.lib().oracle_erfcf(float(x[i])))


def physical_flux(rho, p, ux, uy, T, Sfx, Sfy):
    vn = ux * Sfx + uy * Sfy
    E = CV * T + 0.5 * (ux * ux + uy * uy)
    return np.array([rho * vn, rho * ux * vn + p * Sfx, rho * uy * vn + p * Sfy, rho * vn * (E + p / rho)])


STATES = [(1.0, 1.0, 0.3, -0.2), (0.5323, 0.3, 1.206, 0.0), (0.138, 0.029, 1.206, 1.206), (1.5, 1.5, -2.0, 0.7)]


@pytest.mark.parametrize("state", STATES)
@pytest.mark.parametrize("normal", [(1.0, 0.0), (0.6, -0.8), (-0.3, 0.95)])
def test_flux_consistency(state, normal):
    rho, p, ux, uy = state
    T = p / (rho * R_GAS)
    mag = 0.37
    Sfx, Sfy = normal[0] * mag / np.hypot(*normal), normal[1] * mag / np.hypot(*normal)
    F = physical_flux(rho, p, ux, uy, T, Sfx, Sfy)
    cen = O.central_flux(rho, rho, p, p, ux, ux, uy, uy, T, T, CV, Sfx, Sfy)
    assert np.allclose(cen, F, rtol=1e-14, atol=1e-14)
    roe = O.roe_flux(rho, rho, p, p, ux, ux, uy, uy, T, T, 0.0, 0.0, R_GAS, CV, mag, Sfx, Sfy, base=cen)
    assert np.allclose(roe, F, rtol=1e-13, atol=1e-13)
    kf = O.kfvs_flux(rho, rho, p, p, ux, ux, uy, uy, T, T, R_GAS, CV, mag, Sfx, Sfy, base=cen)
    assert np.allclose(kf, F, rtol=1e-12, atol=1e-12)


def test_roe_upwind_supersonic():
    # left state moving right at Mach ~3; right state slower (no entropy-fix band: eps = 0)
    rhoL, pL, uL, vL = 1.0, 1.0, 3.6, 0.1
    rhoR, pR, uR, vR = 1.1, 1.05, 3.4, 0.05
    TL, TR = pL / (rhoL * R_GAS), pR / (rhoR * R_GAS)
    cL, cR = np.sqrt(GAMMA * pL / rhoL), np.sqrt(GAMMA * pR / rhoR)
    assert uR - cR <= uL - cL and uR <= uL and uR + cR <= uL + cL and uR - cR > 0
    # all wave speeds positive and no entropy fix: the Roe dissipation term is
    # -1/2 sum lambda_k alpha_k r_k = -1/2 (F(UR) - F(UL)) (Roe's property A dU = dF)
    for Sfx, Sfy in ((0.25, 0.0), (0.24, 0.07)):
        mag = np.hypot(Sfx, Sfy)
        FL = physical_flux(rhoL, pL, uL, vL, TL, Sfx, Sfy)
        FR = physical_flux(rhoR, pR, uR, vR, TR, Sfx, Sfy)
        diss = O.roe_flux(rhoL, rhoR, pL, pR, uL, uR, vL, vR, TL, TR, pR - pL, rhoR - rhoL, R_GAS, CV, mag, Sfx, Sfy)
        assert np.allclose(diss, -0.5 * (FR - FL), rtol=1e-11, atol=1e-13)
        # and the full flux is exactly upwind for the continuity equation
        cen = O.central_flux(rhoL, rhoR, pL, pR, uL, uR, vL, vR, TL, TR, CV, Sfx, Sfy)
        assert abs((cen + diss)[0] - FL[0]) < 1e-13
        # a plausible slip (a jump taken with the wrong sign) breaks the property
        bad = O.roe_flux(rhoL, rhoR, pL, pR, uL, uR, vL, vR, TL, TR, pL - pR, rhoR - rhoL, R_GAS, CV, mag, Sfx, Sfy)
        assert not np.allclose(bad, -0.5 * (FR - FL), rtol=1e-6)


def test_kfvs_hypersonic_upwind():
    rhoL, pL, uL, vL = 1.0, 0.01, 20.0, 0.0
    rhoR, pR, uR, vR = 1.0, 0.01, 19.9, 0.0
    TL, TR = pL / (rhoL * R_GAS), pR / (rhoR * R_GAS)
    Sfx, Sfy, mag = 1.0, 0.0, 1.0
    cen = O.central_flux(rhoL, rhoR, pL, pR, uL, uR, vL, vR, TL, TR, CV, Sfx, Sfy)
    kf = O.kfvs_flux(rhoL, rhoR, pL, pR, uL, uR, vL, vR, TL, TR, R_GAS, CV, mag, Sfx, Sfy, base=cen)
    assert np.allclose(kf, physical_flux(rhoL, pL, uL, vL, TL, Sfx, Sfy), rtol=1e-10)


def test_roe_entropy_fix_transonic_rarefaction():
    """The entropy fix of the Roe flux (cudaRiemannSolver.h:120-130) on the
    problem that needs it: Toro's test 1 (left 1, 0.75, 1; right 0.125, 0,
    0.1 as rho, u, p) has a left rarefaction through the sonic point u = c.
    A first-order finite-volume solve with the oracle's Roe flux on 200 cells
    to t = 0.2 stays smooth through the fan (largest density jump between
    neighbouring cells 0.017); without the fix on the acoustic waves the
    scheme keeps an entropy-violating expansion shock at the sonic point (a
    jump of 0.2)."""
    N, cfl = 200, 0.4
    x = (np.arange(N) + 0.5) / N - 0.5
    rho = np.where(x < 0, 1.0, 0.125)
    u = np.where(x < 0, 0.75, 0.0)
    p = np.where(x < 0, 1.0, 0.1)
    dx = 1.0 / N
    U = np.stack([rho, rho * u, np.zeros(N), p / (GAMMA - 1) + 0.5 * rho * u * u], 1)
    t = 0.0
    while t < 0.2:
        rho = U[:, 0]
        u = U[:, 1] / rho
        p = (GAMMA - 1) * (U[:, 3] - 0.5 * rho * u * u)
        T = p / (rho * R_GAS)
        with np.errstate(invalid="ignore"):  # a broken flux may drive p negative; the asserts below fail then
            dt = min(cfl * dx / np.max(np.abs(u) + np.sqrt(GAMMA * p / rho)), 0.2 - t)
        F = np.zeros((N + 1, 4))
        for i in range(1, N):
            a, b = i - 1, i
            cen = O.central_flux(rho[a], rho[b], p[a], p[b], u[a], u[b], 0.0, 0.0, T[a], T[b], CV, 1.0, 0.0)
            F[i] = O.roe_flux(rho[a], rho[b], p[a], p[b], u[a], u[b], 0.0, 0.0, T[a], T[b], p[b] - p[a],
                              rho[b] - rho[a], R_GAS, CV, 1.0, 1.0, 0.0, base=cen)
        for i, c in ((0, 0), (N, N - 1)):  # transmissive ends
            F[i] = physical_flux(rho[c], p[c], u[c], 0.0, T[c], 1.0, 0.0)
        U = U - dt / dx * (F[1:] - F[:-1])
        t += dt
    jump = np.abs(np.diff(U[:, 0]))
    fan = (x[:-1] > -0.3) & (x[:-1] < 0.05)  # the rarefaction, sonic point near x = 0
    assert jump[fan].max() < 0.05, jump[fan].max()
    assert np.all(U[:, 0] > 0)


def _one_cell_update(y0, lam, dt, betas):
    """Drive the oracle's update kernel on a one-cell field with Res = -lam*y*V."""
    lib = O.lib()
    f = O.OracleFields()
    f.nc, f.nf, f.nb, f.maxS, f.maxC = 1, 0, 0, 1, 1
    V = np.array([2.0])
    U = np.array([[y0, 0.0, 0.0, 10.0 * y0]])
    Uold = U.copy()
    Res = np.zeros((1, 4))
    FD = np.zeros((1, 4))
    f.cellVolume = O._p(V)
    f.cFD = O._p(U)
    f.cFDold = O._p(Uold)
    f.Res = O._p(Res)
    f.FD = O._p(FD)
    for b1, b2, b3 in zip(*betas):
        Res[:] = -lam * U * V[0]
        lib.oracle_update(C.byref(f), R_GAS, R_GAS / CV + 1.0, dt, b1, b2, b3)
    return U[0, 0]


def test_rk3_amplification_polynomial():
    exact = ((1.0, 0.75, 1.0 / 3.0), (0.0, 0.25, 2.0 / 3.0), (1.0, 0.25, 2.0 / 3.0))
    literal = O.OracleState.RK_LITERAL
    for z in (-0.1, -0.5, -1.0, -2.0, 0.3):
        poly = 1 + z + z * z / 2 + z ** 3 / 6
        assert abs(_one_cell_update(1.0, z, 1.0, exact) - poly) < 1e-14
        # literal coefficients: the same form with 0.333333 / 0.666667
        lit = _one_cell_update(1.0, z, 1.0, literal)
        y1 = 1 + z
        y2 = 0.75 + 0.25 * y1 + 0.25 * z * y1
        assert abs(lit - (0.333333 + 0.666667 * y2 + 0.666667 * z * y2)) < 1e-14
        assert abs(lit - poly) > 1e-9 * abs(z) ** 2 or z == 0
