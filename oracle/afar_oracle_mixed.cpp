// ORACLE -- test infrastructure only.  Nothing in the product links, loads or
// runs this file; the tests use it as the checker of the FP32-mixed device
// path.
//
// The FP32-mixed path (afar_schemes.precision 1) restated on the CPU: the
// reference's reconstruction, limiter, BVD switch and face fluxes
// (cudaLimiter.h:9-464, cudaRiemannSolver.h:7-271, the same loops and
// expressions as afar_oracle.cpp) evaluated in IEEE single precision, with the
// library's mixed storage: primitive states rounded to float where a kernel
// reads them, geometry, least-squares weights, centre distances and RBF
// coefficients formed in double and rounded to float, face fluxes kept as the
// floats they are computed as.  The conserved state, the residual gather, the
// RK update, the boundary ghosts (formed from the double primitive state) and
// the CFL time step (over float states) stay in double; those parts are the
// FP64 oracle's own functions.  KFVS uses the library's float exp and erfc
// (afar_expf, afar_erfcf), as the device's mixed path does.  Built with
// -ffp-contract=off.  With flux points (ngp > 0, the faceQuadrature
// extension) the point states and fluxes are float as well and their weighted
// sum is formed in double, as on the device.
#include <cmath>

#include "afar_oracle.h"

extern "C" {
void oracle_evaluate_residual(oracle_fields* f);
void oracle_update(oracle_fields* f, double R, double gamma, double deltaT, double beta1, double beta2, double beta3);
void oracle_update_boundary(oracle_fields* f, double R, double Cv);
void oracle_store_conserved(oracle_fields* f);
}

namespace {
using namespace oracle_detail;
using std::fabs;
using std::fmax;
using std::fmin;
using std::sqrt;

typedef float R;
const R SMALL = (R)1e-5, VSMALL = (R)1e-10, BVD_FACTOR = (R)0.5, VALID_LIM_FACTOR = (R)0.9;

inline R w_of(const oracle_fields* f, int c, int k) { return (R)f->FD[4 * (size_t)c + k]; }  // float state

R kexp(R x) { return afar_expf(x); }
R kerfc(R x) { return afar_erfcf(x); }

void limiter_process(R& limiter, R& phif, R phic, R phiMax, R phiMin, R gx, R gy, R dx, R dy) {
    phif = phic + (gx * dx + gy * dy);
    const R d2 = phif - phic;
    const R d1 = (R)(phif < phic) * (phiMin - phic) + (R)(phif >= phic) * (phiMax - phic);
    R yy = fmin((R)2, ((R)0.5 * d1) / ((d2 + (R)(d2 >= (R)0) * VSMALL) - (R)(d2 < (R)0) * VSMALL));
    const int valid = fabs(d2) >= SMALL;
    yy = (R)valid * yy + (R)(1 - valid) * (R)1;
    limiter = fmin(limiter, yy);
}

R to_bj(R y) { return fmin((R)2 * y, (R)1); }
R to_venkat(R y) { return fmin((y * y + y + VSMALL) / (y * y + (R)0.5 + (R)0.5 * y + VSMALL), (R)1); }
R transform_limiter(R lim, int isShock) { return (R)isShock * to_venkat(lim) + (R)(1 - isShock) * to_bj(lim); }

R minimize_delta(R L, R Rv, R LT, R RT) {
    R d = fabs(Rv - L);
    d = fmin(d, fabs(Rv - LT));
    d = fmin(d, fabs(RT - L));
    d = fmin(d, fabs(RT - LT));
    return (R)(Rv >= L) * d - (R)(Rv < L) * d;
}

void kfvs_flux(R* F, R rhoL, R rhoR, R pL, R pR, R uxL, R uxR, R uyL, R uyR, R TL, R TR, R Rg, R Cv, R magSf, R Sfx,
               R Sfy) {
    const R nfx = Sfx / magSf, nfy = Sfy / magSf;
    const R uxLl = uxL * nfx + uyL * nfy, uyLl = uyL * nfx - uxL * nfy;
    const R uxRl = uxR * nfx + uyR * nfy, uyRl = uyR * nfx - uxR * nfy;
    const R lamL = (R)0.5 * rhoL / pL, lamR = (R)0.5 * rhoR / pR;
    const R PI = (R)3.141593;
    const R Mu0p = (R)0.5 * kerfc(-sqrt(lamL) * uxLl);
    const R Mu1p = uxLl * Mu0p + (R)0.5 * kexp(-lamL * uxLl * uxLl) / (sqrt(PI * lamL));
    const R Mu2p = uxLl * Mu1p + (R)0.5 * Mu0p / lamL;
    const R Mu3p = uxLl * Mu2p + Mu1p / lamL;
    const R Mu0n = (R)0.5 * kerfc(sqrt(lamR) * uxRl);
    const R Mu1n = uxRl * Mu0n - (R)0.5 * kexp(-lamR * uxRl * uxRl) / (sqrt(PI * lamR));
    const R Mu2n = uxRl * Mu1n + (R)0.5 * Mu0n / lamR;
    const R Mu3n = uxRl * Mu2n + Mu1n / lamR;
    const R Mv1L = uyLl, Mv1R = uyRl;
    const R Kd = (R)2 * Cv / Rg - (R)2;
    const R trho = rhoL * Mu1p + rhoR * Mu1n;
    const R trhoUx = rhoL * Mu2p + rhoR * Mu2n;
    const R trhoUy = rhoL * Mu1p * Mv1L + rhoR * Mu1n * Mv1R;
    const R trhoE = (R)0.5 * rhoL * (Mu3p + Mu1p * (uyLl * Mv1L + (R)0.5 / lamL) + Mu1p * Kd * (R)0.5 / lamL) +
                    (R)0.5 * rhoR * (Mu3n + Mu1n * (uyRl * Mv1R + (R)0.5 / lamR) + Mu1n * Kd * (R)0.5 / lamR);
    F[1] += trhoUx * Sfx - trhoUy * Sfy - (R)0.5 * (F[0] * uxL + pL * Sfx) - (R)0.5 * (F[0] * uxR + pR * Sfx);
    F[2] += trhoUx * Sfy + trhoUy * Sfx - (R)0.5 * (F[0] * uyL + pL * Sfy) - (R)0.5 * (F[0] * uyR + pR * Sfy);
    F[3] += trhoE * magSf - (R)0.5 * F[0] * (Cv * TL + (R)0.5 * (uxL * uxL + uyL * uyL) + pL / rhoL) -
            (R)0.5 * F[0] * (Cv * TR + (R)0.5 * (uxR * uxR + uyR * uyR) + pR / rhoR);
    F[0] += trho * magSf - (R)0.5 * rhoL * (uxL * Sfx + uyL * Sfy) - (R)0.5 * rhoR * (uxR * Sfx + uyR * Sfy);
}

void roe_flux(R* F, R rhoL, R rhoR, R pL, R pR, R uxL, R uxR, R uyL, R uyR, R TL, R TR, R dP, R dRho, R gamma, R Cv,
              R magSf, R Sfx, R Sfy) {
    const R nfx = Sfx / magSf, nfy = Sfy / magSf;
    const R VL = uxL * nfx + uyL * nfy, VR = uxR * nfx + uyR * nfy;
    const R hL = Cv * TL + (R)0.5 * (uxL * uxL + uyL * uyL) + pL / rhoL;
    const R hR = Cv * TR + (R)0.5 * (uxR * uxR + uyR * uyR) + pR / rhoR;
    const R rhoT = sqrt(fmax(rhoL * rhoR, (R)1e-20));
    const R wL = sqrt(fmax(rhoL, (R)1e-20)) / (sqrt(fmax(rhoL, (R)1e-20)) + sqrt(fmax(rhoR, (R)1e-20)));
    const R wR = (R)1 - wL;
    const R uxT = uxL * wL + uxR * wR, uyT = uyL * wL + uyR * wR, hT = hL * wL + hR * wR;
    const R q2 = uxT * uxT + uyT * uyT;
    const R cT = sqrt(fmax((gamma - (R)1) * (hT - (R)0.5 * q2), (R)1e-20));
    const R VT_ = uxT * nfx + uyT * nfy;
    const R dUx = uxR - uxL, dUy = uyR - uyL;
    const R dV = dUx * nfx + dUy * nfy;
    const R r1 = (dP - rhoT * cT * dV) / ((R)2 * cT * cT);
    const R r2 = dRho - dP / (cT * cT);
    const R r3 = (dP + rhoT * cT * dV) / ((R)2 * cT * cT);
    const R cL = sqrt(fmax((gamma - (R)1) * (hL - (R)0.5 * (uxL * uxL + uyL * uyL)), (R)1e-20));
    const R cR = sqrt(fmax((gamma - (R)1) * (hR - (R)0.5 * (uxR * uxR + uyR * uyR)), (R)1e-20));
    R eps = (R)2 * fmax((R)0, (VR - cR) - (VL - cL));
    R l1 = fabs(VT_ - cT);
    l1 = (R)(l1 >= eps) * l1 + (R)(l1 < eps) * ((l1 * l1 + eps * eps) / ((R)2 * eps + (R)1e-30));
    eps = (R)2 * fmax((R)0, VR - VL);
    R l2 = fabs(VT_);
    l2 = (R)(l2 >= eps) * l2 + (R)(l2 < eps) * ((l2 * l2 + eps * eps) / ((R)2 * eps + (R)1e-30));
    eps = (R)2 * fmax((R)0, (VR + cR) - (VL + cL));
    R l3 = fabs(VT_ + cT);
    l3 = (R)(l3 >= eps) * l3 + (R)(l3 < eps) * ((l3 * l3 + eps * eps) / ((R)2 * eps + (R)1e-30));
    F[0] -= (R)0.5 * (l1 * r1 + l2 * r2 + l3 * r3) * magSf;
    F[1] -= (R)0.5 * (l1 * r1 * (uxT - cT * nfx) + l2 * (r2 * uxT + rhoT * (dUx - dV * nfx)) + l3 * r3 * (uxT + cT * nfx)) * magSf;
    F[2] -= (R)0.5 * (l1 * r1 * (uyT - cT * nfy) + l2 * (r2 * uyT + rhoT * (dUy - dV * nfy)) + l3 * r3 * (uyT + cT * nfy)) * magSf;
    F[3] -= (R)0.5 * (l1 * r1 * (hT - cT * VT_) + l2 * (r2 * (R)0.5 * q2 + rhoT * (uxT * dUx + uyT * dUy - VT_ * dV)) +
                      l3 * r3 * (hT + cT * VT_)) * magSf;
}

// Central part of the face flux (cudaRiemannSolver.h:225-229).
void central_flux(R* out, R rhoL, R rhoR, R pL, R pR, R uxL, R uxR, R uyL, R uyR, R TL, R TR, R CvR, R Sfx, R Sfy) {
    out[0] = (R)0.5 * rhoL * (uxL * Sfx + uyL * Sfy) + (R)0.5 * rhoR * (uxR * Sfx + uyR * Sfy);
    out[1] = (R)0.5 * (out[0] * uxL + pL * Sfx) + (R)0.5 * (out[0] * uxR + pR * Sfx);
    out[2] = (R)0.5 * (out[0] * uyL + pL * Sfy) + (R)0.5 * (out[0] * uyR + pR * Sfy);
    out[3] = (R)0.5 * out[0] * (CvR * TL + (R)0.5 * (uxL * uxL + uyL * uyL) + pL / rhoL) +
             (R)0.5 * out[0] * (CvR * TR + (R)0.5 * (uxR * uxR + uyR * uyR) + pR / rhoR);
}

}  // namespace

extern "C" {

// The face-flux parts in float on their own (the FP64 oracle exports the same
// three): gas constants in double, rounded as evaluate_flux rounds them.
void oracle_mixed_central_flux(float* flux4, float rhoL, float rhoR, float pL, float pR, float uxL, float uxR,
                               float uyL, float uyR, float TL, float TR, double Cv, float Sfx, float Sfy) {
    central_flux(flux4, rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, (R)Cv, Sfx, Sfy);
}

void oracle_mixed_roe_flux(float* flux4, float rhoL, float rhoR, float pL, float pR, float uxL, float uxR, float uyL,
                           float uyR, float TL, float TR, float dP, float dRho, double Rgas, double Cv, float magSf,
                           float Sfx, float Sfy) {
    roe_flux(flux4, rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, dP, dRho, (R)((Cv + Rgas) / Cv), (R)Cv, magSf,
             Sfx, Sfy);
}

void oracle_mixed_kfvs_flux(float* flux4, float rhoL, float rhoR, float pL, float pR, float uxL, float uxR,
                            float uyL, float uyR, float TL, float TR, double Rgas, double Cv, float magSf, float Sfx,
                            float Sfy) {
    kfvs_flux(flux4, rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, (R)Rgas, (R)Cv, magSf, Sfx, Sfy);
}

// CFL time step over the float primitive states (cudaFieldOperation.h:124-180).
double oracle_mixed_adjust_time_step(const oracle_fields* f, double Rgas, double Cv, double CFL) {
    double dtmin = 1e5;
#pragma omp parallel for schedule(static) reduction(min : dtmin)
    for (int faceI = 0; faceI < f->nf; ++faceI) {
        const int own = f->neighbour[2 * faceI];
        const int nei = f->neighbour[2 * faceI + 1];
        const double* F = &f->FACE[5 * (size_t)faceI];
        const double xn = f->CELL[2 * nei] + shx(f, faceI), yn = f->CELL[2 * nei + 1] + shy(f, faceI);
        const double deltaR = afar_hypot(xn - f->CELL[2 * own], yn - f->CELL[2 * own + 1]);
        const double fraction = afar_hypot(F[0] - f->CELL[2 * own], F[1] - f->CELL[2 * own + 1]) / deltaR;
        const double Sfx = F[3], Sfy = F[4], magSf = F[2];
        const double uxo = w_of(f, own, 1), uyo = w_of(f, own, 2), To = w_of(f, own, 3);
        const double uxn = w_of(f, nei, 1), uyn = w_of(f, nei, 2), Tn = w_of(f, nei, 3);
        const double dt = CFL * deltaR * magSf /
                          ((1.0 - fraction) * fabs(uxo * Sfx + uyo * Sfy) + fraction * fabs(uxn * Sfx + uyn * Sfy) +
                           sqrt((Rgas / Cv + 1.0) * Rgas * ((1.0 - fraction) * To + fraction * Tn)) * magSf);
        dtmin = dtmin > dt ? dt : dtmin;
    }
    return dtmin;
}

void oracle_mixed_reconstruct(oracle_fields* f) {
    const int maxS = f->maxS, maxC = f->maxC;
#pragma omp parallel for schedule(static)
    for (int celli = 0; celli < f->nc; ++celli) {
        const int S = f->stencilSize[celli], C = f->compactStencilSize[celli];
        const int32_t* st = &f->stencil[(size_t)maxS * celli];
        R wc[4];
        for (int k = 0; k < 4; ++k) wc[k] = w_of(f, celli, k);
        R g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = 0; i < C; ++i) {
            const R M1 = (R)f->matrix[2 * (size_t)maxC * celli + 2 * i];
            const R M2 = (R)f->matrix[2 * (size_t)maxC * celli + 2 * i + 1];
            R dphi = w_of(f, st[i], 0) - wc[0];
            g[GPX] += M1 * dphi;
            g[GPY] += M2 * dphi;
            dphi = w_of(f, st[i], 1) - wc[1];
            g[GUXX] += M1 * dphi;
            g[GUYX] += M2 * dphi;
            dphi = w_of(f, st[i], 2) - wc[2];
            g[GUXY] += M1 * dphi;
            g[GUYY] += M2 * dphi;
            dphi = w_of(f, st[i], 3) - wc[3];
            g[GTX] += M1 * dphi;
            g[GTY] += M2 * dphi;
        }
        for (int k = 0; k < 8; ++k) f->gradFD[8 * (size_t)celli + k] = g[k];
        R pMax = wc[0], pMin = pMax, UxMax = wc[1], UxMin = UxMax, UyMax = wc[2], UyMin = UyMax, TMax = wc[3], TMin = TMax;
        for (int i = 0; i < C; ++i) {
            pMax = fmax(w_of(f, st[i], 0), pMax);
            pMin = fmin(w_of(f, st[i], 0), pMin);
            UxMax = fmax(w_of(f, st[i], 1), UxMax);
            UxMin = fmin(w_of(f, st[i], 1), UxMin);
            UyMax = fmax(w_of(f, st[i], 2), UyMax);
            UyMin = fmin(w_of(f, st[i], 2), UyMin);
            TMax = fmax(w_of(f, st[i], 3), TMax);
            TMin = fmin(w_of(f, st[i], 3), TMin);
        }
        double* lim = &f->limiter[4 * (size_t)celli];
        if (fabs(TMax - TMin) < VSMALL) {
            lim[0] = lim[1] = lim[2] = lim[3] = 1.0;
            for (int fi = 0; fi < C; ++fi) {
                const Slot s = slot(f, celli, fi);
                if (s.isBoundary) continue;
                const R dx = (R)s.dx, dy = (R)s.dy;
                double* o = &f->faceFD[4 * (2 * (size_t)s.face + s.isNei)];
                o[0] = wc[0] + g[GPX] * dx + g[GPY] * dy;
                o[1] = wc[1] + g[GUXX] * dx + g[GUYX] * dy;
                o[2] = wc[2] + g[GUXY] * dx + g[GUYY] * dy;
                o[3] = wc[3] + g[GTX] * dx + g[GTY] * dy;
                for (int q = 0; q < f->ngp; ++q) {  // the same linear state at the flux points
                    const double* d = &f->gpOffsets[4 * ((size_t)q * f->nf + s.face)];
                    const R gx = (R)(s.isNei ? d[2] : d[0]), gy = (R)(s.isNei ? d[3] : d[1]);
                    double* og = &f->faceGP[4 * (((size_t)q * f->nf + s.face) * 2 + s.isNei)];
                    og[0] = wc[0] + g[GPX] * gx + g[GPY] * gy;
                    og[1] = wc[1] + g[GUXX] * gx + g[GUYX] * gy;
                    og[2] = wc[2] + g[GUXY] * gx + g[GUYY] * gy;
                    og[3] = wc[3] + g[GTX] * gx + g[GTY] * gy;
                }
            }
            f->shockIndicator[2 * (size_t)celli] = 0;
            continue;
        }
        R tmpP = 2, tmpUx = 2, tmpUy = 2, tmpT = 2;
        R smoothnessValue = 0;
        int isNotTrouble = 1;
        const R rP = fabs(pMax - pMin), rUx = fabs(UxMax - UxMin), rUy = fabs(UyMax - UyMin), rT = fabs(TMax - TMin);
        for (int fi = 0; fi < C; ++fi) {
            const Slot s = slot(f, celli, fi);
            if (s.isBoundary) continue;
            const R dx = (R)s.dx, dy = (R)s.dy;
            const R deltaR = (R)afar_hypot(s.ex, s.ey);
            const R fraction = sqrt(dx * dx + dy * dy) / deltaR;
            const double* a = &f->RBFbasis[2 * (size_t)(maxS + 1) * s.face + (size_t)s.isNei * (maxS + 1)];
            R coef = (R)a[0];
            R pf = coef * wc[0], Uxf = coef * wc[1], Uyf = coef * wc[2], Tf = coef * wc[3];
            for (int ci = 0; ci < S; ++ci) {
                coef = (R)a[ci + 1];
                pf += coef * w_of(f, st[ci], 0);
                Uxf += coef * w_of(f, st[ci], 1);
                Uyf += coef * w_of(f, st[ci], 2);
                Tf += coef * w_of(f, st[ci], 3);
            }
            R phif;
            double* o = &f->faceFD[4 * (2 * (size_t)s.face + s.isNei)];
            limiter_process(tmpP, phif, wc[0], pMax, pMin, g[GPX], g[GPY], dx, dy);
            o[0] = (R)(rP >= SMALL) * pf + (R)(rP < SMALL) * phif;
            smoothnessValue = fmax(smoothnessValue, fabs(((R)1 - fraction) * wc[0] + fraction * w_of(f, st[fi], 0) - phif) / deltaR +
                                                        (R)(pf <= (R)0) * (R)1e20 + (R)(phif <= (R)0) * (R)1e20);
            limiter_process(tmpUx, phif, wc[1], UxMax, UxMin, g[GUXX], g[GUYX], dx, dy);
            o[1] = (R)(rUx >= SMALL) * Uxf + (R)(rUx < SMALL) * phif;
            limiter_process(tmpUy, phif, wc[2], UyMax, UyMin, g[GUXY], g[GUYY], dx, dy);
            o[2] = (R)(rUy >= SMALL) * Uyf + (R)(rUy < SMALL) * phif;
            limiter_process(tmpT, phif, wc[3], TMax, TMin, g[GTX], g[GTY], dx, dy);
            o[3] = (R)(rT >= SMALL) * Tf + (R)(rT < SMALL) * phif;
            isNotTrouble = phif > (R)0;
            // flux points: RBF sums with each point's coefficients, or the
            // unlimited linear state where the compact range is flat
            for (int q = 0; q < f->ngp; ++q) {
                const double* ag = &f->RBFgp[(((size_t)q * f->nf + s.face) * 2 + s.isNei) * (size_t)(maxS + 1)];
                R cg = (R)ag[0];
                R pg = cg * wc[0], Uxg = cg * wc[1], Uyg = cg * wc[2], Tg = cg * wc[3];
                for (int ci = 0; ci < S; ++ci) {
                    cg = (R)ag[ci + 1];
                    pg += cg * w_of(f, st[ci], 0);
                    Uxg += cg * w_of(f, st[ci], 1);
                    Uyg += cg * w_of(f, st[ci], 2);
                    Tg += cg * w_of(f, st[ci], 3);
                }
                const double* d = &f->gpOffsets[4 * ((size_t)q * f->nf + s.face)];
                const R gx = (R)(s.isNei ? d[2] : d[0]), gy = (R)(s.isNei ? d[3] : d[1]);
                double* og = &f->faceGP[4 * (((size_t)q * f->nf + s.face) * 2 + s.isNei)];
                og[0] = (R)(rP >= SMALL) * pg + (R)(rP < SMALL) * (wc[0] + (g[GPX] * gx + g[GPY] * gy));
                og[1] = (R)(rUx >= SMALL) * Uxg + (R)(rUx < SMALL) * (wc[1] + (g[GUXX] * gx + g[GUYX] * gy));
                og[2] = (R)(rUy >= SMALL) * Uyg + (R)(rUy < SMALL) * (wc[2] + (g[GUXY] * gx + g[GUYY] * gy));
                og[3] = (R)(rT >= SMALL) * Tg + (R)(rT < SMALL) * (wc[3] + (g[GTX] * gx + g[GTY] * gy));
            }
        }
        const R nt = (R)isNotTrouble;
        lim[0] = fmax(tmpP, (R)0) * nt;
        lim[1] = fmax(tmpUx, (R)0) * nt;
        lim[2] = fmax(tmpUy, (R)0) * nt;
        lim[3] = fmax(tmpT, (R)0) * nt;
        smoothnessValue = smoothnessValue / ((R)0.5 * (pMax + pMin) * nt + VSMALL);
        f->shockIndicator[2 * (size_t)celli] = smoothnessValue > (R)1;
    }
}

void oracle_mixed_bvd(oracle_fields* f) {
    const int maxS = f->maxS;
#pragma omp parallel for schedule(static)
    for (int celli = 0; celli < f->nc; ++celli) {
        int BVDtype = 0, bfacesNum = 0;
        const int C = f->compactStencilSize[celli], S = f->stencilSize[celli];
        const int32_t* st = &f->stencil[(size_t)maxS * celli];
        int8_t shockType = f->shockIndicator[2 * (size_t)celli];
        int8_t shockType_extend = shockType;
        for (int ci = 0; ci < C; ++ci) {
            shockType += f->shockIndicator[2 * (size_t)st[ci]];
            shockType_extend = shockType;
        }
        for (int ci = C; ci < S; ++ci) shockType_extend += f->shockIndicator[2 * (size_t)st[ci]];
        const R Tc = w_of(f, celli, 3);
        const R gTx = (R)f->gradFD[8 * (size_t)celli + GTX], gTy = (R)f->gradFD[8 * (size_t)celli + GTY];
        double* lim = &f->limiter[4 * (size_t)celli];
        for (int fi = 0; fi < C; ++fi) {
            const Slot s = slot(f, celli, fi);
            if (s.isBoundary) {
                bfacesNum += 1;
                if (bfacesNum >= 2) {
                    BVDtype = 1;
                    break;
                }
                continue;
            }
            const R TfOwn = (R)f->faceFD[4 * (2 * (size_t)s.face) + 3];
            const R TfNei = (R)f->faceFD[4 * (2 * (size_t)s.face + 1) + 3];
            const R Tf_cur = (R)s.isNei * TfNei + (R)(1 - s.isNei) * TfOwn;
            const R Tf_nei = (R)(1 - s.isNei) * TfNei + (R)s.isNei * TfOwn;
            const R l = to_bj((R)lim[3]);
            const R Tf_cur_TVD = Tc + l * (gTx * (R)s.dx + gTy * (R)s.dy);
            const int typeval = fabs(Tf_cur_TVD - Tf_nei) < BVD_FACTOR * fabs(Tf_cur - Tf_nei);
            const int isNotFlat = fabs(Tf_cur - Tf_nei) > VSMALL;
            const int isLimited = (l < VALID_LIM_FACTOR) + (shockType_extend > 0);
            BVDtype = BVDtype > isNotFlat * isLimited * typeval ? BVDtype : isNotFlat * isLimited * typeval;
        }
        int finalType = (shockType > 0) * (shockType_extend > 1) * (-2) + (BVDtype > 0);
        if (f->boundaryFirstOrder && bfacesNum > 0) {
            finalType = finalType ? finalType : 1;
            lim[0] = lim[1] = lim[2] = lim[3] = 0.0;
            f->shockIndicator[2 * (size_t)celli + 1] = (int8_t)finalType;
            continue;
        }
        f->shockIndicator[2 * (size_t)celli + 1] = (int8_t)finalType;
        const int isShock = finalType < 0;
        for (int k = 0; k < 4; ++k) lim[k] = transform_limiter((R)lim[k], isShock);
    }
}

// The float flux of face faceI from the states reconstructed at a point with
// offsets (dxL, dyL) from the owner and (dxR, dyR) from the neighbour; fL, fR
// are the high-order states there (cudaRiemannSolver.h:155-251).
static void point_flux_f(const oracle_fields* f, int faceI, R Rg, R CvR, R gamma, R dxL, R dyL, R dxR, R dyR,
                         const R* fL, const R* fR, R* out) {
    const int own = f->neighbour[2 * faceI];
    const int nei = f->neighbour[2 * faceI + 1];
    const double* F = &f->FACE[5 * (size_t)faceI];
    const int8_t to = f->shockIndicator[2 * (size_t)own + 1], tn = f->shockIndicator[2 * (size_t)nei + 1];
    const int isTVD_L = to != 0, isTVD_R = tn != 0, isShock = to < 0 || tn < 0;
    const R magSf = (R)F[2], Sfx = (R)F[3], Sfy = (R)F[4];
    R wL[4], wR[4], lL[4], lR[4], gL[8], gR[8];
    for (int k = 0; k < 4; ++k) {
        wL[k] = w_of(f, own, k);
        wR[k] = w_of(f, nei, k);
        lL[k] = (R)f->limiter[4 * (size_t)own + k];
        lR[k] = (R)f->limiter[4 * (size_t)nei + k];
    }
    for (int k = 0; k < 8; ++k) {
        gL[k] = (R)f->gradFD[8 * (size_t)own + k];
        gR[k] = (R)f->gradFD[8 * (size_t)nei + k];
    }
    const R pLT = wL[0] + lL[0] * (gL[GPX] * dxL + gL[GPY] * dyL);
    const R uxLT = wL[1] + lL[1] * (gL[GUXX] * dxL + gL[GUYX] * dyL);
    const R uyLT = wL[2] + lL[2] * (gL[GUXY] * dxL + gL[GUYY] * dyL);
    const R TLT = wL[3] + lL[3] * (gL[GTX] * dxL + gL[GTY] * dyL);
    const R rhoLT = pLT / (Rg * TLT);
    const R pRT = wR[0] + lR[0] * (gR[GPX] * dxR + gR[GPY] * dyR);
    const R uxRT = wR[1] + lR[1] * (gR[GUXX] * dxR + gR[GUYX] * dyR);
    const R uyRT = wR[2] + lR[2] * (gR[GUXY] * dxR + gR[GUYY] * dyR);
    const R TRT = wR[3] + lR[3] * (gR[GTX] * dxR + gR[GTY] * dyR);
    const R rhoRT = pRT / (Rg * TRT);
    const R pL = (R)(1 - isTVD_L) * fL[0] + (R)isTVD_L * pLT;
    const R uxL = (R)(1 - isTVD_L) * fL[1] + (R)isTVD_L * uxLT;
    const R uyL = (R)(1 - isTVD_L) * fL[2] + (R)isTVD_L * uyLT;
    const R TL = (R)(1 - isTVD_L) * fL[3] + (R)isTVD_L * TLT;
    const R rhoL = pL / (Rg * TL);
    const R pR = (R)(1 - isTVD_R) * fR[0] + (R)isTVD_R * pRT;
    const R uxR = (R)(1 - isTVD_R) * fR[1] + (R)isTVD_R * uxRT;
    const R uyR = (R)(1 - isTVD_R) * fR[2] + (R)isTVD_R * uyRT;
    const R TR = (R)(1 - isTVD_R) * fR[3] + (R)isTVD_R * TRT;
    const R rhoR = pR / (Rg * TR);
    const R dP = minimize_delta(pL, pR, pLT, pRT);
    const R dRho = minimize_delta(rhoL, rhoR, rhoLT, rhoRT);
    central_flux(out, rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, CvR, Sfx, Sfy);
    if (isShock)
        kfvs_flux(out, rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, Rg, CvR, magSf, Sfx, Sfy);
    else
        roe_flux(out, rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, dP, dRho, gamma, CvR, magSf, Sfx, Sfy);
}

// Face fluxes in float (face centre: stored exactly in Flux; flux points:
// their weighted sum in double, in point order), boundary faces in double
// from the float ghost states.
void oracle_mixed_evaluate_flux(oracle_fields* f, double Rgas, double Cv) {
    const R Rg = (R)Rgas, CvR = (R)Cv, gamma = (R)((Cv + Rgas) / Cv);
#pragma omp parallel for schedule(static)
    for (int faceI = 0; faceI < f->nf; ++faceI) {
        double* out = &f->Flux[4 * (size_t)faceI];
        if (f->ngp == 0) {
            const int own = f->neighbour[2 * faceI];
            const int nei = f->neighbour[2 * faceI + 1];
            const double* F = &f->FACE[5 * (size_t)faceI];
            const double xn = f->CELL[2 * nei] + shx(f, faceI), yn = f->CELL[2 * nei + 1] + shy(f, faceI);
            const R dxL = (R)(F[0] - f->CELL[2 * own]), dyL = (R)(F[1] - f->CELL[2 * own + 1]);
            const R dxR = (R)(F[0] - xn), dyR = (R)(F[1] - yn);
            R fL[4], fR[4], fq[4];
            for (int k = 0; k < 4; ++k) {
                fL[k] = (R)f->faceFD[4 * (2 * (size_t)faceI) + k];
                fR[k] = (R)f->faceFD[4 * (2 * (size_t)faceI + 1) + k];
            }
            point_flux_f(f, faceI, Rg, CvR, gamma, dxL, dyL, dxR, dyR, fL, fR, fq);
            for (int k = 0; k < 4; ++k) out[k] = fq[k];
            continue;
        }
        for (int q = 0; q < f->ngp; ++q) {  // extension: weighted sum over the face's quadrature points
            const double* d = &f->gpOffsets[4 * ((size_t)q * f->nf + faceI)];
            R fL[4], fR[4], fq[4];
            for (int k = 0; k < 4; ++k) {
                fL[k] = (R)f->faceGP[4 * (((size_t)q * f->nf + faceI) * 2) + k];
                fR[k] = (R)f->faceGP[4 * (((size_t)q * f->nf + faceI) * 2 + 1) + k];
            }
            point_flux_f(f, faceI, Rg, CvR, gamma, (R)d[0], (R)d[1], (R)d[2], (R)d[3], fL, fR, fq);
            const double w = f->gpWeights[q];
            for (int k = 0; k < 4; ++k) out[k] = q == 0 ? w * (double)fq[k] : out[k] + w * (double)fq[k];
        }
    }
    for (int b = 0; b < f->nb; ++b) {
        const double* F = &f->FACE[5 * ((size_t)f->nf + b)];
        const double Sfx = F[3], Sfy = F[4];
        const int c = f->nc + b;
        const double bp = w_of(f, c, 0), bUx = w_of(f, c, 1), bUy = w_of(f, c, 2), bT = w_of(f, c, 3);
        const double rho = bp / (Rgas * bT);
        const double rhoFlux = rho * (bUx * Sfx + bUy * Sfy);
        double* out = &f->Flux[4 * ((size_t)f->nf + b)];
        out[0] = rhoFlux;
        out[1] = rhoFlux * bUx + bp * Sfx;
        out[2] = rhoFlux * bUy + bp * Sfy;
        out[3] = rhoFlux * (Cv * bT + 0.5 * (bUx * bUx + bUy * bUy) + bp / rho);
    }
}

float oracle_expf(float x) { return afar_expf(x); }
float oracle_erfcf(float x) { return afar_erfcf(x); }

void oracle_mixed_evolve(oracle_fields* f, double Rgas, double Cv, double deltaT, int exactRK) {
    const double gamma = Rgas / Cv + 1.0;
    const double lit1[3] = {1.0, 0.75, 0.333333}, lit2[3] = {0.0, 0.25, 0.666667}, lit3[3] = {1.0, 0.25, 0.666667};
    const double ex1[3] = {1.0, 0.75, 1.0 / 3.0}, ex2[3] = {0.0, 0.25, 2.0 / 3.0}, ex3[3] = {1.0, 0.25, 2.0 / 3.0};
    for (int i = 0; i < 3; ++i) {
        oracle_mixed_reconstruct(f);
        oracle_mixed_bvd(f);
        oracle_mixed_evaluate_flux(f, Rgas, Cv);
        oracle_evaluate_residual(f);
        oracle_update(f, Rgas, gamma, deltaT, exactRK ? ex1[i] : lit1[i], exactRK ? ex2[i] : lit2[i],
                      exactRK ? ex3[i] : lit3[i]);
        oracle_update_boundary(f, Rgas, Cv);
    }
    oracle_store_conserved(f);
}

}  // extern "C"
