// ORACLE -- test infrastructure only.  Nothing in the product links, loads or
// runs this file; tests/, __graft_entry__.smoke() and bench.py's CPU baseline
// use it as the checker.
//
// FP64 loop-for-loop restatement of the reference's per-step kernels on the
// reference's host layout (fieldPointer, AoS), one loop iteration per CUDA
// thread, with every `float` of the reference promoted to `double` and the
// literal thresholds kept (SMALL 1e-5, VSMALL 1e-10, BVD_FACTOR 0.5,
// VALID_LIM_FACTOR 0.9, pi = 3.141593).  Reference files (under
// src/CUDA/cudaEuler2DmixO3RK3/ unless noted):
//   oracle_least_square_matrix  cudaFieldOperation.h:51-121   (K3)
//   oracle_adjust_time_step     cudaFieldOperation.h:124-180 + GPUevolve.cu:177-204 (K4)
//   oracle_reconstruct          cudaLimiter.h:35-50, 53-316    (K5)
//   oracle_bvd                  cudaLimiter.h:9-32, 319-464    (K6)
//   oracle_evaluate_flux        cudaRiemannSolver.h:7-271      (K7; all nb boundary faces)
//   oracle_evaluate_residual    cudaFieldOperation.h:183-207   (K8)
//   oracle_update               cudaFieldOperation.h:210-250   (K9)
//   oracle_update_boundary      src/CUDA/BCs/2D/cudaEulerBCs.h:42-73 and BC headers (K10)
//   oracle_store_conserved      cudaFieldOperation.h:253-268   (K11)
//   oracle_evolve               GPUevolve.cu:207-306 (stage order, RK3 literals :223-225)
// exp/erfc come from the library's shared header (own Cody-Waite exp and
// Schonfelder erfc, pinned against libm by the tests); hypot is sqrt(x*x+y*y).
// Periodic meshes: the neighbour centre seen from the owner is CELL[nei] +
// FSHIFT[f] (zero for ordinary meshes, where every expression reduces to the
// reference's).  Built with -ffp-contract=off (see Makefile).
// Library extensions restated here as the checker of the device path:
// boundaryFirstOrder (oracle_bvd) and faceQuadrature (ngp > 0: face states
// and fluxes at the faces' quadrature points, oracle_reconstruct and
// oracle_evaluate_flux); with ngp = 0 every function is the reference's.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../afar_cudaeuler2dmixo3rk3solver_amd/csrc/afar_math.h"

#include "afar_oracle.h"

namespace {
using namespace oracle_detail;

const double SMALL = 1e-5, VSMALL = 1e-10, BVD_FACTOR = 0.5, VALID_LIM_FACTOR = 0.9;

void limiter_process(double& limiter, double& phif, double& d1, double& d2, double& yy, double phic, double phiMax,
                     double phiMin, double gradPhix, double gradPhiy, double dx, double dy) {
    phif = phic;
    phif += gradPhix * dx + gradPhiy * dy;
    d2 = phif - phic;
    d1 = (phif < phic) * (phiMin - phic) + (phif >= phic) * (phiMax - phic);
    yy = fmin(2.0, 0.5 * d1 / (d2 + (d2 >= 0.0) * VSMALL - (d2 < 0.0) * VSMALL));
    const bool valid = (fabs(d2) >= SMALL);
    yy = valid * yy + (1 - valid) * 1.0;
    limiter = fmin(limiter, yy);
}

double to_bj(double yy) { return fmin(2.0 * yy, 1.0); }
double to_venkat(double yy) { return fmin((yy * yy + yy + VSMALL) / (yy * yy + 0.5 + 0.5 * yy + VSMALL), 1.0); }
double transform_limiter(double lim, bool isShock) { return isShock * (to_venkat(lim)) + (1 - isShock) * (to_bj(lim)); }

void minimize_delta(double& deltaPhi, double phiLeft, double phiRight, double phiLeftTVD, double phiRightTVD) {
    deltaPhi = fabs(phiRight - phiLeft);
    deltaPhi = fmin(deltaPhi, fabs(phiRight - phiLeftTVD));
    deltaPhi = fmin(deltaPhi, fabs(phiRightTVD - phiLeft));
    deltaPhi = fmin(deltaPhi, fabs(phiRightTVD - phiLeftTVD));
    deltaPhi = (phiRight >= phiLeft) * deltaPhi - (phiRight < phiLeft) * deltaPhi;
}

}  // namespace

extern "C" {

void oracle_kfvs_flux(double* flux4, double rhoLeft, double rhoRight, double pLeft, double pRight, double UxLeft,
                      double UxRight, double UyLeft, double UyRight, double TLeft, double TRight, double R, double Cv,
                      double magSf, double Sfx, double Sfy) {
    double& rhoFlux = flux4[0];
    double& rhoUxFlux = flux4[1];
    double& rhoUyFlux = flux4[2];
    double& rhoEFlux = flux4[3];
    const double nfx = Sfx / magSf;
    const double nfy = Sfy / magSf;
    const double UxLeftLocal = UxLeft * nfx + UyLeft * nfy;
    const double UyLeftLocal = UyLeft * nfx - UxLeft * nfy;
    const double UxRightLocal = UxRight * nfx + UyRight * nfy;
    const double UyRightLocal = UyRight * nfx - UxRight * nfy;
    const double lamLeft = 0.5 * rhoLeft / pLeft;
    const double lamRight = 0.5 * rhoRight / pRight;
    const double M_u0_pos = 0.5 * afar_erfc(-sqrt(lamLeft) * UxLeftLocal);
    const double M_u1_pos = UxLeftLocal * M_u0_pos + 0.5 * afar_exp(-lamLeft * UxLeftLocal * UxLeftLocal) / (sqrt(3.141593 * lamLeft));
    const double M_u2_pos = UxLeftLocal * M_u1_pos + 0.5 * M_u0_pos / lamLeft;
    const double M_u3_pos = UxLeftLocal * M_u2_pos + M_u1_pos / lamLeft;
    const double M_u0_neg = 0.5 * afar_erfc(sqrt(lamRight) * UxRightLocal);
    const double M_u1_neg = UxRightLocal * M_u0_neg - 0.5 * afar_exp(-lamRight * UxRightLocal * UxRightLocal) / (sqrt(3.141593 * lamRight));
    const double M_u2_neg = UxRightLocal * M_u1_neg + 0.5 * M_u0_neg / lamRight;
    const double M_u3_neg = UxRightLocal * M_u2_neg + M_u1_neg / lamRight;
    const double M_v1_L = UyLeftLocal;
    const double M_v1_R = UyRightLocal;
    const double K = 2.0 * Cv / R - 2.0;
    const double tmprhoFlux = rhoLeft * M_u1_pos + rhoRight * M_u1_neg;
    const double tmprhoUxFlux = rhoLeft * M_u2_pos + rhoRight * M_u2_neg;
    const double tmprhoUyFlux = rhoLeft * M_u1_pos * M_v1_L + rhoRight * M_u1_neg * M_v1_R;
    const double tmprhoEFlux =
        0.5 * rhoLeft * (M_u3_pos + M_u1_pos * (UyLeftLocal * M_v1_L + 0.5 / lamLeft) + M_u1_pos * K * 0.5 / lamLeft) +
        0.5 * rhoRight * (M_u3_neg + M_u1_neg * (UyRightLocal * M_v1_R + 0.5 / lamRight) + M_u1_neg * K * 0.5 / lamRight);
    rhoUxFlux += tmprhoUxFlux * Sfx - tmprhoUyFlux * Sfy - 0.5 * (rhoFlux * UxLeft + pLeft * Sfx) -
                 0.5 * (rhoFlux * UxRight + pRight * Sfx);
    rhoUyFlux += tmprhoUxFlux * Sfy + tmprhoUyFlux * Sfx - 0.5 * (rhoFlux * UyLeft + pLeft * Sfy) -
                 0.5 * (rhoFlux * UyRight + pRight * Sfy);
    rhoEFlux += tmprhoEFlux * magSf - 0.5 * rhoFlux * (Cv * TLeft + 0.5 * (UxLeft * UxLeft + UyLeft * UyLeft) + pLeft / rhoLeft) -
                0.5 * rhoFlux * (Cv * TRight + 0.5 * (UxRight * UxRight + UyRight * UyRight) + pRight / rhoRight);
    rhoFlux += tmprhoFlux * magSf - 0.5 * rhoLeft * (UxLeft * Sfx + UyLeft * Sfy) - 0.5 * rhoRight * (UxRight * Sfx + UyRight * Sfy);
}

void oracle_roe_flux(double* flux4, double rhoLeft, double rhoRight, double pLeft, double pRight, double UxLeft,
                     double UxRight, double UyLeft, double UyRight, double TLeft, double TRight, double deltaP,
                     double deltaRho, double R, double Cv, double magSf, double Sfx, double Sfy) {
    double& rhoFlux = flux4[0];
    double& rhoUxFlux = flux4[1];
    double& rhoUyFlux = flux4[2];
    double& rhoEFlux = flux4[3];
    const double gamma = (Cv + R) / Cv;
    const double nfx = Sfx / magSf;
    const double nfy = Sfy / magSf;
    const double contrVLeft = UxLeft * nfx + UyLeft * nfy;
    const double contrVRight = UxRight * nfx + UyRight * nfy;
    const double hLeft = Cv * TLeft + 0.5 * (UxLeft * UxLeft + UyLeft * UyLeft) + pLeft / rhoLeft;
    const double hRight = Cv * TRight + 0.5 * (UxRight * UxRight + UyRight * UyRight) + pRight / rhoRight;
    const double rhoTilde = sqrt(fmax(rhoLeft * rhoRight, 1e-20));
    const double wLeft = sqrt(fmax(rhoLeft, 1e-20)) / (sqrt(fmax(rhoLeft, 1e-20)) + sqrt(fmax(rhoRight, 1e-20)));
    const double wRight = 1.0 - wLeft;
    const double UxTilde = UxLeft * wLeft + UxRight * wRight;
    const double UyTilde = UyLeft * wLeft + UyRight * wRight;
    const double hTilde = hLeft * wLeft + hRight * wRight;
    const double qTildeSquare = UxTilde * UxTilde + UyTilde * UyTilde;
    const double cTilde = sqrt(fmax((gamma - 1.0) * (hTilde - 0.5 * qTildeSquare), 1e-20));
    const double contrVTilde = UxTilde * nfx + UyTilde * nfy;
    const double deltaUx = UxRight - UxLeft;
    const double deltaUy = UyRight - UyLeft;
    const double deltaContrV = deltaUx * nfx + deltaUy * nfy;
    const double r1 = (deltaP - rhoTilde * cTilde * deltaContrV) / (2.0 * cTilde * cTilde);
    const double r2 = deltaRho - deltaP / (cTilde * cTilde);
    const double r3 = (deltaP + rhoTilde * cTilde * deltaContrV) / (2.0 * cTilde * cTilde);
    const double UL = contrVLeft;
    const double UR = contrVRight;
    const double cLeft = sqrt(fmax((gamma - 1.0) * (hLeft - 0.5 * (UxLeft * UxLeft + UyLeft * UyLeft)), 1e-20));
    const double cRight = sqrt(fmax((gamma - 1.0) * (hRight - 0.5 * (UxRight * UxRight + UyRight * UyRight)), 1e-20));
    double eps = 2.0 * fmax(0.0,(UR - cRight) - (UL - cLeft));
    double lambda1 = fabs(contrVTilde - cTilde);
    lambda1 = (lambda1 >= eps) * lambda1 + (lambda1 < eps) * ((lambda1 * lambda1 + eps * eps) / (2.0 * eps + 1e-30));
    eps = 2.0 * fmax(0.0,UR - UL);
    double lambda2 = fabs(contrVTilde);
    lambda2 = (lambda2 >= eps) * lambda2 + (lambda2 < eps) * ((lambda2 * lambda2 + eps * eps) / (2.0 * eps + 1e-30));
    eps = 2.0 * fmax(0.0,(UR + cRight) - (UL + cLeft));
    double lambda3 = fabs(contrVTilde + cTilde);
    lambda3 = (lambda3 >= eps) * lambda3 + (lambda3 < eps) * ((lambda3 * lambda3 + eps * eps) / (2.0 * eps + 1e-30));
    rhoFlux -= 0.5 * (lambda1 * r1 + lambda2 * r2 + lambda3 * r3) * magSf;
    rhoUxFlux -= 0.5 * (lambda1 * r1 * (UxTilde - cTilde * nfx) + lambda2 * (r2 * UxTilde + rhoTilde * (deltaUx - deltaContrV * nfx)) +
                        lambda3 * r3 * (UxTilde + cTilde * nfx)) * magSf;
    rhoUyFlux -= 0.5 * (lambda1 * r1 * (UyTilde - cTilde * nfy) + lambda2 * (r2 * UyTilde + rhoTilde * (deltaUy - deltaContrV * nfy)) +
                        lambda3 * r3 * (UyTilde + cTilde * nfy)) * magSf;
    rhoEFlux -= 0.5 * (lambda1 * r1 * (hTilde - cTilde * contrVTilde) +
                       lambda2 * (r2 * 0.5 * qTildeSquare + rhoTilde * (UxTilde * deltaUx + UyTilde * deltaUy - contrVTilde * deltaContrV)) +
                       lambda3 * r3 * (hTilde + cTilde * contrVTilde)) * magSf;
}

// Central part of the face flux (cudaRiemannSolver.h:225-229).
void oracle_central_flux(double* flux4, double rhoLeft, double rhoRight, double pLeft, double pRight, double UxLeft,
                         double UxRight, double UyLeft, double UyRight, double TLeft, double TRight, double Cv,
                         double Sfx, double Sfy) {
    const double rhoFlux = 0.5 * rhoLeft * (UxLeft * Sfx + UyLeft * Sfy) + 0.5 * rhoRight * (UxRight * Sfx + UyRight * Sfy);
    flux4[0] = rhoFlux;
    flux4[1] = 0.5 * (rhoFlux * UxLeft + pLeft * Sfx) + 0.5 * (rhoFlux * UxRight + pRight * Sfx);
    flux4[2] = 0.5 * (rhoFlux * UyLeft + pLeft * Sfy) + 0.5 * (rhoFlux * UyRight + pRight * Sfy);
    flux4[3] = 0.5 * rhoFlux * (Cv * TLeft + 0.5 * (UxLeft * UxLeft + UyLeft * UyLeft) + pLeft / rhoLeft) +
               0.5 * rhoFlux * (Cv * TRight + 0.5 * (UxRight * UxRight + UyRight * UyRight) + pRight / rhoRight);
}

void oracle_least_square_matrix(oracle_fields* f) {
#pragma omp parallel for schedule(static)
    for (int celli = 0; celli < f->nc; ++celli) {
        const int m = f->compactStencilSize[celli];
        double Sxx = 0.0, Sxy = 0.0, Syy = 0.0;
        for (int i = 0; i < m; ++i) {
            const Slot s = slot(f, celli, i);
            Sxx += s.ex * s.ex;
            Sxy += s.ex * s.ey;
            Syy += s.ey * s.ey;
        }
        for (int i = 0; i < m; ++i) {
            const Slot s = slot(f, celli, i);
            f->matrix[2 * (size_t)f->maxC * celli + 2 * i] = (Syy * s.ex - Sxy * s.ey) / (Sxx * Syy - Sxy * Sxy);
            f->matrix[2 * (size_t)f->maxC * celli + 2 * i + 1] = (-Sxy * s.ex + Sxx * s.ey) / (Sxx * Syy - Sxy * Sxy);
        }
    }
}

double oracle_adjust_time_step(const oracle_fields* f, double R, double Cv, double CFL) {
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
        const double* Wo = &f->FD[4 * (size_t)own];
        const double* Wn = &f->FD[4 * (size_t)nei];
        const double dt = CFL * deltaR * magSf /
                          ((1.0 - fraction) * fabs(Wo[1] * Sfx + Wo[2] * Sfy) + fraction * fabs(Wn[1] * Sfx + Wn[2] * Sfy) +
                           sqrt((R / Cv + 1.0) * R * ((1.0 - fraction) * Wo[3] + fraction * Wn[3])) * magSf);
        dtmin = dtmin > dt ? dt : dtmin;
    }
    return dtmin;
}

void oracle_reconstruct(oracle_fields* f) {
    const int maxS = f->maxS, maxC = f->maxC;
#pragma omp parallel for schedule(static)
    for (int celli = 0; celli < f->nc; ++celli) {
        const double* Wc = &f->FD[4 * (size_t)celli];
        const int S = f->stencilSize[celli], C = f->compactStencilSize[celli];
        const int32_t* st = &f->stencil[(size_t)maxS * celli];
        double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = 0; i < C; ++i) {
            const double M1 = f->matrix[2 * (size_t)maxC * celli + 2 * i];
            const double M2 = f->matrix[2 * (size_t)maxC * celli + 2 * i + 1];
            const double* Wj = &f->FD[4 * (size_t)st[i]];
            double dphi = Wj[0] - Wc[0];
            g[GPX] += M1 * dphi;
            g[GPY] += M2 * dphi;
            dphi = Wj[1] - Wc[1];
            g[GUXX] += M1 * dphi;
            g[GUYX] += M2 * dphi;
            dphi = Wj[2] - Wc[2];
            g[GUXY] += M1 * dphi;
            g[GUYY] += M2 * dphi;
            dphi = Wj[3] - Wc[3];
            g[GTX] += M1 * dphi;
            g[GTY] += M2 * dphi;
        }
        memcpy(&f->gradFD[8 * (size_t)celli], g, sizeof g);
        double pMax = Wc[0], pMin = pMax, UxMax = Wc[1], UxMin = UxMax, UyMax = Wc[2], UyMin = UyMax, TMax = Wc[3], TMin = TMax;
        for (int i = 0; i < C; ++i) {
            const double* Wj = &f->FD[4 * (size_t)st[i]];
            pMax = fmax(Wj[0], pMax);
            pMin = fmin(Wj[0], pMin);
            UxMax = fmax(Wj[1], UxMax);
            UxMin = fmin(Wj[1], UxMin);
            UyMax = fmax(Wj[2], UyMax);
            UyMin = fmin(Wj[2], UyMin);
            TMax = fmax(Wj[3], TMax);
            TMin = fmin(Wj[3], TMin);
        }
        double* lim = &f->limiter[4 * (size_t)celli];
        if (fabs(TMax - TMin) < VSMALL) {
            lim[0] = lim[1] = lim[2] = lim[3] = 1.0;
            for (int fi = 0; fi < C; ++fi) {
                const Slot s = slot(f, celli, fi);
                if (s.isBoundary) continue;
                double* o = &f->faceFD[4 * (2 * (size_t)s.face + s.isNei)];
                o[0] = Wc[0] + g[GPX] * s.dx + g[GPY] * s.dy;
                o[1] = Wc[1] + g[GUXX] * s.dx + g[GUYX] * s.dy;
                o[2] = Wc[2] + g[GUXY] * s.dx + g[GUYY] * s.dy;
                o[3] = Wc[3] + g[GTX] * s.dx + g[GTY] * s.dy;
                for (int q = 0; q < f->ngp; ++q) {  // the same linear state at the flux points
                    const double* d = &f->gpOffsets[4 * ((size_t)q * f->nf + s.face)];
                    const double gx = s.isNei ? d[2] : d[0], gy = s.isNei ? d[3] : d[1];
                    double* og = &f->faceGP[4 * (((size_t)q * f->nf + s.face) * 2 + s.isNei)];
                    og[0] = Wc[0] + g[GPX] * gx + g[GPY] * gy;
                    og[1] = Wc[1] + g[GUXX] * gx + g[GUYX] * gy;
                    og[2] = Wc[2] + g[GUXY] * gx + g[GUYY] * gy;
                    og[3] = Wc[3] + g[GTX] * gx + g[GTY] * gy;
                }
            }
            f->shockIndicator[2 * (size_t)celli] = 0;
            continue;
        }
        double tmpP = 2.0, tmpUx = 2.0, tmpUy = 2.0, tmpT = 2.0;
        double smoothnessValue = 0.0;
        bool isNotTrouble = true;
        const double pc = Wc[0], Uxc = Wc[1], Uyc = Wc[2], Tc = Wc[3];
        for (int fi = 0; fi < C; ++fi) {
            const Slot s = slot(f, celli, fi);
            if (s.isBoundary) continue;
            const double dx = s.dx, dy = s.dy;
            const double deltaR = afar_hypot(s.ex, s.ey);
            const double fraction = sqrt(dx * dx + dy * dy) / deltaR;
            const double* a = &f->RBFbasis[2 * (size_t)(maxS + 1) * s.face + (size_t)s.isNei * (maxS + 1)];
            double pf = a[0] * pc, Uxf = a[0] * Uxc, Uyf = a[0] * Uyc, Tf = a[0] * Tc;
            for (int ci = 0; ci < S; ++ci) {
                const double* Wj = &f->FD[4 * (size_t)st[ci]];
                pf += a[ci + 1] * Wj[0];
                Uxf += a[ci + 1] * Wj[1];
                Uyf += a[ci + 1] * Wj[2];
                Tf += a[ci + 1] * Wj[3];
            }
            double phif, d1, d2, yy;
            double* o = &f->faceFD[4 * (2 * (size_t)s.face + s.isNei)];
            limiter_process(tmpP, phif, d1, d2, yy, pc, pMax, pMin, g[GPX], g[GPY], dx, dy);
            o[0] = (fabs(pMax - pMin) >= SMALL) * pf + (fabs(pMax - pMin) < SMALL) * phif;
            isNotTrouble = phif > 0.0;
            const double pnb = f->FD[4 * (size_t)st[fi]];
            smoothnessValue = fmax(smoothnessValue, fabs((1.0 - fraction) * pc + fraction * pnb - phif) / deltaR +
                                                        (pf <= 0.0) * 1e20 + (phif <= 0.0) * 1e20);
            limiter_process(tmpUx, phif, d1, d2, yy, Uxc, UxMax, UxMin, g[GUXX], g[GUYX], dx, dy);
            o[1] = (fabs(UxMax - UxMin) >= SMALL) * Uxf + (fabs(UxMax - UxMin) < SMALL) * phif;
            limiter_process(tmpUy, phif, d1, d2, yy, Uyc, UyMax, UyMin, g[GUXY], g[GUYY], dx, dy);
            o[2] = (fabs(UyMax - UyMin) >= SMALL) * Uyf + (fabs(UyMax - UyMin) < SMALL) * phif;
            limiter_process(tmpT, phif, d1, d2, yy, Tc, TMax, TMin, g[GTX], g[GTY], dx, dy);
            o[3] = (fabs(TMax - TMin) >= SMALL) * Tf + (fabs(TMax - TMin) < SMALL) * phif;
            isNotTrouble = phif > 0.0;
            // flux points: RBF sums with each point's coefficients, or the
            // unlimited linear state where the compact range is flat
            for (int q = 0; q < f->ngp; ++q) {
                const double* ag = &f->RBFgp[(((size_t)q * f->nf + s.face) * 2 + s.isNei) * (size_t)(maxS + 1)];
                double pg = ag[0] * pc, Uxg = ag[0] * Uxc, Uyg = ag[0] * Uyc, Tg = ag[0] * Tc;
                for (int ci = 0; ci < S; ++ci) {
                    const double* Wj = &f->FD[4 * (size_t)st[ci]];
                    pg += ag[ci + 1] * Wj[0];
                    Uxg += ag[ci + 1] * Wj[1];
                    Uyg += ag[ci + 1] * Wj[2];
                    Tg += ag[ci + 1] * Wj[3];
                }
                const double* d = &f->gpOffsets[4 * ((size_t)q * f->nf + s.face)];
                const double gx = s.isNei ? d[2] : d[0], gy = s.isNei ? d[3] : d[1];
                double* og = &f->faceGP[4 * (((size_t)q * f->nf + s.face) * 2 + s.isNei)];
                og[0] = (fabs(pMax - pMin) >= SMALL) * pg + (fabs(pMax - pMin) < SMALL) * (pc + (g[GPX] * gx + g[GPY] * gy));
                og[1] = (fabs(UxMax - UxMin) >= SMALL) * Uxg +
                        (fabs(UxMax - UxMin) < SMALL) * (Uxc + (g[GUXX] * gx + g[GUYX] * gy));
                og[2] = (fabs(UyMax - UyMin) >= SMALL) * Uyg +
                        (fabs(UyMax - UyMin) < SMALL) * (Uyc + (g[GUXY] * gx + g[GUYY] * gy));
                og[3] = (fabs(TMax - TMin) >= SMALL) * Tg + (fabs(TMax - TMin) < SMALL) * (Tc + (g[GTX] * gx + g[GTY] * gy));
            }
        }
        lim[0] = fmax(tmpP, 0.0) * isNotTrouble;
        lim[1] = fmax(tmpUx, 0.0) * isNotTrouble;
        lim[2] = fmax(tmpUy, 0.0) * isNotTrouble;
        lim[3] = fmax(tmpT, 0.0) * isNotTrouble;
        smoothnessValue = smoothnessValue / (0.5 * (pMax + pMin) * isNotTrouble + VSMALL);
        f->shockIndicator[2 * (size_t)celli] = (smoothnessValue > 1.0);
    }
}

void oracle_bvd(oracle_fields* f) {
    const int maxS = f->maxS;
    // the BVD loop reads raw limiters of this cell only, so one pass over cells is race-free
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
        const double Tc = f->FD[4 * (size_t)celli + 3];
        const double gTx = f->gradFD[8 * (size_t)celli + GTX], gTy = f->gradFD[8 * (size_t)celli + GTY];
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
            const double TfOwn = f->faceFD[4 * (2 * (size_t)s.face) + 3];
            const double TfNei = f->faceFD[4 * (2 * (size_t)s.face + 1) + 3];
            const double Tf_cur = s.isNei * TfNei + (1 - s.isNei) * TfOwn;
            const double Tf_nei = (1 - s.isNei) * TfNei + s.isNei * TfOwn;
            const double l = to_bj(lim[3]);
            const double Tf_cur_TVD = Tc + l * (gTx * s.dx + gTy * s.dy);
            const int typeval = fabs(Tf_cur_TVD - Tf_nei) < BVD_FACTOR * fabs(Tf_cur - Tf_nei);
            const int isNotFlat = fabs(Tf_cur - Tf_nei) > VSMALL;
            const int isLimited = (l < VALID_LIM_FACTOR) + (shockType_extend > 0);
            BVDtype = std::max(BVDtype, isNotFlat * isLimited * typeval);
        }
        int finalType = (shockType > 0) * (shockType_extend > 1) * (-2) + (BVDtype > 0);
        if (f->boundaryFirstOrder && bfacesNum > 0) {
            // extension: first-order face states in cells with a boundary face
            finalType = finalType ? finalType : 1;
            lim[0] = lim[1] = lim[2] = lim[3] = 0.0;
            f->shockIndicator[2 * (size_t)celli + 1] = (int8_t)finalType;
            continue;
        }
        f->shockIndicator[2 * (size_t)celli + 1] = (int8_t)finalType;
        // limiter transform (only this cell reads its own limiter in K6)
        const bool isShock = finalType < 0;
        for (int k = 0; k < 4; ++k) lim[k] = transform_limiter(lim[k], isShock);
    }
}

// The flux of face faceI from the states reconstructed at a point with
// offsets (dxL, dyL) from the owner and (dxR, dyR) from the neighbour; fo, fn
// are the high-order states there (cudaRiemannSolver.h:155-251).
static void point_flux(const oracle_fields* f, int faceI, double R, double Cv, double deltaRxLeft, double deltaRyLeft,
                       double deltaRxRight, double deltaRyRight, const double* fo, const double* fn, double* out) {
    const int own = f->neighbour[2 * faceI];
    const int nei = f->neighbour[2 * faceI + 1];
    const double* F = &f->FACE[5 * (size_t)faceI];
    const bool isTVD_left = f->shockIndicator[2 * (size_t)own + 1] != 0;
    const bool isTVD_right = f->shockIndicator[2 * (size_t)nei + 1] != 0;
    const bool isShock = f->shockIndicator[2 * (size_t)own + 1] < 0 || f->shockIndicator[2 * (size_t)nei + 1] < 0;
    const double magSf = F[2], Sfx = F[3], Sfy = F[4];
    const double* Wo = &f->FD[4 * (size_t)own];
    const double* Wn = &f->FD[4 * (size_t)nei];
    const double* Lo = &f->limiter[4 * (size_t)own];
    const double* Ln = &f->limiter[4 * (size_t)nei];
    const double* Go = &f->gradFD[8 * (size_t)own];
    const double* Gn = &f->gradFD[8 * (size_t)nei];
    const double pLeftTVD = Wo[0] + Lo[0] * (Go[GPX] * deltaRxLeft + Go[GPY] * deltaRyLeft);
    const double UxLeftTVD = Wo[1] + Lo[1] * (Go[GUXX] * deltaRxLeft + Go[GUYX] * deltaRyLeft);
    const double UyLeftTVD = Wo[2] + Lo[2] * (Go[GUXY] * deltaRxLeft + Go[GUYY] * deltaRyLeft);
    const double TLeftTVD = Wo[3] + Lo[3] * (Go[GTX] * deltaRxLeft + Go[GTY] * deltaRyLeft);
    const double rhoLeftTVD = pLeftTVD / (R * TLeftTVD);
    const double pRightTVD = Wn[0] + Ln[0] * (Gn[GPX] * deltaRxRight + Gn[GPY] * deltaRyRight);
    const double UxRightTVD = Wn[1] + Ln[1] * (Gn[GUXX] * deltaRxRight + Gn[GUYX] * deltaRyRight);
    const double UyRightTVD = Wn[2] + Ln[2] * (Gn[GUXY] * deltaRxRight + Gn[GUYY] * deltaRyRight);
    const double TRightTVD = Wn[3] + Ln[3] * (Gn[GTX] * deltaRxRight + Gn[GTY] * deltaRyRight);
    const double rhoRightTVD = pRightTVD / (R * TRightTVD);
    const double pLeft = (1 - isTVD_left) * fo[0] + isTVD_left * pLeftTVD;
    const double UxLeft = (1 - isTVD_left) * fo[1] + isTVD_left * UxLeftTVD;
    const double UyLeft = (1 - isTVD_left) * fo[2] + isTVD_left * UyLeftTVD;
    const double TLeft = (1 - isTVD_left) * fo[3] + isTVD_left * TLeftTVD;
    const double rhoLeft = pLeft / (R * TLeft);
    const double pRight = (1 - isTVD_right) * fn[0] + isTVD_right * pRightTVD;
    const double UxRight = (1 - isTVD_right) * fn[1] + isTVD_right * UxRightTVD;
    const double UyRight = (1 - isTVD_right) * fn[2] + isTVD_right * UyRightTVD;
    const double TRight = (1 - isTVD_right) * fn[3] + isTVD_right * TRightTVD;
    const double rhoRight = pRight / (R * TRight);
    double deltaP, deltaRho;
    minimize_delta(deltaP, pLeft, pRight, pLeftTVD, pRightTVD);
    minimize_delta(deltaRho, rhoLeft, rhoRight, rhoLeftTVD, rhoRightTVD);
    oracle_central_flux(out, rhoLeft, rhoRight, pLeft, pRight, UxLeft, UxRight, UyLeft, UyRight, TLeft, TRight, Cv,
                        Sfx, Sfy);
    if (isShock)
        oracle_kfvs_flux(out, rhoLeft, rhoRight, pLeft, pRight, UxLeft, UxRight, UyLeft, UyRight, TLeft, TRight, R, Cv,
                         magSf, Sfx, Sfy);
    else
        oracle_roe_flux(out, rhoLeft, rhoRight, pLeft, pRight, UxLeft, UxRight, UyLeft, UyRight, TLeft, TRight, deltaP,
                        deltaRho, R, Cv, magSf, Sfx, Sfy);
}

void oracle_evaluate_flux(oracle_fields* f, double R, double Cv) {
#pragma omp parallel for schedule(static)
    for (int faceI = 0; faceI < f->nf; ++faceI) {
        const int own = f->neighbour[2 * faceI];
        const int nei = f->neighbour[2 * faceI + 1];
        const double* F = &f->FACE[5 * (size_t)faceI];
        double* out = &f->Flux[4 * (size_t)faceI];
        if (f->ngp == 0) {
            const double xn = f->CELL[2 * nei] + shx(f, faceI), yn = f->CELL[2 * nei + 1] + shy(f, faceI);
            point_flux(f, faceI, R, Cv, F[0] - f->CELL[2 * own], F[1] - f->CELL[2 * own + 1], F[0] - xn, F[1] - yn,
                       &f->faceFD[4 * (2 * (size_t)faceI)], &f->faceFD[4 * (2 * (size_t)faceI + 1)], out);
            continue;
        }
        // extension: weighted sum over the face's quadrature points, in point order
        for (int q = 0; q < f->ngp; ++q) {
            const double* d = &f->gpOffsets[4 * ((size_t)q * f->nf + faceI)];
            double fq[4];
            point_flux(f, faceI, R, Cv, d[0], d[1], d[2], d[3], &f->faceGP[4 * (((size_t)q * f->nf + faceI) * 2)],
                       &f->faceGP[4 * (((size_t)q * f->nf + faceI) * 2 + 1)], fq);
            const double w = f->gpWeights[q];
            for (int k = 0; k < 4; ++k) out[k] = q == 0 ? w * fq[k] : out[k] + w * fq[k];
        }
    }
    // every boundary face (the reference only covers faceI < nf of its grid)
    for (int b = 0; b < f->nb; ++b) {
        const double* F = &f->FACE[5 * ((size_t)f->nf + b)];
        const double Sfx = F[3], Sfy = F[4];
        const double* G = &f->FD[4 * ((size_t)f->nc + b)];
        const double bp = G[0], bUx = G[1], bUy = G[2], bT = G[3];
        const double rho = bp / (R * bT);
        const double rhoFlux = rho * (bUx * Sfx + bUy * Sfy);
        double* out = &f->Flux[4 * ((size_t)f->nf + b)];
        out[0] = rhoFlux;
        out[1] = rhoFlux * bUx + bp * Sfx;
        out[2] = rhoFlux * bUy + bp * Sfy;
        out[3] = rhoFlux * (Cv * bT + 0.5 * (bUx * bUx + bUy * bUy) + bp / rho);
    }
}

void oracle_evaluate_residual(oracle_fields* f) {
#pragma omp parallel for schedule(static)
    for (int idx = 0; idx < f->nc; ++idx) {
        double* Res = &f->Res[4 * (size_t)idx];
        for (int fi = 0; fi < f->compactStencilSize[idx]; ++fi) {
            int faceI = f->cellFaces[(size_t)f->maxC * idx + fi];
            const int direction = (faceI > 0) - (faceI < 0);
            faceI = abs(faceI) - 1;
            const double* Fl = &f->Flux[4 * (size_t)faceI];
            Res[0] += direction * Fl[0];
            Res[1] += direction * Fl[1];
            Res[2] += direction * Fl[2];
            Res[3] += direction * Fl[3];
        }
    }
}

void oracle_update(oracle_fields* f, double R, double gamma, double deltaT, double beta1, double beta2, double beta3) {
#pragma omp parallel for schedule(static)
    for (int celli = 0; celli < f->nc; ++celli) {
        double* U = &f->cFD[4 * (size_t)celli];
        const double* Uo = &f->cFDold[4 * (size_t)celli];
        double* Res = &f->Res[4 * (size_t)celli];
        const double V = f->cellVolume[celli];
        const double rho = beta1 * Uo[0] + beta2 * U[0] + beta3 * deltaT * (-Res[0] / V);
        const double rhoUx = beta1 * Uo[1] + beta2 * U[1] + beta3 * deltaT * (-Res[1] / V);
        const double rhoUy = beta1 * Uo[2] + beta2 * U[2] + beta3 * deltaT * (-Res[2] / V);
        const double rhoE = beta1 * Uo[3] + beta2 * U[3] + beta3 * deltaT * (-Res[3] / V);
        U[0] = rho;
        U[1] = rhoUx;
        U[2] = rhoUy;
        U[3] = rhoE;
        double* W = &f->FD[4 * (size_t)celli];
        W[0] = (gamma - 1.0) * (rhoE - 0.5 * (rhoUx * rhoUx + rhoUy * rhoUy) / rho);
        W[1] = rhoUx / rho;
        W[2] = rhoUy / rho;
        W[3] = W[0] / (R * rho);
        Res[0] = Res[1] = Res[2] = Res[3] = 0.0;
    }
}

void oracle_update_boundary(oracle_fields* f, double R, double Cv) {
    (void)R;
    (void)Cv;
    for (int b = 0; b < f->nb; ++b) {
        const int curFC = f->boundaryFaceNeiLabel[b];
        const double* w = &f->FD[4 * (size_t)curFC];
        double* g = &f->FD[4 * ((size_t)f->nc + b)];
        const double* F = &f->FACE[5 * ((size_t)f->nf + b)];
        const double magSf = F[2], Sfx = F[3], Sfy = F[4];
        switch (f->boundaryFacesType[b]) {
            case 0:  // constantTemperatureWallBoundary.h
                g[0] = w[0]; g[1] = 0.0; g[2] = 0.0; g[3] = 1.0;
                break;
            case 1:  // adiabaticWallBoundary.h
                g[0] = w[0]; g[1] = 0.0; g[2] = 0.0; g[3] = w[3];
                break;
            case 2: {  // symmetryBoundary.h
                const double nfx = Sfx / magSf, nfy = Sfy / magSf;
                g[0] = w[0];
                g[1] = w[1] - (w[1] * nfx + w[2] * nfy) * nfx;
                g[2] = w[2] - (w[1] * nfx + w[2] * nfy) * nfy;
                g[3] = w[3];
                break;
            }
            case 3:  // fixedValueBoundary.h
                g[0] = 1.0; g[1] = 3.0; g[2] = 0.0; g[3] = 1.0;
                break;
            case 4:  // outletBoundary.h
                g[0] = w[0]; g[1] = w[1]; g[2] = w[2]; g[3] = w[3];
                break;
            case 5:  // fixedValue2Boundary.h
                g[0] = 30.05945; g[1] = 4.07795; g[2] = 0.0; g[3] = 5.97677;
                break;
            default:  // 6: per-patch fixed value
                for (int k = 0; k < 4; ++k) g[k] = f->BVALUE ? f->BVALUE[4 * (size_t)b + k] : 0.0;
        }
    }
}

void oracle_store_conserved(oracle_fields* f) {
    memcpy(f->cFDold, f->cFD, sizeof(double) * 4 * (size_t)f->nc);
}

void oracle_evolve(oracle_fields* f, double R, double Cv, double deltaT, int exactRK) {
    const double gamma = R / Cv + 1.0;
    const double lit1[3] = {1.0, 0.75, 0.333333}, lit2[3] = {0.0, 0.25, 0.666667}, lit3[3] = {1.0, 0.25, 0.666667};
    const double ex1[3] = {1.0, 0.75, 1.0 / 3.0}, ex2[3] = {0.0, 0.25, 2.0 / 3.0}, ex3[3] = {1.0, 0.25, 2.0 / 3.0};
    for (int i = 0; i < 3; ++i) {
        oracle_reconstruct(f);
        oracle_bvd(f);
        oracle_evaluate_flux(f, R, Cv);
        oracle_evaluate_residual(f);
        oracle_update(f, R, gamma, deltaT, exactRK ? ex1[i] : lit1[i], exactRK ? ex2[i] : lit2[i],
                      exactRK ? ex3[i] : lit3[i]);
        oracle_update_boundary(f, R, Cv);
    }
    oracle_store_conserved(f);
}

// n steps with the CFL time step of each (the application loop of
// cudaEuler2DmixO3RK3solver.cpp:53-64 with adjustTimeStep true); dts[n] out.
void oracle_run(oracle_fields* f, double R, double Cv, double CFL, int nsteps, int exactRK, double* dts) {
    for (int s = 0; s < nsteps; ++s) {
        const double dt = oracle_adjust_time_step(f, R, Cv, CFL);
        if (dts) dts[s] = dt;
        oracle_evolve(f, R, Cv, dt, exactRK);
    }
}

double oracle_exp(double x) { return afar_exp(x); }
double oracle_erfc(double x) { return afar_erfc(x); }

}  // extern "C"
