// The per-cell arithmetic of the RBF precompute
// (src/CUDA/cudaEuler2DmixO3RK3/basicCUDAsolver.cpp:461-619), written once for
// the host (precompute.cpp) and the device (precompute_gpu.hip): the same
// operations in the same order, so both give the same coefficients bit for
// bit (IEEE double, no contraction, correctly rounded sqrt and division).
//
//  rbf_moments   P_ij = sum_k w_k^(i) sqrt(shapeFactor + |g_k^(i) - x_j|^2/ref^2)
//                (cell averages of the multiquadric centred on member j over
//                member i), stored transposed into the SVD work matrix.
//  rbf_pinv      Moore-Penrose pseudo-inverse G by one-sided (Hestenes)
//                Jacobi SVD, singular values equal to zero dropped (OpenFOAM's
//                SVDinv with minCondition 0, basicCUDAsolver.cpp:509).
//  rbf_trouble   any |G_ij| > 1e10 (basicCUDAsolver.cpp:510-529), or NaN.
//  rbf_face      a = G^T b - mean(G^T b) + 1/n for the face centre
//                (basicCUDAsolver.cpp:568-618).
#pragma once
#include <math.h>

#include "afar_math.h"

namespace afar {

constexpr int kRbfMaxN = 21;  // stencil (at most 20) plus the cell

// Row-major P into U from the members' quadrature: pt(i, k, x, y, w) gives
// point k < count[i] of member i, already in the cell's frame.  X: 2n member
// centres.
template <class PointFn>
AFAR_HD void rbf_moments(int n, const double* X, const int* count, PointFn pt, double ref, double sf, double* U) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const double x0 = X[2 * j], y0 = X[2 * j + 1];
            double acc = 0.0;
            for (int k = 0; k < count[i]; ++k) {
                double gx, gy, gw;
                pt(i, k, gx, gy, gw);
                const double dx = (gx - x0) / ref, dy = (gy - y0) / ref;
                const double r2 = dx * dx + dy * dy;
                acc += gw * sqrt(sf + r2);
            }
            U[i * n + j] = acc;
        }
}

// One-sided Jacobi on the columns of A^T: U holds A^T column-major on entry
// (U[j*n + i] = A[i][j]); V, sig are work arrays; G = pinv(A), row-major.
AFAR_HD void rbf_pinv(int n, double* U, double* V, double* sig, double* G) {
    // U[j*n + i] must equal A[i*n + j]: transpose the moment matrix in place
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            const double t = U[i * n + j];
            U[i * n + j] = U[j * n + i];
            U[j * n + i] = t;
        }
    for (int i = 0; i < n * n; ++i) V[i] = 0.0;
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 80; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double* up = &U[p * n];
                double* uq = &U[q * n];
                double alpha = 0.0, beta = 0.0, gamma = 0.0;
                for (int k = 0; k < n; ++k) {
                    alpha += up[k] * up[k];
                    beta += uq[k] * uq[k];
                    gamma += up[k] * uq[k];
                }
                if (gamma == 0.0 || fabs(gamma) <= 1e-15 * sqrt(alpha * beta)) continue;
                rotated = true;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                double t;
                if (fabs(zeta) > 1e150)
                    t = 0.5 / zeta;
                else
                    t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int k = 0; k < n; ++k) {
                    const double a = up[k], b = uq[k];
                    up[k] = c * a - s * b;
                    uq[k] = s * a + c * b;
                }
                double* vp = &V[p * n];
                double* vq = &V[q * n];
                for (int k = 0; k < n; ++k) {
                    const double a = vp[k], b = vq[k];
                    vp[k] = c * a - s * b;
                    vq[k] = s * a + c * b;
                }
            }
        if (!rotated) break;
    }
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += U[j * n + k] * U[j * n + k];
        sig[j] = sqrt(s);
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double g = 0.0;
            for (int k = 0; k < n; ++k) {
                if (!(sig[k] > 0.0)) continue;
                // V[i][k] * (1/sigma_k) * u_k[j], u_k = U[:,k] / sigma_k
                g += V[k * n + i] * (U[k * n + j] / sig[k]) / sig[k];
            }
            G[i * n + j] = g;
        }
}

AFAR_HD bool rbf_trouble(int n, const double* G) {
    for (int i = 0; i < n * n; ++i)
        if (fabs(G[i]) > 1e10 || G[i] != G[i]) return true;
    return false;
}

// Coefficients of the face at (xf, yf) (in the cell's frame), into out[0..n).
AFAR_HD void rbf_face(int n, const double* G, const double* X, double xf, double yf, double ref, double sf,
                      double* b, double* out) {
    for (int i = 0; i < n; ++i) {
        const double dx = (xf - X[2 * i]) / ref, dy = (yf - X[2 * i + 1]) / ref;
        b[i] = sqrt(sf + (dx * dx + dy * dy));
    }
    double sum = 0.0;
    for (int i = 0; i < n; ++i) {
        double v = 0.0;
        for (int j = 0; j < n; ++j) v += G[j * n + i] * b[j];
        out[i] = v;
        sum += v;
    }
    for (int i = 0; i < n; ++i) out[i] = (out[i] - sum / (double)n) + 1.0 / (double)n;
}

}  // namespace afar
