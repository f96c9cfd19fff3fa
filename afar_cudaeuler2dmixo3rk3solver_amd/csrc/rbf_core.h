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
//                Jacobi SVD with round-robin rotation order (a round's
//                rotations touch disjoint columns, so a wavefront applies them
//                at once and the host in sequence, with the same results),
//                singular values equal to zero dropped (OpenFOAM's SVDinv with
//                minCondition 0, basicCUDAsolver.cpp:509).
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

// Round-robin (circle) schedule of the Jacobi rotations: m = n rounded up to
// even; in round r (0 <= r < m - 1) position i < m/2 pairs the columns at
// positions i and m - 1 - i, column 0 fixed and the others rotated by r.  A
// round's pairs are disjoint, so its rotations commute; a pair holding the
// padding column n (odd n) is skipped.  Returns p < q.
AFAR_HD void rbf_rr_pair(int m, int r, int i, int& p, int& q) {
    auto col = [&](int pos) { return pos == 0 ? 0 : (pos - 1 + r) % (m - 1) + 1; };
    const int a = col(i), b = col(m - 1 - i);
    p = a < b ? a : b;
    q = a < b ? b : a;
}

// The rotation that orthogonalises columns up and uq (length n): false when
// they already are (to the tolerance), else its cosine and sine.  The three
// dot products are summed in index order.
AFAR_HD bool rbf_rotation(int n, const double* up, const double* uq, double& c, double& s) {
    double alpha = 0.0, beta = 0.0, gamma = 0.0;
    for (int k = 0; k < n; ++k) {
        alpha += up[k] * up[k];
        beta += uq[k] * uq[k];
        gamma += up[k] * uq[k];
    }
    if (gamma == 0.0 || fabs(gamma) <= 1e-15 * sqrt(alpha * beta)) return false;
    const double zeta = (beta - alpha) / (2.0 * gamma);
    double t;
    if (fabs(zeta) > 1e150)
        t = 0.5 / zeta;
    else
        t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    c = 1.0 / sqrt(1.0 + t * t);
    s = c * t;
    return true;
}

AFAR_HD void rbf_rotate(double& a, double& b, double c, double s) {
    const double x = a, y = b;
    a = c * x - s * y;
    b = s * x + c * y;
}

// Entry (i, j) of the pseudo-inverse from the converged columns: the sum over
// k with sigma_k > 0 of V[i][k] (u_k[j] / sigma_k) / sigma_k, in index order.
AFAR_HD double rbf_pinv_entry(int n, const double* U, const double* V, const double* sig, int i, int j) {
    double g = 0.0;
    for (int k = 0; k < n; ++k) {
        if (!(sig[k] > 0.0)) continue;
        g += V[k * n + i] * (U[k * n + j] / sig[k]) / sig[k];
    }
    return g;
}

// Moore-Penrose pseudo-inverse by one-sided (Hestenes) Jacobi on the columns
// of A^T, rotations in the round-robin order above.  U holds A row-major on
// entry and is transposed in place (U[j*n + i] = A[i][j]); V, sig are work
// arrays; G = pinv(A), row-major.
AFAR_HD void rbf_pinv(int n, double* U, double* V, double* sig, double* G) {
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            const double t = U[i * n + j];
            U[i * n + j] = U[j * n + i];
            U[j * n + i] = t;
        }
    for (int i = 0; i < n * n; ++i) V[i] = 0.0;
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    const int m = n + (n & 1);
    for (int sweep = 0; sweep < 80; ++sweep) {
        bool rotated = false;
        for (int r = 0; r < m - 1; ++r)
            for (int i = 0; i < m / 2; ++i) {
                int p, q;
                rbf_rr_pair(m, r, i, p, q);
                if (q >= n) continue;
                double c, s;
                if (!rbf_rotation(n, &U[p * n], &U[q * n], c, s)) continue;
                rotated = true;
                for (int k = 0; k < n; ++k) rbf_rotate(U[p * n + k], U[q * n + k], c, s);
                for (int k = 0; k < n; ++k) rbf_rotate(V[p * n + k], V[q * n + k], c, s);
            }
        if (!rotated) break;
    }
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += U[j * n + k] * U[j * n + k];
        sig[j] = sqrt(s);
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) G[i * n + j] = rbf_pinv_entry(n, U, V, sig, i, j);
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
