// Host precompute of the reconstruction: stencils, Gauss quadrature, the RBF
// moment matrices and their pseudo-inverses, the per-face RBF coefficients and
// the basisOwnList/basisNeiList cache files.
//
// Follows src/CUDA/cudaEuler2DmixO3RK3/basicCUDAsolver.cpp:261-694 and
// src/gaussQuadrature/d2GaussQuadrature.C:23-547.
#include <math.h>
#include <sched.h>
#include <stdio.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <fstream>
#include <sstream>
#include <thread>

#include "afar_host.h"
#include "rbf_core.h"

namespace afar {

// Threads of a parallel host loop: AFAR_HOST_THREADS if set, else the CPUs
// this process may run on (its affinity mask, which a container's CPU share
// narrows; hardware_concurrency counts the whole machine), capped by
// OMP_NUM_THREADS when that is set.
static unsigned host_threads() {
    if (const char* e = getenv("AFAR_HOST_THREADS")) return (unsigned)std::max(1, atoi(e));
    unsigned hw = std::thread::hardware_concurrency();
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) hw = std::min(hw, (unsigned)std::max(1, CPU_COUNT(&set)));
    if (const char* e = getenv("OMP_NUM_THREADS")) {
        const int k = atoi(e);
        if (k > 0) hw = std::min(hw, (unsigned)k);
    }
    return std::max(1u, hw);
}

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& fn) {
    if (n <= 0) return;
    unsigned hw = host_threads();
    int nt = (int)std::max(1u, std::min(hw, 64u));
    const int64_t chunk = std::max<int64_t>(256, n / (int64_t)(nt * 16) + 1);
    if (nt == 1 || n <= chunk) {
        fn(0, n);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    std::exception_ptr err = nullptr;
    std::atomic<bool> failed{false};
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&]() {
            for (;;) {
                const int64_t b = next.fetch_add(chunk);
                if (b >= n || failed.load()) break;
                try {
                    fn(b, std::min(n, b + chunk));
                } catch (...) {
                    if (!failed.exchange(true)) err = std::current_exception();
                }
            }
        });
    for (auto& t : th) t.join();
    if (err) std::rethrow_exception(err);
}

// ------------------------------------------------------------------ stencils

// Compact stencil = face neighbours in face order, then the boundary faces'
// pseudo-cells (basicCUDAsolver.cpp:275-299).  Extended stencil either adds
// the second face-neighbour ring (faceNeighbour, :333-355, skipping the rings
// of pseudo-cells) or the cells sharing a vertex (edgeNeighbour, :356-375: in
// a one-cell-thick mesh OpenFOAM's cellEdges/edgeCells reach exactly the
// vertex neighbours).  The vertex walk follows the polygon order and, at each
// vertex, ascending cell index.
Stencils build_stencils(const Mesh& m, int stencilType) {
    const int nc = m.nc, nf = m.nf, nb = m.nb;
    if (stencilType != 0 && stencilType != 1)
        throw Error(1, "Invalid stencil type setting. Valid settings are: faceNeighbour, edgeNeighbour");
    std::vector<int32_t> cc_ptr(nc + 1, 0);
    for (int f = 0; f < nf; ++f) {
        cc_ptr[m.own[f] + 1]++;
        cc_ptr[m.nei[f] + 1]++;
    }
    for (int b = 0; b < nb; ++b) cc_ptr[m.own[nf + b] + 1]++;
    for (int c = 0; c < nc; ++c) cc_ptr[c + 1] += cc_ptr[c];
    std::vector<int32_t> cc(cc_ptr[nc]), cf(cc_ptr[nc]);
    std::vector<Vec2> csh(cc_ptr[nc]);
    std::vector<int32_t> fill(cc_ptr.begin(), cc_ptr.end() - 1);
    for (int f = 0; f < nf; ++f) {
        const int a = m.own[f], b = m.nei[f];
        int k = fill[a]++;
        cc[k] = b;
        cf[k] = f + 1;
        csh[k] = m.fshift[f];
        k = fill[b]++;
        cc[k] = a;
        cf[k] = -(f + 1);
        csh[k] = Vec2{-m.fshift[f].x, -m.fshift[f].y};
    }
    for (int b = 0; b < nb; ++b) {
        const int c = m.own[nf + b];
        const int k = fill[c]++;
        cc[k] = nc + b;
        cf[k] = nf + b + 1;
        csh[k] = Vec2{};
    }
    // point -> (cell, vertex slot) incidence, cells ascending
    std::vector<int32_t> pc_ptr(m.npts + 1, 0);
    for (int c = 0; c < nc; ++c)
        for (int k = m.cell_ptr[c]; k < m.cell_ptr[c + 1]; ++k) pc_ptr[m.cell_pts[k] + 1]++;
    for (int p = 0; p < m.npts; ++p) pc_ptr[p + 1] += pc_ptr[p];
    std::vector<int32_t> pc(pc_ptr[m.npts]), pk(pc_ptr[m.npts]);
    {
        std::vector<int32_t> pf(pc_ptr.begin(), pc_ptr.end() - 1);
        for (int c = 0; c < nc; ++c)
            for (int k = m.cell_ptr[c]; k < m.cell_ptr[c + 1]; ++k) {
                const int s = pf[m.cell_pts[k]]++;
                pc[s] = c;
                pk[s] = k;
            }
    }

    Stencils st;
    st.S.assign(nc, 0);
    st.Cc.assign(nc, 0);
    st.sten.assign((size_t)nc * kMaxValidStencil, -1);
    st.sten_shift.assign((size_t)nc * kMaxValidStencil, Vec2{});
    st.cellFaces.assign((size_t)nc * kMaxValidStencil, 0);
    std::atomic<int> maxS{0}, maxC{0};
    std::atomic<int> bad{-1};
    parallel_for(nc, [&](int64_t c0, int64_t c1) {
        std::vector<int32_t> ids;
        std::vector<Vec2> sh;
        int lS = 0, lC = 0;
        for (int64_t c = c0; c < c1; ++c) {
            ids.clear();
            sh.clear();
            auto has = [&](int id) {
                if (id == (int)c) return true;
                for (size_t i = 0; i < ids.size(); ++i)
                    if (ids[i] == id) return true;
                return false;
            };
            const int C = cc_ptr[c + 1] - cc_ptr[c];
            for (int k = cc_ptr[c]; k < cc_ptr[c + 1]; ++k) {
                ids.push_back(cc[k]);
                sh.push_back(csh[k]);
            }
            if (stencilType == 0) {
                for (int i = 0; i < C; ++i) {
                    const int nbc = ids[i];
                    if (nbc >= nc) continue;
                    const Vec2 s0 = sh[i];
                    for (int k = cc_ptr[nbc]; k < cc_ptr[nbc + 1]; ++k) {
                        const int x = cc[k];
                        if (has(x)) continue;
                        ids.push_back(x);
                        sh.push_back(Vec2{s0.x + csh[k].x, s0.y + csh[k].y});
                    }
                }
            } else {
                for (int k = m.cell_ptr[c]; k < m.cell_ptr[c + 1]; ++k) {
                    const int p = m.cell_pts[k];
                    for (int s = pc_ptr[p]; s < pc_ptr[p + 1]; ++s) {
                        const int x = pc[s];
                        if (has(x)) continue;
                        const int kx = pk[s];
                        ids.push_back(x);
                        sh.push_back(Vec2{(m.cell_vs[2 * (size_t)k] - m.cell_vs[2 * (size_t)kx]) * m.period.x,
                                          (m.cell_vs[2 * (size_t)k + 1] - m.cell_vs[2 * (size_t)kx + 1]) * m.period.y});
                    }
                }
            }
            const int S = (int)ids.size();
            if (S > kMaxValidStencil) {
                bad = (int)c;
                continue;
            }
            st.S[c] = (uint8_t)S;
            st.Cc[c] = (uint8_t)C;
            for (int i = 0; i < S; ++i) {
                st.sten[(size_t)c * kMaxValidStencil + i] = ids[i];
                st.sten_shift[(size_t)c * kMaxValidStencil + i] = sh[i];
            }
            for (int i = 0; i < C; ++i) st.cellFaces[(size_t)c * kMaxValidStencil + i] = cf[cc_ptr[c] + i];
            lS = std::max(lS, S);
            lC = std::max(lC, C);
        }
        int v = maxS.load();
        while (lS > v && !maxS.compare_exchange_weak(v, lS)) {}
        v = maxC.load();
        while (lC > v && !maxC.compare_exchange_weak(v, lC)) {}
    });
    if (bad >= 0) throw Error(3, "Too large stencil! (more than 20 cells) at cell " + std::to_string(bad.load()));
    st.maxS = maxS;
    st.maxC = maxC;
    return st;
}

// ------------------------------------------------------------------ quadrature

namespace {

void tri_rule(int cOrder, std::vector<Vec2>& p, std::vector<double>& w) {
    // standard triangle A(0,0) B(1,0) C(0,1), area 0.5 (d2GaussQuadrature.C:421-486)
    if (cOrder == 1) {
        p = {{1.0 / 3.0, 1.0 / 3.0}};
        w = {0.5};
    } else if (cOrder > 1 && cOrder <= 3) {
        p = {{0.5, 0.5}, {0.5, 0.0}, {0.0, 0.5}};
        w = {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
    } else if (cOrder > 3 && cOrder <= 5) {
        // kept as the reference writes it: centroid weight +21/96, weights sum to 1
        p = {{1.0 / 3.0, 1.0 / 3.0}, {3.0 / 5.0, 1.0 / 5.0}, {1.0 / 5.0, 1.0 / 5.0}, {1.0 / 5.0, 3.0 / 5.0}};
        w = {21.0 / 96.0, 25.0 / 96.0, 25.0 / 96.0, 25.0 / 96.0};
    } else if (cOrder > 5 && cOrder <= 7) {
        const double a = 0.816847572980459, b = 0.091576213509771, c = 0.108103018168070, d = 0.445948490915965;
        const double wa = 0.109951743655322 * 0.5, wc = 0.223381589678011 * 0.5;
        p = {{a, b}, {b, a}, {b, b}, {c, d}, {d, c}, {d, d}};
        w = {wa, wa, wa, wc, wc, wc};
    } else {
        throw Error(1, "No stored information for cOrder = " + std::to_string(cOrder));
    }
}

void rec_rule(int cOrder, std::vector<Vec2>& p, std::vector<double>& w) {
    // standard square [-1,1]^2, area 4 (d2GaussQuadrature.C:333-417)
    if (cOrder == 1) {
        p = {{0.0, 0.0}};
        w = {4.0};
    } else if (cOrder > 1 && cOrder <= 3) {
        // The reference assigns ws[0] four times and leaves ws[1..3] unset
        // (d2GaussQuadrature.C:358-361): undefined, so it is refused here.
        throw Error(1, "cOrder 2-3 on quadrilaterals is undefined in the reference (uninitialised weights)");
    } else if (cOrder > 3 && cOrder <= 5) {
        const double s = sqrt(0.6), a = 5.0 / 9.0, b = 8.0 / 9.0;
        p = {{-s, -s}, {-s, 0.0}, {-s, s}, {0.0, -s}, {0.0, 0.0}, {0.0, s}, {s, -s}, {s, 0.0}, {s, s}};
        w = {a * a, a * b, a * a, b * a, b * b, b * a, a * a, a * b, a * a};
    } else if (cOrder > 5 && cOrder <= 7) {
        const double g[4] = {0.861136311594053, 0.339981043584856, -0.339981043584856, -0.861136311594053};
        const double h[4] = {0.347854845137454, 0.652145154862546, 0.652145154862546, 0.347854845137454};
        // point k = (g[k%4], g[k/4]) with weight h[k%4]*h[k/4], the reference's table order
        p.clear();
        w.clear();
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                p.push_back({g[i], g[j]});
                w.push_back(h[i] * h[j]);
            }
    } else {
        throw Error(1, "No stored information for cOrder = " + std::to_string(cOrder));
    }
}

void bar_rule(int fOrder, std::vector<double>& p, std::vector<double>& w) {
    // standard bar [-1,1] (d2GaussQuadrature.C:489-547)
    if (fOrder == 1) {
        p = {0.0};
        w = {2.0};
    } else if (fOrder > 1 && fOrder <= 3) {
        p = {-1 / sqrt(3.0), 1 / sqrt(3.0)};
        w = {1.0, 1.0};
    } else if (fOrder > 3 && fOrder <= 5) {
        p = {-sqrt(0.6), 0.0, sqrt(0.6)};
        w = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    } else if (fOrder > 5 && fOrder <= 7) {
        p = {0.861136311594053, 0.339981043584856, -0.339981043584856, -0.861136311594053};
        w = {0.347854845137454, 0.652145154862546, 0.652145154862546, 0.347854845137454};
    } else {
        throw Error(1, "No stored information for fOrder = " + std::to_string(fOrder));
    }
}

}  // namespace

// Gauss points of a triangle (vertices from cell_xy[b..b+2]) from the unit
// rule (sp, sw), weights divided by the cell's 2D area.
static void tri_points(const Mesh& m, int cell, const Vec2* sp, const double* sw, int n, Vec2* pts, double* w) {
    const int b = m.cell_ptr[cell];
    const double faceArea = m.V[cell] / m.depth;
    const Vec2 p1 = m.cell_xy[b], p2 = m.cell_xy[b + 1], p3 = m.cell_xy[b + 2];
    const double jacobi = fabs((p2.x - p1.x) * (p3.y - p1.y) - (p3.x - p1.x) * (p2.y - p1.y));
    for (int k = 0; k < n; ++k) {
        const double xi = sp[k].x, eta = sp[k].y;
        pts[k] = Vec2{p1.x + (p2.x - p1.x) * xi + (p3.x - p1.x) * eta, p1.y + (p2.y - p1.y) * xi + (p3.y - p1.y) * eta};
        w[k] = sw[k] * jacobi / faceArea;
    }
}

// The same for a quadrilateral, bilinear map.
static void quad_points(const Mesh& m, int cell, const Vec2* sp, const double* sw, int n, Vec2* pts, double* w) {
    const int b = m.cell_ptr[cell];
    const double faceArea = m.V[cell] / m.depth;
    // sortRecPoints (d2GaussQuadrature.C:283-331): polar angle about the
    // vertex mean, largest first, bubble sort
    Vec2 pt[4];
    double ref_x = 0.0, ref_y = 0.0;
    for (int i = 0; i < 4; ++i) {
        pt[i] = m.cell_xy[b + i];
        ref_x += pt[i].x;
        ref_y += pt[i].y;
    }
    ref_x /= 4.0;
    ref_y /= 4.0;
    double polar[4];
    int idx[4] = {0, 1, 2, 3};
    for (int i = 0; i < 4; ++i) {
        const double x = pt[i].x - ref_x, y = pt[i].y - ref_y;
        const double r = sqrt(x * x + y * y);
        polar[i] = y >= 0 ? acos(x / r) : -acos(x / r);
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3 - i; ++j)
            if (polar[j] < polar[j + 1]) {
                std::swap(polar[j], polar[j + 1]);
                std::swap(idx[j], idx[j + 1]);
            }
    const double x1 = pt[idx[0]].x, x2 = pt[idx[1]].x, x3 = pt[idx[2]].x, x4 = pt[idx[3]].x;
    const double y1 = pt[idx[0]].y, y2 = pt[idx[1]].y, y3 = pt[idx[2]].y, y4 = pt[idx[3]].y;
    for (int k = 0; k < n; ++k) {
        const double xi = sp[k].x, eta = sp[k].y;
        const double b1 = 0.25 * (1.0 - xi) * (1.0 - eta), b2 = 0.25 * (1.0 - xi) * (1.0 + eta);
        const double b3 = 0.25 * (1.0 + xi) * (1.0 + eta), b4 = 0.25 * (1.0 + xi) * (1.0 - eta);
        pts[k] = Vec2{x1 * b1 + x2 * b2 + x3 * b3 + x4 * b4, y1 * b1 + y2 * b2 + y3 * b3 + y4 * b4};
        const double dxdxi = -0.25 * x1 * (1.0 - eta) - 0.25 * x2 * (1.0 + eta) + 0.25 * x3 * (1.0 + eta) + 0.25 * x4 * (1.0 - eta);
        const double dxdeta = -0.25 * x1 * (1.0 - xi) + 0.25 * x2 * (1.0 - xi) + 0.25 * x3 * (1.0 + xi) - 0.25 * x4 * (1.0 + xi);
        const double dydxi = -0.25 * y1 * (1.0 - eta) - 0.25 * y2 * (1.0 + eta) + 0.25 * y3 * (1.0 + eta) + 0.25 * y4 * (1.0 - eta);
        const double dydeta = -0.25 * y1 * (1.0 - xi) + 0.25 * y2 * (1.0 - xi) + 0.25 * y3 * (1.0 + xi) - 0.25 * y4 * (1.0 + xi);
        const double jacobi = fabs(dxdxi * dydeta - dxdeta * dydxi);
        w[k] = sw[k] * jacobi / faceArea;
    }
}

static int cell_vertices(const Mesh& m, int cell) {
    const int nv = m.cell_ptr[cell + 1] - m.cell_ptr[cell];
    if (nv != 3 && nv != 4) throw Error(1, "problem cell appear with 2D face points number = " + std::to_string(nv));
    return nv;
}

// Cell Gauss points in the cell's own frame, weights divided by the cell's 2D
// area so that sum(w f) is the cell average (d2GaussQuadrature.C:23-150).
void gauss_cell(const Mesh& m, int cell, int cOrder, std::vector<Vec2>& pts, std::vector<double>& w) {
    std::vector<Vec2> sp;
    std::vector<double> sw;
    const int nv = cell_vertices(m, cell);
    if (nv == 3)
        tri_rule(cOrder, sp, sw);
    else
        rec_rule(cOrder, sp, sw);
    pts.resize(sp.size());
    w.resize(sp.size());
    if (nv == 3)
        tri_points(m, cell, sp.data(), sw.data(), (int)sp.size(), pts.data(), w.data());
    else
        quad_points(m, cell, sp.data(), sw.data(), (int)sp.size(), pts.data(), w.data());
}

void face_gauss_rule(int fOrder, std::vector<double>& xi, std::vector<double>& w) { bar_rule(fOrder, xi, w); }

Vec2 face_gauss_point(const Mesh& m, int f, double xi) {
    const Vec2 p0 = m.fp0[f], p1 = m.fp1[f];
    return Vec2{0.5 * p0.x * (1 - xi) + 0.5 * p1.x * (1 + xi), 0.5 * p0.y * (1 - xi) + 0.5 * p1.y * (1 + xi)};
}

// Boundary-face Gauss points; weights include the face height and are then
// divided by |S_f| so that they average over the face
// (d2GaussQuadrature.C:217-279 with basicCUDAsolver.cpp:446-453).
static void bface_points(const Mesh& m, int b, const double* sp, const double* sw, int n, Vec2* pts, double* w) {
    const int f = m.nf + b;
    const Vec2 p0 = m.fp0[f], p1 = m.fp1[f];
    const double jacobi = sqrt((0.5 * (p0.x - p1.x)) * (0.5 * (p0.x - p1.x)) + (0.5 * (p0.y - p1.y)) * (0.5 * (p0.y - p1.y)));
    for (int k = 0; k < n; ++k) {
        const double xi = sp[k];
        pts[k] = Vec2{0.5 * p0.x * (1 - xi) + 0.5 * p1.x * (1 + xi), 0.5 * p0.y * (1 - xi) + 0.5 * p1.y * (1 + xi)};
        w[k] = (sw[k] * jacobi * m.depth) / m.magSf[f];
    }
}

void gauss_boundary_face(const Mesh& m, int b, int fOrder, std::vector<Vec2>& pts, std::vector<double>& w) {
    std::vector<double> sp, sw;
    bar_rule(fOrder, sp, sw);
    pts.resize(sp.size());
    w.resize(sp.size());
    bface_points(m, b, sp.data(), sw.data(), (int)sp.size(), pts.data(), w.data());
}

// ------------------------------------------------------------------ pseudo-inverse

void pinv_jacobi(int n, const double* A, double* G) {
    if (n < 1 || n > kRbfMaxN) throw Error(1, "pinv_jacobi: matrix size out of range");
    std::vector<double> U((size_t)n * n), V((size_t)n * n), sig(n);
    for (int i = 0; i < n * n; ++i) U[i] = A[i];
    rbf_pinv(n, U.data(), V.data(), sig.data(), G);
}

// ------------------------------------------------------------------ RBF precompute

// Quadrature of every cell (cOrder rule) and boundary face (fOrder rule) as
// flat tables: points 2 per entry, weights, and nc+nb+1 offsets.
void gauss_tables(const Mesh& m, const Schemes& sc, std::vector<double>& pts, std::vector<double>& w,
                  std::vector<int64_t>& ptr) {
    const int nc = m.nc, nb = m.nb;
    // the unit rules once; a quadrilateral rule is only asked for when the mesh has one
    std::vector<Vec2> tp, qp;
    std::vector<double> tw, qw, bp, bw;
    bool tri = false, quad = false;
    ptr.assign(nc + nb + 1, 0);
    for (int c = 0; c < nc; ++c) (cell_vertices(m, c) == 3 ? tri : quad) = true;
    if (tri) tri_rule(sc.cOrder, tp, tw);
    if (quad) rec_rule(sc.cOrder, qp, qw);
    if (nb) bar_rule(sc.fOrder, bp, bw);
    for (int i = 0; i < nc + nb; ++i) {
        const int64_t n = i < nc ? (int64_t)(m.cell_ptr[i + 1] - m.cell_ptr[i] == 3 ? tp.size() : qp.size())
                                 : (int64_t)bp.size();
        ptr[i + 1] = ptr[i] + n;
    }
    std::vector<Vec2> P((size_t)ptr[nc + nb]);
    w.resize((size_t)ptr[nc + nb]);
    parallel_for(nc + nb, [&](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) {
            Vec2* p = P.data() + ptr[i];
            double* wi = w.data() + ptr[i];
            if (i >= nc)
                bface_points(m, (int)(i - nc), bp.data(), bw.data(), (int)bp.size(), p, wi);
            else if (m.cell_ptr[i + 1] - m.cell_ptr[i] == 3)
                tri_points(m, (int)i, tp.data(), tw.data(), (int)tp.size(), p, wi);
            else
                quad_points(m, (int)i, qp.data(), qw.data(), (int)qp.size(), p, wi);
        }
    });
    pts.resize(2 * P.size());
    for (size_t k = 0; k < P.size(); ++k) {
        pts[2 * k] = P[k].x;
        pts[2 * k + 1] = P[k].y;
    }
}

Precomputed precompute(const Mesh& m, const Schemes& sc) {
    Precomputed pc;
    pc.st = build_stencils(m, sc.stencilType);
    const Stencils& st = pc.st;
    const int nc = m.nc, nf = m.nf, maxS = st.maxS;
    const int W = maxS + 1;
    pc.rbf.assign((size_t)2 * W * nf, 0.0);
    pc.trouble.assign(nc, 0);
    pc.refDistance.assign(nc, 0.0);
    if (sc.faceQuadrature) {
        std::vector<double> sw;
        face_gauss_rule(sc.fOrder, pc.gp_xi, sw);
        pc.ngp = (int)pc.gp_xi.size();
        pc.gp_w.resize(pc.ngp);
        for (int g = 0; g < pc.ngp; ++g) pc.gp_w[g] = 0.5 * sw[g];  // the bar rule's weights sum to 2
        pc.rbf_gp.assign((size_t)pc.ngp * 2 * W * nf, 0.0);
    }
    if (!sc.recalcBasis) return pc;  // coefficients come from read_basis_lists
    const double sf = sc.shapeFactor;
    std::vector<double> tp, tw;
    std::vector<int64_t> tptr;
    gauss_tables(m, sc, tp, tw, tptr);
    std::atomic<int> troubles{0};
    parallel_for(nc, [&](int64_t c0, int64_t c1) {
        double X[2 * kRbfMaxN], U[kRbfMaxN * kRbfMaxN], V[kRbfMaxN * kRbfMaxN], G[kRbfMaxN * kRbfMaxN];
        double sig[kRbfMaxN], b[kRbfMaxN], a[kRbfMaxN];
        int gptr[kRbfMaxN + 1];
        std::vector<double> gp, gw;
        for (int64_t c = c0; c < c1; ++c) {
            const int S = st.S[c], n = S + 1;
            rbf_cell_inputs(m, st, tp.data(), tw.data(), tptr.data(), (int)c, X, gp, gw, gptr);
            double ref = 0.0;
            for (int i = 1; i < n; ++i) {
                const double refdx = X[2 * i] - m.C[c].x, refdy = X[2 * i + 1] - m.C[c].y;
                ref += sqrt(refdx * refdx + refdy * refdy);
            }
            ref /= (double)S;
            pc.refDistance[c] = ref;
            int count[kRbfMaxN];
            for (int i = 0; i < n; ++i) count[i] = gptr[i + 1] - gptr[i];
            rbf_moments(n, X, count,
                        [&](int i, int k, double& x, double& y, double& w) {
                            x = gp[2 * (gptr[i] + k)];
                            y = gp[2 * (gptr[i] + k) + 1];
                            w = gw[gptr[i] + k];
                        },
                        ref, sf, U);
            rbf_pinv(n, U, V, sig, G);
            const bool trouble = rbf_trouble(n, G);
            pc.trouble[c] = trouble;
            if (trouble) troubles++;
            // coefficients of every internal face of c on c's side
            for (int fi = 0; fi < st.Cc[c]; ++fi) {
                const int sf_ = st.cellFaces[(size_t)c * kMaxValidStencil + fi];
                const int f = (sf_ > 0 ? sf_ : -sf_) - 1;
                if (f >= nf) continue;
                const int side = sf_ < 0;
                double* out = &pc.rbf[(size_t)2 * W * f + (size_t)side * W];
                if (trouble) {
                    out[0] = 1.0;
                    for (int i = 1; i < n; ++i) out[i] = 0.0;
                    for (int g = 0; g < pc.ngp; ++g) {
                        double* og = &pc.rbf_gp[((size_t)g * nf + f) * 2 * W + (size_t)side * W];
                        og[0] = 1.0;
                        for (int i = 1; i < n; ++i) og[i] = 0.0;
                    }
                    continue;
                }
                // face centre in c's frame
                Vec2 xf = m.Cf[f];
                if (side) xf = Vec2{xf.x - m.fshift[f].x, xf.y - m.fshift[f].y};
                rbf_face(n, G, X, xf.x, xf.y, ref, sf, b, a);
                for (int i = 0; i < n; ++i) out[i] = a[i];
                // flux quadrature points in c's frame (extension)
                for (int g = 0; g < pc.ngp; ++g) {
                    Vec2 xg = face_gauss_point(m, f, pc.gp_xi[g]);
                    if (side) xg = Vec2{xg.x - m.fshift[f].x, xg.y - m.fshift[f].y};
                    rbf_face(n, G, X, xg.x, xg.y, ref, sf, b, a);
                    double* og = &pc.rbf_gp[((size_t)g * nf + f) * 2 * W + (size_t)side * W];
                    for (int i = 0; i < n; ++i) og[i] = a[i];
                }
            }
        }
    });
    pc.troubleCount = troubles;
    return pc;
}

// Member centres (in the cell's frame) and member quadrature of cell c, from
// the flat quadrature tables (the shift of periodic members applied).
void rbf_cell_inputs(const Mesh& m, const Stencils& st, const double* tp, const double* tw, const int64_t* tptr, int c,
                     double* X, std::vector<double>& gp, std::vector<double>& gw, int* gptr) {
    const int S = st.S[c], n = S + 1;
    gp.clear();
    gw.clear();
    gptr[0] = 0;
    for (int i = 0; i < n; ++i) {
        int id = c;
        Vec2 sh{0.0, 0.0};
        if (i > 0) {
            id = st.sten[(size_t)c * kMaxValidStencil + i - 1];
            sh = st.sten_shift[(size_t)c * kMaxValidStencil + i - 1];
            X[2 * i] = m.C[id].x + sh.x;
            X[2 * i + 1] = m.C[id].y + sh.y;
        } else {
            X[0] = m.C[c].x;
            X[1] = m.C[c].y;
        }
        for (int64_t k = tptr[id]; k < tptr[id + 1]; ++k) {
            gp.push_back(i > 0 ? tp[2 * k] + sh.x : tp[2 * k]);
            gp.push_back(i > 0 ? tp[2 * k + 1] + sh.y : tp[2 * k + 1]);
            gw.push_back(tw[k]);
        }
        gptr[i + 1] = (int)gw.size();
    }
}

// ------------------------------------------------------------------ basis cache I/O

// system/basisOwnList and system/basisNeiList: OpenFOAM IOList<List<scalar>>
// ASCII files, one list of stencil-size+1 coefficients per internal face
// (basicCUDAsolver.cpp:626-694).  Written with 17 significant digits so the
// FP64 path reads back exactly what it computed.
void write_basis_lists(const Mesh& m, const Precomputed& pc, const std::string& dir) {
    const int nf = m.nf, W = pc.st.maxS + 1;
    for (int side = 0; side < 2; ++side) {
        const char* name = side ? "basisNeiList" : "basisOwnList";
        std::ofstream os(dir + "/" + name);
        if (!os) throw Error(1, "cannot write " + dir + "/" + name);
        os << "FoamFile\n{\n    version     2.0;\n    format      ascii;\n    class       scalarListList;\n"
           << "    location    \"system\";\n    object      " << name << ";\n}\n\n";
        os << nf << "\n(\n";
        char buf[40];
        for (int f = 0; f < nf; ++f) {
            const int cell = side ? m.nei[f] : m.own[f];
            const int n = pc.st.S[cell] + 1;
            os << n << '(';
            for (int i = 0; i < n; ++i) {
                snprintf(buf, sizeof buf, "%.17g", pc.rbf[(size_t)2 * W * f + (size_t)side * W + i]);
                os << (i ? " " : "") << buf;
            }
            os << ")\n";
        }
        os << ")\n";
    }
}

namespace {

std::string strip_foam(const std::string& text) {
    std::string s;
    s.reserve(text.size());
    for (size_t i = 0; i < text.size();) {
        if (text.compare(i, 2, "//") == 0) {
            while (i < text.size() && text[i] != '\n') ++i;
        } else if (text.compare(i, 2, "/*") == 0) {
            const size_t e = text.find("*/", i + 2);
            i = (e == std::string::npos) ? text.size() : e + 2;
        } else {
            s.push_back(text[i++]);
        }
    }
    const size_t h = s.find("FoamFile");
    if (h != std::string::npos) {
        const size_t e = s.find('}', h);
        if (e != std::string::npos) s = s.substr(e + 1);
    }
    return s;
}

}  // namespace

void read_basis_lists(const Mesh& m, Precomputed& pc, const std::string& dir) {
    const int nf = m.nf, W = pc.st.maxS + 1;
    for (int side = 0; side < 2; ++side) {
        const char* name = side ? "basisNeiList" : "basisOwnList";
        std::ifstream is(dir + "/" + name);
        if (!is) throw Error(1, "cannot read " + dir + "/" + name);
        std::stringstream ss;
        ss << is.rdbuf();
        std::string s = strip_foam(ss.str());
        size_t pos = 0;
        auto skip_ws = [&]() {
            while (pos < s.size() && isspace((unsigned char)s[pos])) ++pos;
        };
        auto read_num = [&]() {
            skip_ws();
            const char* b = s.c_str() + pos;
            char* e = nullptr;
            const double v = strtod(b, &e);
            if (e == b) throw Error(1, std::string("parse error in ") + name);
            pos += (size_t)(e - b);
            return v;
        };
        auto expect = [&](char ch) {
            skip_ws();
            if (pos >= s.size() || s[pos] != ch) throw Error(1, std::string("parse error in ") + name);
            ++pos;
        };
        const int n_list = (int)read_num();
        if (n_list != nf) throw Error(1, std::string(name) + " has a different number of faces");
        expect('(');
        for (int f = 0; f < nf; ++f) {
            const int cell = side ? m.nei[f] : m.own[f];
            const int want = pc.st.S[cell] + 1;
            const int n = (int)read_num();
            if (n != want) throw Error(1, std::string(name) + " stencil size does not match the mesh");
            double* out = &pc.rbf[(size_t)2 * W * f + (size_t)side * W];
            skip_ws();
            if (pos < s.size() && s[pos] == '{') {  // uniform list n{v}
                ++pos;
                const double v = read_num();
                expect('}');
                for (int i = 0; i < n; ++i) out[i] = v;
            } else {
                expect('(');
                for (int i = 0; i < n; ++i) out[i] = read_num();
                expect(')');
            }
        }
        expect(')');
    }
}

}  // namespace afar
