// Mesh construction: synthetic generators for the benchmark cases, the
// finite-volume topology (faces, owner/neighbour, patches), reverse
// Cuthill-McKee renumbering and the geometry of cells and faces.
//
// The reference receives all of this from OpenFOAM (mesh.C(), mesh.V(),
// mesh.Cf(), mesh.Sf(), mesh.magSf(), owner/neighbour; consumed at
// src/CUDA/cudaEuler2DmixO3RK3/basicCUDAsolver.cpp:122-258,269-299) and from
// the offline `renumberMesh -overwrite` step (README.md:76).  Here the same
// quantities are computed for 2D polygon meshes extruded by `depth`:
// polygon area centroid, edge mid-point face centres and S_f = outward normal
// times edge length times depth.
#include <math.h>

#include <algorithm>
#include <numeric>
#include <unordered_map>

#include "afar_host.h"

namespace afar {

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double sym(double a) { return (2.0 * uniform() - 1.0) * a; }                         // [-a,a)
};

inline uint64_t edge_key(int32_t a, int32_t b) {
    const uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
    return ((uint64_t)lo << 32) | hi;
}

}  // namespace

// ---------------------------------------------------------------- generators

PolyMesh generate_corner(int nx, int ny, double lx, double ly, double jitter, uint64_t seed,
                         double corner_x, double angle_deg, const double inflow[4], int top_label) {
    if (nx < 2 || ny < 2) throw Error(1, "corner mesh needs nx, ny >= 2");
    PolyMesh pm;
    const int px = nx + 1;
    std::vector<double> xi((size_t)px * (ny + 1)), eta(xi.size());
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            xi[(size_t)j * px + i] = (double)i;
            eta[(size_t)j * px + i] = (double)j;
        }
    SplitMix64 rng(seed);
    for (int j = 1; j < ny; ++j)
        for (int i = 1; i < nx; ++i) {
            xi[(size_t)j * px + i] += rng.sym(jitter);
            eta[(size_t)j * px + i] += rng.sym(jitter);
        }
    const double tana = tan(angle_deg * 3.14159265358979323846 / 180.0);
    pm.pts.resize(xi.size());
    for (size_t p = 0; p < xi.size(); ++p) {
        const double x = lx * xi[p] / nx;
        const double yw = (x <= corner_x) ? 0.0 : -(x - corner_x) * tana;
        pm.pts[p].x = x;
        pm.pts[p].y = yw + (ly - yw) * (eta[p] / ny);
    }
    auto P = [&](int i, int j) { return (int32_t)(j * px + i); };
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            pm.cell_pts.insert(pm.cell_pts.end(), {P(i, j), P(i + 1, j), P(i + 1, j + 1), P(i, j + 1)});
            pm.cell_ptr.push_back((int64_t)pm.cell_pts.size());
        }
    pm.patches.resize(4);
    pm.patches[0].name = "inlet";
    pm.patches[0].label = 6;
    for (int k = 0; k < 4; ++k) pm.patches[0].value[k] = inflow[k];
    pm.patches[1].name = "outlet";
    pm.patches[1].label = 4;
    pm.patches[2].name = "lowerWall";
    pm.patches[2].label = 2;
    pm.patches[3].name = "upperWall";
    pm.patches[3].label = top_label;
    for (int j = 0; j < ny; ++j) {
        pm.bedges.push_back({P(0, j), P(0, j + 1)});
        pm.bedge_patch.push_back(0);
        pm.bedges.push_back({P(nx, j), P(nx, j + 1)});
        pm.bedge_patch.push_back(1);
    }
    for (int i = 0; i < nx; ++i) {
        pm.bedges.push_back({P(i, 0), P(i + 1, 0)});
        pm.bedge_patch.push_back(2);
        pm.bedges.push_back({P(i, ny), P(i + 1, ny)});
        pm.bedge_patch.push_back(3);
    }
    return pm;
}

PolyMesh generate_riemann(int n, double jitter, uint64_t seed, double split_p) {
    if (n < 2) throw Error(1, "riemann mesh needs n >= 2");
    PolyMesh pm;
    const int px = n + 1;
    pm.pts.resize((size_t)px * px);
    SplitMix64 rng(seed);
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
            double a = i, b = j;
            if (i > 0 && i < n && j > 0 && j < n) {
                a += rng.sym(jitter);
                b += rng.sym(jitter);
            }
            pm.pts[(size_t)j * px + i] = Vec2{a / n, b / n};
        }
    auto P = [&](int i, int j) { return (int32_t)(j * px + i); };
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const bool split = rng.uniform() < split_p;
            const bool diag = rng.uniform() < 0.5;
            const int32_t a = P(i, j), b = P(i + 1, j), c = P(i + 1, j + 1), d = P(i, j + 1);
            if (!split) {
                pm.cell_pts.insert(pm.cell_pts.end(), {a, b, c, d});
                pm.cell_ptr.push_back((int64_t)pm.cell_pts.size());
            } else if (!diag) {  // diagonal a-c
                pm.cell_pts.insert(pm.cell_pts.end(), {a, b, c});
                pm.cell_ptr.push_back((int64_t)pm.cell_pts.size());
                pm.cell_pts.insert(pm.cell_pts.end(), {a, c, d});
                pm.cell_ptr.push_back((int64_t)pm.cell_pts.size());
            } else {             // diagonal b-d
                pm.cell_pts.insert(pm.cell_pts.end(), {a, b, d});
                pm.cell_ptr.push_back((int64_t)pm.cell_pts.size());
                pm.cell_pts.insert(pm.cell_pts.end(), {b, c, d});
                pm.cell_ptr.push_back((int64_t)pm.cell_pts.size());
            }
        }
    const char* names[4] = {"left", "right", "bottom", "top"};
    pm.patches.resize(4);
    for (int k = 0; k < 4; ++k) {
        pm.patches[k].name = names[k];
        pm.patches[k].label = 4;
    }
    for (int j = 0; j < n; ++j) {
        pm.bedges.push_back({P(0, j), P(0, j + 1)});
        pm.bedge_patch.push_back(0);
        pm.bedges.push_back({P(n, j), P(n, j + 1)});
        pm.bedge_patch.push_back(1);
    }
    for (int i = 0; i < n; ++i) {
        pm.bedges.push_back({P(i, 0), P(i + 1, 0)});
        pm.bedge_patch.push_back(2);
        pm.bedges.push_back({P(i, n), P(i + 1, n)});
        pm.bedge_patch.push_back(3);
    }
    return pm;
}

PolyMesh generate_periodic(int n, double L, double jitter, uint64_t seed, double split_p) {
    if (n < 3) throw Error(1, "periodic mesh needs n >= 3");
    if (!(split_p >= 0.0 && split_p <= 1.0)) throw Error(1, "split_p must lie in [0, 1]");
    PolyMesh pm;
    pm.period = Vec2{L, L};
    pm.pts.resize((size_t)n * n);
    const double h = L / n;
    SplitMix64 rng(seed);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const double a = i + rng.sym(jitter), b = j + rng.sym(jitter);
            pm.pts[(size_t)j * n + i] = Vec2{a * h, b * h};
        }
    auto add = [&](std::initializer_list<std::array<int, 2>> v) {
        for (auto ij : v) {
            const int i = ij[0], j = ij[1];
            pm.cell_pts.push_back((int32_t)((j % n) * n + (i % n)));
            pm.cell_vshift.push_back((int8_t)(i / n));
            pm.cell_vshift.push_back((int8_t)(j / n));
        }
        pm.cell_ptr.push_back((int64_t)pm.cell_pts.size());
    };
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            // the split draw is skipped when every quad is split, so the
            // all-triangle mesh of a seed does not depend on split_p
            const bool split = split_p >= 1.0 || rng.uniform() < split_p;
            if (!split) {
                add({{i, j}, {i + 1, j}, {i + 1, j + 1}, {i, j + 1}});
                continue;
            }
            const bool diag = rng.uniform() < 0.5;
            if (!diag) {
                add({{i, j}, {i + 1, j}, {i + 1, j + 1}});
                add({{i, j}, {i + 1, j + 1}, {i, j + 1}});
            } else {
                add({{i, j}, {i + 1, j}, {i, j + 1}});
                add({{i + 1, j}, {i + 1, j + 1}, {i, j + 1}});
            }
        }
    return pm;
}

// ---------------------------------------------------------------- RCM

// Reverse Cuthill-McKee (E. Cuthill and J. McKee, 1969; A. George, 1971) with a
// George-Liu pseudo-peripheral start node per connected component.  Ties are
// broken by (degree, index) so the order is a pure function of the graph.
std::vector<int32_t> rcm_order(int n, const std::vector<int32_t>& adj_ptr, const std::vector<int32_t>& adj) {
    std::vector<int32_t> deg(n);
    for (int i = 0; i < n; ++i) deg[i] = adj_ptr[i + 1] - adj_ptr[i];
    std::vector<char> done(n, 0);
    std::vector<int32_t> order;
    order.reserve(n);
    std::vector<int32_t> level(n, -1), queue, touched;
    // BFS from s over the unvisited component; returns the nodes of the last level.
    auto bfs_levels = [&](int s, int& ecc) {
        for (int v : touched) level[v] = -1;
        touched.clear();
        queue.clear();
        queue.push_back(s);
        level[s] = 0;
        touched.push_back(s);
        for (size_t h = 0; h < queue.size(); ++h) {
            const int v = queue[h];
            for (int k = adj_ptr[v]; k < adj_ptr[v + 1]; ++k) {
                const int w = adj[k];
                if (done[w] || level[w] >= 0) continue;
                level[w] = level[v] + 1;
                touched.push_back(w);
                queue.push_back(w);
            }
        }
        ecc = level[queue.back()];
        std::vector<int32_t> last;
        for (size_t h = queue.size(); h-- > 0;) {
            if (level[queue[h]] != ecc) break;
            last.push_back(queue[h]);
        }
        return last;
    };
    int next_unvisited = 0;
    std::vector<int32_t> nb;
    while ((int)order.size() < n) {
        // lowest-degree unvisited node (lowest index among ties)
        while (done[next_unvisited]) ++next_unvisited;
        int s = -1;
        {
            // the component of next_unvisited: pick its min-degree node
            int ecc0;
            bfs_levels(next_unvisited, ecc0);
            for (int v : queue)
                if (s < 0 || deg[v] < deg[s] || (deg[v] == deg[s] && v < s)) s = v;
        }
        int ecc;
        auto last = bfs_levels(s, ecc);
        for (int it = 0; it < 32; ++it) {
            int best = -1;
            for (int v : last)
                if (best < 0 || deg[v] < deg[best] || (deg[v] == deg[best] && v < best)) best = v;
            int ecc2;
            auto last2 = bfs_levels(best, ecc2);
            if (ecc2 <= ecc) break;
            s = best;
            ecc = ecc2;
            last = std::move(last2);
        }
        // Cuthill-McKee BFS from s
        const size_t base = order.size();
        order.push_back(s);
        done[s] = 1;
        for (size_t h = base; h < order.size(); ++h) {
            const int v = order[h];
            nb.clear();
            for (int k = adj_ptr[v]; k < adj_ptr[v + 1]; ++k)
                if (!done[adj[k]]) nb.push_back(adj[k]);
            std::sort(nb.begin(), nb.end(), [&](int a, int b) {
                return deg[a] != deg[b] ? deg[a] < deg[b] : a < b;
            });
            nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
            for (int w : nb) {
                done[w] = 1;
                order.push_back(w);
            }
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// ---------------------------------------------------------------- topology

Mesh build_mesh(const PolyMesh& pm, bool renumber) {
    const int ncell = (int)pm.cell_ptr.size() - 1;
    if (ncell <= 0) throw Error(1, "mesh has no cells");
    if (pm.bedges.size() != pm.bedge_patch.size()) throw Error(1, "bedges/bedge_patch size mismatch");
    const bool periodic = !pm.cell_vshift.empty();
    if (periodic && pm.cell_vshift.size() != 2 * pm.cell_pts.size()) throw Error(1, "cell_vshift size mismatch");
    const int npts = (int)pm.pts.size();
    for (int32_t p : pm.cell_pts)
        if (p < 0 || p >= npts) throw Error(1, "cell point index out of range");

    auto vshift = [&](int64_t k, int d) -> int { return periodic ? pm.cell_vshift[2 * k + d] : 0; };
    auto vcoord = [&](int64_t k) {
        const Vec2& p = pm.pts[pm.cell_pts[k]];
        return Vec2{p.x + vshift(k, 0) * pm.period.x, p.y + vshift(k, 1) * pm.period.y};
    };

    // half-edges (key, cell, local edge) sorted by key
    struct HE {
        uint64_t key;
        int32_t cell, k;
    };
    std::vector<HE> he;
    he.reserve(pm.cell_pts.size());
    for (int c = 0; c < ncell; ++c) {
        const int64_t b = pm.cell_ptr[c], e = pm.cell_ptr[c + 1];
        if (e - b < 3) throw Error(1, "cell with fewer than 3 vertices");
        for (int64_t k = b; k < e; ++k) {
            const int64_t k1 = (k + 1 < e) ? k + 1 : b;
            he.push_back(HE{edge_key(pm.cell_pts[k], pm.cell_pts[k1]), c, (int32_t)(k - b)});
        }
    }
    std::sort(he.begin(), he.end(), [](const HE& a, const HE& b) {
        return a.key != b.key ? a.key < b.key : (a.cell != b.cell ? a.cell < b.cell : a.k < b.k);
    });
    std::vector<std::pair<uint64_t, int32_t>> bkeys(pm.bedges.size());
    for (size_t i = 0; i < pm.bedges.size(); ++i) bkeys[i] = {edge_key(pm.bedges[i][0], pm.bedges[i][1]), (int32_t)i};
    std::sort(bkeys.begin(), bkeys.end());

    struct IF {
        int32_t a, ka, b, kb;
    };
    struct BF {
        int32_t c, k, patch;
    };
    std::vector<IF> ifaces;
    std::vector<BF> bfaces;
    for (size_t h = 0; h < he.size();) {
        size_t e = h + 1;
        while (e < he.size() && he[e].key == he[h].key) ++e;
        if (e - h == 2) {
            if (he[h].cell == he[h + 1].cell) throw Error(1, "cell shares an edge with itself");
            ifaces.push_back(IF{he[h].cell, he[h].k, he[h + 1].cell, he[h + 1].k});
        } else if (e - h == 1) {
            auto it = std::lower_bound(bkeys.begin(), bkeys.end(), std::make_pair(he[h].key, (int32_t)-1));
            if (it == bkeys.end() || it->first != he[h].key)
                throw Error(1, "boundary edge without a patch");
            const int patch = pm.bedge_patch[it->second];
            if (patch < 0 || patch >= (int)pm.patches.size()) throw Error(1, "bad patch index");
            bfaces.push_back(BF{he[h].cell, he[h].k, patch});
        } else {
            throw Error(1, "non-manifold edge shared by more than two cells");
        }
        h = e;
    }
    for (const auto& p : pm.patches)
        if (p.label < 0 || p.label > kMaxBCLabel) throw Error(1, "boundary label out of range 0..6");

    // renumbering
    std::vector<int32_t> perm(ncell);
    if (renumber) {
        std::vector<int32_t> cnt(ncell + 1, 0), adj(2 * ifaces.size());
        for (auto& f : ifaces) {
            cnt[f.a + 1]++;
            cnt[f.b + 1]++;
        }
        std::partial_sum(cnt.begin(), cnt.end(), cnt.begin());
        std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
        for (auto& f : ifaces) {
            adj[fill[f.a]++] = f.b;
            adj[fill[f.b]++] = f.a;
        }
        for (int c = 0; c < ncell; ++c) std::sort(adj.begin() + cnt[c], adj.begin() + cnt[c + 1]);
        perm = rcm_order(ncell, cnt, adj);
    } else {
        std::iota(perm.begin(), perm.end(), 0);
    }
    Mesh m;
    m.nc = ncell;
    m.depth = pm.depth;
    m.perm = perm;
    m.iperm.assign(ncell, 0);
    for (int i = 0; i < ncell; ++i) m.iperm[perm[i]] = i;
    m.npts = npts;
    m.period = pm.period;
    m.patches = pm.patches;

    // cells in new order (coordinates in each cell's frame)
    m.cell_ptr.assign(ncell + 1, 0);
    for (int i = 0; i < ncell; ++i) {
        const int c = perm[i];
        m.cell_ptr[i + 1] = m.cell_ptr[i] + (int32_t)(pm.cell_ptr[c + 1] - pm.cell_ptr[c]);
    }
    m.cell_pts.resize(m.cell_ptr[ncell]);
    m.cell_xy.resize(m.cell_ptr[ncell]);
    m.cell_vs.assign(2 * (size_t)m.cell_ptr[ncell], 0);
    for (int i = 0; i < ncell; ++i) {
        const int c = perm[i];
        for (int64_t k = pm.cell_ptr[c]; k < pm.cell_ptr[c + 1]; ++k) {
            const int32_t dst = m.cell_ptr[i] + (int32_t)(k - pm.cell_ptr[c]);
            m.cell_pts[dst] = pm.cell_pts[k];
            m.cell_xy[dst] = vcoord(k);
            m.cell_vs[2 * (size_t)dst] = (int8_t)vshift(k, 0);
            m.cell_vs[2 * (size_t)dst + 1] = (int8_t)vshift(k, 1);
        }
    }
    // per cell vertex shift (integers) in new order, for face shifts
    auto vs_new = [&](int cell_new, int k, int d) -> int {
        return vshift(pm.cell_ptr[perm[cell_new]] + k, d);
    };

    // cell geometry: fan triangulation about the first vertex
    m.V.assign(ncell, 0.0);
    m.C.assign(ncell, Vec2{});
    for (int i = 0; i < ncell; ++i) {
        const Vec2* v = &m.cell_xy[m.cell_ptr[i]];
        const int nv = m.cell_ptr[i + 1] - m.cell_ptr[i];
        double a2 = 0.0, cx = 0.0, cy = 0.0;
        for (int k = 1; k + 1 < nv; ++k) {
            const double ux = v[k].x - v[0].x, uy = v[k].y - v[0].y;
            const double wx = v[k + 1].x - v[0].x, wy = v[k + 1].y - v[0].y;
            const double t = ux * wy - uy * wx;
            a2 += t;
            cx += t * (ux + wx);
            cy += t * (uy + wy);
        }
        if (!(a2 > 0.0)) throw Error(1, "cell with non-positive area (polygons must be counter-clockwise)");
        m.C[i] = Vec2{v[0].x + cx / (3.0 * a2), v[0].y + cy / (3.0 * a2)};
        m.V[i] = 0.5 * a2 * m.depth;
    }

    // internal faces, upper-triangular order
    struct NF {
        int32_t own, nei, kown, knei;
    };
    std::vector<NF> nfv(ifaces.size());
    for (size_t f = 0; f < ifaces.size(); ++f) {
        int32_t a = m.iperm[ifaces[f].a], b = m.iperm[ifaces[f].b];
        int32_t ka = ifaces[f].ka, kb = ifaces[f].kb;
        if (a > b) {
            std::swap(a, b);
            std::swap(ka, kb);
        }
        nfv[f] = NF{a, b, ka, kb};
    }
    std::sort(nfv.begin(), nfv.end(), [](const NF& x, const NF& y) {
        return x.own != y.own ? x.own < y.own : (x.nei != y.nei ? x.nei < y.nei : x.kown < y.kown);
    });
    for (size_t f = 1; f < nfv.size(); ++f)
        if (nfv[f].own == nfv[f - 1].own && nfv[f].nei == nfv[f - 1].nei)
            throw Error(1, "two cells share more than one face");
    m.nf = (int)nfv.size();

    // boundary faces grouped by patch, then by cell, then by local edge
    std::vector<BF> bnew(bfaces.size());
    for (size_t b = 0; b < bfaces.size(); ++b) bnew[b] = BF{m.iperm[bfaces[b].c], bfaces[b].k, bfaces[b].patch};
    std::sort(bnew.begin(), bnew.end(), [](const BF& x, const BF& y) {
        return x.patch != y.patch ? x.patch < y.patch : (x.c != y.c ? x.c < y.c : x.k < y.k);
    });
    m.nb = (int)bnew.size();
    m.patch_start.assign(pm.patches.size(), 0);
    m.patch_size.assign(pm.patches.size(), 0);
    for (auto& b : bnew) m.patch_size[b.patch]++;
    for (size_t p = 1; p < pm.patches.size(); ++p) m.patch_start[p] = m.patch_start[p - 1] + m.patch_size[p - 1];

    const int nft = m.nf + m.nb;
    m.own.assign(nft, 0);
    m.nei.assign(m.nf, 0);
    m.Cf.assign(nft, Vec2{});
    m.Sf.assign(nft, Vec2{});
    m.magSf.assign(nft, 0.0);
    m.fp0.assign(nft, Vec2{});
    m.fp1.assign(nft, Vec2{});
    m.fshift.assign(m.nf, Vec2{});
    auto face_geom = [&](int f, int cell, int k) {
        const int nv = m.cell_ptr[cell + 1] - m.cell_ptr[cell];
        const Vec2 p0 = m.cell_xy[m.cell_ptr[cell] + k];
        const Vec2 p1 = m.cell_xy[m.cell_ptr[cell] + (k + 1) % nv];
        const double dx = p1.x - p0.x, dy = p1.y - p0.y;
        m.fp0[f] = p0;
        m.fp1[f] = p1;
        m.Cf[f] = Vec2{0.5 * (p0.x + p1.x), 0.5 * (p0.y + p1.y)};
        m.Sf[f] = Vec2{dy * m.depth, -dx * m.depth};
        m.magSf[f] = sqrt(dx * dx + dy * dy) * m.depth;
    };
    for (int f = 0; f < m.nf; ++f) {
        m.own[f] = nfv[f].own;
        m.nei[f] = nfv[f].nei;
        face_geom(f, nfv[f].own, nfv[f].kown);
        // the owner's edge runs p0->p1, the neighbour's edge p1->p0
        const int nvn = m.cell_ptr[nfv[f].nei + 1] - m.cell_ptr[nfv[f].nei];
        const int kn1 = (nfv[f].knei + 1) % nvn;  // neighbour's vertex matching owner's p0
        const int dsx = vs_new(nfv[f].own, nfv[f].kown, 0) - vs_new(nfv[f].nei, kn1, 0);
        const int dsy = vs_new(nfv[f].own, nfv[f].kown, 1) - vs_new(nfv[f].nei, kn1, 1);
        m.fshift[f] = Vec2{dsx * pm.period.x, dsy * pm.period.y};
    }
    m.bf_label.assign(m.nb, 0);
    m.bf_patch.assign(m.nb, 0);
    for (int b = 0; b < m.nb; ++b) {
        const int f = m.nf + b;
        m.own[f] = bnew[b].c;
        face_geom(f, bnew[b].c, bnew[b].k);
        m.bf_patch[b] = bnew[b].patch;
        m.bf_label[b] = (uint8_t)pm.patches[bnew[b].patch].label;
    }
    m.C.resize(m.nc + m.nb);
    for (int b = 0; b < m.nb; ++b) m.C[m.nc + b] = m.Cf[m.nf + b];
    return m;
}

}  // namespace afar
