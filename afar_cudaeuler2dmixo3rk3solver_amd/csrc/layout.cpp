// Reference-layout host arrays and the gfx950 device layout built from them.
#include "layout.h"

#include <math.h>

#include <algorithm>
#include <unordered_map>

namespace afar {

void conserved_from_primitive(RefArrays& ra, double R, double Cv) {
    // basicCUDAsolver.cpp:220-223
    ra.cFD.assign(4 * (size_t)ra.nc, 0.0);
    for (int i = 0; i < ra.nc; ++i) {
        const double p = ra.FD[4 * i], ux = ra.FD[4 * i + 1], uy = ra.FD[4 * i + 2], T = ra.FD[4 * i + 3];
        const double rho = p / (R * T);
        ra.cFD[4 * i] = rho;
        ra.cFD[4 * i + 1] = rho * ux;
        ra.cFD[4 * i + 2] = rho * uy;
        ra.cFD[4 * i + 3] = rho * (Cv * T + 0.5 * (ux * ux + uy * uy));
    }
}

void apply_bc_host(RefArrays& ra) {
    // Same formulas as the device boundary kernel (src/CUDA/BCs/2D/*.h).
    for (int b = 0; b < ra.nb; ++b) {
        const int c = ra.boundaryFaceNeiLabel[b];
        const double* w = &ra.FD[4 * (size_t)c];
        double* g = &ra.FD[4 * ((size_t)ra.nc + b)];
        const double* F = &ra.FACE[5 * ((size_t)ra.nf + b)];
        const double magSf = F[2], Sfx = F[3], Sfy = F[4];
        switch (ra.boundaryFacesType[b]) {
            case 0: g[0] = w[0]; g[1] = 0.0; g[2] = 0.0; g[3] = 1.0; break;
            case 1: g[0] = w[0]; g[1] = 0.0; g[2] = 0.0; g[3] = w[3]; break;
            case 2: {
                const double nfx = Sfx / magSf, nfy = Sfy / magSf;
                g[0] = w[0];
                g[1] = w[1] - (w[1] * nfx + w[2] * nfy) * nfx;
                g[2] = w[2] - (w[1] * nfx + w[2] * nfy) * nfy;
                g[3] = w[3];
                break;
            }
            case 3: g[0] = 1.0; g[1] = 3.0; g[2] = 0.0; g[3] = 1.0; break;
            case 4: g[0] = w[0]; g[1] = w[1]; g[2] = w[2]; g[3] = w[3]; break;
            case 5: g[0] = 30.05945; g[1] = 4.07795; g[2] = 0.0; g[3] = 5.97677; break;
            case 6: for (int k = 0; k < 4; ++k) g[k] = ra.BVALUE[4 * (size_t)b + k]; break;
            default: throw Error(1, "boundary label out of range");
        }
    }
}

RefArrays make_ref_arrays(const Mesh& m, const Precomputed& pc, const double* W0, double R, double Cv) {
    RefArrays ra;
    const int nc = m.nc, nf = m.nf, nb = m.nb;
    const Stencils& st = pc.st;
    ra.nc = nc;
    ra.nf = nf;
    ra.nb = nb;
    ra.maxS = st.maxS;
    ra.maxC = st.maxC;
    ra.CELL.resize(2 * (size_t)(nc + nb));
    for (int i = 0; i < nc + nb; ++i) {
        ra.CELL[2 * i] = m.C[i].x;
        ra.CELL[2 * i + 1] = m.C[i].y;
    }
    ra.FACE.resize(5 * (size_t)(nf + nb));
    for (int f = 0; f < nf + nb; ++f) {
        double* F = &ra.FACE[5 * (size_t)f];
        F[0] = m.Cf[f].x;
        F[1] = m.Cf[f].y;
        F[2] = m.magSf[f];
        F[3] = m.Sf[f].x;
        F[4] = m.Sf[f].y;
    }
    ra.cellVolume = m.V;
    ra.FD.assign(4 * (size_t)(nc + nb), 0.0);
    for (size_t i = 0; i < 4 * (size_t)nc; ++i) ra.FD[i] = W0[i];
    ra.neighbour.resize(2 * (size_t)nf);
    for (int f = 0; f < nf; ++f) {
        ra.neighbour[2 * f] = m.own[f];
        ra.neighbour[2 * f + 1] = m.nei[f];
    }
    for (size_t p = 0; p < m.patches.size(); ++p)
        if (m.patch_size[p] > 0) ra.boundaryFacesNum.push_back(m.patch_size[p]);
    ra.npatch = (int)ra.boundaryFacesNum.size();
    ra.boundaryFaceNeiLabel.resize(nb);
    ra.boundaryFacesType = m.bf_label;
    ra.BVALUE.assign(4 * (size_t)nb, 0.0);
    for (int b = 0; b < nb; ++b) {
        ra.boundaryFaceNeiLabel[b] = m.own[nf + b];
        for (int k = 0; k < 4; ++k) ra.BVALUE[4 * (size_t)b + k] = m.patches[m.bf_patch[b]].value[k];
    }
    ra.stencilSize = st.S;
    ra.compactStencilSize = st.Cc;
    ra.stencil.assign((size_t)st.maxS * nc, 0);
    ra.cellFaces.assign((size_t)st.maxC * nc, 0);
    for (int c = 0; c < nc; ++c) {
        for (int i = 0; i < st.S[c]; ++i) ra.stencil[(size_t)st.maxS * c + i] = st.sten[(size_t)c * kMaxValidStencil + i];
        for (int i = 0; i < st.Cc[c]; ++i)
            ra.cellFaces[(size_t)st.maxC * c + i] = st.cellFaces[(size_t)c * kMaxValidStencil + i];
    }
    ra.RBFbasis = pc.rbf;
    ra.FSHIFT.resize(2 * (size_t)nf);
    for (int f = 0; f < nf; ++f) {
        ra.FSHIFT[2 * f] = m.fshift[f].x;
        ra.FSHIFT[2 * f + 1] = m.fshift[f].y;
    }
    if (pc.ngp) {
        const int ngp = pc.ngp;
        ra.ngp = ngp;
        ra.GPW = pc.gp_w;
        ra.RBFgp = pc.rbf_gp;
        ra.GPD.resize(4 * (size_t)ngp * nf);
        for (int g = 0; g < ngp; ++g)
            for (int f = 0; f < nf; ++f) {
                const Vec2 x = face_gauss_point(m, f, pc.gp_xi[g]);
                const int own = m.own[f], nei = m.nei[f];
                const double xn = m.C[nei].x + m.fshift[f].x, yn = m.C[nei].y + m.fshift[f].y;
                double* d = &ra.GPD[4 * ((size_t)g * nf + f)];
                d[0] = x.x - m.C[own].x;
                d[1] = x.y - m.C[own].y;
                d[2] = x.x - xn;
                d[3] = x.y - yn;
            }
    }
    apply_bc_host(ra);
    conserved_from_primitive(ra, R, Cv);
    return ra;
}

namespace {

// Index of grid point (x, y) along the Hilbert curve through an n x n grid
// (n a power of two): the quadrant walk of Hilbert (1891), with the sub-square
// rotated or reflected so that consecutive indices are neighbours.
uint64_t hilbert_index(uint32_t n, uint32_t x, uint32_t y) {
    uint64_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

}  // namespace

Relabel compact_tile_order(const RefArrays& ra, const BatchIndex& bi, int group) {
    const int nc = ra.nc, nf = ra.nf;
    Relabel r;
    r.cell_old.resize(nc);
    r.cell_new.resize(nc);
    std::vector<uint64_t> key(nc);
    const uint32_t n = 1u << 16;
    for (int k = 0; k < bi.ninst(); ++k) {
        const int c0 = bi.cell_off[k], c1 = bi.cell_off[k + 1];
        if (c1 <= c0) continue;
        double x0 = ra.CELL[2 * c0], x1 = x0, y0 = ra.CELL[2 * c0 + 1], y1 = y0;
        for (int c = c0; c < c1; ++c) {
            x0 = std::min(x0, ra.CELL[2 * c]);
            x1 = std::max(x1, ra.CELL[2 * c]);
            y0 = std::min(y0, ra.CELL[2 * c + 1]);
            y1 = std::max(y1, ra.CELL[2 * c + 1]);
        }
        // one scale for both axes keeps tiles square on elongated domains
        const double ext = std::max(x1 - x0, y1 - y0);
        const double sc = ext > 0.0 ? (double)(n - 1) / ext : 0.0;
        for (int c = c0; c < c1; ++c) {
            const uint32_t gx = (uint32_t)std::min((double)(n - 1), (ra.CELL[2 * c] - x0) * sc);
            const uint32_t gy = (uint32_t)std::min((double)(n - 1), (ra.CELL[2 * c + 1] - y0) * sc);
            key[c] = hilbert_index(n, gx, gy);
            r.cell_old[c] = c;
        }
        std::sort(r.cell_old.begin() + c0, r.cell_old.begin() + c1,
                  [&](int a, int b) { return key[a] != key[b] ? key[a] < key[b] : a < b; });
        if (group > 0)
            for (int b0 = c0; b0 < c1;) {
                const int b1 = std::min(c1, (b0 / group + 1) * group);
                std::stable_sort(r.cell_old.begin() + b0, r.cell_old.begin() + b1, [&](int a, int b) {
                    const int ca = ra.compactStencilSize[a], cb = ra.compactStencilSize[b];
                    return ca != cb ? ca < cb : ra.stencilSize[a] < ra.stencilSize[b];
                });
                b0 = b1;
            }
    }
    for (int j = 0; j < nc; ++j) r.cell_new[r.cell_old[j]] = j;
    r.face_old.resize(nf);
    r.face_new.resize(nf);
    for (int f = 0; f < nf; ++f) r.face_old[f] = f;
    for (int k = 0; k < bi.ninst(); ++k) {
        const int f0 = bi.face_off[k], f1 = bi.face_off[k + 1];
        std::sort(r.face_old.begin() + f0, r.face_old.begin() + f1, [&](int a, int b) {
            const int oa = r.cell_new[ra.neighbour[2 * a]], ob = r.cell_new[ra.neighbour[2 * b]];
            return oa != ob ? oa < ob : a < b;
        });
    }
    for (int j = 0; j < nf; ++j) r.face_new[r.face_old[j]] = j;
    return r;
}

RefArrays relabel(const RefArrays& ra, const Relabel& r) {
    const int nc = ra.nc, nf = ra.nf, nb = ra.nb, maxS = ra.maxS, maxC = ra.maxC;
    RefArrays o;
    o.nc = nc;
    o.nf = nf;
    o.nb = nb;
    o.npatch = ra.npatch;
    o.maxS = maxS;
    o.maxC = maxC;
    auto cell = [&](int32_t c) { return c < nc ? r.cell_new[c] : c; };  // pseudo-cells keep their ids
    auto sface = [&](int32_t sf) {  // +-(f+1), boundary faces keep their ids
        if (sf == 0) return sf;
        const int f = (sf > 0 ? sf : -sf) - 1;
        if (f >= nf) return sf;
        const int32_t g = r.face_new[f] + 1;
        return sf > 0 ? g : -g;
    };
    o.CELL = ra.CELL;
    o.cellVolume.resize(nc);
    o.FD = ra.FD;
    o.cFD.resize(ra.cFD.size());
    o.stencilSize.resize(nc);
    o.compactStencilSize.resize(nc);
    o.stencil.assign(ra.stencil.size(), 0);
    o.cellFaces.assign(ra.cellFaces.size(), 0);
    for (int j = 0; j < nc; ++j) {
        const int c = r.cell_old[j];
        o.CELL[2 * j] = ra.CELL[2 * c];
        o.CELL[2 * j + 1] = ra.CELL[2 * c + 1];
        o.cellVolume[j] = ra.cellVolume[c];
        for (int k = 0; k < 4; ++k) o.FD[4 * (size_t)j + k] = ra.FD[4 * (size_t)c + k];
        if (!ra.cFD.empty())
            for (int k = 0; k < 4; ++k) o.cFD[4 * (size_t)j + k] = ra.cFD[4 * (size_t)c + k];
        o.stencilSize[j] = ra.stencilSize[c];
        o.compactStencilSize[j] = ra.compactStencilSize[c];
        for (int i = 0; i < maxS; ++i)
            o.stencil[(size_t)maxS * j + i] =
                i < ra.stencilSize[c] ? cell(ra.stencil[(size_t)maxS * c + i]) : ra.stencil[(size_t)maxS * c + i];
        for (int i = 0; i < maxC; ++i)
            o.cellFaces[(size_t)maxC * j + i] =
                i < ra.compactStencilSize[c] ? sface(ra.cellFaces[(size_t)maxC * c + i]) : 0;
    }
    o.neighbour.resize(2 * (size_t)nf);
    o.FACE = ra.FACE;
    o.FSHIFT.resize(ra.FSHIFT.size());
    const size_t W = 2 * (size_t)(maxS + 1);
    o.RBFbasis.resize(ra.RBFbasis.size());
    for (int g = 0; g < nf; ++g) {
        const int f = r.face_old[g];
        o.neighbour[2 * g] = cell(ra.neighbour[2 * f]);
        o.neighbour[2 * g + 1] = cell(ra.neighbour[2 * f + 1]);
        for (int k = 0; k < 5; ++k) o.FACE[5 * (size_t)g + k] = ra.FACE[5 * (size_t)f + k];
        if (!ra.FSHIFT.empty()) {
            o.FSHIFT[2 * g] = ra.FSHIFT[2 * f];
            o.FSHIFT[2 * g + 1] = ra.FSHIFT[2 * f + 1];
        }
        if (!ra.RBFbasis.empty())
            std::copy(ra.RBFbasis.begin() + W * f, ra.RBFbasis.begin() + W * (f + 1), o.RBFbasis.begin() + W * g);
    }
    o.ngp = ra.ngp;
    o.GPW = ra.GPW;
    o.GPD.resize(ra.GPD.size());
    o.RBFgp.resize(ra.RBFgp.size());
    for (int g = 0; g < ra.ngp; ++g)
        for (int h = 0; h < nf; ++h) {
            const int f = r.face_old[h];
            for (int k = 0; k < 4; ++k) o.GPD[4 * ((size_t)g * nf + h) + k] = ra.GPD[4 * ((size_t)g * nf + f) + k];
            std::copy(ra.RBFgp.begin() + W * ((size_t)g * nf + f), ra.RBFgp.begin() + W * ((size_t)g * nf + f + 1),
                      o.RBFgp.begin() + W * ((size_t)g * nf + h));
        }
    o.boundaryFacesNum = ra.boundaryFacesNum;
    o.boundaryFaceNeiLabel.resize(nb);
    for (int b = 0; b < nb; ++b) o.boundaryFaceNeiLabel[b] = cell(ra.boundaryFaceNeiLabel[b]);
    o.boundaryFacesType = ra.boundaryFacesType;
    o.BVALUE = ra.BVALUE;
    return o;
}

RefArrays union_ref_arrays(const std::vector<const RefArrays*>& parts, BatchIndex& idx) {
    const int n = (int)parts.size();
    if (n == 0) throw Error(1, "empty batch");
    idx.cell_off.assign(n + 1, 0);
    idx.face_off.assign(n + 1, 0);
    idx.bface_off.assign(n + 1, 0);
    RefArrays u;
    int64_t nc = 0, nf = 0, nb = 0;
    for (int k = 0; k < n; ++k) {
        const RefArrays& a = *parts[k];
        nc += a.nc;
        nf += a.nf;
        nb += a.nb;
        if (nc + nb > INT32_MAX / 2 || nf + nb > INT32_MAX / 2) throw Error(3, "batch too large for 32-bit indices");
        idx.cell_off[k + 1] = (int32_t)nc;
        idx.face_off[k + 1] = (int32_t)nf;
        idx.bface_off[k + 1] = (int32_t)nb;
        u.maxS = std::max(u.maxS, a.maxS);
        u.maxC = std::max(u.maxC, a.maxC);
        u.npatch += a.npatch;
        if (a.ngp != parts[0]->ngp || a.GPW != parts[0]->GPW)
            throw Error(1, "every mesh of a batch needs the same face quadrature");
    }
    u.ngp = parts[0]->ngp;
    u.GPW = parts[0]->GPW;
    u.nc = (int)nc;
    u.nf = (int)nf;
    u.nb = (int)nb;
    const int maxS = u.maxS, maxC = u.maxC;
    u.CELL.resize(2 * (size_t)(nc + nb));
    u.FACE.resize(5 * (size_t)(nf + nb));
    u.FD.resize(4 * (size_t)(nc + nb));
    u.cellVolume.resize(nc);
    u.cFD.resize(4 * (size_t)nc);
    u.neighbour.resize(2 * (size_t)nf);
    u.boundaryFaceNeiLabel.resize(nb);
    u.boundaryFacesType.resize(nb);
    u.stencilSize.resize(nc);
    u.compactStencilSize.resize(nc);
    u.stencil.assign((size_t)maxS * nc, 0);
    u.cellFaces.assign((size_t)maxC * nc, 0);
    u.RBFbasis.assign(2 * (size_t)(maxS + 1) * nf, 0.0);
    u.FSHIFT.resize(2 * (size_t)nf);
    u.BVALUE.resize(4 * (size_t)nb);
    u.GPD.assign(4 * (size_t)u.ngp * nf, 0.0);
    u.RBFgp.assign((size_t)u.ngp * 2 * (maxS + 1) * nf, 0.0);
    for (int k = 0; k < n; ++k) {
        const RefArrays& a = *parts[k];
        const int co = idx.cell_off[k], fo = idx.face_off[k], bo = idx.bface_off[k];
        // cell id (real or pseudo) of part k -> batch id
        auto cell = [&](int32_t id) -> int32_t { return id < a.nc ? co + id : (int32_t)nc + bo + (id - a.nc); };
        // face id (internal or boundary) of part k -> batch id
        auto face = [&](int32_t f) -> int32_t { return f < a.nf ? fo + f : (int32_t)nf + bo + (f - a.nf); };
        for (int i = 0; i < a.nc + a.nb; ++i) {
            const size_t d = cell(i);
            u.CELL[2 * d] = a.CELL[2 * (size_t)i];
            u.CELL[2 * d + 1] = a.CELL[2 * (size_t)i + 1];
            for (int q = 0; q < 4; ++q) u.FD[4 * d + q] = a.FD[4 * (size_t)i + q];
        }
        for (int f = 0; f < a.nf + a.nb; ++f) {
            const size_t d = face(f);
            for (int q = 0; q < 5; ++q) u.FACE[5 * d + q] = a.FACE[5 * (size_t)f + q];
        }
        for (int c = 0; c < a.nc; ++c) {
            const size_t d = co + c;
            u.cellVolume[d] = a.cellVolume[c];
            for (int q = 0; q < 4; ++q) u.cFD[4 * d + q] = a.cFD[4 * (size_t)c + q];
            u.stencilSize[d] = a.stencilSize[c];
            u.compactStencilSize[d] = a.compactStencilSize[c];
            for (int i = 0; i < a.stencilSize[c]; ++i)
                u.stencil[(size_t)maxS * d + i] = cell(a.stencil[(size_t)a.maxS * c + i]);
            for (int i = 0; i < a.compactStencilSize[c]; ++i) {
                const int32_t sf = a.cellFaces[(size_t)a.maxC * c + i];
                const int32_t f = (sf < 0 ? -sf : sf) - 1;
                const int32_t g = face(f) + 1;
                u.cellFaces[(size_t)maxC * d + i] = sf < 0 ? -g : g;
            }
        }
        for (int f = 0; f < a.nf; ++f) {
            const size_t d = fo + f;
            u.neighbour[2 * d] = cell(a.neighbour[2 * (size_t)f]);
            u.neighbour[2 * d + 1] = cell(a.neighbour[2 * (size_t)f + 1]);
            u.FSHIFT[2 * d] = a.FSHIFT[2 * (size_t)f];
            u.FSHIFT[2 * d + 1] = a.FSHIFT[2 * (size_t)f + 1];
            for (int side = 0; side < 2; ++side)
                for (int i = 0; i <= a.maxS; ++i)
                    u.RBFbasis[(2 * d + side) * (size_t)(maxS + 1) + i] =
                        a.RBFbasis[(2 * (size_t)f + side) * (size_t)(a.maxS + 1) + i];
            for (int g = 0; g < u.ngp; ++g) {
                for (int q = 0; q < 4; ++q)
                    u.GPD[4 * ((size_t)g * nf + d) + q] = a.GPD[4 * ((size_t)g * a.nf + f) + q];
                for (int side = 0; side < 2; ++side)
                    for (int i = 0; i <= a.maxS; ++i)
                        u.RBFgp[(((size_t)g * nf + d) * 2 + side) * (size_t)(maxS + 1) + i] =
                            a.RBFgp[(((size_t)g * a.nf + f) * 2 + side) * (size_t)(a.maxS + 1) + i];
            }
        }
        for (int b = 0; b < a.nb; ++b) {
            const size_t d = bo + b;
            u.boundaryFaceNeiLabel[d] = cell(a.boundaryFaceNeiLabel[b]);
            u.boundaryFacesType[d] = a.boundaryFacesType[b];
            for (int q = 0; q < 4; ++q) u.BVALUE[4 * d + q] = a.BVALUE[4 * (size_t)b + q];
        }
        u.boundaryFacesNum.insert(u.boundaryFacesNum.end(), a.boundaryFacesNum.begin(), a.boundaryFacesNum.end());
    }
    return u;
}

BlockTables make_block_tables(const RefArrays& ra) {
    const int B = 128, nc = ra.nc, maxS = ra.maxS;
    const int nblk = (nc + B - 1) / B;
    BlockTables bt;
    bt.extendStencilSize.assign(nblk, 0);
    bt.compactExtendStencilSize.assign(nblk, 0);
    bt.localStencil.assign((size_t)maxS * nc, 0);
    std::vector<std::vector<int32_t>> ext(nblk);
    auto outside = [&](int id, int blk) { return id < blk * B || id > B - 1 + blk * B || id >= nc; };
    for (int blk = 0; blk < nblk; ++blk) {
        std::unordered_map<int32_t, int32_t> pos;
        auto& e = ext[blk];
        for (int pass = 0; pass < 2; ++pass) {
            for (int t = 0; t < B; ++t) {
                const int c = t + blk * B;
                if (c >= nc) break;
                const int i0 = pass ? ra.compactStencilSize[c] : 0;
                const int i1 = pass ? ra.stencilSize[c] : ra.compactStencilSize[c];
                for (int i = i0; i < i1; ++i) {
                    const int id = ra.stencil[(size_t)maxS * c + i];
                    if (outside(id, blk) && pos.find(id) == pos.end()) {
                        pos.emplace(id, (int32_t)e.size());
                        e.push_back(id);
                    }
                }
            }
            if (pass == 0) {
                bt.compactExtendStencilSize[blk] = (uint16_t)e.size();
                bt.maxCLB = std::max(bt.maxCLB, (int)e.size() + B);
            }
        }
        bt.extendStencilSize[blk] = (uint16_t)e.size();
        bt.maxLB = std::max(bt.maxLB, (int)e.size() + B);
        for (int t = 0; t < B; ++t) {
            const int c = t + blk * B;
            if (c >= nc) break;
            for (int i = 0; i < ra.stencilSize[c]; ++i) {
                const int id = ra.stencil[(size_t)maxS * c + i];
                bt.localStencil[(size_t)maxS * c + i] =
                    (uint16_t)(outside(id, blk) ? B + pos[id] : id - blk * B);
            }
        }
    }
    bt.maxLB = std::max(bt.maxLB, B);
    bt.maxCLB = std::max(bt.maxCLB, B);
    const int stride = bt.maxLB - B;
    bt.extendStencil.assign((size_t)std::max(stride, 1) * nblk, 0);
    for (int blk = 0; blk < nblk; ++blk)
        for (size_t i = 0; i < ext[blk].size(); ++i) bt.extendStencil[(size_t)stride * blk + i] = ext[blk][i];
    return bt;
}

void global_stencil_from_blocks(int nc, int maxS, int maxLB, const uint8_t* S, const uint16_t* localStencil,
                                const int32_t* extendStencil, int32_t* stencil) {
    const int B = 128, stride = maxLB - B;
    for (int c = 0; c < nc; ++c) {
        const int blk = c / B;
        for (int i = 0; i < S[c]; ++i) {
            const int l = localStencil[(size_t)maxS * c + i];
            stencil[(size_t)maxS * c + i] = l < B ? blk * B + l : extendStencil[(size_t)stride * blk + (l - B)];
        }
    }
}

DevLayout make_dev_layout(const RefArrays& ra, int tile) {
    if (tile != 64 && tile != 128 && tile != 256) throw Error(1, "tile must be 64, 128 or 256");
    DevLayout L;
    const int nc = ra.nc, nf = ra.nf, nb = ra.nb, maxS = ra.maxS, maxC = ra.maxC;
    L.nc = nc;
    L.nf = nf;
    L.nb = nb;
    L.maxS = maxS;
    L.maxC = maxC;
    L.tile = tile;
    L.nwaves = (nc + 63) / 64;
    L.ntiles = (nc + tile - 1) / tile;
    const size_t nslot = (size_t)L.nwaves * maxC * 64;
    const int KR = maxC * (maxS + 1);
    L.scnt = ra.stencilSize;
    L.ccnt = ra.compactStencilSize;
    L.sten.assign((size_t)L.nwaves * std::max(maxS, 1) * 64, 0);
    L.slot_face.assign(nslot, 0);
    L.slot_d.assign(2 * nslot, 0.0);
    L.slot_e.assign(2 * nslot, 0.0);
    L.rbf.assign((size_t)L.nwaves * KR * 64, 0.0);
    L.ngp = ra.ngp;
    L.gpw = ra.GPW;
    L.fgd = ra.GPD;
    L.rbf_gp.assign((size_t)L.nwaves * ra.ngp * KR * 64, 0.0);
    L.vol = ra.cellVolume;
    const double* CELL = ra.CELL.data();
    const double* FACE = ra.FACE.data();
    const double* FS = ra.FSHIFT.data();

    // tile halos and LDS-local stencil ids
    std::vector<std::vector<int32_t>> halos(L.ntiles);
    std::vector<int> overflow(L.ntiles, 0);
    parallel_for(L.ntiles, [&](int64_t t0, int64_t t1) {
        std::unordered_map<int32_t, int32_t> loc;
        for (int64_t t = t0; t < t1; ++t) {
            loc.clear();
            auto& h = halos[t];
            const int c0 = (int)t * tile, c1 = std::min(nc, c0 + tile);
            for (int c = c0; c < c1; ++c)
                for (int i = 0; i < ra.stencilSize[c]; ++i) {
                    const int id = ra.stencil[(size_t)maxS * c + i];
                    int local;
                    if (id >= c0 && id < c1) {
                        local = id - c0;
                    } else {
                        auto it = loc.find(id);
                        if (it == loc.end()) {
                            local = tile + (int)h.size();
                            loc.emplace(id, local);
                            h.push_back(id);
                        } else {
                            local = it->second;
                        }
                    }
                    if (local > 65535) overflow[t] = 1;
                    L.sten[L.sell(c, i, maxS)] = (uint16_t)local;
                }
        }
    });
    for (int t = 0; t < L.ntiles; ++t)
        if (overflow[t]) throw Error(3, "Too large local block stencil size (LDS ids exceed 16 bits)");
    L.halo_ptr.assign(L.ntiles + 1, 0);
    for (int t = 0; t < L.ntiles; ++t) {
        L.halo_ptr[t + 1] = L.halo_ptr[t] + (int32_t)halos[t].size();
        L.maxHalo = std::max(L.maxHalo, (int)halos[t].size());
    }
    L.halo.resize(L.halo_ptr[L.ntiles]);
    for (int t = 0; t < L.ntiles; ++t) std::copy(halos[t].begin(), halos[t].end(), L.halo.begin() + L.halo_ptr[t]);

    // slots: face ids, geometry, RBF coefficients
    parallel_for(nc, [&](int64_t c0, int64_t c1) {
        for (int64_t cc = c0; cc < c1; ++cc) {
            const int c = (int)cc;
            for (int fi = 0; fi < ra.compactStencilSize[c]; ++fi) {
                const int sf = ra.cellFaces[(size_t)maxC * c + fi];
                const int f = (sf > 0 ? sf : -sf) - 1;
                const size_t s = L.sell(c, fi, maxC);
                L.slot_face[s] = sf;
                double* d = &L.slot_d[2 * s];
                double* e = &L.slot_e[2 * s];
                const double* F = &FACE[5 * (size_t)f];
                if (f < nf) {
                    const int own = ra.neighbour[2 * f], nei = ra.neighbour[2 * f + 1];
                    const double xn = CELL[2 * nei] + FS[2 * f], yn = CELL[2 * nei + 1] + FS[2 * f + 1];
                    if (sf > 0) {
                        d[0] = F[0] - CELL[2 * own];
                        d[1] = F[1] - CELL[2 * own + 1];
                        e[0] = xn - CELL[2 * own];
                        e[1] = yn - CELL[2 * own + 1];
                    } else {
                        d[0] = F[0] - xn;
                        d[1] = F[1] - yn;
                        e[0] = CELL[2 * own] - xn;
                        e[1] = CELL[2 * own + 1] - yn;
                    }
                    const int side = sf < 0;
                    const double* a = &ra.RBFbasis[(size_t)2 * (maxS + 1) * f + (size_t)side * (maxS + 1)];
                    for (int i = 0; i <= ra.stencilSize[c]; ++i) L.rbf[L.sell(c, fi * (maxS + 1) + i, KR)] = a[i];
                    for (int g = 0; g < ra.ngp; ++g) {
                        const double* ag =
                            &ra.RBFgp[(((size_t)g * nf + f) * 2 + side) * (size_t)(maxS + 1)];
                        for (int i = 0; i <= ra.stencilSize[c]; ++i)
                            L.rbf_gp[L.sell(c, (g * maxC + fi) * (maxS + 1) + i, ra.ngp * KR)] = ag[i];
                    }
                } else {
                    const int b = f - nf;
                    d[0] = F[0] - CELL[2 * c];
                    d[1] = F[1] - CELL[2 * c + 1];
                    e[0] = CELL[2 * ((size_t)nc + b)] - CELL[2 * c];
                    e[1] = CELL[2 * ((size_t)nc + b) + 1] - CELL[2 * c + 1];
                }
            }
        }
    });

    // internal faces
    L.fcell.resize(2 * (size_t)nf);
    L.fgeo.resize(8 * (size_t)nf);
    parallel_for(nf, [&](int64_t f0, int64_t f1) {
        for (int64_t f = f0; f < f1; ++f) {
            const int own = ra.neighbour[2 * f], nei = ra.neighbour[2 * f + 1];
            L.fcell[2 * f] = own;
            L.fcell[2 * f + 1] = nei;
            const double* F = &FACE[5 * f];
            const double xn = CELL[2 * nei] + FS[2 * f], yn = CELL[2 * nei + 1] + FS[2 * f + 1];
            const double ex = xn - CELL[2 * own], ey = yn - CELL[2 * own + 1];
            double* g = &L.fgeo[8 * f];
            g[0] = F[3];
            g[1] = F[4];
            g[2] = F[2];
            g[3] = sqrt(ex * ex + ey * ey);
            g[4] = F[0] - CELL[2 * own];
            g[5] = F[1] - CELL[2 * own + 1];
            g[6] = F[0] - xn;
            g[7] = F[1] - yn;
        }
    });
    // boundary faces
    L.bcell = ra.boundaryFaceNeiLabel;
    L.blabel = ra.boundaryFacesType;
    L.bgeo.assign(4 * (size_t)nb, 0.0);
    L.bval = ra.BVALUE;
    if (L.bval.size() != 4 * (size_t)nb) L.bval.assign(4 * (size_t)nb, 0.0);
    for (int b = 0; b < nb; ++b) {
        const double* F = &FACE[5 * ((size_t)nf + b)];
        L.bgeo[4 * b] = F[3];
        L.bgeo[4 * b + 1] = F[4];
        L.bgeo[4 * b + 2] = F[2];
    }
    return L;
}

}  // namespace afar
