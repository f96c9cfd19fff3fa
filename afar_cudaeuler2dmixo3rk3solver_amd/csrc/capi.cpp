// extern "C" entry points declared in include/afar.h.
#include "../../include/afar.h"

#include <hip/hip_runtime.h>
#include <string.h>

#include <memory>
#include <string>
#include <vector>

#include "afar_host.h"
#include "afar_math.h"
#include "kernels.h"
#include "foam_io.h"
#include "layout.h"
#include "solver.h"

using namespace afar;

namespace {
thread_local std::string g_last_error;

template <typename F>
int guard(F&& f) {
    try {
        f();
        return AFAR_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return AFAR_ERR_INVALID;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return AFAR_ERR_INVALID;
    }
}

void need(bool ok, const char* msg) {
    if (!ok) throw Error(AFAR_ERR_INVALID, msg);
}

Schemes to_schemes(const afar_schemes* s) {
    Schemes sc;
    if (!s) return sc;
    sc.R = s->R;
    sc.Cv = s->Cv;
    sc.maxCo = s->maxCo;
    sc.shapeFactor = s->shapeFactor;
    sc.stencilType = s->stencilType;
    sc.recalcBasis = s->recalcBasis;
    sc.cOrder = s->cOrder;
    sc.fOrder = s->fOrder;
    sc.precision = s->precision;
    sc.exactRK = s->exactRK;
    sc.faceQuadrature = s->faceQuadrature;
    need(sc.R > 0 && sc.Cv > 0, "R and Cv must be positive");
    need(sc.faceQuadrature == 0 || sc.faceQuadrature == 1, "faceQuadrature must be 0 or 1");
    need(sc.precision == 0 || sc.precision == 1, "precision must be 0 (FP64) or 1 (FP32-mixed)");
    return sc;
}

int tile_of(const afar_schemes* s) {
    const int t = (s && s->tile) ? s->tile : 64;
    need(t == 64 || t == 128 || t == 256, "tile must be 64, 128 or 256");
    return t;
}
}  // namespace

struct afar_mesh {
    Mesh m;
    Precomputed pc;
    bool precomputed = false;
    Schemes sc;
    int maxLB = 0, maxCLB = 0;
};

struct afar_solver {
    std::unique_ptr<DeviceSolver> dev;
    std::vector<int32_t> perm;  // mesh cell -> caller cell
    int nc = 0, nb = 0;
};

struct afar_batch {
    std::unique_ptr<DeviceSolver> dev;
    BatchIndex idx;
    std::vector<std::vector<int32_t>> perm;  // per instance: mesh cell -> caller cell
};

struct afar_device {
    RefArrays ra;
    SolverConfig cfg;
    std::unique_ptr<DeviceSolver> dev;
};

extern "C" {

const char* afar_last_error(void) { return g_last_error.c_str(); }
const char* afar_version(void) { return "afar-mi355x 0.1"; }

void afar_schemes_default(afar_schemes* s) {
    if (!s) return;
    memset(s, 0, sizeof(*s));
    s->R = 0.7142857;
    s->Cv = 1.7857142857;
    s->maxCo = 0.5;
    s->shapeFactor = 10.0;
    s->stencilType = 1;
    s->recalcBasis = 1;
    s->cOrder = 5;
    s->fOrder = 4;
    s->tile = 64;
}

// ---------------------------------------------------------------- meshes

static int finish_mesh(const PolyMesh& pm, bool renumber, afar_mesh** out) {
    need(out != nullptr, "out is NULL");
    auto h = std::make_unique<afar_mesh>();
    h->m = build_mesh(pm, renumber);
    *out = h.release();
    return AFAR_OK;
}

int afar_mesh_from_polygons(int64_t npoints, const double* xy, int64_t ncells, const int64_t* cell_ptr,
                            const int32_t* cell_pts, const int8_t* cell_vshift, double period_x, double period_y,
                            int64_t nbedges, const int32_t* bedge_pts, const int32_t* bedge_patch, int32_t npatches,
                            const int32_t* patch_labels, const double* patch_values, double depth, int32_t renumber,
                            afar_mesh** out) {
    return guard([&] {
        need(npoints > 0 && xy && ncells > 0 && cell_ptr && cell_pts, "empty mesh");
        need(nbedges == 0 || (bedge_pts && bedge_patch), "boundary edges missing");
        need(npatches == 0 || patch_labels, "patch labels missing");
        need(depth > 0, "depth must be positive");
        need(npoints < 2147483647 && ncells < 2147483647, "mesh too large for 32-bit ids");
        PolyMesh pm;
        pm.pts.resize(npoints);
        for (int64_t i = 0; i < npoints; ++i) pm.pts[i] = Vec2{xy[2 * i], xy[2 * i + 1]};
        pm.cell_ptr.assign(cell_ptr, cell_ptr + ncells + 1);
        need(pm.cell_ptr[0] == 0, "cell_ptr[0] must be 0");
        for (int64_t c = 0; c < ncells; ++c) need(pm.cell_ptr[c + 1] > pm.cell_ptr[c], "cell_ptr must increase");
        pm.cell_pts.assign(cell_pts, cell_pts + pm.cell_ptr[ncells]);
        if (cell_vshift) pm.cell_vshift.assign(cell_vshift, cell_vshift + 2 * pm.cell_ptr[ncells]);
        pm.period = Vec2{period_x, period_y};
        pm.bedges.resize(nbedges);
        pm.bedge_patch.resize(nbedges);
        for (int64_t i = 0; i < nbedges; ++i) {
            pm.bedges[i] = {bedge_pts[2 * i], bedge_pts[2 * i + 1]};
            pm.bedge_patch[i] = bedge_patch[i];
        }
        pm.patches.resize(npatches);
        for (int p = 0; p < npatches; ++p) {
            pm.patches[p].name = "patch" + std::to_string(p);
            pm.patches[p].label = patch_labels[p];
            if (patch_values)
                for (int k = 0; k < 4; ++k) pm.patches[p].value[k] = patch_values[4 * p + k];
        }
        pm.depth = depth;
        finish_mesh(pm, renumber != 0, out);
    });
}

int afar_mesh_read_foam(const char* case_dir, int32_t renumber, afar_mesh** out) {
    return guard([&] {
        need(case_dir && out, "NULL argument");
        PolyMesh pm;
        try {
            pm = read_polymesh_2d(case_dir, nullptr, nullptr);
        } catch (const Error& e) {
            throw Error(AFAR_ERR_IO, e.what());
        }
        finish_mesh(pm, renumber != 0, out);
    });
}

int afar_mesh_write_foam(const afar_mesh* h, const char* case_dir) {
    return guard([&] {
        need(h && case_dir, "NULL argument");
        try {
            write_polymesh_2d(h->m, case_dir);
        } catch (const Error& e) {
            throw Error(AFAR_ERR_IO, e.what());
        }
    });
}

int afar_mesh_patch_info(const afar_mesh* h, int32_t patch, char* name, int32_t cap, int32_t* nfaces) {
    return guard([&] {
        need(h != nullptr, "NULL mesh");
        need(patch >= 0 && patch < (int)h->m.patches.size(), "patch index out of range");
        if (name && cap > 0) snprintf(name, cap, "%s", h->m.patches[patch].name.c_str());
        if (nfaces) *nfaces = h->m.patch_size[patch];
    });
}

int afar_foam_read_fields(const afar_mesh* h, const char* case_dir, const char* time_dir, double* W0,
                          double* patch_values, int32_t* patch_fixed) {
    return guard([&] {
        need(h && case_dir && time_dir && W0, "NULL argument");
        const Mesh& m = h->m;
        std::vector<std::string> names;
        for (const Patch& p : m.patches) names.push_back(p.name);
        const std::string d = std::string(case_dir) + "/" + time_dir + "/";
        FoamField fp, fU, fT;
        try {
            fp = read_field(d + "p", 1, m.nc, names);
            fU = read_field(d + "U", 2, m.nc, names);
            fT = read_field(d + "T", 1, m.nc, names);
        } catch (const Error& e) {
            throw Error(AFAR_ERR_IO, e.what());
        }
        need(fp.have_internal && fU.have_internal && fT.have_internal, "a field has no internalField");
        for (int i = 0; i < m.nc; ++i) {
            W0[4 * (size_t)i] = fp.internal[i];
            W0[4 * (size_t)i + 1] = fU.internal[2 * (size_t)i];
            W0[4 * (size_t)i + 2] = fU.internal[2 * (size_t)i + 1];
            W0[4 * (size_t)i + 3] = fT.internal[i];
        }
        for (size_t p = 0; p < names.size(); ++p) {
            if (patch_values) {
                patch_values[4 * p] = fp.patch_value[p];
                patch_values[4 * p + 1] = fU.patch_value[2 * p];
                patch_values[4 * p + 2] = fU.patch_value[2 * p + 1];
                patch_values[4 * p + 3] = fT.patch_value[p];
            }
            if (patch_fixed) patch_fixed[p] = fp.patch_fixed[p] && fU.patch_fixed[p] && fT.patch_fixed[p];
        }
    });
}

int afar_mesh_generate_corner(int32_t nx, int32_t ny, double lx, double ly, double jitter, uint64_t seed,
                              double corner_x, double angle_deg, const double* inflow, int32_t top_label,
                              afar_mesh** out) {
    return guard([&] {
        need(inflow != nullptr, "inflow is NULL");
        need(top_label >= 0 && top_label <= kMaxBCLabel, "top_label out of range");
        finish_mesh(generate_corner(nx, ny, lx, ly, jitter, seed, corner_x, angle_deg, inflow, top_label), true, out);
    });
}

int afar_mesh_generate_riemann(int32_t n, double jitter, uint64_t seed, double split_p, afar_mesh** out) {
    return guard([&] { finish_mesh(generate_riemann(n, jitter, seed, split_p), true, out); });
}

int afar_mesh_generate_periodic(int32_t n, double L, double jitter, uint64_t seed, double split_p, afar_mesh** out) {
    return guard([&] { finish_mesh(generate_periodic(n, L, jitter, seed, split_p), true, out); });
}

void afar_mesh_destroy(afar_mesh* m) { delete m; }

int afar_mesh_set_patch(afar_mesh* h, int32_t patch, int32_t label, const double* value) {
    return guard([&] {
        need(h != nullptr, "NULL mesh");
        need(label >= 0 && label <= kMaxBCLabel, "boundary label out of range 0..6");
        Mesh& m = h->m;
        // the patch index counts every patch, empty ones included, as the
        // reference's patch<i> names do (basicCUDAsolver.cpp:157-172)
        need(patch >= 0 && patch < (int)m.patches.size(), "patch index out of range");
        m.patches[patch].label = label;
        if (value)
            for (int k = 0; k < 4; ++k) m.patches[patch].value[k] = value[k];
        for (int b = m.patch_start[patch]; b < m.patch_start[patch] + m.patch_size[patch]; ++b)
            m.bf_label[b] = (uint8_t)label;
    });
}

int afar_mesh_sizes(const afar_mesh* h, int64_t* out) {
    return guard([&] {
        need(h && out, "NULL argument");
        out[0] = h->m.nc;
        out[1] = h->m.nf;
        out[2] = h->m.nb;
        int np = 0;
        for (int s : h->m.patch_size) np += s > 0;
        out[3] = np;
        out[4] = h->precomputed ? h->pc.st.maxS : 0;
        out[5] = h->precomputed ? h->pc.st.maxC : 0;
        out[6] = h->maxLB;
        out[7] = h->maxCLB;
        out[8] = (int64_t)h->m.patches.size();
    });
}

int afar_mesh_cells(const afar_mesh* h, double* centres, double* volumes) {
    return guard([&] {
        need(h != nullptr, "NULL mesh");
        const Mesh& m = h->m;
        if (centres)
            for (int i = 0; i < m.nc + m.nb; ++i) {
                centres[2 * i] = m.C[i].x;
                centres[2 * i + 1] = m.C[i].y;
            }
        if (volumes) memcpy(volumes, m.V.data(), sizeof(double) * m.nc);
    });
}

int afar_mesh_permutation(const afar_mesh* h, int32_t* perm) {
    return guard([&] {
        need(h && perm, "NULL argument");
        memcpy(perm, h->m.perm.data(), sizeof(int32_t) * h->m.nc);
    });
}

int afar_mesh_faces(const afar_mesh* h, int32_t* own, int32_t* nei, double* face, double* fshift, int32_t* bpatch) {
    return guard([&] {
        need(h != nullptr, "NULL mesh");
        const Mesh& m = h->m;
        const int nft = m.nf + m.nb;
        if (own) memcpy(own, m.own.data(), sizeof(int32_t) * nft);
        if (nei) memcpy(nei, m.nei.data(), sizeof(int32_t) * m.nf);
        if (face)
            for (int f = 0; f < nft; ++f) {
                face[5 * f] = m.Cf[f].x;
                face[5 * f + 1] = m.Cf[f].y;
                face[5 * f + 2] = m.magSf[f];
                face[5 * f + 3] = m.Sf[f].x;
                face[5 * f + 4] = m.Sf[f].y;
            }
        if (fshift)
            for (int f = 0; f < m.nf; ++f) {
                fshift[2 * f] = m.fshift[f].x;
                fshift[2 * f + 1] = m.fshift[f].y;
            }
        if (bpatch) memcpy(bpatch, m.bf_patch.data(), sizeof(int32_t) * m.nb);
    });
}

int afar_mesh_polygons(const afar_mesh* h, int32_t* vptr, double* vxy) {
    return guard([&] {
        need(h != nullptr, "NULL mesh");
        const Mesh& m = h->m;
        if (vptr) memcpy(vptr, m.cell_ptr.data(), sizeof(int32_t) * (m.nc + 1));
        if (vxy)
            for (size_t k = 0; k < m.cell_xy.size(); ++k) {
                vxy[2 * k] = m.cell_xy[k].x;
                vxy[2 * k + 1] = m.cell_xy[k].y;
            }
    });
}

// ---------------------------------------------------------------- precompute

static void finish_precompute(afar_mesh* h, const Schemes& sc);

int afar_mesh_precompute_gpu(afar_mesh* h, const afar_schemes* s) {
    return guard([&] {
        need(h && s, "NULL argument");
        Schemes sc = to_schemes(s);
        need(sc.recalcBasis, "the GPU precompute computes the basis (recalcBasis true)");
        Schemes st_only = sc;
        st_only.recalcBasis = 0;  // stencils on the host, coefficients on the GPU
        Precomputed pc = precompute(h->m, st_only);
        precompute_rbf_gpu(h->m, sc, pc);  // stream-ordered buffers on a stream of its own
        h->pc = std::move(pc);
        finish_precompute(h, sc);
    });
}

int afar_mesh_precompute(afar_mesh* h, const afar_schemes* s, const char* basis_dir) {
    return guard([&] {
        need(h && s, "NULL argument");
        Schemes sc = to_schemes(s);
        need(!(sc.faceQuadrature && !sc.recalcBasis),
             "faceQuadrature needs recalcBasis true (the basis cache holds face-centre coefficients only)");
        h->pc = precompute(h->m, sc);
        if (!sc.recalcBasis) {
            if (!basis_dir) throw Error(AFAR_ERR_IO, "recalcBasis false needs the basis directory");
            try {
                read_basis_lists(h->m, h->pc, basis_dir);
            } catch (const Error& e) {
                throw Error(AFAR_ERR_IO, e.what());
            }
        } else if (basis_dir) {
            try {
                write_basis_lists(h->m, h->pc, basis_dir);
            } catch (const Error& e) {
                throw Error(AFAR_ERR_IO, e.what());
            }
        }
        finish_precompute(h, sc);
    });
}

static void finish_precompute(afar_mesh* h, const Schemes& sc) {
    {
        h->sc = sc;
        h->precomputed = true;
        // reference block-table sizes (reported, and used by the seam-2 export)
        RefArrays ra;
        ra.nc = h->m.nc;
        ra.maxS = h->pc.st.maxS;
        ra.stencilSize = h->pc.st.S;
        ra.compactStencilSize = h->pc.st.Cc;
        ra.stencil.assign((size_t)ra.maxS * ra.nc, 0);
        for (int c = 0; c < ra.nc; ++c)
            for (int i = 0; i < h->pc.st.S[c]; ++i)
                ra.stencil[(size_t)ra.maxS * c + i] = h->pc.st.sten[(size_t)c * kMaxValidStencil + i];
        BlockTables bt = make_block_tables(ra);
        h->maxLB = bt.maxLB;
        h->maxCLB = bt.maxCLB;
    }
}

int afar_mesh_stencils(const afar_mesh* h, uint8_t* S, uint8_t* C, int32_t* stencil, int32_t* cellFaces) {
    return guard([&] {
        need(h && h->precomputed, "mesh not precomputed");
        const Stencils& st = h->pc.st;
        const int nc = h->m.nc;
        if (S) memcpy(S, st.S.data(), nc);
        if (C) memcpy(C, st.Cc.data(), nc);
        for (int c = 0; c < nc; ++c) {
            if (stencil)
                for (int i = 0; i < st.maxS; ++i)
                    stencil[(size_t)st.maxS * c + i] = i < st.S[c] ? st.sten[(size_t)c * kMaxValidStencil + i] : -1;
            if (cellFaces)
                for (int i = 0; i < st.maxC; ++i)
                    cellFaces[(size_t)st.maxC * c + i] = i < st.Cc[c] ? st.cellFaces[(size_t)c * kMaxValidStencil + i] : 0;
        }
    });
}

int afar_mesh_rbf_basis(const afar_mesh* h, double* rbf, uint8_t* trouble, double* refDistance) {
    return guard([&] {
        need(h && h->precomputed, "mesh not precomputed");
        if (rbf) memcpy(rbf, h->pc.rbf.data(), sizeof(double) * h->pc.rbf.size());
        if (trouble) memcpy(trouble, h->pc.trouble.data(), h->m.nc);
        if (refDistance) memcpy(refDistance, h->pc.refDistance.data(), sizeof(double) * h->m.nc);
    });
}

int afar_mesh_face_quadrature(const afar_mesh* h, int32_t* ngp, double* xi, double* weights, double* offsets,
                              double* rbf) {
    return guard([&] {
        need(h && h->precomputed && ngp, "mesh not precomputed or ngp NULL");
        const Precomputed& pc = h->pc;
        *ngp = pc.ngp;
        if (!pc.ngp) return;
        if (xi) memcpy(xi, pc.gp_xi.data(), sizeof(double) * pc.ngp);
        if (weights) memcpy(weights, pc.gp_w.data(), sizeof(double) * pc.ngp);
        if (rbf) memcpy(rbf, pc.rbf_gp.data(), sizeof(double) * pc.rbf_gp.size());
        if (offsets) {
            // the offsets the solver uses (make_ref_arrays), without a state
            const Mesh& m = h->m;
            for (int g = 0; g < pc.ngp; ++g)
                for (int f = 0; f < m.nf; ++f) {
                    const Vec2 x = face_gauss_point(m, f, pc.gp_xi[g]);
                    const int own = m.own[f], nei = m.nei[f];
                    const double xn = m.C[nei].x + m.fshift[f].x, yn = m.C[nei].y + m.fshift[f].y;
                    double* d = &offsets[4 * ((size_t)g * m.nf + f)];
                    d[0] = x.x - m.C[own].x;
                    d[1] = x.y - m.C[own].y;
                    d[2] = x.x - xn;
                    d[3] = x.y - yn;
                }
        }
    });
}

int afar_mesh_gauss_points(const afar_mesh* h, int32_t member, int32_t order, double* xy, double* w, int32_t cap,
                           int32_t* count) {
    return guard([&] {
        need(h && count, "NULL argument");
        const Mesh& m = h->m;
        need(member >= 0 && member < m.nc + m.nb, "member out of range");
        std::vector<Vec2> p;
        std::vector<double> ww;
        if (member < m.nc)
            gauss_cell(m, member, order, p, ww);
        else
            gauss_boundary_face(m, member - m.nc, order, p, ww);
        *count = (int32_t)p.size();
        for (int k = 0; k < (int)p.size() && k < cap; ++k) {
            if (xy) {
                xy[2 * k] = p[k].x;
                xy[2 * k + 1] = p[k].y;
            }
            if (w) w[k] = ww[k];
        }
    });
}

int afar_mesh_write_basis(const afar_mesh* h, const char* dir) {
    return guard([&] {
        need(h && h->precomputed && dir, "mesh not precomputed or dir NULL");
        try {
            write_basis_lists(h->m, h->pc, dir);
        } catch (const Error& e) {
            throw Error(AFAR_ERR_IO, e.what());
        }
    });
}

int afar_pinv(int32_t n, const double* A, double* G) {
    return guard([&] {
        need(n > 0 && A && G, "bad arguments");
        pinv_jacobi(n, A, G);
    });
}

// ---------------------------------------------------------------- reference arrays

static RefArrays ref_arrays_of(const afar_mesh* h, const afar_schemes* s, const double* W0) {
    need(h && h->precomputed, "mesh not precomputed");
    need(W0 != nullptr, "W0 is NULL");
    Schemes sc = to_schemes(s);
    need((sc.faceQuadrature != 0) == (h->pc.ngp > 0),
         "faceQuadrature of the solver differs from the mesh precompute's; precompute with the same schemes");
    const Mesh& m = h->m;
    std::vector<double> w(4 * (size_t)m.nc);
    for (int i = 0; i < m.nc; ++i)
        for (int k = 0; k < 4; ++k) w[4 * (size_t)i + k] = W0[4 * (size_t)m.perm[i] + k];
    return make_ref_arrays(m, h->pc, w.data(), sc.R, sc.Cv);
}

int afar_mesh_export_reference(const afar_mesh* h, const afar_schemes* s, const double* W0, afar_field_pointer* o) {
    return guard([&] {
        need(o != nullptr, "out is NULL");
        RefArrays ra = ref_arrays_of(h, s, W0);
        auto cp = [](auto* dst, const auto& src) {
            if (dst) memcpy(dst, src.data(), sizeof(src[0]) * src.size());
        };
        cp(o->CELL, ra.CELL);
        cp(o->FACE, ra.FACE);
        cp(o->cellVolume, ra.cellVolume);
        cp(o->FD, ra.FD);
        cp(o->cFD, ra.cFD);
        cp(o->neighbour, ra.neighbour);
        cp(o->boundaryFacesNum, ra.boundaryFacesNum);
        cp(o->boundaryFaceNeiLabel, ra.boundaryFaceNeiLabel);
        cp(o->boundaryFacesType, ra.boundaryFacesType);
        cp(o->stencilSize, ra.stencilSize);
        cp(o->compactStencilSize, ra.compactStencilSize);
        cp(o->cellFaces, ra.cellFaces);
        cp(o->globalStencil, ra.stencil);
        cp(o->RBFbasis64, ra.RBFbasis);
        cp(o->faceShift, ra.FSHIFT);
        cp(o->boundaryValue, ra.BVALUE);
        if (o->RBFbasis)
            for (size_t i = 0; i < ra.RBFbasis.size(); ++i) o->RBFbasis[i] = (float)ra.RBFbasis[i];
        if (o->shockIndicator) memset(o->shockIndicator, 0, 2 * (size_t)(ra.nc + ra.nb));
        if (o->localStencil || o->extendStencil) {
            BlockTables bt = make_block_tables(ra);
            cp(o->localStencil, bt.localStencil);
            cp(o->extendStencil, bt.extendStencil);
            cp(o->extendStencilSize, bt.extendStencilSize);
            cp(o->compactExtendStencilSize, bt.compactExtendStencilSize);
        }
    });
}

// ---------------------------------------------------------------- seam 1

static int solver_create(const afar_mesh* h, const afar_schemes* s, const double* W0, hipStream_t stream,
                         afar_solver** out) {
    return guard([&] {
        need(out != nullptr, "out is NULL");
        RefArrays ra = ref_arrays_of(h, s, W0);
        SolverConfig cfg;
        cfg.stream = stream;
        cfg.R = s->R;
        cfg.Cv = s->Cv;
        cfg.CFL = s->maxCo;
        cfg.precision = s->precision;
        cfg.exactRK = s->exactRK;
        cfg.tile = tile_of(s);
        cfg.boundaryFirstOrder = s->boundaryFirstOrder != 0;
        need(s->reconLanes == 0 || s->reconLanes == 1 || s->reconLanes == 4, "reconLanes must be 0, 1 or 4");
        cfg.reconLanes = s->reconLanes;
        need(s->cellOrder >= 0 && s->cellOrder <= 2, "cellOrder must be 0, 1 or 2");
        cfg.cellOrder = s->cellOrder;
        auto sol = std::make_unique<afar_solver>();
        sol->perm = h->m.perm;
        sol->nc = h->m.nc;
        sol->nb = h->m.nb;
        sol->dev = make_device_solver(ra, cfg);
        *out = sol.release();
    });
}

int afar_solver_create(const afar_mesh* h, const afar_schemes* s, const double* W0, afar_solver** out) {
    return solver_create(h, s, W0, nullptr, out);
}

int afar_solver_create_on_stream(const afar_mesh* h, const afar_schemes* s, const double* W0, void* stream,
                                 afar_solver** out) {
    if (!stream) return guard([] { need(false, "stream is NULL (use afar_solver_create for a stream of its own)"); });
    return solver_create(h, s, W0, static_cast<hipStream_t>(stream), out);
}

// ---------------------------------------------------------------- batches

int afar_batch_create(int32_t n, const afar_mesh* const* meshes, const afar_schemes* s, const double* const* W0,
                      afar_batch** out) {
    return guard([&] {
        need(out && meshes && W0 && s && n > 0, "bad arguments");
        auto b = std::make_unique<afar_batch>();
        RefArrays u;
        {
            std::vector<RefArrays> parts(n);
            std::vector<const RefArrays*> pp(n);
            for (int k = 0; k < n; ++k) {
                need(meshes[k] != nullptr && W0[k] != nullptr, "NULL mesh or state in the batch");
                parts[k] = ref_arrays_of(meshes[k], s, W0[k]);
                pp[k] = &parts[k];
                b->perm.push_back(meshes[k]->m.perm);
            }
            u = union_ref_arrays(pp, b->idx);
        }
        SolverConfig cfg;
        cfg.R = s->R;
        cfg.Cv = s->Cv;
        cfg.CFL = s->maxCo;
        cfg.precision = s->precision;
        cfg.exactRK = s->exactRK;
        cfg.tile = tile_of(s);
        cfg.boundaryFirstOrder = s->boundaryFirstOrder != 0;
        need(s->reconLanes == 0 || s->reconLanes == 1 || s->reconLanes == 4, "reconLanes must be 0, 1 or 4");
        cfg.reconLanes = s->reconLanes;
        need(s->cellOrder >= 0 && s->cellOrder <= 2, "cellOrder must be 0, 1 or 2");
        cfg.cellOrder = s->cellOrder;
        b->dev = make_device_solver(u, cfg, &b->idx);
        *out = b.release();
    });
}

int afar_batch_evolve(afar_batch* b, int32_t nsteps) {
    return guard([&] {
        need(b && nsteps >= 0, "bad arguments");
        b->dev->evolve_auto(nsteps);
    });
}

int afar_batch_time_steps(afar_batch* b, int32_t nsteps, double* ms) {
    return guard([&] {
        need(b && ms && nsteps >= 0, "bad arguments");
        *ms = b->dev->time_steps(nsteps, true);
    });
}

int afar_batch_synchronize(afar_batch* b) {
    return guard([&] {
        need(b != nullptr, "NULL batch");
        b->dev->synchronize();
    });
}

int afar_batch_reset(afar_batch* b) {
    return guard([&] {
        need(b != nullptr, "NULL batch");
        b->dev->reset_state();
    });
}

int afar_batch_sizes(afar_batch* b, int64_t* out) {
    return guard([&] {
        need(b && out, "NULL argument");
        const DevLayout& L = b->dev->layout();
        out[0] = b->idx.ninst();
        out[1] = L.nc;
        out[2] = L.nf;
        out[3] = L.nb;
        out[4] = L.tile;
        out[5] = L.maxS;
        out[6] = L.maxC;
        out[7] = L.maxHalo;
    });
}

int afar_batch_cell_offsets(afar_batch* b, int32_t* off) {
    return guard([&] {
        need(b && off, "NULL argument");
        memcpy(off, b->idx.cell_off.data(), sizeof(int32_t) * b->idx.cell_off.size());
    });
}

int afar_batch_last_dt(afar_batch* b, double* dt) {
    return guard([&] {
        need(b && dt, "NULL argument");
        b->dev->last_dts(dt);
    });
}

int afar_batch_get_conserved(afar_batch* b, int32_t inst, double* U) {
    return guard([&] {
        need(b && U && inst >= 0 && inst < b->idx.ninst(), "bad arguments");
        const int c0 = b->idx.cell_off[inst], n = b->idx.cell_off[inst + 1] - c0;
        std::vector<double> u(4 * (size_t)n);
        b->dev->get_conserved_range(c0, n, u.data());
        const auto& perm = b->perm[inst];
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 4; ++k) U[4 * (size_t)perm[i] + k] = u[4 * (size_t)i + k];
    });
}

int afar_batch_write_field(afar_batch* b, int32_t inst, double* W, int8_t* shock) {
    return guard([&] {
        need(b && inst >= 0 && inst < b->idx.ninst(), "bad arguments");
        const int c0 = b->idx.cell_off[inst], n = b->idx.cell_off[inst + 1] - c0;
        std::vector<double> w(4 * (size_t)n);
        std::vector<int8_t> t(n);
        b->dev->get_primitive_range(c0, n, w.data(), t.data());
        const auto& perm = b->perm[inst];
        for (int i = 0; i < n; ++i) {
            if (W)
                for (int k = 0; k < 4; ++k) W[4 * (size_t)perm[i] + k] = w[4 * (size_t)i + k];
            if (shock) shock[perm[i]] = t[i];
        }
    });
}

int afar_batch_profile(afar_batch* b, int32_t reps, double* ms, int32_t* count) {
    return guard([&] {
        need(b && ms && count && reps > 0, "bad arguments");
        int c[kKernKinds];
        b->dev->profile(ms, c, reps);
        for (int k = 0; k < kKernKinds; ++k) count[k] = c[k];
    });
}

void afar_batch_destroy(afar_batch* b) {
    std::lock_guard<std::mutex> lk(runtime_mutex());  // device frees
    delete b;
}

int afar_solver_query(afar_solver* h, int32_t* idle) {
    return guard([&] {
        need(h && idle, "NULL argument");
        *idle = h->dev->idle() ? 1 : 0;
    });
}

int afar_solver_set_time(afar_solver* h, double t, double t_end) {
    return guard([&] {
        need(h != nullptr, "NULL solver");
        h->dev->set_time(-1, t, t_end);
    });
}

int afar_solver_get_time(afar_solver* h, double* t, double* steps) {
    return guard([&] {
        need(h && (t || steps), "NULL argument");
        h->dev->get_time(t, steps);
    });
}

int afar_solver_residual_norm(afar_solver* h, double* norm) {
    return guard([&] {
        need(h && norm, "NULL argument");
        *norm = h->dev->residual_norm();
    });
}

int afar_batch_set_time(afar_batch* b, int32_t inst, double t, double t_end) {
    return guard([&] {
        need(b != nullptr, "NULL batch");
        b->dev->set_time(inst, t, t_end);
    });
}

int afar_batch_get_time(afar_batch* b, double* t, double* steps) {
    return guard([&] {
        need(b && (t || steps), "NULL argument");
        b->dev->get_time(t, steps);
    });
}

int afar_solver_set_delta_t(afar_solver* h, double* dt) {
    return guard([&] {
        need(h && dt, "NULL argument");
        *dt = h->dev->compute_dt();
    });
}

int afar_solver_evolve(afar_solver* h, double dt) {
    return guard([&] {
        need(h != nullptr, "NULL solver");
        h->dev->evolve(dt);
    });
}

int afar_solver_evolve_auto(afar_solver* h, int32_t nsteps) {
    return guard([&] {
        need(h != nullptr && nsteps >= 0, "bad arguments");
        h->dev->evolve_auto(nsteps);
    });
}

int afar_solver_last_dt(afar_solver* h, double* dt) {
    return guard([&] {
        need(h && dt, "NULL argument");
        *dt = h->dev->last_dt();
    });
}

int afar_solver_synchronize(afar_solver* h) {
    return guard([&] {
        need(h != nullptr, "NULL solver");
        h->dev->synchronize();
    });
}

int afar_solver_write_field(afar_solver* h, double* W, double* ghost, int8_t* shock) {
    return guard([&] {
        need(h != nullptr, "NULL solver");
        const int nc = h->nc, nb = h->nb;
        std::vector<double> fd(4 * (size_t)(nc + nb));
        std::vector<int8_t> sh(2 * (size_t)(nc + nb));
        h->dev->copy_out(fd.data(), sh.data());
        for (int i = 0; i < nc; ++i) {
            const int o = h->perm[i];
            if (W)
                for (int k = 0; k < 4; ++k) W[4 * (size_t)o + k] = fd[4 * (size_t)i + k];
            if (shock) shock[o] = sh[2 * (size_t)i + 1];
        }
        if (ghost) memcpy(ghost, fd.data() + 4 * (size_t)nc, sizeof(double) * 4 * nb);
    });
}

int afar_solver_get_conserved(afar_solver* h, double* U) {
    return guard([&] {
        need(h && U, "NULL argument");
        std::vector<double> u(4 * (size_t)h->nc);
        h->dev->get_conserved(u.data());
        for (int i = 0; i < h->nc; ++i)
            for (int k = 0; k < 4; ++k) U[4 * (size_t)h->perm[i] + k] = u[4 * (size_t)i + k];
    });
}

int afar_solver_set_conserved(afar_solver* h, const double* U) {
    return guard([&] {
        need(h && U, "NULL argument");
        std::vector<double> u(4 * (size_t)h->nc);
        for (int i = 0; i < h->nc; ++i)
            for (int k = 0; k < 4; ++k) u[4 * (size_t)i + k] = U[4 * (size_t)h->perm[i] + k];
        h->dev->set_conserved(u.data());
    });
}

int afar_solver_time_steps(afar_solver* h, int32_t nsteps, int32_t use_graph, double* ms) {
    return guard([&] {
        need(h && ms && nsteps >= 0, "bad arguments");
        *ms = h->dev->time_steps(nsteps, use_graph != 0);
    });
}

int afar_solver_time_steps_kernels(afar_solver* h, int32_t nsteps, double* ms_total, double* ms_by_kind,
                                   int32_t* launches_by_kind) {
    return guard([&] {
        need(h && ms_total && ms_by_kind && launches_by_kind && nsteps >= 0, "bad arguments");
        int c[kKernKinds];
        *ms_total = h->dev->time_steps_kernels(nsteps, ms_by_kind, c);
        for (int k = 0; k < kKernKinds; ++k) launches_by_kind[k] = c[k];
    });
}

int afar_solver_profile(afar_solver* h, int32_t reps, double* ms, int32_t* count) {
    return guard([&] {
        need(h && ms && count && reps > 0, "bad arguments");
        int c[kKernKinds];
        h->dev->profile(ms, c, reps);
        for (int k = 0; k < kKernKinds; ++k) count[k] = c[k];
    });
}

int afar_solver_sizes(afar_solver* h, int64_t* out) {
    return guard([&] {
        need(h && out, "NULL argument");
        const DevLayout& L = h->dev->layout();
        out[0] = L.nc;
        out[1] = L.nf;
        out[2] = L.nb;
        out[3] = L.maxS;
        out[4] = L.maxC;
        out[5] = L.tile;
        out[6] = L.ntiles;
        out[7] = L.maxHalo;
    });
}

int afar_solver_run_stage(afar_solver* h, int32_t upto, double dt, int32_t stage) {
    return guard([&] {
        need(h != nullptr && upto >= 0 && upto <= 3 && stage >= 0 && stage < 3, "bad arguments");
        h->dev->run_stage_kernels(upto, dt, stage);
        h->dev->synchronize();
    });
}

int afar_solver_fetch(afar_solver* h, const char* name, void* out, size_t bytes) {
    return guard([&] {
        need(h && name && out, "NULL argument");
        h->dev->debug_fetch(name, out, bytes);
    });
}

void afar_solver_destroy(afar_solver* h) {
    std::lock_guard<std::mutex> lk(runtime_mutex());  // device frees
    delete h;
}

// ---------------------------------------------------------------- seam 2

// Counts arrive as 64-bit integers; the device layout indexes cells and faces
// with 32-bit ids (the reference's int), so larger meshes are refused.
static void need_id_range(int64_t nc, int64_t nf, int64_t nb) {
    if (nc + nb > INT32_MAX || nf + nb > INT32_MAX)
        throw Error(AFAR_ERR_LIMIT, "mesh too large for 32-bit cell and face ids");
}

int afar_set_device_field_data(afar_device** dev, const afar_field_pointer* host, int64_t cellsNum,
                               int64_t facesNum, int64_t totalBoundaryFacesNum) {
    return guard([&] {
        need(dev && host, "NULL argument");
        need(cellsNum > 0 && facesNum >= 0 && totalBoundaryFacesNum >= 0, "bad sizes");
        need_id_range(cellsNum, facesNum, totalBoundaryFacesNum);
        need(host->CELL && host->FACE && host->cellVolume && host->FD && host->cFD, "geometry/state arrays missing");
        auto d = std::make_unique<afar_device>();
        RefArrays& ra = d->ra;
        const size_t nc = cellsNum, nf = facesNum, nb = totalBoundaryFacesNum;
        ra.nc = (int)cellsNum;
        ra.nf = (int)facesNum;
        ra.nb = (int)totalBoundaryFacesNum;
        ra.CELL.assign(host->CELL, host->CELL + 2 * (nc + nb));
        ra.FACE.assign(host->FACE, host->FACE + 5 * (nf + nb));
        ra.cellVolume.assign(host->cellVolume, host->cellVolume + nc);
        ra.FD.assign(host->FD, host->FD + 4 * (nc + nb));
        ra.cFD.assign(host->cFD, host->cFD + 4 * nc);
        *dev = d.release();
    });
}

int afar_device_configure(afar_device* dev, int32_t precision, int32_t tile) {
    return guard([&] {
        need(dev != nullptr, "NULL device");
        need(precision == 0 || precision == 1, "precision must be 0 or 1");
        need(tile == 64 || tile == 128 || tile == 256, "tile must be 64, 128 or 256");
        need(!dev->dev, "configure before setDeviceFieldData #2");
        dev->cfg.precision = precision;
        dev->cfg.tile = tile;
    });
}

int afar_device_set_stream(afar_device* dev, void* stream) {
    return guard([&] {
        need(dev != nullptr, "NULL device");
        need(!dev->dev, "set the stream before setDeviceFieldData #2");
        dev->cfg.stream = static_cast<hipStream_t>(stream);
    });
}

int afar_set_device_stencil_data(afar_device* dev, const afar_field_pointer* host, int64_t cellsNum,
                                 int64_t facesNum, int64_t patchesNum, int64_t totalBoundaryFacesNum,
                                 int64_t maxStencilSize, int64_t maxCompactStencilSize,
                                 int64_t maxLocalBlockStencilSize, int64_t maxCompactLocalBlockStencilSize) {
    (void)maxCompactLocalBlockStencilSize;
    return guard([&] {
        need(dev && host, "NULL argument");
        RefArrays& ra = dev->ra;
        need(cellsNum == ra.nc && facesNum == ra.nf && totalBoundaryFacesNum == ra.nb, "sizes differ from upload #1");
        need(patchesNum >= 0 && patchesNum <= INT32_MAX, "patchesNum out of range");
        need(maxStencilSize > 0 && maxStencilSize <= kMaxValidStencil, "maxStencilSize out of range");
        need(maxCompactStencilSize > 0 && maxCompactStencilSize <= maxStencilSize, "maxCompactStencilSize out of range");
        need(host->neighbour && host->boundaryFaceNeiLabel && host->boundaryFacesType && host->stencilSize &&
                 host->compactStencilSize && host->cellFaces,
             "connectivity arrays missing");
        need(host->RBFbasis || host->RBFbasis64, "RBF basis missing");
        const size_t nc = ra.nc, nf = ra.nf, nb = ra.nb;
        ra.npatch = (int)patchesNum;
        ra.maxS = (int)maxStencilSize;
        ra.maxC = (int)maxCompactStencilSize;
        ra.neighbour.assign(host->neighbour, host->neighbour + 2 * nf);
        if (host->boundaryFacesNum) ra.boundaryFacesNum.assign(host->boundaryFacesNum, host->boundaryFacesNum + patchesNum);
        ra.boundaryFaceNeiLabel.assign(host->boundaryFaceNeiLabel, host->boundaryFaceNeiLabel + nb);
        ra.boundaryFacesType.assign(host->boundaryFacesType, host->boundaryFacesType + nb);
        for (uint8_t t : ra.boundaryFacesType) need(t <= kMaxBCLabel, "boundary label out of range");
        ra.stencilSize.assign(host->stencilSize, host->stencilSize + nc);
        ra.compactStencilSize.assign(host->compactStencilSize, host->compactStencilSize + nc);
        for (size_t c = 0; c < nc; ++c)
            need(ra.stencilSize[c] <= maxStencilSize && ra.compactStencilSize[c] <= maxCompactStencilSize &&
                     ra.compactStencilSize[c] <= ra.stencilSize[c],
                 "stencil sizes exceed the declared maxima");
        ra.cellFaces.assign(host->cellFaces, host->cellFaces + (size_t)maxCompactStencilSize * nc);
        ra.stencil.assign((size_t)maxStencilSize * nc, 0);
        if (host->globalStencil) {
            ra.stencil.assign(host->globalStencil, host->globalStencil + (size_t)maxStencilSize * nc);
        } else {
            need(host->localStencil && host->extendStencil, "stencil tables missing");
            need(maxLocalBlockStencilSize >= 128 && maxLocalBlockStencilSize <= INT32_MAX,
                 "maxLocalBlockStencilSize out of range");
            global_stencil_from_blocks((int)nc, (int)maxStencilSize, (int)maxLocalBlockStencilSize, ra.stencilSize.data(),
                                       host->localStencil, host->extendStencil, ra.stencil.data());
        }
        for (size_t c = 0; c < nc; ++c)
            for (int i = 0; i < ra.stencilSize[c]; ++i) {
                const int id = ra.stencil[(size_t)maxStencilSize * c + i];
                need(id >= 0 && id < (int)(nc + nb), "stencil id out of range");
            }
        const size_t nrbf = (size_t)2 * (maxStencilSize + 1) * nf;
        if (host->RBFbasis64)
            ra.RBFbasis.assign(host->RBFbasis64, host->RBFbasis64 + nrbf);
        else {
            ra.RBFbasis.resize(nrbf);
            for (size_t i = 0; i < nrbf; ++i) ra.RBFbasis[i] = host->RBFbasis[i];
        }
        ra.FSHIFT.assign(2 * nf, 0.0);
        if (host->faceShift) ra.FSHIFT.assign(host->faceShift, host->faceShift + 2 * nf);
        ra.BVALUE.assign(4 * nb, 0.0);
        if (host->boundaryValue) ra.BVALUE.assign(host->boundaryValue, host->boundaryValue + 4 * nb);
        for (size_t f = 0; f < nf; ++f)
            need(ra.neighbour[2 * f] >= 0 && ra.neighbour[2 * f] < (int)nc && ra.neighbour[2 * f + 1] >= 0 &&
                     ra.neighbour[2 * f + 1] < (int)nc,
                 "neighbour out of range");
        for (size_t c = 0; c < nc; ++c)
            for (int i = 0; i < ra.compactStencilSize[c]; ++i) {
                const int sf = ra.cellFaces[(size_t)maxCompactStencilSize * c + i];
                const int f = (sf > 0 ? sf : -sf) - 1;
                need(sf != 0 && f < (int)(nf + nb), "cellFaces entry out of range");
            }
        for (size_t b = 0; b < nb; ++b)
            need(ra.boundaryFaceNeiLabel[b] >= 0 && ra.boundaryFaceNeiLabel[b] < (int)nc, "boundary cell out of range");
        dev->dev = make_device_solver(ra, dev->cfg);
    });
}

int afar_free_field_data(afar_device* dev) {
    return guard([&] {
        std::lock_guard<std::mutex> lk(runtime_mutex());  // device frees
        delete dev;
    });
}

int afar_adjust_time_step(afar_device* dev, int64_t facesNum, double R, double Cv, double CFL, double* deltaT) {
    return guard([&] {
        need(dev && dev->dev && deltaT, "device not initialised");
        need(facesNum == dev->ra.nf, "facesNum differs from the upload");
        dev->dev->set_constants(R, Cv, CFL);
        *deltaT = dev->dev->compute_dt();
    });
}

int afar_gpu_evolve(afar_device* dev, double R, double Cv, double deltaT, int64_t cellsNum, int64_t facesNum,
                    int64_t totalBoundaryFacesNum, int64_t maxStencilSize, int64_t maxCompactStencilSize,
                    int64_t maxLocalBlockStencilSize, int64_t maxCompactLocalBlockStencilSize) {
    (void)maxLocalBlockStencilSize;
    (void)maxCompactLocalBlockStencilSize;
    return guard([&] {
        need(dev && dev->dev, "device not initialised");
        need(cellsNum == dev->ra.nc && facesNum == dev->ra.nf && totalBoundaryFacesNum == dev->ra.nb &&
                 maxStencilSize == dev->ra.maxS && maxCompactStencilSize == dev->ra.maxC,
             "sizes differ from the upload");
        dev->dev->set_constants(R, Cv, dev->dev->config().CFL);
        dev->dev->evolve(deltaT);
        dev->dev->synchronize();
    });
}

int afar_copy_field_data_device_to_host(afar_field_pointer* host, afar_device* dev, int64_t cellsNum,
                                        int64_t totalBoundaryFacesNum) {
    return guard([&] {
        need(dev && dev->dev && host, "device not initialised");
        need(cellsNum == dev->ra.nc && totalBoundaryFacesNum == dev->ra.nb, "sizes differ from the upload");
        dev->dev->copy_out(host->FD, host->shockIndicator);
    });
}

int afar_adjust_time_step_noalloc(const double* FD, const double* CELL, const double* FACE, const int32_t* neighbour,
                                  const double* FSHIFT, int64_t facesNum, double R, double Cv, double CFL,
                                  double* deltaT, void* stream) {
    return guard([&] {
        need(FD && CELL && FACE && deltaT && (facesNum == 0 || neighbour), "NULL device array");
        need(facesNum >= 0, "facesNum must not be negative");
        if (facesNum > INT32_MAX) throw Error(AFAR_ERR_LIMIT, "more faces than 32-bit face ids address");
        adjust_time_step_noalloc(FD, CELL, FACE, neighbour, FSHIFT, (int)facesNum, R, Cv, CFL, deltaT,
                                 static_cast<hipStream_t>(stream));
    });
}

// ---------------------------------------------------------------- diagnostics

int afar_selftest_math(const double* x, double* out, int32_t n) {
    return guard([&] {
        need(n >= 0 && x && out, "bad arguments");
        std::lock_guard<std::mutex> lk(runtime_mutex());  // allocations, a device-wide wait, frees
        double *dx = nullptr, *dout = nullptr;
        check_hip(hipMalloc(&dx, sizeof(double) * (n + 1)), "hipMalloc");
        hipError_t e = hipMalloc(&dout, sizeof(double) * 4 * (n + 1));
        if (e != hipSuccess) {
            (void)hipFree(dx);
            check_hip(e, "hipMalloc");
        }
        e = hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            launch_math_selftest(dx, dout, n, nullptr);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(double) * 4 * n, hipMemcpyDeviceToHost);
        (void)hipFree(dx);
        (void)hipFree(dout);
        check_hip(e, "selftest");
    });
}

int afar_selftest_math_f32(const float* x, float* out, int32_t n) {
    return guard([&] {
        need(n >= 0 && x && out, "bad arguments");
        std::lock_guard<std::mutex> lk(runtime_mutex());  // allocations, a device-wide wait, frees
        float *dx = nullptr, *dout = nullptr;
        check_hip(hipMalloc(&dx, sizeof(float) * (n + 1)), "hipMalloc");
        hipError_t e = hipMalloc(&dout, sizeof(float) * 4 * (n + 1));
        if (e != hipSuccess) {
            (void)hipFree(dx);
            check_hip(e, "hipMalloc");
        }
        e = hipMemcpy(dx, x, sizeof(float) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            launch_math_selftest_f32(dx, dout, n, nullptr);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(float) * 4 * n, hipMemcpyDeviceToHost);
        (void)hipFree(dx);
        (void)hipFree(dout);
        check_hip(e, "selftest");
    });
}

int afar_host_math_f32(const float* x, float* out, int32_t n) {
    return guard([&] {
        need(n >= 0 && x && out, "bad arguments");
        for (int i = 0; i < n; ++i) {
            const float v = x[i];
            out[4 * i] = afar_expf(v);
            out[4 * i + 1] = afar_erfcf(v);
            out[4 * i + 2] = sqrtf(v < 0 ? -v : v);
            out[4 * i + 3] = 1.0f / v;
        }
    });
}

int afar_host_math(const double* x, double* out, int32_t n) {
    return guard([&] {
        need(n >= 0 && x && out, "bad arguments");
        for (int i = 0; i < n; ++i) {
            const double v = x[i];
            out[4 * i] = afar_exp(v);
            out[4 * i + 1] = afar_erfc(v);
            out[4 * i + 2] = sqrt(v < 0 ? -v : v);
            out[4 * i + 3] = 1.0 / v;
        }
    });
}

int afar_device_info(char* name, int32_t cap, int32_t* cus, int64_t* hbm_bytes) {
    return guard([&] {
        int dev = 0;
        check_hip(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t p;
        check_hip(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
        if (name && cap > 0) {
            snprintf(name, cap, "%s (%s)", p.name, p.gcnArchName);
        }
        if (cus) *cus = p.multiProcessorCount;
        if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
    });
}

}  // extern "C"
