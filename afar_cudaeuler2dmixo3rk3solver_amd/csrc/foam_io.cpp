// OpenFOAM case I/O for one-cell-thick 2D cases (ASCII format):
//
//  read_polymesh_2d   constant/polyMesh/{points,faces,owner,neighbour,boundary}
//                     -> the 2D polygons (the cells' faces on the lower z plane),
//                     the boundary edges of every non-empty patch, the depth.
//  write_polymesh_2d  a Mesh as a one-cell-thick polyMesh: two point layers,
//                     the mesh's internal faces (already in OpenFOAM's
//                     upper-triangular order), its boundary faces patch by
//                     patch, then one "frontAndBack" empty patch.
//  read_field         internalField of a 0/ (or time) field file, uniform or
//                     nonuniform, and the uniform fixedValue of each patch.
//
// The reference reads the same files through Foam::fvMesh and volFields
// (src/CUDA/cudaEuler2DmixO3RK3/basicCUDAsolver.cpp:9-54,122-258); patch
// indices count every polyMesh patch, empty ones included, as the
// reference's patch<i> names do (basicCUDAsolver.cpp:171-172).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <map>
#include <sstream>
#include <unordered_map>

#include "afar_host.h"
#include "foam_io.h"

namespace afar {

namespace {

std::string slurp(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(4, "cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// Tokens of an OpenFOAM file with comments removed: words, numbers and the
// punctuation ( ) { } ; as single-character tokens.
class Lexer {
   public:
    explicit Lexer(const std::string& text) : s_(text) {}
    bool next(std::string& tok) {
        skip();
        if (i_ >= s_.size()) return false;
        const char c = s_[i_];
        if (strchr("(){};[]", c)) {
            tok.assign(1, c);
            ++i_;
            return true;
        }
        if (c == '"') {
            const size_t j = s_.find('"', i_ + 1);
            if (j == std::string::npos) throw Error(4, "unterminated string");
            tok = s_.substr(i_ + 1, j - i_ - 1);
            i_ = j + 1;
            return true;
        }
        const size_t b = i_;
        while (i_ < s_.size() && !isspace((unsigned char)s_[i_]) && !strchr("(){};[]\"", s_[i_])) ++i_;
        tok = s_.substr(b, i_ - b);
        return true;
    }
    std::string need(const char* what) {
        std::string t;
        if (!next(t)) throw Error(4, std::string("unexpected end of file, expected ") + what);
        return t;
    }
    void expect(const char* t) {
        const std::string g = need(t);
        if (g != t) throw Error(4, std::string("expected '") + t + "', found '" + g + "'");
    }
    // Skips a FoamFile { ... } header if present; leaves the cursor at the first data token.
    void skip_header() {
        const size_t save = i_;
        std::string t;
        if (next(t) && t == "FoamFile") {
            expect("{");
            int depth = 1;
            while (depth > 0) {
                t = need("}");
                if (t == "{") ++depth;
                if (t == "}") --depth;
            }
        } else {
            i_ = save;
        }
    }
    double number() {
        const std::string t = need("number");
        char* end = nullptr;
        const double v = strtod(t.c_str(), &end);
        if (end == t.c_str() || *end) throw Error(4, "expected a number, found '" + t + "'");
        return v;
    }
    long label() {
        const std::string t = need("integer");
        char* end = nullptr;
        const long v = strtol(t.c_str(), &end, 10);
        if (end == t.c_str() || *end) throw Error(4, "expected an integer, found '" + t + "'");
        return v;
    }
    // "N(" where the count may be glued to the parenthesis ("4(1 2 3 4)").
    long list_count() {
        std::string t = need("list size");
        const size_t p = t.find('(');
        long n;
        if (p != std::string::npos) throw Error(4, "malformed list size");
        char* end = nullptr;
        n = strtol(t.c_str(), &end, 10);
        if (end == t.c_str() || *end) throw Error(4, "expected a list size, found '" + t + "'");
        return n;
    }

   private:
    void skip() {
        for (;;) {
            while (i_ < s_.size() && isspace((unsigned char)s_[i_])) ++i_;
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
                while (i_ < s_.size() && s_[i_] != '\n') ++i_;
                continue;
            }
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
                const size_t j = s_.find("*/", i_ + 2);
                i_ = j == std::string::npos ? s_.size() : j + 2;
                continue;
            }
            return;
        }
    }
    const std::string& s_;
    size_t i_ = 0;
};

std::vector<double> read_points(const std::string& path) {
    const std::string text = slurp(path);
    Lexer L(text);
    L.skip_header();
    const long n = L.list_count();
    L.expect("(");
    std::vector<double> p(3 * (size_t)n);
    for (long i = 0; i < n; ++i) {
        L.expect("(");
        for (int k = 0; k < 3; ++k) p[3 * i + k] = L.number();
        L.expect(")");
    }
    L.expect(")");
    return p;
}

void read_faces(const std::string& path, std::vector<int64_t>& ptr, std::vector<int32_t>& pts) {
    const std::string text = slurp(path);
    Lexer L(text);
    L.skip_header();
    const long n = L.list_count();
    L.expect("(");
    ptr.assign(1, 0);
    ptr.reserve(n + 1);
    for (long i = 0; i < n; ++i) {
        const long k = L.list_count();
        L.expect("(");
        for (long j = 0; j < k; ++j) pts.push_back((int32_t)L.label());
        L.expect(")");
        ptr.push_back((int64_t)pts.size());
    }
    L.expect(")");
}

std::vector<int32_t> read_labels(const std::string& path) {
    const std::string text = slurp(path);
    Lexer L(text);
    L.skip_header();
    const long n = L.list_count();
    L.expect("(");
    std::vector<int32_t> v(n);
    for (long i = 0; i < n; ++i) v[i] = (int32_t)L.label();
    L.expect(")");
    return v;
}

struct FoamPatch {
    std::string name, type;
    long nFaces = 0, startFace = 0;
};

std::vector<FoamPatch> read_boundary(const std::string& path) {
    const std::string text = slurp(path);
    Lexer L(text);
    L.skip_header();
    const long n = L.list_count();
    L.expect("(");
    std::vector<FoamPatch> out(n);
    for (long i = 0; i < n; ++i) {
        out[i].name = L.need("patch name");
        L.expect("{");
        std::string key;
        for (;;) {
            key = L.need("}");
            if (key == "}") break;
            std::vector<std::string> val;
            std::string t;
            int depth = 0;
            for (;;) {
                t = L.need(";");
                if (t == "(") ++depth;
                if (t == ")") --depth;
                if (t == ";" && depth == 0) break;
                val.push_back(t);
            }
            if (key == "type" && !val.empty()) out[i].type = val[0];
            if (key == "nFaces" && !val.empty()) out[i].nFaces = atol(val[0].c_str());
            if (key == "startFace" && !val.empty()) out[i].startFace = atol(val[0].c_str());
        }
    }
    L.expect(")");
    return out;
}

}  // namespace

PolyMesh read_polymesh_2d(const std::string& case_dir, std::vector<std::string>* patch_names,
                          std::vector<std::string>* patch_types) {
    const std::string d = case_dir + "/constant/polyMesh/";
    const std::vector<double> P = read_points(d + "points");
    std::vector<int64_t> fptr;
    std::vector<int32_t> fpts;
    read_faces(d + "faces", fptr, fpts);
    const std::vector<int32_t> owner = read_labels(d + "owner");
    const std::vector<int32_t> neighbour = read_labels(d + "neighbour");
    const std::vector<FoamPatch> bnd = read_boundary(d + "boundary");
    const int64_t np3 = (int64_t)P.size() / 3, nfaces = (int64_t)fptr.size() - 1;
    if ((int64_t)owner.size() != nfaces) throw Error(4, "owner size differs from the number of faces");
    if (np3 == 0 || nfaces == 0) throw Error(4, "empty polyMesh");
    int32_t ncell = 0;
    for (int32_t o : owner) ncell = std::max(ncell, o + 1);
    for (int32_t n : neighbour) ncell = std::max(ncell, n + 1);

    double zmin = HUGE_VAL, zmax = -HUGE_VAL;
    for (int64_t i = 0; i < np3; ++i) {
        zmin = std::min(zmin, P[3 * i + 2]);
        zmax = std::max(zmax, P[3 * i + 2]);
    }
    const double depth = zmax - zmin;
    if (!(depth > 0)) throw Error(4, "the polyMesh is not one cell thick in z");
    const double tol = 1e-9 * depth;
    // points of the lower plane -> 2D ids
    std::vector<int32_t> id2(np3, -1);
    PolyMesh pm;
    for (int64_t i = 0; i < np3; ++i)
        if (fabs(P[3 * i + 2] - zmin) <= tol) {
            id2[i] = (int32_t)pm.pts.size();
            pm.pts.push_back(Vec2{P[3 * i], P[3 * i + 1]});
        } else if (fabs(P[3 * i + 2] - zmax) > tol) {
            throw Error(4, "a point lies between the two z planes: the mesh is not one cell thick");
        }
    // each cell's face on the lower plane is its 2D polygon
    std::vector<int64_t> cface(ncell, -1);
    for (int64_t f = 0; f < nfaces; ++f) {
        bool lower = true;
        for (int64_t k = fptr[f]; k < fptr[f + 1] && lower; ++k) lower = id2[fpts[k]] >= 0;
        if (!lower) continue;
        if (f < (int64_t)neighbour.size()) throw Error(4, "an internal face lies in the z plane");
        const int32_t c = owner[f];
        if (cface[c] >= 0) throw Error(4, "a cell has two faces on the lower plane");
        cface[c] = f;
    }
    pm.cell_ptr.assign(1, 0);
    for (int32_t c = 0; c < ncell; ++c) {
        const int64_t f = cface[c];
        if (f < 0) throw Error(4, "cell " + std::to_string(c) + " has no face on the lower plane");
        std::vector<int32_t> poly;
        for (int64_t k = fptr[f]; k < fptr[f + 1]; ++k) poly.push_back(id2[fpts[k]]);
        double a2 = 0.0;  // twice the signed area
        for (size_t k = 0; k < poly.size(); ++k) {
            const Vec2 p = pm.pts[poly[k]], q = pm.pts[poly[(k + 1) % poly.size()]];
            a2 += p.x * q.y - q.x * p.y;
        }
        if (a2 < 0) std::reverse(poly.begin(), poly.end());  // counter-clockwise
        pm.cell_pts.insert(pm.cell_pts.end(), poly.begin(), poly.end());
        pm.cell_ptr.push_back((int64_t)pm.cell_pts.size());
    }
    // boundary edges of the non-empty patches; patch index = polyMesh index
    pm.patches.resize(bnd.size());
    for (size_t p = 0; p < bnd.size(); ++p) {
        pm.patches[p].name = bnd[p].name;
        if (patch_names) patch_names->push_back(bnd[p].name);
        if (patch_types) patch_types->push_back(bnd[p].type);
        if (bnd[p].type == "empty") continue;
        if (bnd[p].type == "cyclic" || bnd[p].type == "cyclicAMI")
            throw Error(4, "cyclic patches are not read (periodic meshes come from the generators)");
        for (long f = bnd[p].startFace; f < bnd[p].startFace + bnd[p].nFaces; ++f) {
            if (f < 0 || f >= nfaces) throw Error(4, "patch faces out of range");
            int32_t e[2], ne = 0;
            for (int64_t k = fptr[f]; k < fptr[f + 1]; ++k)
                if (id2[fpts[k]] >= 0) {
                    if (ne == 2) throw Error(4, "a boundary face has more than two lower points");
                    e[ne++] = id2[fpts[k]];
                }
            if (ne != 2) throw Error(4, "a boundary face of patch " + bnd[p].name + " is not a side face");
            pm.bedges.push_back({e[0], e[1]});
            pm.bedge_patch.push_back((int32_t)p);
        }
    }
    pm.depth = depth;
    pm.period = Vec2{0, 0};
    return pm;
}

void write_polymesh_2d(const Mesh& m, const std::string& case_dir) {
    if (m.period.x != 0.0 || m.period.y != 0.0) throw Error(4, "periodic meshes cannot be written as polyMesh");
    const int nc = m.nc, nf = m.nf, nb = m.nb, N = m.npts;
    std::vector<Vec2> pt(N);
    std::vector<char> have(N, 0);
    for (size_t k = 0; k < m.cell_pts.size(); ++k) {
        pt[m.cell_pts[k]] = m.cell_xy[k];
        have[m.cell_pts[k]] = 1;
    }
    for (int i = 0; i < N; ++i)
        if (!have[i]) throw Error(4, "a point belongs to no cell");
    // point id of a face end: the owner's vertex at that position
    auto vid = [&](int cell, Vec2 p) -> int32_t {
        for (int32_t k = m.cell_ptr[cell]; k < m.cell_ptr[cell + 1]; ++k)
            if (m.cell_xy[k].x == p.x && m.cell_xy[k].y == p.y) return m.cell_pts[k];
        throw Error(4, "face end is not a vertex of its owner");
    };
    const std::string d = case_dir + "/constant/polyMesh/";
    auto header = [](FILE* f, const char* cls, const char* obj) {
        fprintf(f,
                "FoamFile\n{\n    version     2.0;\n    format      ascii;\n    class       %s;\n"
                "    location    \"constant/polyMesh\";\n    object      %s;\n}\n\n",
                cls, obj);
    };
    auto open = [&](const char* name) {
        FILE* f = fopen((d + name).c_str(), "w");
        if (!f) throw Error(4, "cannot write " + d + name + " (does constant/polyMesh exist?)");
        return f;
    };
    FILE* f = open("points");
    header(f, "vectorField", "points");
    fprintf(f, "%d\n(\n", 2 * N);
    for (int layer = 0; layer < 2; ++layer)
        for (int i = 0; i < N; ++i)
            fprintf(f, "(%.17g %.17g %.17g)\n", pt[i].x, pt[i].y, layer ? m.depth : 0.0);
    fprintf(f, ")\n");
    fclose(f);

    const int nfaces = nf + nb + 2 * nc;
    f = open("faces");
    header(f, "faceList", "faces");
    fprintf(f, "%d\n(\n", nfaces);
    for (int g = 0; g < nf + nb; ++g) {  // side faces: normal out of the owner
        const int a = vid(m.own[g], m.fp0[g]), b = vid(m.own[g], m.fp1[g]);
        fprintf(f, "4(%d %d %d %d)\n", a, b, b + N, a + N);
    }
    for (int c = 0; c < nc; ++c) {  // lower plane, normal -z
        fprintf(f, "%d(", m.cell_ptr[c + 1] - m.cell_ptr[c]);
        for (int32_t k = m.cell_ptr[c + 1] - 1; k >= m.cell_ptr[c]; --k)
            fprintf(f, k == m.cell_ptr[c] ? "%d" : "%d ", m.cell_pts[k]);
        fprintf(f, ")\n");
    }
    for (int c = 0; c < nc; ++c) {  // upper plane, normal +z
        fprintf(f, "%d(", m.cell_ptr[c + 1] - m.cell_ptr[c]);
        for (int32_t k = m.cell_ptr[c]; k < m.cell_ptr[c + 1]; ++k)
            fprintf(f, k + 1 == m.cell_ptr[c + 1] ? "%d" : "%d ", m.cell_pts[k] + N);
        fprintf(f, ")\n");
    }
    fprintf(f, ")\n");
    fclose(f);

    f = open("owner");
    header(f, "labelList", "owner");
    fprintf(f, "%d\n(\n", nfaces);
    for (int g = 0; g < nf + nb; ++g) fprintf(f, "%d\n", m.own[g]);
    for (int layer = 0; layer < 2; ++layer)
        for (int c = 0; c < nc; ++c) fprintf(f, "%d\n", c);
    fprintf(f, ")\n");
    fclose(f);

    f = open("neighbour");
    header(f, "labelList", "neighbour");
    fprintf(f, "%d\n(\n", nf);
    for (int g = 0; g < nf; ++g) fprintf(f, "%d\n", m.nei[g]);
    fprintf(f, ")\n");
    fclose(f);

    f = open("boundary");
    header(f, "polyBoundaryMesh", "boundary");
    fprintf(f, "%zu\n(\n", m.patches.size() + 1);
    for (size_t p = 0; p < m.patches.size(); ++p)
        fprintf(f, "    %s\n    {\n        type patch;\n        nFaces %d;\n        startFace %d;\n    }\n",
                m.patches[p].name.c_str(), m.patch_size[p], nf + m.patch_start[p]);
    fprintf(f, "    frontAndBack\n    {\n        type empty;\n        nFaces %d;\n        startFace %d;\n    }\n",
            2 * nc, nf + nb);
    fprintf(f, ")\n");
    fclose(f);
}

FoamField read_field(const std::string& path, int ncomp, int64_t ncells, const std::vector<std::string>& patches) {
    const std::string text = slurp(path);
    Lexer L(text);
    L.skip_header();
    FoamField out;
    out.internal.assign((size_t)ncomp * ncells, 0.0);
    out.patch_value.assign((size_t)ncomp * patches.size(), 0.0);
    out.patch_fixed.assign(patches.size(), 0);
    auto read_value = [&](double* v) {  // scalar or (x y z): first ncomp components
        if (ncomp == 1) {
            v[0] = L.number();
            return;
        }
        L.expect("(");
        double t[3];
        for (int k = 0; k < 3; ++k) t[k] = L.number();
        L.expect(")");
        for (int k = 0; k < ncomp; ++k) v[k] = t[k];
    };
    auto read_uniform_or_list = [&](std::vector<double>& dst, int64_t n) {
        const std::string kind = L.need("uniform or nonuniform");
        if (kind == "uniform") {
            double v[3];
            read_value(v);
            for (int64_t i = 0; i < n; ++i)
                for (int k = 0; k < ncomp; ++k) dst[(size_t)ncomp * i + k] = v[k];
        } else if (kind == "nonuniform") {
            L.need("List<...>");
            const long cnt = L.list_count();
            if (cnt != n) throw Error(4, path + ": list size " + std::to_string(cnt) + " differs from " +
                                             std::to_string(n));
            L.expect("(");
            for (int64_t i = 0; i < n; ++i) read_value(&dst[(size_t)ncomp * i]);
            L.expect(")");
        } else {
            throw Error(4, path + ": unknown field form " + kind);
        }
        L.expect(";");
    };
    std::string key;
    while (L.next(key)) {
        if (key == "internalField") {
            read_uniform_or_list(out.internal, ncells);
            out.have_internal = true;
        } else if (key == "boundaryField") {
            L.expect("{");
            for (;;) {
                const std::string name = L.need("}");
                if (name == "}") break;
                L.expect("{");
                int pi = -1;
                for (size_t p = 0; p < patches.size(); ++p)
                    if (patches[p] == name) pi = (int)p;
                std::string type;
                for (;;) {
                    const std::string k = L.need("}");
                    if (k == "}") break;
                    if (k == "value" && pi >= 0) {
                        std::vector<double> tmp(ncomp);
                        const std::string kind = L.need("uniform");
                        if (kind == "uniform") {
                            read_value(tmp.data());
                            L.expect(";");
                            if (type == "fixedValue") {
                                for (int c = 0; c < ncomp; ++c) out.patch_value[(size_t)ncomp * pi + c] = tmp[c];
                                out.patch_fixed[pi] = 1;
                            }
                            continue;
                        }
                    }
                    // any other entry: skip to its ';' at depth 0
                    std::string t;
                    int depth = 0;
                    bool first = true;
                    for (;;) {
                        t = L.need(";");
                        if (t == "(") ++depth;
                        if (t == ")") --depth;
                        if (t == ";" && depth == 0) break;
                        if (k == "type" && first) type = t;
                        first = false;
                    }
                }
            }
        } else if (key == "dimensions") {
            while (L.need(";") != ";") {
            }
        } else {
            // other top-level entries: skip to ';'
            std::string t;
            int depth = 0;
            for (;;) {
                if (!L.next(t)) break;
                if (t == "(" || t == "{") ++depth;
                if (t == ")" || t == "}") --depth;
                if (t == ";" && depth == 0) break;
            }
        }
    }
    return out;
}

}  // namespace afar
