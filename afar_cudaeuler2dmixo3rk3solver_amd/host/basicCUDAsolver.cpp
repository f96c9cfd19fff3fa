#include "basicCUDAsolver.h"

#include <math.h>
#include <sys/stat.h>

#include <cstdio>
#include <fstream>
#include <sstream>

namespace afarhost {

namespace {

void check(int rc, const char* what) {
    if (rc != AFAR_OK) throw solverError(rc, std::string(what) + ": " + afar_last_error());
}

std::vector<double> vec_of(const dictionary& d, const std::string& key, size_t n) {
    std::vector<double> v;
    for (const auto& t : d.tokens(key))
        if (t != "(" && t != ")") v.push_back(strtod(t.c_str(), nullptr));
    if (v.size() != n) throw solverError(AFAR_ERR_INVALID, key + " needs " + std::to_string(n) + " values");
    return v;
}

}  // namespace

// ------------------------------------------------------------------ runTime

runTime::runTime(const std::string& caseDir) : dir_(caseDir) {
    control_ = dictionary::fromFile(caseDir + "/system/controlDict");
    t_ = control_.lookupOrDefault("startTime", 0.0);
    end_ = control_.lookupScalar("endTime");
    dt_ = control_.lookupScalar("deltaT");
    write_interval_ = control_.lookupScalar("writeInterval");
    next_write_ = t_ + write_interval_;
}

void runTime::setDeltaT(double dt) {
    double d = dt;
    const double to_write = next_write_ - t_, to_end = end_ - t_;
    if (to_write > 0 && d > to_write) d = to_write;
    if (to_end > 0 && d > to_end) d = to_end;
    dt_ = d;
}

runTime& runTime::operator++() {
    t_ += dt_;
    ++index_;
    const double tol = 1e-12 * (write_interval_ > 0 ? write_interval_ : 1.0);
    write_now_ = false;
    if (write_interval_ > 0 && t_ >= next_write_ - tol) {
        write_now_ = true;
        while (next_write_ <= t_ + tol) next_write_ += write_interval_;
    }
    if (t_ >= end_ - tol) write_now_ = true;
    return *this;
}

std::string runTime::timeName() const {
    char buf[64];
    snprintf(buf, sizeof buf, "%.6g", t_);
    return buf;
}

// ------------------------------------------------------------------ caseMesh

static bool file_exists(const std::string& path) {
    FILE* f = fopen(path.c_str(), "r");
    if (f) fclose(f);
    return f != nullptr;
}

caseMesh::caseMesh(const std::string& caseDir) {
    schemes_ = dictionary::fromFile(caseDir + "/system/fvSchemes");
    const bool haveMd = file_exists(caseDir + "/system/afarMeshDict");
    const bool havePoly = file_exists(caseDir + "/constant/polyMesh/faces");
    const dictionary md = haveMd ? dictionary::fromFile(caseDir + "/system/afarMeshDict") : dictionary();
    if (havePoly && (!haveMd || md.lookupWord("mesh.type") == "polyMesh")) {
        // an OpenFOAM case: constant/polyMesh and the 0/ fields, as the reference reads them
        // (basicCUDAsolver.cpp:9-54,122-258); renumbering is left to the case (renumberMesh)
        // unless afarMeshDict asks for it
        const bool renumber = haveMd && md.lookupOrDefault("mesh.renumber", false);
        check(afar_mesh_read_foam(caseDir.c_str(), renumber ? 1 : 0, &mesh_), "read polyMesh");
        int64_t s[9];
        check(afar_mesh_sizes(mesh_, s), "afar_mesh_sizes");
        nc_ = s[0];
        W0_.assign(4 * (size_t)nc_, 0.0);
        patchValues_.assign(4 * (size_t)s[8], 0.0);
        std::vector<int32_t> fixed(s[8], 0);
        check(afar_foam_read_fields(mesh_, caseDir.c_str(), "0", W0_.data(), patchValues_.data(), fixed.data()),
              "read 0/ fields");
        patchFixed_.assign(fixed.begin(), fixed.end());
        return;
    }
    const dictionary& m = md.subDict("mesh");
    const std::string type = m.lookupWord("type");
    const long seed = m.lookupOrDefault("seed", 0L);
    const double jitter = m.lookupOrDefault("jitter", 0.0);
    if (type == "corner") {
        const auto inflow = vec_of(m, "inflow", 4);
        check(afar_mesh_generate_corner((int32_t)m.lookupLabel("nx"), (int32_t)m.lookupLabel("ny"),
                                        m.lookupOrDefault("lx", 1.0), m.lookupOrDefault("ly", 0.5), jitter,
                                        (uint64_t)seed, m.lookupOrDefault("cornerX", 0.5),
                                        m.lookupOrDefault("angle", 15.0), inflow.data(),
                                        (int32_t)m.lookupOrDefault("topLabel", 2L), &mesh_),
              "afar_mesh_generate_corner");
    } else if (type == "riemann") {
        check(afar_mesh_generate_riemann((int32_t)m.lookupLabel("n"), jitter, (uint64_t)seed,
                                         m.lookupOrDefault("splitP", 0.5), &mesh_),
              "afar_mesh_generate_riemann");
    } else if (type == "periodic") {
        check(afar_mesh_generate_periodic((int32_t)m.lookupLabel("n"), m.lookupOrDefault("L", 10.0), jitter,
                                          (uint64_t)seed, m.lookupOrDefault("splitP", 1.0), &mesh_),
              "afar_mesh_generate_periodic");
    } else {
        throw solverError(AFAR_ERR_INVALID, "unknown mesh type " + type);
    }
    int64_t s[9];
    check(afar_mesh_sizes(mesh_, s), "afar_mesh_sizes");
    nc_ = s[0];
    std::vector<double> c(2 * (size_t)(s[0] + s[2]));
    std::vector<int32_t> perm(nc_);
    check(afar_mesh_cells(mesh_, c.data(), nullptr), "afar_mesh_cells");
    check(afar_mesh_permutation(mesh_, perm.data()), "afar_mesh_permutation");
    centres_.assign(2 * (size_t)nc_, 0.0);
    for (int64_t i = 0; i < nc_; ++i) {
        centres_[2 * (size_t)perm[i]] = c[2 * (size_t)i];
        centres_[2 * (size_t)perm[i] + 1] = c[2 * (size_t)i + 1];
    }
    // initial conditions (the 0/ fields of the reference)
    const dictionary& ic = md.subDict("initialConditions");
    const std::string it = ic.lookupWord("type");
    const double R = schemes_.lookupScalar("CUDA.R");
    W0_.assign(4 * (size_t)nc_, 0.0);
    for (int64_t i = 0; i < nc_; ++i) {
        const double x = centres_[2 * (size_t)i], y = centres_[2 * (size_t)i + 1];
        double* w = &W0_[4 * (size_t)i];
        if (it == "uniform") {
            const auto U = vec_of(ic, "U", 2);
            w[0] = ic.lookupScalar("p");
            w[1] = U[0];
            w[2] = U[1];
            w[3] = ic.lookupScalar("T");
        } else if (it == "riemann") {
            // four-quadrant states (rho, u, v, p), quadrants counter-clockwise from x > .5, y > .5
            static const double q[4][4] = {{1.5, 0.0, 0.0, 1.5}, {0.5323, 1.206, 0.0, 0.3},
                                           {0.138, 1.206, 1.206, 0.029}, {0.5323, 0.0, 1.206, 0.3}};
            const int k = x > 0.5 ? (y > 0.5 ? 0 : 3) : (y > 0.5 ? 1 : 2);
            w[0] = q[k][3];
            w[1] = q[k][1];
            w[2] = q[k][2];
            w[3] = q[k][3] / (q[k][0] * R);
        } else if (it == "vortex") {
            const double beta = ic.lookupOrDefault("beta", 5.0), g = ic.lookupOrDefault("gamma", 1.4);
            const double L = ic.lookupOrDefault("L", 10.0), x0 = 0.5 * L, y0 = 0.5 * L;
            double dx = fmod(x - x0 + 1.5 * L, L) - 0.5 * L, dy = fmod(y - y0 + 1.5 * L, L) - 0.5 * L;
            const double r2 = dx * dx + dy * dy, e = exp(0.5 * (1.0 - r2));
            const double Ts = 1.0 - (g - 1.0) * beta * beta / (8 * g * M_PI * M_PI) * exp(1.0 - r2);
            const double rho = pow(Ts, 1.0 / (g - 1.0)), p = rho * Ts;
            w[0] = p;
            w[1] = 1.0 - beta / (2 * M_PI) * e * dy;
            w[2] = 1.0 + beta / (2 * M_PI) * e * dx;
            w[3] = p / (rho * R);
        } else {
            throw solverError(AFAR_ERR_INVALID, "unknown initial condition " + it);
        }
    }
}

caseMesh::~caseMesh() { afar_mesh_destroy(mesh_); }

// ------------------------------------------------------------------ schemes

afar_schemes readSchemes(const dictionary& fv, const dictionary& control) {
    afar_schemes s;
    afar_schemes_default(&s);
    // basicCUDAsolver.cpp:55-68; d2GaussQuadrature.C:12-13
    s.R = fv.lookupScalar("CUDA.R");
    s.Cv = fv.lookupScalar("CUDA.Cv");
    s.maxCo = control.lookupScalar("maxCo");
    s.shapeFactor = fv.lookupOrDefault("shapeFactor", 10.0);
    s.recalcBasis = fv.lookupOrDefault("CUDA.recalcBasis", true);
    const std::string st = fv.lookupWord("CUDA.stencilType");
    if (st == "faceNeighbour")
        s.stencilType = 0;
    else if (st == "edgeNeighbour")
        s.stencilType = 1;
    else
        throw solverError(AFAR_ERR_INVALID,
                          "Error: Invalid stencil type setting. Valid settings are: faceNeighbour, edgeNeighbour");
    s.cOrder = (int32_t)fv.lookupOrDefault("gaussQuadrature.cOrder", 5L);
    s.fOrder = (int32_t)fv.lookupOrDefault("gaussQuadrature.fOrder", 4L);
    const std::string prec = fv.found("CUDA.precision") ? fv.lookupWord("CUDA.precision") : "fp64";
    s.precision = prec == "mixed" ? 1 : 0;
    s.exactRK = fv.lookupOrDefault("CUDA.exactRK", false);
    s.tile = (int32_t)fv.lookupOrDefault("CUDA.tile", 64L);
    s.boundaryFirstOrder = fv.lookupOrDefault("CUDA.boundaryFirstOrder", false);
    s.faceQuadrature = fv.lookupOrDefault("CUDA.faceQuadrature", false);  // extension: flux at the fOrder points
    // device cell order: grouped compact tiles (default), "mesh", or "curve" (compact tiles, ungrouped)
    const std::string order = fv.found("CUDA.cellOrder") ? fv.lookupWord("CUDA.cellOrder") : "tiles";
    s.cellOrder = order == "mesh" ? 1 : order == "curve" ? 2 : 0;
    return s;
}

// ------------------------------------------------------------------ basicCUDAsolver

basicCUDAsolver::basicCUDAsolver(const runTime& rt, const caseMesh& mesh) {
    sc_ = readSchemes(mesh.schemes(), rt.controlDict());
    afar_mesh* m = mesh.mesh();
    // boundary labels CUDA.boundary.patch<i> for every non-empty patch, i counting all
    // patches (basicCUDAsolver.cpp:157-172); label 6 takes CUDA.boundary.value<i>, else the
    // patch's fixedValue values of the 0/ fields
    int64_t s[9];
    check(afar_mesh_sizes(m, s), "afar_mesh_sizes");
    for (int p = 0; p < s[8]; ++p) {
        int32_t nfaces = 0;
        check(afar_mesh_patch_info(m, p, nullptr, 0, &nfaces), "afar_mesh_patch_info");
        if (nfaces == 0) continue;
        const std::string key = "CUDA.boundary.patch" + std::to_string(p);
        const long label = mesh.schemes().lookupLabel(key);
        std::vector<double> val(4, 0.0);
        const std::string vkey = "CUDA.boundary.value" + std::to_string(p);
        bool haveVal = false;
        if (mesh.schemes().found(vkey)) {
            val = vec_of(mesh.schemes(), vkey, 4);
            haveVal = true;
        } else if (mesh.patchFixed(p)) {
            val = mesh.patchValue(p);
            haveVal = true;
        }
        check(afar_mesh_set_patch(m, p, (int32_t)label, label == 6 && haveVal ? val.data() : nullptr),
              "afar_mesh_set_patch");
    }
    const std::string basisDir = rt.path() + "/system";
    check(afar_mesh_precompute(m, &sc_, sc_.recalcBasis ? nullptr : basisDir.c_str()), "initStencilData");
    check(afar_solver_create(m, &sc_, mesh.initialW().data(), &solver_), "basicCUDAsolver");
    nc_ = s[0];
    nb_ = s[2];
}

basicCUDAsolver::~basicCUDAsolver() { afar_solver_destroy(solver_); }

void basicCUDAsolver::setDeltaT(runTime& rt) {
    double dt = 0.0;
    check(afar_solver_set_delta_t(solver_, &dt), "adjustTimeStep");
    rt.setDeltaT(dt);
}

void basicCUDAsolver::evolve(double deltaT) { check(afar_solver_evolve(solver_, deltaT), "GPUevolve"); }

void basicCUDAsolver::writeField(const runTime& rt) {
    std::vector<double> W(4 * (size_t)nc_);
    std::vector<int8_t> shock(nc_);
    check(afar_solver_write_field(solver_, W.data(), nullptr, shock.data()), "copyFieldDataDeviceToHost");
    const std::string dir = rt.path() + "/" + rt.timeName();
    mkdir(dir.c_str(), 0755);
    auto header = [](std::ostream& os, const char* cls, const char* obj, const char* dims) {
        os << "FoamFile\n{\n    version     2.0;\n    format      ascii;\n    class       " << cls
           << ";\n    object      " << obj << ";\n}\n\ndimensions      " << dims << ";\n\n";
    };
    auto scalar_field = [&](const char* name, const char* dims, auto value) {
        std::ofstream os(dir + "/" + name);
        header(os, "volScalarField", name, dims);
        os.precision(17);
        os << "internalField   nonuniform List<scalar>\n" << nc_ << "\n(\n";
        for (int64_t i = 0; i < nc_; ++i) os << value(i) << "\n";
        os << ")\n;\n\nboundaryField\n{\n    \".*\"\n    {\n        type            calculated;\n    }\n}\n";
    };
    scalar_field("p", "[1 -1 -2 0 0 0 0]", [&](int64_t i) { return W[4 * i]; });
    scalar_field("T", "[0 0 0 1 0 0 0]", [&](int64_t i) { return W[4 * i + 3]; });
    scalar_field("shockIndicator", "[0 0 0 0 0 0 0]", [&](int64_t i) { return (double)shock[i]; });
    std::ofstream os(dir + "/U");
    header(os, "volVectorField", "U", "[0 1 -1 0 0 0 0]");
    os.precision(17);
    os << "internalField   nonuniform List<vector>\n" << nc_ << "\n(\n";
    for (int64_t i = 0; i < nc_; ++i) os << "(" << W[4 * i + 1] << " " << W[4 * i + 2] << " 0)\n";
    os << ")\n;\n\nboundaryField\n{\n    \".*\"\n    {\n        type            calculated;\n    }\n}\n";
}

}  // namespace afarhost
