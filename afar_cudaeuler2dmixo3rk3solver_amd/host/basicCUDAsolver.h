// C++ host side above the C-ABI, mirroring the reference's solver class
// basicCUDAsolver (src/CUDA/cudaEuler2DmixO3RK3/basicCUDAsolver.h:9-53):
// same constructor role, setDeltaT / evolve / writeField, same dictionary
// keys (fvSchemes CUDA{R; Cv; stencilType; recalcBasis; boundary{patch<i>}},
// top-level shapeFactor, gaussQuadrature{cOrder; fOrder}, controlDict maxCo /
// adjustTimeStep).  Foam::Time and Foam::fvMesh are replaced by runTime and
// caseMesh below; errors are thrown as afarhost::solverError (the reference
// calls exit(-1)).
#pragma once
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "afar.h"
#include "foamDict.h"

namespace afarhost {

struct solverError : std::runtime_error {
    int code;
    solverError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// The time control the reference takes from Foam::Time and controlDict
// (cudaEuler2DmixO3RK3solver.cpp:23,43,53-69): startTime, endTime, deltaT,
// writeInterval (adjustableRunTime), adjustTimeStep, maxCo.
class runTime {
   public:
    explicit runTime(const std::string& caseDir);
    const dictionary& controlDict() const { return control_; }
    const std::string& path() const { return dir_; }
    double value() const { return t_; }
    double deltaT() const { return dt_; }
    // Sets the step and shortens it so the run lands on the next write time
    // and on endTime.
    void setDeltaT(double dt);
    bool run() const { return t_ < end_ - 1e-12 * (end_ > 0 ? end_ : 1.0); }
    runTime& operator++();
    bool write() const { return write_now_; }
    std::string timeName() const;
    long timeIndex() const { return index_; }

   private:
    std::string dir_;
    dictionary control_;
    double t_ = 0, dt_ = 0, end_ = 0, write_interval_ = 0, next_write_ = 0;
    long index_ = 0;
    bool write_now_ = false;
};

// The mesh, schemes and initial fields the reference takes from
// Foam::fvMesh and the 0/ directory: an OpenFOAM case (constant/polyMesh and
// 0/{p,U,T}, ASCII, one cell thick), or system/afarMeshDict (a generator and
// an initial condition).
class caseMesh {
   public:
    explicit caseMesh(const std::string& caseDir);
    ~caseMesh();
    caseMesh(const caseMesh&) = delete;
    caseMesh& operator=(const caseMesh&) = delete;
    afar_mesh* mesh() const { return mesh_; }
    const dictionary& schemes() const { return schemes_; }
    const std::vector<double>& initialW() const { return W0_; }
    int64_t nCells() const { return nc_; }
    const std::vector<double>& cellCentres() const { return centres_; }  // caller order

    // fixedValue values (p, Ux, Uy, T) of patch p in the 0/ fields of an OpenFOAM case
    bool patchFixed(int p) const { return p < (int)patchFixed_.size() && patchFixed_[p]; }
    std::vector<double> patchValue(int p) const {
        return std::vector<double>(patchValues_.begin() + 4 * p, patchValues_.begin() + 4 * p + 4);
    }

   private:
    afar_mesh* mesh_ = nullptr;
    dictionary schemes_;
    std::vector<double> W0_, centres_, patchValues_;
    std::vector<char> patchFixed_;
    int64_t nc_ = 0;
};

class basicCUDAsolver {
   public:
    basicCUDAsolver(const runTime& rt, const caseMesh& mesh);
    ~basicCUDAsolver();
    basicCUDAsolver(const basicCUDAsolver&) = delete;
    basicCUDAsolver& operator=(const basicCUDAsolver&) = delete;
    // adjustTimeStep on the device, then runTime.setDeltaT (basicCUDAsolver.cpp:854-859)
    void setDeltaT(runTime& rt);
    // one SSP-RK3 step (basicCUDAsolver.cpp:862-871)
    void evolve(double deltaT);
    // p, U, T, shockIndicator of the current state into <case>/<time>/ (basicCUDAsolver.cpp:874-913)
    void writeField(const runTime& rt);
    const afar_schemes& schemes() const { return sc_; }

   private:
    afar_solver* solver_ = nullptr;
    afar_schemes sc_{};
    int64_t nc_ = 0, nb_ = 0;
};

// Reads the afar_schemes of a case: the reference keys and this library's
// extensions (CUDA.precision fp64|mixed, CUDA.exactRK, CUDA.tile).
afar_schemes readSchemes(const dictionary& fvSchemes, const dictionary& controlDict);

}  // namespace afarhost
