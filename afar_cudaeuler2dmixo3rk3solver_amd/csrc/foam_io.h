// OpenFOAM ASCII case I/O for one-cell-thick 2D cases (see foam_io.cpp).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "afar_host.h"

namespace afar {

// The 2D polygon mesh of case_dir/constant/polyMesh.  Patches keep the
// polyMesh order (empty ones included, with no edges).
PolyMesh read_polymesh_2d(const std::string& case_dir, std::vector<std::string>* patch_names,
                          std::vector<std::string>* patch_types);
// Writes m as case_dir/constant/polyMesh (the directory must exist).
void write_polymesh_2d(const Mesh& m, const std::string& case_dir);

struct FoamField {
    bool have_internal = false;
    std::vector<double> internal;      // ncomp per cell, file order
    std::vector<double> patch_value;   // ncomp per patch: the uniform value of fixedValue patches
    std::vector<char> patch_fixed;     // per patch
};
// A volScalarField (ncomp 1) or the x, y components of a volVectorField (ncomp 2).
FoamField read_field(const std::string& path, int ncomp, int64_t ncells, const std::vector<std::string>& patches);

}  // namespace afar
