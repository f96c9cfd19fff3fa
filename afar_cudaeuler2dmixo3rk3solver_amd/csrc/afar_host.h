// Host-side data model of the AFAR pipeline: 2D polygon meshes, their
// renumbered finite-volume topology and geometry, the reconstruction stencils
// and the precomputed reconstruction coefficients.  Everything here is plain
// C++ (FP64); the device layout is built from it in layout.cpp.
#pragma once
#include <stdint.h>

#include <array>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace afar {

// Compile-time limits of the reference solver
// (src/CUDA/cudaEuler2DmixO3RK3/basicDataStructure.h:6-8).
constexpr int kMaxValidStencil = 20;

struct Vec2 {
    double x = 0.0, y = 0.0;
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Boundary-condition labels kept from the reference dispatch table
// (src/CUDA/BCs/2D/cudaEulerBCs.h:34-39): 0 constantTemperatureWall,
// 1 adiabaticWall, 2 symmetry (slip), 3 fixedValue, 4 outlet, 5 fixedValue2.
// Label 6 is this library's parameterised fixed value (p, Ux, Uy, T per patch).
constexpr int kMaxBCLabel = 6;

struct Patch {
    std::string name;
    int label = 4;
    double value[4] = {0, 0, 0, 0};  // p, Ux, Uy, T for label 6
};

// A mesh as the caller describes it: canonical points, CCW polygons, and the
// boundary edges with their patch.  Periodic meshes give, per cell vertex, the
// integer multiple of the period that places the vertex next to the cell.
struct PolyMesh {
    std::vector<Vec2> pts;
    std::vector<int64_t> cell_ptr{0};
    std::vector<int32_t> cell_pts;
    std::vector<int8_t> cell_vshift;  // 2 per cell vertex, or empty
    Vec2 period;
    std::vector<std::array<int32_t, 2>> bedges;
    std::vector<int32_t> bedge_patch;
    std::vector<Patch> patches;
    double depth = 1.0;
};

// The finite-volume mesh after renumbering and face ordering.  Internal faces
// come first in upper-triangular order (owner < neighbour, sorted by owner then
// neighbour, as OpenFOAM's renumberMesh leaves them); boundary faces follow,
// grouped by patch.  Boundary faces act as pseudo-cells nc..nc+nb-1 located at
// the face centres (src/CUDA/cudaEuler2DmixO3RK3/basicCUDAsolver.cpp:203-235).
struct Mesh {
    int nc = 0, nf = 0, nb = 0;
    double depth = 1.0;
    std::vector<Patch> patches;
    std::vector<int32_t> patch_start, patch_size;  // into boundary faces [0, nb)

    std::vector<int32_t> perm;   // new cell id -> caller's cell id
    std::vector<int32_t> iperm;  // caller's cell id -> new cell id

    // cells: polygon vertices in the cell's own frame
    std::vector<int32_t> cell_ptr, cell_pts;  // canonical point ids
    std::vector<Vec2> cell_xy;                // coordinates, same indexing as cell_pts
    std::vector<int8_t> cell_vs;              // periodic vertex shifts (2 per vertex, zeros if none)
    std::vector<Vec2> C;                      // nc + nb centres
    std::vector<double> V;                    // nc volumes (area * depth)

    // faces [0, nf + nb): centre, area vector, |Sf| in the owner's frame
    std::vector<int32_t> own, nei;            // nei only for internal faces
    std::vector<Vec2> Cf, Sf;
    std::vector<double> magSf;
    std::vector<Vec2> fp0, fp1;               // end points (owner frame, owner's CCW order)
    std::vector<Vec2> fshift;                 // nf: neighbour centre in owner frame = C[nei] + fshift
    std::vector<uint8_t> bf_label;            // nb
    std::vector<int32_t> bf_patch;            // nb

    // canonical point -> (cell, shift) incidence, used by the edge stencil
    int npts = 0;
    Vec2 period;
};

struct Schemes {
    double R = 0.7142857, Cv = 1.7857142857, maxCo = 0.5, shapeFactor = 10.0;
    int stencilType = 1;  // 0 faceNeighbour, 1 edgeNeighbour
    int recalcBasis = 1;
    int cOrder = 5, fOrder = 4;  // reference defaults (d2GaussQuadrature.C:12-13)
    int precision = 0;           // 0 FP64, 1 FP32-mixed
    int exactRK = 0;             // 0 literal reference coefficients, 1 exact 1/3, 2/3
    int faceQuadrature = 0;      // extension: 0 flux at the face centre (reference), 1 at the fOrder points
};

// Reconstruction stencils (src/CUDA/cudaEuler2DmixO3RK3/basicCUDAsolver.cpp:269-417).
struct Stencils {
    int maxS = 0, maxC = 0;
    std::vector<uint8_t> S, Cc;            // per cell: stencil and compact sizes
    std::vector<int32_t> sten;             // nc * kMaxValidStencil, global ids (pseudo-cells >= nc)
    std::vector<Vec2> sten_shift;          // same indexing: member position in the cell's frame = C[id] + shift
    std::vector<int32_t> cellFaces;        // nc * kMaxValidStencil, +-(face+1) in compact order
};

struct Precomputed {
    Stencils st;
    std::vector<double> rbf;     // 2*(maxS+1)*nf, reference layout (RBFbasis)
    std::vector<uint8_t> trouble;  // per cell: singular moment matrix
    std::vector<double> refDistance;
    int troubleCount = 0;
    // extension (faceQuadrature 1): flux quadrature points of every internal
    // face, xi on [-1, 1] with weights summing to one, and each face side's
    // RBF coefficients at each point: rbf_gp[(g*nf + f)*2*(maxS+1) + side*(maxS+1) + i]
    int ngp = 0;
    std::vector<double> gp_xi, gp_w;
    std::vector<double> rbf_gp;
};

// ---- mesh.cpp
Mesh build_mesh(const PolyMesh& pm, bool renumber);
PolyMesh generate_corner(int nx, int ny, double lx, double ly, double jitter, uint64_t seed,
                         double corner_x, double angle_deg, const double inflow[4], int top_label);
PolyMesh generate_riemann(int n, double jitter, uint64_t seed, double split_p);
PolyMesh generate_periodic(int n, double L, double jitter, uint64_t seed, double split_p);
std::vector<int32_t> rcm_order(int n, const std::vector<int32_t>& adj_ptr, const std::vector<int32_t>& adj);

// ---- precompute.cpp
Stencils build_stencils(const Mesh& m, int stencilType);
void gauss_cell(const Mesh& m, int cell, int cOrder, std::vector<Vec2>& pts, std::vector<double>& w);
void gauss_boundary_face(const Mesh& m, int b, int fOrder, std::vector<Vec2>& pts, std::vector<double>& w);
// The fOrder bar rule on [-1, 1] (d2GaussQuadrature.C:489-547) and a point of
// face f at xi, in the owner's frame (d2GaussQuadrature.C:199-206).
void face_gauss_rule(int fOrder, std::vector<double>& xi, std::vector<double>& w);
Vec2 face_gauss_point(const Mesh& m, int f, double xi);
void pinv_jacobi(int n, const double* A, double* G);  // row-major n x n, n <= 21
void gauss_tables(const Mesh& m, const Schemes& sc, std::vector<double>& pts, std::vector<double>& w,
                  std::vector<int64_t>& ptr);
void rbf_cell_inputs(const Mesh& m, const Stencils& st, const double* tp, const double* tw, const int64_t* tptr, int c,
                     double* X, std::vector<double>& gp, std::vector<double>& gw, int* gptr);
Precomputed precompute(const Mesh& m, const Schemes& sc);
// ---- precompute_gpu.hip: the RBF part of precompute on the GPU, from pc.st
// (rbf, trouble, refDistance replaced); equal to the host's bit for bit.
void precompute_rbf_gpu(const Mesh& m, const Schemes& sc, Precomputed& pc);
void write_basis_lists(const Mesh& m, const Precomputed& pc, const std::string& dir);
void read_basis_lists(const Mesh& m, Precomputed& pc, const std::string& dir);
void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& fn);

}  // namespace afar
