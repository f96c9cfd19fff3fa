// Two host-side layouts:
//
//  RefArrays  the reference solver's host arrays (fieldPointer,
//             src/CUDA/cudaEuler2DmixO3RK3/basicDataStructure.h:75-160) with
//             the block-local stencil tables replaced by the global stencil
//             they encode, RBF coefficients held in FP64, plus two extension
//             arrays (periodic face shifts, per-face fixed values of label 6).
//             Seam 2 of the C-ABI accepts the reference's own arrays and turns
//             them into this; seam 1 builds it from a Mesh and its precompute.
//
//  DevLayout  what the gfx950 kernels read: per-cell slot data in a sliced
//             ELL layout of width 64 (one wavefront: element k of cell c sits at
//             ((c/64)*K + k)*64 + c%64, so every load of a wave is one
//             contiguous 512-byte row), per-face records, LDS tile halos.
#pragma once
#include <stdint.h>

#include <vector>

#include "afar_host.h"

namespace afar {

struct RefArrays {
    int nc = 0, nf = 0, nb = 0, npatch = 0, maxS = 0, maxC = 0;
    std::vector<double> CELL;        // 2*(nc+nb): x, y
    std::vector<double> FACE;        // 5*(nf+nb): x, y, magSf, Sfx, Sfy
    std::vector<double> cellVolume;  // nc
    std::vector<double> FD;          // 4*(nc+nb): p, Ux, Uy, T (ghost values after nc)
    std::vector<double> cFD;         // 4*nc: rho, rhoUx, rhoUy, rhoE
    std::vector<int32_t> neighbour;  // 2*nf
    std::vector<int32_t> boundaryFacesNum;      // npatch
    std::vector<int32_t> boundaryFaceNeiLabel;  // nb
    std::vector<uint8_t> boundaryFacesType;     // nb
    std::vector<uint8_t> stencilSize, compactStencilSize;  // nc
    std::vector<int32_t> stencil;    // maxS*nc, global ids
    std::vector<double> RBFbasis;    // 2*(maxS+1)*nf
    std::vector<int32_t> cellFaces;  // maxC*nc
    std::vector<double> FSHIFT;      // 2*nf: neighbour centre seen from the owner = CELL[nei] + FSHIFT
    std::vector<double> BVALUE;      // 4*nb: fixed values for label 6
    // extension (faceQuadrature): ngp flux points per internal face
    int ngp = 0;
    std::vector<double> GPW;         // ngp weights, sum 1
    std::vector<double> GPD;         // 4*ngp*nf: per (g, f) point - owner centre, point - neighbour centre
    std::vector<double> RBFgp;       // ngp*2*(maxS+1)*nf: RBFbasis at each point
};

// Builds the reference arrays of a mesh.  W0 holds p, Ux, Uy, T per cell in
// the mesh's (renumbered) order; ghost values are set from the BC table.
RefArrays make_ref_arrays(const Mesh& m, const Precomputed& pc, const double* W0, double R, double Cv);
// Applies the boundary table to FD's ghost entries from the adjacent cells
// (src/CUDA/BCs/2D/cudaEulerBCs.h:42-73), on the host.
void apply_bc_host(RefArrays& ra);
// Conserved variables from FD (basicCUDAsolver.cpp:220-223).
void conserved_from_primitive(RefArrays& ra, double R, double Cv);

struct DevLayout {
    int nc = 0, nf = 0, nb = 0, maxS = 0, maxC = 0, tile = 256;
    int nwaves = 0, ntiles = 0, maxHalo = 0;
    // per cell
    std::vector<uint8_t> scnt, ccnt;      // nc
    std::vector<uint16_t> sten;           // nwaves*maxS*64, LDS-local ids
    std::vector<int32_t> slot_face;       // nwaves*maxC*64, +-(f+1), 0 = padding
    std::vector<double> slot_d;           // 2 per slot: face centre - cell centre
    std::vector<double> slot_e;           // 2 per slot: compact neighbour - cell centre
    std::vector<double> rbf;              // nwaves*maxC*(maxS+1)*64
    std::vector<int32_t> halo_ptr, halo;  // ntiles+1, global ids of halo members
    // per internal face
    std::vector<int32_t> fcell;           // 2*nf: own, nei
    std::vector<double> fgeo;             // 8*nf: Sfx, Sfy, magSf, deltaR, dxo, dyo, dxn, dyn
    // per boundary face
    std::vector<int32_t> bcell;           // nb
    std::vector<uint8_t> blabel;          // nb
    std::vector<double> bgeo;             // 4*nb: Sfx, Sfy, magSf, 0
    std::vector<double> bval;             // 4*nb
    std::vector<double> vol;              // nc
    // extension (faceQuadrature): flux points
    int ngp = 0;
    std::vector<double> gpw;              // ngp
    std::vector<double> rbf_gp;           // per slot and point: element (g*maxC + k)*(maxS+1) + i, width 64
    std::vector<double> fgd;              // 4*ngp*nf as RefArrays::GPD
    size_t sell(int c, int k, int K) const { return ((size_t)(c >> 6) * K + k) * 64 + (c & 63); }
};

DevLayout make_dev_layout(const RefArrays& ra, int tile);

// A batch of independent meshes as one: the instances' cells, internal
// faces, boundary faces (pseudo-cells) and patches each follow one another,
// every index renumbered accordingly.  No stencil or face crosses instances,
// so each cell computes exactly what it computes in its own mesh.
struct BatchIndex {
    std::vector<int32_t> cell_off;   // ninst+1, real cells
    std::vector<int32_t> face_off;   // ninst+1, internal faces
    std::vector<int32_t> bface_off;  // ninst+1, boundary faces
    int ninst() const { return (int)cell_off.size() - 1; }
};
RefArrays union_ref_arrays(const std::vector<const RefArrays*>& parts, BatchIndex& idx);

// Device order of cells and internal faces.  The mesh's order (reverse
// Cuthill-McKee) makes consecutive cells a thin front, so a 64-cell tile has
// about four halo cells per cell for an extended stencil.  Cells ordered along
// a Hilbert curve through their centres form compact tiles instead; internal
// faces then follow their owner's new index.  Every instance of a batch keeps
// its cell and face ranges.  Relabelling changes no arithmetic: each cell
// keeps its stencil and face lists in their order.
struct Relabel {
    std::vector<int32_t> cell_new, cell_old;  // nc: new index of mesh cell i; mesh cell at new index j
    std::vector<int32_t> face_new, face_old;  // nf internal faces, likewise
};
// group > 0: within each aligned block of `group` cells (inside one
// instance) cells are further ordered by face count, then stencil size, so
// that the cells of a wavefront run loops of equal length and read
// coefficient rows that all of them use.  The price is that cells close in
// space are no longer close in memory within a block, which the face kernel's
// gathers of both sides' states pay for.
Relabel compact_tile_order(const RefArrays& ra, const BatchIndex& bi, int group = 0);
// The same problem with cells and internal faces renamed by r (pseudo-cells
// and boundary faces keep their indices).
RefArrays relabel(const RefArrays& ra, const Relabel& r);

// The reference's 128-cell block tables (basicCUDAsolver.cpp:697-850): per
// block the halo cells (compact-stencil halo first, then the rest of the
// extended stencil) and block-local stencil ids.
struct BlockTables {
    int maxLB = 0, maxCLB = 0;                  // 128 + largest halo, 128 + largest compact halo
    std::vector<int32_t> extendStencil;         // (maxLB-128) per block
    std::vector<uint16_t> extendStencilSize;    // per block
    std::vector<uint16_t> compactExtendStencilSize;
    std::vector<uint16_t> localStencil;         // maxS*nc
};
BlockTables make_block_tables(const RefArrays& ra);
// Inverse: global stencil ids from the block tables.
void global_stencil_from_blocks(int nc, int maxS, int maxLB, const uint8_t* S, const uint16_t* localStencil,
                                const int32_t* extendStencil, int32_t* stencil);

}  // namespace afar
