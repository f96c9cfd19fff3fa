// ORACLE -- test infrastructure only (see afar_oracle.cpp): the field
// struct of the reference layout and the slot geometry shared by the FP64
// and the FP32-mixed restatements.
#pragma once
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../afar_cudaeuler2dmixo3rk3solver_amd/csrc/afar_math.h"

extern "C" {

typedef struct oracle_fields {
    int32_t nc, nf, nb, maxS, maxC;
    const double* CELL;              // 2*(nc+nb)
    const double* FACE;              // 5*(nf+nb)
    const double* cellVolume;        // nc
    double* FD;                      // 4*(nc+nb)
    double* cFD;                     // 4*nc
    double* cFDold;                  // 4*nc
    const int32_t* neighbour;        // 2*nf
    const int32_t* boundaryFaceNeiLabel;  // nb
    const uint8_t* boundaryFacesType;     // nb
    const uint8_t* stencilSize;      // nc
    const uint8_t* compactStencilSize;  // nc
    const int32_t* stencil;          // maxS*nc global ids
    const int32_t* cellFaces;        // maxC*nc
    const double* RBFbasis;          // 2*(maxS+1)*nf
    const double* FSHIFT;            // 2*nf or NULL
    const double* BVALUE;            // 4*nb or NULL
    double* matrix;                  // 2*maxC*nc
    double* gradFD;                  // 8*nc: px, py, Uxx, Uxy, Uyx, Uyy, Tx, Ty (gradientFieldData order)
    double* limiter;                 // 4*nc
    int8_t* shockIndicator;          // 2*(nc+nb)
    double* faceFD;                  // 8*nf: [2f+side] * 4
    double* Flux;                    // 4*(nf+nb)
    double* Res;                     // 4*nc
    int32_t boundaryFirstOrder;      // library extension (0 = reference behaviour)
    // library extension (faceQuadrature): flux at ngp points per internal face
    int32_t ngp;                     // 0 = face centre (reference)
    const double* gpWeights;         // ngp, sum 1
    const double* gpOffsets;         // 4*ngp*nf: point - owner centre, point - neighbour centre
    const double* RBFgp;             // ngp*2*(maxS+1)*nf
    double* faceGP;                  // 8*ngp*nf: [(g*nf + f)*2 + side] * 4
} oracle_fields;

}  // extern "C"

namespace oracle_detail {

// gradientFieldData member order (basicDataStructure.h:41-49)
enum { GPX = 0, GPY, GUXX, GUXY, GUYX, GUYY, GTX, GTY };

inline double shx(const oracle_fields* f, int face) { return f->FSHIFT ? f->FSHIFT[2 * face] : 0.0; }
inline double shy(const oracle_fields* f, int face) { return f->FSHIFT ? f->FSHIFT[2 * face + 1] : 0.0; }

// Position of compact neighbour fi of cell c in c's frame, and the face
// centre offset, following the owner-frame convention above.
struct Slot {
    int face, isNei, isBoundary, nb;
    double ex, ey;  // neighbour centre - cell centre
    double dx, dy;  // face centre - cell centre
};

inline Slot slot(const oracle_fields* f, int c, int fi) {
    Slot s;
    int sf = f->cellFaces[(size_t)f->maxC * c + fi];
    s.isNei = sf < 0;
    s.face = abs(sf) - 1;
    s.isBoundary = s.face >= f->nf;
    const double* F = &f->FACE[5 * (size_t)s.face];
    const double xc = f->CELL[2 * c], yc = f->CELL[2 * c + 1];
    if (s.isBoundary) {
        s.nb = f->nc + (s.face - f->nf);
        s.ex = f->CELL[2 * (size_t)s.nb] - xc;
        s.ey = f->CELL[2 * (size_t)s.nb + 1] - yc;
        s.dx = F[0] - xc;
        s.dy = F[1] - yc;
    } else {
        const int own = f->neighbour[2 * s.face], nei = f->neighbour[2 * s.face + 1];
        const double xn = f->CELL[2 * nei] + shx(f, s.face), yn = f->CELL[2 * nei + 1] + shy(f, s.face);
        if (!s.isNei) {
            s.nb = nei;
            s.ex = xn - f->CELL[2 * own];
            s.ey = yn - f->CELL[2 * own + 1];
            s.dx = F[0] - f->CELL[2 * own];
            s.dy = F[1] - f->CELL[2 * own + 1];
        } else {
            s.nb = own;
            s.ex = f->CELL[2 * own] - xn;
            s.ey = f->CELL[2 * own + 1] - yn;
            s.dx = F[0] - xn;
            s.dy = F[1] - yn;
        }
    }
    return s;
}

}  // namespace oracle_detail
