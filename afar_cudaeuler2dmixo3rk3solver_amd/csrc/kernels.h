// Device pointer bundle and kernel launchers of the AFAR stage pipeline.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace afar {

constexpr int kCflBlocks = 1024;
constexpr int kMaxSten = 20;  // largest stencil (the reference's limit, basicCUDAsolver.cpp:393-397)

template <typename R> struct VT;
template <> struct VT<double> {
    using v4 = double4;
    using v2 = double2;
};
template <> struct VT<float> {
    using v4 = float4;
    using v2 = float2;
};

// Everything a kernel reads or writes, passed by value.  Cell arrays hold nc
// real cells followed (where noted) by nb boundary pseudo-cells.
template <typename R>
struct DevPtrs {
    int nc, nf, nb, maxS, maxC, ntiles;
    int lds_sten_off;          // byte offset of the stencil ids in the reconstruction kernels' LDS
    int lds_m_off;             // byte offset of the tile's least-squares weights (four-lane kernel)
    int lds_bvd_sten_off;      // byte offset of the stencil ids in the BVD kernel's LDS
    int boundaryFirstOrder;    // extension: first-order faces in cells with a boundary face
    typename VT<R>::v4* W;     // nc+nb primitive p, Ux, Uy, T
    double4* U;                // nc conserved rho, rhoUx, rhoUy, rhoE
    double4* Uold;             // nc
    const double* vol;         // nc
    typename VT<R>::v2 *gp, *gu, *gv, *gT;  // nc: grad p; (dUx/dx, dUx/dy); (dUy/dx, dUy/dy); grad T
    typename VT<R>::v4* lim;   // nc limiters p, Ux, Uy, T
    int8_t* smooth;            // nc+nb smoothness flag (shockIndicator[2c])
    int8_t* ftype;             // nc final reconstruction type (shockIndicator[2c+1])
    uint8_t* flat_prev;        // nc: the cell took the flat branch at the last reconstruction (load scheduling only)
    const uint8_t* scnt;       // nc stencil size
    const uint8_t* ccnt;       // nc compact stencil size
    const uint16_t* sten;      // sliced ELL, LDS-local stencil ids
    const int32_t* slot_face;  // sliced ELL, +-(face+1)
    const typename VT<R>::v2* slot_d;  // sliced ELL, face centre - cell centre, in the kernels' precision
    const double2* slot_e;     // sliced ELL, compact neighbour - cell centre (K3 only)
    // FP64 four-lane reconstruction: the slot geometry and least-squares
    // weights are formed from these instead of read from slot_d, slot_dr, lsq
    const double2* cen;        // nc+nb cell and pseudo-cell centres
    const double2* fcen;       // nf+nb face centres
    const double2* fshift;     // nf periodic shift of the neighbour centre, or null when every shift is zero
    // Face-centre flux passes with the FP64 geometry (FP64; FP32-mixed with
    // the CFL bound): the face's offsets from both cell centres (and the
    // centre distance of the CFL bound) are formed from cen, fcen, fshift;
    // the face area vector and its magnitude are read from these
    const double2* fsf;        // nf Sf
    const double* fmag;        // nf magSf
    R* slot_dr;                // sliced ELL, |slot_e| in the kernels' precision (written by K3)
    typename VT<R>::v2* lsq;   // sliced ELL, least-squares weights
    const R* rbf;              // sliced ELL, RBF coefficients (face slot, stencil member)
    const int32_t* halo_ptr;
    const int32_t* halo;
    const int2* fcell;         // nf own, nei
    const double4* fgeo;       // 2*nf: (Sfx, Sfy, magSf, deltaR), (dxo, dyo, dxn, dyn)
    const float4* fgeo32;      // FP32-mixed: fgeo rounded to float (the flux pass without the CFL bound
                               // uses the geometry in float only); else null
    // 2*nf reconstructed face values [2f + side].  FP64: one array per
    // variable (fvp, fvu, fvv, fvT; fv null).  FP32-mixed: (p, Ux, Uy, T)
    // records in fv and the temperatures again in fvT (fvp, fvu, fvv null):
    // 4-byte scattered stores into four arrays measured slower than one
    // 16-byte record.  The BVD test reads fvT alone.
    typename VT<R>::v4* fv;
    R *fvp, *fvu, *fvv, *fvT;
    R* tTVD;                   // 2*nf the BVD test's limited linear face temperature of that side's cell,
                               // Tc + BJ(raw T limiter) * (grad T . d), formed by the reconstruction
    int8_t* tlim;              // nc: BJ(raw T limiter) < VALID_LIM_FACTOR (the BVD test's limiter condition)
    // extension (faceQuadrature): flux at ngp points per internal face
    int ngp;                   // 0: face centre only (reference)
    double gpw[4];             // point weights, sum 1
    const R* rbf_gp;           // sliced ELL, RBF coefficients (point, face slot, stencil member)
    const typename VT<R>::v4* fgd;  // ngp*nf: point - owner centre (x, y), point - neighbour centre (z, w),
                                    // in the kernels' precision
    typename VT<R>::v4* fvg;   // ngp*2*nf reconstructed states at the points, [(g*nf + f)*2 + side]
    double4* flux;             // nf + nb face fluxes (FP32-mixed face-centre path: boundary faces only)
    float4* flux32;            // FP32-mixed face-centre path: the nf internal faces' fluxes as computed,
                               // in float (exact); else null
    const int32_t* bcell;      // nb
    const uint8_t* blabel;     // nb
    const double4* bgeo;       // nb (Sfx, Sfy, magSf, 0)
    const double4* bval;       // nb fixed values of label 6
    double* dt;                // device time step of each instance (ninst)
    double* dt_part;           // dt_cap partial minima of the CFL pass
    unsigned int* dt_count;    // blocks of the CFL pass done
    int dt_cap;                // capacity of dt_part: kCflBlocks or two per flux block, whichever is larger
    // Batches: ninst independent meshes stored one after the other (cells,
    // internal faces, boundary faces each contiguous per instance).  Each has
    // its own CFL time step.  A single mesh is ninst = 1 with cinst = nullptr.
    int ninst;
    const int32_t* cinst;      // nc instance of each cell, or nullptr
    const int32_t* finst_off;  // ninst+1 first internal face of each instance
    // Device clock of each instance: tctl[4i] = t, tctl[4i+1] = end time,
    // tctl[4i+2] = steps taken.  The device time step is clipped to land on
    // the end time; once there, steps leave the state unchanged (time step 0).
    double* tctl;
    double* res_part;          // partial sums of the residual norm (one per update block)
};

template <typename R> void launch_lsq(const DevPtrs<R>& p, hipStream_t s);
template <typename R> void launch_cfl(const DevPtrs<R>& p, double Rg, double Cv, double CFL, hipStream_t s);
// lanes: 1 = one thread per cell (any mesh); 4 = four threads per cell;
// 2 = four threads per cell with the face work compacted (both: meshes whose
// cells have at most four faces; 2 also tiles of at most 128 cells).
template <typename R> void launch_reconstruct(const DevPtrs<R>& p, int tile, size_t lds, hipStream_t s, int lanes);
template <typename R> void launch_bvd(const DevPtrs<R>& p, int tile, size_t lds, hipStream_t s);
template <typename R> void launch_flux(const DevPtrs<R>& p, double Rg, double Cv, hipStream_t s);
// Flux pass that also reduces the CFL bound of the current state (the state
// the step starts from) to per-block minima in p.dt_part; launch_dt_reduce
// then stores their minimum in *p.dt.
template <typename R> void launch_flux_dt(const DevPtrs<R>& p, double Rg, double Cv, double CFL, hipStream_t s);
// Grid size of the flux kernels (blocks of 256 faces, a multiple of 8).
int flux_blocks(int nf, int nb);
template <typename R> void launch_dt_reduce(const DevPtrs<R>& p, hipStream_t s);
// Sum over cells of (density residual / volume)^2 from the current fluxes
// into *out (device); res_part needs residual_blocks(nc) entries.
int residual_blocks(int nc);
template <typename R> void launch_residual_norm(const DevPtrs<R>& p, double* out, hipStream_t s);
template <typename R>
void launch_update(const DevPtrs<R>& p, double Rg, double gamma, double dt, const double* dt_ptr, double b1, double b2,
                   double b3, int last, hipStream_t s);
// Raises each reconstruction kernel's dynamic-LDS limit to lds minus its static LDS.
template <typename R> hipError_t set_lds_limits(size_t lds);
void launch_math_selftest(const double* x, double* out, int n, hipStream_t s);
void launch_math_selftest_f32(const float* x, float* out, int n, hipStream_t s);
// The CFL time step over the caller's device arrays in the reference's layout
// (FD, CELL, FACE, neighbour, optional FSHIFT) into *deltaT (device memory),
// on stream s; allocates nothing (its partial minima are a static device array).
void adjust_time_step_noalloc(const double* FD, const double* CELL, const double* FACE, const int32_t* neighbour,
                              const double* FSHIFT, int nf, double R, double Cv, double CFL, double* deltaT,
                              hipStream_t s);

}  // namespace afar
