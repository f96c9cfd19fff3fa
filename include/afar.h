/*
 * afar.h -- C-ABI of the MI355X-native AFAR SSP-RK3 pipeline.
 *
 * Two seams of the reference solver are reproduced:
 *
 *   Seam 1, application <-> solver object: the C++ class basicCUDAsolver
 *   (src/CUDA/cudaEuler2DmixO3RK3/basicCUDAsolver.h:9-53) called from main()
 *   (applications/solvers/compressible/CUDA/cudaEuler2DmixO3RK3solver/
 *   cudaEuler2DmixO3RK3solver.cpp:46,58,64,68,85).  Here: afar_mesh_* builds
 *   the mesh and its precompute, afar_solver_* is the solver object.
 *
 *   Seam 2, host <-> device: the free functions over fieldPointer of
 *   src/CUDA/cudaEuler2DmixO3RK3/GPUevolve.h:6-59, called from
 *   basicCUDAsolver.cpp:80-96,115,856,864-870,876.  Here: afar_set_device_*,
 *   afar_free_field_data, afar_adjust_time_step, afar_gpu_evolve,
 *   afar_copy_field_data_device_to_host over an opaque afar_device.
 *
 * Every function returns AFAR_OK (0) or an error code and never exits the
 * process; afar_last_error() gives the message of the calling thread's last
 * failure.
 *
 * Threading.  Handles are independent: different handles may be created,
 * stepped and destroyed on different host threads at the same time.  The
 * calls that allocate or free device memory, create or destroy a solver,
 * batch or device, or capture a step graph take one process-wide lock, so
 * they never overlap each other (a stream being captured must not be waited
 * on, and some of these calls wait on the device); kernel launches, copies
 * and queries on a handle's own stream do not take it.  One handle must not
 * be used by two threads at once.  Each solver runs on one HIP stream: its
 * own (afar_solver_create), or the caller's (afar_solver_create_on_stream,
 * afar_device_set_stream), which the library uses but never destroys.
 *
 * Sizes.  Cell and face counts of seam 2 are 64-bit; the device layout
 * indexes cells and faces with 32-bit ids (the reference's int), so meshes
 * with more than 2^31 - 1 cells + boundary faces are refused (AFAR_ERR_LIMIT).
 */
#ifndef AFAR_H
#define AFAR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AFAR_OK 0
#define AFAR_ERR_INVALID 1 /* bad argument, mesh or dictionary value */
#define AFAR_ERR_HIP 2     /* HIP runtime error */
#define AFAR_ERR_LIMIT 3   /* stencil or LDS limit exceeded */
#define AFAR_ERR_IO 4      /* basis cache file error */

typedef struct afar_mesh afar_mesh;
typedef struct afar_solver afar_solver;
typedef struct afar_device afar_device;

/* Dictionary values of the reference case:
 *   fvSchemes CUDA{R; Cv; stencilType; recalcBasis}  (basicCUDAsolver.cpp:55-68)
 *   fvSchemes shapeFactor (top level, default 10)     (basicCUDAsolver.cpp:58)
 *   fvSchemes gaussQuadrature{cOrder 5; fOrder 4}     (d2GaussQuadrature.C:12-13)
 *   controlDict maxCo                                 (basicCUDAsolver.cpp:57)
 * plus this library's extensions: precision, exactRK, tile. */
typedef struct afar_schemes {
    double R;
    double Cv;
    double maxCo;
    double shapeFactor;
    int32_t stencilType; /* 0 faceNeighbour, 1 edgeNeighbour */
    int32_t recalcBasis; /* 1 compute the RBF basis, 0 read system/basis{Own,Nei}List */
    int32_t cOrder;
    int32_t fOrder;
    int32_t precision;   /* 0 all-FP64, 1 FP32 reconstruction/limiter/flux with FP64 state */
    int32_t exactRK;     /* 0 the reference's RK3 literals (GPUevolve.cu:223-225), 1 exact 1/3, 2/3 */
    int32_t tile;        /* cells per LDS tile: 64, 128 or 256 (0 = 64) */
    int32_t boundaryFirstOrder; /* extension, 0 = reference behaviour: cells with a boundary face
                                   use first-order face states (the reference does this only for
                                   cells with two boundary faces, cudaLimiter.h:419-429) */
    int32_t reconLanes;  /* reconstruction kernel, same results: 0 automatic, 1 one thread per
                            cell, 4 four threads per cell (cells of at most four faces;
                            automatic picks 4, else 1) */
    int32_t cellOrder;   /* device order of cells, same results: 0 compact tiles (cells along a
                            Hilbert curve through their centres, faces by owner) whose blocks of
                            32 (FP64) or 64 (mixed) cells are ordered by face count and stencil
                            size, 1 the mesh's order, 2 compact tiles in curve order throughout;
                            every array the library returns is in the mesh's order */
    int32_t faceQuadrature; /* extension, 0 = reference: the flux of an internal face is evaluated
                            at its centre (cudaRiemannSolver.h:178-182).  1: at the points of the
                            fOrder bar rule (d2GaussQuadrature.C:153-214), weights summing to
                            one; face states there come from per-point RBF coefficients (O3) or
                            the limited linear reconstruction (TVD) with the limiters, flags and
                            switch of the centre.  Set it for the mesh precompute and the solver
                            alike; fOrder 1 reproduces 0 bit for bit */
} afar_schemes;

const char* afar_last_error(void);
const char* afar_version(void);
/* Fills the reference defaults (R, Cv of the cornerFlow case, shapeFactor 10,
 * edgeNeighbour, recalcBasis true, cOrder 5, fOrder 4, maxCo 0.5). */
void afar_schemes_default(afar_schemes* s);

/* ------------------------------------------------------------------ meshes
 * Replaces the OpenFOAM fvMesh the reference reads (basicCUDAsolver.cpp:
 * 122-258) and the offline `renumberMesh` (README.md:76).  Polygons are
 * counter-clockwise in the x-y plane, extruded by `depth`.  Boundary edges
 * carry a patch index; patch_labels are the CUDA.boundary.patch<i> labels
 * 0..5 of the reference (src/CUDA/BCs/2D/cudaEulerBCs.h:34-39) or 6, this
 * library's fixed value taken from patch_values (p, Ux, Uy, T per patch).
 * cell_vshift (optional) gives per cell vertex the multiple of the period
 * (period_x, period_y) added to the point, for periodic meshes. */
int afar_mesh_from_polygons(int64_t npoints, const double* xy, int64_t ncells, const int64_t* cell_ptr,
                            const int32_t* cell_pts, const int8_t* cell_vshift, double period_x, double period_y,
                            int64_t nbedges, const int32_t* bedge_pts, const int32_t* bedge_patch, int32_t npatches,
                            const int32_t* patch_labels, const double* patch_values, double depth, int32_t renumber,
                            afar_mesh** out);
/* Supersonic corner: nx x ny quads on [0,lx] x [wall(x), ly], wall bending
 * down by angle_deg at corner_x; interior nodes jittered by +-jitter spacing.
 * Patches: 0 inlet (label 6, inflow p,Ux,Uy,T), 1 outlet (4), 2 lower wall (2), 3 upper (top_label). */
int afar_mesh_generate_corner(int32_t nx, int32_t ny, double lx, double ly, double jitter, uint64_t seed,
                              double corner_x, double angle_deg, const double* inflow, int32_t top_label,
                              afar_mesh** out);
/* Unit square, n x n quads, interior nodes jittered, each quad split into two
 * triangles with probability split_p (diagonal chosen by a second draw).
 * Four patches (left, right, bottom, top), all label 4. */
int afar_mesh_generate_riemann(int32_t n, double jitter, uint64_t seed, double split_p, afar_mesh** out);
/* Doubly periodic [0,L]^2, n x n quads, all nodes jittered; each quad split
 * into two triangles with probability split_p (1: 2n^2 triangles). */
int afar_mesh_generate_periodic(int32_t n, double L, double jitter, uint64_t seed, double split_p, afar_mesh** out);
void afar_mesh_destroy(afar_mesh* m);
/* OpenFOAM case I/O (ASCII), one-cell-thick 2D cases as the reference runs
 * (basicCUDAsolver.cpp:9-54 reads them through Foam::fvMesh):
 * afar_mesh_read_foam reads case_dir/constant/polyMesh (points, faces, owner,
 * neighbour, boundary); the 2D cells are the cells' faces on the lower z
 * plane, the depth is the z extent, patches keep the polyMesh order (empty
 * ones included) and start with label 4 (set them with afar_mesh_set_patch
 * from CUDA.boundary.patch<i>).  renumber 1 applies RCM (as renumberMesh),
 * 0 keeps the file's cell order.  afar_mesh_write_foam writes a mesh as a
 * one-cell-thick polyMesh into case_dir/constant/polyMesh (which must exist),
 * with an extra empty patch frontAndBack.  afar_foam_read_fields reads p, U, T
 * of case_dir/time_dir: W0 (4 per cell, file order = caller order) and per
 * patch the uniform fixedValue values (4 per patch, patch_fixed 1 when p, U
 * and T are all fixedValue); the last two may be NULL. */
int afar_mesh_read_foam(const char* case_dir, int32_t renumber, afar_mesh** out);
int afar_mesh_write_foam(const afar_mesh* m, const char* case_dir);
/* Name (truncated to cap bytes) and boundary faces of patch `patch`; either output may be NULL. */
int afar_mesh_patch_info(const afar_mesh* m, int32_t patch, char* name, int32_t cap, int32_t* nfaces);
int afar_foam_read_fields(const afar_mesh* m, const char* case_dir, const char* time_dir, double* W0,
                          double* patch_values, int32_t* patch_fixed);
/* Sets the boundary label of patch `patch` (counting every patch, empty ones
 * included, as the reference's CUDA.boundary.patch<i> names do,
 * basicCUDAsolver.cpp:157-172); value (p, Ux,
 * Uy, T) is used by label 6 and may be NULL to keep the current one. */
int afar_mesh_set_patch(afar_mesh* m, int32_t patch, int32_t label, const double* value);
/* out[0..8] = nc, nf, nb, non-empty patches, maxS, maxC,
 * maxLocalBlockStencilSize, maxCompactLocalBlockStencilSize (these four are 0
 * before precompute; the block sizes are those of the reference's 128-cell
 * tables, basicCUDAsolver.cpp:697-850), all patches (empty ones included). */
int afar_mesh_sizes(const afar_mesh* m, int64_t* out);
/* Cell and pseudo-cell centres 2*(nc+nb), volumes nc (mesh order). */
int afar_mesh_cells(const afar_mesh* m, double* centres, double* volumes);
/* perm[i] = caller's index of mesh cell i (the RCM order). */
int afar_mesh_permutation(const afar_mesh* m, int32_t* perm);
/* own nf+nb, nei nf, face 5*(nf+nb) (x, y, magSf, Sfx, Sfy), fshift 2*nf, patch nb. */
int afar_mesh_faces(const afar_mesh* m, int32_t* own, int32_t* nei, double* face, double* fshift, int32_t* bpatch);
/* Cell polygons in mesh order: vptr nc+1, vxy 2*vptr[nc] (coordinates in the cell's frame). */
int afar_mesh_polygons(const afar_mesh* m, int32_t* vptr, double* vxy);

/* ------------------------------------------------------------------ precompute
 * Stencils, Gauss quadrature, RBF moment matrices, pseudo-inverses and face
 * coefficients of basicCUDAsolver::initStencilData (basicCUDAsolver.cpp:
 * 261-694).  basis_dir: when recalcBasis is 1 and basis_dir is not NULL the
 * coefficients are written there as basisOwnList/basisNeiList; when
 * recalcBasis is 0 they are read from there (basicCUDAsolver.cpp:626-694). */
int afar_mesh_precompute(afar_mesh* m, const afar_schemes* s, const char* basis_dir);
/* Same precompute with the RBF moment matrices, their pseudo-inverses and the
 * face coefficients computed on the GPU (one thread per cell); equal to
 * afar_mesh_precompute bit for bit (shared arithmetic, IEEE double).
 * Extension without a reference counterpart, for meshes whose host
 * precompute dominates the time to solution. */
int afar_mesh_precompute_gpu(afar_mesh* m, const afar_schemes* s);
/* S nc, C nc, stencil maxS*nc (global ids, pseudo-cells >= nc), cellFaces maxC*nc (+-(face+1)). */
int afar_mesh_stencils(const afar_mesh* m, uint8_t* S, uint8_t* C, int32_t* stencil, int32_t* cellFaces);
/* rbf 2*(maxS+1)*nf (RBFbasis layout), trouble nc, refDistance nc; NULL to skip. */
int afar_mesh_rbf_basis(const afar_mesh* m, double* rbf, uint8_t* trouble, double* refDistance);
/* Flux quadrature of a mesh precomputed with faceQuadrature 1 (extension, no
 * reference counterpart; *ngp = 0 otherwise): xi ngp points on [-1, 1],
 * weights ngp (sum 1), offsets 4*ngp*nf (per point g and internal face f:
 * point - owner centre, point - neighbour centre), rbf ngp*2*(maxS+1)*nf
 * (RBFbasis layout per point).  NULL to skip an output. */
int afar_mesh_face_quadrature(const afar_mesh* m, int32_t* ngp, double* xi, double* weights, double* offsets,
                              double* rbf);
/* Gauss points of a stencil member (cell < nc: cell rule cOrder; pseudo-cell: face rule fOrder). */
int afar_mesh_gauss_points(const afar_mesh* m, int32_t member, int32_t order, double* xy, double* w, int32_t cap,
                           int32_t* count);
int afar_mesh_write_basis(const afar_mesh* m, const char* dir);
/* Moore-Penrose pseudo-inverse (one-sided Jacobi SVD), row-major n x n; the
 * precompute's replacement of Foam::SVDinv (basicCUDAsolver.cpp:509). */
int afar_pinv(int32_t n, const double* A, double* G);

/* ------------------------------------------------------------------ seam 1: solver object */
/* basicCUDAsolver(runTime, mesh) (basicCUDAsolver.cpp:9-110): W0 = p, Ux, Uy, T
 * per cell in the caller's order (the 0/ fields); ghost states come from the
 * boundary table.  The mesh must have been precomputed with the same schemes. */
int afar_solver_create(const afar_mesh* m, const afar_schemes* s, const double* W0, afar_solver** out);
/* The same solver on the caller's stream (a hipStream_t passed as void*; not
 * NULL): every kernel, graph launch and copy of the solver is ordered on it,
 * so the caller can order the solver's work with its own.  The library never
 * destroys the stream; it must outlive the solver.  The reference has no
 * counterpart (its kernels run on the legacy default stream,
 * GPUevolve.cu:193,299); for callers that schedule several solvers, as
 * batching and serving do. */
int afar_solver_create_on_stream(const afar_mesh* m, const afar_schemes* s, const double* W0, void* stream,
                                 afar_solver** out);
/* basicCUDAsolver::setDeltaT (basicCUDAsolver.cpp:854-859): global CFL time step. */
int afar_solver_set_delta_t(afar_solver* h, double* dt);
/* basicCUDAsolver::evolve (basicCUDAsolver.cpp:862-871): one SSP-RK3 step of size dt.
 * Asynchronous: returns once the work is queued on the solver's stream. */
int afar_solver_evolve(afar_solver* h, double dt);
/* nsteps of (CFL time step + RK3 step) with the time step kept on the
 * device; one captured hipGraph per step.  Asynchronous, with at most 512
 * steps left queued: a longer call waits for earlier steps as it goes. */
int afar_solver_evolve_auto(afar_solver* h, int32_t nsteps);
/* Time step used by the last evolve_auto step (synchronises). */
int afar_solver_last_dt(afar_solver* h, double* dt);
/* Device clock of evolve_auto: the solver is at time t and its device time
 * steps are clipped to end exactly at t_end (dt = min(dt, t_end - t), as the
 * reference application's Foam::Time does); steps past t_end leave the state
 * unchanged.  Initially t = 0, t_end = +inf.  get_time gives the time and the
 * number of steps taken since the last set_time (either pointer may be NULL). */
int afar_solver_set_time(afar_solver* h, double t, double t_end);
int afar_solver_get_time(afar_solver* h, double* t, double* steps);
/* sqrt(mean over cells of (sum_f F_rho / V)^2) for the current state (runs
 * the stage kernels; the switch flags become those of the current state). */
int afar_solver_residual_norm(afar_solver* h, double* norm);
int afar_solver_synchronize(afar_solver* h);
/* *idle = 1 when all work enqueued on the solver's stream has finished, else 0; does not wait. */
int afar_solver_query(afar_solver* h, int32_t* idle);
/* basicCUDAsolver::writeField (basicCUDAsolver.cpp:874-913): W 4*nc and
 * shock nc (final reconstruction type, shockIndicator[2c+1]) in the caller's
 * order, ghost 4*nb in boundary-face order.  NULL pointers are skipped. */
int afar_solver_write_field(afar_solver* h, double* W, double* ghost, int8_t* shock);
/* Conserved state rho, rhoUx, rhoUy, rhoE (4*nc, caller's order): exact
 * checkpoint / restart of the FP64 state (the reference rebuilds it from p, U, T). */
int afar_solver_get_conserved(afar_solver* h, double* U);
int afar_solver_set_conserved(afar_solver* h, const double* U);
/* hipEvent time of nsteps auto-dt steps on the solver's stream (graph replay if use_graph). */
int afar_solver_time_steps(afar_solver* h, int32_t nsteps, int32_t use_graph, double* ms);
/* nsteps auto-dt steps replayed as a hipGraph with an event after every
 * kernel (two graph copies alternate, so the host reads one step's events
 * while the next runs): *ms_total for the nsteps, and per kernel kind (the
 * order of afar_solver_profile) the summed kernel time ms_by_kind[7] and the
 * number of launches launches_by_kind[7] over those same steps.  Extension
 * without a reference counterpart: per-kernel timing of the timed steps. */
int afar_solver_time_steps_kernels(afar_solver* h, int32_t nsteps, double* ms_total, double* ms_by_kind,
                                   int32_t* launches_by_kind);
/* Average time of `reps` back-to-back launches of each kernel: ms[7] and
 * launches per device-time-step RK3 step count[7] for {cfl (standalone, used
 * by set_delta_t), reconstruct, bvd, flux, update, flux with CFL partials,
 * CFL partial reduction}. */
int afar_solver_profile(afar_solver* h, int32_t reps, double* ms, int32_t* count);
/* Device sizes: out[0..7] = nc, nf, nb, maxS, maxC, tile, ntiles, maxHalo. */
int afar_solver_sizes(afar_solver* h, int64_t* out);
/* Test hooks: run stage kernels (0 reconstruct, 1 +bvd, 2 +flux, 3 +update of
 * RK stage `stage` with step dt) and copy named device arrays (mesh order):
 * W, U, gp, gu, gv, gT, lim (the raw K5 limiter: the flux pass applies K6's
 * transform as it reads it; -1 marks a first-order cell), smooth, ftype, fv,
 * fvg, flux (FP64 for every face in both precisions), lsq. */
int afar_solver_run_stage(afar_solver* h, int32_t upto, double dt, int32_t stage);
int afar_solver_fetch(afar_solver* h, const char* name, void* out, size_t bytes);
void afar_solver_destroy(afar_solver* h);

/* ------------------------------------------------------------------ batches
 * Extension without a reference counterpart (the reference solves one mesh
 * per process): n independent precomputed meshes advanced together.  The
 * meshes are stored one after another in one device layout, so every kernel
 * launch of a step covers all of them and one graph launch advances the whole
 * batch by one RK3 step; each mesh keeps its own CFL time step.  Each mesh
 * gets exactly the trajectory it gets alone (same arithmetic per cell).  One
 * schemes (gas constants, CFL, precision, tile) for all; W0[k] is the initial
 * state of mesh k in its caller order.  Every mesh needs >= 256 internal faces. */
typedef struct afar_batch afar_batch;
int afar_batch_create(int32_t n, const afar_mesh* const* meshes, const afar_schemes* s, const double* const* W0,
                      afar_batch** out);
/* nsteps RK3 steps of every mesh, device time steps, asynchronous. */
int afar_batch_evolve(afar_batch* b, int32_t nsteps);
/* Same, synchronous, hipEvent milliseconds in *ms. */
int afar_batch_time_steps(afar_batch* b, int32_t nsteps, double* ms);
int afar_batch_synchronize(afar_batch* b);
/* Back to the initial states. */
int afar_batch_reset(afar_batch* b);
/* out[0..7] = meshes, cells, internal faces, boundary faces, tile, maxS, maxC, maxHalo. */
int afar_batch_sizes(afar_batch* b, int64_t* out);
/* off[n+1]: first cell of each mesh in the batch layout. */
int afar_batch_cell_offsets(afar_batch* b, int32_t* off);
/* dt[n]: the time step each mesh took in its last step. */
int afar_batch_last_dt(afar_batch* b, double* dt);
/* Device clocks as afar_solver_set_time, for mesh inst (-1: all); t[n], steps[n]. */
int afar_batch_set_time(afar_batch* b, int32_t inst, double t, double t_end);
int afar_batch_get_time(afar_batch* b, double* t, double* steps);
/* Conserved state (4 per cell) of mesh inst, its caller order. */
int afar_batch_get_conserved(afar_batch* b, int32_t inst, double* U);
/* Primitive state and final reconstruction types of mesh inst; either may be NULL. */
int afar_batch_write_field(afar_batch* b, int32_t inst, double* W, int8_t* shock);
/* Per-kernel times as afar_solver_profile. */
int afar_batch_profile(afar_batch* b, int32_t reps, double* ms, int32_t* count);
void afar_batch_destroy(afar_batch* b);

/* Reference-layout arrays of a precomputed mesh (what seam 2 receives), for
 * checking against a host implementation.  The pointers in `out` must be
 * allocated by the caller with the sizes listed at afar_field_pointer. */
typedef struct afar_field_pointer afar_field_pointer;
int afar_mesh_export_reference(const afar_mesh* m, const afar_schemes* s, const double* W0, afar_field_pointer* out);

/* ------------------------------------------------------------------ seam 2: GPUevolve.h
 * Host arrays of fieldPointer (src/CUDA/cudaEuler2DmixO3RK3/basicDataStructure.h:
 * 75-160), same meaning and layout.  Device-only members (gradFD, Flux, Res,
 * faceFD, matrix, minDeltaT) are internal.  The extension members may be NULL. */
struct afar_field_pointer {
    double* CELL;                      /* 2*(nc+nb): x, y (cells, then boundary-face centres) */
    double* FACE;                      /* 5*(nf+nb): x, y, magSf, Sfx, Sfy */
    double* cellVolume;                /* nc */
    double* FD;                        /* 4*(nc+nb): p, Ux, Uy, T */
    int8_t* shockIndicator;            /* 2*(nc+nb) */
    double* cFD;                       /* 4*nc: rho, rhoUx, rhoUy, rhoE */
    int32_t* neighbour;                /* 2*nf: own, nei */
    int32_t* boundaryFacesNum;         /* npatches */
    int32_t* boundaryFaceNeiLabel;     /* nb */
    uint8_t* boundaryFacesType;        /* nb */
    uint8_t* stencilSize;              /* nc */
    uint8_t* compactStencilSize;       /* nc */
    float* RBFbasis;                   /* 2*(maxS+1)*nf */
    int32_t* extendStencil;            /* (maxLB-128)*ceil(nc/128) */
    uint16_t* extendStencilSize;       /* ceil(nc/128) */
    uint16_t* compactExtendStencilSize;/* ceil(nc/128) */
    uint16_t* localStencil;            /* maxS*nc, 128-cell block-local ids */
    int32_t* cellFaces;                /* maxC*nc */
    /* extensions */
    int32_t* globalStencil;            /* maxS*nc global ids; replaces the block tables when set */
    double* RBFbasis64;                /* 2*(maxS+1)*nf FP64 coefficients; replaces RBFbasis when set */
    double* faceShift;                 /* 2*nf periodic shift of the neighbour centre */
    double* boundaryValue;             /* 4*nb fixed values of label 6 */
};

/* setDeviceFieldData #1 (GPUevolve.cu:13-64): geometry and state.  The counts
 * are the reference's int arguments widened to 64 bits (a C/C++ caller
 * passing int compiles unchanged). */
int afar_set_device_field_data(afar_device** dev, const afar_field_pointer* host, int64_t cellsNum,
                               int64_t facesNum, int64_t totalBoundaryFacesNum);
/* Optional, between the two uploads: precision (0 FP64, 1 FP32-mixed), tile, R, Cv. */
int afar_device_configure(afar_device* dev, int32_t precision, int32_t tile);
/* Optional, between the two uploads: run on the caller's stream (a hipStream_t
 * passed as void*) instead of one of the device's own; NULL keeps the
 * default.  No reference counterpart (the reference uses the legacy default
 * stream); the stream must outlive the device. */
int afar_device_set_stream(afar_device* dev, void* stream);
/* setDeviceFieldData #2 (GPUevolve.cu:67-148): connectivity, stencils,
 * coefficients; computes the least-squares weights on the device (K3). */
int afar_set_device_stencil_data(afar_device* dev, const afar_field_pointer* host, int64_t cellsNum,
                                 int64_t facesNum, int64_t patchesNum, int64_t totalBoundaryFacesNum,
                                 int64_t maxStencilSize, int64_t maxCompactStencilSize,
                                 int64_t maxLocalBlockStencilSize, int64_t maxCompactLocalBlockStencilSize);
/* freeFieldData (GPUevolve.cu:151-174): releases the device and the handle. */
int afar_free_field_data(afar_device* dev);
/* adjustTimeStep (GPUevolve.cu:177-204). */
int afar_adjust_time_step(afar_device* dev, int64_t facesNum, double R, double Cv, double CFL, double* deltaT);
/* GPUevolve (GPUevolve.cu:207-306): one SSP-RK3 step; synchronous like the reference. */
int afar_gpu_evolve(afar_device* dev, double R, double Cv, double deltaT, int64_t cellsNum, int64_t facesNum,
                    int64_t totalBoundaryFacesNum, int64_t maxStencilSize, int64_t maxCompactStencilSize,
                    int64_t maxLocalBlockStencilSize, int64_t maxCompactLocalBlockStencilSize);
/* copyFieldDataDeviceToHost (GPUevolve.cu:309-318): FD and shockIndicator. */
int afar_copy_field_data_device_to_host(afar_field_pointer* host, afar_device* dev, int64_t cellsNum,
                                        int64_t totalBoundaryFacesNum);

/* adjustTimeStep without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The CFL time step
 * over the caller's DEVICE arrays in the reference's layout -- FD 4*(nc+nb),
 * CELL 2*(nc+nb), FACE 5*(nf+nb), neighbour 2*nf, FSHIFT 2*nf or NULL (the
 * periodic shift of the neighbour centre) -- written to *deltaT in device
 * memory, asynchronously on `stream` (a hipStream_t, NULL = the null
 * stream).  Its partial minima live in an 8 KiB static device array (1024
 * partials) per GPU, held until the process ends: calls that overlap on one
 * GPU race; facesNum above 2^31 - 1 is refused.  Equal bit for bit to
 * afar_adjust_time_step on the same state. */
int afar_adjust_time_step_noalloc(const double* FD, const double* CELL, const double* FACE, const int32_t* neighbour,
                                  const double* FSHIFT, int64_t facesNum, double R, double Cv, double CFL,
                                  double* deltaT, void* stream);

/* ------------------------------------------------------------------ diagnostics */
/* Evaluates exp, erfc, |x|^(1/2) and 1/x of the FP64 flux path on the device: out 4*n. */
int afar_selftest_math(const double* x, double* out, int32_t n);
/* Same functions on the host. */
int afar_host_math(const double* x, double* out, int32_t n);
/* The same four in single precision, as the FP32-mixed flux path evaluates
 * them (exp and erfc: afar_expf, afar_erfcf), on the device and on the host. */
int afar_selftest_math_f32(const float* x, float* out, int32_t n);
int afar_host_math_f32(const float* x, float* out, int32_t n);
/* Name and compute units of the current device. */
int afar_device_info(char* name, int32_t cap, int32_t* cus, int64_t* hbm_bytes);

#ifdef __cplusplus
}
#endif

#endif /* AFAR_H */
