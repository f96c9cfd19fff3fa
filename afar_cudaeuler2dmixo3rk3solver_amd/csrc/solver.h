// Device solver: owns the HIP buffers of one mesh, its stream and the
// captured step graph, and drives the stage pipeline.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "layout.h"

namespace afar {

struct SolverConfig {
    double R = 0.7142857, Cv = 1.7857142857, CFL = 0.5;
    int precision = 0;  // 0 FP64, 1 FP32-mixed
    int exactRK = 0;
    int tile = 64;
    int boundaryFirstOrder = 0;
    int reconLanes = 0;  // reconstruction kernel: 0 automatic, 1 one lane per cell, 4 four lanes
    int cellOrder = 0;   // device cell order: 0 compact tiles (Hilbert curve) with blocks of 32
                         // (FP64) or 64 (mixed) cells ordered by face count and stencil size,
                         // 1 the mesh's order, 2 compact tiles in Hilbert order throughout
    hipStream_t stream = nullptr;  // the caller's stream (not owned), or nullptr: the solver creates its own
};

// Held by every library call that allocates or frees device memory, creates
// or destroys a solver, or captures or destroys a graph.  A stream under
// capture must not be waited on, and some of these calls wait on the device
// (a plain hipFree, the synchronous copies of an upload): serialising them
// with every capture makes solvers safe to create, step and destroy from
// different host threads.  Solver buffers are stream-ordered (hipMallocAsync /
// hipFreeAsync on the solver's stream), so a destroy waits for its own stream
// only.  Kernel and graph launches, copies and queries on a solver's own
// stream do not take the lock.
std::mutex& runtime_mutex();

// kernel kinds reported by profile()
// Kernel kinds of profile(): the standalone CFL pass (host-requested time
// steps), the four stage kernels, the flux pass that also reduces the CFL
// bound (first stage of a device-time-step RK3 step) and its final reduction.
enum {
    kKernCfl = 0,
    kKernRecon = 1,
    kKernBvd = 2,
    kKernFlux = 3,
    kKernUpdate = 4,
    kKernFluxDt = 5,
    kKernDtReduce = 6,
    kKernKinds = 7
};

class DeviceSolver {
   public:
    virtual ~DeviceSolver() = default;
    virtual double compute_dt() = 0;                    // CFL pass, result copied to the host
    virtual void evolve(double dt) = 0;                 // one SSP-RK3 step with a host time step
    virtual void evolve_auto(int nsteps) = 0;           // CFL + step on the device, graph replay
    virtual void synchronize() = 0;
    virtual bool idle() = 0;  // all enqueued work done (does not wait)
    virtual void copy_out(double* FD, int8_t* shock) = 0;  // FD 4*(nc+nb); shock 2*(nc+nb) interleaved
    virtual void get_conserved(double* U) = 0;          // 4*nc
    virtual void set_conserved(const double* U) = 0;    // also recomputes W and ghosts
    virtual double last_dt() = 0;
    // Instances of a batch (1 for one mesh) and the time step of each.
    virtual int ninst() const = 0;
    virtual void last_dts(double* out) = 0;
    // Cells [c0, c0+n): conserved state (4n); primitive state (4n) and final
    // reconstruction types (n); either output may be NULL.
    virtual void get_conserved_range(int c0, int n, double* U) = 0;
    virtual void get_primitive_range(int c0, int n, double* W, int8_t* type) = 0;
    // Back to the state the solver was created with (or last given to
    // set_conserved); the device clock is not changed.
    virtual void reset_state() = 0;
    // Device clock of the device-time-step path: instance inst (-1: all) is
    // at time t and stops at t_end (steps past it leave the state alone);
    // its count of steps taken restarts at 0.
    virtual void set_time(int inst, double t, double t_end) = 0;
    virtual void get_time(double* t, double* steps) = 0;  // ninst values each; either may be NULL
    // Root mean square over cells of the density residual per unit volume,
    // |sum_f F_rho| / V, of the current state (runs the stage kernels, so the
    // switch flags become those of the current state).
    virtual double residual_norm() = 0;
    virtual double time_steps(int nsteps, bool graph) = 0;  // hipEvent milliseconds for nsteps
    // nsteps graph-replayed steps with events around every kernel: total ms,
    // and per kernel kind the summed ms and launch count over those steps
    virtual double time_steps_kernels(int nsteps, double* ms_kind, int* cnt_kind) = 0;
    virtual void profile(double* ms, int* count, int reps) = 0;
    virtual const DevLayout& layout() const = 0;
    virtual void debug_fetch(const char* what, void* out, size_t bytes) = 0;
    // Runs reconstruct, then (upto >= 1) bvd, (>= 2) flux, (>= 3) update of RK stage `stage`.
    virtual void run_stage_kernels(int upto, double dt, int stage) = 0;
    const SolverConfig& config() const { return cfg_; }
    // Gas constants and CFL number; a captured step graph is rebuilt on change.
    virtual void set_constants(double R, double Cv, double CFL) = 0;

   protected:
    SolverConfig cfg_;
};

// bi: instance offsets when ra is a batch (union_ref_arrays), else nullptr.
// A batch steps with the device time step only (evolve_auto, time_steps).
std::unique_ptr<DeviceSolver> make_device_solver(const RefArrays& ra, const SolverConfig& cfg,
                                                 const BatchIndex* bi = nullptr);
void check_hip(hipError_t e, const char* what);

}  // namespace afar
