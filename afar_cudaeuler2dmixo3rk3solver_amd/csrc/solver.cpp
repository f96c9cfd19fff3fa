// Device solver implementation: buffer management, the RK3 stage loop
// (src/CUDA/cudaEuler2DmixO3RK3/GPUevolve.cu:177-306) and graph capture.
#include "solver.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <mutex>

#include "kernels.h"

namespace afar {

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(2, std::string(what) + ": " + hipGetErrorString(e));
}

std::mutex& runtime_mutex() {
    static std::mutex m;
    return m;
}

namespace {

// Device buffer, allocated and freed in stream order on its solver's stream
// (a plain hipFree waits for the whole device, so tearing one solver down
// would stall every other solver's work).  The owner synchronises the stream
// before the buffer is used elsewhere.
struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    hipStream_t s = nullptr;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() {
        if (p) (void)hipFreeAsync(p, s);
    }
    void alloc(size_t n, hipStream_t st) {
        if (p) check_hip(hipFreeAsync(p, s), "hipFreeAsync");
        p = nullptr;
        bytes = n;
        s = st;
        check_hip(hipMallocAsync(&p, n ? n : 16, st), "hipMallocAsync");
    }
    template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

// The solver's stream: created here unless the caller gave one; destroyed
// after the buffers (members declared after it) have enqueued their frees.
struct StreamHolder {
    hipStream_t s = nullptr;
    bool own = false;
    ~StreamHolder() {
        if (s) {
            (void)hipStreamSynchronize(s);
            if (own) (void)hipStreamDestroy(s);
        }
    }
};

template <typename R>
class SolverImpl final : public DeviceSolver {
    using v4 = typename VT<R>::v4;
    using v2 = typename VT<R>::v2;

   public:
    // ra and its device layout L (make_dev_layout(ra, cfg.tile)) are built by
    // the caller, outside the library's lock, and moved in.
    SolverImpl(RefArrays&& ra, DevLayout&& L, const SolverConfig& cfg, const BatchIndex* bi)
        : ra_(std::move(ra)), L_(std::move(L)) {
        cfg_ = cfg;
        if (bi) {
            bi_ = *bi;
            for (int k = 0; k < bi_.ninst(); ++k)
                if (bi_.face_off[k + 1] - bi_.face_off[k] < 256)
                    throw Error(1, "every mesh of a batch needs at least 256 internal faces");
        } else {
            bi_.cell_off = {0, ra_.nc};
            bi_.face_off = {0, ra_.nf};
            bi_.bface_off = {0, ra_.nb};
        }
        if (cfg.stream) {
            sh_.s = cfg.stream;  // the caller's stream: used, never destroyed
        } else {
            check_hip(hipStreamCreateWithFlags(&sh_.s, hipStreamNonBlocking), "hipStreamCreate");
            sh_.own = true;
        }
        stream_ = sh_.s;
        check_hip(hipEventCreate(&ev0_), "hipEventCreate");
        check_hip(hipEventCreate(&ev1_), "hipEventCreate");
        for (hipEvent_t& e : evq_) check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        // LDS of the reconstruction: tile + halo states, then the tile's stencil ids
        lds_recon_ = (size_t)(L_.tile + L_.maxHalo) * sizeof(v4) + (size_t)L_.tile * L_.maxS * sizeof(uint16_t);
        lds_m_off_ = (lds_recon_ + 15) & ~(size_t)15;
        if (cfg.reconLanes == 4 && L_.maxC > 4) throw Error(1, "reconLanes 4 needs cells with at most four faces");
        if (cfg.reconLanes != 0)
            lanes_ = cfg.reconLanes;
        else
            lanes_ = L_.maxC > 4 ? 1 : 4;  // four lanes measured fastest at 64 and 128-cell tiles
        // the four-lane kernel also stages the tile's least-squares weights
        // (FP64 at the face centre: the tile's and halo's centres in their place)
        if (lanes_ == 4)
            lds_recon_ = lds_m_off_ + std::max((size_t)L_.tile * 4 * sizeof(v2),
                                               sizeof(R) == 8 ? (size_t)(L_.tile + L_.maxHalo) * 16 : (size_t)0);
        // BVD: the tile's and halo's smoothness flags, then the stencil ids
        lds_bvd_sten_off_ = ((size_t)(L_.tile + L_.maxHalo) + 15) & ~(size_t)15;
        lds_bvd_ = lds_bvd_sten_off_ + (size_t)L_.tile * L_.maxS * sizeof(uint16_t);
        int dev = 0;
        check_hip(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t prop;
        check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        size_t lds_cap = prop.sharedMemPerBlock;
        if (lds_recon_ > lds_cap)
            throw Error(3, "Too large local block stencil size: reconstruct needs " +
                               std::to_string(lds_recon_) + " bytes of LDS, the device allows " +
                               std::to_string(lds_cap));
        // The reconstruction kernels' dynamic-LDS limit is raised to the
        // device's per-block maximum once per process: the attribute is
        // global per kernel, so it is never changed while other solvers (on
        // other host threads) may be launching those kernels.
        static std::once_flag lds_once;
        static hipError_t lds_err = hipSuccess;
        std::call_once(lds_once, [&] { lds_err = set_lds_limits<R>(lds_cap); });
        check_hip(lds_err, "hipFuncSetAttribute");
        upload();
    }
    ~SolverImpl() override {
        // no launch of the graph is in flight when it is destroyed; the
        // buffers' frees follow, in stream order
        if (stream_) (void)hipStreamSynchronize(stream_);
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        (void)hipEventDestroy(ev0_);
        (void)hipEventDestroy(ev1_);
        for (hipEvent_t e : evq_)
            if (e) (void)hipEventDestroy(e);
    }

    double compute_dt() override {
        if (ninst() > 1) throw Error(1, "a batch steps with the device time step only");
        launch_cfl<R>(P_, cfg_.R, cfg_.Cv, cfg_.CFL, stream_);
        check_hip(hipGetLastError(), "k_cfl");
        double dt = 0.0;
        check_hip(hipMemcpyAsync(&dt, P_.dt, sizeof(double), hipMemcpyDeviceToHost, stream_), "dt copy");
        check_hip(hipStreamSynchronize(stream_), "dt sync");
        return dt;
    }

    void evolve(double dt) override {
        if (ninst() > 1) throw Error(1, "a batch steps with the device time step only");
        enqueue_stages(dt, nullptr);
        state_in_uold_ = true;
        check_hip(hipGetLastError(), "evolve");
    }

    void evolve_auto(int nsteps) override {
        if (!graph_exec_) capture();
        for (int i = 0; i < nsteps; ++i) launch_graph();
    }

    void synchronize() override { check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

    bool idle() override {
        const hipError_t e = hipStreamQuery(stream_);
        if (e == hipErrorNotReady) return false;
        check_hip(e, "hipStreamQuery");
        return true;
    }

    int ninst() const override { return bi_.ninst(); }

    void last_dts(double* out) override {
        check_hip(hipMemcpyAsync(out, P_.dt, sizeof(double) * ninst(), hipMemcpyDeviceToHost, stream_), "dt copy");
        synchronize();
    }

    double last_dt() override {
        double dt = 0.0;
        check_hip(hipMemcpyAsync(&dt, P_.dt, sizeof(double), hipMemcpyDeviceToHost, stream_), "dt copy");
        synchronize();
        return dt;
    }

    void copy_out(double* FD, int8_t* shock) override {
        const int nc = L_.nc, nb = L_.nb;
        std::vector<v4> w(nc + nb);
        std::vector<int8_t> s(nc + nb), t(nc);
        check_hip(hipMemcpyAsync(w.data(), P_.W, sizeof(v4) * (nc + nb), hipMemcpyDeviceToHost, stream_), "W copy");
        check_hip(hipMemcpyAsync(s.data(), P_.smooth, nc + nb, hipMemcpyDeviceToHost, stream_), "flag copy");
        check_hip(hipMemcpyAsync(t.data(), P_.ftype, nc, hipMemcpyDeviceToHost, stream_), "type copy");
        synchronize();
        if (FD)
            for (int i = 0; i < nc + nb; ++i) {
                FD[4 * i] = w[i].x;
                FD[4 * i + 1] = w[i].y;
                FD[4 * i + 2] = w[i].z;
                FD[4 * i + 3] = w[i].w;
            }
        if (shock)
            for (int i = 0; i < nc + nb; ++i) {
                shock[2 * i] = s[i];
                shock[2 * i + 1] = i < nc ? t[i] : 0;
            }
    }

    void get_conserved(double* U) override {
        check_hip(hipMemcpyAsync(U, state_U(), sizeof(double4) * L_.nc, hipMemcpyDeviceToHost, stream_), "U copy");
        synchronize();
    }

    void get_conserved_range(int c0, int n, double* U) override {
        if (c0 < 0 || n < 0 || c0 + n > L_.nc) throw Error(1, "cell range out of bounds");
        check_hip(hipMemcpyAsync(U, state_U() + c0, sizeof(double4) * n, hipMemcpyDeviceToHost, stream_), "U copy");
        synchronize();
    }

    void get_primitive_range(int c0, int n, double* W, int8_t* type) override {
        if (c0 < 0 || n < 0 || c0 + n > L_.nc) throw Error(1, "cell range out of bounds");
        std::vector<v4> w(W ? n : 0);
        if (W) check_hip(hipMemcpyAsync(w.data(), P_.W + c0, sizeof(v4) * n, hipMemcpyDeviceToHost, stream_), "W copy");
        if (type) check_hip(hipMemcpyAsync(type, P_.ftype + c0, n, hipMemcpyDeviceToHost, stream_), "type copy");
        synchronize();
        for (int i = 0; i < (W ? n : 0); ++i) {
            W[4 * (size_t)i] = w[i].x;
            W[4 * (size_t)i + 1] = w[i].y;
            W[4 * (size_t)i + 2] = w[i].z;
            W[4 * (size_t)i + 3] = w[i].w;
        }
    }

    void reset_state() override { upload_state(); }

    void set_time(int inst, double t, double t_end) override {
        if (inst < -1 || inst >= ninst()) throw Error(1, "instance out of range");
        if (!(t_end >= t)) throw Error(1, "end time before the current time");
        std::vector<double> tc(4 * (size_t)ninst());
        check_hip(hipMemcpyAsync(tc.data(), P_.tctl, sizeof(double) * tc.size(), hipMemcpyDeviceToHost, stream_),
                  "clock copy");
        synchronize();
        for (int k = 0; k < ninst(); ++k)
            if (inst < 0 || inst == k) {
                tc[4 * k] = t;
                tc[4 * k + 1] = t_end;
                tc[4 * k + 2] = 0.0;
            }
        check_hip(hipMemcpyAsync(P_.tctl, tc.data(), sizeof(double) * tc.size(), hipMemcpyHostToDevice, stream_),
                  "clock copy");
        synchronize();
    }

    void get_time(double* t, double* steps) override {
        std::vector<double> tc(4 * (size_t)ninst());
        check_hip(hipMemcpyAsync(tc.data(), P_.tctl, sizeof(double) * tc.size(), hipMemcpyDeviceToHost, stream_),
                  "clock copy");
        synchronize();
        for (int k = 0; k < ninst(); ++k) {
            if (t) t[k] = tc[4 * k];
            if (steps) steps[k] = tc[4 * k + 2];
        }
    }

    double residual_norm() override {
        launch_reconstruct<R>(P_, L_.tile, lds_recon_, stream_, lanes_);
        launch_bvd<R>(P_, L_.tile, lds_bvd_, stream_);
        launch_flux<R>(P_, cfg_.R, cfg_.Cv, stream_);
        launch_residual_norm<R>(P_, P_.res_part + residual_blocks(L_.nc), stream_);
        check_hip(hipGetLastError(), "residual");
        double s = 0.0;
        check_hip(hipMemcpyAsync(&s, P_.res_part + residual_blocks(L_.nc), sizeof(double), hipMemcpyDeviceToHost,
                                 stream_),
                  "residual copy");
        synchronize();
        return sqrt(s / L_.nc);
    }

    void set_conserved(const double* U) override {
        // primitive state from the conserved one, as the update kernel does it
        const double gamma = cfg_.R / cfg_.Cv + 1.0;
        for (int i = 0; i < L_.nc; ++i) {
            const double rho = U[4 * i], rhoUx = U[4 * i + 1], rhoUy = U[4 * i + 2], rhoE = U[4 * i + 3];
            const double p = (gamma - 1.0) * (rhoE - 0.5 * (rhoUx * rhoUx + rhoUy * rhoUy) / rho);
            ra_.FD[4 * i] = p;
            ra_.FD[4 * i + 1] = rhoUx / rho;
            ra_.FD[4 * i + 2] = rhoUy / rho;
            ra_.FD[4 * i + 3] = p / (cfg_.R * rho);
        }
        apply_bc_host(ra_);
        ra_.cFD.assign(U, U + 4 * (size_t)L_.nc);
        upload_state();
    }

    double time_steps(int nsteps, bool graph) override {
        if (graph && !graph_exec_) capture();
        synchronize();
        check_hip(hipEventRecord(ev0_, stream_), "hipEventRecord");
        if (graph) {
            for (int i = 0; i < nsteps; ++i) launch_graph();
        } else {
            for (int i = 0; i < nsteps; ++i) enqueue_stages(0.0, P_.dt);
            if (nsteps > 0) state_in_uold_ = true;
        }
        check_hip(hipEventRecord(ev1_, stream_), "hipEventRecord");
        check_hip(hipEventSynchronize(ev1_), "hipEventSynchronize");
        float ms = 0.f;
        check_hip(hipEventElapsedTime(&ms, ev0_, ev1_), "hipEventElapsedTime");
        return ms;
    }

    // nsteps device-time-step RK3 steps as graph replays, with an event before
    // and after every kernel inside the graph: the total time (returned, ms)
    // and each kind's summed kernel time and launch count over these very
    // steps.  Two graph copies with their own events alternate, so the events
    // of one step are read while the next runs; the device never waits for
    // the host.
    double time_steps_kernels(int nsteps, double* ms_kind, int* cnt_kind) override {
        constexpr int NK = 13;
        static const int kind[NK] = {kKernRecon, kKernBvd, kKernFluxDt, kKernDtReduce, kKernUpdate,
                                     kKernRecon, kKernBvd, kKernFlux,   kKernUpdate,   kKernRecon,
                                     kKernBvd,   kKernFlux, kKernUpdate};
        for (int k = 0; k < kKernKinds; ++k) {
            ms_kind[k] = 0.0;
            cnt_kind[k] = 0;
        }
        if (nsteps <= 0) return 0.0;
        struct Copy {
            hipEvent_t ev[NK + 1] = {};
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
        } cp[2];
        struct Cleanup {
            Copy* c;
            hipStream_t s;
            ~Cleanup() {
                (void)hipStreamSynchronize(s);  // an error may leave launches of the copies in flight
                std::lock_guard<std::mutex> lk(runtime_mutex());
                for (int b = 0; b < 2; ++b) {
                    if (c[b].ge) (void)hipGraphExecDestroy(c[b].ge);
                    if (c[b].g) (void)hipGraphDestroy(c[b].g);
                    for (hipEvent_t e : c[b].ev)
                        if (e) (void)hipEventDestroy(e);
                }
            }
        } cleanup{cp, stream_};
        synchronize();
        std::unique_lock<std::mutex> build_lock(runtime_mutex());
        for (int b = 0; b < 2; ++b) {
            for (hipEvent_t& e : cp[b].ev) check_hip(hipEventCreate(&e), "hipEventCreate");
            check_hip(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            check_hip(hipEventRecordWithFlags(cp[b].ev[0], stream_, hipEventRecordExternal), "hipEventRecord");
            enqueue_stages(0.0, P_.dt, cp[b].ev + 1);
            check_hip(hipStreamEndCapture(stream_, &cp[b].g), "hipStreamEndCapture");
            check_hip(hipGraphInstantiate(&cp[b].ge, cp[b].g, nullptr, nullptr, 0), "hipGraphInstantiate");
        }
        build_lock.unlock();
        auto collect = [&](Copy& c) {
            check_hip(hipEventSynchronize(c.ev[NK]), "hipEventSynchronize");
            for (int j = 0; j < NK; ++j) {
                float ms = 0.f;
                check_hip(hipEventElapsedTime(&ms, c.ev[j], c.ev[j + 1]), "hipEventElapsedTime");
                ms_kind[kind[j]] += ms;
                cnt_kind[kind[j]] += 1;
            }
        };
        check_hip(hipEventRecord(ev0_, stream_), "hipEventRecord");
        for (int i = 0; i < nsteps; ++i) {
            if (i >= 2) collect(cp[i & 1]);  // step i-2 used this copy
            check_hip(hipGraphLaunch(cp[i & 1].ge, stream_), "hipGraphLaunch");
            state_in_uold_ = true;
        }
        check_hip(hipEventRecord(ev1_, stream_), "hipEventRecord");
        for (int i = std::max(0, nsteps - 2); i < nsteps; ++i) collect(cp[i & 1]);
        check_hip(hipEventSynchronize(ev1_), "hipEventSynchronize");
        float total = 0.f;
        check_hip(hipEventElapsedTime(&total, ev0_, ev1_), "hipEventElapsedTime");
        return total;
    }

    // Average duration of each kernel: `reps` back-to-back launches of that
    // kernel between two events on the solver's stream (the bvd and update
    // launches keep transforming the state, so call this after timed work).
    void profile(double* ms, int* count, int reps) override {
        synchronize();
        for (int k = 0; k < kKernKinds; ++k) {
            check_hip(hipEventRecord(ev0_, stream_), "hipEventRecord");
            for (int r = 0; r < reps; ++r) {
                switch (k) {
                    case kKernCfl: launch_cfl<R>(P_, cfg_.R, cfg_.Cv, cfg_.CFL, stream_); break;
                    case kKernRecon: launch_reconstruct<R>(P_, L_.tile, lds_recon_, stream_, lanes_); break;
                    case kKernBvd: launch_bvd<R>(P_, L_.tile, lds_bvd_, stream_); break;
                    case kKernFlux: launch_flux<R>(P_, cfg_.R, cfg_.Cv, stream_); break;
                    case kKernFluxDt: launch_flux_dt<R>(P_, cfg_.R, cfg_.Cv, cfg_.CFL, stream_); break;
                    case kKernDtReduce: launch_dt_reduce<R>(P_, stream_); break;
                    default: launch_update<R>(P_, cfg_.R, gamma(), 0.0, P_.dt, b1(1), b2(1), b3(1), 0, stream_);
                }
            }
            check_hip(hipEventRecord(ev1_, stream_), "hipEventRecord");
            check_hip(hipEventSynchronize(ev1_), "hipEventSynchronize");
            float t = 0.f;
            check_hip(hipEventElapsedTime(&t, ev0_, ev1_), "hipEventElapsedTime");
            ms[k] = t / reps;
        }
        state_in_uold_ = false;  // the update launches above wrote U
        // launches per RK3 step with the device-side time step
        count[kKernCfl] = 0;
        count[kKernRecon] = count[kKernBvd] = count[kKernUpdate] = 3;
        count[kKernFlux] = 2;
        count[kKernFluxDt] = count[kKernDtReduce] = 1;
        for (int k = 0; k < kKernKinds; ++k) {
        }
    }

    const DevLayout& layout() const override { return L_; }

    void set_constants(double Rgas, double Cv, double CFL) override {
        if (Rgas == cfg_.R && Cv == cfg_.Cv && CFL == cfg_.CFL) return;
        cfg_.R = Rgas;
        cfg_.Cv = Cv;
        cfg_.CFL = CFL;
        synchronize();
        std::lock_guard<std::mutex> lk(runtime_mutex());
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        graph_exec_ = nullptr;
        graph_ = nullptr;
    }

    void debug_fetch(const char* what, void* out, size_t bytes) override {
        const void* src = nullptr;
        size_t have = 0;
        auto pick = [&](const char* name, const void* p, size_t n) {
            if (strcmp(what, name) == 0) {
                src = p;
                have = n;
            }
        };
        const size_t nc = L_.nc, nf = L_.nf, nb = L_.nb;
        pick("W", P_.W, sizeof(v4) * (nc + nb));
        pick("U", state_U(), sizeof(double4) * nc);
        pick("gp", P_.gp, sizeof(v2) * nc);
        pick("gu", P_.gu, sizeof(v2) * nc);
        pick("gv", P_.gv, sizeof(v2) * nc);
        pick("gT", P_.gT, sizeof(v2) * nc);
        pick("lim", P_.lim, sizeof(v4) * nc);
        pick("smooth", P_.smooth, nc + nb);
        pick("ftype", P_.ftype, nc);
        pick("fv", P_.fv ? (const void*)P_.fv : (const void*)P_.fvp, sizeof(v4) * 2 * nf);
        pick("fvg", P_.fvg, sizeof(v4) * 2 * nf * L_.ngp);
        pick("flux", P_.flux, sizeof(double4) * (nf + nb));
        pick("lsq", P_.lsq, sizeof(v2) * (size_t)L_.nwaves * L_.maxC * 64);
        if (!src) throw Error(1, std::string("unknown device array ") + what);
        if (bytes != have) throw Error(1, std::string("size mismatch fetching ") + what);
        synchronize();
        if (strcmp(what, "fv") == 0 && !P_.fv) {  // (p, Ux, Uy, T) records from the per-variable arrays
            std::vector<R> a[4];
            const R* srcs[4] = {P_.fvp, P_.fvu, P_.fvv, P_.fvT};
            for (int k = 0; k < 4; ++k) {
                a[k].resize(2 * nf);
                if (nf)
                    check_hip(hipMemcpyAsync(a[k].data(), srcs[k], sizeof(R) * 2 * nf, hipMemcpyDeviceToHost, stream_),
                              "debug copy");
            }
            synchronize();
            v4* o = static_cast<v4*>(out);
            for (size_t i = 0; i < 2 * nf; ++i) o[i] = v4{a[0][i], a[1][i], a[2][i], a[3][i]};
            return;
        }
        if (P_.flux32 && strcmp(what, "flux") == 0) {  // internal faces from their float store
            std::vector<float4> q(nf);
            if (nf)
                check_hip(hipMemcpyAsync(q.data(), P_.flux32, sizeof(float4) * nf, hipMemcpyDeviceToHost, stream_),
                          "debug copy");
            synchronize();
            double4* o = static_cast<double4*>(out);
            for (size_t f = 0; f < nf; ++f) o[f] = double4{q[f].x, q[f].y, q[f].z, q[f].w};
            if (nb)
                check_hip(hipMemcpyAsync(o + nf, P_.flux + nf, sizeof(double4) * nb, hipMemcpyDeviceToHost, stream_),
                          "debug copy");
            synchronize();
            return;
        }
        check_hip(hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, stream_), "debug copy");
        synchronize();
    }

    void run_stage_kernels(int upto, double dt, int stage) override {
        launch_reconstruct<R>(P_, L_.tile, lds_recon_, stream_, lanes_);
        if (upto > 0) launch_bvd<R>(P_, L_.tile, lds_bvd_, stream_);
        if (upto > 1) launch_flux<R>(P_, cfg_.R, cfg_.Cv, stream_);
        if (upto > 2) {
            launch_update<R>(P_, cfg_.R, gamma(), dt, nullptr, b1(stage), b2(stage), b3(stage), stage == 2, stream_);
            state_in_uold_ = stage == 2;
        }
        check_hip(hipGetLastError(), "stage kernels");
    }

   private:
    double gamma() const { return cfg_.R / cfg_.Cv + 1.0; }
    // SSP-RK3 Shu-Osher coefficients: the reference's truncated literals
    // (GPUevolve.cu:223-225) or the exact fractions.
    double b1(int s) const {
        static const double lit[3] = {1.0, 0.75, 0.333333}, ex[3] = {1.0, 0.75, 1.0 / 3.0};
        return cfg_.exactRK ? ex[s] : lit[s];
    }
    double b2(int s) const {
        static const double lit[3] = {0.0, 0.25, 0.666667}, ex[3] = {0.0, 0.25, 2.0 / 3.0};
        return cfg_.exactRK ? ex[s] : lit[s];
    }
    double b3(int s) const {
        static const double lit[3] = {1.0, 0.25, 0.666667}, ex[3] = {1.0, 0.25, 2.0 / 3.0};
        return cfg_.exactRK ? ex[s] : lit[s];
    }

    // One RK3 step.  dt_ptr == nullptr: the host's dt.  Otherwise the step's
    // time step is evaluated on the device by the first flux pass (from the
    // state the step starts from, exactly as the standalone CFL pass would)
    // and read by the updates from dt_ptr.
    // One RK3 step.  ev (optional): an event recorded after every launch, in
    // launch order (13 with a device time step).
    void enqueue_stages(double dt, const double* dt_ptr, hipEvent_t* ev = nullptr) {
        int j = 0;
        auto mark = [&] {
            if (ev) check_hip(hipEventRecordWithFlags(ev[j++], stream_, hipEventRecordExternal), "hipEventRecord");
        };
        for (int s = 0; s < 3; ++s) {
            launch_reconstruct<R>(P_, L_.tile, lds_recon_, stream_, lanes_);
            mark();
            launch_bvd<R>(P_, L_.tile, lds_bvd_, stream_);
            mark();
            if (s == 0 && dt_ptr) {
                launch_flux_dt<R>(P_, cfg_.R, cfg_.Cv, cfg_.CFL, stream_);
                mark();
                launch_dt_reduce<R>(P_, stream_);
                mark();
            } else {
                launch_flux<R>(P_, cfg_.R, cfg_.Cv, stream_);
                mark();
            }
            launch_update<R>(P_, cfg_.R, gamma(), dt, dt_ptr, b1(s), b2(s), b3(s), s == 2, stream_);
            mark();
        }
    }

    // One graph launch.  At most 2 * kGraphBlock launches are left queued: after
    // every kGraphBlock launches an event is recorded, and before a third
    // block starts the host waits for the first to finish (a long evolve_auto
    // otherwise queues tens of thousands of kernel dispatches at once).
    static constexpr int kGraphBlock = 256;
    void launch_graph() {
        check_hip(hipGraphLaunch(graph_exec_, stream_), "hipGraphLaunch");
        state_in_uold_ = true;
        if (++launches_ % kGraphBlock) return;
        hipEvent_t& e = evq_[blocks_ & 1];
        if (blocks_ >= 2) check_hip(hipEventSynchronize(e), "hipEventSynchronize");
        check_hip(hipEventRecord(e, stream_), "hipEventRecord");
        ++blocks_;
    }

    void capture() {
        synchronize();
        std::lock_guard<std::mutex> lk(runtime_mutex());
        check_hip(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        enqueue_stages(0.0, P_.dt);
        check_hip(hipStreamEndCapture(stream_, &graph_), "hipStreamEndCapture");
        check_hip(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
    }

    void dalloc(DevMem& m, size_t bytes) { m.alloc(bytes, stream_); }
    void dzero(void* p, size_t bytes) { check_hip(hipMemsetAsync(p, 0, bytes, stream_), "hipMemsetAsync"); }
    // n doubles uploaded in the kernels' precision: as they are in FP64,
    // rounded to float through a host copy in FP32-mixed
    R* put_as_R(DevMem& m, const double* src, size_t n) {
        if constexpr (sizeof(R) == sizeof(double)) {
            return reinterpret_cast<R*>(put(m, src, n));
        } else {
            std::vector<R> c(n);
            for (size_t i = 0; i < n; ++i) c[i] = (R)src[i];
            return put(m, c.data(), n);
        }
    }
    // allocation and upload in stream order; returns once src may be reused
    template <typename T> T* put(DevMem& m, const T* src, size_t n) {
        dalloc(m, sizeof(T) * n);
        if (n && src) check_hip(hipMemcpyAsync(m.p, src, sizeof(T) * n, hipMemcpyHostToDevice, stream_), "upload");
        synchronize();
        return m.as<T>();
    }

    void upload_state() {
        state_in_uold_ = false;  // U and Uold both receive the state
        const int nc = L_.nc, nb = L_.nb;
        std::vector<v4> w(nc + nb);
        for (int i = 0; i < nc + nb; ++i) {
            w[i].x = (R)ra_.FD[4 * i];
            w[i].y = (R)ra_.FD[4 * i + 1];
            w[i].z = (R)ra_.FD[4 * i + 2];
            w[i].w = (R)ra_.FD[4 * i + 3];
        }
        check_hip(hipMemcpyAsync(P_.W, w.data(), sizeof(v4) * (nc + nb), hipMemcpyHostToDevice, stream_), "W upload");
        check_hip(hipMemcpyAsync(P_.U, ra_.cFD.data(), sizeof(double4) * nc, hipMemcpyHostToDevice, stream_),
                  "U upload");
        check_hip(hipMemcpyAsync(P_.Uold, ra_.cFD.data(), sizeof(double4) * nc, hipMemcpyHostToDevice, stream_),
                  "Uold upload");
        synchronize();
    }

    void upload() {
        const size_t nc = L_.nc, nf = L_.nf, nb = L_.nb;
        const size_t nslot = (size_t)L_.nwaves * L_.maxC * 64;
        P_.nc = L_.nc;
        P_.nf = L_.nf;
        P_.nb = L_.nb;
        P_.maxS = L_.maxS;
        P_.maxC = L_.maxC;
        P_.ntiles = L_.ntiles;
        P_.lds_sten_off = (int)((size_t)(L_.tile + L_.maxHalo) * sizeof(v4));
        P_.lds_m_off = (int)lds_m_off_;
        P_.lds_bvd_sten_off = (int)lds_bvd_sten_off_;
        P_.boundaryFirstOrder = cfg_.boundaryFirstOrder;
        dalloc(m_W, sizeof(v4) * (nc + nb));
        P_.W = m_W.as<v4>();
        dalloc(m_U, sizeof(double4) * nc);
        P_.U = m_U.as<double4>();
        dalloc(m_Uold, sizeof(double4) * nc);
        P_.Uold = m_Uold.as<double4>();
        P_.vol = put(m_vol, L_.vol.data(), nc);
        for (DevMem* m : {&m_gp, &m_gu, &m_gv, &m_gT}) dalloc(*m, sizeof(v2) * nc);
        P_.gp = m_gp.as<v2>();
        P_.gu = m_gu.as<v2>();
        P_.gv = m_gv.as<v2>();
        P_.gT = m_gT.as<v2>();
        dalloc(m_lim, sizeof(v4) * nc);
        P_.lim = m_lim.as<v4>();
        dalloc(m_smooth, nc + nb);
        P_.smooth = m_smooth.as<int8_t>();
        dzero(P_.smooth, nc + nb);
        dalloc(m_ftype, nc);
        P_.ftype = m_ftype.as<int8_t>();
        dzero(P_.ftype, nc);
        dalloc(m_flat_prev, nc);
        P_.flat_prev = m_flat_prev.as<uint8_t>();
        dzero(P_.flat_prev, nc);
        dalloc(m_tlim, nc);
        P_.tlim = m_tlim.as<int8_t>();
        dzero(P_.tlim, nc);
        P_.scnt = put(m_scnt, L_.scnt.data(), nc);
        P_.ccnt = put(m_ccnt, L_.ccnt.data(), nc);
        P_.sten = put(m_sten, L_.sten.data(), L_.sten.size());
        P_.slot_face = put(m_slot_face, L_.slot_face.data(), nslot);
        // face offsets in the kernels' precision (they use them in it only)
        P_.slot_d = reinterpret_cast<const v2*>(put_as_R(m_slot_d, L_.slot_d.data(), 2 * nslot));
        P_.slot_e = reinterpret_cast<const double2*>(put(m_slot_e, L_.slot_e.data(), 2 * nslot));
        dalloc(m_slot_dr, sizeof(R) * nslot);
        P_.slot_dr = m_slot_dr.as<R>();
        dzero(P_.slot_dr, sizeof(R) * nslot);
        dalloc(m_lsq, sizeof(v2) * nslot);
        P_.lsq = m_lsq.as<v2>();
        P_.rbf = put_as_R(m_rbf, L_.rbf.data(), L_.rbf.size());
        P_.halo_ptr = put(m_halo_ptr, L_.halo_ptr.data(), L_.halo_ptr.size());
        {
            // one trailing id (cell 0) so that a tile without halo may read
            // its first id unconditionally
            std::vector<int32_t> halo(L_.halo.begin(), L_.halo.end());
            halo.push_back(0);
            P_.halo = put(m_halo, halo.data(), halo.size());
        }
        P_.fcell = reinterpret_cast<const int2*>(put(m_fcell, L_.fcell.data(), 2 * nf));
        P_.fgeo = reinterpret_cast<const double4*>(put(m_fgeo, L_.fgeo.data(), 8 * nf));
        // Cell and face centres (and periodic shifts): the FP64 four-lane
        // reconstruction forms its slot geometry from them, and the
        // face-centre flux passes their offsets and the CFL bound's distance
        P_.cen = P_.fcen = P_.fshift = nullptr;
        {
            P_.cen = reinterpret_cast<const double2*>(put(m_cen, ra_.CELL.data(), 2 * (nc + nb)));
            std::vector<double> fc(2 * (nf + nb));
            for (size_t i = 0; i < nf + nb; ++i) {
                fc[2 * i] = ra_.FACE[5 * i];
                fc[2 * i + 1] = ra_.FACE[5 * i + 1];
            }
            P_.fcen = reinterpret_cast<const double2*>(put(m_fcen, fc.data(), fc.size()));
            bool shifted = false;
            for (double v : ra_.FSHIFT) shifted = shifted || v != 0.0 || std::signbit(v);  // -0.0 counts
            if (shifted) P_.fshift = reinterpret_cast<const double2*>(put(m_fshift, ra_.FSHIFT.data(), 2 * nf));
        }
        // The face-centre flux passes that use the FP64 geometry (every FP64
        // pass, the FP32-mixed pass with the CFL bound) form their offsets
        // from the centres above and read Sf and magSf alone (40 instead of
        // 64 B per face)
        P_.fsf = nullptr;
        P_.fmag = nullptr;
        {
            std::vector<double> sf(2 * (size_t)nf), mag(nf);
            for (size_t i = 0; i < nf; ++i) {
                mag[i] = ra_.FACE[5 * i + 2];
                sf[2 * i] = ra_.FACE[5 * i + 3];
                sf[2 * i + 1] = ra_.FACE[5 * i + 4];
            }
            P_.fsf = reinterpret_cast<const double2*>(put(m_fsf, sf.data(), sf.size()));
            P_.fmag = put(m_fmag, mag.data(), mag.size());
        }
        P_.fgeo32 = nullptr;
        if (sizeof(R) == 4) {
            std::vector<float> g(L_.fgeo.size());
            for (size_t i = 0; i < g.size(); ++i) g[i] = (float)L_.fgeo[i];
            P_.fgeo32 = reinterpret_cast<const float4*>(put(m_fgeo32, g.data(), g.size()));
        }
        P_.fv = nullptr;
        P_.fvp = P_.fvu = P_.fvv = nullptr;
        if (sizeof(R) == 8) {  // one array per variable
            dalloc(m_fvp, sizeof(R) * 2 * nf);
            P_.fvp = m_fvp.as<R>();
            dalloc(m_fvu, sizeof(R) * 2 * nf);
            P_.fvu = m_fvu.as<R>();
            dalloc(m_fvv, sizeof(R) * 2 * nf);
            P_.fvv = m_fvv.as<R>();
        } else {  // records, the temperatures again in fvT
            dalloc(m_fv, sizeof(v4) * 2 * nf);
            P_.fv = m_fv.as<v4>();
        }
        dalloc(m_fvT, sizeof(R) * 2 * nf);
        P_.fvT = m_fvT.as<R>();
        dalloc(m_tTVD, sizeof(R) * 2 * nf);
        P_.tTVD = m_tTVD.as<R>();
        P_.ngp = L_.ngp;
        if (L_.ngp > 4) throw Error(3, "more than four flux quadrature points");
        for (int g = 0; g < L_.ngp; ++g) P_.gpw[g] = L_.gpw[g];
        if (L_.ngp) {
            P_.rbf_gp = put_as_R(m_rbf_gp, L_.rbf_gp.data(), L_.rbf_gp.size());
            P_.fgd = reinterpret_cast<const v4*>(put_as_R(m_fgd, L_.fgd.data(), L_.fgd.size()));
            dalloc(m_fvg, sizeof(v4) * 2 * nf * L_.ngp);
            P_.fvg = m_fvg.as<v4>();
        }
        dalloc(m_flux, sizeof(double4) * (nf + nb));
        P_.flux = m_flux.as<double4>();
        P_.flux32 = nullptr;
        if (sizeof(R) == 4 && L_.ngp == 0) {  // face-centre fluxes are floats: stored as such
            dalloc(m_flux32, sizeof(float4) * (nf ? nf : 1));
            P_.flux32 = m_flux32.as<float4>();
        }
        P_.bcell = put(m_bcell, L_.bcell.data(), nb);
        P_.blabel = put(m_blabel, L_.blabel.data(), nb);
        P_.bgeo = reinterpret_cast<const double4*>(put(m_bgeo, L_.bgeo.data(), 4 * nb));
        P_.bval = reinterpret_cast<const double4*>(put(m_bval, L_.bval.data(), 4 * nb));
        // per-instance time steps, the ticket counter (own 256-byte line) and
        // the CFL partials (two per flux tile, or kCflBlocks for the
        // standalone pass)
        const int ni = bi_.ninst();
        P_.ninst = ni;
        if (ni > 1) {
            std::vector<int32_t> ci(nc);
            for (int k = 0; k < ni; ++k)
                for (int c = bi_.cell_off[k]; c < bi_.cell_off[k + 1]; ++c) ci[c] = k;
            P_.cinst = put(m_cinst, ci.data(), nc);
        } else {
            P_.cinst = nullptr;
        }
        P_.finst_off = put(m_finst, bi_.face_off.data(), ni + 1);
        {
            std::vector<double> tc(4 * (size_t)ni, 0.0);
            for (int k = 0; k < ni; ++k) tc[4 * k + 1] = HUGE_VAL;
            P_.tctl = put(m_tctl, tc.data(), tc.size());
        }
        dalloc(m_res, sizeof(double) * (residual_blocks(L_.nc) + 1));
        P_.res_part = m_res.as<double>();
        const int dt_cap = std::max(kCflBlocks, 2 * flux_blocks(L_.nf, L_.nb));
        const size_t dt_bytes = (sizeof(double) * ni + 255) / 256 * 256;
        dalloc(m_dt, dt_bytes + 256 + sizeof(double) * dt_cap);
        P_.dt_cap = dt_cap;
        P_.dt = m_dt.as<double>();
        P_.dt_count = reinterpret_cast<unsigned int*>(m_dt.as<char>() + dt_bytes);
        P_.dt_part = reinterpret_cast<double*>(m_dt.as<char>() + dt_bytes + 256);
        {
            const std::vector<double> init_dt(ni, 1e5);
            unsigned int zero = 0;
            check_hip(hipMemcpyAsync(P_.dt, init_dt.data(), sizeof(double) * ni, hipMemcpyHostToDevice, stream_),
                      "dt init");
            check_hip(hipMemcpyAsync(P_.dt_count, &zero, 4, hipMemcpyHostToDevice, stream_), "dt init");
            synchronize();
        }
        upload_state();
        // K3: least-squares weights, once (GPUevolve.cu:131-140)
        launch_lsq<R>(P_, stream_);
        check_hip(hipGetLastError(), "k_lsq");
        synchronize();
        // the host keeps only what set_conserved needs (state, faces, boundary table)
        release_host(ra_.RBFbasis, ra_.RBFgp, ra_.stencil, ra_.cellFaces, ra_.neighbour, ra_.GPD);
        release_host(L_.rbf, L_.rbf_gp, L_.slot_d, L_.slot_e, L_.fgeo, L_.fgd);
        release_host(L_.sten, L_.slot_face, L_.fcell, L_.halo);
    }

    template <typename... V> static void release_host(V&... v) { ((void)std::decay_t<V>().swap(v), ...); }

    RefArrays ra_;
    BatchIndex bi_;
    DevLayout L_;
    DevPtrs<R> P_{};
    StreamHolder sh_;  // before the buffers: destroyed after them
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    hipEvent_t evq_[2] = {nullptr, nullptr};  // ends of the last two blocks of graph launches
    long long launches_ = 0, blocks_ = 0;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    // Where the conserved state is: a full step's last stage writes it to
    // Uold only (the next step starts from there); a first or second stage
    // writes U.  Host reads take it from state_U().
    bool state_in_uold_ = false;
    double4* state_U() const { return state_in_uold_ ? P_.Uold : P_.U; }
    size_t lds_recon_ = 0, lds_bvd_ = 0, lds_m_off_ = 0, lds_bvd_sten_off_ = 0;
    int lanes_ = 1;  // threads per cell of the reconstruction kernel
    DevMem m_W, m_U, m_Uold, m_vol, m_gp, m_gu, m_gv, m_gT, m_lim, m_smooth, m_ftype, m_scnt, m_ccnt, m_sten,
        m_slot_face, m_slot_d, m_slot_e, m_slot_dr, m_lsq, m_rbf, m_halo_ptr, m_halo, m_fcell, m_fgeo, m_fgeo32,
        m_fv, m_fvp, m_fvu, m_fvv, m_flux, m_flux32, m_bcell, m_blabel, m_bgeo, m_bval, m_dt, m_fvT, m_cinst, m_finst, m_tctl, m_res,
        m_rbf_gp, m_fgd, m_fvg, m_flat_prev, m_tlim, m_tTVD, m_cen, m_fcen, m_fshift, m_fsf, m_fmag;
};

// A solver on relabelled cells and faces (compact_tile_order) that reads and
// writes every cell- and face-indexed array in the caller's order.
class RelabeledSolver final : public DeviceSolver {
   public:
    RelabeledSolver(std::unique_ptr<DeviceSolver> in, Relabel r, int nc, int nf, int nb)
        : in_(std::move(in)), r_(std::move(r)), nc_(nc), nf_(nf), nb_(nb) {
        cfg_ = in_->config();
    }
    double compute_dt() override { return in_->compute_dt(); }
    void evolve(double dt) override { in_->evolve(dt); }
    void evolve_auto(int nsteps) override { in_->evolve_auto(nsteps); }
    void synchronize() override { in_->synchronize(); }
    bool idle() override { return in_->idle(); }
    void copy_out(double* FD, int8_t* shock) override {
        std::vector<double> f(FD ? 4 * (size_t)(nc_ + nb_) : 0);
        std::vector<int8_t> s(shock ? 2 * (size_t)(nc_ + nb_) : 0);
        in_->copy_out(FD ? f.data() : nullptr, shock ? s.data() : nullptr);
        for (int i = 0; i < nc_ + nb_; ++i) {
            const size_t j = i < nc_ ? (size_t)r_.cell_new[i] : (size_t)i;
            if (FD)
                for (int k = 0; k < 4; ++k) FD[4 * (size_t)i + k] = f[4 * j + k];
            if (shock) {
                shock[2 * (size_t)i] = s[2 * j];
                shock[2 * (size_t)i + 1] = s[2 * j + 1];
            }
        }
    }
    void get_conserved(double* U) override {
        std::vector<double> u(4 * (size_t)nc_);
        in_->get_conserved(u.data());
        for (int i = 0; i < nc_; ++i)
            for (int k = 0; k < 4; ++k) U[4 * (size_t)i + k] = u[4 * (size_t)r_.cell_new[i] + k];
    }
    void set_conserved(const double* U) override {
        std::vector<double> u(4 * (size_t)nc_);
        for (int i = 0; i < nc_; ++i)
            for (int k = 0; k < 4; ++k) u[4 * (size_t)r_.cell_new[i] + k] = U[4 * (size_t)i + k];
        in_->set_conserved(u.data());
    }
    double last_dt() override { return in_->last_dt(); }
    int ninst() const override { return in_->ninst(); }
    void last_dts(double* out) override { in_->last_dts(out); }
    void get_conserved_range(int c0, int n, double* U) override {
        int lo, hi;
        span(c0, n, lo, hi);
        std::vector<double> u(4 * (size_t)(hi - lo));
        if (hi > lo) in_->get_conserved_range(lo, hi - lo, u.data());
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 4; ++k) U[4 * (size_t)i + k] = u[4 * (size_t)(r_.cell_new[c0 + i] - lo) + k];
    }
    void get_primitive_range(int c0, int n, double* W, int8_t* type) override {
        int lo, hi;
        span(c0, n, lo, hi);
        std::vector<double> w(W ? 4 * (size_t)(hi - lo) : 0);
        std::vector<int8_t> t(type ? (size_t)(hi - lo) : 0);
        if (hi > lo) in_->get_primitive_range(lo, hi - lo, W ? w.data() : nullptr, type ? t.data() : nullptr);
        for (int i = 0; i < n; ++i) {
            const size_t j = (size_t)(r_.cell_new[c0 + i] - lo);
            if (W)
                for (int k = 0; k < 4; ++k) W[4 * (size_t)i + k] = w[4 * j + k];
            if (type) type[i] = t[j];
        }
    }
    void reset_state() override { in_->reset_state(); }
    void set_time(int inst, double t, double t_end) override { in_->set_time(inst, t, t_end); }
    void get_time(double* t, double* steps) override { in_->get_time(t, steps); }
    double residual_norm() override { return in_->residual_norm(); }
    double time_steps(int nsteps, bool graph) override { return in_->time_steps(nsteps, graph); }
    double time_steps_kernels(int nsteps, double* ms_kind, int* cnt_kind) override {
        return in_->time_steps_kernels(nsteps, ms_kind, cnt_kind);
    }
    void profile(double* ms, int* count, int reps) override { in_->profile(ms, count, reps); }
    const DevLayout& layout() const override { return in_->layout(); }
    // Cell arrays (W, U, gradients, lim, smooth, ftype) and face arrays (fv,
    // flux) come back in the caller's order; the slot arrays (lsq) stay in
    // the device's.
    void debug_fetch(const char* what, void* out, size_t bytes) override {
        std::vector<unsigned char> b(bytes);
        in_->debug_fetch(what, b.data(), bytes);
        unsigned char* o = static_cast<unsigned char*>(out);
        const std::string w(what);
        auto cells = [&](int n) {  // n entries: nc cells (then pseudo-cells)
            const size_t e = bytes / n;
            for (int i = 0; i < n; ++i) {
                const size_t j = i < nc_ ? (size_t)r_.cell_new[i] : (size_t)i;
                memcpy(o + e * i, b.data() + e * j, e);
            }
        };
        auto faces = [&](int n, int per) {  // n entries of `per` records: nf faces (then boundary faces)
            const size_t e = bytes / ((size_t)n * per) * per;
            for (int i = 0; i < n; ++i) {
                const size_t j = i < nf_ ? (size_t)r_.face_new[i] : (size_t)i;
                memcpy(o + e * i, b.data() + e * j, e);
            }
        };
        if (w == "W" || w == "smooth")
            cells(nc_ + nb_);
        else if (w == "U" || w == "gp" || w == "gu" || w == "gv" || w == "gT" || w == "lim" || w == "ftype")
            cells(nc_);
        else if (w == "fv")
            faces(nf_, 2);
        else if (w == "flux")
            faces(nf_ + nb_, 1);
        else if (w == "fvg" && nf_ > 0) {  // ngp blocks of nf faces, two states each
            const size_t rec = 2 * 4 * (cfg_.precision == 1 ? sizeof(float) : sizeof(double));
            const size_t blocks = bytes / ((size_t)nf_ * rec);
            for (size_t g = 0; g < blocks; ++g)
                for (int i = 0; i < nf_; ++i)
                    memcpy(o + rec * (g * nf_ + i), b.data() + rec * (g * nf_ + r_.face_new[i]), rec);
        } else
            memcpy(out, b.data(), bytes);
    }
    void run_stage_kernels(int upto, double dt, int stage) override { in_->run_stage_kernels(upto, dt, stage); }
    void set_constants(double R, double Cv, double CFL) override {
        in_->set_constants(R, Cv, CFL);
        cfg_ = in_->config();
    }

   private:
    // the device range holding mesh cells [c0, c0+n)
    void span(int c0, int n, int& lo, int& hi) const {
        if (c0 < 0 || n < 0 || c0 + n > nc_) throw Error(1, "cell range out of bounds");
        lo = nc_;
        hi = 0;
        for (int i = c0; i < c0 + n; ++i) {
            lo = std::min(lo, r_.cell_new[i]);
            hi = std::max(hi, r_.cell_new[i] + 1);
        }
        if (hi < lo) lo = hi = 0;
    }
    std::unique_ptr<DeviceSolver> in_;
    Relabel r_;
    int nc_, nf_, nb_;
};

std::unique_ptr<DeviceSolver> make_impl(RefArrays&& ra, DevLayout&& L, const SolverConfig& cfg, const BatchIndex* bi) {
    if (cfg.precision == 1)
        return std::unique_ptr<DeviceSolver>(new SolverImpl<float>(std::move(ra), std::move(L), cfg, bi));
    return std::unique_ptr<DeviceSolver>(new SolverImpl<double>(std::move(ra), std::move(L), cfg, bi));
}

}  // namespace

std::unique_ptr<DeviceSolver> make_device_solver(const RefArrays& ra, const SolverConfig& cfg,
                                                 const BatchIndex* bi) {
    if (ra.nc <= 0) throw Error(1, "empty mesh");
    // host-side work (copies, relabelling, the device layout) runs before the
    // library's lock is taken, so that solvers prepared on several threads
    // overlap in it; the lock covers the device allocation and upload
    if (cfg.cellOrder == 1) {
        RefArrays copy = ra;
        DevLayout L = make_dev_layout(copy, cfg.tile);
        std::lock_guard<std::mutex> lk(runtime_mutex());
        return make_impl(std::move(copy), std::move(L), cfg, bi);
    }
    BatchIndex one;
    if (!bi) {
        one.cell_off = {0, ra.nc};
        one.face_off = {0, ra.nf};
        one.bface_off = {0, ra.nb};
    }
    // grouped blocks: 32 cells in FP64, 64 in FP32-mixed (fastest per step on
    // the 1.5M-cell mixed triangle/quad mesh; 16, 32 and 64 were measured)
    const int group = cfg.cellOrder == 2 ? 0 : cfg.precision == 0 ? 32 : 64;
    Relabel r = compact_tile_order(ra, bi ? *bi : one, group);
    std::unique_ptr<DeviceSolver> in;
    {
        RefArrays rl = relabel(ra, r);
        DevLayout L = make_dev_layout(rl, cfg.tile);
        std::lock_guard<std::mutex> lk(runtime_mutex());
        in = make_impl(std::move(rl), std::move(L), cfg, bi);
    }
    return std::unique_ptr<DeviceSolver>(new RelabeledSolver(std::move(in), std::move(r), ra.nc, ra.nf, ra.nb));
}

}  // namespace afar
