// Time step of seam 2 (adjustTimeStep) that allocates nothing, over device
// arrays the caller owns, for callers that cannot allocate: code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc.
//
// The per-block partial minima live in a file-scope __device__ array; each
// device holds its own copy, whose address is looked up once per device with
// hipGetSymbolAddress and kept.  Two launches: every block writes its
// partial (so nothing has to start at zero), then one block reduces them and
// writes the time step.  Calls that overlap on one device race on the
// partials.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "../../include/afar.h"
#include "afar_host.h"
#include "afar_math.h"
#include "kernels.h"
#include "solver.h"

namespace afar {

constexpr int kNoAllocParts = 1024;  // partial minima: at most this many blocks (8 KiB per device)
static __device__ __align__(256) double scratch[kNoAllocParts];

namespace {

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// The device address of `scratch` on the current device (looked up once per device).
double* scratch_ptr() {
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw Error(AFAR_ERR_LIMIT, "device index beyond 64");
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

// CFL bound of every internal face (cudaFieldOperation.h:124-180) on the
// reference's layout: FD (p, Ux, Uy, T) and CELL (x, y) per cell and
// pseudo-cell, FACE (x, y, magSf, Sfx, Sfy), neighbour (own, nei); FSHIFT
// (optional) the periodic shift of the neighbour centre.  Block b reduces
// faces b, b + grid, ... to parts[b].
__global__ __launch_bounds__(256) void k_cfl_parts(const double* FD, const double* CELL, const double* FACE,
                                                   const int32_t* neighbour, const double* FSHIFT, int nf,
                                                   double R, double Cv, double CFL, double* parts) {
    double m = 1e5;
    for (int f = blockIdx.x * 256 + threadIdx.x; f < nf; f += gridDim.x * 256) {
        const int own = neighbour[2 * f], nei = neighbour[2 * f + 1];
        const double xo = CELL[2 * own], yo = CELL[2 * own + 1];
        const double xn = CELL[2 * nei] + (FSHIFT ? FSHIFT[2 * f] : 0.0);
        const double yn = CELL[2 * nei + 1] + (FSHIFT ? FSHIFT[2 * f + 1] : 0.0);
        const double* F = FACE + 5 * (size_t)f;
        const double deltaR = afar_hypot(xn - xo, yn - yo);
        const double fraction = afar_hypot(F[0] - xo, F[1] - yo) / deltaR;
        const double magSf = F[2], Sfx = F[3], Sfy = F[4];
        const double* wo = FD + 4 * (size_t)own;
        const double* wn = FD + 4 * (size_t)nei;
        const double d = CFL * deltaR * magSf /
                         ((1.0 - fraction) * fabs(wo[1] * Sfx + wo[2] * Sfy) +
                          fraction * fabs(wn[1] * Sfx + wn[2] * Sfy) +
                          sqrt((R / Cv + 1.0) * R * ((1.0 - fraction) * wo[3] + fraction * wn[3])) * magSf);
        m = d < m ? d : m;  // a NaN bound never replaces the minimum
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off);
        m = o < m ? o : m;
    }
    __shared__ double wmin[4];
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = wmin[0];
        for (int w = 1; w < 4; ++w) v = wmin[w] < v ? wmin[w] : v;
        parts[blockIdx.x] = v;
    }
}

__global__ __launch_bounds__(256) void k_cfl_min(const double* parts, int nparts, double* out) {
    double m = 1e5;
    for (int i = threadIdx.x; i < nparts; i += 256) m = parts[i] < m ? parts[i] : m;
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off);
        m = o < m ? o : m;
    }
    __shared__ double wmin[4];
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = wmin[0];
        for (int w = 1; w < 4; ++w) v = wmin[w] < v ? wmin[w] : v;
        *out = v;
    }
}

}  // namespace

void adjust_time_step_noalloc(const double* FD, const double* CELL, const double* FACE, const int32_t* neighbour,
                              const double* FSHIFT, int nf, double R, double Cv, double CFL, double* deltaT,
                              hipStream_t s) {
    double* parts = scratch_ptr();  // never the array's name: that is a host-side stand-in
    const int nblk = std::max(1, std::min(kNoAllocParts, (nf + 255) / 256));
    hipLaunchKernelGGL(k_cfl_parts, dim3(nblk), dim3(256), 0, s, FD, CELL, FACE, neighbour, FSHIFT, nf, R, Cv, CFL,
                       parts);
    hipLaunchKernelGGL(k_cfl_min, dim3(1), dim3(256), 0, s, (const double*)parts, nblk, deltaT);
    check_hip(hipGetLastError(), "time-step kernels");
}

}  // namespace afar
