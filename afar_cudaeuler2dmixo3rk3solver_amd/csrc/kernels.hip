// gfx950 kernels of the AFAR SSP-RK3 stage pipeline.
//
// Per RK stage: reconstruct (cell) -> bvd (cell) -> flux (face) -> update
// (cell: residual gather + RK update + boundary ghosts); one CFL pass per
// step.  The kernels are templated on the reconstruction/flux type R: R =
// double is the all-FP64 transcription of the reference kernels, R = float is
// the mixed-precision path (state, residual and RK update stay FP64).
//
// Reference kernels: src/CUDA/cudaEuler2DmixO3RK3/cudaLimiter.h:35-464,
// cudaRiemannSolver.h:7-271, cudaFieldOperation.h:51-268,
// src/CUDA/BCs/2D/cudaEulerBCs.h:42-73, launch order GPUevolve.cu:207-306.
//
// Arithmetic is written operation-for-operation in the reference's order and
// the file is built with -ffp-contract=off, so the FP64 path reproduces the
// host oracle bit for bit.
#include <hip/hip_runtime.h>

#include "afar_math.h"
#include "kernels.h"

namespace afar {

// Thresholds of the reference (cudaLimiter.h:4-7), in the working type.
template <typename R> struct K {
    static constexpr R SMALL = (R)1e-5;
    static constexpr R VSMALL = (R)1e-10;
    static constexpr R BVD_FACTOR = (R)0.5;
    static constexpr R VALID_LIM_FACTOR = (R)0.9;
};

__device__ __forceinline__ size_t sell(int c, int k, int Kw) {
    return ((size_t)(c >> 6) * Kw + k) * 64 + (c & 63);
}

// Block -> tile mapping.  Blocks are dispatched round-robin over the 8 XCDs,
// each with its own L2.  Each XCD takes chunks of XCD_CHUNK consecutive tiles
// (cells in Hilbert order, so a chunk is a compact region whose halos are
// mostly read through the same L2), and the 8 chunks running at a time are
// neighbours in memory.  Measured on the 1.5M-cell bench: chunks of 16-64
// tiles are 8% faster per step than one contiguous eighth of the grid per
// XCD, and 4% faster than plain round-robin.  XCD_CHUNK is a build-time knob
// (1: round-robin).
#ifndef AFAR_XCD_CHUNK
#define AFAR_XCD_CHUNK 32
#endif
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
    constexpr int G = AFAR_XCD_CHUNK;
    const int span = 8 * G;
    if (b >= nblocks - nblocks % span) return b;  // the last partial span in dispatch order
    return (b / span) * span + (b & 7) * G + (b >> 3) % G;
}

// exp and erfc of KFVS: the library's own functions (afar_math.h) in the
// working precision, which depend on no vendor math library, so the host
// restatements reproduce them bit for bit.
template <typename R> __device__ __forceinline__ R kexp(R x);
template <> __device__ __forceinline__ double kexp<double>(double x) { return afar_exp(x); }
template <> __device__ __forceinline__ float kexp<float>(float x) { return afar_expf(x); }
template <typename R> __device__ __forceinline__ R kerfc(R x);
template <> __device__ __forceinline__ double kerfc<double>(double x) { return afar_erfc(x); }
template <> __device__ __forceinline__ float kerfc<float>(float x) { return afar_erfcf(x); }

// ---------------------------------------------------------------- K3 least squares

template <typename R>
__global__ void k_lsq(DevPtrs<R> p) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= p.nc) return;
    const int C = p.ccnt[c];
    double Sxx = 0.0, Sxy = 0.0, Syy = 0.0;
    for (int i = 0; i < C; ++i) {
        const double2 e = p.slot_e[sell(c, i, p.maxC)];
        Sxx += e.x * e.x;
        Sxy += e.x * e.y;
        Syy += e.y * e.y;
    }
    for (int i = 0; i < C; ++i) {
        const double2 e = p.slot_e[sell(c, i, p.maxC)];
        typename VT<R>::v2 m;
        m.x = (R)((Syy * e.x - Sxy * e.y) / (Sxx * Syy - Sxy * Sxy));
        m.y = (R)((-Sxy * e.x + Sxx * e.y) / (Sxx * Syy - Sxy * Sxy));
        p.lsq[sell(c, i, p.maxC)] = m;
        // the centre distance K5 divides by (cudaLimiter.h:280), once here
        p.slot_dr[sell(c, i, p.maxC)] = (R)afar_hypot(e.x, e.y);
    }
}

// ---------------------------------------------------------------- K4 time step

__device__ __forceinline__ double block_min_256(double v, double* wmin) {
    // wavefront min through cross-lane swizzles, then across the 4 waves
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = wmin[0];
    for (int w = 1; w < 4; ++w) m = wmin[w] < m ? wmin[w] : m;
    __syncthreads();
    return m;
}

// Time-step bound of one internal face (cudaFieldOperation.h:124-180).
template <typename V4>
__device__ __forceinline__ double cfl_face_dt(const double4& g0, const double4& g1, const V4& wo, const V4& wn,
                                              double R_, double Cv, double CFL) {
    const double Sfx = g0.x, Sfy = g0.y, magSf = g0.z, deltaR = g0.w;
    const double fraction = afar_hypot(g1.x, g1.y) / deltaR;
    const double uxo = wo.y, uyo = wo.z, To = wo.w, uxn = wn.y, uyn = wn.z, Tn = wn.w;
    return CFL * deltaR * magSf /
           ((1.0 - fraction) * fabs(uxo * Sfx + uyo * Sfy) + fraction * fabs(uxn * Sfx + uyn * Sfy) +
            sqrt((R_ / Cv + 1.0) * R_ * ((1.0 - fraction) * To + fraction * Tn)) * magSf);
}

// Each block reduces a contiguous chunk of faces (stride loop) to one
// partial minimum; the last block to finish reduces the partials and stores
// the time step on the device.  At most kCflBlocks partials, so only that
// many same-address atomics per step.
template <typename R>
__global__ __launch_bounds__(256) void k_cfl(DevPtrs<R> p, double R_, double Cv, double CFL, int chunk) {
    const int blk = xcd_tile(blockIdx.x, gridDim.x);
    const int f0 = blk * chunk;
    const int f1 = min(p.nf, f0 + chunk);
    double dtf = 1e5;
    for (int f = f0 + threadIdx.x; f < f1; f += 256) {
        const int2 oc = p.fcell[f];
        const double d = cfl_face_dt(p.fgeo[2 * f], p.fgeo[2 * f + 1], p.W[oc.x], p.W[oc.y], R_, Cv, CFL);
        dtf = d < dtf ? d : dtf;  // a NaN face value never replaces the minimum
    }
    __shared__ double wmin[4];
    __shared__ int last;
    const double m = block_min_256(dtf, wmin);
    if (threadIdx.x == 0) {
        __hip_atomic_store(&p.dt_part[blockIdx.x], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const unsigned t = atomicAdd(p.dt_count, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double v = 1e5;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += 256) {
        const double o = __hip_atomic_load(&p.dt_part[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v = o < v ? o : v;
    }
    v = block_min_256(v, wmin);
    if (threadIdx.x == 0) {
        *p.dt = v;
        atomicExch(p.dt_count, 0u);
    }
}

// ---------------------------------------------------------------- K5 reconstruct

// Copies the stencil ids (sliced ELL, width 64) of the tile's cells into LDS:
// the tile's waves are consecutive slices, so the block is one contiguous
// range of (waves of the tile) * maxS * 64 ids, copied 16 bytes at a time.
__device__ __forceinline__ void stage_stencil(uint16_t* sSt, const uint16_t* sten, int tile, int tileCells, int nc,
                                              int maxS, int t, int nthreads) {
    const int wpt = tileCells >> 6;  // waves per tile
    const int w0 = tile * wpt;
    const int nw = min(wpt, ((nc + 63) >> 6) - w0);
    const int n16 = nw * maxS * 8;  // 16-byte words (64 ids of 2 bytes = 8 words per slice row)
    const uint4* src = reinterpret_cast<const uint4*>(sten + (size_t)w0 * maxS * 64);
    uint4* dst = reinterpret_cast<uint4*>(sSt);
    for (int i = t; i < n16; i += nthreads) dst[i] = src[i];
}

// The reconstructed state of face side i (= 2f + side): FP64 one array per
// variable, FP32 a record plus the temperature again (the layouts of DevPtrs).
template <typename R>
__device__ __forceinline__ void store_face_state(const DevPtrs<R>& p, int i, const typename VT<R>::v4& o) {
    if constexpr (sizeof(R) == 8) {
        p.fvp[i] = o.x;
        p.fvu[i] = o.y;
        p.fvv[i] = o.z;
    } else {
        p.fv[i] = o;
    }
    p.fvT[i] = o.w;
}

template <typename R>
__device__ __forceinline__ void limiter_process(R& lim, R& phif, R phic, R phiMax, R phiMin, R gx, R gy, R dx, R dy) {
    // cudaLimiter.h:35-50
    phif = phic + (gx * dx + gy * dy);
    const R d2 = phif - phic;
    const R d1 = (R)(phif < phic) * (phiMin - phic) + (R)(phif >= phic) * (phiMax - phic);
    R yy = fmin((R)2, ((R)0.5 * d1) / ((d2 + (R)(d2 >= (R)0) * K<R>::VSMALL) - (R)(d2 < (R)0) * K<R>::VSMALL));
    const int valid = fabs(d2) >= K<R>::SMALL;
    yy = (R)valid * yy + (R)(1 - valid) * (R)1;
    lim = fmin(lim, yy);
}

// Limiter transforms of K6 (cudaLimiter.h:9-32).
template <typename R> __device__ __forceinline__ R to_bj(R y) { return fmin((R)2 * y, (R)1); }
template <typename R> __device__ __forceinline__ R to_venkat(R y) {
    return fmin((y * y + y + K<R>::VSMALL) / (y * y + (R)0.5 + (R)0.5 * y + K<R>::VSMALL), (R)1);
}

// GP: the flux-quadrature extension's point states (a separate instantiation,
// so the reference path carries none of its registers)
template <typename R, int TILE, bool GP>
__global__ __launch_bounds__(TILE) void k_reconstruct(DevPtrs<R> p) {
    using v4 = typename VT<R>::v4;
    using v2 = typename VT<R>::v2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    v4* sW = reinterpret_cast<v4*>(smem_raw);
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= p.ntiles) return;
    const int tid = threadIdx.x, c = tile * TILE + tid;
    const int hb = p.halo_ptr[tile], hn = p.halo_ptr[tile + 1] - hb;
    uint16_t* sSt = reinterpret_cast<uint16_t*>(smem_raw + p.lds_sten_off);
    if (c < p.nc) sW[tid] = p.W[c];
    for (int i = tid; i < hn; i += TILE) sW[TILE + i] = p.W[p.halo[hb + i]];
    stage_stencil(sSt, p.sten, tile, TILE, p.nc, p.maxS, tid, TILE);
    __syncthreads();
    if (c >= p.nc) return;
    const int S = p.scnt[c], C = p.ccnt[c];
    const int maxS = p.maxS, maxC = p.maxC, KR = maxC * (maxS + 1);
    const v4 wc = sW[tid];

    R gpx = 0, gpy = 0, guxx = 0, guyx = 0, guxy = 0, guyy = 0, gTx = 0, gTy = 0;
    for (int i = 0; i < C; ++i) {
        const v2 M = p.lsq[sell(c, i, maxC)];
        const v4 wj = sW[sSt[sell(tid, i, maxS)]];
        R dphi = wj.x - wc.x;
        gpx += M.x * dphi;
        gpy += M.y * dphi;
        dphi = wj.y - wc.y;
        guxx += M.x * dphi;
        guyx += M.y * dphi;
        dphi = wj.z - wc.z;
        guxy += M.x * dphi;
        guyy += M.y * dphi;
        dphi = wj.w - wc.w;
        gTx += M.x * dphi;
        gTy += M.y * dphi;
    }
    p.gp[c] = v2{gpx, gpy};
    p.gu[c] = v2{guxx, guyx};
    p.gv[c] = v2{guxy, guyy};
    p.gT[c] = v2{gTx, gTy};

    R pMax = wc.x, pMin = wc.x, uxMax = wc.y, uxMin = wc.y, uyMax = wc.z, uyMin = wc.z, TMax = wc.w, TMin = wc.w;
    for (int i = 0; i < C; ++i) {
        const v4 wj = sW[sSt[sell(tid, i, maxS)]];
        pMax = fmax(wj.x, pMax);
        pMin = fmin(wj.x, pMin);
        uxMax = fmax(wj.y, uxMax);
        uxMin = fmin(wj.y, uxMin);
        uyMax = fmax(wj.z, uyMax);
        uyMin = fmin(wj.z, uyMin);
        TMax = fmax(wj.w, TMax);
        TMin = fmin(wj.w, TMin);
    }

    const int nf = p.nf;
    if (fabs(TMax - TMin) < K<R>::VSMALL) {
        p.lim[c] = v4{(R)1, (R)1, (R)1, (R)1};
        const R lT = to_bj((R)1);  // the BVD test's limiter of this cell
        p.tlim[c] = lT < K<R>::VALID_LIM_FACTOR;
        for (int fi = 0; fi < C; ++fi) {
            const int sf = p.slot_face[sell(c, fi, maxC)];
            const int isNei = sf < 0;
            const int f = (sf < 0 ? -sf : sf) - 1;
            if (f >= nf) continue;
            const v2 d = p.slot_d[sell(c, fi, maxC)];
            const R dx = d.x, dy = d.y;
            v4 o;
            o.x = wc.x + gpx * dx + gpy * dy;
            o.y = wc.y + guxx * dx + guyx * dy;
            o.z = wc.z + guxy * dx + guyy * dy;
            o.w = wc.w + gTx * dx + gTy * dy;
            store_face_state(p, 2 * f + isNei, o);
            p.tTVD[2 * f + isNei] = wc.w + lT * (gTx * dx + gTy * dy);
            if constexpr (GP)
                for (int g = 0; g < p.ngp; ++g) {  // flux points (extension): the same linear state there
                const v4 dg = p.fgd[(size_t)g * nf + f];
                const R gx = (R)(isNei ? dg.z : dg.x), gy = (R)(isNei ? dg.w : dg.y);
                v4 og;
                og.x = wc.x + gpx * gx + gpy * gy;
                og.y = wc.y + guxx * gx + guyx * gy;
                og.z = wc.z + guxy * gx + guyy * gy;
                og.w = wc.w + gTx * gx + gTy * gy;
                p.fvg[((size_t)g * nf + f) * 2 + isNei] = og;
            }
        }
        p.smooth[c] = 0;
        return;
    }

    R limP = 2, limUx = 2, limUy = 2, limT = 2;
    R smoothness = 0;
    int isNotTrouble = 1;
    const R rP = fabs(pMax - pMin), rUx = fabs(uxMax - uxMin), rUy = fabs(uyMax - uyMin), rT = fabs(TMax - TMin);
    for (int fi = 0; fi < C; ++fi) {
        const int sf = p.slot_face[sell(c, fi, maxC)];
        const int isNei = sf < 0;
        const int f = (sf < 0 ? -sf : sf) - 1;
        if (f >= nf) continue;
        const v2 d = p.slot_d[sell(c, fi, maxC)];
        const R dx = d.x, dy = d.y;
        const R deltaR = p.slot_dr[sell(c, fi, maxC)];
        const R fraction = sqrt(dx * dx + dy * dy) / deltaR;
        const int nbl = sSt[sell(tid, fi, maxS)];

        const R* a = p.rbf + sell(c, fi * (maxS + 1), KR);  // row stride 64
        R coef = a[0];
        R pf = coef * wc.x, uxf = coef * wc.y, uyf = coef * wc.z, Tf = coef * wc.w;
        for (int ci = 0; ci < S; ++ci) {
            const v4 wj = sW[sSt[sell(tid, ci, maxS)]];
            coef = a[(size_t)(ci + 1) * 64];
            pf += coef * wj.x;
            uxf += coef * wj.y;
            uyf += coef * wj.z;
            Tf += coef * wj.w;
        }
        v4 o;
        R phif;
        limiter_process(limP, phif, wc.x, pMax, pMin, gpx, gpy, dx, dy);
        o.x = (R)(rP >= K<R>::SMALL) * pf + (R)(rP < K<R>::SMALL) * phif;
        isNotTrouble = phif > (R)0;
        smoothness = fmax(smoothness, fabs(((R)1 - fraction) * wc.x + fraction * sW[nbl].x - phif) / deltaR +
                                          (R)(pf <= (R)0) * (R)1e20 + (R)(phif <= (R)0) * (R)1e20);
        limiter_process(limUx, phif, wc.y, uxMax, uxMin, guxx, guyx, dx, dy);
        o.y = (R)(rUx >= K<R>::SMALL) * uxf + (R)(rUx < K<R>::SMALL) * phif;
        limiter_process(limUy, phif, wc.z, uyMax, uyMin, guxy, guyy, dx, dy);
        o.z = (R)(rUy >= K<R>::SMALL) * uyf + (R)(rUy < K<R>::SMALL) * phif;
        limiter_process(limT, phif, wc.w, TMax, TMin, gTx, gTy, dx, dy);
        o.w = (R)(rT >= K<R>::SMALL) * Tf + (R)(rT < K<R>::SMALL) * phif;
        isNotTrouble = phif > (R)0;
        store_face_state(p, 2 * f + isNei, o);
        if constexpr (GP)
        for (int g = 0; g < p.ngp; ++g) {  // flux points (extension), as in k_reconstruct_q
            const R* ag = p.rbf_gp + sell(c, (g * maxC + fi) * (maxS + 1), p.ngp * KR);
            R cg = ag[0];
            R pg = cg * wc.x, uxg = cg * wc.y, uyg = cg * wc.z, Tg = cg * wc.w;
            for (int ci = 0; ci < S; ++ci) {
                const v4 wj = sW[sSt[sell(tid, ci, maxS)]];
                cg = ag[(size_t)(ci + 1) * 64];
                pg += cg * wj.x;
                uxg += cg * wj.y;
                uyg += cg * wj.z;
                Tg += cg * wj.w;
            }
            const v4 dg = p.fgd[(size_t)g * nf + f];
            const R gx = (R)(isNei ? dg.z : dg.x), gy = (R)(isNei ? dg.w : dg.y);
            v4 og;
            og.x = (R)(rP >= K<R>::SMALL) * pg + (R)(rP < K<R>::SMALL) * (wc.x + (gpx * gx + gpy * gy));
            og.y = (R)(rUx >= K<R>::SMALL) * uxg + (R)(rUx < K<R>::SMALL) * (wc.y + (guxx * gx + guyx * gy));
            og.z = (R)(rUy >= K<R>::SMALL) * uyg + (R)(rUy < K<R>::SMALL) * (wc.z + (guxy * gx + guyy * gy));
            og.w = (R)(rT >= K<R>::SMALL) * Tg + (R)(rT < K<R>::SMALL) * (wc.w + (gTx * gx + gTy * gy));
            p.fvg[((size_t)g * nf + f) * 2 + isNei] = og;
        }
    }
    const R nt = (R)isNotTrouble;
    const R limTraw = fmax(limT, (R)0) * nt;
    p.lim[c] = v4{fmax(limP, (R)0) * nt, fmax(limUx, (R)0) * nt, fmax(limUy, (R)0) * nt, limTraw};
    smoothness = smoothness / ((R)0.5 * (pMax + pMin) * nt + K<R>::VSMALL);
    p.smooth[c] = smoothness > (R)1;
    // the BVD test's inputs of this cell (cudaLimiter.h:431-440): its limiter
    // condition and the limited linear temperature on each internal face
    const R lT = to_bj(limTraw);
    p.tlim[c] = lT < K<R>::VALID_LIM_FACTOR;
    for (int fi = 0; fi < C; ++fi) {
        const int sf = p.slot_face[sell(c, fi, maxC)];
        const int f = (sf < 0 ? -sf : sf) - 1;
        if (f >= nf) continue;
        const v2 d = p.slot_d[sell(c, fi, maxC)];
        p.tTVD[2 * f + (sf < 0)] = wc.w + lT * (gTx * d.x + gTy * d.y);
    }
}

// Moves within groups of four consecutive lanes (DPP quad permutations, full
// rate VALU); every lane of the group must be active.
template <int CTRL> __device__ __forceinline__ int dpp_move(int x) {
    return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ float dpp_move(float x) {
    return __int_as_float(dpp_move<CTRL>(__float_as_int(x)));
}
template <int CTRL> __device__ __forceinline__ double dpp_move(double x) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)dpp_move<CTRL>((int)(unsigned)u);
    const unsigned hi = (unsigned)dpp_move<CTRL>((int)(unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// the value of lane J of this lane's group
template <int J, typename T> __device__ __forceinline__ T quad_bcast(T x) { return dpp_move<J * 0x55>(x); }
// the value of lane (this lane xor X) of the group, X = 1 or 2
template <int X, typename T> __device__ __forceinline__ T quad_xor(T x) {
    return dpp_move<X == 1 ? 0xB1 : 0x4E>(x);  // quad_perm [1,0,3,2] and [2,3,0,1]
}

// K5 with four lanes per cell, for meshes whose cells have at most four
// faces (triangles and quads): lane k of a cell's group handles face slot k
// (its RBF sum and limiter candidates) and variable k of the gradient and the
// compact min/max, which the group then shares.  The per-cell results are
// exact reductions over the group (the limiter is a minimum, the smoothness a
// maximum, the trouble test that of the last internal face), so the values
// equal the one-lane-per-cell kernel's bit for bit.
// Waves per SIMD the register allocator must allow for the four-lane kernel,
// per working type (build-time tuning knobs; 1 = no constraint).  5 (FP64, 96
// VGPRs) and 8 (FP32, 64 VGPRs) fit without spills and were measured faster
// than the unconstrained 4 and 7 waves.
#ifndef AFAR_RECONQ_WPE_D
#define AFAR_RECONQ_WPE_D 5
#endif
#ifndef AFAR_RECONQ_WPE_F
#define AFAR_RECONQ_WPE_F 8
#endif
// The flux-quadrature instantiation holds more values and asks for one wave
// fewer, so that it does not spill.
template <typename R, bool GP> struct ReconQWaves {
    static constexpr int value = (sizeof(R) == 8 ? AFAR_RECONQ_WPE_D : AFAR_RECONQ_WPE_F) - (GP ? 1 : 0);
};
#define AFAR_RECONQ_ATTR __attribute__((amdgpu_waves_per_eu(ReconQWaves<R, GP>::value, 8)))

template <typename R, int TILE, bool GP>
__global__ __launch_bounds__(4 * TILE) AFAR_RECONQ_ATTR void k_reconstruct_q(DevPtrs<R> p) {
    using v4 = typename VT<R>::v4;
    using v2 = typename VT<R>::v2;
    constexpr int NT = 4 * TILE;
    // FP64 at the face centre: lane k forms its slot's geometry (face offset,
    // neighbour offset and distance) and least-squares weight from the staged
    // cell centres and the face centre, with k_lsq's and the host layout's
    // operations in their order, instead of reading 40 bytes of stored slot
    // data per slot; the weights are shared within the group by DPP.
    constexpr bool GEO = sizeof(R) == 8 && !GP;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    v4* sW = reinterpret_cast<v4*>(smem_raw);
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= p.ntiles) return;
    const int t = threadIdx.x;
    const int lc = t >> 2, k = t & 3;
    const int c = tile * TILE + lc;
    const int maxS = p.maxS, maxC = p.maxC, KR = maxC * (maxS + 1), nf = p.nf;
    // Every load that needs no LDS is issued before the staging barrier, in
    // two dependent rounds: (1) the cell's counts, the tile's own states,
    // halo ids and stencil block; (2) lane k's least-squares weight and slot
    // data, and the halo states.  Their latencies overlap.
    const bool inb = c < p.nc;
    int S = 0, C = 0, sf = 0, wasFlat = 0;
    if (inb) {
        S = p.scnt[c];
        C = p.ccnt[c];
        wasFlat = p.flat_prev[c];
    }
    const int hb = p.halo_ptr[tile], hn = p.halo_ptr[tile + 1] - hb;
    const int wpt = TILE >> 6;  // stencil block: (waves of the tile) * maxS * 64 ids
    const int nw = min(wpt, ((p.nc + 63) >> 6) - tile * wpt);
    const int n16 = nw * maxS * 8;
    const uint4* stSrc = reinterpret_cast<const uint4*>(p.sten + (size_t)tile * wpt * maxS * 64);
    const bool ownW = t < TILE && tile * TILE + t < p.nc;
    v4 wOwn;
    if (ownW) wOwn = p.W[tile * TILE + t];
    uint4 st0;
    if (t < n16) st0 = stSrc[t];
    int h0 = 0;
    if (t < hn) h0 = p.halo[hb + t];
    v2 Mk = v2{(R)0, (R)0};
    v2 d = v2{(R)0, (R)0};
    if (inb && k < C) {
        sf = p.slot_face[sell(c, k, maxC)];
        if constexpr (!GEO) {
            Mk = p.lsq[sell(c, k, maxC)];
            d = p.slot_d[sell(c, k, maxC)];
        }
    }
    v4 wHalo;
    if (t < hn) wHalo = p.W[h0];
    double2 cOwn = double2{0.0, 0.0}, cHalo = double2{0.0, 0.0};  // GEO: staged centres
    double2 fc = double2{0.0, 0.0}, fs = double2{0.0, 0.0};        // GEO: face centre, periodic shift
    if constexpr (GEO) {
        if (ownW) cOwn = p.cen[tile * TILE + t];
        if (t < hn) cHalo = p.cen[h0];
        if (inb && k < C) {
            const int ff = (sf < 0 ? -sf : sf) - 1;
            fc = p.fcen[ff];
            if (p.fshift && ff < nf) fs = p.fshift[ff];
        }
    }
    // A cell that took the full path at the previous launch most likely takes
    // it again: its RBF coefficients and centre distance are requested in this
    // round already (flat_prev only schedules loads; results do not depend on it).
    R cf[kMaxSten + 1];
    R deltaR = 0;
    const R* a = p.rbf + sell(c, k * (maxS + 1), KR);  // row stride 64
    const bool early = inb && k < C && !wasFlat;
    if (early) {
        if constexpr (!GEO) deltaR = p.slot_dr[sell(c, k, maxC)];
#pragma unroll
        for (int i = 0; i <= kMaxSten; ++i)
            if (i <= S) cf[i] = a[(size_t)i * 64];
    }
    uint16_t* sSt = reinterpret_cast<uint16_t*>(smem_raw + p.lds_sten_off);
    v2* sM = reinterpret_cast<v2*>(smem_raw + p.lds_m_off);  // 4 weights per cell
    double2* sC = reinterpret_cast<double2*>(smem_raw + p.lds_m_off);  // GEO: tile + halo centres instead
    uint4* stDst = reinterpret_cast<uint4*>(sSt);
    if (ownW) sW[t] = wOwn;
    if (t < hn) sW[TILE + t] = wHalo;
    if (t < n16) stDst[t] = st0;
    if constexpr (GEO) {
        if (ownW) sC[t] = cOwn;
        if (t < hn) sC[TILE + t] = cHalo;
        for (int i = t + NT; i < hn; i += NT) sC[TILE + i] = p.cen[p.halo[hb + i]];
    } else {
        sM[t] = Mk;
    }
    for (int i = t + NT; i < hn; i += NT) sW[TILE + i] = p.W[p.halo[hb + i]];  // halos beyond NT cells
    for (int i = t + NT; i < n16; i += NT) stDst[i] = stSrc[i];
    __syncthreads();
    if (!inb) return;  // the four lanes of a cell leave together
    const R* sWr = reinterpret_cast<const R*>(sW);

    // the temperature range over the compact stencil selects the branch
    R TMax = sWr[4 * lc + 3], TMin = TMax;
    for (int i = 0; i < C; ++i) {
        const R Tj = sWr[4 * sSt[sell(lc, i, maxS)] + 3];
        TMax = fmax(Tj, TMax);
        TMin = fmin(Tj, TMin);
    }
    // Lane k forms variable k's least-squares gradient and compact-stencil
    // extrema (the reference's per-variable sums, in the same order; the
    // extrema are exact) and stores the gradient; the group then shares them.
    // The full path does this after its RBF sums, so that fewer values are
    // live while the coefficients are in registers.
    const int isNei = sf < 0;
    const int f = (sf < 0 ? -sf : sf) - 1;
    const bool live = k < C && f < nf;
    double ex = 0.0, ey = 0.0;  // GEO: compact neighbour - cell centre of lane k's slot
    if constexpr (GEO) {
        // the host layout's expressions (face offset d, neighbour offset e)
        if (k < C) {
            const double2 cc = sC[lc];
            const double2 cn = sC[sSt[sell(lc, k, maxS)]];
            if (f < nf && !isNei) {
                d = v2{fc.x - cc.x, fc.y - cc.y};
                ex = (cn.x + fs.x) - cc.x;
                ey = (cn.y + fs.y) - cc.y;
            } else if (f < nf) {
                const double xn = cc.x + fs.x, yn = cc.y + fs.y;
                d = v2{fc.x - xn, fc.y - yn};
                ex = cn.x - xn;
                ey = cn.y - yn;
            } else {
                d = v2{fc.x - cc.x, fc.y - cc.y};
                ex = cn.x - cc.x;
                ey = cn.y - cc.y;
            }
        }
        // k_lsq's sums over the compact slots in order, then lane k's weight
        double Sxx = 0.0, Sxy = 0.0, Syy = 0.0;
        const double bx[4] = {quad_bcast<0>(ex), quad_bcast<1>(ex), quad_bcast<2>(ex), quad_bcast<3>(ex)};
        const double by[4] = {quad_bcast<0>(ey), quad_bcast<1>(ey), quad_bcast<2>(ey), quad_bcast<3>(ey)};
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (i < C) {
                Sxx += bx[i] * bx[i];
                Sxy += bx[i] * by[i];
                Syy += by[i] * by[i];
            }
        if (k < C)
            Mk = v2{(Syy * ex - Sxy * ey) / (Sxx * Syy - Sxy * Sxy), (-Sxy * ex + Sxx * ey) / (Sxx * Syy - Sxy * Sxy)};
    }
    R gxk = 0, gyk = 0, mxk = 0, mnk = 0;
    auto gradient = [&]() {
        const R wck = sWr[4 * lc + k];
        mxk = wck;
        mnk = wck;
        if constexpr (GEO) {
            // every slot's weight from its lane, in slot order
            const double mx[4] = {quad_bcast<0>(Mk.x), quad_bcast<1>(Mk.x), quad_bcast<2>(Mk.x), quad_bcast<3>(Mk.x)};
            const double my[4] = {quad_bcast<0>(Mk.y), quad_bcast<1>(Mk.y), quad_bcast<2>(Mk.y), quad_bcast<3>(Mk.y)};
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (i < C) {
                    const R cj = sWr[4 * sSt[sell(lc, i, maxS)] + k];
                    const R dphi = cj - wck;
                    gxk += mx[i] * dphi;
                    gyk += my[i] * dphi;
                    mxk = fmax(cj, mxk);
                    mnk = fmin(cj, mnk);
                }
        } else {
            for (int i = 0; i < C; ++i) {
                const v2 M = sM[4 * lc + i];
                const R cj = sWr[4 * sSt[sell(lc, i, maxS)] + k];
                const R dphi = cj - wck;
                gxk += M.x * dphi;
                gyk += M.y * dphi;
                mxk = fmax(cj, mxk);
                mnk = fmin(cj, mnk);
            }
        }
        v2* g = k == 0 ? p.gp : k == 1 ? p.gu : k == 2 ? p.gv : p.gT;
        g[c] = v2{gxk, gyk};
    };
    const bool flat = fabs(TMax - TMin) < K<R>::VSMALL;

    if (k == 0) p.flat_prev[c] = flat;
    // Full path not foreseen: every RBF coefficient of the face and its centre
    // distance are requested now, at once (one memory round trip).
    if (!flat && live && !early) {
        if constexpr (!GEO) deltaR = p.slot_dr[sell(c, k, maxC)];
#pragma unroll
        for (int i = 0; i <= kMaxSten; ++i)
            if (i <= S) cf[i] = a[(size_t)i * 64];
    }
    if constexpr (GEO) {
        if (!flat && live) deltaR = afar_hypot(ex, ey);  // k_lsq's centre distance
    }
    const R dx = d.x, dy = d.y;
    R fraction = 0;
    if (!flat && live) fraction = sqrt(dx * dx + dy * dy) / deltaR;
    const v4 wc = sW[lc];

    if (flat) {
        gradient();
        const R gpx = quad_bcast<0>(gxk), gpy = quad_bcast<0>(gyk), guxx = quad_bcast<1>(gxk),
                guyx = quad_bcast<1>(gyk), guxy = quad_bcast<2>(gxk), guyy = quad_bcast<2>(gyk),
                gTx = quad_bcast<3>(gxk), gTy = quad_bcast<3>(gyk);
        const R lT = to_bj((R)1);  // the BVD test's limiter of this cell
        if (k == 0) {
            p.lim[c] = v4{(R)1, (R)1, (R)1, (R)1};
            p.smooth[c] = 0;
            p.tlim[c] = lT < K<R>::VALID_LIM_FACTOR;
        }
        if (live) {
            v4 o;
            o.x = wc.x + gpx * dx + gpy * dy;
            o.y = wc.y + guxx * dx + guyx * dy;
            o.z = wc.z + guxy * dx + guyy * dy;
            o.w = wc.w + gTx * dx + gTy * dy;
            store_face_state(p, 2 * f + isNei, o);
            p.tTVD[2 * f + isNei] = wc.w + lT * (gTx * dx + gTy * dy);
            if constexpr (GP)
                for (int g = 0; g < p.ngp; ++g) {  // flux points (extension): the same linear state there
                const v4 dg = p.fgd[(size_t)g * nf + f];
                const R gx = (R)(isNei ? dg.z : dg.x), gy = (R)(isNei ? dg.w : dg.y);
                v4 og;
                og.x = wc.x + gpx * gx + gpy * gy;
                og.y = wc.y + guxx * gx + guyx * gy;
                og.z = wc.z + guxy * gx + guyy * gy;
                og.w = wc.w + gTx * gx + gTy * gy;
                p.fvg[((size_t)g * nf + f) * 2 + isNei] = og;
            }
        }
        return;
    }

    R pf = 0, uxf = 0, uyf = 0, Tf = 0;
    if (live) {
        R coef = cf[0];
        pf = coef * wc.x;
        uxf = coef * wc.y;
        uyf = coef * wc.z;
        Tf = coef * wc.w;
#pragma unroll
        for (int ci = 0; ci < kMaxSten; ++ci) {
            if (ci < S) {
                const v4 wj = sW[sSt[sell(lc, ci, maxS)]];
                coef = cf[ci + 1];
                pf += coef * wj.x;
                uxf += coef * wj.y;
                uyf += coef * wj.z;
                Tf += coef * wj.w;
            }
        }
    }
    gradient();
    const R gpx = quad_bcast<0>(gxk), gpy = quad_bcast<0>(gyk), guxx = quad_bcast<1>(gxk),
            guyx = quad_bcast<1>(gyk), guxy = quad_bcast<2>(gxk), guyy = quad_bcast<2>(gyk),
            gTx = quad_bcast<3>(gxk), gTy = quad_bcast<3>(gyk);
    const R pMax = quad_bcast<0>(mxk), pMin = quad_bcast<0>(mnk), uxMax = quad_bcast<1>(mxk),
            uxMin = quad_bcast<1>(mnk), uyMax = quad_bcast<2>(mxk), uyMin = quad_bcast<2>(mnk);

    R limP = 2, limUx = 2, limUy = 2, limT = 2;
    R smoothness = 0;
    int faceNotTrouble = 1;
    if (live) {
        const R rP = fabs(pMax - pMin), rUx = fabs(uxMax - uxMin), rUy = fabs(uyMax - uyMin), rT = fabs(TMax - TMin);
        const int nbl = sSt[sell(lc, k, maxS)];
        v4 o;
        R phif;
        limiter_process(limP, phif, wc.x, pMax, pMin, gpx, gpy, dx, dy);
        o.x = (R)(rP >= K<R>::SMALL) * pf + (R)(rP < K<R>::SMALL) * phif;
        smoothness = fmax(smoothness, fabs(((R)1 - fraction) * wc.x + fraction * sW[nbl].x - phif) / deltaR +
                                          (R)(pf <= (R)0) * (R)1e20 + (R)(phif <= (R)0) * (R)1e20);
        limiter_process(limUx, phif, wc.y, uxMax, uxMin, guxx, guyx, dx, dy);
        o.y = (R)(rUx >= K<R>::SMALL) * uxf + (R)(rUx < K<R>::SMALL) * phif;
        limiter_process(limUy, phif, wc.z, uyMax, uyMin, guxy, guyy, dx, dy);
        o.z = (R)(rUy >= K<R>::SMALL) * uyf + (R)(rUy < K<R>::SMALL) * phif;
        limiter_process(limT, phif, wc.w, TMax, TMin, gTx, gTy, dx, dy);
        o.w = (R)(rT >= K<R>::SMALL) * Tf + (R)(rT < K<R>::SMALL) * phif;
        faceNotTrouble = phif > (R)0;
        store_face_state(p, 2 * f + isNei, o);
        // flux points (extension): RBF sums with the point's coefficients,
        // or the unlimited linear state where the compact range is flat
        if constexpr (GP)
        for (int g = 0; g < p.ngp; ++g) {
            const R* ag = p.rbf_gp + sell(c, (g * maxC + k) * (maxS + 1), p.ngp * KR);
            R cg = ag[0];
            R pg = cg * wc.x, uxg = cg * wc.y, uyg = cg * wc.z, Tg = cg * wc.w;
            for (int ci = 0; ci < S; ++ci) {
                const v4 wj = sW[sSt[sell(lc, ci, maxS)]];
                cg = ag[(size_t)(ci + 1) * 64];
                pg += cg * wj.x;
                uxg += cg * wj.y;
                uyg += cg * wj.z;
                Tg += cg * wj.w;
            }
            const v4 dg = p.fgd[(size_t)g * nf + f];
            const R gx = (R)(isNei ? dg.z : dg.x), gy = (R)(isNei ? dg.w : dg.y);
            v4 og;
            og.x = (R)(rP >= K<R>::SMALL) * pg + (R)(rP < K<R>::SMALL) * (wc.x + (gpx * gx + gpy * gy));
            og.y = (R)(rUx >= K<R>::SMALL) * uxg + (R)(rUx < K<R>::SMALL) * (wc.y + (guxx * gx + guyx * gy));
            og.z = (R)(rUy >= K<R>::SMALL) * uyg + (R)(rUy < K<R>::SMALL) * (wc.z + (guxy * gx + guyy * gy));
            og.w = (R)(rT >= K<R>::SMALL) * Tg + (R)(rT < K<R>::SMALL) * (wc.w + (gTx * gx + gTy * gy));
            p.fvg[((size_t)g * nf + f) * 2 + isNei] = og;
        }
    }
    // group reductions (lanes 4j..4j+3 of the wavefront)
    limP = fmin(limP, quad_xor<1>(limP));
    limUx = fmin(limUx, quad_xor<1>(limUx));
    limUy = fmin(limUy, quad_xor<1>(limUy));
    limT = fmin(limT, quad_xor<1>(limT));
    smoothness = fmax(smoothness, quad_xor<1>(smoothness));
    limP = fmin(limP, quad_xor<2>(limP));
    limUx = fmin(limUx, quad_xor<2>(limUx));
    limUy = fmin(limUy, quad_xor<2>(limUy));
    limT = fmin(limT, quad_xor<2>(limT));
    smoothness = fmax(smoothness, quad_xor<2>(smoothness));
    // the trouble test of the last internal face in slot order
    const unsigned long long lm = __ballot(live);
    const int base = (t & 63) & ~3;
    const unsigned grp = (unsigned)(lm >> base) & 0xFu;
    const int lastLane = grp ? 31 - __builtin_clz(grp) : 0;
    const int nt0 = __shfl(faceNotTrouble, base + lastLane);
    const int isNotTrouble = grp ? nt0 : 1;
    const R nt = (R)isNotTrouble;
    const R limTraw = fmax(limT, (R)0) * nt;
    // the BVD test's inputs (cudaLimiter.h:431-440): the limited linear
    // temperature on lane k's face, and the cell's limiter condition
    const R lT = to_bj(limTraw);
    if (live) p.tTVD[2 * f + isNei] = wc.w + lT * (gTx * dx + gTy * dy);
    if (k == 0) {
        p.lim[c] = v4{fmax(limP, (R)0) * nt, fmax(limUx, (R)0) * nt, fmax(limUy, (R)0) * nt, limTraw};
        smoothness = smoothness / ((R)0.5 * (pMax + pMin) * nt + K<R>::VSMALL);
        p.smooth[c] = smoothness > (R)1;
        p.tlim[c] = lT < K<R>::VALID_LIM_FACTOR;
    }
}

// ---------------------------------------------------------------- K6 BVD indicator


template <typename R, int TILE>
__global__ __launch_bounds__(TILE) void k_bvd(DevPtrs<R> p) {
    using v4 = typename VT<R>::v4;
    constexpr int PF = 4;  // face slots whose data is requested in the first load rounds
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int8_t* sS = reinterpret_cast<int8_t*>(smem_raw);
    uint16_t* sSt = reinterpret_cast<uint16_t*>(smem_raw + p.lds_bvd_sten_off);
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= p.ntiles) return;
    const int tid = threadIdx.x, c = tile * TILE + tid;
    const int maxS = p.maxS, maxC = p.maxC, nf = p.nf;
    const bool inb = c < p.nc;
    // Two rounds of independent loads before the tile's flag vote: (1) the cell's
    // counts, flag and limiter condition, the first slots' faces and halo
    // ids; (2) halo flags and, in cells meeting the limiter condition, the
    // slots' face temperatures (both sides' O3 values and this side's limited
    // linear value, formed by the reconstruction).  Halo and stencil indices
    // are clamped rather than branched on, and values held in plain scalars,
    // so that no load is waited for before its round is complete.  The
    // stencil block is read only by tiles with a flagged cell among their own
    // and halo cells: elsewhere every flag sum is zero.
    const int cc = inb ? c : p.nc - 1;
    const int S = p.scnt[cc], C = inb ? (int)p.ccnt[cc] : 0;
    const int8_t own = p.smooth[cc];
    const int limited = p.tlim[cc];
    int sfs[PF];
#pragma unroll
    for (int fi = 0; fi < PF; ++fi) sfs[fi] = fi < maxC ? p.slot_face[sell(cc, fi, maxC)] : 0;
    const int hb = p.halo_ptr[tile], hn = p.halo_ptr[tile + 1] - hb;
    const int wpt = TILE >> 6;
    const int nw = min(wpt, ((p.nc + 63) >> 6) - tile * wpt);
    const int n16 = nw * maxS * 8;
    const uint4* stSrc = reinterpret_cast<const uint4*>(p.sten + (size_t)tile * wpt * maxS * 64);
    const int hmax = hn > 0 ? hn - 1 : 0;  // a tile without halo reads its first id harmlessly
    const unsigned hid0 = (unsigned)p.halo[hb + min(tid, hmax)];
    const unsigned hid1 = (unsigned)p.halo[hb + min(tid + TILE, hmax)];
    const int8_t hs0 = p.smooth[hid0];
    const int8_t hs1 = p.smooth[hid1];
    // A face's test counts only in a limited cell (the limiter condition, or
    // a flagged extended stencil, known after the flag sums): the
    // temperatures are requested here for cells meeting the condition and,
    // after the sums, for the others that turn out limited.  (Requesting
    // them in a flagged tile for every cell, with the stencil block, measured
    // slower.)
    R tfo[PF], tfn[PF], tvd[PF];
    auto load_temps = [&](int fi) {
        const int sf = sfs[fi];
        const int f = (sf < 0 ? -sf : sf) - 1;
        if (fi < C && f < nf) {
            tfo[fi] = p.fvT[2 * f];
            tfn[fi] = p.fvT[2 * f + 1];
            tvd[fi] = p.tTVD[2 * f + (sf < 0)];
        }
    };
#pragma unroll
    for (int fi = 0; fi < PF; ++fi) {
        tfo[fi] = tfn[fi] = tvd[fi] = (R)0;
        if (limited) load_temps(fi);
    }
    if (inb) sS[tid] = own;
    if (tid < hn) sS[TILE + tid] = hs0;
    if (tid + TILE < hn) sS[2 * TILE + tid] = hs1;
    int flagged = (inb && own != 0) || (tid < hn && hs0 != 0) || (tid + TILE < hn && hs1 != 0);
    for (int i = tid + 2 * TILE; i < hn; i += TILE) {
        const int8_t h = p.smooth[p.halo[hb + i]];
        sS[TILE + i] = h;
        flagged |= h != 0;
    }
    int8_t shockType = own;
    int8_t shockTypeExt = own;
    // block-uniform: one wavefront's vote, or a workgroup reduction
    const bool anyFlagged = TILE == 64 ? __any(flagged) != 0 : __syncthreads_or(flagged) != 0;
    if (anyFlagged) {
        uint4* stDst = reinterpret_cast<uint4*>(sSt);
        const uint4 st0 = stSrc[min(tid, n16 - 1)];
        const uint4 st1 = stSrc[min(tid + TILE, n16 - 1)];
        const uint4 st2 = stSrc[min(tid + 2 * TILE, n16 - 1)];
        if (tid < n16) stDst[tid] = st0;
        if (tid + TILE < n16) stDst[tid + TILE] = st1;
        if (tid + 2 * TILE < n16) stDst[tid + 2 * TILE] = st2;
        for (int i = tid + 3 * TILE; i < n16; i += TILE) stDst[i] = stSrc[i];
        __syncthreads();
        if (inb) {
            for (int ci = 0; ci < C; ++ci) {
                shockType += sS[sSt[sell(tid, ci, maxS)]];
                shockTypeExt = shockType;
            }
            for (int ci = C; ci < S; ++ci) shockTypeExt += sS[sSt[sell(tid, ci, maxS)]];
        }
    }
    if (!inb) return;

    // Every face is visited; the reference leaves the loop at a cell's second
    // boundary face with bvd = 1, the largest value, so the result is the same.
    // cudaLimiter.h:431-448, with Tf_cur_TVD = Tc + BJ(lim) * (grad T . d) and
    // the condition lim < VALID_LIM_FACTOR taken from the reconstruction.
    int bvd = 0, bfaces = 0;
    const int isLimited = limited + (shockTypeExt > 0);
    if (isLimited && !limited) {
#pragma unroll
        for (int fi = 0; fi < PF; ++fi) load_temps(fi);
    }
    auto face = [&](int sf, R TfOwn, R TfNei, R Tf_cur_TVD) {
        const int isNei = sf < 0;
        const R Tf_cur = (R)isNei * TfNei + (R)(1 - isNei) * TfOwn;
        const R Tf_nei = (R)(1 - isNei) * TfNei + (R)isNei * TfOwn;
        const int typeval = fabs(Tf_cur_TVD - Tf_nei) < K<R>::BVD_FACTOR * fabs(Tf_cur - Tf_nei);
        const int isNotFlat = fabs(Tf_cur - Tf_nei) > K<R>::VSMALL;
        bvd = max(bvd, isNotFlat * isLimited * typeval);
    };
#pragma unroll
    for (int fi = 0; fi < PF; ++fi) {
        if (fi < C) {
            const int sf = sfs[fi];
            const int f = (sf < 0 ? -sf : sf) - 1;
            if (f >= nf)
                bfaces += 1;
            else if (isLimited)
                face(sf, tfo[fi], tfn[fi], tvd[fi]);
        }
    }
    for (int fi = PF; fi < C; ++fi) {  // cells with more than PF faces
        const int sf = p.slot_face[sell(c, fi, maxC)];
        const int f = (sf < 0 ? -sf : sf) - 1;
        if (f >= nf)
            bfaces += 1;
        else if (isLimited)
            face(sf, p.fvT[2 * f], p.fvT[2 * f + 1], p.tTVD[2 * f + (sf < 0)]);
    }
    if (bfaces >= 2) bvd = 1;  // corner cells use the second-order reconstruction
    int finalType = (shockType > 0) * (shockTypeExt > 1) * (-2) + (bvd > 0);
    if (p.boundaryFirstOrder && bfaces > 0) {
        // extension: a cell with a boundary face takes the first-order state
        // on its faces (TVD with zero limiter), generalising the reference's
        // corner-cell rule; shock types keep their flux choice.  A raw
        // limiter is never negative, so -1 marks the zero limiter.
        finalType = finalType ? finalType : 1;
        p.ftype[c] = (int8_t)finalType;
        p.lim[c] = v4{(R)-1, (R)-1, (R)-1, (R)-1};
        return;
    }
    // The limiter transform (cudaLimiter.h:453-461) is applied where the
    // limiter is used, by the face pass, from this type and the raw limiter:
    // the same function of the same values, without a write of the array.
    p.ftype[c] = (int8_t)finalType;
}

// The limiter the face pass uses: the transform of K6 applied to the raw K5
// limiter with the cell's final type, or zero for a first-order cell.  The
// reference blends both transforms, isShock * venkat + (1 - isShock) * bj
// (cudaLimiter.h:453-461); only the selected one is evaluated here, so that a
// wavefront without shock cells skips the Venkatakrishnan divisions.  The
// value is the blend's bit for bit: both transforms of a raw limiter >= 0 are
// finite (fmin(., 1) of a finite or NaN value), the Venkatakrishnan value is
// positive, and adding zero turns a BJ value of -0, where the blend gives
// +0 + -0 = +0, into +0.
template <typename R> __device__ __forceinline__ R face_limiter(R raw, int isShock) {
    if (raw < (R)0) return (R)0;
    const R t = isShock ? to_venkat(raw) : to_bj(raw);
    return t + (R)0;
}

// ---------------------------------------------------------------- K7 flux

template <typename R>
__device__ __forceinline__ R minimize_delta(R L, R Rv, R LT, R RT) {
    // cudaRiemannSolver.h:7-14
    R d = fabs(Rv - L);
    d = fmin(d, fabs(Rv - LT));
    d = fmin(d, fabs(RT - L));
    d = fmin(d, fabs(RT - LT));
    return (R)(Rv >= L) * d - (R)(Rv < L) * d;
}

template <typename R>
__device__ __forceinline__ void kfvs_flux(R& rhoFlux, R& rhoUxFlux, R& rhoUyFlux, R& rhoEFlux, R rhoL, R rhoR,
                                          R pL, R pR, R uxL, R uxR, R uyL, R uyR, R TL, R TR, R Rg, R Cv, R magSf,
                                          R Sfx, R Sfy) {
    // cudaRiemannSolver.h:16-68
    const R nfx = Sfx / magSf, nfy = Sfy / magSf;
    const R uxLl = uxL * nfx + uyL * nfy, uyLl = uyL * nfx - uxL * nfy;
    const R uxRl = uxR * nfx + uyR * nfy, uyRl = uyR * nfx - uxR * nfy;
    const R lamL = (R)0.5 * rhoL / pL, lamR = (R)0.5 * rhoR / pR;
    const R PI = (R)3.141593;
    const R Mu0p = (R)0.5 * kerfc(-sqrt(lamL) * uxLl);
    const R Mu1p = uxLl * Mu0p + (R)0.5 * kexp(-lamL * uxLl * uxLl) / (sqrt(PI * lamL));
    const R Mu2p = uxLl * Mu1p + (R)0.5 * Mu0p / lamL;
    const R Mu3p = uxLl * Mu2p + Mu1p / lamL;
    const R Mu0n = (R)0.5 * kerfc(sqrt(lamR) * uxRl);
    const R Mu1n = uxRl * Mu0n - (R)0.5 * kexp(-lamR * uxRl * uxRl) / (sqrt(PI * lamR));
    const R Mu2n = uxRl * Mu1n + (R)0.5 * Mu0n / lamR;
    const R Mu3n = uxRl * Mu2n + Mu1n / lamR;
    const R Mv1L = uyLl, Mv1R = uyRl;
    const R Kd = (R)2 * Cv / Rg - (R)2;
    const R trho = rhoL * Mu1p + rhoR * Mu1n;
    const R trhoUx = rhoL * Mu2p + rhoR * Mu2n;
    const R trhoUy = rhoL * Mu1p * Mv1L + rhoR * Mu1n * Mv1R;
    const R trhoE = (R)0.5 * rhoL * (Mu3p + Mu1p * (uyLl * Mv1L + (R)0.5 / lamL) + Mu1p * Kd * (R)0.5 / lamL) +
                    (R)0.5 * rhoR * (Mu3n + Mu1n * (uyRl * Mv1R + (R)0.5 / lamR) + Mu1n * Kd * (R)0.5 / lamR);
    rhoUxFlux += trhoUx * Sfx - trhoUy * Sfy - (R)0.5 * (rhoFlux * uxL + pL * Sfx) - (R)0.5 * (rhoFlux * uxR + pR * Sfx);
    rhoUyFlux += trhoUx * Sfy + trhoUy * Sfx - (R)0.5 * (rhoFlux * uyL + pL * Sfy) - (R)0.5 * (rhoFlux * uyR + pR * Sfy);
    rhoEFlux += trhoE * magSf - (R)0.5 * rhoFlux * (Cv * TL + (R)0.5 * (uxL * uxL + uyL * uyL) + pL / rhoL) -
                (R)0.5 * rhoFlux * (Cv * TR + (R)0.5 * (uxR * uxR + uyR * uyR) + pR / rhoR);
    rhoFlux += trho * magSf - (R)0.5 * rhoL * (uxL * Sfx + uyL * Sfy) - (R)0.5 * rhoR * (uxR * Sfx + uyR * Sfy);
}

template <typename R>
__device__ __forceinline__ void roe_flux(R& rhoFlux, R& rhoUxFlux, R& rhoUyFlux, R& rhoEFlux, R rhoL, R rhoR, R pL,
                                         R pR, R uxL, R uxR, R uyL, R uyR, R TL, R TR, R dP, R dRho, R gamma, R Cv,
                                         R magSf, R Sfx, R Sfy) {
    // cudaRiemannSolver.h:71-136
    const R nfx = Sfx / magSf, nfy = Sfy / magSf;
    const R VL = uxL * nfx + uyL * nfy, VR = uxR * nfx + uyR * nfy;
    const R hL = Cv * TL + (R)0.5 * (uxL * uxL + uyL * uyL) + pL / rhoL;
    const R hR = Cv * TR + (R)0.5 * (uxR * uxR + uyR * uyR) + pR / rhoR;
    const R rhoT = sqrt(fmax(rhoL * rhoR, (R)1e-20));
    const R wL = sqrt(fmax(rhoL, (R)1e-20)) / (sqrt(fmax(rhoL, (R)1e-20)) + sqrt(fmax(rhoR, (R)1e-20)));
    const R wR = (R)1 - wL;
    const R uxT = uxL * wL + uxR * wR, uyT = uyL * wL + uyR * wR, hT = hL * wL + hR * wR;
    const R q2 = uxT * uxT + uyT * uyT;
    const R cT = sqrt(fmax((gamma - (R)1) * (hT - (R)0.5 * q2), (R)1e-20));
    const R VT_ = uxT * nfx + uyT * nfy;
    const R dUx = uxR - uxL, dUy = uyR - uyL;
    const R dV = dUx * nfx + dUy * nfy;
    const R r1 = (dP - rhoT * cT * dV) / ((R)2 * cT * cT);
    const R r2 = dRho - dP / (cT * cT);
    const R r3 = (dP + rhoT * cT * dV) / ((R)2 * cT * cT);
    const R cL = sqrt(fmax((gamma - (R)1) * (hL - (R)0.5 * (uxL * uxL + uyL * uyL)), (R)1e-20));
    const R cR = sqrt(fmax((gamma - (R)1) * (hR - (R)0.5 * (uxR * uxR + uyR * uyR)), (R)1e-20));
    R eps = (R)2 * fmax((R)0, (VR - cR) - (VL - cL));
    R l1 = fabs(VT_ - cT);
    l1 = (R)(l1 >= eps) * l1 + (R)(l1 < eps) * ((l1 * l1 + eps * eps) / ((R)2 * eps + (R)1e-30));
    eps = (R)2 * fmax((R)0, VR - VL);
    R l2 = fabs(VT_);
    l2 = (R)(l2 >= eps) * l2 + (R)(l2 < eps) * ((l2 * l2 + eps * eps) / ((R)2 * eps + (R)1e-30));
    eps = (R)2 * fmax((R)0, (VR + cR) - (VL + cL));
    R l3 = fabs(VT_ + cT);
    l3 = (R)(l3 >= eps) * l3 + (R)(l3 < eps) * ((l3 * l3 + eps * eps) / ((R)2 * eps + (R)1e-30));
    rhoFlux -= (R)0.5 * (l1 * r1 + l2 * r2 + l3 * r3) * magSf;
    rhoUxFlux -= (R)0.5 * (l1 * r1 * (uxT - cT * nfx) + l2 * (r2 * uxT + rhoT * (dUx - dV * nfx)) + l3 * r3 * (uxT + cT * nfx)) * magSf;
    rhoUyFlux -= (R)0.5 * (l1 * r1 * (uyT - cT * nfy) + l2 * (r2 * uyT + rhoT * (dUy - dV * nfy)) + l3 * r3 * (uyT + cT * nfy)) * magSf;
    rhoEFlux -= (R)0.5 * (l1 * r1 * (hT - cT * VT_) + l2 * (r2 * (R)0.5 * q2 + rhoT * (uxT * dUx + uyT * dUy - VT_ * dV)) +
                          l3 * r3 * (hT + cT * VT_)) * magSf;
}

// Waves per SIMD the register allocator must allow for the face-centre flux
// pass, per working type (build-time tuning knobs; 1 = no constraint).  The
// FP64 passes take 104 VGPRs (4 waves) unconstrained.  At 5 waves the pass
// without the CFL bound fits in 92 VGPRs without spills, yet measured 4%
// slower per launch (the pass with it spills 20 bytes per lane); 6 and 8
// waves spill (36 and 132 bytes per lane) and measured 26% and 108% slower.
#ifndef AFAR_FLUX_WPE_D
#define AFAR_FLUX_WPE_D 1
#endif
#ifndef AFAR_FLUX_DT_WPE_D
#define AFAR_FLUX_DT_WPE_D 1
#endif
#ifndef AFAR_FLUX_WPE_F
#define AFAR_FLUX_WPE_F 1
#endif
template <typename R, bool DT, bool GP> struct FluxWaves {
    static constexpr int value =
        GP ? 1 : (sizeof(R) == 8 ? (DT ? AFAR_FLUX_DT_WPE_D : AFAR_FLUX_WPE_D) : AFAR_FLUX_WPE_F);
};

// DT: also reduce the CFL bound of every internal face to one partial minimum
// per block (dt_part[blockIdx.x]); the bound is a function of the
// primitive state alone, which no kernel of the stage has changed yet.
template <typename R, bool DT, bool GP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FluxWaves<R, DT, GP>::value, 8))) void k_flux(
    DevPtrs<R> p, double Rgas, double Cv, double CFL) {
    using v4 = typename VT<R>::v4;
    using v2 = typename VT<R>::v2;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int i = tile * 256 + threadIdx.x;
    const int nf = p.nf;
    double dtf = 1e5;
    if (i < nf) {
        const int f = i;
        const int2 oc = p.fcell[f];
        const int own = oc.x, nei = oc.y;
        double4 g0, g1;
        if constexpr (sizeof(R) == 4 && !DT) {  // the geometry as the float arithmetic uses it
            const float4 h0 = p.fgeo32[2 * f], h1 = p.fgeo32[2 * f + 1];
            g0 = double4{h0.x, h0.y, h0.z, h0.w};
            g1 = double4{h1.x, h1.y, h1.z, h1.w};
        } else if constexpr (!GP) {
            // the host layout's expressions in FP64: offsets of the face
            // centre from the owner's centre and from the (shifted)
            // neighbour's centre, and the distance between the two centres
            // (FP32-mixed rounds the offsets to float below, as it rounds the
            // stored record)
            const double2 sf = p.fsf[f], fc = p.fcen[f], co = p.cen[own], cn = p.cen[nei];
            const double2 fs = p.fshift ? p.fshift[f] : double2{0.0, 0.0};
            const double xn = cn.x + fs.x, yn = cn.y + fs.y;
            double dR = 0.0;
            if (DT) {
                const double ex = xn - co.x, ey = yn - co.y;
                dR = afar_hypot(ex, ey);
            }
            g0 = double4{sf.x, sf.y, p.fmag[f], dR};
            g1 = double4{fc.x - co.x, fc.y - co.y, fc.x - xn, fc.y - yn};
        } else {
            g0 = p.fgeo[2 * f];
            g1 = p.fgeo[2 * f + 1];
        }
        const R dxL = (R)g1.x, dyL = (R)g1.y, dxR = (R)g1.z, dyR = (R)g1.w;
        const int8_t to = p.ftype[own], tn = p.ftype[nei];
        const int isTVD_L = to != 0, isTVD_R = tn != 0;
        const int isShock = to < 0 || tn < 0;
        const R magSf = (R)g0.z, Sfx = (R)g0.x, Sfy = (R)g0.y;
        const R Rg = (R)Rgas;
        const v4 wL = p.W[own], wR = p.W[nei];
        const v4 rL = p.lim[own], rR = p.lim[nei];
        const int shL = to < 0, shR = tn < 0;
        const v4 lL = v4{face_limiter(rL.x, shL), face_limiter(rL.y, shL), face_limiter(rL.z, shL),
                         face_limiter(rL.w, shL)};
        const v4 lR = v4{face_limiter(rR.x, shR), face_limiter(rR.y, shR), face_limiter(rR.z, shR),
                         face_limiter(rR.w, shR)};
        // face-centre O3 states of both sides
        v4 fL0{}, fR0{};
        if constexpr (!GP) {
            if constexpr (sizeof(R) == 8) {  // one 2-vector per variable
                const v2 sp = reinterpret_cast<const v2*>(p.fvp)[f], su = reinterpret_cast<const v2*>(p.fvu)[f];
                const v2 sv = reinterpret_cast<const v2*>(p.fvv)[f], sT = reinterpret_cast<const v2*>(p.fvT)[f];
                fL0 = v4{sp.x, su.x, sv.x, sT.x};
                fR0 = v4{sp.y, su.y, sv.y, sT.y};
            } else {
                fL0 = p.fv[2 * f];
                fR0 = p.fv[2 * f + 1];
            }
        }
        const v2 gpL = p.gp[own], gTL = p.gT[own];
        const v2 gpR = p.gp[nei], gTR = p.gT[nei];
        // The limited linear velocities enter the flux only on a TVD side: on
        // an O3 side they are blended with weight 0, which adds a zero to the
        // O3 velocity, so they change nothing unless that velocity is itself a
        // zero (whose sign the added zero decides).  Their gradients are read
        // only where they can matter (at the face centre).  A non-finite
        // velocity gradient beside a finite O3 state (a degenerate
        // least-squares stencil) would turn that blend into NaN in the
        // reference; here it is not read.
        v2 guL{}, gvL{}, guR{}, gvR{};
        if (GP || isTVD_L || fL0.y == (R)0 || fL0.z == (R)0) {
            guL = p.gu[own];
            gvL = p.gv[own];
        }
        if (GP || isTVD_R || fR0.y == (R)0 || fR0.z == (R)0) {
            guR = p.gu[nei];
            gvR = p.gv[nei];
        }
        if (DT) {
            const double d = cfl_face_dt(g0, g1, wL, wR, Rgas, Cv, CFL);
            dtf = d < dtf ? d : dtf;
        }

        // the flux of the two states reconstructed at a point with offsets
        // (dxL, dyL) from the owner and (dxR, dyR) from the neighbour; fL, fR
        // are the high-order (O3) states there
        auto point_flux = [&](R dxL, R dyL, R dxR, R dyR, const v4& fL, const v4& fR, R& out0, R& out1, R& out2,
                              R& out3) {
            const R pLT = wL.x + lL.x * (gpL.x * dxL + gpL.y * dyL);
            const R uxLT = wL.y + lL.y * (guL.x * dxL + guL.y * dyL);
            const R uyLT = wL.z + lL.z * (gvL.x * dxL + gvL.y * dyL);
            const R TLT = wL.w + lL.w * (gTL.x * dxL + gTL.y * dyL);
            const R rhoLT = pLT / (Rg * TLT);
            const R pRT = wR.x + lR.x * (gpR.x * dxR + gpR.y * dyR);
            const R uxRT = wR.y + lR.y * (guR.x * dxR + guR.y * dyR);
            const R uyRT = wR.z + lR.z * (gvR.x * dxR + gvR.y * dyR);
            const R TRT = wR.w + lR.w * (gTR.x * dxR + gTR.y * dyR);
            const R rhoRT = pRT / (Rg * TRT);

            const R pL = (R)(1 - isTVD_L) * fL.x + (R)isTVD_L * pLT;
            const R uxL = (R)(1 - isTVD_L) * fL.y + (R)isTVD_L * uxLT;
            const R uyL = (R)(1 - isTVD_L) * fL.z + (R)isTVD_L * uyLT;
            const R TL = (R)(1 - isTVD_L) * fL.w + (R)isTVD_L * TLT;
            const R rhoL = pL / (Rg * TL);
            const R pR = (R)(1 - isTVD_R) * fR.x + (R)isTVD_R * pRT;
            const R uxR = (R)(1 - isTVD_R) * fR.y + (R)isTVD_R * uxRT;
            const R uyR = (R)(1 - isTVD_R) * fR.z + (R)isTVD_R * uyRT;
            const R TR = (R)(1 - isTVD_R) * fR.w + (R)isTVD_R * TRT;
            const R rhoR = pR / (Rg * TR);

            const R dP = minimize_delta(pL, pR, pLT, pRT);
            const R dRho = minimize_delta(rhoL, rhoR, rhoLT, rhoRT);

            const R CvR = (R)Cv;
            R rhoFlux = (R)0.5 * rhoL * (uxL * Sfx + uyL * Sfy) + (R)0.5 * rhoR * (uxR * Sfx + uyR * Sfy);
            R rhoUxFlux = (R)0.5 * (rhoFlux * uxL + pL * Sfx) + (R)0.5 * (rhoFlux * uxR + pR * Sfx);
            R rhoUyFlux = (R)0.5 * (rhoFlux * uyL + pL * Sfy) + (R)0.5 * (rhoFlux * uyR + pR * Sfy);
            R rhoEFlux = (R)0.5 * rhoFlux * (CvR * TL + (R)0.5 * (uxL * uxL + uyL * uyL) + pL / rhoL) +
                         (R)0.5 * rhoFlux * (CvR * TR + (R)0.5 * (uxR * uxR + uyR * uyR) + pR / rhoR);
            if (isShock)
                kfvs_flux<R>(rhoFlux, rhoUxFlux, rhoUyFlux, rhoEFlux, rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, Rg,
                             CvR, magSf, Sfx, Sfy);
            else
                roe_flux<R>(rhoFlux, rhoUxFlux, rhoUyFlux, rhoEFlux, rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, dP,
                            dRho, (R)((Cv + Rgas) / Cv), CvR, magSf, Sfx, Sfy);
            out0 = rhoFlux;
            out1 = rhoUxFlux;
            out2 = rhoUyFlux;
            out3 = rhoEFlux;
        };
        if constexpr (!GP) {
            R rhoFlux, rhoUxFlux, rhoUyFlux, rhoEFlux;
            point_flux(dxL, dyL, dxR, dyR, fL0, fR0, rhoFlux, rhoUxFlux, rhoUyFlux, rhoEFlux);
            if constexpr (sizeof(R) == 4)
                p.flux32[f] = float4{rhoFlux, rhoUxFlux, rhoUyFlux, rhoEFlux};
            else
                p.flux[f] = double4{(double)rhoFlux, (double)rhoUxFlux, (double)rhoUyFlux, (double)rhoEFlux};
        } else {
            // extension: weighted sum over the face's quadrature points
            double4 acc = double4{0.0, 0.0, 0.0, 0.0};
            for (int g = 0; g < p.ngp; ++g) {
                const v4 d = p.fgd[(size_t)g * nf + f];
                R rhoFlux, rhoUxFlux, rhoUyFlux, rhoEFlux;
                point_flux((R)d.x, (R)d.y, (R)d.z, (R)d.w, p.fvg[((size_t)g * nf + f) * 2],
                           p.fvg[((size_t)g * nf + f) * 2 + 1], rhoFlux, rhoUxFlux, rhoUyFlux, rhoEFlux);
                const double w = p.gpw[g];
                if (g == 0) {
                    acc = double4{w * (double)rhoFlux, w * (double)rhoUxFlux, w * (double)rhoUyFlux,
                                  w * (double)rhoEFlux};
                } else {
                    acc.x += w * (double)rhoFlux;
                    acc.y += w * (double)rhoUxFlux;
                    acc.z += w * (double)rhoUyFlux;
                    acc.w += w * (double)rhoEFlux;
                }
            }
            p.flux[f] = acc;
        }
    } else if (i < nf + p.nb) {
        // boundary faces: physical flux of the ghost state (cudaRiemannSolver.h:254-270)
        const int b = i - nf;
        const double4 g = p.bgeo[b];
        const typename VT<R>::v4 w = p.W[p.nc + b];
        const double bp = w.x, bUx = w.y, bUy = w.z, bT = w.w;
        const double rho = bp / (Rgas * bT);
        const double rhoFlux = rho * (bUx * g.x + bUy * g.y);
        p.flux[i] = double4{rhoFlux, rhoFlux * bUx + bp * g.x, rhoFlux * bUy + bp * g.y,
                            rhoFlux * (Cv * bT + 0.5 * (bUx * bUx + bUy * bUy) + bp / rho)};
    }
    if (DT) {
        // Two partials per tile of 256 faces: slot 0 for the instance of the
        // tile's first face, slot 1 for the next one (every instance of a
        // batch has at least 256 internal faces, so a tile meets at most two).
        const int f0 = min(tile * 256, nf - 1);
        const int i0 = (p.cinst && nf > 0) ? p.cinst[p.fcell[f0].x] : 0;
        const int mine = (i < nf && p.cinst) ? p.cinst[p.fcell[i].x] : i0;
        __shared__ double wmin[4];
        const double m0 = block_min_256(mine == i0 ? dtf : 1e5, wmin);
        const double m1 = block_min_256(mine == i0 ? 1e5 : dtf, wmin);
        if (threadIdx.x == 0) {
            p.dt_part[2 * tile] = m0;
            p.dt_part[2 * tile + 1] = m1;
        }
    }
}

// Density residual norm: per block the sum over its cells of (Res_rho/V)^2
// (the residual gathered as in the update), then one block adds the partials
// in index order.  Fixed launch shapes make the sum order, and so the value,
// reproducible.
// The flux of face f (internal or boundary) as the residual gather reads it.
template <typename R> __device__ __forceinline__ double4 face_flux(const DevPtrs<R>& p, int f) {
    if (p.flux32 && f < p.nf) {
        const float4 q = p.flux32[f];
        return double4{(double)q.x, (double)q.y, (double)q.z, (double)q.w};
    }
    return p.flux[f];
}

template <typename R>
__global__ __launch_bounds__(256) void k_residual_part(DevPtrs<R> p) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (c < p.nc) {
        const int C = p.ccnt[c], maxC = p.maxC;
        double r0 = 0.0;
        for (int fi = 0; fi < C; ++fi) {
            const int sf = p.slot_face[sell(c, fi, maxC)];
            const int dir = (sf > 0) - (sf < 0);
            const int f = (sf < 0 ? -sf : sf) - 1;
            r0 += dir * face_flux(p, f).x;
        }
        const double q = r0 / p.vol[c];
        v = q * q;
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) p.res_part[blockIdx.x] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

__global__ __launch_bounds__(1024) void k_residual_sum(const double* part, int n, double* out) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) v += part[i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __shared__ double ws[16];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < 16; ++w) s += ws[w];
        *out = s;
    }
}

// Time step of instance blockIdx.x: minimum of the flux kernel's partials
// over the tiles holding its internal faces (one block per instance).
__global__ __launch_bounds__(1024) void k_dt_reduce(double* dt, const double* part, const int32_t* finst_off,
                                                    double* tctl) {
    const int inst = blockIdx.x;
    const int fa = finst_off[inst], fb = finst_off[inst + 1];
    double v = 1e5;
    if (fb > fa) {
        const int t0 = fa >> 8, t1 = (fb - 1) >> 8;
        for (int t = t0 + (int)threadIdx.x; t <= t1; t += 1024) {
            const double o = part[2 * t + ((t == t0 && (fa & 255)) ? 1 : 0)];
            v = o < v ? o : v;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    __shared__ double wmin[16];
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmin[0];
        for (int w = 1; w < 16; ++w) m = wmin[w] < m ? wmin[w] : m;
        // clip to the end time as the reference's caller does with
        // Foam::Time (dt = min(dt, t_end - t)); 0 once the end is reached
        double* tc = tctl + 4 * inst;
        const double t = tc[0], te = tc[1];
        double d = 0.0;
        if (t < te) {
            const double rem = te - t;
            d = m < rem ? m : rem;
            tc[0] = t + d;
            tc[2] += 1.0;  // steps taken
        }
        dt[inst] = d;
    }
}

// ---------------------------------------------------------------- K8-K11 update

__device__ __forceinline__ double4 boundary_value(int label, double4 w, double4 g, double4 val) {
    // src/CUDA/BCs/2D/*.h, dispatch table cudaEulerBCs.h:34-39; label 6 = per-patch fixed value
    switch (label) {
        case 0: return double4{w.x, 0.0, 0.0, 1.0};
        case 1: return double4{w.x, 0.0, 0.0, w.w};
        case 2: {
            const double nfx = g.x / g.z, nfy = g.y / g.z;
            return double4{w.x, w.y - (w.y * nfx + w.z * nfy) * nfx, w.z - (w.y * nfx + w.z * nfy) * nfy, w.w};
        }
        case 3: return double4{1.0, 3.0, 0.0, 1.0};
        case 4: return w;
        case 5: return double4{30.05945, 4.07795, 0.0, 5.97677};
        default: return val;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_update(DevPtrs<R> p, double Rgas, double gamma, double dt_val,
                                                 const double* dt_ptr, double b1, double b2, double b3, int last) {
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int c = tile * 256 + threadIdx.x;
    if (c >= p.nc) return;
    const int C = p.ccnt[c], maxC = p.maxC, nf = p.nf;
    const double dt = dt_ptr ? dt_ptr[p.cinst ? p.cinst[c] : 0] : dt_val;
    if (dt_ptr && dt == 0.0) return;  // device path past its end time: the state stays
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    for (int fi = 0; fi < C; ++fi) {
        const int sf = p.slot_face[sell(c, fi, maxC)];
        const int dir = (sf > 0) - (sf < 0);
        const int f = (sf < 0 ? -sf : sf) - 1;
        const double4 F = face_flux(p, f);
        r0 += dir * F.x;
        r1 += dir * F.y;
        r2 += dir * F.z;
        r3 += dir * F.w;
    }
    const double V = p.vol[c];
    // b2 = 0 (first stage): U equals Uold at the start of every step (both
    // are stored from one value at the end of a step and at upload), so the
    // stage reads it once; b2 * u is then b2 * uo, the same value
    const double4 uo = p.Uold[c], u = b2 == 0.0 ? uo : p.U[c];
    const double rho = b1 * uo.x + b2 * u.x + b3 * dt * (-r0 / V);
    const double rhoUx = b1 * uo.y + b2 * u.y + b3 * dt * (-r1 / V);
    const double rhoUy = b1 * uo.z + b2 * u.z + b3 * dt * (-r2 / V);
    const double rhoE = b1 * uo.w + b2 * u.w + b3 * dt * (-r3 / V);
    const double4 un = double4{rho, rhoUx, rhoUy, rhoE};
    // The last stage's result is the next step's starting state, which the
    // next first stage reads from Uold alone; U would only be overwritten by
    // that stage, so it is not written here (the host reads the state from
    // Uold after a full step).
    if (last)
        p.Uold[c] = un;
    else
        p.U[c] = un;
    const double pr = (gamma - 1.0) * (rhoE - 0.5 * (rhoUx * rhoUx + rhoUy * rhoUy) / rho);
    const double4 w = double4{pr, rhoUx / rho, rhoUy / rho, pr / (Rgas * rho)};
    typename VT<R>::v4 wr;
    wr.x = (R)w.x;
    wr.y = (R)w.y;
    wr.z = (R)w.z;
    wr.w = (R)w.w;
    p.W[c] = wr;
    // ghost states of this cell's boundary faces (K10), from the updated state
    for (int fi = 0; fi < C; ++fi) {
        const int sf = p.slot_face[sell(c, fi, maxC)];
        const int f = sf - 1;
        if (f < nf) continue;
        const int b = f - nf;
        const double4 g = boundary_value(p.blabel[b], w, p.bgeo[b], p.bval[b]);
        typename VT<R>::v4 gr;
        gr.x = (R)g.x;
        gr.y = (R)g.y;
        gr.z = (R)g.z;
        gr.w = (R)g.w;
        p.W[p.nc + b] = gr;
    }
}

// ---------------------------------------------------------------- launchers

template <typename R>
void launch_lsq(const DevPtrs<R>& p, hipStream_t s) {
    if (p.nc == 0) return;
    hipLaunchKernelGGL(k_lsq<R>, dim3((p.nc + 255) / 256), dim3(256), 0, s, p);
}

static inline int pad8(int n) { return ((n + 7) / 8) * 8; }

template <typename R>
void launch_cfl(const DevPtrs<R>& p, double Rg, double Cv, double CFL, hipStream_t s) {
    const int nt = pad8(std::min(kCflBlocks, std::max(1, (p.nf + 255) / 256)));
    const int chunk = (p.nf + nt - 1) / nt;
    hipLaunchKernelGGL(k_cfl<R>, dim3(nt), dim3(256), 0, s, p, Rg, Cv, CFL, chunk);
}

template <typename R, bool GP>
void launch_reconstruct_t(const DevPtrs<R>& p, int nb, int tile, size_t lds, hipStream_t s, int lanes) {
    if (lanes == 4) {
        switch (tile) {
            case 64: hipLaunchKernelGGL((k_reconstruct_q<R, 64, GP>), dim3(nb), dim3(256), lds, s, p); break;
            case 128: hipLaunchKernelGGL((k_reconstruct_q<R, 128, GP>), dim3(nb), dim3(512), lds, s, p); break;
            default: hipLaunchKernelGGL((k_reconstruct_q<R, 256, GP>), dim3(nb), dim3(1024), lds, s, p); break;
        }
        return;
    }
    switch (tile) {
        case 64: hipLaunchKernelGGL((k_reconstruct<R, 64, GP>), dim3(nb), dim3(64), lds, s, p); break;
        case 128: hipLaunchKernelGGL((k_reconstruct<R, 128, GP>), dim3(nb), dim3(128), lds, s, p); break;
        default: hipLaunchKernelGGL((k_reconstruct<R, 256, GP>), dim3(nb), dim3(256), lds, s, p); break;
    }
}

template <typename R>
void launch_reconstruct(const DevPtrs<R>& p, int tile, size_t lds, hipStream_t s, int lanes) {
    const int nb = pad8(p.ntiles);
    if (p.ngp > 0) {
        launch_reconstruct_t<R, true>(p, nb, tile, lds, s, lanes);
        return;
    }
    launch_reconstruct_t<R, false>(p, nb, tile, lds, s, lanes);
}

template <typename R>
void launch_bvd(const DevPtrs<R>& p, int tile, size_t lds, hipStream_t s) {
    const int nb = pad8(p.ntiles);
    switch (tile) {
        case 64: hipLaunchKernelGGL((k_bvd<R, 64>), dim3(nb), dim3(64), lds, s, p); break;
        case 128: hipLaunchKernelGGL((k_bvd<R, 128>), dim3(nb), dim3(128), lds, s, p); break;
        default: hipLaunchKernelGGL((k_bvd<R, 256>), dim3(nb), dim3(256), lds, s, p); break;
    }
}

int flux_blocks(int nf, int nb) { return pad8(std::max(1, (nf + nb + 255) / 256)); }

template <typename R>
void launch_flux(const DevPtrs<R>& p, double Rg, double Cv, hipStream_t s) {
    if (p.ngp > 0)
        hipLaunchKernelGGL((k_flux<R, false, true>), dim3(flux_blocks(p.nf, p.nb)), dim3(256), 0, s, p, Rg, Cv, 0.0);
    else
        hipLaunchKernelGGL((k_flux<R, false, false>), dim3(flux_blocks(p.nf, p.nb)), dim3(256), 0, s, p, Rg, Cv, 0.0);
}

template <typename R>
void launch_dt_reduce(const DevPtrs<R>& p, hipStream_t s) {
    hipLaunchKernelGGL(k_dt_reduce, dim3(p.ninst), dim3(1024), 0, s, p.dt, (const double*)p.dt_part, p.finst_off,
                       p.tctl);
}

int residual_blocks(int nc) { return std::max(1, (nc + 255) / 256); }

template <typename R>
void launch_residual_norm(const DevPtrs<R>& p, double* out, hipStream_t s) {
    const int nb = residual_blocks(p.nc);
    hipLaunchKernelGGL(k_residual_part<R>, dim3(nb), dim3(256), 0, s, p);
    hipLaunchKernelGGL(k_residual_sum, dim3(1), dim3(1024), 0, s, (const double*)p.res_part, nb, out);
}

template <typename R>
void launch_flux_dt(const DevPtrs<R>& p, double Rg, double Cv, double CFL, hipStream_t s) {
    if (p.ngp > 0)
        hipLaunchKernelGGL((k_flux<R, true, true>), dim3(flux_blocks(p.nf, p.nb)), dim3(256), 0, s, p, Rg, Cv, CFL);
    else
        hipLaunchKernelGGL((k_flux<R, true, false>), dim3(flux_blocks(p.nf, p.nb)), dim3(256), 0, s, p, Rg, Cv, CFL);
}

template <typename R>
void launch_update(const DevPtrs<R>& p, double Rg, double gamma, double dt, const double* dt_ptr, double b1,
                   double b2, double b3, int last, hipStream_t s) {
    const int nt = pad8(std::max(1, (p.nc + 255) / 256));
    hipLaunchKernelGGL(k_update<R>, dim3(nt), dim3(256), 0, s, p, Rg, gamma, dt, dt_ptr, b1, b2, b3, last);
}

template <typename R>
hipError_t set_lds_limits(size_t lds) {
    hipError_t e = hipSuccess;
    const void* fns[] = {(const void*)k_reconstruct<R, 64, false>,    (const void*)k_reconstruct<R, 128, false>,
                         (const void*)k_reconstruct<R, 256, false>,   (const void*)k_reconstruct_q<R, 64, false>,
                         (const void*)k_reconstruct_q<R, 128, false>, (const void*)k_reconstruct_q<R, 256, false>,
                         (const void*)k_reconstruct<R, 64, true>,     (const void*)k_reconstruct<R, 128, true>,
                         (const void*)k_reconstruct<R, 256, true>,    (const void*)k_reconstruct_q<R, 64, true>,
                         (const void*)k_reconstruct_q<R, 128, true>,  (const void*)k_reconstruct_q<R, 256, true>};
    for (const void* fn : fns) {
        // the dynamic limit plus the kernel's static LDS may not exceed the block limit
        hipFuncAttributes attr;
        e = hipFuncGetAttributes(&attr, fn);
        if (e != hipSuccess) return e;
        const size_t dyn = lds > attr.sharedSizeBytes ? lds - attr.sharedSizeBytes : 0;
        e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
    }
    return e;
}

#define AFAR_INSTANTIATE(R)                                                                                        \
    template void launch_lsq<R>(const DevPtrs<R>&, hipStream_t);                                                  \
    template void launch_cfl<R>(const DevPtrs<R>&, double, double, double, hipStream_t);                          \
    template void launch_reconstruct<R>(const DevPtrs<R>&, int, size_t, hipStream_t, int);                        \
    template void launch_bvd<R>(const DevPtrs<R>&, int, size_t, hipStream_t);                                     \
    template void launch_flux<R>(const DevPtrs<R>&, double, double, hipStream_t);                                 \
    template void launch_flux_dt<R>(const DevPtrs<R>&, double, double, double, hipStream_t);                      \
    template void launch_dt_reduce<R>(const DevPtrs<R>&, hipStream_t);                                            \
    template void launch_residual_norm<R>(const DevPtrs<R>&, double*, hipStream_t);                               \
    template void launch_update<R>(const DevPtrs<R>&, double, double, double, const double*, double, double,      \
                                   double, int, hipStream_t);                                                     \
    template hipError_t set_lds_limits<R>(size_t);

AFAR_INSTANTIATE(double)
AFAR_INSTANTIATE(float)

// Self-test entry: evaluates the shared elementary functions on the device.
__global__ void k_math_selftest(const double* x, double* out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    out[4 * i] = afar_exp(v);
    out[4 * i + 1] = afar_erfc(v);
    out[4 * i + 2] = sqrt(v < 0 ? -v : v);
    out[4 * i + 3] = 1.0 / v;
}

__global__ void k_math_selftest_f32(const float* x, float* out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = x[i];
    out[4 * i] = afar_expf(v);
    out[4 * i + 1] = afar_erfcf(v);
    out[4 * i + 2] = sqrtf(v < 0 ? -v : v);
    out[4 * i + 3] = 1.0f / v;
}

void launch_math_selftest(const double* x, double* out, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_math_selftest, dim3((n + 255) / 256), dim3(256), 0, s, x, out, n);
}

void launch_math_selftest_f32(const float* x, float* out, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_math_selftest_f32, dim3((n + 255) / 256), dim3(256), 0, s, x, out, n);
}

}  // namespace afar
