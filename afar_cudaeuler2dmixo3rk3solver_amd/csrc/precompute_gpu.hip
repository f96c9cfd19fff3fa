// RBF precompute on the GPU (basicCUDAsolver.cpp:461-619): one wavefront per
// cell assembles the moment matrix from the members' quadrature, inverts it
// by the Jacobi SVD and writes the coefficients of the cell's side of each of
// its internal faces.  The arithmetic is rbf_core.h, shared with the host
// precompute, so the coefficients equal the host's bit for bit.  Stencils,
// quadrature tables and geometry come from the host; the work matrices (three
// of at most 21 x 21 doubles per cell) live in LDS.
#include <hip/hip_runtime.h>

#include <vector>

#include "afar_host.h"
#include "rbf_core.h"
#include "solver.h"

namespace afar {

namespace {

struct RbfArgs {
    int nc, nf, W;  // W = maxS + 1
    double sf;
    const uint8_t* S;
    const uint8_t* Cc;
    const int32_t* sten;        // nc * kMaxValidStencil
    const double2* sten_shift;  // nc * kMaxValidStencil
    const int32_t* cellFaces;   // nc * kMaxValidStencil
    const double2* C;           // nc + nb centres
    const double* tp;           // quadrature points, 2 per entry
    const double* tw;           // quadrature weights
    const int64_t* tptr;        // nc + nb + 1
    const double2* Cf;          // nf face centres
    const double2* fshift;      // nf
    double* rbf;                // 2 * W * nf
    uint8_t* trouble;           // nc
    double* ref;                // nc
    // flux quadrature points (extension): ngp points at xi[g] of each face's
    // end points fp0, fp1 (owner frame); rbf_gp: ngp * 2 * W * nf
    int ngp;
    double xi[4];
    const double2* fp0;
    const double2* fp1;
    double* rbf_gp;
};

// One wavefront per cell (a block of 64 threads), its work matrices in LDS.
// Every value is formed by one lane with the host's operations in the host's
// order (rbf_core.h); the lanes share the independent parts: the moment
// entries, the disjoint rotations of a round-robin round and their column
// updates, the pseudo-inverse entries, and per face the basis values and
// coefficients.
__global__ __launch_bounds__(64) void k_rbf_precompute(RbfArgs a) {
    constexpr int N = kRbfMaxN;
    __shared__ double sU[N * N], sV[N * N], sG[N * N];
    __shared__ double sX[2 * N], sShx[N], sShy[N], sSig[N], sB[N], sOut[N], sCs[2 * (N + 1) / 2 * 2];
    __shared__ int sId[N], sCount[N], sDo[(N + 1) / 2];
    __shared__ double sRef, sSum;
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    const int S = a.S[c], n = S + 1, nn = n * n;
    if (lane < n) {
        int id = c;
        double shx = 0.0, shy = 0.0;
        if (lane > 0) {
            const size_t s = (size_t)c * kMaxValidStencil + lane - 1;
            id = a.sten[s];
            shx = a.sten_shift[s].x;
            shy = a.sten_shift[s].y;
        }
        sId[lane] = id;
        sShx[lane] = shx;
        sShy[lane] = shy;
        sX[2 * lane] = lane > 0 ? a.C[id].x + shx : a.C[c].x;
        sX[2 * lane + 1] = lane > 0 ? a.C[id].y + shy : a.C[c].y;
        sCount[lane] = (int)(a.tptr[id + 1] - a.tptr[id]);
    }
    __syncthreads();
    if (lane == 0) {  // mean distance to the members, summed in order
        double ref = 0.0;
        for (int i = 1; i < n; ++i) {
            const double refdx = sX[2 * i] - a.C[c].x, refdy = sX[2 * i + 1] - a.C[c].y;
            ref += sqrt(refdx * refdx + refdy * refdy);
        }
        ref /= (double)S;
        sRef = ref;
        a.ref[c] = ref;
    }
    __syncthreads();
    const double ref = sRef, sf = a.sf;
    // moment entries P_ij, stored transposed (U[j*n + i] = P_ij) as rbf_pinv works on them
    for (int e = lane; e < nn; e += 64) {
        const int i = e / n, j = e - i * n;
        const double x0 = sX[2 * j], y0 = sX[2 * j + 1];
        const int64_t e0 = a.tptr[sId[i]];
        double acc = 0.0;
        for (int k = 0; k < sCount[i]; ++k) {
            const double gx = i > 0 ? a.tp[2 * (e0 + k)] + sShx[i] : a.tp[2 * (e0 + k)];
            const double gy = i > 0 ? a.tp[2 * (e0 + k) + 1] + sShy[i] : a.tp[2 * (e0 + k) + 1];
            const double dx = (gx - x0) / ref, dy = (gy - y0) / ref;
            const double r2 = dx * dx + dy * dy;
            acc += a.tw[e0 + k] * sqrt(sf + r2);
        }
        sU[j * n + i] = acc;
        sV[e] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    // Jacobi sweeps in the round-robin order: lane i < m/2 finds pair i's
    // rotation, then all lanes apply the round's rotations to U and V columns.
    const int m = n + (n & 1), half = m / 2;
    for (int sweep = 0; sweep < 80; ++sweep) {
        bool rotated = false;
        for (int r = 0; r < m - 1; ++r) {
            if (lane < half) {
                int p, q;
                rbf_rr_pair(m, r, lane, p, q);
                double cs = 0.0, sn = 0.0;
                const bool d = q < n && rbf_rotation(n, &sU[p * n], &sU[q * n], cs, sn);
                sDo[lane] = d;
                sCs[2 * lane] = cs;
                sCs[2 * lane + 1] = sn;
            }
            __syncthreads();
            bool any = false;
            for (int i = 0; i < half; ++i) any = any || sDo[i];
            rotated = rotated || any;
            for (int e = lane; e < half * 2 * n; e += 64) {
                const int i = e / (2 * n), rem = e - i * 2 * n;
                if (!sDo[i]) continue;
                int p, q;
                rbf_rr_pair(m, r, i, p, q);
                double* M = rem < n ? sU : sV;
                const int k = rem < n ? rem : rem - n;
                rbf_rotate(M[p * n + k], M[q * n + k], sCs[2 * i], sCs[2 * i + 1]);
            }
            __syncthreads();
        }
        if (!rotated) break;
    }
    if (lane < n) {
        double s2 = 0.0;
        for (int k = 0; k < n; ++k) s2 += sU[lane * n + k] * sU[lane * n + k];
        sSig[lane] = sqrt(s2);
    }
    __syncthreads();
    bool bad = false;
    for (int e = lane; e < nn; e += 64) {
        const int i = e / n, j = e - i * n;
        const double g = rbf_pinv_entry(n, sU, sV, sSig, i, j);
        sG[e] = g;
        bad = bad || fabs(g) > 1e10 || g != g;
    }
    const bool trouble = __any(bad);
    if (lane == 0) a.trouble[c] = trouble;
    __syncthreads();
    // coefficients of the cell's side of each of its internal faces
    auto face = [&](double xf, double yf, double* o) {
        if (lane < n) {
            const double dx = (xf - sX[2 * lane]) / ref, dy = (yf - sX[2 * lane + 1]) / ref;
            sB[lane] = sqrt(sf + (dx * dx + dy * dy));
        }
        __syncthreads();
        if (lane < n) {
            double v = 0.0;
            for (int j = 0; j < n; ++j) v += sG[j * n + lane] * sB[j];
            sOut[lane] = v;
        }
        __syncthreads();
        if (lane == 0) {
            double sum = 0.0;
            for (int i = 0; i < n; ++i) sum += sOut[i];
            sSum = sum;
        }
        __syncthreads();
        if (lane < n) o[lane] = (sOut[lane] - sSum / (double)n) + 1.0 / (double)n;
        __syncthreads();
    };
    for (int fi = 0; fi < a.Cc[c]; ++fi) {
        const int sfc = a.cellFaces[(size_t)c * kMaxValidStencil + fi];
        const int f = (sfc > 0 ? sfc : -sfc) - 1;
        if (f >= a.nf) continue;
        const int side = sfc < 0;
        double* o = a.rbf + (size_t)2 * a.W * f + (size_t)side * a.W;
        if (trouble) {
            if (lane < n) o[lane] = lane == 0 ? 1.0 : 0.0;
            for (int g = 0; g < a.ngp; ++g) {
                double* og = a.rbf_gp + ((size_t)g * a.nf + f) * 2 * a.W + (size_t)side * a.W;
                if (lane < n) og[lane] = lane == 0 ? 1.0 : 0.0;
            }
            continue;
        }
        double xf = a.Cf[f].x, yf = a.Cf[f].y;
        if (side) {
            xf = xf - a.fshift[f].x;
            yf = yf - a.fshift[f].y;
        }
        face(xf, yf, o);
        for (int g = 0; g < a.ngp; ++g) {  // the host's face_gauss_point, same operations
            const double t = a.xi[g];
            double xg = 0.5 * a.fp0[f].x * (1 - t) + 0.5 * a.fp1[f].x * (1 + t);
            double yg = 0.5 * a.fp0[f].y * (1 - t) + 0.5 * a.fp1[f].y * (1 + t);
            if (side) {
                xg = xg - a.fshift[f].x;
                yg = yg - a.fshift[f].y;
            }
            face(xg, yg, a.rbf_gp + ((size_t)g * a.nf + f) * 2 * a.W + (size_t)side * a.W);
        }
    }
}

// The precompute's own stream, destroyed after the buffers below (declared
// after it) have enqueued their frees.
struct OwnStream {
    hipStream_t s = nullptr;
    OwnStream() { check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"); }
    ~OwnStream() {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
};

// Device buffer allocated, filled and freed in stream order (a plain hipFree
// would wait for every other solver's work on the device).
struct Buf {
    void* p = nullptr;
    hipStream_t s = nullptr;
    explicit Buf(hipStream_t st) : s(st) {}
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() {
        if (p) (void)hipFreeAsync(p, s);
    }
    // the host source must stay alive until the stream is synchronised
    template <typename T> T* put(const T* src, size_t n) {
        check_hip(hipMallocAsync(&p, sizeof(T) * (n ? n : 1), s), "hipMallocAsync");
        if (n && src) check_hip(hipMemcpyAsync(p, src, sizeof(T) * n, hipMemcpyHostToDevice, s), "upload");
        return static_cast<T*>(p);
    }
};

}  // namespace

void precompute_rbf_gpu(const Mesh& m, const Schemes& sc, Precomputed& pc) {
    const Stencils& st = pc.st;
    const int nc = m.nc, nf = m.nf, nb = m.nb, W = st.maxS + 1;
    if (W > kRbfMaxN) throw Error(3, "stencil larger than the RBF work arrays");
    std::vector<double> tp, tw;
    std::vector<int64_t> tptr;
    gauss_tables(m, sc, tp, tw, tptr);
    std::vector<double2> C(nc + nb), Cf(nf), fsh(nf), shift((size_t)nc * kMaxValidStencil), p0(nf), p1(nf);
    for (int i = 0; i < nc + nb; ++i) C[i] = double2{m.C[i].x, m.C[i].y};
    for (int f = 0; f < nf; ++f) {
        Cf[f] = double2{m.Cf[f].x, m.Cf[f].y};
        fsh[f] = double2{m.fshift[f].x, m.fshift[f].y};
        p0[f] = double2{m.fp0[f].x, m.fp0[f].y};
        p1[f] = double2{m.fp1[f].x, m.fp1[f].y};
    }
    if (pc.ngp > 4) throw Error(3, "more than four flux quadrature points");
    for (size_t i = 0; i < shift.size(); ++i) shift[i] = double2{st.sten_shift[i].x, st.sten_shift[i].y};
    // own stream: waits for this work only, not for other solvers on the device
    OwnStream os;
    hipStream_t s = os.s;
    Buf bS(s), bCc(s), bSten(s), bShift(s), bCF(s), bC(s), bTp(s), bTw(s), bTptr(s), bCf(s), bFsh(s), bRbf(s),
        bTr(s), bRef(s), bP0(s), bP1(s), bGp(s);
    RbfArgs a{};
    a.nc = nc;
    a.nf = nf;
    a.W = W;
    a.sf = sc.shapeFactor;
    a.S = bS.put(st.S.data(), st.S.size());
    a.Cc = bCc.put(st.Cc.data(), st.Cc.size());
    a.sten = bSten.put(st.sten.data(), st.sten.size());
    a.sten_shift = bShift.put(shift.data(), shift.size());
    a.cellFaces = bCF.put(st.cellFaces.data(), st.cellFaces.size());
    a.C = bC.put(C.data(), C.size());
    a.tp = bTp.put(tp.data(), tp.size());
    a.tw = bTw.put(tw.data(), tw.size());
    a.tptr = bTptr.put(tptr.data(), tptr.size());
    a.Cf = bCf.put(Cf.data(), Cf.size());
    a.fshift = bFsh.put(fsh.data(), fsh.size());
    a.rbf = bRbf.put<double>(nullptr, (size_t)2 * W * nf);
    a.trouble = bTr.put<uint8_t>(nullptr, nc);
    a.ref = bRef.put<double>(nullptr, nc);
    a.ngp = pc.ngp;
    for (int g = 0; g < pc.ngp; ++g) a.xi[g] = pc.gp_xi[g];
    a.fp0 = pc.ngp ? bP0.put(p0.data(), p0.size()) : nullptr;
    a.fp1 = pc.ngp ? bP1.put(p1.data(), p1.size()) : nullptr;
    a.rbf_gp = pc.ngp ? bGp.put<double>(nullptr, (size_t)pc.ngp * 2 * W * nf) : nullptr;
    check_hip(hipMemsetAsync(a.rbf, 0, sizeof(double) * 2 * W * (size_t)nf, s), "hipMemsetAsync");
    if (pc.ngp)
        check_hip(hipMemsetAsync(a.rbf_gp, 0, sizeof(double) * pc.ngp * 2 * W * (size_t)nf, s), "hipMemsetAsync");
    hipLaunchKernelGGL(k_rbf_precompute, dim3(nc), dim3(64), 0, s, a);
    check_hip(hipGetLastError(), "k_rbf_precompute");
    pc.rbf.assign((size_t)2 * W * nf, 0.0);
    pc.trouble.assign(nc, 0);
    pc.refDistance.assign(nc, 0.0);
    check_hip(hipMemcpyAsync(pc.rbf.data(), a.rbf, sizeof(double) * pc.rbf.size(), hipMemcpyDeviceToHost, s),
              "download");
    check_hip(hipMemcpyAsync(pc.trouble.data(), a.trouble, nc, hipMemcpyDeviceToHost, s), "download");
    check_hip(hipMemcpyAsync(pc.refDistance.data(), a.ref, sizeof(double) * nc, hipMemcpyDeviceToHost, s),
              "download");
    pc.rbf_gp.assign((size_t)pc.ngp * 2 * W * nf, 0.0);
    if (pc.ngp)
        check_hip(hipMemcpyAsync(pc.rbf_gp.data(), a.rbf_gp, sizeof(double) * pc.rbf_gp.size(), hipMemcpyDeviceToHost,
                                 s),
                  "download");
    check_hip(hipStreamSynchronize(s), "k_rbf_precompute");
    int t = 0;
    for (uint8_t v : pc.trouble) t += v;
    pc.troubleCount = t;
}

}  // namespace afar
