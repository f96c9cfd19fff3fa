"""ORACLE wrapper -- test infrastructure only (tests/, __graft_entry__.smoke(),
bench.py's CPU baseline).  ctypes binding of liboracle.so, the FP64 host
restatement of the reference kernels (afar_oracle.cpp) and the FP32-mixed
restatement of the library's mixed path (afar_oracle_mixed.cpp).

``OracleState`` owns reference-layout arrays (as exported by
Mesh.reference_arrays) plus the device-only work arrays of the reference
(matrix, gradFD, limiter, faceFD, Flux, Res).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
P = C.POINTER


class OracleFields(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nc", "nf", "nb", "maxS", "maxC")] + [
        ("CELL", P(C.c_double)), ("FACE", P(C.c_double)), ("cellVolume", P(C.c_double)), ("FD", P(C.c_double)),
        ("cFD", P(C.c_double)), ("cFDold", P(C.c_double)), ("neighbour", P(C.c_int32)),
        ("boundaryFaceNeiLabel", P(C.c_int32)), ("boundaryFacesType", P(C.c_uint8)), ("stencilSize", P(C.c_uint8)),
        ("compactStencilSize", P(C.c_uint8)), ("stencil", P(C.c_int32)), ("cellFaces", P(C.c_int32)),
        ("RBFbasis", P(C.c_double)), ("FSHIFT", P(C.c_double)), ("BVALUE", P(C.c_double)),
        ("matrix", P(C.c_double)), ("gradFD", P(C.c_double)), ("limiter", P(C.c_double)),
        ("shockIndicator", P(C.c_int8)), ("faceFD", P(C.c_double)), ("Flux", P(C.c_double)), ("Res", P(C.c_double)),
        ("boundaryFirstOrder", C.c_int32),
        ("ngp", C.c_int32), ("gpWeights", P(C.c_double)), ("gpOffsets", P(C.c_double)), ("RBFgp", P(C.c_double)),
        ("faceGP", P(C.c_double)),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(path):
    """The oracle library at `path` with its ctypes signatures."""
    L = C.CDLL(path)
    pf = P(OracleFields)
    d = C.c_double
    L.oracle_least_square_matrix.argtypes = [pf]
    L.oracle_adjust_time_step.argtypes = [pf, d, d, d]
    L.oracle_adjust_time_step.restype = d
    L.oracle_reconstruct.argtypes = [pf]
    L.oracle_bvd.argtypes = [pf]
    L.oracle_evaluate_flux.argtypes = [pf, d, d]
    L.oracle_evaluate_residual.argtypes = [pf]
    L.oracle_update.argtypes = [pf, d, d, d, d, d, d]
    L.oracle_update_boundary.argtypes = [pf, d, d]
    L.oracle_store_conserved.argtypes = [pf]
    L.oracle_evolve.argtypes = [pf, d, d, d, C.c_int]
    L.oracle_run.argtypes = [pf, d, d, d, C.c_int, C.c_int, P(d)]
    L.oracle_roe_flux.argtypes = [P(d)] + [d] * 17
    L.oracle_kfvs_flux.argtypes = [P(d)] + [d] * 15
    L.oracle_central_flux.argtypes = [P(d)] + [d] * 13
    L.oracle_exp.argtypes = [d]
    L.oracle_exp.restype = d
    L.oracle_erfc.argtypes = [d]
    L.oracle_erfc.restype = d
    if hasattr(L, "oracle_mixed_evolve"):  # absent from single-file builds (the mutation tests')
        L.oracle_mixed_adjust_time_step.argtypes = [pf, d, d, d]
        L.oracle_mixed_adjust_time_step.restype = d
        L.oracle_mixed_reconstruct.argtypes = [pf]
        L.oracle_mixed_bvd.argtypes = [pf]
        L.oracle_mixed_evaluate_flux.argtypes = [pf, d, d]
        L.oracle_mixed_evolve.argtypes = [pf, d, d, d, C.c_int]
        f = C.c_float
        L.oracle_mixed_central_flux.argtypes = [P(f)] + [f] * 10 + [d, f, f]
        L.oracle_mixed_roe_flux.argtypes = [P(f)] + [f] * 12 + [d, d, f, f, f]
        L.oracle_mixed_kfvs_flux.argtypes = [P(f)] + [f] * 10 + [d, d, f, f, f]
        for name in ("oracle_expf", "oracle_erfcf"):
            getattr(L, name).argtypes = [f]
            getattr(L, name).restype = f
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(P(t))


class OracleState:
    """Reference-layout state advanced by the oracle kernels: the FP64
    restatement, or with ``mixed`` the FP32-mixed one."""

    RK_LITERAL = ((1.0, 0.75, 0.333333), (0.0, 0.25, 0.666667), (1.0, 0.25, 0.666667))

    def __init__(self, ref, R, Cv, boundary_first_order=False, mixed=False):
        self.mixed = bool(mixed)
        s = ref["sizes"]
        nc, nf, nb, maxS, maxC = s["nc"], s["nf"], s["nb"], s["maxS"], s["maxC"]
        self.R, self.Cv = R, Cv
        self.s = s
        self.a = {k: np.ascontiguousarray(v).copy() for k, v in ref.items() if isinstance(v, np.ndarray)}
        a = self.a
        a["cFDold"] = a["cFD"].copy()
        a["matrix"] = np.zeros((nc, maxC, 2))
        a["gradFD"] = np.zeros((nc, 8))
        a["limiter"] = np.zeros((nc, 4))
        a["shockIndicator"] = np.zeros(2 * (nc + nb), np.int8)
        a["faceFD"] = np.zeros((nf, 2, 4))
        a["Flux"] = np.zeros((nf + nb, 4))
        a["Res"] = np.zeros((nc, 4))
        f = OracleFields()
        f.nc, f.nf, f.nb, f.maxS, f.maxC = nc, nf, nb, maxS, maxC
        for name in ("CELL", "FACE", "cellVolume", "FD", "cFD", "cFDold", "matrix", "gradFD", "limiter", "faceFD",
                     "Flux", "Res"):
            setattr(f, name, _p(a[name]))
        f.neighbour = _p(a["neighbour"], C.c_int32)
        f.boundaryFaceNeiLabel = _p(a["boundaryFaceNeiLabel"], C.c_int32)
        f.boundaryFacesType = _p(a["boundaryFacesType"], C.c_uint8)
        f.stencilSize = _p(a["stencilSize"], C.c_uint8)
        f.compactStencilSize = _p(a["compactStencilSize"], C.c_uint8)
        f.stencil = _p(a["globalStencil"], C.c_int32)
        f.cellFaces = _p(a["cellFaces"], C.c_int32)
        f.RBFbasis = _p(a["RBFbasis64"])
        f.FSHIFT = _p(a["faceShift"])
        f.BVALUE = _p(a["boundaryValue"])
        f.shockIndicator = _p(a["shockIndicator"], C.c_int8)
        f.boundaryFirstOrder = int(bool(boundary_first_order))
        if "RBFgp" in a:  # flux quadrature points (library extension)
            ngp = a["RBFgp"].shape[0]
            a["faceGP"] = np.zeros((ngp, nf, 2, 4))
            f.ngp = ngp
            f.gpWeights = _p(a["gpWeights"])
            f.gpOffsets = _p(a["gpOffsets"])
            f.RBFgp = _p(a["RBFgp"])
            f.faceGP = _p(a["faceGP"])
        self.f = f
        lib().oracle_least_square_matrix(C.byref(f))

    def _fn(self, name):
        return getattr(lib(), ("oracle_mixed_" if self.mixed else "oracle_") + name)

    def adjust_time_step(self, CFL):
        return self._fn("adjust_time_step")(C.byref(self.f), self.R, self.Cv, CFL)

    def evolve(self, dt, exact_rk=False):
        self._fn("evolve")(C.byref(self.f), self.R, self.Cv, dt, int(exact_rk))

    def run(self, nsteps, CFL, exact_rk=False):
        dts = np.zeros(nsteps)
        if not self.mixed:
            lib().oracle_run(C.byref(self.f), self.R, self.Cv, CFL, nsteps, int(exact_rk), _p(dts))
            return dts
        for i in range(nsteps):  # the same loop as oracle_run
            dts[i] = self.adjust_time_step(CFL)
            self.evolve(dts[i], exact_rk)
        return dts

    # single kernels, for kernel-level checks
    def reconstruct(self):
        self._fn("reconstruct")(C.byref(self.f))

    def bvd(self):
        self._fn("bvd")(C.byref(self.f))

    def flux(self):
        self._fn("evaluate_flux")(C.byref(self.f), self.R, self.Cv)

    def residual_update(self, dt, stage=0):
        lib().oracle_evaluate_residual(C.byref(self.f))
        b1, b2, b3 = (c[stage] for c in self.RK_LITERAL)
        lib().oracle_update(C.byref(self.f), self.R, self.R / self.Cv + 1.0, dt, b1, b2, b3)
        lib().oracle_update_boundary(C.byref(self.f), self.R, self.Cv)

    @property
    def W(self):
        return self.a["FD"][: self.s["nc"]]

    @property
    def U(self):
        return self.a["cFD"]

    @property
    def smooth_flags(self):
        return self.a["shockIndicator"][0: 2 * self.s["nc"]: 2]

    @property
    def final_types(self):
        return self.a["shockIndicator"][1: 2 * self.s["nc"]: 2]


def roe_flux(rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, dP, dRho, R, Cv, magSf, Sfx, Sfy, base=None):
    out = np.zeros(4) if base is None else np.array(base, dtype=np.float64)
    lib().oracle_roe_flux(_p(out), rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, dP, dRho, R, Cv, magSf, Sfx, Sfy)
    return out


def kfvs_flux(rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, R, Cv, magSf, Sfx, Sfy, base=None):
    out = np.zeros(4) if base is None else np.array(base, dtype=np.float64)
    lib().oracle_kfvs_flux(_p(out), rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, R, Cv, magSf, Sfx, Sfy)
    return out


def central_flux(rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, Cv, Sfx, Sfy):
    out = np.zeros(4)
    lib().oracle_central_flux(_p(out), rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, Cv, Sfx, Sfy)
    return out


# The same three in float (the FP32-mixed oracle); float arguments are rounded
# to float by ctypes, R and Cv stay double.
def mixed_central_flux(rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, Cv, Sfx, Sfy):
    out = np.zeros(4, np.float32)
    lib().oracle_mixed_central_flux(_p(out, C.c_float), rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, Cv, Sfx, Sfy)
    return out


def mixed_roe_flux(rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, dP, dRho, R, Cv, magSf, Sfx, Sfy, base):
    out = np.array(base, dtype=np.float32)
    lib().oracle_mixed_roe_flux(_p(out, C.c_float), rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, dP, dRho, R, Cv,
                                magSf, Sfx, Sfy)
    return out


def mixed_kfvs_flux(rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, R, Cv, magSf, Sfx, Sfy, base):
    out = np.array(base, dtype=np.float32)
    lib().oracle_mixed_kfvs_flux(_p(out, C.c_float), rhoL, rhoR, pL, pR, uxL, uxR, uyL, uyR, TL, TR, R, Cv, magSf, Sfx,
                                 Sfy)
    return out
