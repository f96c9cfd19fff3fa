"""GPU: afar_adjust_time_step_noalloc against afar_adjust_time_step.

Three calls in a row, each on a smaller mesh than the one before: 400^2 mixed
triangles/quads (more faces than the 1024 partials hold at 256 faces per
block, so the blocks loop), 60^2, then a periodic 12^2 triangle mesh with
its neighbour shifts.  Each time step equals the seam-2 entry point's (which
the library computes on its own device arrays) and the oracle's bit for bit.
The device arrays are the caller's: this test allocates them; the entry
point allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def hip():
    L = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


def to_device(L, a, keep):
    a = np.ascontiguousarray(a)
    p = C.c_void_p()
    assert L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
    assert L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
    keep.append(p)
    return p


def test_noalloc_time_step_equals_seam2(afar):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    from afar_cudaeuler2dmixo3rk3solver_amd._lib import check, lib
    from afar_cudaeuler2dmixo3rk3solver_amd.solver import DeviceFields
    L = hip()
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    meshes = []
    for n in (400, 60):
        m, W0 = cases.riemann_case(n=n, seed=1)
        meshes.append(("riemann %d" % n, m, W0, False))
    m, W0 = cases.vortex_case(n=12)
    meshes.append(("periodic vortex 12", m, W0, True))
    for label, m, W0, periodic in meshes:
        print("case: %s, noalloc time step vs seam 2 and the oracle" % label, flush=True)
        m.precompute(sc)
        a = m.reference_arrays(sc, W0)
        s = a["sizes"]
        dev = DeviceFields(a)
        dt_seam2 = dev.adjust_time_step(sc.R, sc.Cv, sc.maxCo)
        dt_oracle = O.OracleState(a, sc.R, sc.Cv).adjust_time_step(sc.maxCo)
        keep = []
        try:
            ptrs = [to_device(L, a[k], keep) for k in ("FD", "CELL", "FACE", "neighbour")]
            shift = to_device(L, a["faceShift"], keep) if periodic else None
            out = to_device(L, np.full(1, -1.0), keep)
            check(lib().afar_adjust_time_step_noalloc(*ptrs, shift, s["nf"], sc.R, sc.Cv, sc.maxCo, out, None))
            assert L.hipDeviceSynchronize() == 0
            dt = np.zeros(1)
            assert L.hipMemcpy(dt.ctypes.data, out, 8, 2) == 0  # device to host
        finally:
            for p in keep:
                L.hipFree(p)
        print("  dt noalloc %.17g, seam 2 %.17g, oracle %.17g" % (dt[0], dt_seam2, dt_oracle), flush=True)
        assert dt[0] == dt_seam2 == dt_oracle
    # refused sizes: nothing launched
    rc = lib().afar_adjust_time_step_noalloc(None, None, None, None, None, 10, sc.R, sc.Cv, 0.5, None, None)
    assert rc == 1
