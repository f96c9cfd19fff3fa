"""GPU: solvers on the caller's stream (afar_solver_create_on_stream,
afar_device_set_stream).  The results equal those of a solver on a stream
of its own bit for bit, the caller can order its own work after the
solver's on that stream, and the library leaves the stream alive.  The
stream comes from the HIP runtime the library itself links (a process holds
one HIP runtime; torch's wheel carries another)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def hip():
    """The HIP runtime libafar.so is linked against (already loaded)."""
    L = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    L.hipStreamDestroy.argtypes = [C.c_void_p]
    L.hipStreamSynchronize.argtypes = [C.c_void_p]
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    return L


def stream_round_trip(L, stream, values):
    """Host -> device -> host copies ordered on `stream`; returns the copy."""
    buf = C.c_void_p()
    assert L.hipMalloc(C.byref(buf), values.nbytes) == 0
    out = np.zeros_like(values)
    assert L.hipMemcpyAsync(buf, values.ctypes.data, values.nbytes, 1, stream) == 0  # hipMemcpyHostToDevice
    assert L.hipMemcpyAsync(out.ctypes.data, buf, values.nbytes, 2, stream) == 0  # hipMemcpyDeviceToHost
    assert L.hipStreamSynchronize(stream) == 0
    assert L.hipFree(buf) == 0
    return out


def _case(afar):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    mesh, W0 = cases.corner_case(nx=60, ny=30)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    mesh.precompute(sc)
    return mesh, W0, sc


def test_solver_on_caller_stream(afar):
    from afar_cudaeuler2dmixo3rk3solver_amd._lib import check, dptr, lib
    print("case: corner 60x30 on a caller's stream vs its own stream, 40 device-dt steps", flush=True)
    mesh, W0, sc = _case(afar)
    own = afar.BasicCUDASolver(mesh, sc, W0)
    own.evolve_auto(40)
    ref = own.conserved()
    L = hip()
    stream = C.c_void_p()
    assert L.hipStreamCreate(C.byref(stream)) == 0
    h = C.c_void_p()
    w0 = np.ascontiguousarray(W0)
    check(lib().afar_solver_create_on_stream(mesh.handle, C.byref(sc), dptr(w0), stream, C.byref(h)))
    check(lib().afar_solver_evolve_auto(h, 40))
    # the caller's own work, ordered after the solver's steps on the same stream
    v = np.arange(256, dtype=np.float64)
    assert np.array_equal(stream_round_trip(L, stream, v), v)
    U = np.zeros((mesh.nc, 4))
    check(lib().afar_solver_get_conserved(h, dptr(U)))
    assert np.array_equal(U, ref)
    lib().afar_solver_destroy(h)
    # the stream survives the solver
    assert np.array_equal(stream_round_trip(L, stream, 2 * v), 2 * v)
    assert L.hipStreamDestroy(stream) == 0


def test_seam2_on_caller_stream(afar):
    from afar_cudaeuler2dmixo3rk3solver_amd._lib import check, lib
    from afar_cudaeuler2dmixo3rk3solver_amd.solver import DeviceFields, field_pointer
    print("case: seam 2 on a caller's stream vs its own stream, 10 host-dt steps", flush=True)
    mesh, W0, sc = _case(afar)
    a = mesh.reference_arrays(sc, W0)
    b = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in a.items()}
    ref = DeviceFields(a)
    s = b["sizes"]
    L = hip()
    stream = C.c_void_p()
    assert L.hipStreamCreate(C.byref(stream)) == 0
    fp = field_pointer(b)
    h = C.c_void_p()
    check(lib().afar_set_device_field_data(C.byref(h), C.byref(fp), s["nc"], s["nf"], s["nb"]))
    check(lib().afar_device_set_stream(h, stream))
    check(lib().afar_set_device_stencil_data(h, C.byref(fp), s["nc"], s["nf"], s["npatches"], s["nb"], s["maxS"],
                                             s["maxC"], s["maxLB"], s["maxCLB"]))
    for _ in range(10):
        dt0 = ref.adjust_time_step(sc.R, sc.Cv, sc.maxCo)
        dt1 = C.c_double()
        check(lib().afar_adjust_time_step(h, s["nf"], sc.R, sc.Cv, sc.maxCo, C.byref(dt1)))
        assert dt0 == dt1.value
        ref.gpu_evolve(sc.R, sc.Cv, dt0)
        check(lib().afar_gpu_evolve(h, sc.R, sc.Cv, dt0, s["nc"], s["nf"], s["nb"], s["maxS"], s["maxC"], s["maxLB"],
                                    s["maxCLB"]))
    FD0, sh0 = ref.copy_to_host()
    check(lib().afar_copy_field_data_device_to_host(C.byref(fp), h, s["nc"], s["nb"]))
    assert np.array_equal(b["FD"], FD0) and np.array_equal(b["shockIndicator"], sh0)
    check(lib().afar_free_field_data(h))
    assert L.hipStreamSynchronize(stream) == 0 and L.hipStreamDestroy(stream) == 0
