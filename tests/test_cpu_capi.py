"""The C-ABI library loads and exports every function include/afar.h declares,
and the host-side argument checks fail cleanly (no GPU calls here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "afar.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(afar_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(afar):
    names = declared_functions()
    assert len(names) >= 40
    L = afar.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_header(afar):
    from afar_cudaeuler2dmixo3rk3solver_amd._lib import SIGNATURES
    assert set(declared_functions()) == set(SIGNATURES)


def test_errors_are_reported_not_fatal(afar):
    L = afar.lib()
    h = C.c_void_p()
    rc = L.afar_mesh_generate_riemann(1, 0.2, 1, 0.5, C.byref(h))
    assert rc == 1 and b"n >= 2" in L.afar_last_error()
    with pytest.raises(afar.AfarError):
        afar.schemes(stencilType=5)  # unknown type -> rejected at precompute
        m = afar.Mesh.riemann(n=4)
        m.precompute(afar.schemes(stencilType=5))


def test_bad_polygon_mesh_rejected(afar):
    # clockwise square -> non-positive area
    pts = [[0, 0], [0, 1], [1, 1], [1, 0]]
    with pytest.raises(afar.AfarError, match="counter-clockwise"):
        afar.Mesh.from_polygons(pts, [0, 4], [0, 1, 2, 3], [[0, 1], [1, 2], [2, 3], [3, 0]], [0, 0, 0, 0], [4])
    # an edge with no patch
    with pytest.raises(afar.AfarError, match="boundary edge without a patch"):
        afar.Mesh.from_polygons([[0, 0], [1, 0], [1, 1], [0, 1]], [0, 4], [0, 1, 2, 3], [[0, 1]], [0], [4])
    # no cells at all
    with pytest.raises(afar.AfarError, match="empty mesh"):
        afar.Mesh.from_polygons(np.zeros((0, 2)), [0], np.zeros(0, int), np.zeros((0, 2), int), np.zeros(0, int), [4])


def test_polygon_mesh_equals_generator(afar):
    """A caller-built mesh (same polygons as the generator) renumbers and
    precomputes identically."""
    n = 6
    g = afar.Mesh.riemann(n=n, jitter=0.0, split_p=0.0, seed=1)
    pts = np.array([[i / n, j / n] for j in range(n + 1) for i in range(n + 1)])
    P = lambda i, j: j * (n + 1) + i
    cells = []
    for j in range(n):
        for i in range(n):
            cells += [P(i, j), P(i + 1, j), P(i + 1, j + 1), P(i, j + 1)]
    be, bp = [], []
    for j in range(n):
        be += [[P(0, j), P(0, j + 1)], [P(n, j), P(n, j + 1)]]
        bp += [0, 1]
    for i in range(n):
        be += [[P(i, 0), P(i + 1, 0)], [P(i, n), P(i + 1, n)]]
        bp += [2, 3]
    m = afar.Mesh.from_polygons(pts, np.arange(0, 4 * n * n + 1, 4), cells, be, bp, [4, 4, 4, 4])
    assert np.array_equal(m.permutation(), g.permutation())
    assert np.array_equal(m.faces()["face"], g.faces()["face"])


def test_stencil_over_twenty_is_an_error(afar):
    """A vertex shared by 26 triangles gives their edgeNeighbour stencils more
    than 20 cells: the precompute reports it (the reference exits,
    basicCUDAsolver.cpp:393-397) and the library stays usable."""
    k = 26
    ang = np.arange(k) * 2 * np.pi / k
    pts = np.vstack([[0.0, 0.0], np.stack([np.cos(ang), np.sin(ang)], 1)])
    cells = []
    for i in range(k):
        cells += [0, 1 + i, 1 + (i + 1) % k]
    be = [[1 + i, 1 + (i + 1) % k] for i in range(k)]
    m = afar.Mesh.from_polygons(pts, np.arange(0, 3 * k + 1, 3), cells, be, [0] * k, [4])
    with pytest.raises(afar.AfarError, match="Too large stencil"):
        m.precompute(afar.schemes(cOrder=7, fOrder=2))
    # a valid mesh still builds afterwards
    ok = afar.Mesh.riemann(n=6, seed=2)
    ok.precompute(afar.schemes(cOrder=7, fOrder=2))


def test_stream_and_64bit_count_checks(afar):
    """The caller-stream entry points and the 64-bit seam-2 counts reject bad
    input before any device call."""
    import afar_cudaeuler2dmixo3rk3solver_amd._lib as L_
    L = afar.lib()
    m = afar.Mesh.riemann(n=4)
    sc = afar.schemes(cOrder=7, fOrder=2)
    m.precompute(sc)
    W0 = np.tile([1.0, 0.0, 0.0, 1.0], (m.nc, 1))
    h = C.c_void_p()
    rc = L.afar_solver_create_on_stream(m.handle, C.byref(sc), W0.ctypes.data_as(C.POINTER(C.c_double)), None,
                                        C.byref(h))
    assert rc == L_.AFAR_OK + 1 and b"stream is NULL" in L.afar_last_error()
    assert L.afar_device_set_stream(None, None) == 1
    # more cells than 32-bit ids can address: refused with AFAR_ERR_LIMIT, host arrays not read
    fp = L_.FieldPointer()
    dummy = np.zeros(16)
    for name in ("CELL", "FACE", "cellVolume", "FD", "cFD"):
        setattr(fp, name, dummy.ctypes.data_as(C.POINTER(C.c_double)))
    dev = C.c_void_p()
    rc = L.afar_set_device_field_data(C.byref(dev), C.byref(fp), 2 ** 31, 10, 4)
    assert rc == 3 and b"32-bit" in L.afar_last_error()
    assert not dev.value
