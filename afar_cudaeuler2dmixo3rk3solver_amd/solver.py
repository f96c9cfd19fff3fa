"""Python mirror of the reference's solver interfaces over the C-ABI.

Seam 1 -- ``BasicCUDASolver`` mirrors the C++ class basicCUDAsolver
(src/CUDA/cudaEuler2DmixO3RK3/basicCUDAsolver.h:43-52): ``set_delta_t``,
``evolve``, ``write_field`` and destruction; the mesh and its precompute are
built by ``Mesh``.

Seam 2 -- ``DeviceFields`` mirrors the free functions of
src/CUDA/cudaEuler2DmixO3RK3/GPUevolve.h:6-59 over reference-layout arrays.
"""
import ctypes as C

import numpy as np

from ._lib import FieldPointer, Schemes, check, dptr, lib, ptr

STENCIL_TYPES = {"faceNeighbour": 0, "edgeNeighbour": 1}


def schemes(R=0.7142857, Cv=1.7857142857, maxCo=0.5, shapeFactor=10.0, stencilType="edgeNeighbour",
            recalcBasis=True, cOrder=5, fOrder=4, precision="fp64", exactRK=False, tile=64,
            boundaryFirstOrder=False, reconLanes=0, cellOrder=0, faceQuadrature=0):
    """afar_schemes with the reference's key names and defaults (the last
    seven are this library's extensions)."""
    s = Schemes()
    lib().afar_schemes_default(C.byref(s))
    s.R, s.Cv, s.maxCo, s.shapeFactor = R, Cv, maxCo, shapeFactor
    s.stencilType = STENCIL_TYPES[stencilType] if isinstance(stencilType, str) else int(stencilType)
    s.recalcBasis = int(bool(recalcBasis))
    s.cOrder, s.fOrder = cOrder, fOrder
    s.precision = {"fp64": 0, "mixed": 1, 0: 0, 1: 1}[precision]
    s.exactRK = int(bool(exactRK))
    s.tile = tile
    s.boundaryFirstOrder = int(bool(boundaryFirstOrder))
    s.reconLanes = int(reconLanes)
    s.cellOrder = int(cellOrder)
    s.faceQuadrature = int(faceQuadrature)
    return s


class Mesh:
    """A renumbered 2D finite-volume mesh (afar_mesh)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self.precomputed = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().afar_mesh_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    # -- constructors
    @classmethod
    def corner(cls, nx=200, ny=100, lx=1.0, ly=0.5, jitter=0.15, seed=0, corner_x=0.5, angle_deg=15.0,
               inflow=(1.0 / 1.4, 2.0, 0.0, 1.0), top_label=2):
        h = C.c_void_p()
        infl = np.ascontiguousarray(inflow, dtype=np.float64)
        check(lib().afar_mesh_generate_corner(nx, ny, lx, ly, jitter, seed, corner_x, angle_deg, dptr(infl),
                                              top_label, C.byref(h)))
        return cls(h)

    @classmethod
    def riemann(cls, n=1000, jitter=0.2, seed=1, split_p=0.5):
        h = C.c_void_p()
        check(lib().afar_mesh_generate_riemann(n, jitter, seed, split_p, C.byref(h)))
        return cls(h)

    @classmethod
    def periodic(cls, n=64, L=10.0, jitter=0.15, seed=0, split_p=1.0):
        h = C.c_void_p()
        check(lib().afar_mesh_generate_periodic(n, L, jitter, seed, split_p, C.byref(h)))
        return cls(h)

    @classmethod
    def read_foam(cls, case_dir, renumber=False):
        """The 2D mesh of an OpenFOAM case (constant/polyMesh, ASCII, one cell
        thick); patches in polyMesh order, label 4 until set_patch."""
        h = C.c_void_p()
        check(lib().afar_mesh_read_foam(str(case_dir).encode(), int(bool(renumber)), C.byref(h)))
        return cls(h)

    def write_foam(self, case_dir):
        """Writes case_dir/constant/polyMesh (created if missing)."""
        import os
        os.makedirs(os.path.join(str(case_dir), "constant", "polyMesh"), exist_ok=True)
        check(lib().afar_mesh_write_foam(self._h, str(case_dir).encode()))

    def patches(self):
        """[(name, boundary faces)] of every patch, empty ones included."""
        out = []
        for p in range(self.sizes()["npatches_all"]):
            buf = C.create_string_buffer(256)
            n = C.c_int32()
            check(lib().afar_mesh_patch_info(self._h, p, buf, 256, C.byref(n)))
            out.append((buf.value.decode(), n.value))
        return out

    def set_patch(self, patch, label, value=None):
        """Boundary label of patch `patch` (CUDA.boundary.patch<i>); value
        (p, Ux, Uy, T) for label 6."""
        v = None if value is None else np.ascontiguousarray(value, dtype=np.float64)
        check(lib().afar_mesh_set_patch(self._h, patch, label, None if v is None else dptr(v)))

    def read_foam_fields(self, case_dir, time_dir="0"):
        """(W0 per cell in file order, per-patch fixedValue values, fixed flags)
        from the case's p, U and T."""
        n = self.sizes()["npatches_all"]
        W0 = np.zeros((self.nc, 4))
        pv = np.zeros((n, 4))
        pf = np.zeros(n, dtype=np.int32)
        check(lib().afar_foam_read_fields(self._h, str(case_dir).encode(), str(time_dir).encode(), dptr(W0),
                                          dptr(pv), ptr(pf, C.c_int32)))
        return W0, pv, pf.astype(bool)

    @classmethod
    def from_polygons(cls, points, cell_ptr, cell_pts, bedges, bedge_patch, patch_labels, patch_values=None,
                      cell_vshift=None, period=(0.0, 0.0), depth=1.0, renumber=True):
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
        cptr = np.ascontiguousarray(cell_ptr, dtype=np.int64)
        cpts = np.ascontiguousarray(cell_pts, dtype=np.int32)
        be = np.ascontiguousarray(bedges, dtype=np.int32).reshape(-1, 2)
        bp = np.ascontiguousarray(bedge_patch, dtype=np.int32)
        pl = np.ascontiguousarray(patch_labels, dtype=np.int32)
        pv = None if patch_values is None else np.ascontiguousarray(patch_values, dtype=np.float64).reshape(-1, 4)
        vs = None if cell_vshift is None else np.ascontiguousarray(cell_vshift, dtype=np.int8)
        h = C.c_void_p()
        check(lib().afar_mesh_from_polygons(len(pts), dptr(pts), len(cptr) - 1, ptr(cptr, C.c_int64),
                                            ptr(cpts, C.c_int32), ptr(vs, C.c_int8), period[0], period[1], len(be),
                                            ptr(be, C.c_int32), ptr(bp, C.c_int32), len(pl), ptr(pl, C.c_int32),
                                            dptr(pv), depth, int(renumber), C.byref(h)))
        return cls(h)

    # -- queries
    def sizes(self):
        o = np.zeros(9, dtype=np.int64)
        check(lib().afar_mesh_sizes(self._h, ptr(o, C.c_int64)))
        keys = ["nc", "nf", "nb", "npatches", "maxS", "maxC", "maxLB", "maxCLB", "npatches_all"]
        return dict(zip(keys, (int(v) for v in o)))

    @property
    def nc(self):
        return self.sizes()["nc"]

    def cells(self):
        s = self.sizes()
        xy = np.zeros((s["nc"] + s["nb"], 2))
        v = np.zeros(s["nc"])
        check(lib().afar_mesh_cells(self._h, dptr(xy), dptr(v)))
        return xy, v

    def permutation(self):
        p = np.zeros(self.nc, dtype=np.int32)
        check(lib().afar_mesh_permutation(self._h, ptr(p, C.c_int32)))
        return p

    def faces(self):
        s = self.sizes()
        own = np.zeros(s["nf"] + s["nb"], dtype=np.int32)
        nei = np.zeros(s["nf"], dtype=np.int32)
        face = np.zeros((s["nf"] + s["nb"], 5))
        fsh = np.zeros((s["nf"], 2))
        bp = np.zeros(s["nb"], dtype=np.int32)
        check(lib().afar_mesh_faces(self._h, ptr(own, C.c_int32), ptr(nei, C.c_int32), dptr(face), dptr(fsh),
                                    ptr(bp, C.c_int32)))
        return dict(own=own, nei=nei, face=face, fshift=fsh, bpatch=bp)

    def polygons(self):
        nc = self.nc
        vptr = np.zeros(nc + 1, dtype=np.int32)
        check(lib().afar_mesh_polygons(self._h, ptr(vptr, C.c_int32), None))
        vxy = np.zeros((int(vptr[-1]), 2))
        check(lib().afar_mesh_polygons(self._h, None, dptr(vxy)))
        return vptr, vxy

    # -- precompute (basicCUDAsolver::initStencilData)
    def precompute(self, sc, basis_dir=None, gpu=False):
        """initStencilData: stencils, quadrature, RBF basis (gpu: the RBF part
        on the GPU, bitwise equal to the host's)."""
        if gpu:
            if basis_dir is not None:
                raise ValueError("the GPU precompute does not read or write the basis cache")
            check(lib().afar_mesh_precompute_gpu(self._h, C.byref(sc)))
        else:
            check(lib().afar_mesh_precompute(self._h, C.byref(sc),
                                             None if basis_dir is None else basis_dir.encode()))
        self.precomputed = True
        return self

    def stencils(self):
        s = self.sizes()
        nc, maxS, maxC = s["nc"], s["maxS"], s["maxC"]
        S = np.zeros(nc, dtype=np.uint8)
        Cc = np.zeros(nc, dtype=np.uint8)
        st = np.zeros((nc, maxS), dtype=np.int32)
        cf = np.zeros((nc, maxC), dtype=np.int32)
        check(lib().afar_mesh_stencils(self._h, ptr(S, C.c_uint8), ptr(Cc, C.c_uint8), ptr(st, C.c_int32),
                                       ptr(cf, C.c_int32)))
        return S, Cc, st, cf

    def rbf_basis(self):
        s = self.sizes()
        rbf = np.zeros((s["nf"], 2, s["maxS"] + 1))
        tr = np.zeros(s["nc"], dtype=np.uint8)
        ref = np.zeros(s["nc"])
        check(lib().afar_mesh_rbf_basis(self._h, dptr(rbf), ptr(tr, C.c_uint8), dptr(ref)))
        return rbf, tr, ref

    def face_quadrature(self):
        """Flux quadrature of a mesh precomputed with faceQuadrature 1, or
        None: dict(xi (ngp,), weights (ngp,), offsets (ngp, nf, 4): point -
        owner centre and point - neighbour centre, rbf (ngp, nf, 2, maxS+1))."""
        s = self.sizes()
        n = C.c_int32()
        check(lib().afar_mesh_face_quadrature(self._h, C.byref(n), None, None, None, None))
        ngp = n.value
        if ngp == 0:
            return None
        q = dict(xi=np.zeros(ngp), weights=np.zeros(ngp), offsets=np.zeros((ngp, s["nf"], 4)),
                 rbf=np.zeros((ngp, s["nf"], 2, s["maxS"] + 1)))
        check(lib().afar_mesh_face_quadrature(self._h, C.byref(n), dptr(q["xi"]), dptr(q["weights"]),
                                              dptr(q["offsets"]), dptr(q["rbf"])))
        return q

    def gauss_points(self, member, order):
        xy = np.zeros((32, 2))
        w = np.zeros(32)
        n = C.c_int32()
        check(lib().afar_mesh_gauss_points(self._h, member, order, dptr(xy), dptr(w), 32, C.byref(n)))
        return xy[: n.value].copy(), w[: n.value].copy()

    def write_basis(self, directory):
        check(lib().afar_mesh_write_basis(self._h, directory.encode()))

    def reference_arrays(self, sc, W0, block_tables=False):
        """Reference-layout host arrays (fieldPointer) in mesh order, FP64 coefficients."""
        s = self.sizes()
        nc, nf, nb, maxS, maxC = s["nc"], s["nf"], s["nb"], s["maxS"], s["maxC"]
        a = dict(
            CELL=np.zeros((nc + nb, 2)), FACE=np.zeros((nf + nb, 5)), cellVolume=np.zeros(nc),
            FD=np.zeros((nc + nb, 4)), cFD=np.zeros((nc, 4)), shockIndicator=np.zeros(2 * (nc + nb), np.int8),
            neighbour=np.zeros((nf, 2), np.int32), boundaryFacesNum=np.zeros(s["npatches"], np.int32),
            boundaryFaceNeiLabel=np.zeros(nb, np.int32), boundaryFacesType=np.zeros(nb, np.uint8),
            stencilSize=np.zeros(nc, np.uint8), compactStencilSize=np.zeros(nc, np.uint8),
            cellFaces=np.zeros((nc, maxC), np.int32), globalStencil=np.zeros((nc, maxS), np.int32),
            RBFbasis64=np.zeros((nf, 2, maxS + 1)), RBFbasis=np.zeros((nf, 2, maxS + 1), np.float32),
            faceShift=np.zeros((nf, 2)), boundaryValue=np.zeros((nb, 4)))
        if block_tables:
            nblk = (nc + 127) // 128
            a["extendStencil"] = np.zeros(max(s["maxLB"] - 128, 1) * nblk, np.int32)
            a["extendStencilSize"] = np.zeros(nblk, np.uint16)
            a["compactExtendStencilSize"] = np.zeros(nblk, np.uint16)
            a["localStencil"] = np.zeros((nc, maxS), np.uint16)
        fp = field_pointer(a)
        w0 = np.ascontiguousarray(W0, dtype=np.float64)
        check(lib().afar_mesh_export_reference(self._h, C.byref(sc), dptr(w0), C.byref(fp)))
        q = self.face_quadrature()
        if q is not None:  # extension: flux quadrature points
            a["gpWeights"], a["gpOffsets"], a["RBFgp"] = q["weights"], q["offsets"], q["rbf"]
        a["sizes"] = s
        return a


_FP_TYPES = {name: t for name, t in FieldPointer._fields_}


def field_pointer(arrays):
    """afar_field_pointer over a dict of numpy arrays (missing keys stay NULL)."""
    fp = FieldPointer()
    for name, t in FieldPointer._fields_:
        if name in arrays and arrays[name] is not None:
            arr = arrays[name]
            assert arr.flags["C_CONTIGUOUS"], name
            setattr(fp, name, arr.ctypes.data_as(t))
    return fp


class BasicCUDASolver:
    """basicCUDAsolver over afar_solver: create, set_delta_t, evolve, write_field."""

    def __init__(self, mesh, sc, W0):
        if not mesh.precomputed:
            mesh.precompute(sc)
        self.mesh = mesh
        self.schemes = sc
        w0 = np.ascontiguousarray(W0, dtype=np.float64).reshape(-1, 4)
        h = C.c_void_p()
        check(lib().afar_solver_create(mesh.handle, C.byref(sc), dptr(w0), C.byref(h)))
        self._h = h
        s = mesh.sizes()
        self.nc, self.nb = s["nc"], s["nb"]
        self.time = 0.0
        self.steps = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().afar_solver_destroy(h)
            self._h = None

    def close(self):
        self.__del__()

    def set_delta_t(self):
        dt = C.c_double()
        check(lib().afar_solver_set_delta_t(self._h, C.byref(dt)))
        return dt.value

    def evolve(self, dt):
        check(lib().afar_solver_evolve(self._h, dt))
        self.time += dt
        self.steps += 1

    def evolve_auto(self, nsteps):
        check(lib().afar_solver_evolve_auto(self._h, nsteps))
        self.steps += nsteps

    def last_dt(self):
        dt = C.c_double()
        check(lib().afar_solver_last_dt(self._h, C.byref(dt)))
        return dt.value

    def synchronize(self):
        check(lib().afar_solver_synchronize(self._h))

    def idle(self):
        """True when every enqueued step has finished (does not wait)."""
        v = C.c_int32()
        check(lib().afar_solver_query(self._h, C.byref(v)))
        return bool(v.value)

    def set_time(self, t, t_end=float("inf")):
        check(lib().afar_solver_set_time(self._h, t, t_end))

    def device_time(self):
        t = C.c_double()
        check(lib().afar_solver_get_time(self._h, C.byref(t), None))
        return t.value

    def device_steps(self):
        """Steps taken on the device clock since the last set_time."""
        n = C.c_double()
        check(lib().afar_solver_get_time(self._h, None, C.byref(n)))
        return int(n.value)

    def run_until(self, t_end, chunk=64, max_steps=None):
        """Device-clocked steps from the current device time to t_end (the
        last step clipped), `chunk` graph launches between clock reads.
        Returns the number of graph launches (the last chunk may hold
        steps past t_end, which change nothing)."""
        t = self.device_time()
        self.set_time(t, t_end)
        n = 0
        while t < t_end and (max_steps is None or n < max_steps):
            k = chunk if max_steps is None else min(chunk, max_steps - n)
            self.evolve_auto(k)
            n += k
            t = self.device_time()
        return n

    def residual_norm(self):
        r = C.c_double()
        check(lib().afar_solver_residual_norm(self._h, C.byref(r)))
        return r.value

    def write_field(self):
        W = np.zeros((self.nc, 4))
        g = np.zeros((self.nb, 4))
        sh = np.zeros(self.nc, dtype=np.int8)
        check(lib().afar_solver_write_field(self._h, dptr(W), dptr(g), ptr(sh, C.c_int8)))
        return W, g, sh

    def conserved(self):
        U = np.zeros((self.nc, 4))
        check(lib().afar_solver_get_conserved(self._h, dptr(U)))
        return U

    def set_conserved(self, U):
        u = np.ascontiguousarray(U, dtype=np.float64)
        check(lib().afar_solver_set_conserved(self._h, dptr(u)))

    def time_steps(self, nsteps, graph=True):
        ms = C.c_double()
        check(lib().afar_solver_time_steps(self._h, nsteps, int(graph), C.byref(ms)))
        return ms.value

    def time_steps_kernels(self, nsteps):
        """nsteps graph-replayed steps with events around every kernel:
        (total ms, {kernel: (average ms per launch, launches per step)})."""
        names = ["cfl", "reconstruct", "bvd", "flux", "update", "flux_dt", "dt_reduce"]
        tot = C.c_double()
        ms = np.zeros(len(names))
        cnt = np.zeros(len(names), dtype=np.int32)
        check(lib().afar_solver_time_steps_kernels(self._h, nsteps, C.byref(tot), dptr(ms),
                                                   cnt.ctypes.data_as(C.POINTER(C.c_int32))))
        self.steps += nsteps
        return tot.value, {n: (ms[i] / cnt[i] if cnt[i] else 0.0, int(cnt[i]) // max(nsteps, 1))
                           for i, n in enumerate(names)}

    def profile(self, reps=1):
        names = ["cfl", "reconstruct", "bvd", "flux", "update", "flux_dt", "dt_reduce"]
        ms = np.zeros(len(names))
        cnt = np.zeros(len(names), dtype=np.int32)
        check(lib().afar_solver_profile(self._h, reps, dptr(ms), ptr(cnt, C.c_int32)))
        return dict(zip(names, zip(ms.tolist(), cnt.tolist())))

    def device_sizes(self):
        o = np.zeros(8, dtype=np.int64)
        check(lib().afar_solver_sizes(self._h, ptr(o, C.c_int64)))
        return dict(zip(["nc", "nf", "nb", "maxS", "maxC", "tile", "ntiles", "maxHalo"], (int(v) for v in o)))

    def run_stage(self, upto, dt=0.0, stage=0):
        check(lib().afar_solver_run_stage(self._h, upto, dt, stage))

    def fetch(self, name, dtype, shape):
        out = np.zeros(shape, dtype=dtype)
        check(lib().afar_solver_fetch(self._h, name.encode(), out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out


class DeviceFields:
    """The reference's host<->device functions (GPUevolve.h) over afar_device."""

    def __init__(self, a, precision=0, tile=64):
        s = a["sizes"]
        self.a = a
        self.s = s
        self._fp = field_pointer(a)
        h = C.c_void_p()
        check(lib().afar_set_device_field_data(C.byref(h), C.byref(self._fp), s["nc"], s["nf"], s["nb"]))
        self._h = h
        check(lib().afar_device_configure(h, precision, tile))
        check(lib().afar_set_device_stencil_data(h, C.byref(self._fp), s["nc"], s["nf"], s["npatches"], s["nb"],
                                                 s["maxS"], s["maxC"], s["maxLB"], s["maxCLB"]))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().afar_free_field_data(h)
            self._h = None

    def adjust_time_step(self, R, Cv, CFL):
        dt = C.c_double()
        check(lib().afar_adjust_time_step(self._h, self.s["nf"], R, Cv, CFL, C.byref(dt)))
        return dt.value

    def gpu_evolve(self, R, Cv, dt):
        s = self.s
        check(lib().afar_gpu_evolve(self._h, R, Cv, dt, s["nc"], s["nf"], s["nb"], s["maxS"], s["maxC"], s["maxLB"],
                                    s["maxCLB"]))

    def copy_to_host(self):
        check(lib().afar_copy_field_data_device_to_host(C.byref(self._fp), self._h, self.s["nc"], self.s["nb"]))
        return self.a["FD"], self.a["shockIndicator"]


def device_info():
    name = C.create_string_buffer(256)
    cus = C.c_int32()
    mem = C.c_int64()
    check(lib().afar_device_info(name, 256, C.byref(cus), C.byref(mem)))
    return name.value.decode(), cus.value, mem.value
