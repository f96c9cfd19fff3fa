"""OpenFOAM case I/O (ASCII, one-cell-thick 2D): a mesh written as a polyMesh
reads back as the same mesh (geometry, connectivity, stencils and solution
bit for bit); 0/ fields are read in their uniform and nonuniform forms."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as O


def same_mesh(a, b):
    assert a.sizes()["nc"] == b.sizes()["nc"] and a.sizes()["nf"] == b.sizes()["nf"]
    assert a.sizes()["nb"] == b.sizes()["nb"]
    ca, va = a.cells()
    cb, vb = b.cells()
    assert np.array_equal(ca, cb) and np.array_equal(va, vb)
    fa, fb = a.faces(), b.faces()
    for k in ("own", "nei", "face", "bpatch"):
        assert np.array_equal(fa[k], fb[k]), k
    assert np.array_equal(a.permutation(), b.permutation())


@pytest.mark.parametrize("kind", ["corner", "riemann"])
def test_polymesh_round_trip(afar, tmp_path, kind):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    if kind == "corner":
        m, W0 = cases.corner_case(nx=24, ny=12)
    else:
        m, W0 = cases.riemann_case(n=16, seed=3)
    m.write_foam(tmp_path)
    names = [n for n, _ in m.patches()]
    r = afar.Mesh.read_foam(tmp_path, renumber=False)
    pr = r.patches()
    assert [n for n, _ in pr[:-1]] == names and pr[-1] == ("frontAndBack", 0)
    assert [k for _, k in pr[:-1]] == [k for _, k in m.patches()]
    # the written cell order is the mesh order; reading keeps it
    assert np.array_equal(r.permutation(), np.arange(r.nc))
    mo = m.permutation()
    cm, _ = m.cells()
    cr, _ = r.cells()
    assert np.array_equal(cm[: m.nc], cr[: r.nc])
    for p, (name, n) in enumerate(m.patches()):
        lab = 6 if kind == "corner" and p == 0 else (2 if kind == "corner" and p >= 2 else 4)
        val = [1 / 1.4, 2.0, 0.0, (1 / 1.4) / cases.R_REF] if lab == 6 else None
        r.set_patch(p, lab, val)
    sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    m.precompute(sc)
    r.precompute(sc)
    Sm, Cm, stm, cfm = m.stencils()
    Sr, Cr, str_, cfr = r.stencils()
    assert np.array_equal(Sm, Sr) and np.array_equal(stm, str_) and np.array_equal(cfm, cfr)
    Wr = np.empty_like(W0)
    Wr[:] = W0[mo]  # file order = mesh order of m
    om = O.OracleState(m.reference_arrays(sc, W0), sc.R, sc.Cv)
    orr = O.OracleState(r.reference_arrays(sc, Wr), sc.R, sc.Cv)
    assert np.array_equal(om.a["FD"], orr.a["FD"])
    om.run(3, sc.maxCo)
    orr.run(3, sc.maxCo)
    assert np.array_equal(om.U, orr.U)


def test_polymesh_renumbered_read(afar, tmp_path):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    m, _ = cases.corner_case(nx=20, ny=10)
    m.write_foam(tmp_path)
    r = afar.Mesh.read_foam(tmp_path, renumber=True)
    # RCM of an RCM-ordered mesh: same cells, possibly another order
    cm, vm = m.cells()
    cr, vr = r.cells()
    assert np.allclose(np.sort(vm), np.sort(vr), rtol=0, atol=0)
    assert abs(vm.sum() - vr.sum()) < 1e-12


def write_field(path, name, cls, internal, boundary):
    with open(path, "w") as f:
        f.write("FoamFile\n{\n    version 2.0;\n    format ascii;\n    class %s;\n    object %s;\n}\n" % (cls, name))
        f.write("// a comment\ndimensions      [0 2 -2 0 0 0 0];\n\n")
        f.write("internalField   %s;\n\n/* block\ncomment */\nboundaryField\n{\n" % internal)
        for pname, body in boundary:
            f.write("    %s\n    {\n%s    }\n" % (pname, body))
        f.write("}\n")


def test_read_fields(afar, tmp_path):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    m, _ = cases.corner_case(nx=8, ny=4)
    m.write_foam(tmp_path)
    r = afar.Mesh.read_foam(tmp_path)
    nc = r.nc
    d = tmp_path / "0"
    d.mkdir()
    names = [n for n, _ in r.patches()]
    fixed_p = "        type fixedValue;\n        value uniform 0.7142857;\n"
    zg = "        type zeroGradient;\n"
    empty = "        type empty;\n"
    bnd = lambda fx: [(names[0], fx)] + [(n, zg) for n in names[1:-1]] + [(names[-1], empty)]
    pvals = np.linspace(0.5, 1.5, nc)
    write_field(d / "p", "p", "volScalarField",
                "nonuniform List<scalar> %d\n(\n%s\n)" % (nc, "\n".join("%.17g" % v for v in pvals)), bnd(fixed_p))
    write_field(d / "U", "U", "volVectorField", "uniform (2 0.25 0)",
                bnd("        type fixedValue;\n        value uniform (2 0 0);\n"))
    write_field(d / "T", "T", "volScalarField", "uniform 1", bnd("        type fixedValue;\n        value uniform 1;\n"))
    W0, pv, pf = r.read_foam_fields(tmp_path, "0")
    assert np.array_equal(W0[:, 0], pvals)
    assert np.all(W0[:, 1] == 2.0) and np.all(W0[:, 2] == 0.25) and np.all(W0[:, 3] == 1.0)
    assert pf[0] and not pf[1:].any()
    assert np.array_equal(pv[0], [0.7142857, 2.0, 0.0, 1.0])


def test_foam_errors(afar, tmp_path):
    from afar_cudaeuler2dmixo3rk3solver_amd._lib import AfarError
    with pytest.raises(AfarError):
        afar.Mesh.read_foam(tmp_path / "nowhere")
    p = afar.Mesh.periodic(n=6)
    with pytest.raises(AfarError):
        p.write_foam(tmp_path)
