"""Pins of the host precompute (pseudo-inverse, RBF coefficients, LSQ weights,
basis cache, reference block tables) and of the oracle's reconstruction on
fields whose answer is known."""
import os

import numpy as np
import pytest

from oracle import precompute_ref as PR
from oracle import pyoracle as O


def pinv(afar, A):
    from afar_cudaeuler2dmixo3rk3solver_amd._lib import check, dptr, lib
    A = np.ascontiguousarray(A, dtype=np.float64)
    G = np.zeros_like(A)
    check(lib().afar_pinv(A.shape[0], dptr(A), dptr(G)))
    return G


@pytest.mark.parametrize("n", [3, 9, 13, 21])
def test_pinv_moore_penrose(afar, n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n))
    G = pinv(afar, A)
    assert np.allclose(A @ G @ A, A, rtol=1e-10, atol=1e-10)
    assert np.allclose(G @ A @ G, G, rtol=1e-10, atol=1e-10)
    assert np.allclose(G, np.linalg.inv(A), rtol=1e-9, atol=1e-9)


def test_pinv_rank_deficient(afar):
    rng = np.random.default_rng(4)
    B = rng.normal(size=(8, 5))
    A = B @ B.T  # rank 5
    G = pinv(afar, A)
    # exact zeros of a rank-deficient matrix come out at round-off level; the
    # reference's SVDinv (minCondition 0) inverts them too, and the precompute
    # then flags the cell as trouble (|G| > 1e10)
    assert np.abs(G).max() > 1e10 or np.allclose(A @ G @ A, A, atol=1e-8)


@pytest.fixture(scope="module")
def tri_mesh(afar):
    m = afar.Mesh.riemann(n=14, jitter=0.2, seed=11, split_p=0.5)
    sc = afar.schemes(cOrder=7, fOrder=2)
    m.precompute(sc)
    return m, sc


def test_rbf_against_numpy_restatement(tri_mesh):
    m, sc = tri_mesh
    s = m.sizes()
    nc, nf, nb = s["nc"], s["nf"], s["nb"]
    S, Cc, st, cf = m.stencils()
    rbf, trouble, refd = m.rbf_basis()
    xy, _ = m.cells()
    f = m.faces()
    gauss = {k: m.gauss_points(k, 7 if k < nc else 2) for k in range(nc + nb)}
    assert trouble.sum() == 0
    for fc in range(nf):
        for side, cell in ((0, f["own"][fc]), (1, f["nei"][fc])):
            a, tr, _ = PR.rbf_coefficients(cell, list(st[cell, : S[cell]]), xy, gauss, f["face"][fc, :2], 10.0)
            assert not tr
            got = rbf[fc, side, : S[cell] + 1]
            assert np.abs(got - a).max() < 1e-7
            assert abs(got.sum() - 1.0) < 1e-12  # constants reproduced exactly
            assert np.all(rbf[fc, side, S[cell] + 1:] == 0.0)


def test_basis_cache_round_trip(afar, tri_mesh, tmp_path):
    m, sc = tri_mesh
    d = str(tmp_path)
    m.write_basis(d)
    head = open(os.path.join(d, "basisOwnList")).read(400)
    assert "FoamFile" in head and "basisOwnList" in head
    m2 = afar.Mesh.riemann(n=14, jitter=0.2, seed=11, split_p=0.5)
    sc2 = afar.schemes(cOrder=7, fOrder=2, recalcBasis=False)
    m2.precompute(sc2, basis_dir=d)
    assert np.array_equal(m2.rbf_basis()[0], m.rbf_basis()[0])
    # a missing file is an I/O error, not a crash
    m3 = afar.Mesh.riemann(n=14, jitter=0.2, seed=11, split_p=0.5)
    with pytest.raises(afar.AfarError) as e:
        m3.precompute(sc2, basis_dir=str(tmp_path / "nowhere"))
    assert e.value.code == 4


def test_block_tables_encode_the_stencil(tri_mesh):
    m, sc = tri_mesh
    nc = m.nc
    W0 = np.tile([1.0, 0.0, 0.0, 1.4], (nc, 1))
    a = m.reference_arrays(sc, W0, block_tables=True)
    s = a["sizes"]
    stride = s["maxLB"] - 128
    S = a["stencilSize"]
    for c in range(nc):
        blk = c // 128
        for i in range(S[c]):
            loc = a["localStencil"][c, i]
            g = blk * 128 + loc if loc < 128 else a["extendStencil"][stride * blk + loc - 128]
            assert g == a["globalStencil"][c, i]
    # compact halo first, then the rest (basicCUDAsolver.cpp:703-780)
    assert np.all(a["compactExtendStencilSize"] <= a["extendStencilSize"])


def _linear_field(m, sc, grad):
    """Reference arrays whose cells and pseudo-cells hold a linear field
    (cell averages of a linear field are its centroid values)."""
    nc = m.nc
    xy, _ = m.cells()
    perm = m.permutation()
    lin = lambda p, k: 1.0 + grad[k][0] * p[..., 0] + grad[k][1] * p[..., 1]
    Wm = np.stack([lin(xy[:nc], k) for k in range(4)], axis=1)
    W0 = np.zeros_like(Wm)
    W0[perm] = Wm
    ref = m.reference_arrays(sc, W0)
    for k in range(4):
        ref["FD"][nc:, k] = lin(xy[nc:], k)
    return ref


def test_lsq_gradient_exact_for_linear_fields(tri_mesh):
    m, sc = tri_mesh
    grad = [(0.3, -0.2), (1.1, 0.4), (-0.7, 0.25), (0.05, 0.6)]
    orc = O.OracleState(_linear_field(m, sc, grad), sc.R, sc.Cv)
    orc.reconstruct()
    g = orc.a["gradFD"]
    # gradientFieldData: px, py, Uxx(dUx/dx), Uxy(dUy/dx), Uyx(dUx/dy), Uyy(dUy/dy), Tx, Ty
    want = np.array([grad[0][0], grad[0][1], grad[1][0], grad[2][0], grad[1][1], grad[2][1], grad[3][0], grad[3][1]])
    assert np.abs(g - want).max() < 1e-11


def test_reconstruct_constant_field(tri_mesh):
    """A uniform field: zero gradients, face values equal to the state, flat
    temperature branch (limiter 1, smooth flag 0)."""
    m, sc = tri_mesh
    orc = O.OracleState(_linear_field(m, sc, [(0.0, 0.0)] * 4), sc.R, sc.Cv)
    orc.reconstruct()
    orc.bvd()
    assert np.all(orc.a["gradFD"] == 0.0)
    assert np.all(orc.a["faceFD"] == 1.0)
    # cells with two boundary faces take the second-order type 1 (cudaLimiter.h:419-429)
    nb_faces = (np.abs(orc.a["cellFaces"]) - 1 >= m.sizes()["nf"]).sum(axis=1)
    assert np.all(orc.smooth_flags == 0)
    assert np.array_equal(orc.final_types, (nb_faces >= 2).astype(np.int8))
    assert np.all(orc.a["limiter"] == 1.0)  # BJ(1) = 1


def test_precompute_independent_of_host_threads(afar, monkeypatch):
    """The host precompute's parallel loops give the same stencils and
    coefficients bit for bit whatever the thread count (AFAR_HOST_THREADS)."""
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    out = []
    for threads in ("1", "3", "8"):
        monkeypatch.setenv("AFAR_HOST_THREADS", threads)
        m, _ = cases.riemann_case(n=40, seed=4)
        m.precompute(afar.schemes(cOrder=7, fOrder=2))
        out.append((m.stencils(), m.rbf_basis()))
    for (st, rb) in out[1:]:
        for a, b in zip(st + rb, out[0][0] + out[0][1]):
            assert np.array_equal(a, b)
