"""Writes tests/golden/corner_8x4.npz: a small fixture of the FP64 pipeline.

The reference ships no test vectors, so these are produced by this build's
oracle (oracle/afar_oracle.cpp) on an 8x4 jittered corner mesh and reviewed
against the closed-form pins in tests/test_cpu_*.py.  They freeze, for
regression: the RCM permutation, stencils, compact face lists, the RBF
coefficients, and the state, flags and time steps after 3 CFL-limited steps.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def build_fixture():
    import afar_cudaeuler2dmixo3rk3solver_amd as A
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    from oracle import pyoracle as O
    mesh, W0 = cases.corner_case(nx=8, ny=4, seed=0)
    sc = A.schemes(cOrder=7, fOrder=2, maxCo=0.5)
    mesh.precompute(sc)
    S, Cc, st, cf = mesh.stencils()
    rbf, tr, refd = mesh.rbf_basis()
    orc = O.OracleState(mesh.reference_arrays(sc, W0), sc.R, sc.Cv)
    dts = orc.run(3, sc.maxCo)
    return dict(perm=mesh.permutation(), S=S, C=Cc, stencil=st, cellFaces=cf, rbf=rbf, refDistance=refd,
                dts=dts, U=orc.U.copy(), W=orc.a["FD"].copy(), shock=orc.a["shockIndicator"].copy())


if __name__ == "__main__":
    out = os.path.join(HERE, "corner_8x4.npz")
    np.savez_compressed(out, **build_fixture())
    print("wrote", out)
