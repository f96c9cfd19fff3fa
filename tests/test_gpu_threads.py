"""GPU: solvers created, stepped and destroyed on different host threads at
the same time (the threading contract of include/afar.h).  Each thread runs
its own meshes through creation, host-time-step and device-time-step
(graph-captured) steps, a residual norm, a copy-back and destruction, three
rounds over; every result equals the same work done on one thread, bit for
bit.  ctypes releases the GIL during each library call, so the threads'
library calls overlap."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _work(afar, mesh, W0, sc):
    g = afar.BasicCUDASolver(mesh, sc, W0)
    for _ in range(5):
        g.evolve(g.set_delta_t())
    g.evolve_auto(30)
    r = g.residual_norm()
    U = g.conserved()
    g.close()
    return U, r


def test_two_threads(afar):
    from afar_cudaeuler2dmixo3rk3solver_amd import cases
    print("case: 4 host threads x 3 rounds of create / step / capture / destroy vs one thread", flush=True)
    jobs = []
    for k, (nx, ny, prec) in enumerate([(40, 20, "fp64"), (64, 30, "mixed"), (50, 26, "fp64"), (30, 16, "mixed")]):
        mesh, W0 = cases.corner_case(nx=nx, ny=ny, seed=k, mach_u=1.8 + 0.3 * k)
        sc = afar.schemes(cOrder=7, fOrder=2, maxCo=0.5, precision=prec)
        mesh.precompute(sc)
        jobs.append((mesh, W0, sc))
    expect = [_work(afar, *j) for j in jobs]
    errors = []

    def run(k):
        try:
            for _ in range(3):
                U, r = _work(afar, *jobs[k])
                if not (np.array_equal(U, expect[k][0]) and r == expect[k][1]):
                    errors.append("thread %d: result differs" % k)
        except Exception as e:  # reported below
            errors.append("thread %d: %r" % (k, e))

    threads = [threading.Thread(target=run, args=(k,)) for k in range(len(jobs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors
