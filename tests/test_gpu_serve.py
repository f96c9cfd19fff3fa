"""GPU: the continuous-batching request loop.  Requests that run together
(own streams, overlapping on the device) give exactly the result each gives
alone, with the same step counts; the device clock and the residual stop
criterion are honoured."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_requests():
    from afar_cudaeuler2dmixo3rk3solver_amd.serve import Request
    return [Request(0, 0.0, "corner", 3000, 11, res_tol=1e-30, max_steps=60),
            Request(1, 0.0, "riemann", 6000, 12, t_end=0.01),
            Request(2, 0.0, "corner", 9000, 13, res_tol=1e-30, max_steps=100),
            Request(3, 0.0, "riemann", 2500, 14, t_end=0.02)]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_served_equals_alone(afar, precision):
    from afar_cudaeuler2dmixo3rk3solver_amd.serve import Server, build_case
    print("case: 4 requests served together vs alone (%s)" % precision, flush=True)
    srv = Server(precision=precision, max_active=4, chunk=20, prep_workers=2)
    done, wall = srv.run(make_requests(), realtime=False)
    assert len(done) == 4
    for r, a in zip(done, make_requests()):
        alone = Server(precision=precision, max_active=1, chunk=20)
        sc = alone.sc if a.kind == "corner" else alone.sc_riemann
        mesh, W0 = build_case(a, sc)
        alone.solve(a, mesh, W0, time.perf_counter())
        assert a.steps == r.steps, r.rid
        assert np.array_equal(a.result, r.result), r.rid
        if r.kind == "corner":
            assert r.steps == r.max_steps
            assert np.isfinite(r.residual)
        else:
            assert r.steps >= 1


def test_corner_residual_stop(afar):
    from afar_cudaeuler2dmixo3rk3solver_amd.serve import Request, Server, build_case
    print("case: corner 2000 cells, residual tolerance stops the run", flush=True)
    srv = Server(precision="mixed", max_active=1, chunk=10)
    r = Request(0, 0.0, "corner", 2000, 5, res_tol=1e30, max_steps=200)
    mesh, W0 = build_case(r, srv.sc)
    srv.solve(r, mesh, W0, time.perf_counter())
    assert r.steps == 10  # the first check already meets the tolerance


def test_served_with_gpu_precompute_equals_host(afar):
    """Requests precomputed on the GPU at admission give the results of the
    host-precomputed run (the GPU precompute is bitwise the host's)."""
    from afar_cudaeuler2dmixo3rk3solver_amd.serve import Server
    print("case: 4 requests, GPU precompute vs host precompute (mixed)", flush=True)
    a, _ = Server(precision="mixed", max_active=4, chunk=20, gpu_precompute=True).run(make_requests(), realtime=False)
    b, _ = Server(precision="mixed", max_active=4, chunk=20, gpu_precompute=False).run(make_requests(), realtime=False)
    for x, y in zip(a, b):
        assert x.steps == y.steps and np.array_equal(x.result, y.result), x.rid
