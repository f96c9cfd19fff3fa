"""The host-only library code built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++, no GPU code) and run over small meshes of
every generator, with the face-quadrature extension, every tile size, batch
unions and a polyMesh round trip: it must finish with no sanitizer report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "afar_cudaeuler2dmixo3rk3solver_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_paths")
    srcs = [os.path.join(ROOT, "tests", "sanitize", "host_paths.cpp")] + [
        os.path.join(CSRC, f) for f in ("mesh.cpp", "precompute.cpp", "layout.cpp", "foam_io.cpp")]
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-I", CSRC, "-o", exe] + srcs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    case = tmp_path / "case"
    (case / "constant" / "polyMesh").mkdir(parents=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               AFAR_HOST_THREADS="4")
    r = subprocess.run([exe, str(case)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-3000:]
    assert len(r.stdout.strip().splitlines()) == 4, r.stdout
