"""Host code and the oracle under AddressSanitizer + UndefinedBehaviorSanitizer
(CPU only): tests/sanitize/Makefile builds the drivers with gcc/g++; they run
the Gmsh generator / reader / monomial table of libmgm's host code and the
oracle's whole pipeline (reader, stencils, reordering, weights, transfer rows,
colouring, V-cycles) on generated clouds.  Any sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HERE = os.path.join(ROOT, "tests", "sanitize")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    if not shutil.which("gcc") or not shutil.which("g++"):
        pytest.skip("no host compiler")
    out = str(tmp_path_factory.mktemp("san"))
    r = subprocess.run(["make", "-s", "-C", HERE, f"OUT={out}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def _run(cmd, env_extra=None):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, \
        r.stdout + r.stderr
    return r.stdout


def test_host_io_under_sanitizers(built, tmp_path):
    assert "ok" in _run([os.path.join(built, "host_io_driver"), str(tmp_path)])


@pytest.mark.parametrize("dim,n", [(2, 21), (3, 8)])
def test_oracle_under_sanitizers(built, tmp_path, dim, n):
    from multigrid_meshless_amd import Points
    paths = []
    for k, m in enumerate((n, (n + 1) // 2 + 1)):
        p = str(tmp_path / f"l{k}.msh")
        Points.generate(dim, m, 3 + k, 0.2, with_volume_elements=True, gmsh_path=p)
        paths.append(p)
    assert "ok" in _run([os.path.join(built, "oracle_driver"), *paths, str(dim)])
