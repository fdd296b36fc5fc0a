"""The header-only C++ facade (include/grid2grid) used the way the reference's
own examples use the reference: tests/cpp/scalapack2scalapack.cpp."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "scalapack2scalapack")

needs_bin = pytest.mark.skipif(not os.path.exists(BIN), reason="tests/cpp not built (grid2grid_amd.build)")


@needs_bin
def test_facade_submatrix_known_answers():
    """get_scalapack_grid of the facade against the reference's submatrix KATs (host only)."""
    r = subprocess.run([BIN, "--kat-only"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "SUCCESS" in r.stdout


@needs_bin
@pytest.mark.gpu
def test_facade_scalapack2scalapack_gpu():
    """The reference's CTest cases, a transposed case and a transformer batch
    through the facade, bit-exact, 4 virtual ranks on the GPU."""
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL CORRECT" in r.stdout
