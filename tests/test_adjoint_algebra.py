"""The hand-written adjoint algebra of the IdealMHD step (spruce_amd/csrc/adj_cell.hpp) against
central finite differences of the forward per-cell functions, on the host: the transposed time
derivative (faces, limiter branches, Lorentz forces, induction) and the step-size gradient, with
the sound speed and with the Alfven speed dominating.  The same header is compiled into the GPU
adjoint kernels."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("adj") / "adjoint_algebra")
    cmd = [HIPCC, "-O1", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
           "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "adjoint_algebra.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


@pytest.mark.parametrize("field", ["1", "100"])
def test_adjoint_algebra_matches_finite_differences(checker, field):
    r = subprocess.run([checker, field], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    worst = float(r.stdout.strip().splitlines()[-1].split()[1])
    assert worst < 1e-6


def test_conduction_reverse_matches_forward_mode_and_differences(tmp_path):
    """The hand-written reverse of the conduction energy derivative (tc_cell.hpp) against the same
    function evaluated in forward mode (all 15 partials at once) and against central differences,
    at every interior cell of periodic and ghost-bounded non-uniform meshes."""
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path / "tc_reverse_check")
    cmd = [HIPCC, "-O1", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
           "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "tc_reverse_check.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([out], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
