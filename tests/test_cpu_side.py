"""CPU-only checks: the C ABI library exports every declared symbol, the division identity the
stencils rely on, and the text-format readers/writers."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

from spruce_amd import stateio
from spruce_amd.gpu import LIB_PATH, header_functions

HERE = os.path.dirname(os.path.abspath(__file__))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "libspruce_gpu.so not built"
    lib = ctypes.CDLL(LIB_PATH)
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_reciprocal_corrected_division_is_exact():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "arith_check")
        subprocess.run(["gcc", "-O2", "-o", exe, os.path.join(HERE, "arith_check.c"), "-lm"], check=True)
        r = subprocess.run([exe, "20000000"], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout


def test_state_roundtrip():
    rng = np.random.default_rng(0)
    g = {n: rng.standard_normal((7, 5)) * 1e-12 for n in ("d_x", "rho")}
    st = stateio.StateFile(7, 5, 1.6726e-24, 5.0 / 3.0, 0.0, g, ["#c"])
    back = stateio.parse_state(st.to_text(17).replace("\n", "\r\n"))
    assert back.comments == ["#c"]
    for n in g:
        np.testing.assert_array_equal(back.grids[n], g[n])


def test_config_roundtrip_and_traps():
    from spruce_amd import workloads as W
    cfg = W.arcade_config(steps=3)
    back = stateio.parse_config(cfg.to_text())
    assert back.equation_set == "ideal_mhd"
    assert [m for m, _ in back.modules] == ["thermal_conduction", "radiative_losses"]
    with pytest.raises(ValueError):
        stateio.parse_config("ideal_mhd = true\n{\n\n}\n")
    with pytest.raises(ValueError):
        stateio.parse_config("ideal_mhd = true\n{\n}\nrho_min = 3\n")


def test_format_matches_ostream():
    assert stateio.format_double(0.1, 17) == "0.10000000000000001"
    assert stateio.format_double(1e-15, 4) == "1e-15"
    assert stateio.format_double(202020202.020202, -1) == "202020202.020202"
