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
        csrc = os.path.join(os.path.dirname(HERE), "spruce_amd", "csrc")
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-I" + csrc, "-o", exe, os.path.join(HERE, "arith_check.cpp"),
                        "-lm"], check=True)
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


def test_correctly_rounded_powers_match_quad_precision():
    """The powers the conduction, radiation and viscosity kernels use (crpow.hpp: x^1.5, x^2.5,
    x^-1.5, x^(1/3)) equal libquadmath's powq rounded to double, on 3e5 samples per exponent over
    2^-60..2^100 and 8..8e9."""
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "crpow_check")
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "crpow_check.cpp"),
                        "-lquadmath", "-lm"], check=True)
        r = subprocess.run([exe, "300000"], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout
