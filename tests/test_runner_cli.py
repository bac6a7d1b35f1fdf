"""spruce_run's command line and input checks, which run before any GPU call: the reference's
asserts on these paths (utils.cpp:41-45, fileio.cpp:111-115, main.cpp:23) become a message on stderr
and exit status 1."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))
from cases import build_case  # noqa: E402

RUN = os.path.join(ROOT, "spruce_amd", "spruce_run")


def _run(args, cwd):
    if not os.path.exists(RUN):
        pytest.skip("spruce_run is not built")
    r = subprocess.run([RUN] + args, capture_output=True, text=True, timeout=60, cwd=cwd)
    return r.returncode, r.stdout, r.stderr


def _inputs(tmp, extra_lines=(), module_opts=None):
    st, cfg = build_case("mhd_periodic_rk2")
    if module_opts is not None:
        cfg.modules = [("thermal_conduction", module_opts)]
    sp, cp = os.path.join(tmp, "c.state"), os.path.join(tmp, "c.config")
    st.write(sp)
    with open(cp, "w") as f:
        f.write(cfg.to_text() + "".join(l + "\n" for l in extra_lines))
    return sp, cp


def test_missing_output_directory(tmp_path):
    rc, _, err = _run(["-m", "input"], str(tmp_path))
    assert rc == 1 and "output directory must be specified" in err


def test_flag_without_value(tmp_path):
    rc, _, err = _run(["-m", "input", "-o"], str(tmp_path))
    assert rc == 1 and "flag followed by non-flag" in err


def test_unknown_config_name(tmp_path):
    sp, cp = _inputs(str(tmp_path), extra_lines=["not_a_setting = 3"])
    rc, _, err = _run(["-m", "input", "-o", os.path.join(str(tmp_path), "o"), "-s", sp, "-c", cp], str(tmp_path))
    assert rc == 1 and "not_a_setting is not a valid config name." in err
