"""spruce_run (the reference's command-line program on the GPU library) against the reference
executable built from its own sources (oracle/_ref/run): same .state/.config inputs, same command
line, and byte-identical mhd.out, init.state and end.state, in input mode and in continue mode.
Cases whose modules call libm's pow in the reference (conduction, radiative losses) are compared
to 1e-10 instead: the GPU computes those powers correctly rounded."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))
from cases import build_case  # noqa: E402

from spruce_amd.stateio import parse_state  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "run")
RUN = os.path.join(ROOT, "spruce_amd", "spruce_run")

pytestmark = pytest.mark.gpu

EXACT = ["mhd_periodic_rk2", "mhd_open_rk4", "mhd_mixed_euler", "mhd_reflect_rk2", "tf_ucnp_rk2",
         "tf_curl_periodic_rk4", "visc_hyper_rk2", "heat_const_euler", "e2_open_rk4", "moc_mixed_stretch_rk2",
         "moc_all_limits_rk4"]


def _inputs(tmp, name, **settings):
    st, cfg = build_case(name)
    for k, v in settings.items():
        cfg.settings[k] = str(v)
    sp, cp = os.path.join(tmp, "case.state"), os.path.join(tmp, "case.config")
    st.write(sp)
    cfg.write(cp)
    return sp, cp


def _run(exe, args, cwd):
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=300, cwd=cwd)
    return r.returncode, r.stdout, r.stderr


def _both(tmp, args_of):
    out = {}
    for tag, exe in (("ref", REF), ("gpu", RUN)):
        d = os.path.join(tmp, tag)
        rc, so, se = _run(exe, args_of(d), tmp)
        print(tag, "exit", rc, se.strip()[-200:], flush=True)
        out[tag] = (d, rc, so, se)
    return out


def _files(d):
    return sorted(f for f in os.listdir(d) if f.endswith((".out", ".state")))


def _same_files(o):
    (da, ra, _, sa), (db, rb, _, sb) = o["ref"], o["gpu"]
    assert ra in (134, -6) and "successfully reached" in sa, sa[-500:]
    assert rb == 134 and "successfully reached" in sb, sb[-500:]
    assert _files(da) == _files(db)
    for f in _files(da):
        a, b = open(os.path.join(da, f), "rb").read(), open(os.path.join(db, f), "rb").read()
        if a != b:
            pytest.fail("%s differs (%d vs %d bytes): %s" % (f, len(a), len(b), _first_difference(a, b)))
    return _files(da)


def _first_difference(a, b):
    """Where two output files first differ: the line, the grid name above it and both texts."""
    k = next((i for i in range(min(len(a), len(b))) if a[i] != b[i]), min(len(a), len(b)))
    la, lb = a[:k].split(b"\n"), b[:k].split(b"\n")
    at = next((i for i in range(len(la) - 1, -1, -1)
               if la[i] and b"," not in la[i] and not la[i][:1].isdigit() and la[i][:1] != b"-"), -1)
    name = la[at] if at >= 0 else b"?"
    end_a, end_b = a.find(b"\n", k), b.find(b"\n", k)
    return "line %d (grid %s, %d lines below its name): %r vs %r" % (
        len(la), name.decode(errors="replace"), len(la) - 1 - at,
        (la[-1] + a[k:end_a if end_a >= 0 else None])[:120], (lb[-1] + b[k:end_b if end_b >= 0 else None])[:120])


@pytest.mark.parametrize("name", EXACT)
def test_runner_outputs_are_byte_identical(name, tmp_path):
    tmp = str(tmp_path)
    sp, cp = _inputs(tmp, name)
    files = _same_files(_both(tmp, lambda d: ["-m", "input", "-o", d, "-s", sp, "-c", cp]))
    assert "mhd.out" in files and "end.state" in files and "init.state" in files


def test_runner_output_schedule_and_continue_mode(tmp_path):
    """iter output interval, buffered writes, the default 4-digit precision and the standard output
    lines, then a run that continues from end.state (whose t= is written to 6 digits) and appends to
    mhd.out until max_iterations."""
    tmp = str(tmp_path)
    sp, cp = _inputs(tmp, "mhd_open_rk2", max_iterations=9, iter_output_interval=2, write_interval=3,
                     write_precision=4, std_out_interval=2)
    first = _both(tmp, lambda d: ["-m", "input", "-o", d, "-s", sp, "-c", cp])
    _same_files(first)
    ref_out = first["ref"][2].replace(os.path.join(tmp, "ref"), "<out>")
    gpu_out = first["gpu"][2].replace(os.path.join(tmp, "gpu"), "<out>")
    assert ref_out == gpu_out, (ref_out, gpu_out)
    cont = {}
    for tag, exe in (("ref", REF), ("gpu", RUN)):
        rc, so, se = _run(exe, ["-m", "continue", "-o", os.path.join(tmp, tag), "-d", "1e300"], tmp)
        print(tag, "continue exit", rc, se.strip()[-200:], flush=True)
        cont[tag] = (os.path.join(tmp, tag), rc, so, se)
    _same_files(cont)


SOURCE_MODULES = ["lheat_ramp_rk2", "mass_inj_rk4", "mom_inj_osc_euler", "sink_exp_heat_rk2", "divclean_rk2",
                  "ar_flood_time_euler", "ar_flood_masks_rk2", "outflow_gauss_x1_dyn_euler", "outflow_flat_y1_rk4",
                  "moc_visc_outflow_rk2"]


@pytest.mark.parametrize("name", SOURCE_MODULES)
def test_runner_source_modules_byte_identical_with_messages(name, tmp_path):
    """Files and the per-step standard output lines, whose module messages switch from Off to On
    at the injection start times."""
    tmp = str(tmp_path)
    sp, cp = _inputs(tmp, name, std_out_interval=1)
    o = _both(tmp, lambda d: ["-m", "input", "-o", d, "-s", sp, "-c", cp])
    _same_files(o)
    ref_out = o["ref"][2].replace(os.path.join(tmp, "ref"), "<out>")
    gpu_out = o["gpu"][2].replace(os.path.join(tmp, "gpu"), "<out>")
    assert ref_out == gpu_out, (ref_out, gpu_out)


def test_runner_with_libm_modules_matches_to_1e10(tmp_path):
    tmp = str(tmp_path)
    sp, cp = _inputs(tmp, "solar_tc_rl_rk2")
    o = _both(tmp, lambda d: ["-m", "input", "-o", d, "-s", sp, "-c", cp])
    (da, ra, _, _), (db, rb, _, _) = o["ref"], o["gpu"]
    assert ra in (134, -6) and rb == 134
    a = parse_state(open(os.path.join(da, "end.state")).read(), continue_mode=True)
    b = parse_state(open(os.path.join(db, "end.state")).read(), continue_mode=True)
    assert list(a.grids) == list(b.grids) and a.t == b.t
    for k in a.grids:
        x, y = a.grids[k], b.grids[k]
        assert np.all(np.abs(x - y) <= 1e-10 * np.maximum(np.abs(x), 1e-300)), k
