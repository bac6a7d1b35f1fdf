"""GPU parity for the source-term modules: localized heating, mass and momentum injection, the
ambient heating sink, field heating, divergence cleaning, anomalous resistivity and boundary
outflow.

Their templates depend on positions only and are built on the host with the C library's exp, sin,
cos and pow, as the reference builds them, so every case without a power of a grid is
bit-identical to the reference golden vectors (grids after set-up, after step 1 and after the last
step, and the step-size sequence) and to the CPU restatement step by step, subcycle counts
included.  Field heating raises grids to arbitrary powers with the device pow, and the frobenius
template of anomalous resistivity raises it to 1.5 correctly rounded; the reference's C library pow
need not agree in the last place, so those are compared within 1e-10 relative L-infinity, the
tolerance of the plain path, with step sizes within 1e-12 (and the x^1.5 case bit for bit against
the restatement with correctly rounded powers).
"""
import os
import sys

import numpy as np
import pytest

from golden_util import load_golden
from oracle.oracle import OracleDomain, set_pow_mode
from spruce_amd import workloads as W
from spruce_amd.gpu import GpuDomain
from spruce_amd.spec import spec_from_files

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from cases import random_mhd_state  # noqa: E402

pytestmark = pytest.mark.gpu

EXACT = ["lheat_ramp_rk2", "lheat_periodic_euler", "mass_inj_rk4", "mom_inj_rot_rk2", "mom_inj_osc_euler",
         "sink_exp_heat_rk2", "sink_const_euler", "divclean_rk2", "ar_flood_time_euler", "ar_flood_masks_rk2",
         "outflow_exp_y2_rk2", "outflow_gauss_x1_dyn_euler", "outflow_flat_y1_rk4", "outflow_exp_x2_aligned_rk2"]
VARS = ["rho", "temp", "mom_x", "mom_y", "mom_z", "bi_x", "bi_y", "bi_z", "n", "thermal_energy", "press", "dt"]
REL_TOL = 1e-10


def same_bits(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return a.shape == b.shape and (np.array_equal(a.view(np.int64), b.view(np.int64)) or np.array_equal(a, b, equal_nan=True))


def rel_linf(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b)))


@pytest.mark.parametrize("case", EXACT)
def test_source_modules_match_reference_golden_bitwise(case):
    g = load_golden(case)
    dom = GpuDomain(g.spec())
    try:
        for name, ref in g.frame(0).items():
            assert same_bits(dom.grid(name), ref), "%s setup %s" % (case, name)
        step = 0
        for k in g.frames[1:]:
            dts = dom.step(k - step)
            assert same_bits(dts, g.dt_steps[step:k]), "%s dt %s vs %s" % (case, dts, g.dt_steps[step:k])
            step = k
            for name, ref in g.frame(k).items():
                got = dom.grid(name)
                assert same_bits(got, ref), "%s frame %d %s: rel %.3e" % (case, k, name, rel_linf(got, ref))
    finally:
        dom.close()


@pytest.mark.parametrize("case", EXACT)
def test_source_modules_match_restatement_each_step(case):
    spec = load_golden(case).spec()
    gpu = GpuDomain(spec)
    ora = OracleDomain(spec)
    try:
        for s in range(5):
            dg, sg = gpu.step(1, subcycles=True)
            do = ora.step(1)
            assert list(sg[0]) == ora.subcycles(), "step %d subcycles %s vs %s" % (s, sg[0], ora.subcycles())
            assert same_bits(dg, do), "step %d dt %r vs %r" % (s, dg, do)
        for v in VARS:
            assert same_bits(gpu.grid(v), ora.grid(v)), v
        assert gpu.time == ora.time
    finally:
        gpu.close()
        ora.close()


@pytest.mark.parametrize("case", ["fheat_rk2", "ar_frob_ys_rk4"])
def test_pow_modules_match_reference_within_tolerance(case):
    """Field heating (arbitrary powers) and the frobenius template of anomalous resistivity (x^1.5)
    go through the C library's pow in the reference."""
    g = load_golden(case)
    dom = GpuDomain(g.spec())
    try:
        step = 0
        for k in g.frames[1:]:
            dts = dom.step(k - step)
            ref = g.dt_steps[step:k]
            assert np.max(np.abs(dts - ref) / ref) <= 1e-12, (dts, ref)
            step = k
            for name, want in g.frame(k).items():
                got = dom.grid(name)
                assert rel_linf(got, want) <= REL_TOL, "frame %d %s: %.3e" % (k, name, rel_linf(got, want))
    finally:
        dom.close()


def _big_spec(bc, integ, modules, seed):
    st = random_mhd_state(133, 197, seed, stretch=True)
    cfg = W.base_config("ideal_mhd", bc, integ, epsilon=0.2, density_min=1.0, temp_min=1.0,
                        thermal_energy_min=1e-12, max_iterations=4, modules=modules)
    return spec_from_files(cfg, st)


MULTI = [
    ("divclean_open", ("open", "fixed", "reflect", "open"), "rk2",
     [("div_cleaning", [("epsilon", "0.05"), ("time_scale", "0.01")])]),
    ("divclean_periodic", ("periodic",) * 4, "euler",
     [("div_cleaning", [("epsilon", "0.1"), ("time_scale", "0.02")])]),
    ("injections", ("periodic", "periodic", "open", "open"), "rk4",
     [("localized_heating", [("start_time", "0.0"), ("duration", "1.0"), ("max_heating_rate", "3.0"),
                             ("stddev_x", "9.0"), ("stddev_y", "12.0"), ("center_x", "128.0"),
                             ("center_y", "60.0"), ("ramp_time", "0.05")]),
      ("mass_injection", [("start_time", "0.0"), ("duration", "1.0"), ("max_injection_rate", "2e9"),
                          ("stddev_x", "5.0"), ("stddev_y", "7.0"), ("center_x", "3.0"), ("center_y", "100.0")]),
      ("momentum_injection", [("start_time", "0.0"), ("duration", "1.0"), ("max_accel", "1e8"),
                              ("stddev_x", "6.0"), ("stddev_y", "4.0"), ("center_x", "70.0"),
                              ("center_y", "150.0"), ("dir_x", "-0.4"), ("dir_y", "-1.0"),
                              ("template_angle", "-20.0"), ("oscillatory", "true"),
                              ("oscillation_period", "0.3")]),
      ("ambient_heating_sink", [("heating_rate", "0.2")])]),
    # a window of thousands of cells (several reduction blocks), along the field, throttled
    ("outflow_dynamic", ("open", "fixed", "reflect", "open"), "rk2",
     [("boundary_outflow", [("max_accel", "1e9"), ("falloff_length", "6e9"), ("boundary", "y_bound_2"),
                            ("falloff_shape", "gaussian"), ("feather_length", "1e8"), ("field_aligned_mode", "true"),
                            ("dynamic_mode", "true"), ("dynamic_time", "0.5"), ("dynamic_target_speed", "2e7")])]),
    ("outflow_flat_x2", ("periodic", "periodic", "open", "open"), "euler",
     [("boundary_outflow", [("max_accel", "3e8"), ("falloff_length", "4e8"), ("boundary", "x_bound_2"),
                            ("falloff_shape", "flat"), ("feather_length", "5e8")])]),
]


@pytest.mark.parametrize("name,bc,integ,modules", MULTI, ids=[m[0] for m in MULTI])
def test_source_modules_multi_block_match_restatement(name, bc, integ, modules):
    spec = _big_spec(bc, integ, modules, 91)
    gpu = GpuDomain(spec)
    ora = OracleDomain(spec)
    try:
        for s in range(4):
            dg, sg = gpu.step(1, subcycles=True)
            do = ora.step(1)
            assert list(sg[0]) == ora.subcycles(), "step %d subcycles %s vs %s" % (s, sg[0], ora.subcycles())
            assert same_bits(dg, do), "step %d dt %r vs %r" % (s, dg, do)
        for v in VARS:
            assert same_bits(gpu.grid(v), ora.grid(v)), "%s %s" % (name, v)
    finally:
        gpu.close()
        ora.close()


def test_boundary_outflow_report_follows_the_throttle():
    """The reported outflow speed and acceleration: max_accel without dynamic mode, and with it
    (target - outflow) / dynamic_time clipped to [0, max_accel]; the acceleration is 0 before the
    first step."""
    g = load_golden("outflow_gauss_x1_dyn_euler")
    dom = GpuDomain(g.spec())
    try:
        r0 = dom.module_report(0)
        assert r0.shape == (2,) and r0[1] == 0.0 and np.isfinite(r0[0])
        for _ in range(3):
            dom.step(1)
            mean, acc = dom.module_report(0)
            assert acc == min((3e7 - mean) / 1.0, 1e9) and 0.0 < acc < 1e9, (mean, acc)
    finally:
        dom.close()
    g = load_golden("outflow_exp_y2_rk2")
    dom = GpuDomain(g.spec())
    try:
        dom.step(2)
        assert dom.module_report(0)[1] == 1e8
    finally:
        dom.close()


def test_field_heating_multi_block_matches_restatement_within_tolerance():
    mods = [("field_heating", [("coeff", "2e3"), ("current_pow", "1.5"), ("b_pow", "-0.5"), ("n_pow", "0.25"),
                               ("roc_pow", "1.0")])]
    spec = _big_spec(("periodic", "periodic", "fixed", "open"), "rk2", mods, 92)
    gpu = GpuDomain(spec)
    ora = OracleDomain(spec)
    try:
        for s in range(4):
            dg = gpu.step(1)
            do = ora.step(1)
            assert abs(dg[0] - do[0]) <= 1e-12 * do[0], (s, dg, do)
        for v in VARS:
            assert rel_linf(gpu.grid(v), ora.grid(v)) <= REL_TOL, v
        # heating acted: the thermal energy differs from a run without the module
        spec0 = _big_spec(("periodic", "periodic", "fixed", "open"), "rk2", [], 92)
        base = OracleDomain(spec0)
        base.step(4)
        assert rel_linf(gpu.grid("thermal_energy"), base.grid("thermal_energy")) > 1e-6
        base.close()
    finally:
        gpu.close()
        ora.close()


def test_resistivity_frobenius_matches_cr_restatement_bitwise():
    spec = load_golden("ar_frob_ys_rk4").spec()
    set_pow_mode(1)
    gpu = GpuDomain(spec)
    ora = OracleDomain(spec)
    try:
        for s in range(4):
            dg, sg = gpu.step(1, subcycles=True)
            do = ora.step(1)
            assert list(sg[0]) == ora.subcycles(), (s, sg[0], ora.subcycles())
            assert same_bits(dg, do), (s, dg, do)
        for v in VARS:
            assert same_bits(gpu.grid(v), ora.grid(v)), v
    finally:
        set_pow_mode(0)
        gpu.close()
        ora.close()


def test_resistivity_multi_block_matches_restatement():
    """Flood-fill template with both masks, Syntelis-19 resistivity, gradient correction and RK4
    subcycles on a 133x197 stretched mesh (many blocks, a large flood region)."""
    mods = [("anomalous_resistivity", [
        ("resistivity_model", "syntelis_19"), ("resistivity_model_params", "2e13,5e13,150.0"),
        ("template_mode", "flood_fill"), ("flood_fill_threshold", "2.4"), ("flood_fill_argmin_radius", "3e8"),
        ("flood_fill_max_radius", "6e8"), ("flood_fill_min_current", "80.0"),
        ("flood_fill_current_ramp_length", "150.0"), ("smoothing_sigma", "1.5"), ("gradient_correction", "true"),
        ("time_integrator", "rk4")])]
    spec = _big_spec(("open", "fixed", "reflect", "open"), "rk2", mods, 93)
    gpu = GpuDomain(spec)
    ora = OracleDomain(spec)
    try:
        for s in range(3):
            dg, sg = gpu.step(1, subcycles=True)
            do = ora.step(1)
            assert list(sg[0]) == ora.subcycles(), (s, sg[0], ora.subcycles())
            assert same_bits(dg, do), (s, dg, do)
        for v in VARS:
            assert same_bits(gpu.grid(v), ora.grid(v)), v
        spec0 = _big_spec(("open", "fixed", "reflect", "open"), "rk2", [], 93)
        base = OracleDomain(spec0)
        base.step(3)
        assert rel_linf(gpu.grid("bi_x"), base.grid("bi_x")) > 1e-6
        base.close()
    finally:
        gpu.close()
        ora.close()
