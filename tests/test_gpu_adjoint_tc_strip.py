"""The strip form of the adjoint's conduction subcycle reverse (Euler subcycles without flux
saturation, k_adj_tc_strip: one wavefront per 60 columns and R rows, the adjoint of te, rho and b_hat
written without atomics) against the tiled one-thread-per-cell form on the same steps, and against
central differences of the GPU forward.

The two forms differ only in the order their sums are taken, so their gradients agree to rounding;
the shapes leave partial strips and partial 60-column tiles, with periodic, fixed, open and reflect
sides along both axes."""
import os
import sys

import numpy as np
import pytest

from spruce_amd import workloads as W
from spruce_amd.gpu import GpuDomain
from spruce_amd.spec import spec_from_files

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))
from cases import random_mhd_state  # noqa: E402
from test_gpu_adjoint_modules import (HEAT, RL, TC, THETA, _check_directions, _relative_scale,  # noqa: E402
                                      _seed, _weights)

pytestmark = pytest.mark.gpu


def _arcade(nx, ny, seed):
    """A stratified arcade with seeded flows, field and temperature noise, stretched gently in height."""
    st = W.arcade_state(nx, ny, dx=5e6, dy0=5e6, stretch=1.01, t_corona=1e6, depth=4e7)
    u = [W.uniform(seed * 16 + k, (nx, ny)) for k in range(5)]
    rho = st.grids["rho"] * (1.0 + 0.01 * (2 * u[4] - 1))
    st.grids["rho"] = rho
    st.grids["mom_x"] = rho * 2e6 * (2 * u[0] - 1)
    st.grids["mom_y"] = rho * 2e6 * (2 * u[1] - 1)
    st.grids["bi_x"] = 0.5 * (2 * u[2] - 1)
    st.grids["bi_y"] = 0.5 * (2 * u[3] - 1)
    return st


def _cfg(bc, integ, mods, temp_min=1e4, te_min=1e-8):
    return W.base_config("ideal_mhd", bc, integ, epsilon=0.2, density_min=1.0, temp_min=temp_min,
                         thermal_energy_min=te_min, modules=mods)


CASES = [
    # name, state, boundaries, integrator, modules, steps, checkpoint interval, strip heights
    ("arcade_x_periodic_rk2", lambda: _arcade(42, 130, 3), ("periodic", "periodic", "fixed", "open"), "rk2",
     [TC, RL(), HEAT(THETA[:8])], 5, 3, (4, 7, 16)),
    ("arcade_reflect_fixed_x_euler", lambda: _arcade(37, 70, 5), ("reflect", "fixed", "periodic", "periodic"),
     "euler", [TC, HEAT(THETA[:8])], 5, 2, (5, 16)),
    ("random_periodic_rk4", lambda: random_mhd_state(34, 125, 61, stretch=True), ("periodic",) * 4, "rk4",
     [TC, RL()], 4, 2, (4, 16)),
]


@pytest.mark.parametrize("name,make,bc,integ,mods,nsteps,ckpt,rows", CASES, ids=[c[0] for c in CASES])
def test_strip_reverse_matches_tiled_reverse(name, make, bc, integ, mods, nsteps, ckpt, rows):
    print("conduction reverse strips against tiles", name, flush=True)
    st = make()
    shape = (st.xdim, st.ydim)
    rng = np.random.default_rng(7)
    w = _weights(rng, shape, (1e13, 1e6, 1.0, 1.0))
    seed = _seed(w, shape)
    cfg = _cfg(bc, integ, mods, *((1.0, 1e-12) if name.startswith("random") else ()))
    out = {}
    for R in (-1,) + tuple(rows):
        d = GpuDomain(spec_from_files(cfg, st), {"adjoint_tc_rows": R})
        try:
            g, dts = d.adjoint(nsteps, ckpt, seed)
            p = d.adjoint_params()
        finally:
            d.close()
        out[R] = (g, p, dts)
    g0, p0, d0 = out[-1]
    assert np.all(np.isfinite(g0))
    for R in rows:
        g, p, dts = out[R]
        assert np.array_equal(dts, d0), R
        for k in range(8):
            scale = float(np.max(np.abs(g0[k])))
            err = float(np.max(np.abs(g[k] - g0[k])))
            print("rows", R, "component", k, "max |difference| / max |gradient| %.2e" % (err / max(scale, 1e-300)),
                  flush=True)
            assert err <= 1e-10 * scale, (R, k, err, scale)
        if p0.size:
            assert np.max(np.abs(p - p0)) <= 1e-10 * np.max(np.abs(p0)), (R, p, p0)


def test_strip_reverse_matches_finite_differences():
    print("conduction reverse strips: finite differences", flush=True)
    st = _arcade(42, 130, 3)
    rng = np.random.default_rng(31)
    w = _weights(rng, (42, 130), (1e13, 1e6, 1.0, 1.0))

    def cfg():
        return _cfg(("periodic", "periodic", "fixed", "open"), "rk2", [TC, RL(), HEAT(THETA[:8])])
    _check_directions(st, cfg, 5, 3, w, rng, scale=_relative_scale(st), options={"adjoint_tc_rows": 8})


@pytest.mark.parametrize("integ", ["euler", "rk2", "rk4"])
def test_segment_records_match_per_step_replay(integ):
    """The conduction records kept for whole segments (the replay of a step takes its subcycle
    results from them) against replays that run every subcycle again, for each subcycle integrator:
    the same gradients."""
    print("conduction segment records", integ, flush=True)
    st = _arcade(42, 130, 3)
    shape = (st.xdim, st.ydim)
    rng = np.random.default_rng(9)
    w = _weights(rng, shape, (1e13, 1e6, 1.0, 1.0))
    tc = ("thermal_conduction", [(k, v) if k != "time_integrator" else (k, integ) for k, v in TC[1]])
    cfg = _cfg(("periodic", "periodic", "fixed", "open"), "rk2", [tc, RL(), HEAT(THETA[:8])])
    out = []
    for mode in ("auto", "step"):
        d = GpuDomain(spec_from_files(cfg, st), {"adjoint_tc_records": mode})
        try:
            g, _ = d.adjoint(6, 4, _seed(w, shape))
            out.append((g, d.adjoint_params()))
        finally:
            d.close()
    (ga, pa), (gb, pb) = out
    for k in range(8):
        assert np.max(np.abs(ga[k] - gb[k])) <= 1e-10 * np.max(np.abs(gb[k])), k
    assert np.max(np.abs(pa - pb)) <= 1e-10 * np.max(np.abs(pb))


def test_segment_records_with_viscosity_match_per_step_replay():
    """Artificial viscosity's right-hand-side terms beside conduction: with whole-segment records the
    reverse forms the terms' dt grid again from the state the integrator starts from; the gradients
    equal those of per-step replays."""
    from test_gpu_adjoint_modules import VISC
    st = _arcade(42, 130, 3)
    shape = (st.xdim, st.ydim)
    rng = np.random.default_rng(19)
    w = _weights(rng, shape, (1e13, 1e6, 1.0, 1.0))
    mods = [TC, RL(), VISC("local,global", "0.1,0.05", "v_y,temp", "mom_y,thermal_energy"), HEAT(THETA[:8])]
    cfg = _cfg(("periodic", "periodic", "fixed", "open"), "rk2", mods)
    out = []
    for mode in ("auto", "step"):
        d = GpuDomain(spec_from_files(cfg, st), {"adjoint_tc_records": mode})
        try:
            g, _ = d.adjoint(6, 4, _seed(w, shape))
            out.append((g, d.adjoint_params(), d.adjoint_device_bytes()))
        finally:
            d.close()
    (ga, pa, ba), (gb, pb, bb) = out
    assert ba > bb   # the whole-segment records were kept
    for k in range(8):
        assert np.max(np.abs(ga[k] - gb[k])) <= 1e-10 * np.max(np.abs(gb[k])), k
    assert np.max(np.abs(pa - pb)) <= 1e-10 * np.max(np.abs(pb))


def test_segment_records_with_field_heating_match_per_step_replay():
    """Field heating reads b_hat of the step's start state, which the reverse forms again with
    whole-segment records: the gradients equal those of per-step replays."""
    from test_gpu_adjoint_modules import FHEAT
    st = random_mhd_state(40, 48, 97, stretch=True)
    shape = (st.xdim, st.ydim)
    rng = np.random.default_rng(13)
    w = _weights(rng, shape, (1e15, 1e8, 10.0, 1.0))
    cfg = _cfg(("periodic", "periodic", "fixed", "reflect"), "rk2", [FHEAT], temp_min=1.0, te_min=1e-12)
    out = []
    for mode in ("auto", "step"):
        d = GpuDomain(spec_from_files(cfg, st), {"adjoint_tc_records": mode})
        try:
            g, _ = d.adjoint(6, 3, _seed(w, shape))
            out.append((g, d.adjoint_device_bytes()))
        finally:
            d.close()
    (ga, ba), (gb, bb) = out
    assert ba > bb   # the whole-segment records were kept
    for k in range(8):
        assert np.max(np.abs(ga[k] - gb[k])) <= 1e-10 * np.max(np.abs(gb[k])), k


def test_segment_records_with_outflow_match_per_step_replay():
    """Boundary outflow in dynamic mode (its acceleration follows the window's largest outflow speed
    of each step): with whole-segment records the reverse forms that speed, its cell and the
    acceleration again from the state the module acted on; the gradients equal those of per-step
    replays, and after either the module reports the scalars of the last forward step."""
    from test_gpu_adjoint_modules import OUTFLOW
    st = random_mhd_state(40, 48, 97, stretch=True)
    shape = (st.xdim, st.ydim)
    rng = np.random.default_rng(17)
    w = _weights(rng, shape, (1e15, 1e8, 10.0, 1.0))
    mod = OUTFLOW(max_accel="1e9", falloff_length="5e7", boundary="x_bound_2", falloff_shape="gaussian",
                  feather_length="2e7", field_aligned_mode="true", dynamic_mode="true", dynamic_time="1.0",
                  dynamic_target_speed="3e7")
    cfg = _cfg(("reflect", "open", "periodic", "periodic"), "rk2", [mod], temp_min=1.0, te_min=1e-12)
    ref = GpuDomain(spec_from_files(cfg, st))
    try:
        ref.step(6)
        want = ref.module_report(0)
    finally:
        ref.close()
    out = []
    for mode in ("auto", "step"):
        d = GpuDomain(spec_from_files(cfg, st), {"adjoint_tc_records": mode})
        try:
            g, _ = d.adjoint(6, 3, _seed(w, shape))
            out.append((g, d.adjoint_device_bytes()))
            assert np.array_equal(d.module_report(0), want), mode
        finally:
            d.close()
    (ga, ba), (gb, bb) = out
    assert ba > bb   # the whole-segment records were kept
    for k in range(8):
        assert np.max(np.abs(ga[k] - gb[k])) <= 1e-10 * np.max(np.abs(gb[k])), k
