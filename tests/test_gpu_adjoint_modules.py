"""GPU discrete adjoint with non-periodic boundaries and modules, against central finite
differences of the GPU forward.

Covered: the ghost-fill reverse of fixed, reflect, open and open_ucnp boundaries; thermal conduction (Euler,
RK2 and RK4 subcycles, with and without flux saturation), radiative losses (with prevent_subcycling too), the
artificial viscosity (right-hand-side terms with local, global and boundary strengths, any
differentiated variable and the gradient correction; hyperviscous terms under each subcycle
integrator), physical viscosity (heating and force, Euler and RK2 subcycles, gradient correction),
field heating, divergence cleaning, boundary outflow (its dynamic throttle too), the source modules after the integrator (localized heating with its time ramp and window, mass and
momentum injection, oscillating too, the ambient heating sink) and the parametrised heating
profile of the gradient workload, whose coefficients get their own gradient
(spruce_gpu_adjoint_params).  The loss
is a seeded linear functional of the final evolved grids.  It is only piecewise smooth (limiter,
min/max and floor selections, the argmin of the step size, the integer subcycle counts), so every
directional check takes two difference steps and leaves out directions whose two differences
disagree (a branch crossed inside the interval); the rest must agree with the adjoint to 1e-6
relative, the tolerance of the adjoint requirement."""
import os
import sys

import numpy as np
import pytest

from oracle.oracle import OracleDomain, set_pow_mode
from spruce_amd import workloads as W
from spruce_amd.gpu import GpuDomain, SpruceError
from spruce_amd.spec import spec_from_files
from spruce_amd.stateio import StateFile

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from cases import mini_arcade, random_mhd_state  # noqa: E402

pytestmark = pytest.mark.gpu

STATE = ["rho", "temp", "mom_x", "mom_y", "mom_z", "bi_x", "bi_y", "bi_z"]
SCALE = {"rho": 1e-15, "temp": 1e6, "mom_x": 1e-8, "mom_y": 1e-8, "mom_z": 1e-8, "bi_x": 1.0, "bi_y": 1.0,
         "bi_z": 1.0}
def TCI(integ):
    return ("thermal_conduction", [("flux_saturation", "false"), ("epsilon", "0.5"), ("dt_subcycle_min", "1e-6"),
                                   ("time_integrator", integ)])


TC = TCI("euler")


def TCS(integ):
    """Conduction with flux saturation: on the arcade below a quarter of the interior cells carry a
    coefficient under 0.99 at step 0 (down to 0.64), so the coefficient, its derivatives and the
    raw flux all enter the gradient."""
    return ("thermal_conduction", [("flux_saturation", "true"), ("epsilon", "0.5"), ("dt_subcycle_min", "1e-6"),
                                   ("time_integrator", integ)])


def RL(integ="euler", prevent=False):
    return ("radiative_losses", [("cutoff_ramp", "2e4"), ("cutoff_temp", "3e4"), ("epsilon", "0.5"),
                                 ("time_integrator", integ)] + ([("prevent_subcycling", "true")] if prevent else []))


def HEAT(theta):
    return ("parametric_heating", [("num_basis", str(len(theta)))] +
            [("coeff_%d" % k, repr(float(t))) for k, t in enumerate(theta)])


THETA = np.array([1e-2 * (1 + 0.1 * k) for k in range(16)])


def _arcade():
    """A small stratified arcade.  Its density is uniform along x, so at step 0 the x-face limiter
    of the density compares equal values and selects among equal branches: the loss has a kink at
    that state in every direction that moves the density unevenly (difference quotients agree with
    each other there but not with any one gradient).  A 1% seeded density noise removes the ties."""
    st = mini_arcade(24, 40, 7)
    u = W.uniform(4242, (24, 40))
    st.grids["rho"] = st.grids["rho"] * (1.0 + 0.01 * (2 * u - 1))
    return st


def _state(grids, base):
    return StateFile(base.xdim, base.ydim, base.ion_mass, base.adiabatic_index, base.t, grids, list(base.comments))


def _loss(spec, nsteps, w):
    d = GpuDomain(spec)
    try:
        d.step(nsteps)
        out = [d.grid(v) for v in ("rho", "mom_x", "thermal_energy", "bi_y")]
    finally:
        d.close()
    return sum(float(np.sum(wk * o)) for wk, o in zip(w, out))


def _seed(w, shape):
    seed = np.zeros((8,) + shape)
    seed[0], seed[1], seed[4], seed[6] = w
    return seed


def _weights(rng, shape, scales):
    return [rng.uniform(-1, 1, shape) * s for s in scales]


def _relative_scale(base):
    """Direction scales proportional to each grid's own magnitude (with a floor for grids that are
    zero somewhere): the arcade's density spans three decades, and a perturbation of a few ulps of
    a cell is swamped by the re-rounding rho = max(rho/m, n_min) * m of every update."""
    out = {}
    for v in STATE:
        a = np.abs(base.grids[v])
        out[v] = a + 1e-3 * float(np.max(a)) + 1e-3 * SCALE[v]
    return out


def _check_directions(base, make_cfg, nsteps, ckpt, w, rng, trials=6, min_smooth=4, steps=(1e-7, 1e-8),
                      scale=None, options=None):
    shape = (base.xdim, base.ydim)
    scale = scale if scale is not None else {v: SCALE[v] for v in STATE}
    d = GpuDomain(spec_from_files(make_cfg(), base), options)
    try:
        g_ev, _ = d.adjoint(nsteps, ckpt, _seed(w, shape))
        g = d.adjoint_state(g_ev)
    finally:
        d.close()
    smooth, worst = 0, 0.0
    for trial in range(trials):
        dirs = {v: rng.uniform(-1, 1, shape) * scale[v] for v in STATE}
        ad = sum(float(np.sum(g[k] * dirs[v])) for k, v in enumerate(STATE))
        fds = []
        for eps in steps:
            gp = {v: base.grids[v] + eps * dirs[v] if v in dirs else base.grids[v] for v in base.grids}
            gm = {v: base.grids[v] - eps * dirs[v] if v in dirs else base.grids[v] for v in base.grids}
            fds.append((_loss(spec_from_files(make_cfg(), _state(gp, base)), nsteps, w) -
                        _loss(spec_from_files(make_cfg(), _state(gm, base)), nsteps, w)) / (2 * eps))
        consistent = abs(fds[0] - fds[1]) <= 1e-6 * abs(fds[0])
        rel = abs(ad - fds[0]) / (abs(ad) + abs(fds[0]))
        print("direction", trial, "adjoint %.12e fd %.12e / %.12e rel %.2e%s"
              % (ad, fds[0], fds[1], rel, "" if consistent else " (branch crossed)"), flush=True)
        if consistent:
            smooth += 1
            worst = max(worst, rel)
    assert smooth >= min_smooth
    assert worst < 1e-6


BOUNDARIES = [
    ("fixed_reflect_rk2", ("periodic", "periodic", "fixed", "reflect"), "rk2", 12, 5),
    ("reflect_fixed_x_euler", ("reflect", "fixed", "periodic", "periodic"), "euler", 10, 4),
    ("fixed_open_rk2", ("periodic", "periodic", "fixed", "open"), "rk2", 12, 4),
    # RK4: stages 2-4 read the live grids whose ghost cells the intermediate fills rewrote
    ("fixed_reflect_rk4", ("periodic", "periodic", "fixed", "reflect"), "rk4", 10, 4),
    ("open_fixed_x_rk4", ("open", "fixed", "periodic", "periodic"), "rk4", 10, 4),
    # open_ucnp: the final fill and every intermediate stage state's fill copy the first interior cell
    ("ucnp_all_rk2", ("open_ucnp",) * 4, "rk2", 10, 4),
    ("ucnp_x_fixed_reflect_rk4", ("open_ucnp", "open_ucnp", "fixed", "reflect"), "rk4", 8, 3),
    ("ucnp_y_euler", ("periodic", "periodic", "open_ucnp", "open_ucnp"), "euler", 10, 5),
]


@pytest.mark.parametrize("name,bc,integ,nsteps,ckpt", BOUNDARIES, ids=[b[0] for b in BOUNDARIES])
def test_boundary_adjoint_matches_finite_differences(name, bc, integ, nsteps, ckpt):
    print("adjoint", name, flush=True)
    base = random_mhd_state(40, 48, 93, stretch=True)
    rng = np.random.default_rng(11)
    w = _weights(rng, (40, 48), (1e15, 1e8, 10.0, 1.0))

    def cfg():
        return W.base_config("ideal_mhd", bc, integ, epsilon=0.2, density_min=1.0, temp_min=1.0,
                             thermal_energy_min=1e-12)
    _check_directions(base, cfg, nsteps, ckpt, w, rng)


def VISC(opt, strength, diff, evol, length=None, species=None, **extra):
    n = len(opt.split(","))
    return ("artificial_viscosity", [("visc_opt", opt), ("visc_strength", strength), ("visc_vars_to_diff", diff),
                                     ("visc_vars_to_evol", evol), ("visc_length", length or ",".join(["0"] * n)),
                                     ("visc_species", species or ",".join(["i"] * n))] + list(extra.items()))


VISC_CASES = [
    # artificial viscosity as a right-hand-side term: local terms read the dt grid of the state the
    # integrator starts from, global ones its minimum; velocity and temperature are scaled by n
    ("local_global_rk2", ("periodic", "periodic", "fixed", "reflect"), "rk2",
     VISC("local,global", "0.1,0.05", "v_x,temp", "mom_x,thermal_energy"), 8, 3),
    ("boundary_euler", ("periodic", "periodic", "fixed", "open"), "euler",
     VISC("boundary,boundary_global", "0.3,0.2", "rho,v_y", "rho,mom_y", "2e7,3e7"), 8, 4),
    ("fields_rk4", ("periodic", "periodic", "reflect", "fixed"), "rk4",
     VISC("local,global,local", "0.2,0.1,0.1", "bi_x,b_hat_y,press", "bi_x,bi_y,thermal_energy"), 6, 2),
    ("ke_bmag_ucnp_rk2", ("open_ucnp",) * 4, "rk2",
     VISC("local,global", "0.1,0.1", "kinetic_energy,b_mag", "thermal_energy,bi_z"), 6, 3),
    # gradient correction: the derivatives of nu*scale take the neighbours' dt grid, strength and n
    ("gradcorr_rk2", ("periodic", "periodic", "fixed", "open"), "rk2",
     VISC("boundary,local,global", "0.3,0.1,0.1", "rho,temp,v_x", "rho,thermal_energy,mom_x", "2e7,0,0",
          gradient_correction="true"), 8, 3),
    ("gradcorr_rk4", ("reflect", "fixed", "periodic", "periodic"), "rk4",
     VISC("boundary_global,local", "0.2,0.1", "v_y,bi_y", "mom_y,bi_y", "3e7,0", gradient_correction="true"), 6, 2),
    # hyperviscous terms (strength above 1): subcycled after the iterating modules, each evaluation
    # taking the dt grid of the live state it reads, with each subcycle integrator
    ("hyper_euler", ("periodic", "periodic", "fixed", "reflect"), "rk2",
     VISC("local,global", "2.5,0.1", "v_y,bi_x", "mom_y,bi_x", hv_time_integrator="euler", hv_epsilon="1.0"), 6, 3),
    ("hyper_rk2", ("periodic", "periodic", "fixed", "open"), "euler",
     VISC("local,global", "2.5,1.5", "temp,rho", "thermal_energy,rho", hv_time_integrator="rk2", hv_epsilon="1.0"),
     6, 3),
    ("hyper_rk4_grad", ("reflect", "fixed", "periodic", "periodic"), "rk4",
     VISC("boundary,local", "2.0,1.5", "rho,v_x", "rho,mom_x", "3e7,0", hv_time_integrator="rk4", hv_epsilon="0.8",
          gradient_correction="true"), 4, 2),
]


def PV(coeff, **extra):
    return ("physical_viscosity", [("coeff", coeff)] + list(extra.items()))


VISC_CASES += [
    # physical (Braginskii) viscosity: heating and force through the velocity derivatives, T^2.5 and
    # the frozen b_hat of the cell and its neighbours, subcycled with Euler or RK2
    ("pv_euler", ("periodic", "periodic", "fixed", "reflect"), "rk2", PV("5e-17", epsilon="0.15"), 6, 3),
    ("pv_rk2_gc_ramp", ("periodic", "periodic", "fixed", "open"), "euler",
     PV("5e-17", ramp_length="1e8", gradient_correction="true", time_integrator="rk2", epsilon="0.1"), 6, 3),
    ("pv_force_only_rk4", ("reflect", "fixed", "periodic", "periodic"), "rk4",
     PV("5e-17", heating_on="false", epsilon="0.3"), 4, 2),
    ("pv_heat_only_rk2", ("open_ucnp",) * 4, "rk2", PV("1e-16", force_on="false", time_integrator="rk2", epsilon="0.3"), 6, 3),
]

VISC_CASES += [
    # divergence cleaning after the integrator: subcycles linear in b (about 25 per step here)
    ("divclean_rk2", ("periodic", "periodic", "reflect", "fixed"), "rk2",
     ("div_cleaning", [("epsilon", "0.1"), ("time_scale", "0.01")]), 6, 3),
    ("divclean_visc_rk4", ("fixed", "reflect", "periodic", "periodic"), "rk4",
     [VISC("local", "0.1", "v_x", "mom_x"), ("div_cleaning", [("epsilon", "0.2"), ("time_scale", "0.03")])], 4, 2),
]

def OUTFLOW(**o):
    return ("boundary_outflow", list(o.items()))


VISC_CASES += [
    # boundary outflow after the integrator: the acceleration along the boundary normal or the field,
    # and the dynamic throttle, whose acceleration follows the largest outflow speed of its window
    ("outflow_exp_y2_rk2", ("periodic", "periodic", "fixed", "open"), "rk2",
     OUTFLOW(max_accel="1e8", falloff_length="4e7", boundary="y_bound_2", falloff_shape="exp"), 6, 3),
    ("outflow_dyn_aligned_x2_rk2", ("reflect", "open", "periodic", "periodic"), "rk2",
     OUTFLOW(max_accel="1e9", falloff_length="5e7", boundary="x_bound_2", falloff_shape="gaussian",
             feather_length="2e7", field_aligned_mode="true", dynamic_mode="true", dynamic_time="1.0",
             dynamic_target_speed="3e7"), 6, 3),
    ("outflow_flat_y1_rk4", ("periodic", "periodic", "open", "fixed"), "rk4",
     OUTFLOW(max_accel="5e7", falloff_length="3e7", boundary="y_bound_1", falloff_shape="flat",
             feather_length="2e7"), 4, 2),
]

FHEAT = ("field_heating", [("coeff", "3e4"), ("current_pow", "1.0"), ("b_pow", "0.5"), ("n_pow", "-0.5"),
                           ("roc_pow", "0.5")])
VISC_CASES += [
    # field heating: the heating of the step's start state (current, |b|, n and the field lines'
    # curvature) added to the thermal energy; after a hyperviscous term the state it heats differs
    # from the one its heating reads
    ("fheat_rk2", ("periodic", "periodic", "fixed", "reflect"), "rk2", FHEAT, 6, 3),
    ("fheat_after_hyper_euler", ("periodic", "periodic", "fixed", "open"), "euler",
     [VISC("local", "2.5", "v_x", "mom_x", hv_time_integrator="rk2"),
      ("field_heating", [("coeff", "3e-4"), ("current_pow", "0.5"), ("b_pow", "1.0"), ("n_pow", "1.0"),
                         ("roc_pow", "1.0")])], 6, 3),
]


@pytest.mark.parametrize("name,bc,integ,visc,nsteps,ckpt", VISC_CASES, ids=[v[0] for v in VISC_CASES])
def test_viscosity_adjoint_matches_finite_differences(name, bc, integ, visc, nsteps, ckpt):
    print("adjoint viscosity", name, flush=True)
    base = random_mhd_state(40, 48, 97, stretch=True)
    rng = np.random.default_rng(13)
    w = _weights(rng, (40, 48), (1e15, 1e8, 10.0, 1.0))

    def cfg():
        return W.base_config("ideal_mhd", bc, integ, epsilon=0.2, density_min=1.0, temp_min=1.0,
                             thermal_energy_min=1e-12, modules=visc if isinstance(visc, list) else [visc])
    _check_directions(base, cfg, nsteps, ckpt, w, rng)


def _arcade_cfg(theta, bc=("periodic", "periodic", "fixed", "open"), integ="rk2", mods=None):
    mods = mods if mods is not None else [TC, RL(), HEAT(theta)]
    return W.base_config("ideal_mhd", bc, integ, epsilon=0.2, density_min=1.0, temp_min=1e4,
                         thermal_energy_min=1e-8, modules=mods)


def test_parametric_heating_forward_matches_restatement():
    """The gradient workload's physics (conduction, radiation, parametrised heating) on a small
    arcade: bit-identical to the restatement with correctly rounded powers, step by step."""
    st = _arcade()
    spec = spec_from_files(_arcade_cfg(THETA), st)
    set_pow_mode(1)
    gpu = GpuDomain(spec)
    ora = OracleDomain(spec)
    try:
        for s in range(6):
            dg, sg = gpu.step(1, subcycles=True)
            do = ora.step(1)
            assert list(sg[0]) == ora.subcycles(), (s, sg[0], ora.subcycles())
            assert dg[0] == do[0], (s, dg, do)
        for v in ("rho", "thermal_energy", "mom_x", "mom_y", "bi_x", "bi_y", "temp", "dt"):
            assert np.array_equal(gpu.grid(v), ora.grid(v)), v
    finally:
        set_pow_mode(0)
        gpu.close()
        ora.close()


MODULE_CASES = [
    ("tc_rl_heat_rk2", ("periodic", "periodic", "fixed", "open"), "rk2", None, 6, 4),
    ("tc_rl2_ambient_euler", ("periodic", "periodic", "reflect", "fixed"), "euler",
     [TC, RL("rk2"), ("ambient_heating", [("heating_rate", "3e-3")]), HEAT(THETA[:8])], 6, 3),
    ("tc_rl_heat_rk4", ("periodic", "periodic", "fixed", "open"), "rk4", None, 5, 3),
    # conduction with RK2 and RK4 subcycles (the midpoint, or i1..i3 and k1..k3, kept per subcycle)
    ("tc2_rl_heat_rk2", ("periodic", "periodic", "fixed", "open"), "rk2", [TCI("rk2"), RL(), HEAT(THETA[:8])], 6, 3),
    ("tc4_heat_euler", ("periodic", "periodic", "reflect", "fixed"), "euler", [TCI("rk4"), HEAT(THETA[:8])], 6, 3),
    ("tc4_rl4_heat_rk4", ("periodic", "periodic", "fixed", "open"), "rk4", [TCI("rk4"), RL("rk4"), HEAT(THETA[:8])],
     4, 2),
    # prevent_subcycling caps the losses of 65 of the 720 interior cells at step 0 by the live energy
    # over the live dt grid, whose gradient then enters
    ("tc_rlprev_heat_rk2", ("periodic", "periodic", "fixed", "open"), "rk2",
     [TC, RL(prevent=True), HEAT(THETA[:8])], 6, 3),
    ("rlprev2_heat_rk4", ("periodic", "periodic", "reflect", "fixed"), "rk4", [RL("rk2", prevent=True), HEAT(THETA[:8])],
     4, 2),
    # flux saturation under each subcycle integrator
    ("tcsat_rl_heat_rk2", ("periodic", "periodic", "fixed", "open"), "rk2", [TCS("euler"), RL(), HEAT(THETA[:8])], 6, 3),
    ("tcsat2_heat_euler", ("periodic", "periodic", "reflect", "fixed"), "euler", [TCS("rk2"), HEAT(THETA[:8])], 6, 3),
    ("tcsat4_heat_rk4", ("periodic", "periodic", "fixed", "open"), "rk4", [TCS("rk4"), HEAT(THETA[:8])], 4, 2),
    # artificial viscosity beside the iterating modules (its dt grid is that of the state after them)
    ("tc_visc_heat_rk2", ("periodic", "periodic", "fixed", "open"), "rk2",
     [TC, RL(), VISC("local,global", "0.1,0.05", "v_y,temp", "mom_y,thermal_energy"), HEAT(THETA[:8])], 6, 3),
]


@pytest.mark.parametrize("name,bc,integ,mods,nsteps,ckpt", MODULE_CASES, ids=[m[0] for m in MODULE_CASES])
def test_module_adjoint_parameters_match_finite_differences(name, bc, integ, mods, nsteps, ckpt):
    """dJ/dtheta_k of the heating coefficients against central differences in theta_k, at relative
    steps 1e-4 and 1e-5: a coefficient moves the loss by a few parts in 1e3 of its size, so at
    smaller steps the rounding of the loss dominates the difference quotient."""
    print("adjoint params", name, flush=True)
    st = _arcade()
    theta = THETA if mods is None else THETA[:8]

    def cfg(th):
        m = None if mods is None else [x if x[0] != "parametric_heating" else HEAT(th) for x in mods]
        return _arcade_cfg(th, bc, integ, m)
    rng = np.random.default_rng(21)
    w = _weights(rng, (24, 40), (1e13, 1e6, 1.0, 1.0))
    d = GpuDomain(spec_from_files(cfg(theta), st))
    try:
        d.adjoint(nsteps, ckpt, _seed(w, (24, 40)))
        gt = d.adjoint_params()
    finally:
        d.close()
    assert gt.shape == theta.shape
    smooth, worst = 0, 0.0
    for k in range(0, len(theta), 2 if len(theta) <= 8 else 3):
        fds = []
        for rel in (1e-4, 1e-5):
            e = rel * theta[k]
            tp, tm = theta.copy(), theta.copy()
            tp[k] += e
            tm[k] -= e
            fds.append((_loss(spec_from_files(cfg(tp), st), nsteps, w) -
                        _loss(spec_from_files(cfg(tm), st), nsteps, w)) / (tp[k] - tm[k]))
        consistent = abs(fds[0] - fds[1]) <= 1e-5 * abs(fds[0])
        r = abs(gt[k] - fds[0]) / (abs(gt[k]) + abs(fds[0]))
        print("theta", k, "adjoint %.12e fd %.12e / %.12e rel %.2e%s"
              % (gt[k], fds[0], fds[1], r, "" if consistent else " (branch crossed)"), flush=True)
        if consistent:
            smooth += 1
            worst = max(worst, r)
    assert smooth >= 4
    assert worst < 1e-6


@pytest.mark.parametrize("name,bc,integ,mods,nsteps,ckpt", MODULE_CASES, ids=[m[0] for m in MODULE_CASES])
def test_module_adjoint_state_matches_finite_differences(name, bc, integ, mods, nsteps, ckpt):
    print("adjoint state", name, flush=True)
    st = _arcade()
    rng = np.random.default_rng(31)
    w = _weights(rng, (24, 40), (1e13, 1e6, 1.0, 1.0))

    def cfg():
        return _arcade_cfg(THETA, bc, integ, mods)
    _check_directions(st, cfg, nsteps, ckpt, w, rng, scale=_relative_scale(st))


# Source modules after the integrator, in index coordinates of the 24 x 40 arcade (time step about
# 0.04 s).  The localized heating is in its ramp (heating factor (t - start) / ramp_time, 0.125 to
# 0.8 over the steps) and the momentum injection oscillates with a 0.5 s period: both depend on
# the simulation time, the sum of the earlier step sizes, so the step sizes' own dependence on the
# state carries their time derivative back.
LHEAT = ("localized_heating", [("start_time", "-0.05"), ("duration", "1.0"), ("max_heating_rate", "3.0"),
                               ("stddev_x", "4"), ("stddev_y", "6"), ("center_x", "10"), ("center_y", "15"),
                               ("ramp_time", "0.4")])
MASSINJ = ("mass_injection", [("start_time", "0"), ("duration", "10"), ("max_injection_rate", "1e10"),
                              ("stddev_x", "3"), ("stddev_y", "5"), ("center_x", "12"), ("center_y", "20")])
MOMINJ = ("momentum_injection", [("start_time", "0"), ("duration", "10"), ("max_accel", "1e8"), ("dir_x", "1"),
                                 ("dir_y", "0.5"), ("template_angle", "20"), ("stddev_x", "4"), ("stddev_y", "6"),
                                 ("center_x", "12"), ("center_y", "10"), ("oscillatory", "true"),
                                 ("oscillation_period", "0.5")])
SINK = ("ambient_heating_sink", [("exp_mode", "true"), ("exp_base_heating_rate", "2.0"),
                                 ("exp_scale_height", "1e8"), ("center_x", "6e7"), ("half_width", "5e7")])

SOURCE_CASES = [
    ("lheat_sink_tc_rk2", ("periodic", "periodic", "fixed", "open"), "rk2", [TC, LHEAT, SINK], 6, 3),
    ("massinj_heat_euler", ("periodic", "periodic", "reflect", "fixed"), "euler", [MASSINJ, HEAT(THETA[:8])], 6, 3),
    ("mominj_rl_rk4", ("periodic", "periodic", "fixed", "open"), "rk4", [MOMINJ, RL()], 5, 2),
    ("all_sources_rk2", ("periodic", "periodic", "fixed", "open"), "rk2", [LHEAT, MASSINJ, MOMINJ, SINK], 6, 3),
    # a localized heating window that opens after the first steps and closes before the last
    ("lheat_window_euler", ("periodic", "periodic", "fixed", "open"), "euler",
     [("localized_heating", [("start_time", "0.07"), ("duration", "0.11"), ("max_heating_rate", "3.0"),
                             ("stddev_x", "4"), ("stddev_y", "6"), ("center_x", "10"), ("center_y", "15"),
                             ("ramp_time", "0.05")])], 7, 3),
]


@pytest.mark.parametrize("name,bc,integ,mods,nsteps,ckpt", SOURCE_CASES, ids=[m[0] for m in SOURCE_CASES])
def test_source_adjoint_state_matches_finite_differences(name, bc, integ, mods, nsteps, ckpt):
    print("adjoint sources", name, flush=True)
    st = _arcade()
    rng = np.random.default_rng(41)
    w = _weights(rng, (24, 40), (1e13, 1e6, 1.0, 1.0))

    def cfg():
        return _arcade_cfg(THETA, bc, integ, mods)
    _check_directions(st, cfg, nsteps, ckpt, w, rng, scale=_relative_scale(st))


@pytest.mark.parametrize("mods", [None, [TC, LHEAT, MASSINJ, MOMINJ, SINK]], ids=["heat", "sources"])
def test_module_adjoint_leaves_the_same_final_state_as_stepping(mods):
    st = _arcade()
    spec = spec_from_files(_arcade_cfg(THETA, mods=mods), st)
    a = GpuDomain(spec)
    b = GpuDomain(spec)
    try:
        dta = a.step(7)
        _, dtb = b.adjoint(7, 3, np.zeros((8, 24, 40)))
        assert np.array_equal(dta, dtb)
        for v in ("rho", "mom_x", "thermal_energy", "bi_y", "dt"):
            assert np.array_equal(a.grid(v), b.grid(v)), v
        assert np.array_equal(a.step(2), b.step(2))
    finally:
        a.close()
        b.close()


def test_inactive_modules_leave_the_gradient_alone():
    """Modules in inactive_mode compute their update without applying it: the adjoint with inactive
    conduction, radiative losses and field heating equals the adjoint without them."""
    st = _arcade()
    off = [("inactive_mode", "true")]
    inactive = [(TC[0], TC[1] + off), (RL()[0], RL()[1] + off),
                ("field_heating", FHEAT[1] + off), HEAT(THETA[:8])]
    rng = np.random.default_rng(51)
    w = _weights(rng, (24, 40), (1e13, 1e6, 1.0, 1.0))
    out = []
    for mods in (inactive, [HEAT(THETA[:8])]):
        d = GpuDomain(spec_from_files(_arcade_cfg(THETA[:8], mods=mods), st))
        try:
            g, dts = d.adjoint(5, 2, _seed(w, (24, 40)))
            out.append((g, d.adjoint_params(), dts))
        finally:
            d.close()
    (ga, pa, da), (gb, pb, db) = out
    assert np.array_equal(da, db)
    for k in range(8):
        assert np.max(np.abs(ga[k] - gb[k])) <= 1e-12 * np.max(np.abs(gb[k])), k
    assert np.max(np.abs(pa - pb)) <= 1e-12 * np.max(np.abs(pb))


def test_module_adjoint_refuses_what_it_does_not_cover():
    st = _arcade()
    resist = ("anomalous_resistivity", [("resistivity_model", "time_scale"), ("time_scale", "1e-2"),
                                        ("template_mode", "flood_fill"), ("flood_fill_threshold", "1.9"),
                                        ("smoothing_sigma", "1.0"), ("flood_fill_argmin_radius", "6e7")])
    # the resistivity's flood-fill argmin box must stay inside a non-periodic x axis
    for bc, mods in ((("reflect", "open", "fixed", "fixed"), [resist]),):
        d = GpuDomain(spec_from_files(_arcade_cfg(THETA, bc=bc, mods=mods), st))
        try:
            with pytest.raises(SpruceError):
                d.adjoint(2, 1, np.zeros((8, 24, 40)))
        finally:
            d.close()


@pytest.mark.parametrize("active", [[5], [0, 11]])
def test_parametric_heating_closed_form(active):
    """On a static uniform plasma the ideal-MHD step changes nothing, so one step of a heating
    profile with coefficients theta_k raises the thermal energy of every interior cell by
    dt * sum_k theta_k phi_k(y), phi_k(y) = exp(-(y - y_k)^2 / (2 w^2)) with centres y_k spread evenly
    from the lowest to the highest cell centre and w = basis_width * spacing, and leaves the ghost
    cells alone."""
    nx, ny, K = 12, 40, 16
    dy = 4e6 * 1.01 ** np.arange(ny)
    z = np.zeros((nx, ny))
    st = W.mhd_state(nx, ny, [5e6] * nx, dy, np.full((nx, ny), 1e-15), np.full((nx, ny), 1e6), (z, z, z),
                     (z, z, z), be=(np.full((nx, ny), 2.0), z, z))
    te0 = 2.0 * (1e-15 / W.M_PROTON) * W.K_B * 1e6 / (W.GAMMA - 1.0)
    theta = np.zeros(K)
    for k in active:
        theta[k] = 10.0 * te0 * (1 + k)
    cfg = W.base_config("ideal_mhd", ("periodic", "periodic", "fixed", "fixed"), "euler", epsilon=0.2,
                        max_iterations=1, density_min=1.0, temp_min=1.0, thermal_energy_min=1e-12,
                        modules=[("parametric_heating", [("num_basis", str(K)), ("basis_width", "1.5")] +
                                  [("coeff_%d" % k, repr(float(theta[k]))) for k in active])])
    gpu = GpuDomain(spec_from_files(cfg, st))
    try:
        before = gpu.grid("thermal_energy")
        dt = float(gpu.step(1)[0])
        after = gpu.grid("thermal_energy")
    finally:
        gpu.close()
    y = st.grids["pos_y"][0]
    lo, hi = float(np.min(st.grids["pos_y"])), float(np.max(st.grids["pos_y"]))
    sp = (hi - lo) / (K - 1)
    w = 1.5 * sp
    phi = np.exp(-0.5 * ((y[None, :] - (lo + sp * np.arange(K)[:, None])) / w) ** 2)   # [K, ny]
    want = dt * (theta @ phi)
    interior = slice(2, ny - 2)
    got = after - before
    # the difference of two thermal energies carries their rounding: 1e-13 of the largest increment
    np.testing.assert_allclose(got[:, interior], np.broadcast_to(want[interior], (nx, ny - 4)), rtol=1e-12,
                               atol=1e-13 * float(np.max(np.abs(want))))
    assert float(np.max(np.abs(want))) > 0.05 * te0
    assert np.all(got[:, :2] == 0.0) and np.all(got[:, -2:] == 0.0)


def test_parametric_heating_rejects_heights_that_vary_along_x():
    """The heating basis is a function of the height of each row; a mesh whose cell heights differ
    between columns is refused at set-up."""
    st = _arcade()
    st.grids["pos_y"] = st.grids["pos_y"] * (1.0 + 1e-3 * np.arange(st.xdim)[:, None])
    cfg = W.base_config("ideal_mhd", ("periodic", "periodic", "fixed", "open"), "rk2", epsilon=0.2,
                        max_iterations=1, density_min=1.0, temp_min=1e4, thermal_energy_min=1e-8,
                        modules=[HEAT(THETA)])
    with pytest.raises(SpruceError):
        GpuDomain(spec_from_files(cfg, st))


def test_arcade_2048_heating_gradient_matches_finite_differences():
    """The gradient workload at its full size: the 2048x2048 arcade with conduction, radiative
    losses and the 16-coefficient heating, dJ/dtheta from the adjoint against central differences
    (relative step 1e-5) on 4 coefficients, over 5 steps, a window in which the conduction subcycle
    count does not change with theta (a change makes the loss jump)."""
    from spruce_amd.adjoint import ArcadeHeatingProblem

    prob = ArcadeHeatingProblem(2048, 2048, nsteps=5, checkpoint=50)
    th = prob.theta0
    print("arcade 2048: forward", flush=True)
    _, final, _ = prob.forward(th)
    g, _, _ = prob.gradient(th, final)
    for k in (0, 5, 10, 15):
        e = 1e-5 * th[k]
        tp, tm = th.copy(), th.copy()
        tp[k] += e
        tm[k] -= e
        fd = (prob.forward(tp)[0] - prob.forward(tm)[0]) / (tp[k] - tm[k])
        rel = abs(g[k] - fd) / (abs(g[k]) + abs(fd))
        print("arcade 2048: theta_%d adjoint %.9e fd %.9e rel %.2e" % (k, g[k], fd, rel), flush=True)
        assert rel <= 1e-6, (k, g[k], fd)
