"""GPU discrete adjoint of the IdealMHD step against central finite differences of the GPU forward
(the forward itself is bit-identical to the reference), and through a torch.autograd.Function
harness.  The loss is piecewise smooth (limiter and min/max branches), so the directional checks
use seeded random directions over many cells; agreement is stated per check."""
import numpy as np
import pytest

from spruce_amd import workloads as W
from spruce_amd.adjoint import PeriodicWaveProblem
from spruce_amd.gpu import GpuDomain, SpruceError
from spruce_amd.spec import spec_from_files

pytestmark = pytest.mark.gpu

STATE = ["rho", "temp", "mom_x", "mom_y", "mom_z", "bi_x", "bi_y", "bi_z"]


def _spec_from(grids, base, cfg):
    from spruce_amd.stateio import StateFile
    return spec_from_files(cfg, StateFile(base.xdim, base.ydim, base.ion_mass, base.adiabatic_index, base.t, grids,
                                          list(base.comments)))


def _loss(grids, base, cfg, nsteps, w):
    d = GpuDomain(_spec_from(grids, base, cfg))
    try:
        d.step(nsteps)
        out = [d.grid(v) for v in ("rho", "mom_x", "thermal_energy", "bi_y")]
    finally:
        d.close()
    return sum(float(np.sum(wk * o)) for wk, o in zip(w, out))


@pytest.mark.parametrize("integ,nsteps,ckpt", [("rk2", 24, 5), ("euler", 16, 16), ("rk2", 9, 1), ("rk4", 12, 5)])
def test_adjoint_matches_finite_differences(integ, nsteps, ckpt):
    """dJ/dstate . d against central differences at two step sizes (relative 1e-7 and 1e-8 along
    seeded random directions).  The loss is only piecewise smooth: the limiter's min/max selections
    switch where two compared values cross, and a direction whose two differences disagree has
    crossed such a switch inside the difference interval, where no single derivative describes it.
    Those directions are counted and left out; at least four of six must remain."""
    print("adjoint", integ, nsteps, ckpt, flush=True)
    from tests.golden.cases import random_mhd_state
    base = random_mhd_state(40, 48, 91, stretch=True)
    cfg = W.base_config("ideal_mhd", ("periodic",) * 4, integ, epsilon=0.2, density_min=1.0, temp_min=1.0,
                        thermal_energy_min=1e-12)
    rng = np.random.default_rng(5)
    w = [rng.uniform(-1, 1, (40, 48)) * s for s in (1e15, 1e8, 10.0, 1.0)]
    d = GpuDomain(_spec_from(dict(base.grids), base, cfg))
    try:
        seed = np.zeros((8, 40, 48))
        seed[0], seed[1], seed[4], seed[6] = w
        g_ev, _ = d.adjoint(nsteps, ckpt, seed)
        g = d.adjoint_state(g_ev)
    finally:
        d.close()
    scale = {"rho": 1e-15, "temp": 1e6, "mom_x": 1e-8, "mom_y": 1e-8, "mom_z": 1e-8, "bi_x": 1.0, "bi_y": 1.0,
             "bi_z": 1.0}
    smooth, worst = 0, 0.0
    for trial in range(6):
        dirs = {v: rng.uniform(-1, 1, (40, 48)) * scale[v] for v in STATE}
        ad = sum(float(np.sum(g[k] * dirs[v])) for k, v in enumerate(STATE))
        fds = []
        for eps in (1e-7, 1e-8):
            gp = {v: base.grids[v] + eps * dirs[v] if v in dirs else base.grids[v] for v in base.grids}
            gm = {v: base.grids[v] - eps * dirs[v] if v in dirs else base.grids[v] for v in base.grids}
            fds.append((_loss(gp, base, cfg, nsteps, w) - _loss(gm, base, cfg, nsteps, w)) / (2 * eps))
        consistent = abs(fds[0] - fds[1]) <= 1e-6 * abs(fds[0])
        rel = abs(ad - fds[0]) / (abs(ad) + abs(fds[0]))
        print("direction", trial, "adjoint %.12e fd(1e-7) %.12e fd(1e-8) %.12e rel %.2e%s"
              % (ad, fds[0], fds[1], rel, "" if consistent else " (branch crossed)"), flush=True)
        if consistent:
            smooth += 1
            worst = max(worst, rel)
    assert smooth >= 4
    assert worst < 1e-6


@pytest.mark.parametrize("integ", ["rk2", "rk4"])
def test_adjoint_leaves_the_same_final_state_as_stepping(integ):
    base = W.periodic_wave_state(32, 40, seed=2)
    cfg = W.periodic_wave_config(steps=10, integrator=integ)
    a = GpuDomain(spec_from_files(cfg, base))
    b = GpuDomain(spec_from_files(cfg, base))
    try:
        dta = a.step(10)
        _, dtb = b.adjoint(10, 3, np.zeros((8, 32, 40)))
        assert np.array_equal(dta, dtb)
        for v in ("rho", "mom_x", "thermal_energy", "bi_y"):
            assert np.array_equal(a.grid(v), b.grid(v))
    finally:
        a.close()
        b.close()


def test_adjoint_refuses_what_it_does_not_cover():
    from tests.golden.cases import random_mhd_state
    st = random_mhd_state(24, 20, 4)
    cfg = W.base_config("ideal_mhd", ("open",) * 4, "rk2", epsilon=0.2)
    d = GpuDomain(spec_from_files(cfg, st))
    try:
        with pytest.raises(SpruceError):
            d.adjoint(2, 1, np.zeros((8, 24, 20)))
    finally:
        d.close()


def test_periodic_wave_gradient_256_through_torch_harness():
    """The 256x256 periodic wave, 500 steps, checkpoint every 50: gradient of the final density
    misfit with respect to the initial temperature grid and v0, against central differences
    (h = 1e-6 relative) on 32 seeded unit probe directions.

    The initial state has uniform temperature, so the sound speed does not depend on the density
    noise, and |v_x| = |v_y| = v0 cos(pi/256) at the 16 cells next to the four velocity extrema:
    their step sizes tie exactly at step 0 and the loss has a kink there (the step is a minimum over
    cells).  A probe that moves one of those cells' temperature sees different forward and backward
    differences; no gradient reproduces their mean.  For such probes (detected by forward and
    backward differences that disagree) the adjoint must equal one of the two one-sided
    derivatives; every other probe must match the central difference to 1e-6 relative.  Both
    comparisons carry an absolute floor of 1e-9 of the gradient's norm (in relative parameter
    units), the size of the differences' rounding noise (|J| * 2^-52 / h is about 5e-37 here), which
    matters only for probes whose derivative is a thousand times smaller than typical."""
    torch = pytest.importorskip("torch")
    from torch_harness import PeriodicWaveLoss
    prob = PeriodicWaveProblem(256, 256, nsteps=500, checkpoint=50)
    temp = torch.tensor(prob.temp0, dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor(prob.v0, dtype=torch.float64, requires_grad=True)
    loss = PeriodicWaveLoss.apply(temp, v0, prob)
    loss.backward()
    g_temp = temp.grad.numpy()
    g_v0 = float(v0.grad)
    print("cost ratio", PeriodicWaveLoss.last_ratio, flush=True)
    j0, _, _ = prob.forward(prob.temp0, prob.v0)
    floor = 1e-9 * float(np.sqrt(np.sum((g_temp * prob.temp0) ** 2) + (g_v0 * prob.v0) ** 2))
    rng = np.random.default_rng(32)
    worst, kinks = 0.0, 0
    for k in range(32):
        u = rng.standard_normal(257)
        u /= np.linalg.norm(u)
        dt = np.zeros((256, 256))
        cells = rng.integers(0, 256 * 256, 256)
        dt.reshape(-1)[cells] = u[:256]
        dv = u[256]
        h = 1e-6
        ad = float(np.sum(g_temp * dt * prob.temp0)) + g_v0 * dv * prob.v0
        lp, _, _ = prob.forward(prob.temp0 * (1 + h * dt), prob.v0 * (1 + h * dv))
        lm, _, _ = prob.forward(prob.temp0 * (1 - h * dt), prob.v0 * (1 - h * dv))
        fwd, bwd, fd = (lp - j0) / h, (j0 - lm) / h, (lp - lm) / (2 * h)
        if abs(fwd - bwd) <= 1e-5 * (abs(fwd) + abs(bwd)) + 4 * floor:   # one-sided noise is twice the central
            rel = abs(ad - fd) / (abs(ad) + abs(fd) + floor / 1e-6)
            print("probe", k, "adjoint %.9e fd %.9e rel %.2e" % (ad, fd, rel), flush=True)
            worst = max(worst, rel)
        else:
            kinks += 1
            rel = min(abs(ad - fwd) / (abs(ad) + abs(fwd)), abs(ad - bwd) / (abs(ad) + abs(bwd)))
            print("probe", k, "kink: adjoint %.9e forward %.9e backward %.9e rel %.2e" % (ad, fwd, bwd, rel), flush=True)
            assert rel < 1e-5
    assert kinks <= 2
    assert worst < 1e-6
