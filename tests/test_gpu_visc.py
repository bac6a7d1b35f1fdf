"""GPU parity for artificial viscosity (module "artificial_viscosity", viscosity.cpp): RHS terms of
strength <= 1 fused into the stage kernels, hyperviscous subcycling, boundary strength profiles;
and for ambient heating (ambientheating.cpp), the post-integrator heating injection.
No libm power is involved, so every case without thermal conduction is bit-identical to the
reference golden vectors; the case with conduction is bit-identical to the restatement run with
correctly rounded powers and within 1e-10 of the reference."""
import numpy as np
import pytest

from golden_util import golden_names, load_golden
from oracle.oracle import OracleDomain, set_pow_mode
from spruce_amd import workloads as W
from spruce_amd.gpu import GpuDomain
from spruce_amd.spec import spec_from_files

pytestmark = pytest.mark.gpu

VISC = [c for c in golden_names() if c.startswith(("visc_", "heat_")) and "_tc_" not in c]
# cases through a libm power in the reference: conduction, physical viscosity (T^2.5)
VISC_TC = [c for c in golden_names() if (c.startswith(("visc_", "heat_")) and "_tc_" in c) or c.startswith("pv_")]
VARS = ["rho", "temp", "mom_x", "mom_y", "mom_z", "bi_x", "bi_y", "bi_z", "n", "thermal_energy", "press", "dt"]
KERNELS = ["cell", "strip", "tile"]


def same_bits(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return a.shape == b.shape and (np.array_equal(a.view(np.int64), b.view(np.int64)) or np.array_equal(a, b, equal_nan=True))


def describe(a, b):
    d = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    k = np.unravel_index(np.nanargmax(d), d.shape)
    return "max rel diff %.3e at %s (%r vs %r), %d cells differ" % (d[k], k, a[k], b[k], int(np.sum(a != b)))


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("case", VISC)
def test_viscosity_matches_reference_golden(case, kernel):
    print("case", case, kernel, flush=True)
    g = load_golden(case)
    dom = GpuDomain(g.spec(), options={"stage_kernel": kernel})
    try:
        step = 0
        for k in g.frames[1:]:
            dts = dom.step(k - step)
            assert same_bits(dts, g.dt_steps[step:k]), "%s dt %s vs %s" % (case, dts, g.dt_steps[step:k])
            step = k
            for name, ref in g.frame(k).items():
                got = dom.grid(name)
                assert same_bits(got, ref), "%s frame %d %s: %s" % (case, k, name, describe(got, ref))
    finally:
        dom.close()


@pytest.mark.parametrize("case", VISC_TC)
def test_viscosity_with_conduction(case):
    print("case", case, flush=True)
    g = load_golden(case)
    spec = g.spec()
    set_pow_mode(1)
    gpu = GpuDomain(spec)
    ora = OracleDomain(spec)
    try:
        dg = gpu.step(g.nsteps)
        do = ora.step(g.nsteps)
        assert same_bits(dg, do)
        for v in VARS:
            assert same_bits(gpu.grid(v), ora.grid(v)), v
    finally:
        set_pow_mode(0)
        gpu.close()
        ora.close()
    gpu = GpuDomain(spec)
    try:
        dts = gpu.step(g.nsteps)
        assert np.max(np.abs(dts / g.dt_steps - 1)) < 1e-12
        for name, ref in g.frame(g.frames[-1]).items():
            got = gpu.grid(name)
            scale = np.max(np.abs(ref))
            err = np.max(np.abs(got - ref)) / scale if scale > 0 else np.max(np.abs(got - ref))
            assert err < 1e-10, "%s %s rel L-inf %.3e" % (case, name, err)
    finally:
        gpu.close()


def _visc_spec(nx, ny, seed, bc, integ, terms, **extra):
    from tests.golden.cases import random_mhd_state
    st = random_mhd_state(nx, ny, seed, stretch=extra.pop("stretch", False))
    opt = [("visc_opt", terms[0]), ("visc_strength", terms[1]), ("visc_vars_to_diff", terms[2]),
           ("visc_vars_to_evol", terms[3]), ("visc_length", terms[4]), ("visc_species", terms[5])]
    opt += [(k, v) for k, v in extra.pop("visc", {}).items()]
    cfg = W.base_config("ideal_mhd", bc, integ, epsilon=0.2, density_min=1.0, temp_min=1.0,
                        thermal_energy_min=1e-12, modules=[("artificial_viscosity", opt)], **extra)
    return spec_from_files(cfg, st)


BIG = [
    ("rhs_terms_rk4", 96, 80, ("periodic",) * 4, "rk4",
     ("local,global,boundary", "0.2,0.1,0.3", "v_x,temp,b_y", "mom_x,thermal_energy,bi_y", "0,0,6e7", "i,i,i"),
     dict(visc={"gradient_correction": "true"})),
    ("hyper_rk4_open", 72, 88, ("open",) * 4, "rk2",
     ("boundary_global,local", "2.2,0.05", "rho,v_z", "rho,mom_z", "9e7,0", "i,i"),
     dict(visc={"hv_time_integrator": "rk4", "hv_epsilon": "0.8", "boundary_falloff_shape": "exp_elliptical"},
          open_boundary_strength=0.3, stretch=True)),
    ("hyper_euler_fixed", 64, 64, ("fixed", "reflect", "open_ucnp", "open"), "euler",
     ("local", "3.0", "mom_y", "mom_y", "0", "i"), dict(visc={"hv_time_integrator": "euler"})),
]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("case", BIG, ids=[c[0] for c in BIG])
def test_viscosity_matches_oracle(case, kernel):
    name, nx, ny, bc, integ, terms, extra = case
    print("case", name, kernel, flush=True)
    spec = _visc_spec(nx, ny, 61, bc, integ, terms, **dict(extra))
    gpu = GpuDomain(spec, options={"stage_kernel": kernel})
    ora = OracleDomain(spec)
    try:
        rg, ro = gpu.rhs(), ora.rhs()
        for k in range(8):
            assert same_bits(rg[k], ro[k]), "%s rhs %d: %s" % (name, k, describe(rg[k], ro[k]))
        for chunk in (1, 2):
            dg, sg = gpu.step(chunk, subcycles=True)
            do = ora.step(chunk)
            assert same_bits(dg, do), "%s dt %s vs %s" % (name, dg, do)
            assert list(sg[-1]) == ora.subcycles()
            for v in VARS:
                a, b = gpu.grid(v), ora.grid(v)
                assert same_bits(a, b), "%s %s: %s" % (name, v, describe(a, b))
    finally:
        gpu.close()
        ora.close()


PV_BIG = [
    ("pv_euler_periodic", ("periodic",) * 4, "rk2", [("coeff", "2e-16"), ("epsilon", "0.3")]),
    ("pv_rk2_gc_ramp_open", ("open",) * 4, "rk4", [("coeff", "3e-16"), ("ramp_length", "2e8"),
                                                   ("gradient_correction", "true"), ("time_integrator", "rk2")]),
    ("pv_force_only_inactive_heat", ("fixed", "reflect", "periodic", "periodic"), "euler",
     [("coeff", "2e-16"), ("heating_on", "false")]),
]


@pytest.mark.parametrize("case", PV_BIG, ids=[c[0] for c in PV_BIG])
def test_physical_viscosity_matches_cr_oracle(case):
    from tests.golden.cases import random_mhd_state
    name, bc, integ, opts = case
    print("case", name, flush=True)
    st = random_mhd_state(72, 64, 81, stretch=True)
    cfg = W.base_config("ideal_mhd", bc, integ, epsilon=0.2, density_min=1.0, temp_min=1.0, thermal_energy_min=1e-12,
                        open_boundary_strength=0.3, modules=[("physical_viscosity", opts)])
    spec = spec_from_files(cfg, st)
    set_pow_mode(1)
    gpu = GpuDomain(spec)
    ora = OracleDomain(spec)
    try:
        for _ in range(3):
            dg, sg = gpu.step(1, subcycles=True)
            do = ora.step(1)
            assert same_bits(dg, do), "%s dt %s vs %s" % (name, dg, do)
            assert list(sg[-1]) == ora.subcycles(), "%s subcycles %s vs %s" % (name, sg[-1], ora.subcycles())
        for v in VARS:
            a, b = gpu.grid(v), ora.grid(v)
            assert same_bits(a, b), "%s %s: %s" % (name, v, describe(a, b))
    finally:
        set_pow_mode(0)
        gpu.close()
        ora.close()
