"""ideal_mhd_2E on the GPU against the CPU restatement, bit for bit, on every boundary kind and
integrator at sizes above the golden cases; ideal_mhd_cons set-up against the restatement and its
refusal to step (the reference aborts at its first time derivative)."""
import numpy as np
import pytest

from oracle.oracle import OracleDomain
from spruce_amd import workloads as W
from spruce_amd.gpu import GpuDomain, SpruceError
from spruce_amd.spec import spec_from_files

pytestmark = pytest.mark.gpu

E2_VARS = ["rho", "n", "mom_x", "mom_y", "i_thermal_energy", "e_thermal_energy", "bi_x", "bi_y", "i_temp", "e_temp",
           "i_press", "e_press", "press", "v_x", "v_y", "kinetic_energy", "b_x", "b_y", "b_mag", "b_hat_x",
           "b_hat_y", "dt"]
CONS_VARS = ["rho", "temp", "mom_x", "mom_y", "bi_x", "bi_y", "n", "v_x", "v_y", "kinetic_energy", "b_x", "b_y",
             "b_mag", "mag_energy", "b_hat_x", "b_hat_y", "thermal_energy", "press", "press_tot", "energy", "dt"]


def same_bits(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return a.shape == b.shape and (np.array_equal(a.view(np.int64), b.view(np.int64)) or np.array_equal(a, b, equal_nan=True))


def describe(a, b):
    d = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    k = np.unravel_index(np.nanargmax(d), d.shape)
    return "max rel diff %.3e at %s (%r vs %r), %d cells differ" % (d[k], k, a[k], b[k], int(np.sum(a != b)))


def spec_of(eq, nx, ny, seed, bc, integ, **kw):
    st = W.mhd_variant_state(eq, nx, ny, seed, stretch=kw.pop("stretch", False), zero_field=kw.pop("zero_field", False))
    cfg = W.base_config(eq, bc, integ, epsilon=0.2, density_min=1.0, temp_min=1.0, thermal_energy_min=1e-12, **kw)
    return spec_from_files(cfg, st)


E2_CASES = [
    ("periodic_rk2", 96, 80, ("periodic",) * 4, "rk2", {}),
    ("open_rk4", 72, 88, ("open",) * 4, "rk4", dict(open_boundary_strength=0.4, open_boundary_decay_base=0.6,
                                                    stretch=True)),
    ("reflect_euler", 64, 72, ("reflect",) * 4, "euler", {}),
    ("fixed_rk4", 80, 64, ("fixed",) * 4, "rk4", {}),
    ("ucnp_rk2", 70, 90, ("open_ucnp",) * 4, "rk2", {}),
    ("mixed_rk2_nullfield", 64, 64, ("periodic", "periodic", "fixed", "open"), "rk2",
     dict(open_boundary_strength=0.2, zero_field=True)),
]


@pytest.mark.parametrize("case", E2_CASES, ids=[c[0] for c in E2_CASES])
def test_e2_matches_oracle(case):
    name, nx, ny, bc, integ, kw = case
    print("case", name, flush=True)
    spec = spec_of("ideal_mhd_2E", nx, ny, 71, bc, integ, **dict(kw))
    gpu = GpuDomain(spec)
    ora = OracleDomain(spec)
    try:
        for v in E2_VARS:
            assert same_bits(gpu.grid(v), ora.grid(v)), "%s setup %s: %s" % (name, v, describe(gpu.grid(v), ora.grid(v)))
        rg, ro = gpu.rhs(), ora.rhs()
        for k in range(7):
            assert same_bits(rg[k], ro[k]), "%s rhs %d: %s" % (name, k, describe(rg[k], ro[k]))
        for chunk in (1, 4):
            dg, do = gpu.step(chunk), ora.step(chunk)
            assert same_bits(dg, do), "%s dt %s vs %s" % (name, dg, do)
            for v in E2_VARS:
                a, b = gpu.grid(v), ora.grid(v)
                assert same_bits(a, b), "%s %s: %s" % (name, v, describe(a, b))
    finally:
        gpu.close()
        ora.close()


@pytest.mark.parametrize("bc", [("periodic",) * 4, ("open", "reflect", "fixed", "open_ucnp")])
def test_cons_setup_matches_oracle_and_refuses_to_step(bc):
    spec = spec_of("ideal_mhd_cons", 64, 80, 72, bc, "rk2", open_boundary_strength=0.3)
    gpu = GpuDomain(spec)
    ora = OracleDomain(spec)
    try:
        for v in CONS_VARS:
            a, b = gpu.grid(v), ora.grid(v)
            assert same_bits(a, b), "%s: %s" % (v, describe(a, b))
        with pytest.raises(SpruceError):
            gpu.step(1)
        with pytest.raises(SpruceError):
            gpu.rhs()
    finally:
        gpu.close()
        ora.close()
