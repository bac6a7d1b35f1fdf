"""spruce_gpu_rhs_noalloc (the right-hand side staged through the library's static device array,
no device allocation) against spruce_gpu_rhs, bit for bit: two calls in a row on one GPU, the first
on a grid near the array's limit (64,000 of 65,536 cells, open_moc sides and an artificial-viscosity
term, so the module hooks write into the staged grids too), the second on a smaller periodic grid
that leaves most of the array as the first call left it; and a grid over the limit is refused."""
import numpy as np
import pytest

from spruce_amd import workloads as W
from spruce_amd.gpu import GpuDomain, SpruceError
from spruce_amd.spec import spec_from_files

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def _big_spec():
    from tests.golden.cases import random_mhd_state
    st = random_mhd_state(200, 320, 17, stretch=False)
    visc = ("artificial_viscosity", [("visc_opt", "local"), ("visc_strength", "0.2"), ("visc_vars_to_diff", "v_x"),
                                     ("visc_vars_to_evol", "mom_x"), ("visc_length", "0"), ("visc_species", "i")])
    cfg = W.base_config("ideal_mhd", ("periodic", "periodic", "open_moc", "fixed"), "rk2", epsilon=0.2,
                        density_min=1.0, temp_min=1.0, thermal_energy_min=1e-12, modules=[visc])
    return spec_from_files(cfg, st)


def _small_spec():
    return spec_from_files(W.periodic_wave_config(steps=1), W.periodic_wave_state(40, 56, seed=5))


def test_rhs_noalloc_matches_rhs_on_two_calls():
    for k, spec in enumerate((_big_spec(), _small_spec())):
        print("rhs_noalloc call %d: %dx%d" % (k, spec.xdim, spec.ydim), flush=True)
        d = GpuDomain(spec)
        try:
            d.step(2)
            want = d.rhs()
            got = d.rhs_noalloc()
            assert same_bits(got, want), "call %d: %d values differ" % (k, int(np.sum(got != want)))
            assert same_bits(d.rhs_noalloc(), want)
        finally:
            d.close()


def test_rhs_noalloc_refuses_grids_over_its_limit():
    spec = spec_from_files(W.periodic_wave_config(steps=1), W.periodic_wave_state(256, 257, seed=6))
    d = GpuDomain(spec)
    try:
        with pytest.raises(SpruceError):
            d.rhs_noalloc()
        d.rhs()
    finally:
        d.close()
