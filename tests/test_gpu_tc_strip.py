"""The wavefront-strip form of the conduction update against the one-thread-per-cell form and the
CPU restatement (correctly rounded powers), bit for bit.

The strip form shares every face value between the two cells of the face and forms the b_hat
derivatives itself; these cases cover the lane/row edges of that sharing: strips of 1, 3 and 16 rows
(a last strip shorter than the others), column counts below and above one 62-column tile and not a
multiple of it, periodic and closed sides on either axis, stretched meshes, each subcycle integrator
(Euler, RK2 and RK4 use the separate-base and stored-derivative paths) and flux saturation."""
import numpy as np
import pytest

from oracle.oracle import OracleDomain, set_pow_mode
from spruce_amd import workloads as W
from spruce_amd.gpu import GpuDomain
from spruce_amd.spec import spec_from_files

pytestmark = pytest.mark.gpu

VARS = ["rho", "temp", "mom_x", "mom_y", "thermal_energy", "bi_x", "bi_y", "dt"]


def same_bits(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return a.shape == b.shape and (np.array_equal(a.view(np.int64), b.view(np.int64)) or np.array_equal(a, b, equal_nan=True))


def describe(a, b):
    d = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    k = np.unravel_index(np.nanargmax(d), d.shape)
    return "max rel diff %.3e at %s, %d cells differ" % (d[k], k, int(np.sum(a != b)))


def conduction_case(nx, ny, bc, integ, saturation, seed):
    """A stretched arcade (dipole b_hat varying along both axes, stratified temperature) with 1%
    noise in the temperature and a random momentum, under the given sides."""
    st = W.arcade_state(nx, ny, dx=5e6, dy0=5e6, stretch=1.01, t_corona=2e6)
    u = W.uniform(seed, (nx, ny))
    st.grids["temp"] = st.grids["temp"] * (1.0 + 0.01 * (2 * u - 1))
    st.grids["mom_x"] = st.grids["rho"] * 3e6 * (2 * W.uniform(seed + 1, (nx, ny)) - 1)
    mods = [("thermal_conduction", [("flux_saturation", "true" if saturation else "false"), ("epsilon", "0.5"),
                                    ("dt_subcycle_min", "1e-9"), ("time_integrator", integ)])]
    cfg = W.base_config("ideal_mhd", bc, "rk2", epsilon=0.2, max_iterations=4, density_min=1.0, temp_min=1e4,
                        thermal_energy_min=1e-8, modules=mods)
    return spec_from_files(cfg, st)


CASES = [
    # nx, ny, sides (x1, x2, y1, y2), subcycle integrator, saturation
    (40, 70, ("periodic", "periodic", "fixed", "open"), "euler", False),
    (37, 133, ("periodic", "periodic", "fixed", "open"), "rk2", False),
    (45, 48, ("open", "fixed", "periodic", "periodic"), "euler", False),
    (33, 125, ("fixed", "open", "reflect", "open"), "rk4", False),
    (26, 64, ("periodic", "periodic", "periodic", "periodic"), "euler", True),
    (41, 90, ("open", "open", "fixed", "open"), "rk2", True),
]


@pytest.mark.parametrize("nx,ny,bc,integ,sat", CASES)
def test_strip_matches_cell_form(nx, ny, bc, integ, sat):
    spec = conduction_case(nx, ny, bc, integ, sat, seed=nx * ny)
    print("tc strip vs cell", nx, ny, bc, integ, sat, flush=True)
    ref = GpuDomain(spec, options={"tc_kernel": "cell"})
    try:
        dts_ref, sub_ref = ref.step(3, subcycles=True)
        want = {v: ref.grid(v) for v in VARS}
    finally:
        ref.close()
    for rows in (1, 3, 16):
        g = GpuDomain(spec, options={"tc_kernel": "strip", "tc_rows": rows})
        try:
            dts, sub = g.step(3, subcycles=True)
            assert np.array_equal(sub, sub_ref), (rows, sub, sub_ref)
            assert same_bits(dts, dts_ref), (rows, dts, dts_ref)
            for v in VARS:
                got = g.grid(v)
                assert same_bits(got, want[v]), "rows %d %s: %s" % (rows, v, describe(got, want[v]))
        finally:
            g.close()


@pytest.fixture
def cr_oracle():
    set_pow_mode(1)
    yield
    set_pow_mode(0)


@pytest.mark.parametrize("nx,ny,bc,integ,sat", [CASES[1], CASES[3], CASES[5]])
def test_strip_matches_cr_restatement(nx, ny, bc, integ, sat, cr_oracle):
    spec = conduction_case(nx, ny, bc, integ, sat, seed=nx + ny)
    print("tc strip vs restatement", nx, ny, bc, integ, sat, flush=True)
    g = GpuDomain(spec, options={"tc_kernel": "strip"})
    o = OracleDomain(spec)
    try:
        for s in range(2):
            dg, sg = g.step(1, subcycles=True)
            do = o.step(1)
            assert list(sg[0]) == o.subcycles(), (s, sg[0], o.subcycles())
            assert same_bits(dg, do), (s, dg, do)
        for v in VARS:
            a, b = g.grid(v), o.grid(v)
            assert same_bits(a, b), "%s: %s" % (v, describe(a, b))
    finally:
        g.close()
        o.close()
