"""Module file output and multispecies mode in mhd.out (PlasmaDomain::storeGrids, fileio.cpp:142-200):
spruce_run against the reference executable on the same inputs.

With `output_to_file` (thermal conduction, radiative losses, physical viscosity, field heating,
anomalous resistivity) or `visc_output_*` (artificial viscosity) a module appends its fileOutput
grids to every stored frame; with `multispecies_mode` the domain appends the cumulative electron, ion
and Joule heating grids, reset after every store.  Cases whose modules take no libm power are
byte-identical to the reference's files; the others (the reference's libm pow against the GPU's
correctly rounded one) agree frame by frame within 1e-10 of each grid's largest magnitude."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from test_gpu_runner import _both, _inputs, _same_files  # noqa: E402

from spruce_amd.stateio import read_out  # noqa: E402

pytestmark = pytest.mark.gpu


def _with_opts(name, tmp, module_opts=None, **settings):
    """The case's inputs with extra options on some of its modules ({module name: [(key, value)]})."""
    import cases
    st, cfg = cases.build_case(name)
    for k, v in settings.items():
        cfg.settings[k] = str(v)
    if module_opts:
        cfg.modules = [(m, list(o) + list(module_opts.get(m, []))) for m, o in cfg.modules]
    sp, cp = os.path.join(tmp, "case.state"), os.path.join(tmp, "case.config")
    st.write(sp)
    cfg.write(cp)
    return sp, cp


EXACT = [
    # name, extra module options, extra settings
    ("ar_flood_time_euler", {"anomalous_resistivity": [("output_to_file", "true")]}, {}),
    ("visc_local_global_rk2", {"artificial_viscosity": [("visc_output_visc", "true"), ("visc_output_lap", "true"),
                                                        ("visc_output_strength", "true"),
                                                        ("visc_output_timescale", "true")]}, {}),
    ("lheat_ramp_rk2", None, {"multispecies_mode": "true"}),
    ("heat_const_euler", None, {"multispecies_mode": "true"}),
    ("sink_exp_heat_rk2", None, {"multispecies_mode": "true"}),
    ("visc_hyper_rk2", {"artificial_viscosity": [("visc_output_visc", "true"), ("visc_output_lap", "true"),
                                                 ("visc_output_strength", "true"),
                                                 ("visc_output_timescale", "true")]}, {"multispecies_mode": "true"}),
    ("sink_const_euler", {"ambient_heating_sink": [("ms_electron_heating_fraction", "0.3")]},
     {"multispecies_mode": "true"}),
    ("lheat_periodic_euler", {"localized_heating": [("ms_electron_heating_fraction", "0.6")]},
     {"multispecies_mode": "true"}),
    # the two-fluid sets have no module that adds to the cumulative grids: zeros in every frame
    ("tf_ucnp_rk2", None, {"multispecies_mode": "true"}),
    ("e2_open_rk4", None, {"multispecies_mode": "true"}),
]


@pytest.mark.parametrize("name,mods,settings", EXACT, ids=[e[0] for e in EXACT])
def test_module_output_byte_identical(name, mods, settings, tmp_path):
    tmp = str(tmp_path)
    sp, cp = _with_opts(name, tmp, mods, **settings)
    files = _same_files(_both(tmp, lambda d: ["-m", "input", "-o", d, "-s", sp, "-c", cp]))
    assert "mhd.out" in files


TOLERANT = [
    ("solar_tc_rl_rk2", {"thermal_conduction": [("output_to_file", "true"), ("ms_electron_heating_fraction", "0.3")],
                         "radiative_losses": [("output_to_file", "true"), ("ms_electron_heating_fraction", "0.7")]},
     {"multispecies_mode": "true"}),
    ("solar_tcsat_rk2", {"thermal_conduction": [("output_to_file", "true")]}, {}),
    ("pv_gc_ramp_rk2", {"physical_viscosity": [("output_to_file", "true"),
                                                    ("ms_electron_heating_fraction", "0.5")]},
     {"multispecies_mode": "true"}),
    ("fheat_rk2", {"field_heating": [("output_to_file", "true")]}, {}),
    ("ar_frob_ys_rk4", {"anomalous_resistivity": [("output_to_file", "true")]}, {"multispecies_mode": "true"}),
    ("ar_flood_masks_rk2", {"anomalous_resistivity": [("output_to_file", "true")]}, {"multispecies_mode": "true"}),
    ("visc_hyper_ellip_euler", {"artificial_viscosity": [("visc_output_visc", "true"),
                                                         ("visc_output_timescale", "true")]}, {}),
    ("visc_boundary_gc_rk4", {"artificial_viscosity": [("visc_output_lap", "true"),
                                                       ("visc_output_strength", "true"),
                                                       ("visc_output_timescale", "true")]}, {}),
    ("solar_rl_tc_rk4", {"thermal_conduction": [("output_to_file", "true"), ("ms_electron_heating_fraction", "0.5")],
                         "radiative_losses": [("output_to_file", "true"), ("ms_electron_heating_fraction", "0.0")]},
     {"multispecies_mode": "true"}),
    ("solar_rl_prevent_euler", {"radiative_losses": [("output_to_file", "true")]}, {"multispecies_mode": "true"}),
    ("pv_euler", {"physical_viscosity": [("output_to_file", "true"), ("inactive_mode", "true"),
                                         ("ms_electron_heating_fraction", "0.2")]}, {"multispecies_mode": "true"}),
    ("pv_tc_force_only", {"physical_viscosity": [("output_to_file", "true")],
                          "thermal_conduction": [("output_to_file", "true")]}, {}),
    ("heat_split_tc_rk2", {"ambient_heating": [("ms_electron_heating_fraction", "0.25")]},
     {"multispecies_mode": "true"}),
]


@pytest.mark.parametrize("name,mods,settings", TOLERANT, ids=[e[0] for e in TOLERANT])
def test_module_output_within_tolerance(name, mods, settings, tmp_path):
    tmp = str(tmp_path)
    sp, cp = _with_opts(name, tmp, mods, **settings)
    o = _both(tmp, lambda d: ["-m", "input", "-o", d, "-s", sp, "-c", cp])
    (da, ra, _, _), (db, rb, _, _) = o["ref"], o["gpu"]
    assert ra in (134, -6) and rb == 134
    a, b = read_out(os.path.join(da, "mhd.out")), read_out(os.path.join(db, "mhd.out"))
    assert len(a.frames) == len(b.frames) and len(a.frames) > 1
    for fa, fb in zip(a.frames, b.frames):
        assert list(fa) == list(fb), (list(fa), list(fb))
        for k in fa:
            x, y = np.asarray(fa[k]), np.asarray(fb[k])
            scale = max(float(np.max(np.abs(x))), 1e-300)
            assert np.max(np.abs(x - y)) <= 1e-10 * scale, (k, float(np.max(np.abs(x - y)) / scale))
