"""Synthetic input generators for the benchmark workloads and the parity cases.

Every generated input is an ordinary ``.state`` + ``.config`` pair in the reference grammar, so the
same files drive the reference executable, the CPU restatement and the GPU path.

Random numbers come from splitmix64 with ``U = (x >> 11) * 2**-53`` (one draw per cell, row-major,
x outer), which makes every input reproducible from its seed alone.
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .stateio import ConfigFile, StateFile, positions_from_sizes

K_B = 1.3807e-16           # erg/K (source/constants.hpp:8)
M_PROTON = 1.6726e-24      # g, ion mass used by the bundled states (example.state:7)
M_SR = 1.455e-22           # g (constants.hpp:17)
GAMMA = 1.6666666666666667 # examples/solar/solar_gaussian.state:11

_MASK64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive splitmix64 outputs (uint64) starting from ``seed``."""
    out = np.empty(n, dtype=np.uint64)
    x = seed & _MASK64
    for k in range(n):
        x = (x + 0x9E3779B97F4A7C15) & _MASK64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
        out[k] = z ^ (z >> 31)
    return out


def uniform(seed: int, shape: Tuple[int, ...]) -> np.ndarray:
    n = int(np.prod(shape))
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return u.reshape(shape)


def _pd_grids(dx: np.ndarray, dy: np.ndarray) -> Dict[str, np.ndarray]:
    """d_x, d_y, pos_x, pos_y as full grids from 1-D cell sizes."""
    nx, ny = len(dx), len(dy)
    d_x = np.repeat(np.asarray(dx, np.float64)[:, None], ny, axis=1)
    d_y = np.repeat(np.asarray(dy, np.float64)[None, :], nx, axis=0)
    return {"d_x": d_x, "d_y": d_y,
            "pos_x": positions_from_sizes(d_x, 0), "pos_y": positions_from_sizes(d_y, 1)}


def mhd_state(nx: int, ny: int, dx: Sequence[float], dy: Sequence[float],
              rho: np.ndarray, temp: np.ndarray, mom: Tuple[np.ndarray, np.ndarray, np.ndarray],
              bi: Tuple[np.ndarray, np.ndarray, np.ndarray],
              be: Optional[Tuple[np.ndarray, np.ndarray, np.ndarray]] = None,
              grav: Optional[Tuple[np.ndarray, np.ndarray]] = None,
              ion_mass: float = M_PROTON, comments: Sequence[str] = ()) -> StateFile:
    z = np.zeros((nx, ny))
    g = _pd_grids(np.asarray(dx, np.float64), np.asarray(dy, np.float64))
    be = be if be is not None else (z, z, z)
    grav = grav if grav is not None else (z, z)
    grids = dict(g)
    grids.update({"be_x": be[0], "be_y": be[1], "be_z": be[2],
                  "rho": rho, "temp": temp, "mom_x": mom[0], "mom_y": mom[1], "mom_z": mom[2],
                  "bi_x": bi[0], "bi_y": bi[1], "bi_z": bi[2],
                  "grav_x": grav[0], "grav_y": grav[1]})
    return StateFile(nx, ny, ion_mass, GAMMA, 0.0, grids, list(comments))


def base_config(equation_set: str = "ideal_mhd", bc: Sequence[str] = ("periodic",) * 4,
                integrator: str = "rk2", epsilon: float = 0.2, max_iterations: int = 10,
                density_min: float = 1.0, temp_min: float = 1.0, thermal_energy_min: float = 1e-12,
                open_boundary_strength: float = 0.0, open_boundary_decay_base: float = 1.0,
                output_flags: Sequence[str] = (), write_precision: int = 17,
                equation_set_options: Sequence[Tuple[str, str]] = (),
                modules: Sequence[Tuple[str, Sequence[Tuple[str, str]]]] = ()) -> ConfigFile:
    cfg = ConfigFile()
    cfg.equation_set = equation_set
    cfg.equation_set_options = list(equation_set_options)
    s = cfg.settings
    s["x_bound_1"], s["x_bound_2"], s["y_bound_1"], s["y_bound_2"] = bc
    s["epsilon"] = repr(float(epsilon))
    s["density_min"] = repr(float(density_min))
    s["temp_min"] = repr(float(temp_min))
    s["thermal_energy_min"] = repr(float(thermal_energy_min))
    s["max_iterations"] = str(int(max_iterations))
    s["iter_output_interval"] = "1"
    s["time_output_interval"] = "-1"
    s["std_out_interval"] = "-1"
    s["write_interval"] = "1"
    s["open_boundary_strength"] = repr(float(open_boundary_strength))
    s["open_boundary_decay_base"] = repr(float(open_boundary_decay_base))
    s["time_integrator"] = integrator
    s["duration"] = "1e30"
    s["write_precision"] = str(int(write_precision))
    cfg.output_flags = list(output_flags)
    cfg.modules = [(n, list(o)) for n, o in modules]
    return cfg


# --------------------------------------------------------------------------------------------
# Benchmark workloads
# --------------------------------------------------------------------------------------------

def periodic_wave_state(nx: int = 256, ny: int = 256, d: float = 1e7, seed: int = 0,
                        rho0: float = 1e-15, t0: float = 1e6, v0: float = 1e7, b0: float = 1.0) -> StateFile:
    """Plain ideal-MHD periodic workload: noisy density, sheared flow and field.

    rho = rho0 (1 + 0.05 (2U - 1)), T = t0, mom_x = -rho v0 sin(2 pi y/L), mom_y = rho v0 sin(2 pi x/L),
    bi_x = -b0 sin(2 pi y/L), bi_y = b0 sin(4 pi x/L); be = grav = mom_z = bi_z = 0; cell centres
    at (i + 1/2) d; L = n d.
    """
    Lx, Ly = nx * d, ny * d
    x = (np.arange(nx) + 0.5) * d
    y = (np.arange(ny) + 0.5) * d
    X, Y = np.meshgrid(x, y, indexing="ij")
    u = uniform(seed, (nx, ny))
    rho = rho0 * (1.0 + 0.05 * (2.0 * u - 1.0))
    temp = np.full((nx, ny), t0)
    mom_x = -rho * v0 * np.sin(2 * np.pi * Y / Ly)
    mom_y = rho * v0 * np.sin(2 * np.pi * X / Lx)
    z = np.zeros((nx, ny))
    bi_x = -b0 * np.sin(2 * np.pi * Y / Ly)
    bi_y = b0 * np.sin(4 * np.pi * X / Lx)
    return mhd_state(nx, ny, [d] * nx, [d] * ny, rho, temp, (mom_x, mom_y, z), (bi_x, bi_y, z),
                     comments=["#periodic ideal-MHD wave, seed %d" % seed])


def periodic_wave_config(steps: int = 1000, integrator: str = "rk2", **kw) -> ConfigFile:
    return base_config("ideal_mhd", ("periodic",) * 4, integrator, epsilon=0.2, max_iterations=steps,
                       density_min=1.0, temp_min=1.0, thermal_energy_min=1e-12, **kw)


def arcade_state(nx: int = 512, ny: int = 1024, dx: float = 5e6, dy0: float = 5e6, stretch: float = 1.002,
                 b_foot: float = 10.0, t_base: float = 2e4, t_corona: float = 1e6, n_base: float = 3e11,
                 grav_y: float = -2.74e4, depth: Optional[float] = None) -> StateFile:
    """Solar arcade: stretched-y grid, potential line-dipole field, stratified hydrostatic atmosphere.

    The external field is B_x - i B_y = -b_foot h^2 / (z - z0)^2 with z0 = x_c - i h (a line dipole at
    depth h below the mid-point of the base), which is curl- and divergence-free and has |B| = b_foot
    at the footpoint above the dipole.  T(y) rises from t_base to t_corona through a tanh transition;
    the density follows dp/dy = -rho g with p = 2 n k_B T, integrated cell to cell.
    """
    dys = dy0 * stretch ** np.arange(ny)
    x = (np.arange(nx) + 0.5) * dx
    y = np.cumsum(dys) - 0.5 * dys
    Lx = nx * dx
    h = depth if depth is not None else 0.25 * Lx
    X, Y = np.meshgrid(x, y, indexing="ij")
    zz = (X - 0.5 * Lx) + 1j * (Y + h)
    bc = -b_foot * h * h / zz ** 2
    be_x, be_y = bc.real, -bc.imag
    y_tr, w = 2.0e8, 4.0e7
    T1 = t_base + (t_corona - t_base) * 0.5 * (1.0 + np.tanh((y - y_tr) / w))
    g = -grav_y
    lnp = np.zeros(ny)
    for j in range(1, ny):
        dist = y[j] - y[j - 1]
        lnp[j] = lnp[j - 1] - M_PROTON * g * dist / (2 * K_B) * 0.5 * (1.0 / T1[j - 1] + 1.0 / T1[j])
    p = 2 * n_base * K_B * t_base * np.exp(lnp)
    n = p / (2 * K_B * T1)
    rho = np.repeat((n * M_PROTON)[None, :], nx, axis=0)
    temp = np.repeat(T1[None, :], nx, axis=0)
    z = np.zeros((nx, ny))
    return mhd_state(nx, ny, [dx] * nx, dys, rho, temp, (z, z, z), (z, z, z),
                     be=(be_x, be_y, z), grav=(z, np.full((nx, ny), grav_y)),
                     comments=["#solar arcade %dx%d" % (nx, ny)])


def arcade_config(steps: int = 2000, integrator: str = "rk2", conduction: bool = True, radiation: bool = True,
                  **kw) -> ConfigFile:
    mods = []
    if conduction:
        mods.append(("thermal_conduction", [("flux_saturation", "false"), ("epsilon", "0.5"),
                                            ("dt_subcycle_min", "1e-6"), ("time_integrator", "euler")]))
    if radiation:
        mods.append(("radiative_losses", [("cutoff_ramp", "2e4"), ("cutoff_temp", "3e4"),
                                          ("epsilon", "0.5"), ("time_integrator", "euler")]))
    return base_config("ideal_mhd", ("periodic", "periodic", "fixed", "open"), integrator, epsilon=0.2,
                       max_iterations=steps, density_min=1.0, temp_min=1e4, thermal_energy_min=1e-8,
                       modules=mods, **kw)


M_ELECTRON = 9.1094e-28  # g (constants.hpp:9)


def ucnp_state(nx: int, ny: int, n_peak: float = 5e9, sigma: float = 0.1, te: float = 100.0, ti: float = 1.0,
               half_width_sigmas: float = 4.0, n_floor_frac: float = 1e-4, seed: Optional[int] = None,
               v_noise: float = 0.0, e_noise: float = 0.0) -> StateFile:
    """Two-fluid Gaussian plasma (ideal_2F state variables, ideal2F.hpp:36-38): ions of mass M_SR.

    n = n_peak exp(-r^2 / (2 sigma^2)) + n_floor on a uniform mesh spanning +-half_width_sigmas*sigma;
    electrons at temperature te, ions at ti; fields and momenta zero unless noise amplitudes are given
    (seeded), which makes every term of the equations non-zero in tests.
    """
    L = 2 * half_width_sigmas * sigma
    dx, dy = L / nx, L / ny
    x = (np.arange(nx) + 0.5) * dx - 0.5 * L
    y = (np.arange(ny) + 0.5) * dy - 0.5 * L
    X, Y = np.meshgrid(x, y, indexing="ij")
    n = n_peak * np.exp(-(X * X + Y * Y) / (2 * sigma * sigma)) + n_floor_frac * n_peak
    z = np.zeros((nx, ny))
    g = _pd_grids(np.full(nx, dx), np.full(ny, dy))
    grids = dict(g)
    grids.update({"be_x": z, "be_y": z, "be_z": z, "i_rho": n * M_SR, "e_rho": n * M_ELECTRON,
                  "i_mom_x": z, "i_mom_y": z, "e_mom_x": z, "e_mom_y": z,
                  "i_temp": np.full((nx, ny), ti), "e_temp": np.full((nx, ny), te),
                  "bi_x": z, "bi_y": z, "bi_z": z, "E_x": z, "E_y": z, "E_z": z, "grav_x": z, "grav_y": z})
    if seed is not None:
        u = [uniform(seed * 16 + k, (nx, ny)) for k in range(9)]
        if v_noise:
            grids["i_mom_x"] = n * M_SR * v_noise * 1e-2 * (2 * u[0] - 1)
            grids["i_mom_y"] = n * M_SR * v_noise * 1e-2 * (2 * u[1] - 1)
            grids["e_mom_x"] = n * M_ELECTRON * v_noise * (2 * u[2] - 1)
            grids["e_mom_y"] = n * M_ELECTRON * v_noise * (2 * u[3] - 1)
        if e_noise:
            for k, name in enumerate(("E_x", "E_y", "E_z", "bi_x", "bi_y", "bi_z")):
                grids[name] = e_noise * (2 * u[4 + k % 5] - 1) * (1.0 + 0.1 * k)
    return StateFile(nx, ny, M_SR, GAMMA, 0.0, grids, ["#two-fluid Gaussian plasma"])


def ucnp_config(steps: int = 5000, integrator: str = "rk2", remove_curl_terms: bool = True,
                bc: Sequence[str] = ("open_ucnp",) * 4, **kw) -> ConfigFile:
    return base_config("ideal_2F", bc, integrator, epsilon=0.2, max_iterations=steps, density_min=1e2,
                       temp_min=1e-3, thermal_energy_min=1e-30,
                       equation_set_options=[("use_sub_cycling", "false"),
                                             ("remove_curl_terms", "true" if remove_curl_terms else "false")], **kw)


BATCH_EDGES = (128, 192, 256, 384, 512)


def batch_2f_members(n: int = 64, seed: int = 7, integrator: str = "rk2", steps: int = 5000):
    """Independent two-fluid plasma expansions of uneven sizes (the batched-mode workload).

    Member k: an edge drawn from BATCH_EDGES, Gaussian peak density 1e9-1e10 cm^-3, width
    0.05-0.15 cm, electron temperature 50-200 K, ion temperature 1 K, zero fields and momenta, all
    drawn from a seeded stream.  Returns [(StateFile, ConfigFile)].
    """
    u = uniform(seed, (n, 4))
    out = []
    for k in range(n):
        e = BATCH_EDGES[min(len(BATCH_EDGES) - 1, int(u[k, 0] * len(BATCH_EDGES)))]
        st = ucnp_state(e, e, n_peak=1e9 + 9e9 * u[k, 1], sigma=0.05 + 0.1 * u[k, 2], te=50.0 + 150.0 * u[k, 3],
                        ti=1.0)
        out.append((st, ucnp_config(steps=steps, integrator=integrator)))
    return out


def mhd_variant_state(equation_set: str, nx: int, ny: int, seed: int, stretch: bool = False,
                      zero_field: bool = False) -> StateFile:
    """Perturbed uniform plasma in the state variables of ideal_mhd_2E (idealmhd2E.hpp:27-29) or
    ideal_mhd_cons (idealmhdcons.hpp:27-29); every evolved field carries seeded noise.

    ideal_mhd_cons reads a press_tot grid it never computes, so the state supplies it: the thermal
    plus magnetic pressure of the initial state.
    """
    u = [uniform(seed * 16 + k, (nx, ny)) for k in range(10)]
    dx = 1e7 * (1.01 ** np.arange(nx)) if stretch else np.full(nx, 1e7)
    dy = 8e6 * (1.02 ** np.arange(ny)) if stretch else np.full(ny, 8e6)
    rho = 1e-15 * (1.0 + 0.3 * (2 * u[0] - 1))
    t_i = 1e6 * (1.0 + 0.3 * (2 * u[1] - 1))
    t_e = 1.5e6 * (1.0 + 0.3 * (2 * u[9] - 1))
    mom_x = rho * 1e7 * (2 * u[2] - 1)
    mom_y = rho * 1e7 * (2 * u[3] - 1)
    z = np.zeros((nx, ny))
    bi_x = z if zero_field else 1.0 * (2 * u[5] - 1)
    bi_y = z if zero_field else 1.0 * (2 * u[6] - 1)
    be_x = z if zero_field else 2.0 + 0.5 * (2 * u[7] - 1)
    be_y = z if zero_field else 1.0 + 0.5 * (2 * u[8] - 1)
    grids = dict(_pd_grids(dx, dy))
    grids.update({"be_x": be_x, "be_y": be_y, "be_z": z, "rho": rho, "mom_x": mom_x, "mom_y": mom_y,
                  "bi_x": bi_x, "bi_y": bi_y, "grav_x": z, "grav_y": np.full((nx, ny), -2.74e4)})
    if equation_set == "ideal_mhd_2E":
        grids.update({"i_temp": t_i, "e_temp": t_e})
    elif equation_set == "ideal_mhd_cons":
        n = rho / M_PROTON
        b2 = (be_x + bi_x) ** 2 + (be_y + bi_y) ** 2
        grids.update({"temp": t_i, "press_tot": 2 * n * K_B * t_i + b2 / (8 * np.pi)})
    else:
        raise ValueError(equation_set)
    return StateFile(nx, ny, M_PROTON, GAMMA, 0.0, grids, ["#random %s case, seed %d" % (equation_set, seed)])
