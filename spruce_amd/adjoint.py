"""Gradient of a misfit of the final state through many IdealMHD steps, by the GPU discrete adjoint.

``ArcadeHeatingProblem`` is the gradient workload with modules: the solar arcade with thermal
conduction and radiative losses, heated by ``parametric_heating`` (heating(y) = sum_k theta_k
phi_k(y)); the loss is the coronal temperature misfit and the gradient is taken with respect to the
coefficients theta_k (``spruce_gpu_adjoint_params``).

``PeriodicWaveProblem`` is the adjoint workload on the periodic ideal-MHD wave: the loss is
``sum((rho_final - rho_target)^2)`` after ``nsteps`` steps, and its gradient is taken with respect to
the initial temperature grid and the flow amplitude v0 (mom_x = -rho v0 sin(2 pi y/L),
mom_y = rho v0 sin(2 pi x/L)).  The gradient comes from ``spruce_gpu_adjoint`` (hand-written reverse
of every step, checkpointed) and ``spruce_gpu_adjoint_state`` (reverse of the set-up).
"""
from __future__ import annotations

import time
from typing import Tuple

import numpy as np

from . import workloads as W
from .gpu import GpuDomain
from .spec import spec_from_files


class PeriodicWaveProblem:
    def __init__(self, nx: int = 256, ny: int = 256, nsteps: int = 500, checkpoint: int = 50, d: float = 1e7,
                 target_seed: int = 3, integrator: str = "rk2"):
        self.nx, self.ny, self.nsteps, self.checkpoint = nx, ny, nsteps, checkpoint
        self.base = W.periodic_wave_state(nx, ny, d=d, seed=0)
        self.cfg = W.periodic_wave_config(steps=nsteps, integrator=integrator)
        self.rho = self.base.grids["rho"]
        Lx, Ly = nx * d, ny * d
        x = (np.arange(nx) + 0.5) * d
        y = (np.arange(ny) + 0.5) * d
        X, Y = np.meshgrid(x, y, indexing="ij")
        self.sx = -self.rho * np.sin(2 * np.pi * Y / Ly)      # d mom_x / d v0
        self.sy = self.rho * np.sin(2 * np.pi * X / Lx)       # d mom_y / d v0
        self.target = 1e-15 * (1.0 + 0.05 * (2.0 * W.uniform(target_seed, (nx, ny)) - 1.0))
        self.temp0 = self.base.grids["temp"].copy()
        self.v0 = 1e7

    def spec(self, temp: np.ndarray, v0: float):
        st = self.base
        g = dict(st.grids)
        g["temp"] = np.asarray(temp, np.float64)
        g["mom_x"] = v0 * self.sx
        g["mom_y"] = v0 * self.sy
        from .stateio import StateFile
        return spec_from_files(self.cfg, StateFile(st.xdim, st.ydim, st.ion_mass, st.adiabatic_index, st.t, g,
                                                   list(st.comments)))

    def forward(self, temp: np.ndarray, v0: float) -> Tuple[float, np.ndarray, float]:
        """(loss, final rho, seconds of the nsteps steps)."""
        dom = GpuDomain(self.spec(temp, v0))
        try:
            t0 = time.perf_counter()
            dom.step(self.nsteps)
            sec = time.perf_counter() - t0
            rho = dom.grid("rho")
        finally:
            dom.close()
        return float(np.sum((rho - self.target) ** 2)), rho, sec

    def gradient(self, temp: np.ndarray, v0: float, rho_final: np.ndarray) -> Tuple[np.ndarray, float, float]:
        """(dJ/dtemp, dJ/dv0, seconds of the adjoint call: checkpointed forward + reverse sweep)."""
        dom = GpuDomain(self.spec(temp, v0))
        try:
            seed = np.zeros((8, self.nx, self.ny))
            seed[0] = 2.0 * (rho_final - self.target)
            t0 = time.perf_counter()
            g_ev, _ = dom.adjoint(self.nsteps, self.checkpoint, seed)
            sec = time.perf_counter() - t0
            g_state = dom.adjoint_state(g_ev)
        finally:
            dom.close()
        g_temp = g_state[1]
        g_v0 = float(np.sum(g_state[2] * self.sx) + np.sum(g_state[3] * self.sy))
        return g_temp, g_v0, sec


class ArcadeHeatingProblem:
    """Arcade (default 2048x2048, rows stretched by 1.0005, dipole depth 6.4e8 cm), ideal MHD with
    thermal conduction and radiative losses (RK2, FP64), heated by parametric_heating with K
    coefficients.  J = sum over coronal cells (initial T above 0.9 of the coronal temperature, below
    the top eighth of the rows) of (T_final - T_initial)^2 after nsteps steps."""

    def __init__(self, nx: int = 2048, ny: int = 2048, stretch: float = 1.0005, depth: float = 6.4e8,
                 nsteps: int = 1000, checkpoint: int = 50, K: int = 16, theta_scale: float = 1e-4,
                 bc=("periodic", "periodic", "fixed", "open"), options=None):
        self.nx, self.ny, self.nsteps, self.checkpoint, self.K = nx, ny, nsteps, checkpoint, K
        self.options = dict(options or {})   # execution options of the domains (spruce_gpu_set_option)
        self.bc = tuple(bc)
        self.base = W.arcade_state(nx, ny, stretch=stretch, depth=depth)
        self.t0 = self.base.grids["temp"].copy()
        # coronal cells below the top eighth of the domain: the cells next to the open top boundary
        # cool to the temperature floor within tens of steps, where the misfit has no derivative
        self.mask = (self.t0 > 0.9 * float(np.max(self.t0))) & (np.arange(ny)[None, :] < ny - ny // 8)
        self.theta0 = theta_scale * (1.0 + 0.5 * np.sin(np.arange(K)))
        self.m = self.base.ion_mass
        self.gm1 = self.base.adiabatic_index - 1.0
        self.tmin = float(W.arcade_config(steps=1).settings["temp_min"])

    def spec(self, theta: np.ndarray):
        cfg = W.arcade_config(steps=self.nsteps)
        for k, v in zip(("x_bound_1", "x_bound_2", "y_bound_1", "y_bound_2"), self.bc):
            cfg.settings[k] = v
        cfg.modules.append(("parametric_heating", [("num_basis", str(self.K))] +
                            [("coeff_%d" % k, repr(float(t))) for k, t in enumerate(theta)]))
        return spec_from_files(cfg, self.base)

    def forward(self, theta: np.ndarray):
        """(loss, final (temp, n), seconds of the nsteps steps)."""
        dom = GpuDomain(self.spec(theta), self.options)
        try:
            t0 = time.perf_counter()
            dom.step(self.nsteps)
            sec = time.perf_counter() - t0
            T, n = dom.grid("temp"), dom.grid("n")
        finally:
            dom.close()
        return float(np.sum(((T - self.t0) * self.mask) ** 2)), (T, n), sec

    def seed(self, final) -> np.ndarray:
        """dJ/d(evolved grids after the steps) from T = max((gamma-1) te / (2 K_B n), T_min), n = rho/m."""
        T, n = final
        dJ = 2.0 * (T - self.t0) * self.mask * (T > self.tmin)
        seed = np.zeros((8, self.nx, self.ny))
        seed[4] = dJ * (self.gm1 / (2.0 * W.K_B * n))
        seed[0] = dJ * (-T / n) / self.m
        return seed

    def gradient(self, theta: np.ndarray, final):
        """(dJ/dtheta, seconds of the adjoint call, device bytes the adjoint allocated)."""
        dom = GpuDomain(self.spec(theta), self.options)
        try:
            t0 = time.perf_counter()
            dom.adjoint(self.nsteps, self.checkpoint, self.seed(final))
            sec = time.perf_counter() - t0
            g = dom.adjoint_params()
            mem = dom.adjoint_device_bytes()
        finally:
            dom.close()
        return g, sec, mem
