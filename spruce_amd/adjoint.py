"""Gradient of a misfit of the final state through many IdealMHD steps, by the GPU discrete adjoint.

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
