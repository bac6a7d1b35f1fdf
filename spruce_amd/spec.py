"""Domain description shared by every backend: what a .config + .state pair asks for.

The fields mirror PlasmaDomain's configuration (source/mhd/plasmadomain.hpp:93-153) and the
registries of equation sets and modules (equationset.hpp:22; modulehandler.hpp:34-38).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

from .stateio import BOUNDARY_NAMES, INTEGRATOR_NAMES, PD_GRID_NAMES, ConfigFile, StateFile


@dataclass
class DomainSpec:
    xdim: int
    ydim: int
    bc: Tuple[int, int, int, int]
    integrator: int
    epsilon: float
    density_min: float
    temp_min: float
    thermal_energy_min: float
    open_boundary_strength: float
    open_boundary_decay_base: float
    ion_mass: float
    adiabatic_index: float
    equation_set: str
    equation_set_options: List[Tuple[str, str]] = field(default_factory=list)
    modules: List[Tuple[str, List[Tuple[str, str]]]] = field(default_factory=list)
    grids: Dict[str, np.ndarray] = field(default_factory=dict)
    max_iterations: int = 100
    time: float = 0.0

    @property
    def scalars(self) -> List[float]:
        return [self.epsilon, self.density_min, self.temp_min, self.thermal_energy_min,
                self.open_boundary_strength, self.open_boundary_decay_base, self.ion_mass,
                self.adiabatic_index]

    def interior(self) -> Tuple[int, int, int, int]:
        """(xl, xu, yl, yu) as PlasmaDomain::computeIterationBounds (plasmadomain.cpp:138-161)."""
        xl = 0 if self.bc[0] == 0 else 2
        xu = self.xdim - 1 if self.bc[1] == 0 else self.xdim - 3
        yl = 0 if self.bc[2] == 0 else 2
        yu = self.ydim - 1 if self.bc[3] == 0 else self.ydim - 3
        return xl, xu, yl, yu


def spec_from_files(cfg: ConfigFile, st: StateFile) -> DomainSpec:
    s = cfg.settings
    need = ("x_bound_1", "x_bound_2", "y_bound_1", "y_bound_2", "epsilon", "density_min", "temp_min",
            "thermal_energy_min")
    for k in need:
        if k not in s:
            raise ValueError("config key %s must be set" % k)
    bc = tuple(BOUNDARY_NAMES.index(s[k]) for k in need[:4])
    integ = INTEGRATOR_NAMES.index(s.get("time_integrator", "euler"))
    grids = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in st.grids.items()}
    for name in PD_GRID_NAMES:
        if name not in grids:
            raise ValueError("state file lacks the PlasmaDomain grid %s" % name)
    return DomainSpec(
        xdim=st.xdim, ydim=st.ydim, bc=bc, integrator=integ,
        epsilon=float(s["epsilon"]), density_min=float(s["density_min"]), temp_min=float(s["temp_min"]),
        thermal_energy_min=float(s["thermal_energy_min"]),
        open_boundary_strength=float(s.get("open_boundary_strength", "0.0")),
        open_boundary_decay_base=float(s.get("open_boundary_decay_base", "1.0")),
        ion_mass=st.ion_mass, adiabatic_index=st.adiabatic_index,
        equation_set=cfg.equation_set, equation_set_options=list(cfg.equation_set_options),
        modules=[(n, list(o)) for n, o in cfg.modules], grids=grids,
        max_iterations=int(s.get("max_iterations", "100")), time=st.t)
