"""Loading the golden vectors written by tests/golden/make_golden.py."""
from __future__ import annotations

import glob
import json
import os
import re
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np

from spruce_amd.spec import DomainSpec, spec_from_files
from spruce_amd.stateio import parse_config, parse_state

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@dataclass
class Golden:
    name: str
    state_text: str
    config_text: str
    names: List[str]
    frames: List[int]
    data: np.ndarray          # [frame, var, x, y]
    dt_steps: np.ndarray      # step sizes of steps 1..N
    bounds: np.ndarray

    def spec(self) -> DomainSpec:
        return spec_from_files(parse_config(self.config_text), parse_state(self.state_text))

    def frame(self, k: int) -> Dict[str, np.ndarray]:
        f = self.frames.index(k)
        return {n: self.data[f, i] for i, n in enumerate(self.names)}

    @property
    def nsteps(self) -> int:
        return len(self.dt_steps)


def golden_names() -> List[str]:
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN_DIR, "*.npz")))


def load_golden(name: str) -> Golden:
    z = np.load(os.path.join(GOLDEN_DIR, name + ".npz"), allow_pickle=False)
    return Golden(name, bytes(z["state"]).decode(), bytes(z["config"]).decode(), [str(s) for s in z["names"]],
                  [int(f) for f in z["frames"]], z["data"], z["dt_steps"], z["bounds"])


# Modules whose progress message carries their integer subcycle count (thermalconduction.cpp:244,
# radiativelosses.cpp:171, anomalousresistivity.cpp:317, physicalviscosity.cpp:286).
COUNTED_MODULES = ("thermal_conduction", "radiative_losses", "anomalous_resistivity", "physical_viscosity")
_COUNT = re.compile(r"(?:Subcycles: (\d+)|, (\d+) Subcycle\(s\))")


def load_progress(name: str) -> Optional[List[str]]:
    """The reference's per-step progress lines of a golden case (tests/golden/progress.json)."""
    path = os.path.join(GOLDEN_DIR, "progress.json")
    if not os.path.exists(path):
        return None
    return json.load(open(path)).get(name)


def progress_counts(line: str) -> List[int]:
    """The subcycle counts of one progress line, in module order."""
    out = []
    for part in line.split("|")[3:]:
        m = _COUNT.search(part)
        if m:
            out.append(int(m.group(1) or m.group(2)))
    return out


def counted_cases() -> List[str]:
    """Golden cases whose progress lines carry subcycle counts."""
    path = os.path.join(GOLDEN_DIR, "progress.json")
    if not os.path.exists(path):
        return []
    prog = json.load(open(path))
    return sorted(k for k, v in prog.items() if v and any(progress_counts(x) for x in v))


def counted_module_indices(spec) -> List[int]:
    """Indices (config order) of the modules whose counts the progress lines print."""
    return [k for k, (name, _) in enumerate(spec.modules) if name in COUNTED_MODULES]
