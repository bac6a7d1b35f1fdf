"""Loading the golden vectors written by tests/golden/make_golden.py."""
from __future__ import annotations

import glob
import os
from dataclasses import dataclass
from typing import Dict, List

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
