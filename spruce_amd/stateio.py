"""Readers and writers for SPRUCE's text formats: ``.state``, ``.config`` and ``mhd.out``.

The grammar follows the reference's own reader/writer:

* ``.state``  -- PlasmaDomain::readStateFile / writeStateFile (source/mhd/fileio.cpp:14-80, 221-255).
  Optional leading ``#`` comment lines, then ``xdim,ydim`` / ``X,Y`` / ``ion_mass`` / value /
  ``adiabatic_index`` / value / ``t=...``, then repeated ``name`` + X rows of Y comma-separated
  values (rows are x, columns are y).  Whitespace (including CR) is stripped from every line
  (clearWhitespace/getCleanedLine, source/mhd/utils.cpp:8-25).
* ``.config`` -- PlasmaDomain::readConfigFile (fileio.cpp:84-122), EquationSet::configureEquationSet
  (source/equationsets/equationset.cpp:64-85) and Module::configureModule (source/modules/module.cpp:11-35):
  ``key = value`` lines, ``#`` comments, and ``name = true`` blocks followed by ``{`` ... ``}`` whose
  braces sit in column 0.
* ``mhd.out`` -- outputPreamble / storeGrids / writeToOutFile (fileio.cpp:125-215).

Numbers are printed the way ``std::ostream`` prints a double in its default float field with
``precision(p)``, i.e. C's ``%.{p}g`` (Grid::format, source/mhd/grid.cpp:412-427).  Precision
``-1`` means ``digits10 + 1`` = 16 significant digits.
"""
from __future__ import annotations

import io
import os
from dataclasses import dataclass, field
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

# PlasmaDomain's own grids, in the order the reference writes them (plasmadomain.hpp:35-36).
PD_GRID_NAMES = ("d_x", "d_y", "pos_x", "pos_y", "be_x", "be_y", "be_z")

# Keys PlasmaDomain accepts in a .config file (plasmadomain.hpp:46-51).
PD_CONFIG_NAMES = (
    "x_bound_1", "x_bound_2", "y_bound_1", "y_bound_2",
    "epsilon", "density_min", "temp_min", "thermal_energy_min", "max_iterations",
    "iter_output_interval", "time_output_interval", "output_flags", "xdim", "ydim",
    "open_boundary_strength", "std_out_interval", "write_interval",
    "open_boundary_decay_base", "x_origin", "y_origin", "time_integrator", "duration",
    "sg_opt", "write_precision", "multispecies_mode",
)

EQUATION_SET_NAMES = ("ideal_mhd", "ideal_mhd_cons", "ideal_mhd_2E", "ideal_2F")

MODULE_NAMES = (
    "ambient_heating", "radiative_losses", "thermal_conduction", "localized_heating",
    "field_heating", "sg_filtering", "coulomb_explosion", "eic_thermalization",
    "artificial_viscosity", "global_temperature", "anomalous_resistivity",
    "momentum_injection", "tracer_particles", "mass_injection", "ambient_heating_sink",
    "physical_viscosity", "div_cleaning", "boundary_outflow",
    # the parametrised heating profile of the gradient workload (no reference module)
    "parametric_heating",
)

BOUNDARY_NAMES = ("periodic", "open", "fixed", "reflect", "open_moc", "open_ucnp")
INTEGRATOR_NAMES = ("euler", "rk2", "rk4")


def _clean(line: str) -> str:
    """Remove every whitespace character the reference strips (utils.cpp:8-16)."""
    for c in (" ", "\t", "\n", "\v", "\b", "\r", "\f", "\a"):
        line = line.replace(c, "")
    return line


def format_double(x: float, precision: int = 4) -> str:
    """Format like ``std::ostream << double`` with ``precision(p)`` (default float field)."""
    if precision == -1:
        precision = 16
    if precision == 0:
        precision = 1
    s = "%.*g" % (precision, x)
    return s


def format_grid(a: np.ndarray, precision: int = 4) -> str:
    """Grid::format(',', '\\n', precision, '\\n') (grid.cpp:412-427)."""
    a = np.asarray(a, dtype=np.float64)
    rows = [",".join(format_double(float(v), precision) for v in row) for row in a]
    return "\n".join(rows) + "\n"


# --------------------------------------------------------------------------------------------
# .state
# --------------------------------------------------------------------------------------------

@dataclass
class StateFile:
    xdim: int
    ydim: int
    ion_mass: float
    adiabatic_index: float
    t: float = 0.0
    grids: Dict[str, np.ndarray] = field(default_factory=dict)
    comments: List[str] = field(default_factory=list)

    def to_text(self, precision: int = 17, order: Optional[Sequence[str]] = None) -> str:
        out = io.StringIO()
        for c in self.comments:
            assert c.startswith("#")
            out.write(c + "\n")
        out.write("xdim,ydim\n%d,%d\n" % (self.xdim, self.ydim))
        out.write("ion_mass\n%s\n" % format_double(self.ion_mass, precision))
        out.write("adiabatic_index\n%s\n" % format_double(self.adiabatic_index, precision))
        out.write("t=%s\n" % format_double(self.t, precision))
        names = list(order) if order is not None else list(self.grids.keys())
        for name in names:
            g = np.asarray(self.grids[name], dtype=np.float64)
            if g.shape != (self.xdim, self.ydim):
                raise ValueError("grid %s has shape %s, expected %s" % (name, g.shape, (self.xdim, self.ydim)))
            out.write(name + "\n")
            out.write(format_grid(g, precision))
        return out.getvalue()

    def write(self, path: str, precision: int = 17, order: Optional[Sequence[str]] = None) -> None:
        with open(path, "w") as f:
            f.write(self.to_text(precision, order))


def _first_numeric(row: str) -> bool:
    return bool(row) and (row[0].isdigit() or row[0] in "-.ni")


def parse_state(text: str, continue_mode: bool = False) -> StateFile:
    """Parse a .state file with the semantics of PlasmaDomain::readStateFile (fileio.cpp:14-80)."""
    lines = text.splitlines()
    pos = 0
    comments: List[str] = []
    while pos < len(lines) and (lines[pos] == "" or lines[pos].startswith("#")):
        if lines[pos].startswith("#"):
            comments.append(lines[pos])
        pos += 1

    def nxt() -> str:
        nonlocal pos
        if pos >= len(lines):
            raise ValueError("unexpected end of .state file")
        s = _clean(lines[pos])
        pos += 1
        return s

    if nxt() != "xdim,ydim":
        raise ValueError(".state: expected 'xdim,ydim'")
    dims = nxt().split(",")
    xdim, ydim = int(dims[0]), int(dims[1])
    if nxt() != "ion_mass":
        raise ValueError(".state: expected 'ion_mass'")
    ion_mass = float(nxt())
    if nxt() != "adiabatic_index":
        raise ValueError(".state: expected 'adiabatic_index'")
    gamma = float(nxt())
    tline = nxt()
    key, _, val = tline.partition("=")
    if key != "t":
        raise ValueError(".state: expected 't='")
    t = float(val) if continue_mode else 0.0
    grids: Dict[str, np.ndarray] = {}
    while pos < len(lines):
        name = nxt()
        if name == "":
            continue
        g = np.zeros((xdim, ydim), dtype=np.float64)
        for i in range(xdim):
            row = nxt()
            if not _first_numeric(row):
                raise ValueError(".state: non-numerical row in grid %s" % name)
            vals = [v for v in row.split(",")]
            if len(vals) > ydim:
                raise ValueError(".state: row longer than ydim in grid %s" % name)
            g[i, : len(vals)] = [float(v) for v in vals]
        grids[name] = g
    return StateFile(xdim, ydim, ion_mass, gamma, t, grids, comments)


def read_state(path: str, continue_mode: bool = False) -> StateFile:
    with open(path, "r", newline="") as f:
        return parse_state(f.read(), continue_mode)


# --------------------------------------------------------------------------------------------
# .config
# --------------------------------------------------------------------------------------------

@dataclass
class ConfigFile:
    """Parsed .config: PlasmaDomain settings, the equation set block, and ordered module blocks."""
    settings: Dict[str, str] = field(default_factory=dict)
    output_flags: List[str] = field(default_factory=list)
    equation_set: str = ""
    equation_set_options: List[Tuple[str, str]] = field(default_factory=list)
    modules: List[Tuple[str, List[Tuple[str, str]]]] = field(default_factory=list)

    def to_text(self) -> str:
        out = io.StringIO()
        out.write("%s = true\n{\n" % self.equation_set)
        for k, v in self.equation_set_options:
            out.write("%s = %s\n" % (k, v))
        out.write("}\n")
        for k, v in self.settings.items():
            out.write("%s = %s\n" % (k, v))
        if self.output_flags:
            out.write("output_flags = %s\n" % ",".join(self.output_flags))
        for name, opts in self.modules:
            out.write("%s = true\n{\n" % name)
            for k, v in opts:
                out.write("%s = %s\n" % (k, v))
            out.write("}\n")
        return out.getvalue()

    def write(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(self.to_text())

    def get(self, key: str, default=None):
        return self.settings.get(key, default)


def parse_config(text: str) -> ConfigFile:
    """Parse a .config with the semantics of readConfigFile (fileio.cpp:84-122)."""
    cfg = ConfigFile()
    lines = text.splitlines()
    pos = 0

    def block(cleaned_braces: bool) -> List[Tuple[str, str]]:
        # Block bodies: '{' must open the block in column 0 (equationset.cpp:68-72; module.cpp:15-19).
        nonlocal pos
        opts: List[Tuple[str, str]] = []
        if pos >= len(lines) or not lines[pos].startswith("{"):
            raise ValueError(".config: block must be followed by '{' in column 0")
        pos += 1
        while pos < len(lines) and not lines[pos].startswith("}"):
            ln = _clean(lines[pos])
            pos += 1
            if ln == "" or ln.startswith("#"):
                if not cleaned_braces:
                    # An empty/comment line inside an equation-set block never advances the
                    # reference's reader (equationset.cpp:72-74): reject instead of hanging.
                    raise ValueError(".config: empty or comment line inside an equation-set block")
                continue
            k, _, rest = ln.partition("=")
            v = rest.split("#", 1)[0]
            opts.append((k, v))
        pos += 1
        return opts

    while pos < len(lines):
        ln = _clean(lines[pos])
        pos += 1
        if ln == "" or ln.startswith("#"):
            continue
        lhs, _, rest = ln.partition("=")
        rhs = rest.split("#", 1)[0]
        if lhs in EQUATION_SET_NAMES:
            opts = block(cleaned_braces=False)
            if rhs == "true":
                cfg.equation_set = lhs
                cfg.equation_set_options = opts
            continue
        if lhs in MODULE_NAMES:
            if not cfg.equation_set:
                raise ValueError(".config: the equation set must precede modules")
            opts = block(cleaned_braces=True)
            if rhs == "true":
                cfg.modules.append((lhs, opts))
            continue
        if lhs not in PD_CONFIG_NAMES:
            raise ValueError("%s is not a valid config name." % lhs)
        if lhs == "output_flags":
            cfg.output_flags.extend([v for v in rhs.split(",") if v])
        else:
            cfg.settings[lhs] = rhs
    return cfg


def read_config(path: str) -> ConfigFile:
    with open(path, "r", newline="") as f:
        return parse_config(f.read())


# --------------------------------------------------------------------------------------------
# mhd.out
# --------------------------------------------------------------------------------------------

@dataclass
class OutFile:
    xdim: int
    ydim: int
    preamble: Dict[str, np.ndarray]
    times: List[str]
    frames: List[Dict[str, np.ndarray]]


def parse_out(text: str) -> OutFile:
    """Parse mhd.out as written by outputPreamble/storeGrids (fileio.cpp:125-200)."""
    lines = [ln for ln in text.splitlines()]
    pos = 0
    while lines[pos].startswith("#"):
        pos += 1
    assert _clean(lines[pos]) == "xdim,ydim"
    xdim, ydim = (int(v) for v in _clean(lines[pos + 1]).split(","))
    pos += 2

    def grid() -> np.ndarray:
        nonlocal pos
        g = np.array([[float(v) for v in _clean(lines[pos + i]).split(",")] for i in range(xdim)],
                     dtype=np.float64)
        pos += xdim
        return g

    preamble: Dict[str, np.ndarray] = {}
    while pos < len(lines) and not lines[pos].startswith("t="):
        name = _clean(lines[pos])
        pos += 1
        preamble[name] = grid()
    times: List[str] = []
    frames: List[Dict[str, np.ndarray]] = []
    while pos < len(lines):
        ln = _clean(lines[pos])
        pos += 1
        if ln == "":
            continue
        if ln.startswith("t="):
            times.append(ln[2:])
            frames.append({})
            continue
        frames[-1][ln] = grid()
    return OutFile(xdim, ydim, preamble, times, frames)


def read_out(path: str) -> OutFile:
    with open(path, "r") as f:
        return parse_out(f.read())


def write_out_preamble(f, xdim: int, ydim: int, pd_grids: Dict[str, np.ndarray],
                       comments: Iterable[str] = ()) -> None:
    """outputPreamble (fileio.cpp:125-139): comments, dims, then pos/be grids at precision 4."""
    for c in comments:
        f.write(c + "\n")
    f.write("xdim,ydim\n%d,%d\n" % (xdim, ydim))
    for name in ("pos_x", "pos_y", "be_x", "be_y", "be_z"):
        f.write(name + "\n")
        f.write(format_grid(pd_grids[name], 4))


def positions_from_sizes(d: np.ndarray, index: int) -> np.ndarray:
    """Cell-centre positions from sizes with origin 'lower' (plasmadomain.cpp:169-192)."""
    d = np.asarray(d, dtype=np.float64)
    pos = np.zeros_like(d)
    xdim, ydim = d.shape
    for i in range(xdim):
        for j in range(ydim):
            if (i == 0 and index == 0) or (j == 0 and index == 1):
                pos[i, j] = 0.5 * d[i, j]
            else:
                ip, jp = (i - 1, j) if index == 0 else (i, j - 1)
                pos[i, j] = pos[ip, jp] + 0.5 * d[ip, jp] + 0.5 * d[i, j]
    return pos
