"""ctypes binding of libspruce_gpu.so (include/spruce_gpu.h).

``GpuDomain`` mirrors the reference PlasmaDomain's lifecycle (construct from config + state, setup,
advance, read grids).  There is no CPU fallback: if the HIP library is missing or fails to load, the
constructor raises.
"""
from __future__ import annotations

import ctypes
import os
import re
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .spec import DomainSpec

_HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; SPRUCE_GPU_LIB names another build of the same library
LIB_PATH = os.environ.get("SPRUCE_GPU_LIB") or os.path.join(_HERE, "libspruce_gpu.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "spruce_gpu.h")
_LIB: Optional[ctypes.CDLL] = None


class SpruceError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("spruce_gpu error %d: %s" % (code, msg))
        self.code = code


class DomainDesc(ctypes.Structure):
    _fields_ = [("xdim", ctypes.c_int32), ("ydim", ctypes.c_int32), ("bc", ctypes.c_int32 * 4),
                ("time_integrator", ctypes.c_int32), ("epsilon", ctypes.c_double),
                ("density_min", ctypes.c_double), ("temp_min", ctypes.c_double),
                ("thermal_energy_min", ctypes.c_double), ("open_boundary_strength", ctypes.c_double),
                ("open_boundary_decay_base", ctypes.c_double), ("ion_mass", ctypes.c_double),
                ("adiabatic_index", ctypes.c_double), ("time", ctypes.c_double)]


def header_functions() -> List[str]:
    """Names of every function the C ABI header declares."""
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(spruce_gpu_[a-z_0-9]+)\s*\(", text)))


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libspruce_gpu.so is not built; run __graft_entry__.build() or python spruce_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    P, I32, I64, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    CP = ctypes.POINTER(ctypes.c_char_p)
    DP = ctypes.POINTER(D)
    L.spruce_gpu_create.restype = P
    L.spruce_gpu_create.argtypes = [ctypes.POINTER(DomainDesc), ctypes.c_char_p, I32, CP, CP]
    L.spruce_gpu_add_module.argtypes = [P, ctypes.c_char_p, I32, CP, CP]
    L.spruce_gpu_set_grid.argtypes = [P, ctypes.c_char_p, DP, I64]
    L.spruce_gpu_setup.argtypes = [P]
    L.spruce_gpu_step.argtypes = [P, I32, DP, ctypes.POINTER(I32)]
    L.spruce_gpu_get_grid.argtypes = [P, ctypes.c_char_p, DP, I64]
    L.spruce_gpu_time.argtypes = [P, DP, ctypes.POINTER(I32)]
    L.spruce_gpu_rhs.argtypes = [P, DP, I64]
    L.spruce_gpu_dt_min.argtypes = [P, DP]
    L.spruce_gpu_put_grid.argtypes = [P, ctypes.c_char_p, DP, I64]
    L.spruce_gpu_rhs_noalloc.argtypes = [P, DP, I64]
    L.spruce_gpu_ghost_fill.argtypes = [P]
    L.spruce_gpu_propagate.argtypes = [P, DP]
    L.spruce_gpu_num_evolved.argtypes = [P]
    L.spruce_gpu_stream.argtypes = [P, ctypes.POINTER(P)]
    L.spruce_gpu_device_grid.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(DP)]
    L.spruce_gpu_set_option.argtypes = [P, ctypes.c_char_p, ctypes.c_char_p]
    L.spruce_gpu_set_profiling.argtypes = [P, I32]
    L.spruce_gpu_kernel_times.argtypes = [P, ctypes.c_char_p, I32, DP, ctypes.POINTER(I32), I32]
    L.spruce_gpu_arith_probe.argtypes = [DP, DP, I64, DP, DP, DP]
    L.spruce_gpu_pow_probe.argtypes = [DP, I64, I32, DP]
    L.spruce_gpu_last_error.restype = ctypes.c_char_p
    L.spruce_gpu_destroy.argtypes = [P]
    L.spruce_gpu_batch_create.restype = P
    L.spruce_gpu_batch_create.argtypes = [ctypes.POINTER(P), I32]
    L.spruce_gpu_batch_set_option.argtypes = [P, ctypes.c_char_p, ctypes.c_char_p]
    L.spruce_gpu_batch_autotune.argtypes = [P, I32, ctypes.c_char_p, I32, DP]
    L.spruce_gpu_batch_step.argtypes = [P, I32, DP]
    L.spruce_gpu_batch_destroy.argtypes = [P]
    L.spruce_gpu_adjoint.argtypes = [P, I32, I32, DP, DP, DP]
    L.spruce_gpu_adjoint_state.argtypes = [P, DP, DP]
    L.spruce_gpu_adjoint_params.argtypes = [P, DP, I32, ctypes.POINTER(I32)]
    L.spruce_gpu_module_report.argtypes = [P, I32, DP, I32, ctypes.POINTER(I32)]
    L.spruce_gpu_module_output.argtypes = [P, I32, I32, ctypes.c_char_p, I32, DP, ctypes.c_int64,
                                           ctypes.POINTER(I32)]
    L.spruce_gpu_multispecies.argtypes = [P, I32, DP, ctypes.c_int64]
    L.spruce_gpu_multispecies_reset.argtypes = [P]
    L.spruce_gpu_adjoint_device_bytes.argtypes = [P, ctypes.POINTER(ctypes.c_int64)]
    L.spruce_gpu_batch_set_profiling.argtypes = [P, I32]
    L.spruce_gpu_batch_kernel_times.argtypes = [P, ctypes.c_char_p, I32, DP, ctypes.POINTER(I32), I32]
    _LIB = L
    return L


def _err() -> str:
    return lib().spruce_gpu_last_error().decode()


def _chk(rc: int) -> int:
    if rc < 0:
        raise SpruceError(rc, _err())
    return rc


def _strs(items: Sequence[str]):
    arr = (ctypes.c_char_p * max(1, len(items)))()
    for i, s in enumerate(items):
        arr[i] = s.encode()
    return arr


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def arith_probe(a: np.ndarray, b: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    n = a.size
    d, s, q = np.empty(n), np.empty(n), np.empty(n)
    _chk(lib().spruce_gpu_arith_probe(_dp(a), _dp(b), n, _dp(d), _dp(s), _dp(q)))
    return d, s, q


POW_EXPONENTS = (1.5, 2.5, -1.5, 1.0 / 3.0)


def pow_probe(x: np.ndarray, which: int) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty(x.size)
    _chk(lib().spruce_gpu_pow_probe(_dp(x), x.size, which, _dp(out)))
    return out


class GpuDomain:
    """One PlasmaDomain advanced on the GPU through the C ABI."""

    def __init__(self, spec: DomainSpec, options: Optional[Dict[str, object]] = None):
        L = lib()
        self.spec = spec
        desc = DomainDesc()
        desc.xdim, desc.ydim = spec.xdim, spec.ydim
        for k in range(4):
            desc.bc[k] = spec.bc[k]
        desc.time_integrator = spec.integrator
        (desc.epsilon, desc.density_min, desc.temp_min, desc.thermal_energy_min, desc.open_boundary_strength,
         desc.open_boundary_decay_base, desc.ion_mass, desc.adiabatic_index) = spec.scalars
        desc.time = spec.time
        keys = [k for k, _ in spec.equation_set_options]
        vals = [v for _, v in spec.equation_set_options]
        self.h = L.spruce_gpu_create(ctypes.byref(desc), spec.equation_set.encode(), len(keys), _strs(keys), _strs(vals))
        if not self.h:
            raise SpruceError(-1, _err())
        self.modules = [n for n, _ in spec.modules]
        for name, opts in spec.modules:
            k = [a for a, _ in opts]
            v = [b for _, b in opts]
            _chk(L.spruce_gpu_add_module(self.h, name.encode(), len(k), _strs(k), _strs(v)))
        for k, v in (options or {}).items():
            _chk(L.spruce_gpu_set_option(self.h, k.encode(), str(v).encode()))
        n = spec.xdim * spec.ydim
        for name, g in spec.grids.items():
            a = np.ascontiguousarray(g, dtype=np.float64)
            _chk(L.spruce_gpu_set_grid(self.h, name.encode(), _dp(a), n))
        _chk(L.spruce_gpu_setup(self.h))

    def step(self, n: int = 1, subcycles: bool = False):
        dts = np.zeros(n)
        nm = max(1, len(self.modules))
        sub = np.zeros(n * nm, dtype=np.int32)
        _chk(lib().spruce_gpu_step(self.h, n, _dp(dts), sub.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        if subcycles:
            return dts, sub.reshape(n, nm)[:, : len(self.modules)]
        return dts

    def grid(self, name: str) -> np.ndarray:
        out = np.empty((self.spec.xdim, self.spec.ydim))
        _chk(lib().spruce_gpu_get_grid(self.h, name.encode(), _dp(out), out.size))
        return out

    def rhs(self) -> np.ndarray:
        ne = _chk(lib().spruce_gpu_num_evolved(self.h))
        out = np.empty((ne, self.spec.xdim, self.spec.ydim))
        _chk(lib().spruce_gpu_rhs(self.h, _dp(out), out.size))
        return out

    def dt_min(self) -> float:
        v = ctypes.c_double()
        _chk(lib().spruce_gpu_dt_min(self.h, ctypes.byref(v)))
        return v.value

    def rhs_noalloc(self) -> np.ndarray:
        """rhs() staged through the library's static device array (no device allocation)."""
        ne = _chk(lib().spruce_gpu_num_evolved(self.h))
        out = np.empty((ne, self.spec.xdim, self.spec.ydim))
        _chk(lib().spruce_gpu_rhs_noalloc(self.h, _dp(out), out.size))
        return out

    def put_grid(self, name: str, g: np.ndarray) -> None:
        """Overwrite one evolved grid of the live state without propagating."""
        a = np.ascontiguousarray(g, dtype=np.float64)
        _chk(lib().spruce_gpu_put_grid(self.h, name.encode(), _dp(a), a.size))

    def ghost_fill(self) -> None:
        _chk(lib().spruce_gpu_ghost_fill(self.h))

    def propagate(self) -> float:
        """propagateChanges on the live grids; returns the next step size."""
        v = ctypes.c_double()
        _chk(lib().spruce_gpu_propagate(self.h, ctypes.byref(v)))
        return v.value

    @property
    def time(self) -> float:
        t = ctypes.c_double()
        it = ctypes.c_int32()
        _chk(lib().spruce_gpu_time(self.h, ctypes.byref(t), ctypes.byref(it)))
        return t.value

    def set_option(self, key: str, value) -> None:
        _chk(lib().spruce_gpu_set_option(self.h, key.encode(), str(value).encode()))

    def set_profiling(self, on: bool) -> None:
        _chk(lib().spruce_gpu_set_profiling(self.h, 1 if on else 0))

    def kernel_times(self) -> Dict[str, Tuple[float, int]]:
        names = ctypes.create_string_buffer(4096)
        ms = (ctypes.c_double * 64)()
        cnt = (ctypes.c_int32 * 64)()
        n = _chk(lib().spruce_gpu_kernel_times(self.h, names, 4096, ms, cnt, 64))
        keys = names.value.decode().split(",") if n else []
        return {k: (ms[i], cnt[i]) for i, k in enumerate(keys)}

    def adjoint(self, nsteps: int, checkpoint: int, seed: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        """Runs nsteps steps and returns (dJ/d evolved grids before them, step sizes), given
        seed = dJ/d evolved grids after them, both shaped (8, xdim, ydim) in evolved order."""
        seed = np.ascontiguousarray(seed, np.float64).reshape(8, self.spec.xdim, self.spec.ydim)
        grad = np.empty_like(seed)
        dts = np.empty(nsteps)
        _chk(lib().spruce_gpu_adjoint(self.h, nsteps, checkpoint, _dp(seed), _dp(grad), _dp(dts)))
        return grad, dts

    def adjoint_params(self) -> np.ndarray:
        """dJ/dtheta of the last adjoint call for the parametric_heating coefficients, in module order."""
        n = ctypes.c_int32(0)
        _chk(lib().spruce_gpu_adjoint_params(self.h, None, 0, ctypes.byref(n)))
        out = np.empty(n.value)
        if n.value:
            _chk(lib().spruce_gpu_adjoint_params(self.h, _dp(out), n.value, None))
        return out

    def module_report(self, k: int) -> np.ndarray:
        """The device values module k's command-line message prints after the latest step
        (boundary_outflow: largest outflow speed, applied acceleration)."""
        n = ctypes.c_int32(0)
        _chk(lib().spruce_gpu_module_report(self.h, k, None, 0, ctypes.byref(n)))
        out = np.empty(n.value)
        if n.value:
            _chk(lib().spruce_gpu_module_report(self.h, k, _dp(out), n.value, None))
        return out

    def module_output(self, k: int) -> dict:
        """The grids module k writes to every stored frame (its fileOutput), by name, as of the
        latest step."""
        n = ctypes.c_int32(0)
        _chk(lib().spruce_gpu_module_output(self.h, k, 0, None, 0, None, 0, ctypes.byref(n)))
        out = {}
        for i in range(n.value):
            name = ctypes.create_string_buffer(128)
            g = np.empty((self.spec.xdim, self.spec.ydim))
            _chk(lib().spruce_gpu_module_output(self.h, k, i, name, 128, _dp(g), g.size, None))
            out[name.value.decode()] = g
        return out

    def multispecies(self) -> dict:
        """The cumulative electron, ion and Joule heating grids of multispecies mode."""
        out = {}
        for w, name in enumerate(("cumulative_electron_heating", "cumulative_ion_heating",
                                  "cumulative_joule_heating")):
            g = np.empty((self.spec.xdim, self.spec.ydim))
            _chk(lib().spruce_gpu_multispecies(self.h, w, _dp(g), g.size))
            out[name] = g
        return out

    def multispecies_reset(self) -> None:
        _chk(lib().spruce_gpu_multispecies_reset(self.h))

    def adjoint_device_bytes(self) -> int:
        """Device memory the last adjoint call allocated beyond the domain's own grids."""
        n = ctypes.c_int64(0)
        _chk(lib().spruce_gpu_adjoint_device_bytes(self.h, ctypes.byref(n)))
        return int(n.value)

    def adjoint_state(self, grad_evolved: np.ndarray) -> np.ndarray:
        """dJ/d(.state grids rho, temp, mom_x, mom_y, mom_z, bi_x, bi_y, bi_z) from dJ/d(evolved grids
        after setup)."""
        g = np.ascontiguousarray(grad_evolved, np.float64).reshape(8, self.spec.xdim, self.spec.ydim)
        out = np.empty_like(g)
        _chk(lib().spruce_gpu_adjoint_state(self.h, _dp(g), _dp(out)))
        return out

    def stream(self) -> int:
        p = ctypes.c_void_p()
        _chk(lib().spruce_gpu_stream(self.h, ctypes.byref(p)))
        return p.value or 0

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().spruce_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuBatch:
    """Several set-up ideal_2F GpuDomains advanced together (one launch per kernel of the step for all
    members, optionally replayed from a captured HIP graph).  Close it before its members."""

    def __init__(self, domains: Sequence[GpuDomain], graph: bool = True, block: Optional[str] = None):
        L = lib()
        self.domains = list(domains)
        arr = (ctypes.c_void_p * len(self.domains))(*[d.h for d in self.domains])
        self.h = L.spruce_gpu_batch_create(arr, len(self.domains))
        if not self.h:
            raise SpruceError(-1, _err())
        self.set_option("graph", "true" if graph else "false")
        if block:
            self.set_option("block", block)

    def set_option(self, key: str, value) -> None:
        _chk(lib().spruce_gpu_batch_set_option(self.h, key.encode(), str(value).encode()))

    def autotune(self, trial_steps: int = 5) -> Tuple[str, float]:
        buf = ctypes.create_string_buffer(64)
        ms = ctypes.c_double()
        _chk(lib().spruce_gpu_batch_autotune(self.h, trial_steps, buf, 64, ctypes.byref(ms)))
        return buf.value.decode(), ms.value

    def step(self, n: int = 1) -> np.ndarray:
        """n steps of every member; returns the step sizes as [member, step]."""
        dts = np.zeros(len(self.domains) * n)
        _chk(lib().spruce_gpu_batch_step(self.h, n, _dp(dts)))
        return dts.reshape(len(self.domains), n)

    def set_profiling(self, on: bool) -> None:
        _chk(lib().spruce_gpu_batch_set_profiling(self.h, 1 if on else 0))

    def kernel_times(self) -> Dict[str, Tuple[float, int]]:
        names = ctypes.create_string_buffer(4096)
        ms = (ctypes.c_double * 64)()
        cnt = (ctypes.c_int32 * 64)()
        n = _chk(lib().spruce_gpu_batch_kernel_times(self.h, names, 4096, ms, cnt, 64))
        keys = names.value.decode().split(",") if n else []
        return {k: (ms[i], cnt[i]) for i, k in enumerate(keys)}

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().spruce_gpu_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
