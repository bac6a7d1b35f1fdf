"""ctypes binding of the CPU restatement (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU leg import this module; it is the checker
for the GPU path, never a fallback for it.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        CP = ctypes.POINTER(ctypes.c_char_p)
        L.oracle_create.restype = P
        L.oracle_create.argtypes = [I, I, ctypes.POINTER(I), I, ctypes.POINTER(D), ctypes.c_char_p, I, CP, CP]
        L.oracle_add_module.argtypes = [P, ctypes.c_char_p, I, CP, CP]
        L.oracle_set_grid.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(D)]
        L.oracle_setup.argtypes = [P]
        L.oracle_step.argtypes = [P, I, ctypes.POINTER(D)]
        L.oracle_get_grid.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(D)]
        L.oracle_rhs.argtypes = [P, ctypes.POINTER(D)]
        L.oracle_num_evolved.argtypes = [P]
        L.oracle_propagate.argtypes = [P, ctypes.POINTER(D)]
        L.oracle_ghost_fill.argtypes = [P]
        L.oracle_time.restype = D
        L.oracle_time.argtypes = [P]
        L.oracle_subcycles.argtypes = [P, ctypes.POINTER(I), I]
        L.oracle_set_pow_mode.argtypes = [I]
        L.oracle_pow.restype = D
        L.oracle_pow.argtypes = [D, D, I]
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_destroy.argtypes = [P]
        _LIB = L
    return _LIB


def _strs(items: Sequence[str]):
    arr = (ctypes.c_char_p * max(1, len(items)))()
    for i, s in enumerate(items):
        arr[i] = s.encode()
    return arr


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_pow_mode(mode: int) -> None:
    """0: the C library's pow (bit-identical to the reference); 1: correctly rounded pow (as the GPU)."""
    lib().oracle_set_pow_mode(mode)


def cr_pow(x: np.ndarray, y: float) -> np.ndarray:
    f = lib().oracle_pow
    return np.array([f(float(v), float(y), 1) for v in np.ravel(x)]).reshape(np.shape(x))


class OracleDomain:
    """A CPU restatement of one PlasmaDomain, driven through the same spec as the GPU path."""

    def __init__(self, spec):
        L = lib()
        self.spec = spec
        bc = (ctypes.c_int * 4)(*spec.bc)
        sc = (ctypes.c_double * 8)(*spec.scalars)
        keys = [k for k, _ in spec.equation_set_options]
        vals = [v for _, v in spec.equation_set_options]
        self.h = L.oracle_create(spec.xdim, spec.ydim, bc, spec.integrator, sc, spec.equation_set.encode(),
                                 len(keys), _strs(keys), _strs(vals))
        if not self.h:
            raise RuntimeError(L.oracle_last_error().decode())
        for name, opts in spec.modules:
            k = [a for a, _ in opts]
            v = [b for _, b in opts]
            self._chk(L.oracle_add_module(self.h, name.encode(), len(k), _strs(k), _strs(v)))
        for name, g in spec.grids.items():
            a = np.ascontiguousarray(g, dtype=np.float64)
            self._chk(L.oracle_set_grid(self.h, name.encode(), _dptr(a)))
        self._chk(L.oracle_setup(self.h))

    def _chk(self, rc: int) -> None:
        if rc != 0:
            raise RuntimeError(lib().oracle_last_error().decode())

    def step(self, n: int = 1) -> np.ndarray:
        dts = np.zeros(n)
        self._chk(lib().oracle_step(self.h, n, _dptr(dts)))
        return dts

    def grid(self, name: str) -> np.ndarray:
        out = np.empty((self.spec.xdim, self.spec.ydim))
        self._chk(lib().oracle_get_grid(self.h, name.encode(), _dptr(out)))
        return out

    def rhs(self) -> np.ndarray:
        ne = lib().oracle_num_evolved(self.h)
        out = np.empty((ne, self.spec.xdim, self.spec.ydim))
        self._chk(lib().oracle_rhs(self.h, _dptr(out)))
        return out

    def put_grid(self, name: str, g: np.ndarray) -> None:
        """Overwrite one grid of the live state (no propagate)."""
        a = np.ascontiguousarray(g, dtype=np.float64)
        self._chk(lib().oracle_set_grid(self.h, name.encode(), _dptr(a)))

    def ghost_fill(self) -> None:
        self._chk(lib().oracle_ghost_fill(self.h))

    def propagate(self) -> float:
        v = ctypes.c_double()
        self._chk(lib().oracle_propagate(self.h, ctypes.byref(v)))
        return v.value

    def subcycles(self) -> List[int]:
        buf = (ctypes.c_int * 16)()
        n = lib().oracle_subcycles(self.h, buf, 16)
        return [buf[i] for i in range(min(n, 16))]

    @property
    def time(self) -> float:
        return lib().oracle_time(self.h)

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
