"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module: it is the checker, never the product.
The C sources next to this file restate the reference's arithmetic; see the
header of ``mmc_oracle.c`` for the reference file:line each part follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libmmc_oracle.so")

EXCH, DMI, ANIS, ZEEMAN, DEMAG, ALL = 1, 2, 4, 8, 16, 31


class Params(C.Structure):
    """Same field order as ``mmc_params`` (include/mmc.h)."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("pbc_x", C.c_int), ("pbc_y", C.c_int), ("pbc_z", C.c_int),
        ("ms", C.c_double), ("aex", C.c_double), ("dind", C.c_double),
        ("anis_ux", C.c_double), ("anis_uy", C.c_double),
        ("anis_uz", C.c_double), ("ku1", C.c_double),
        ("bext_x", C.c_double), ("bext_y", C.c_double), ("bext_z", C.c_double),
        ("cal_demag", C.c_int), ("reserved", C.c_int),
        ("seed", C.c_ulonglong),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the .so path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []),
                       check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_padded_len.argtypes = [C.c_int, C.c_int]
        L.orc_padded_len.restype = C.c_int
        L.orc_tensor_offset.argtypes = [C.c_int] * 4 + [C.POINTER(C.c_int)]
        L.orc_tensor_offset.restype = C.c_int
        L.orc_newell_f.argtypes = [C.c_double] * 3
        L.orc_newell_f.restype = C.c_double
        L.orc_newell_g.argtypes = [C.c_double] * 3
        L.orc_newell_g.restype = C.c_double
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.orc_philox4x32_10.restype = None
        L.orc_u01.argtypes = [C.c_uint32]
        L.orc_u01.restype = C.c_double
        L.orc_fft1d.argtypes = [dp, C.c_int, C.c_int]
        L.orc_fft1d.restype = C.c_int
        L.orc_create.argtypes = [C.POINTER(Params)]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        for name in ("orc_build_tensor",):
            getattr(L, name).argtypes = [C.c_void_p]
        for name in ("orc_set_tensor", "orc_get_tensor", "orc_mirror",
                     "orc_set_m", "orc_get_m", "orc_demag_fft",
                     "orc_demag_direct"):
            getattr(L, name).argtypes = [C.c_void_p, dp]
        L.orc_init_random_m.argtypes = [C.c_void_p, C.c_ulonglong]
        L.orc_field.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_energy.argtypes = [C.c_void_p, dp]
        L.orc_energy_density.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_delta_e.argtypes = [C.c_void_p, dp, C.c_int, dp]
        L.orc_delta_e.restype = C.c_double
        L.orc_trial_delta_e.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, dp]
        L.orc_trial_delta_e.restype = None
        L.orc_sweep.argtypes = [C.c_void_p, C.c_double, C.c_uint32, C.c_uint32]
        L.orc_sweep.restype = C.c_long
        L.orc_num_colors.argtypes = [C.c_void_p]
        L.orc_num_colors.restype = C.c_int
        L.orc_cell_color.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_cell_color.restype = C.c_int
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(**kw) -> Params:
    p = Params()
    defaults = dict(nx=8, ny=8, nz=1, dx=1e-9, dy=1e-9, dz=0.8e-9,
                    pbc_x=0, pbc_y=0, pbc_z=0, ms=1.6e6, aex=16e-12,
                    dind=6e-3, anis_ux=0.0, anis_uy=0.0, anis_uz=1.0,
                    ku1=2e6, bext_x=0.0, bext_y=0.0, bext_z=0.3,
                    cal_demag=1, reserved=0, seed=256)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def philox(ctr, key):
    a = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(a, k, o)
    return list(o)


def fft1d(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    z = np.ascontiguousarray(x, dtype=np.complex128).copy()
    buf = z.view(np.float64)
    rc = lib().orc_fft1d(_dp(buf), len(z), int(inverse))
    assert rc == 0
    return z


class Oracle:
    """One oracle instance: grid, tensor, magnetisation."""

    def __init__(self, params: Params | None = None, **kw):
        self.params = params if params is not None else make_params(**kw)
        L = lib()
        self._c = L.orc_create(C.byref(self.params))
        if not self._c:
            raise ValueError("orc_create failed")
        d = (C.c_int * 6)()
        L.orc_dims(self._c, d)
        self.n = tuple(d[:3])
        self.P = tuple(d[3:])
        self.N = self.n[0] * self.n[1] * self.n[2]
        self.NP = self.P[0] * self.P[1] * self.P[2]

    def __del__(self):
        if getattr(self, "_c", None):
            lib().orc_destroy(self._c)
            self._c = None

    # tensor -------------------------------------------------------------
    def build_tensor(self):
        lib().orc_build_tensor(self._c)

    def set_tensor(self, K: np.ndarray):
        K = np.ascontiguousarray(K, dtype=np.float64).reshape(6 * self.NP)
        lib().orc_set_tensor(self._c, _dp(K))

    def tensor(self) -> np.ndarray:
        K = np.empty(6 * self.NP)
        lib().orc_get_tensor(self._c, _dp(K))
        return K.reshape(6, self.P[2], self.P[1], self.P[0])

    def mirror(self) -> np.ndarray:
        K0 = np.empty(6)
        lib().orc_mirror(self._c, _dp(K0))
        return K0

    # magnetisation ------------------------------------------------------
    def set_m(self, m: np.ndarray):
        m = np.ascontiguousarray(m, dtype=np.float64).reshape(3 * self.N)
        lib().orc_set_m(self._c, _dp(m))

    def m(self) -> np.ndarray:
        m = np.empty(3 * self.N)
        lib().orc_get_m(self._c, _dp(m))
        return m.reshape(3, self.n[2], self.n[1], self.n[0])

    def init_random_m(self, seed: int):
        lib().orc_init_random_m(self._c, seed)

    # physics ------------------------------------------------------------
    def demag(self, direct: bool = False) -> np.ndarray:
        h = np.empty(3 * self.N)
        fn = lib().orc_demag_direct if direct else lib().orc_demag_fft
        fn(self._c, _dp(h))
        return h.reshape(3, self.n[2], self.n[1], self.n[0])

    def field(self, terms: int = ALL) -> np.ndarray:
        H = np.empty(3 * self.N)
        lib().orc_field(self._c, terms, _dp(H))
        return H.reshape(3, self.n[2], self.n[1], self.n[0])

    def energy(self) -> dict:
        o = np.empty(9)
        lib().orc_energy(self._c, _dp(o))
        return dict(total=o[0], exch=o[1], dmi=o[2], anis=o[3], zeeman=o[4],
                    demag=o[5], m_avg=o[6:9].copy())

    def energy_density(self, terms: int = ALL) -> np.ndarray:
        e = np.empty(self.N)
        lib().orc_energy_density(self._c, terms, _dp(e))
        return e.reshape(self.n[2], self.n[1], self.n[0])

    def delta_e(self, h: np.ndarray, cell: int, mnew) -> float:
        h = np.ascontiguousarray(h, dtype=np.float64).reshape(3 * self.N)
        mn = np.ascontiguousarray(mnew, dtype=np.float64)
        return lib().orc_delta_e(self._c, _dp(h), cell, _dp(mn))

    def trial_delta_e(self, stage: int, sweep: int) -> np.ndarray:
        dE = np.empty(self.N)
        lib().orc_trial_delta_e(self._c, stage, sweep, _dp(dE))
        return dE.reshape(self.n[2], self.n[1], self.n[0])

    def sweep(self, T: float, stage: int, sweep: int) -> int:
        return lib().orc_sweep(self._c, T, stage, sweep)

    def num_colors(self) -> int:
        return lib().orc_num_colors(self._c)

    def cell_color(self, x, y, z) -> int:
        return lib().orc_cell_color(self._c, x, y, z)
