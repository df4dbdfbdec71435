"""MI355X Monte Carlo annealing library: Python view of the C ABI (include/mmc.h).

The compute lives in ``_lib/libmmc.so`` (hand-written gfx950 HIP kernels plus
the C ABI); this module only marshals numpy arrays through ctypes.  There is
no CPU fallback: if the shared library is missing, every entry point raises.

Names and argument meanings follow the reference's ini keys
(README.md:44-85 of chg-hou/Micromagnetic-GPU-Monte-Carlo) and the per-call
boundary documented in include/mmc.h.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmmc.so")

FP64, FP32 = 0, 1
EXCH, DMI, ANIS, ZEEMAN, DEMAG, ALL = 1, 2, 4, 8, 16, 31
KINDS = ("tensor", "xfwd", "zfwd", "ymul", "zinv", "xinv", "metro", "energy", "xinv_metro",
         "metro_xfwd", "yzconv", "row_sweep")


class MMCError(RuntimeError):
    pass


class Params(C.Structure):
    """``mmc_params`` (include/mmc.h)."""

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

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Schedule(C.Structure):
    """``mmc_schedule`` (include/mmc.h)."""

    _fields_ = [
        ("T_start", C.c_double), ("T_end", C.c_double), ("T_step", C.c_double),
        ("use_exp", C.c_int),
        ("sweeps_per_stage", C.c_long),
        ("energy_period", C.c_long),
        ("snapshot_period", C.c_long),
    ]


ENERGY_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_long, C.c_long,
                        C.c_double)
SNAPSHOT_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_long, C.c_long,
                          C.POINTER(C.c_double), C.POINTER(C.c_double),
                          C.POINTER(C.c_double))


class OvfInfo(C.Structure):
    """``mmc_ovf_info`` (include/mmc_io.h)."""

    _fields_ = [("xnodes", C.c_int), ("ynodes", C.c_int), ("znodes", C.c_int),
                ("xstepsize", C.c_double), ("ystepsize", C.c_double),
                ("zstepsize", C.c_double)]


class Config(C.Structure):
    """``mmc_config`` (include/mmc_io.h): every ini key."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("Bextx", C.c_double), ("Bexty", C.c_double), ("Bextz", C.c_double),
        ("pbc_x", C.c_int), ("pbc_y", C.c_int), ("pbc_z", C.c_int),
        ("ms", C.c_double), ("Aex", C.c_double), ("Dind", C.c_double),
        ("anisUx", C.c_double), ("anisUy", C.c_double), ("anisUz", C.c_double),
        ("Ku1", C.c_double),
        ("use_random_init", C.c_int),
        ("randomseed", C.c_longlong),
        ("ovf_filename", C.c_char * 4096),
        ("Temperature_start", C.c_double), ("Temperature_end", C.c_double),
        ("Temperature_step", C.c_double),
        ("Temperature_use_exp", C.c_int),
        ("circle_per_stage", C.c_long),
        ("terminal_output_period", C.c_long),
        ("energy_output_period", C.c_long),
        ("ms_output_period", C.c_long),
        ("rand_block_size", C.c_int),
        ("cal_demag_flag", C.c_int),
        ("precision", C.c_int),
        ("ovf_format", C.c_int),
    ]


# every entry point declared in include/mmc.h and include/mmc_io.h
EXPORTS = (
    "mmc_create", "mmc_batch_create", "mmc_destroy", "mmc_last_error",
    "mmc_dims", "mmc_count", "mmc_build_tensor", "mmc_set_tensor",
    "mmc_get_tensor", "mmc_mirror_tensor", "mmc_set_m", "mmc_get_m",
    "mmc_init_random_m", "mmc_field", "mmc_energy_density", "mmc_trial_delta_e", "mmc_energy",
    "mmc_energy_noalloc",
    "mmc_set_temperature",
    "mmc_sweeps", "mmc_autotune", "mmc_set_schedule", "mmc_sweep", "mmc_sync", "mmc_anneal", "mmc_set_option",
    "mmc_kernel_stats", "mmc_kernel_bytes",
    "mmc_ovf_read", "mmc_ovf_write", "mmc_ovf_mesh_equal", "mmc_ini_parse",
)


def make_params(**kw) -> Params:
    """Parameters with the reference README's example values as defaults."""
    p = Params()
    d = dict(nx=64, ny=64, nz=1, dx=1e-9, dy=1e-9, dz=0.8e-9, pbc_x=0, pbc_y=0,
             pbc_z=0, ms=1.6e6, aex=16e-12, dind=6e-3, anis_ux=0.0,
             anis_uy=0.0, anis_uz=1.0, ku1=2e6, bext_x=0.0, bext_y=0.0,
             bext_z=0.3, cal_demag=1, reserved=0, seed=256)
    d.update(kw)
    for k, v in d.items():
        setattr(p, k, v)
    return p


def build(force: bool = False) -> str:
    """Compile libmmc.so with hipcc for gfx950 (the package Makefile)."""
    args = ["make", "-s", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(args, check=True)
    return LIB_PATH


_lib = None


def lib() -> C.CDLL:
    """Load libmmc.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MMCError(f"{LIB_PATH} is missing: build it with "
                       "__graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    L.mmc_create.argtypes = [C.POINTER(Params), C.c_int, C.POINTER(vp)]
    L.mmc_batch_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int,
                                   C.POINTER(vp)]
    L.mmc_destroy.argtypes = [vp]
    L.mmc_destroy.restype = None
    L.mmc_last_error.argtypes = [vp]
    L.mmc_last_error.restype = C.c_char_p
    L.mmc_dims.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.mmc_count.argtypes = [vp]
    L.mmc_build_tensor.argtypes = [vp, C.c_char_p]
    L.mmc_set_tensor.argtypes = [vp, C.c_int, dp]
    L.mmc_get_tensor.argtypes = [vp, C.c_int, dp]
    L.mmc_mirror_tensor.argtypes = [vp, C.c_int, dp]
    L.mmc_set_m.argtypes = [vp, C.c_int, dp, dp, dp]
    L.mmc_get_m.argtypes = [vp, C.c_int, dp, dp, dp]
    L.mmc_init_random_m.argtypes = [vp, C.c_int]
    L.mmc_field.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp]
    L.mmc_trial_delta_e.argtypes = [vp, C.c_int, C.c_uint32, C.c_uint32, dp]
    L.mmc_energy_density.argtypes = [vp, C.c_int, C.c_int, dp]
    L.mmc_autotune.argtypes = [vp, C.c_int, C.c_char_p, C.c_size_t]
    L.mmc_set_schedule.argtypes = [vp, C.c_char_p]
    L.mmc_energy.argtypes = [vp, C.c_int, dp, dp, dp]
    L.mmc_energy_noalloc.argtypes = [vp, C.c_int, dp, dp, dp]
    L.mmc_set_temperature.argtypes = [vp, C.c_int, C.c_double]
    L.mmc_sweeps.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32]
    L.mmc_sweep.argtypes = [vp, C.c_double, C.c_uint32, C.c_uint32]
    L.mmc_sync.argtypes = [vp]
    L.mmc_anneal.argtypes = [vp, C.POINTER(Schedule), ENERGY_CB, SNAPSHOT_CB,
                             vp, C.POINTER(C.c_int), C.POINTER(C.c_long)]
    L.mmc_set_option.argtypes = [vp, C.c_char_p, C.c_long]
    L.mmc_kernel_stats.argtypes = [vp, C.c_char_p, C.POINTER(C.c_long), dp, dp,
                                   C.c_int]
    L.mmc_kernel_bytes.argtypes = [vp, C.c_char_p, dp]
    L.mmc_ovf_read.argtypes = [C.c_char_p, C.POINTER(OvfInfo), dp, dp, dp, C.c_long,
                               C.c_char_p, C.c_int]
    L.mmc_ovf_write.argtypes = [C.c_char_p, C.POINTER(OvfInfo), dp, dp, dp, C.c_int,
                                C.c_char_p]
    L.mmc_ovf_mesh_equal.argtypes = [C.POINTER(OvfInfo), C.POINTER(OvfInfo)]
    L.mmc_ini_parse.argtypes = [C.c_char_p, C.POINTER(Config), C.c_char_p, C.c_int]
    _lib = L
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ini_parse(path: str) -> Config:
    cfg = Config()
    msg = C.create_string_buffer(1024)
    rc = lib().mmc_ini_parse(path.encode(), C.byref(cfg), msg, 1024)
    if rc != 0:
        raise MMCError(f"ini: {msg.value.decode()}")
    return cfg


def ovf_read(path: str):
    """Returns (OvfInfo, m) with m of shape (3, znodes, ynodes, xnodes)."""
    L = lib()
    info = OvfInfo()
    msg = C.create_string_buffer(1024)
    rc = L.mmc_ovf_read(path.encode(), C.byref(info), None, None, None, 0, msg, 1024)
    if rc != 0:
        raise MMCError(f"ovf: {msg.value.decode()}")
    n = info.xnodes * info.ynodes * info.znodes
    m = np.empty((3, info.znodes, info.ynodes, info.xnodes))
    rc = L.mmc_ovf_read(path.encode(), C.byref(info), _dp(m[0]), _dp(m[1]), _dp(m[2]), n,
                        msg, 1024)
    if rc != 0:
        raise MMCError(f"ovf: {msg.value.decode()}")
    return info, m


def ovf_write(path: str, m: np.ndarray, steps, fmt: int = 8, title: str = "m"):
    m = np.ascontiguousarray(m, dtype=np.float64)
    _, nz, ny, nx = m.shape
    info = OvfInfo(nx, ny, nz, *steps)
    rc = lib().mmc_ovf_write(path.encode(), C.byref(info), _dp(m[0]), _dp(m[1]), _dp(m[2]),
                             fmt, title.encode())
    if rc != 0:
        raise MMCError(f"ovf write failed ({rc})")


class Context:
    """One library context: a single instance or a batch of instances."""

    def __init__(self, params, precision: int = FP64):
        L = lib()
        if isinstance(params, Params):
            params = [params]
        arr = (Params * len(params))(*params)
        self._ptr = C.c_void_p()
        self._arr = arr
        rc = L.mmc_batch_create(arr, len(params), precision, C.byref(self._ptr))
        if rc != 0:
            msg = self._err()
            if self._ptr:
                L.mmc_destroy(self._ptr)
                self._ptr = C.c_void_p()
            raise MMCError(f"mmc_batch_create failed ({rc}): {msg}")
        self.precision = precision
        self.count = len(params)
        self.dims = []
        for i in range(self.count):
            d = (C.c_int * 6)()
            self._check(L.mmc_dims(self._ptr, i, d))
            self.dims.append(tuple(d))

    def _err(self) -> str:
        if not self._ptr:
            return "(no context)"
        return lib().mmc_last_error(self._ptr).decode()

    def _check(self, rc: int):
        if rc != 0:
            raise MMCError(f"libmmc error {rc}: {self._err()}")

    def close(self):
        if getattr(self, "_ptr", None):
            lib().mmc_destroy(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # geometry ------------------------------------------------------------
    def n(self, inst=0):
        return self.dims[inst][:3]

    def P(self, inst=0):
        return self.dims[inst][3:]

    def N(self, inst=0):
        nx, ny, nz = self.n(inst)
        return nx * ny * nz

    # tensor --------------------------------------------------------------
    def build_tensor(self, cache_dir: str | None = None):
        arg = cache_dir.encode() if cache_dir else None
        self._check(lib().mmc_build_tensor(self._ptr, arg))

    def set_tensor(self, K: np.ndarray, inst: int = 0):
        K = np.ascontiguousarray(K, dtype=np.float64)
        Px, Py, Pz = self.P(inst)
        assert K.size == 6 * Px * Py * Pz
        self._check(lib().mmc_set_tensor(self._ptr, inst, _dp(K)))

    def tensor(self, inst: int = 0) -> np.ndarray:
        Px, Py, Pz = self.P(inst)
        K = np.empty((6, Pz, Py, Px))
        self._check(lib().mmc_get_tensor(self._ptr, inst, _dp(K)))
        return K

    def mirror(self, inst: int = 0) -> np.ndarray:
        K0 = np.empty(6)
        self._check(lib().mmc_mirror_tensor(self._ptr, inst, _dp(K0)))
        return K0

    # magnetisation -------------------------------------------------------
    def set_m(self, m: np.ndarray, inst: int = 0):
        nx, ny, nz = self.n(inst)
        m = np.ascontiguousarray(m, dtype=np.float64).reshape(3, -1)
        assert m.shape[1] == nx * ny * nz
        comps = [np.ascontiguousarray(m[k]) for k in range(3)]
        self._check(lib().mmc_set_m(self._ptr, inst, *[_dp(c) for c in comps]))

    def m(self, inst: int = 0) -> np.ndarray:
        nx, ny, nz = self.n(inst)
        out = np.empty((3, nz, ny, nx))
        self._check(lib().mmc_get_m(self._ptr, inst, _dp(out[0]), _dp(out[1]),
                                    _dp(out[2])))
        return out

    def init_random_m(self, inst: int = -1):
        self._check(lib().mmc_init_random_m(self._ptr, inst))

    # observables ---------------------------------------------------------
    def field(self, terms: int = ALL, inst: int = 0) -> np.ndarray:
        nx, ny, nz = self.n(inst)
        out = np.empty((3, nz, ny, nx))
        self._check(lib().mmc_field(self._ptr, inst, terms, _dp(out[0]),
                                    _dp(out[1]), _dp(out[2])))
        return out

    def energy_density(self, terms: int = ALL, inst: int = 0) -> np.ndarray:
        """Energy density [J/m^3] of the selected terms; shape (nz, ny, nx)."""
        nx, ny, nz = self.n(inst)
        out = np.empty((nz, ny, nx))
        self._check(lib().mmc_energy_density(self._ptr, inst, terms, _dp(out)))
        return out

    def trial_delta_e(self, stage: int, sweep: int, inst: int = 0) -> np.ndarray:
        """Energy change [J] of every cell's trial move for (stage, sweep),
        moves not applied; shape (nz, ny, nx)."""
        nx, ny, nz = self.n(inst)
        out = np.empty((nz, ny, nx))
        self._check(lib().mmc_trial_delta_e(self._ptr, inst, stage, sweep, _dp(out)))
        return out

    def energy(self, inst: int = 0) -> dict:
        E = C.c_double()
        t = np.empty(5)
        mav = np.empty(3)
        self._check(lib().mmc_energy(self._ptr, inst, C.byref(E), _dp(t), _dp(mav)))
        return dict(total=E.value, exch=t[0], dmi=t[1], anis=t[2], zeeman=t[3],
                    demag=t[4], m_avg=mav)

    # Monte Carlo ---------------------------------------------------------
    def set_temperature(self, T: float, inst: int = -1):
        self._check(lib().mmc_set_temperature(self._ptr, inst, T))

    def sweeps(self, stage: int, sweep0: int, count: int):
        self._check(lib().mmc_sweeps(self._ptr, stage, sweep0, count))

    def autotune(self, sweeps_per_trial: int = 20) -> str:
        """Time the sweep schedules, keep the fastest; returns the table."""
        buf = C.create_string_buffer(4096)
        self._check(lib().mmc_autotune(self._ptr, sweeps_per_trial, buf, len(buf)))
        return buf.value.decode()

    def set_schedule(self, name: str):
        """Apply a schedule named as in the autotuner's report."""
        self._check(lib().mmc_set_schedule(self._ptr, name.encode()))

    def sweep(self, T: float, stage: int, sweep: int):
        self._check(lib().mmc_sweep(self._ptr, T, stage, sweep))

    def sync(self):
        self._check(lib().mmc_sync(self._ptr))

    def anneal(self, T_start, T_end, T_step, use_exp=True, sweeps_per_stage=100,
               energy_period=0, snapshot_period=0, on_energy=None,
               on_snapshot=None, stop_flag=None):
        s = Schedule(T_start, T_end, T_step, int(use_exp), sweeps_per_stage,
                     energy_period, snapshot_period)
        N = self.N(0)

        def ecb(user, T, stage, sweep, E):
            return int(bool(on_energy(T, stage, sweep, E))) if on_energy else 0

        def scb(user, T, stage, sweep, mx, my, mz):
            if not on_snapshot:
                return 0
            m = np.stack([np.ctypeslib.as_array(p, (N,)).copy() for p in (mx, my, mz)])
            return int(bool(on_snapshot(T, stage, sweep, m)))

        e_cb, s_cb = ENERGY_CB(ecb), SNAPSHOT_CB(scb)
        stop = stop_flag if stop_flag is not None else C.c_int(0)
        stages = C.c_long()
        self._check(lib().mmc_anneal(self._ptr, C.byref(s), e_cb, s_cb, None,
                                     C.byref(stop), C.byref(stages)))
        return stages.value

    # instrumentation -----------------------------------------------------
    def set_option(self, key: str, value: int):
        self._check(lib().mmc_set_option(self._ptr, key.encode(), int(value)))

    def kernel_stats(self, kind: str, reset: bool = False):
        """(launches, total_ms, total_bytes) recorded while timing was on."""
        n = C.c_long()
        ms = C.c_double()
        by = C.c_double()
        self._check(lib().mmc_kernel_stats(self._ptr, kind.encode(), C.byref(n),
                                           C.byref(ms), C.byref(by), int(reset)))
        return n.value, ms.value, by.value

    def kernel_bytes(self, kind: str) -> float:
        b = C.c_double()
        self._check(lib().mmc_kernel_bytes(self._ptr, kind.encode(), C.byref(b)))
        return b.value
