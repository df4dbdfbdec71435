"""mmc_energy_noalloc: the energies of mmc_energy, computed with the
library's fixed device scratch array instead of context memory, through the
C ABI."""
import ctypes as C

import numpy as np
import pytest

from helpers import film, random_m

pytestmark = pytest.mark.gpu


def _noalloc(mmc, c, inst):
    E = C.c_double()
    t = np.empty(5)
    mav = np.empty(3)
    dp = C.POINTER(C.c_double)
    rc = mmc.lib().mmc_energy_noalloc(c._ptr, inst, C.byref(E), t.ctypes.data_as(dp),
                                      mav.ctypes.data_as(dp))
    return rc, dict(total=E.value, exch=t[0], dmi=t[1], anis=t[2], zeeman=t[3], demag=t[4],
                    m_avg=mav)


def _same(a, b):
    for k in ("total", "exch", "dmi", "anis", "zeeman", "demag"):
        assert a[k] == b[k], k
    assert np.array_equal(a["m_avg"], b["m_avg"])


def test_noalloc_energy_equals_energy(mmc):
    """Two calls in a row, the second on a smaller context, then the first
    again after sweeps: bit-identical to mmc_energy each time."""
    kws = [film(nx=64, ny=64, seed=1), film(nx=96, ny=40, pbc_x=1, seed=2),
           film(nx=48, ny=32, nz=3, seed=3)]
    big = mmc.Context([mmc.make_params(**kw) for kw in kws], 1)
    big.build_tensor()
    big.init_random_m()
    for i in range(3):
        rc, e = _noalloc(mmc, big, i)
        assert rc == 0, big._err()
        _same(e, big.energy(i))
    small = mmc.Context(mmc.make_params(**film(nx=16, ny=12, seed=4)), 0)
    small.build_tensor()
    small.set_m(random_m((16, 12, 1), 8))
    rc, e = _noalloc(mmc, small, 0)
    assert rc == 0, small._err()
    _same(e, small.energy(0))
    big.set_temperature(700.0)
    big.sweeps(0, 0, 3)
    for i in (2, 0):
        rc, e = _noalloc(mmc, big, i)
        assert rc == 0, big._err()
        _same(e, big.energy(i))


def test_noalloc_energy_refuses_oversize(mmc):
    """A context whose reduction needs more than the fixed scratch (2048^2
    cells: 16384 blocks of 256, above 15872) gets MMC_EINVAL and a message
    naming the limit; mmc_energy still serves it."""
    c = mmc.Context(mmc.make_params(**film(nx=2048, ny=2048, pbc_x=-1, pbc_y=-1,
                                           cal_demag=0)), 1)
    c.init_random_m()
    rc, _ = _noalloc(mmc, c, 0)
    assert rc == 1
    assert "scratch" in c._err() and "15872" in c._err()
    assert np.isfinite(c.energy()["total"])
