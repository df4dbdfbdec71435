"""The anneal driver's stop paths, the open-axis tensor cache, the energy
cache and the 16-seed anneal statistics at the 64x64 film's schedule."""
import ctypes as C
import re
import signal
import subprocess
import time

import numpy as np
import pytest

from helpers import film, random_m, rel_inf
from test_cli import CLI, _ini

pytestmark = pytest.mark.gpu


def _ctx(mmc, kw, precision=1):
    c = mmc.Context(mmc.make_params(**kw), precision)
    c.build_tensor()
    c.init_random_m()
    return c


def _plain(mmc, kw, runs):
    """The state after (temperature, stage, sweeps) runs through mmc_sweeps."""
    c = _ctx(mmc, kw)
    for T, st, n in runs:
        c.set_temperature(T)
        c.sweeps(st, 0, n)
    return c.m()


def test_anneal_callback_stops_at_once(mmc):
    """An energy callback returning non-zero ends mmc_anneal before the
    sweeps that follow it (src/main.cu:1122-1126 stops the same way on its
    Ctrl-C flag): the state is the one the callback saw."""
    kw = film(nx=48, ny=32, seed=9)
    c = _ctx(mmc, kw)
    seen = []

    def on_energy(T, st, sw, E):
        seen.append((st, sw))
        return st == 1 and sw == 10

    stages = c.anneal(1000.0, 100.0, 0.5, True, 30, energy_period=5, on_energy=on_energy)
    assert stages == 1
    assert seen[-1] == (1, 10) and len(seen) == 6 + 3
    assert np.array_equal(c.m(), _plain(mmc, kw, [(1000.0, 0, 30), (500.0, 1, 10)]))


def test_anneal_stop_flag(mmc):
    """The stop flag (the SIGINT flag of src/main.cu:82-88) is polled between
    runs of sweeps: set during stage 1 at sweep 10, the driver finishes the
    run up to the next output point (sweep 15) and returns, stage 1 counted
    as entered."""
    kw = film(nx=48, ny=32, seed=9)
    c = _ctx(mmc, kw)
    stop = C.c_int(0)

    def on_energy(T, st, sw, E):
        if st == 1 and sw == 10:
            stop.value = 1
        return 0

    stages = c.anneal(1000.0, 100.0, 0.5, True, 30, energy_period=5, on_energy=on_energy,
                      stop_flag=stop)
    assert stages == 2
    assert np.array_equal(c.m(), _plain(mmc, kw, [(1000.0, 0, 30), (500.0, 1, 15)]))
    stop.value = 1
    c2 = _ctx(mmc, kw)
    m0 = c2.m()
    assert c2.anneal(1000.0, 100.0, 0.5, True, 30, stop_flag=stop) == 0
    assert np.array_equal(c2.m(), m0)


def test_cli_ctrl_c(mmc, tmp_path):
    """SIGINT after the first energy line: the annealer stops at its next
    poll, exits 0, and leaves energy.txt and the snapshots complete."""
    f = _ini(tmp_path, name="long", nx=32, ny=32, pbc_x=0, pbc_y=0, circle_per_stage=200000,
             terminal_output_period=50000, energy_output_period=20000,
             ms_output_period=40000, precision="fp64")
    p = subprocess.Popen([str(CLI), str(f)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                         text=True)
    energy = tmp_path / "long.output" / "energy.txt"
    t0 = time.time()
    try:
        while time.time() - t0 < 90:
            if energy.exists() and len(energy.read_text().splitlines()) >= 2:
                break
            if p.poll() is not None:
                break
            time.sleep(0.05)
        assert p.poll() is None, "annealer ended before the interrupt"
        p.send_signal(signal.SIGINT)
        out, err = p.communicate(timeout=60)
    finally:
        if p.poll() is None:
            p.kill()
            p.communicate()
    assert p.returncode == 0, err[-2000:]
    assert "Interrupted." in out
    lines = energy.read_text().splitlines()
    assert 2 <= len(lines) < 34 * 10
    for ln in lines:
        assert re.fullmatch(r"\d+\.\d{6} \d+ \S+", ln), ln
        float(ln.split()[2])
    snaps = sorted((tmp_path / "long.output").glob("*.ovf"))
    assert snaps
    for s in snaps:
        info, m = mmc.ovf_read(str(s))
        assert (info.xnodes, info.ynodes, info.znodes) == (32, 32, 1)
        assert np.allclose(np.linalg.norm(m, axis=0), 1.0, atol=1e-12)


def test_open_axis_tensor_cache(orc, mmc, tmp_path):
    """The tensor cache of a grid open on every axis is written in the
    reference's layout (6 x (2nx-1)(2ny-1)(2nz-1) doubles, z fastest, offset
    o at index o mod (2n-1), src/main.cu:419-573 with the index map of
    src/demag_cu.cu:149-151), read back into the build's own padding
    (2,3,5-smooth lengths), and gives the bit-identical tensor and energy."""
    kw = film(nx=12, ny=10, nz=3)
    c = mmc.Context(mmc.make_params(**kw), 0)
    assert c.P() == (24, 20, 5)
    c.build_tensor(str(tmp_path))
    f = tmp_path / "cache_23_19_5-1e-09_1e-09_8e-10-0_0_0.ovf"
    assert f.stat().st_size == 6 * 23 * 19 * 5 * 8
    raw = np.fromfile(f, dtype="<f8").reshape(6, 23, 19, 5)
    K = c.tensor()
    for ox in range(-11, 12):
        for oy in range(-9, 10):
            for oz in range(-2, 3):
                assert np.array_equal(raw[:, ox % 23, oy % 19, oz % 5],
                                      K[:, oz % 5, oy % 20, ox % 24])
    o = orc.Oracle(**kw)
    o.build_tensor()
    assert rel_inf(K, o.tensor()) <= 1e-9
    c2 = mmc.Context(mmc.make_params(**kw), 0)
    c2.set_option("timing", 1)
    c2.build_tensor(str(tmp_path))
    assert c2.kernel_stats("tensor")[0] == 0          # loaded, not computed
    assert np.array_equal(c2.tensor(), K)
    m = random_m((12, 10, 3), 5)
    c.set_m(m)
    c2.set_m(m)
    assert c2.energy()["total"] == c.energy()["total"]


def test_energy_is_cached_per_state(mmc):
    """Reading the energies of every instance of a batch evaluates the
    batch once; a sweep, a new state or a new tensor evaluates it again."""
    kws = [film(nx=64, ny=64, seed=1), film(nx=96, ny=40, seed=2), film(nx=32, ny=32, seed=3)]
    b = mmc.Context([mmc.make_params(**kw) for kw in kws], 1)
    b.build_tensor()
    b.init_random_m()
    b.set_temperature(800.0)
    b.set_option("timing", 1)
    E = [b.energy(i)["total"] for i in range(3)]
    assert b.kernel_stats("energy", reset=True)[0] == 1
    assert [b.energy(i)["total"] for i in range(3)] == E
    assert b.kernel_stats("energy", reset=True)[0] == 0
    b.sweeps(0, 0, 2)
    E2 = [b.energy(i)["total"] for i in range(3)]
    assert b.kernel_stats("energy", reset=True)[0] == 1
    assert E2 != E
    b.set_m(random_m((96, 40, 1), 3), 1)
    b.energy(0)
    assert b.kernel_stats("energy", reset=True)[0] == 1
    b.set_tensor(b.tensor(2), 2)
    b.energy(2)
    assert b.kernel_stats("energy", reset=True)[0] == 1
    # the cached value is the value a fresh context computes for the state
    c = mmc.Context(mmc.make_params(**kws[1]), 1)
    c.build_tensor()
    c.set_m(b.m(1))
    assert c.energy()["total"] == b.energy(1)["total"]


def _stages(T0=1400.0, T1=40.0, f=0.9):
    out, T = [], T0
    while T > T1:
        out.append(T)
        T *= f
    return out


def test_anneal_statistics_film64_schedule(orc, mmc):
    """16 seeds of the 64x64 open film annealed on its own schedule (1400 K
    x0.9 to 40 K, 34 stages of 1000 sweeps), batched, in FP64 and in FP32
    bulk: the seed means of the final energy density and <m_z> agree within
    four combined standard errors, and the FP32 path's energy of its final
    states is within 1e-4 of the FP64 oracle's evaluation of the same
    states."""
    stages = _stages()
    assert len(stages) == 34
    kws = [film(nx=64, ny=64, nz=1, seed=s) for s in range(1, 17)]
    V = 64 * 64 * kws[0]["dx"] * kws[0]["dy"] * kws[0]["dz"]
    res = {}
    for prec in (0, 1):
        b = mmc.Context([mmc.make_params(**kw) for kw in kws], prec)
        b.build_tensor()
        b.init_random_m()
        b.set_option("graph", 1)
        b.set_option("graph_sweeps", 20)
        for st, T in enumerate(stages):
            b.set_temperature(T)
            b.sweeps(st, 0, 1000)
        e = np.array([b.energy(i)["total"] / V for i in range(16)])
        mz = np.array([b.energy(i)["m_avg"][2] for i in range(16)])
        res[prec] = (e, mz, [b.m(i) for i in range(16)])
        b.close()
    for k, name in ((0, "energy density"), (1, "<m_z>")):
        a, f = res[0][k], res[1][k]
        se = np.sqrt(a.var(ddof=1) / 16 + f.var(ddof=1) / 16)
        print(f"{name}: fp64 {a.mean():.6e} +- {a.std(ddof=1) / 4:.2e}, "
              f"fp32 {f.mean():.6e} +- {f.std(ddof=1) / 4:.2e}")
        assert abs(a.mean() - f.mean()) <= 4 * se + 1e-12 * abs(a.mean())
    worst = 0.0
    for i in (0, 5, 10, 15):
        o = orc.Oracle(**kws[i])
        o.build_tensor()
        o.set_m(res[1][2][i])
        e64 = o.energy()["total"] / V
        worst = max(worst, abs(res[1][0][i] - e64) / abs(e64))
    print(f"fp32 final-energy error vs the FP64 oracle {worst:.3e}")
    assert worst <= 1e-4
