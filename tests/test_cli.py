"""The command-line annealer (drop-in for the reference executable)."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR
from test_io import README_INI

CLI = PKG_DIR / "_lib" / "mmc_run"


def _ini(tmp_path, name="run", **over):
    text = README_INI
    for k, v in over.items():
        text = re.sub(rf"^{k}\s*=.*$", f"{k} = {v}", text, flags=re.M)
        if not re.search(rf"^{k}\s*=", text, flags=re.M):
            text += f"{k} = {v}\n"
    f = tmp_path / f"{name}.ini"
    f.write_text(text)
    return f


def test_cli_usage_and_bad_ini(mmc, tmp_path):
    if not CLI.exists():
        mmc.build()
    r = subprocess.run([str(CLI)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "need input file" in r.stdout
    f = _ini(tmp_path, nx="zero")
    r = subprocess.run([str(CLI), str(f)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "nx" in r.stderr


@pytest.mark.gpu
def test_cli_anneal_outputs_and_resume(mmc, tmp_path):
    """Two stages of 20 sweeps on a 16x16 film: energy.txt lines
    "%f %ld %g" every 10 sweeps, snapshots "%06ld_%.2fK_%ld.ovf" every 20,
    a tensor cache written and reused, then a resume from a snapshot."""
    f = _ini(tmp_path, nx=16, ny=16, pbc_x=2, pbc_y=2, Temperature_start=500,
             Temperature_end=300, Temperature_step=0.7, circle_per_stage=20,
             terminal_output_period=10, energy_output_period=10, ms_output_period=20,
             precision="fp64")
    cache = tmp_path / "cache"
    cache.mkdir()
    r = subprocess.run([str(CLI), str(f), "-gpu=0", f'-tensorcache="{cache}/"'],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = tmp_path / "run.output"
    lines = (out / "energy.txt").read_text().splitlines()
    assert len(lines) == 4                       # stages at 500 K and 350 K, sweeps 0 and 10
    T, s, E = lines[2].split()
    assert T == "350.000000" and s == "0" and float(E) != 0.0
    snaps = sorted(p.name for p in out.glob("*.ovf"))
    assert snaps == ["000000_500.00K_0.ovf", "000001_350.00K_0.ovf"]
    caches = list(cache.glob("cache_*.ovf"))
    assert [c.name for c in caches] == ["cache_16_16_1-1e-09_1e-09_8e-10-2_2_0.ovf"]
    assert caches[0].stat().st_size == 6 * 16 * 16 * 1 * 8
    info, m = mmc.ovf_read(str(out / snaps[1]))
    assert (info.xnodes, info.ynodes, info.znodes) == (16, 16, 1)
    assert np.allclose(np.linalg.norm(m, axis=0), 1.0, atol=1e-12)
    # resume from the second snapshot with the cached tensor
    f2 = _ini(tmp_path, name="resume", nx=16, ny=16, pbc_x=2, pbc_y=2, Temperature_start=300,
              Temperature_end=200, Temperature_step=0.5, circle_per_stage=10,
              terminal_output_period=10, energy_output_period=10, ms_output_period=100,
              use_random_init="false", ovf_filename=str(out / snaps[1]), precision="fp64")
    r = subprocess.run([str(CLI), str(f2), f'-tensorcache="{cache}/"'],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines2 = (tmp_path / "resume.output" / "energy.txt").read_text().splitlines()
    assert len(lines2) == 1
    # the resumed start state is the snapshot: same energy as logged then
    assert lines2[0].split()[2] == E


@pytest.mark.gpu
def test_cli_debug_fields(mmc, tmp_path):
    """-debugfields writes the per-term fields [T] and the energy density of
    each snapshot state (the reference's DEBUG_H_FIELD dumps): they equal the
    library's field and energy density of the snapshot's magnetisation."""
    f = _ini(tmp_path, name="dbg", nx=24, ny=20, pbc_x=2, Temperature_start=500,
             Temperature_end=400, Temperature_step=0.7, circle_per_stage=12,
             terminal_output_period=100, energy_output_period=100, ms_output_period=12,
             precision="fp64")
    r = subprocess.run([str(CLI), str(f), "-gpu=0", "-debugfields"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = tmp_path / "dbg.output"
    names = {p.name for p in out.glob("*.ovf")}
    for n in ("B_demag", "B_anis", "B_exch", "B_dmi", "B_zeeman", "B_effective", "E_density"):
        assert f"{n}.ovf" in names
    snap = sorted(n for n in names if n[0].isdigit())[-1]
    info, m = mmc.ovf_read(str(out / snap))
    cfg = mmc.ini_parse(str(f))
    c = mmc.Context(mmc.make_params(nx=24, ny=20, nz=1, pbc_x=cfg.pbc_x, pbc_y=cfg.pbc_y,
                                    pbc_z=cfg.pbc_z, dx=cfg.dx, dy=cfg.dy,
                                    dz=cfg.dz, ms=cfg.ms, aex=cfg.Aex, dind=cfg.Dind,
                                    anis_ux=cfg.anisUx, anis_uy=cfg.anisUy, anis_uz=cfg.anisUz,
                                    ku1=cfg.Ku1, bext_x=cfg.Bextx, bext_y=cfg.Bexty,
                                    bext_z=cfg.Bextz, seed=cfg.randomseed), 0)
    c.build_tensor()
    c.set_m(m)
    mu0 = 4e-7 * 3.14159265358
    _, B = mmc.ovf_read(str(out / "B_effective.ovf"))
    H = c.field(31)
    assert np.abs(B - mu0 * H).max() <= 1e-12 * np.abs(mu0 * H).max()
    _, Ed = mmc.ovf_read(str(out / "E_density.ovf"))
    e = c.energy_density(31)
    assert np.abs(Ed[2] - e).max() <= 1e-12 * np.abs(e).max()
