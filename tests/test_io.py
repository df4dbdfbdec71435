"""ini and OVF boundary formats (host code, no GPU)."""
import struct

import numpy as np
import pytest

from conftest import ROOT

README_INI = """nx = 256
ny = 256
nz = 1
dx = 1e-9
dy = 1e-9
dz = 0.8e-9

Bextx = 0
Bexty = 0
Bextz = 0.3

pbc_x = 20
pbc_y = 20
pbc_z = 0

ms = 1.6e6
Aex = 16e-12
Dind = 6e-3

anisUx = 0
anisUy = 0
anisUz = 1
Ku1 = 2e6

use_random_init = true
randomseed = 256
ovf_filename = B_0_175_D_5_500_nx_100_ny_58_r_15_1.ovf

Temperature_start = 1400
Temperature_end = 40
Temperature_step = 0.90
Temperature_use_exp = true

circle_per_stage  = 100000
terminal_output_period = 10001
energy_output_period = 10001
ms_output_period = 10001

rand_block_size = 40

cal_demag_flag = true
"""


def test_ini_readme_example(mmc, tmp_path):
    """The example input of README.md:44-85 parses to the stated values."""
    f = tmp_path / "a.ini"
    f.write_text(README_INI)
    c = mmc.ini_parse(str(f))
    assert (c.nx, c.ny, c.nz) == (256, 256, 1)
    assert (c.dx, c.dy, c.dz) == (1e-9, 1e-9, 0.8e-9)
    assert (c.Bextx, c.Bexty, c.Bextz) == (0, 0, 0.3)
    assert (c.pbc_x, c.pbc_y, c.pbc_z) == (20, 20, 0)
    assert (c.ms, c.Aex, c.Dind, c.Ku1) == (1.6e6, 16e-12, 6e-3, 2e6)
    assert (c.anisUx, c.anisUy, c.anisUz) == (0, 0, 1)
    assert c.use_random_init == 1 and c.randomseed == 256
    assert c.ovf_filename == b"B_0_175_D_5_500_nx_100_ny_58_r_15_1.ovf"
    assert (c.Temperature_start, c.Temperature_end, c.Temperature_step) == (1400, 40, 0.90)
    assert c.Temperature_use_exp == 1 and c.circle_per_stage == 100000
    assert (c.terminal_output_period, c.energy_output_period, c.ms_output_period) == (10001,) * 3
    assert c.rand_block_size == 40 and c.cal_demag_flag == 1
    assert c.precision == mmc.FP32 and c.ovf_format == 8


def test_ini_comments_and_options(mmc, tmp_path):
    text = "; comment\n# another\n[section]\n" + README_INI.replace(
        "nx = 256", "nx = 64 ; inline").replace("cal_demag_flag = true", "cal_demag_flag = 0")
    text += "precision = fp64\novf_format = 4\n"
    f = tmp_path / "b.ini"
    f.write_text(text)
    c = mmc.ini_parse(str(f))
    assert c.nx == 64 and c.cal_demag_flag == 0
    assert c.precision == mmc.FP64 and c.ovf_format == 4


@pytest.mark.parametrize("edit,needle", [
    (("Aex = 16e-12\n", ""), "Aex"),
    (("nx = 256", "nx = abc"), "nx"),
    (("use_random_init = true", "use_random_init = maybe"), "use_random_init"),
    (("Temperature_step = 0.90", "Temperature_step = 1.5"), "Temperature_step"),
    (("nz = 1", "nz"), "key = value"),
])
def test_ini_errors(mmc, tmp_path, edit, needle):
    f = tmp_path / "c.ini"
    f.write_text(README_INI.replace(*edit))
    with pytest.raises(mmc.MMCError, match=needle):
        mmc.ini_parse(str(f))


@pytest.mark.parametrize("fmt", [0, 4, 8])
def test_ovf_round_trip(mmc, tmp_path, fmt):
    rng = np.random.default_rng(fmt)
    m = rng.standard_normal((3, 2, 3, 5))
    f = tmp_path / f"m{fmt}.ovf"
    mmc.ovf_write(str(f), m, (1e-9, 2e-9, 0.5e-9), fmt)
    info, back = mmc.ovf_read(str(f))
    assert (info.xnodes, info.ynodes, info.znodes) == (5, 3, 2)
    assert (info.xstepsize, info.ystepsize, info.zstepsize) == (1e-9, 2e-9, 0.5e-9)
    if fmt == 4:
        assert np.array_equal(back, m.astype(np.float32).astype(np.float64))
    else:
        assert np.array_equal(back, m)
    # data order: x fastest, then y, then z (src/ovf_io.h:156-164)
    raw = f.read_bytes()
    if fmt == 8:
        start = raw.index(b"# Begin: Data Binary 8\n") + len(b"# Begin: Data Binary 8\n")
        vals = struct.unpack("<4d", raw[start:start + 32])
        assert vals[0] == 123456789012345.0
        assert vals[1:] == (m[0, 0, 0, 0], m[1, 0, 0, 0], m[2, 0, 0, 0])
        second = struct.unpack("<3d", raw[start + 32:start + 56])
        assert second == (m[0, 0, 0, 1], m[1, 0, 0, 1], m[2, 0, 0, 1])


def test_ovf1_big_endian_binary(mmc, tmp_path):
    """OVF 1.0 binary data is big-endian (OOMMF format spec)."""
    m = np.arange(3 * 4, dtype=np.float64).reshape(3, 1, 2, 2) / 7.0
    hdr = ("# OOMMF: rectangular mesh v1.0\n# Segment count: 1\n# Begin: Segment\n"
           "# Begin: Header\n# xnodes: 2\n# ynodes: 2\n# znodes: 1\n"
           "# xstepsize: 1e-9\n# ystepsize: 1e-9\n# zstepsize: 1e-9\n"
           "# End: Header\n# Begin: Data Binary 8\n").encode()
    data = struct.pack(">d", 123456789012345.0)
    flat = m.reshape(3, -1)
    for i in range(4):
        data += struct.pack(">3d", *flat[:, i])
    f = tmp_path / "old.ovf"
    f.write_bytes(hdr + data + b"\n# End: Data Binary 8\n# End: Segment\n")
    info, back = mmc.ovf_read(str(f))
    assert np.array_equal(back, m)


def test_ovf_rejects_bad_check_value(mmc, tmp_path):
    m = np.zeros((3, 1, 1, 2))
    f = tmp_path / "bad.ovf"
    mmc.ovf_write(str(f), m, (1e-9, 1e-9, 1e-9), 4)
    raw = bytearray(f.read_bytes())
    i = raw.index(b"Binary 4\n") + len(b"Binary 4\n")
    raw[i:i + 4] = struct.pack("<f", 1.0)
    f.write_bytes(bytes(raw))
    with pytest.raises(mmc.MMCError, match="check value"):
        mmc.ovf_read(str(f))


def test_mesh_equality_tolerance(mmc):
    a = mmc.OvfInfo(4, 4, 1, 1e-9, 1e-9, 1e-9)
    b = mmc.OvfInfo(4, 4, 1, 1e-9 * (1 + 1e-6), 1e-9, 1e-9)
    c = mmc.OvfInfo(4, 4, 1, 1e-9 * (1 + 1e-4), 1e-9, 1e-9)
    d = mmc.OvfInfo(4, 5, 1, 1e-9, 1e-9, 1e-9)
    L = mmc.lib()
    assert L.mmc_ovf_mesh_equal(a, b) == 1
    assert L.mmc_ovf_mesh_equal(a, c) == 0
    assert L.mmc_ovf_mesh_equal(a, d) == 0
