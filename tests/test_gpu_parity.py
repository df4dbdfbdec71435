"""GPU parity: the HIP library against the FP64 oracle on identical inputs.

FP64 mode: fields and energies within 1e-12 (norm-wise relative), using the
oracle's tensor uploaded to the device so the comparison isolates the
convolution, stencil and reductions from the Newell cancellation noise.
FP32-bulk mode: within the bounds written in each test.
"""
import numpy as np
import pytest

from helpers import film, FILM3D, pair, random_m, rel_inf

pytestmark = pytest.mark.gpu

TERMS = [("exch", 1), ("dmi", 2), ("anis", 4), ("zeeman", 8), ("demag", 16), ("all", 31)]

CASES = [
    ("film64_open", dict(nx=64, ny=64, nz=1)),
    ("film48x40_pbc", dict(nx=48, ny=40, nz=1, pbc_x=2, pbc_y=3)),
    ("odd_pbc", dict(nx=15, ny=12, nz=1, pbc_x=1)),
    ("box3d_open", dict(nx=12, ny=10, nz=4)),
    ("box3d_pbcz", dict(nx=10, ny=8, nz=3, pbc_z=2)),
    ("ragged_3d", dict(nx=7, ny=9, nz=5, pbc_y=1)),          # Pz = 9, runtime z plan
    ("two_layer_3d_zconv", dict(nx=32, ny=32, nz=2)),
    ("deep_z_3d", dict(nx=16, ny=16, nz=16)),                # Pz = 32, compile-time z plan
    ("deep_z_pbcz_3d", dict(nx=12, ny=10, nz=16, pbc_z=1)),  # Pz = 16, compile-time z plan
    ("deep_z_ragged_y_3d", dict(nx=8, ny=40, nz=16)),        # z tiles of 32 + 8 y columns
]


@pytest.mark.parametrize("name,geo", CASES, ids=[c[0] for c in CASES])
def test_field_fp64(orc, mmc, name, geo):
    kw = film(**geo) if "3d" not in name else dict(FILM3D, **geo)
    o, c = pair(orc, mmc, **kw)
    m = random_m((kw["nx"], kw["ny"], kw["nz"]), 11)
    o.set_m(m)
    c.set_m(m)
    for tname, bits in TERMS:
        Ho = o.field(bits)
        Hg = c.field(bits)
        err = rel_inf(Hg, Ho)
        assert err <= 1e-12, f"{name} {tname}: rel err {err:.3e}"


@pytest.mark.parametrize("name,geo", CASES, ids=[c[0] for c in CASES])
def test_energy_fp64(orc, mmc, name, geo):
    kw = film(**geo) if "3d" not in name else dict(FILM3D, **geo)
    o, c = pair(orc, mmc, **kw)
    m = random_m((kw["nx"], kw["ny"], kw["nz"]), 12)
    o.set_m(m)
    c.set_m(m)
    Eo = o.energy()
    Eg = c.energy()
    scale = sum(abs(Eo[k]) for k in ("exch", "dmi", "anis", "zeeman", "demag"))
    for k in ("total", "exch", "dmi", "anis", "zeeman", "demag"):
        assert abs(Eg[k] - Eo[k]) <= 1e-12 * scale, f"{name} {k}: {Eg[k]} vs {Eo[k]}"
    assert np.allclose(Eg["m_avg"], Eo["m_avg"], rtol=0, atol=1e-14)


def test_demag_impulse_response(orc, mmc):
    """One-hot m (cell 0 along x, the reference's USE_ONE_HOT_MX_INIT,
    src/main.cu:302-310): the demag field is Ms times one tensor column."""
    kw = film(nx=20, ny=12, nz=1)
    o, c = pair(orc, mmc, **kw)
    m = np.zeros((3, 1, 12, 20))
    m[0, 0, 0, 0] = 1.0
    c.set_m(m)
    H = c.field(16) / kw["ms"]
    K = o.tensor()
    # offsets (x, y) >= 0 sit at padded index (x, y); the xz entries of a
    # one-layer film are zero up to rounding, so errors are measured against
    # the column's scale
    scale = max(np.abs(K[k, 0, :12, :20]).max() for k in (0, 1, 2))
    for comp, kidx in ((0, 0), (1, 1), (2, 2)):
        ref = K[kidx, 0, :12, :20]
        assert np.abs(H[comp, 0] - ref).max() <= 1e-12 * scale


def test_tensor_gpu_vs_oracle(orc, mmc):
    """GPU Newell tensor against the oracle's (same formula, same addition
    order; device vs host libm differ by ulps, amplified by cancellation)."""
    for geo in (dict(nx=8, ny=8, nz=1), dict(nx=6, ny=5, nz=3, pbc_x=2, pbc_z=1),
                dict(nx=16, ny=16, nz=1, pbc_x=3, pbc_y=3)):
        kw = film(**geo)
        o = orc.Oracle(**kw)
        o.build_tensor()
        c = mmc.Context(mmc.make_params(**kw))
        c.build_tensor()
        Kg = c.tensor()
        Ko = o.tensor()
        assert rel_inf(Kg, Ko) <= 1e-9, geo
        assert np.allclose(c.mirror(), o.mirror(), rtol=0, atol=1e-9)


def test_init_random_matches_oracle(orc, mmc):
    kw = film(nx=33, ny=17, nz=2, seed=99)
    o = orc.Oracle(**kw)
    o.init_random_m(99)
    c = mmc.Context(mmc.make_params(**kw))
    c.init_random_m()
    assert rel_inf(c.m(), o.m()) <= 1e-15


@pytest.mark.parametrize("geo", [dict(nx=16, ny=16, nz=1),
                                 dict(nx=12, ny=10, nz=1, pbc_x=1, pbc_y=1),
                                 dict(nx=9, ny=8, nz=1, pbc_x=2),
                                 dict(nx=8, ny=6, nz=3, pbc_z=1),
                                 dict(nx=64, ny=9, nz=1),
                                 dict(nx=64, ny=6, nz=2, pbc_y=1),
                                 dict(nx=64, ny=32, nz=2),
                                 dict(nx=16, ny=64, nz=3, pbc_y=1, pbc_z=1)],
                         ids=["open", "pbc", "odd_pbc_3colors", "3d", "fused_open",
                              "fused_3d", "zconv_fused", "zconv_pbc"])
def test_sweep_trajectory_fp64(orc, mmc, geo):
    """Same Philox stream, same colour order, frozen demag: identical
    accept/reject decisions, so m agrees to rounding after several sweeps."""
    kw = film(**geo)
    o, c = pair(orc, mmc, **kw)
    m = random_m((kw["nx"], kw["ny"], kw["nz"]), 5)
    o.set_m(m)
    c.set_m(m)
    for s, T in enumerate((900.0, 300.0, 60.0)):
        o.sweep(T, 2, s)
        c.sweep(T, 2, s)
    assert rel_inf(c.m(), o.m()) <= 1e-10


def test_fp32_bulk_bounds(orc, mmc):
    """FP32 FFT/stencil against the FP64 oracle (FP64 tensor, downcast K^)."""
    kw = film(nx=64, ny=64, nz=1)
    o, c = pair(orc, mmc, precision=1, **kw)
    m = random_m((64, 64, 1), 3)
    o.set_m(m)
    c.set_m(m)
    assert rel_inf(c.field(31), o.field(31)) <= 2e-6
    assert rel_inf(c.field(16), o.field(16)) <= 2e-6
    Eo, Eg = o.energy(), c.energy()
    scale = sum(abs(Eo[k]) for k in ("exch", "dmi", "anis", "zeeman", "demag"))
    assert abs(Eg["total"] - Eo["total"]) <= 1e-6 * scale


def test_determinism_and_batch(orc, mmc):
    """Bit-identical reruns; a batch of uneven instances gives each
    instance exactly what its own single context gives."""
    geos = [film(nx=16, ny=16, nz=1, seed=1), film(nx=24, ny=20, nz=1, pbc_x=1, seed=2),
            film(nx=12, ny=12, nz=2, bext_z=0.5, seed=3), film(nx=64, ny=10, nz=1, seed=4)]
    singles = []
    for kw in geos:
        c = mmc.Context(mmc.make_params(**kw), 1)
        c.build_tensor()
        c.init_random_m()
        c.set_temperature(500.0)
        c.sweeps(0, 0, 5)
        singles.append(c.m())
    b = mmc.Context([mmc.make_params(**kw) for kw in geos], 1)
    b.build_tensor()
    b.init_random_m()
    b.set_temperature(500.0)
    b.sweeps(0, 0, 5)
    for i, ref in enumerate(singles):
        assert np.array_equal(b.m(i), ref), f"instance {i}"
    b2 = mmc.Context([mmc.make_params(**kw) for kw in geos], 1)
    b2.build_tensor()
    b2.init_random_m()
    b2.set_temperature(500.0)
    b2.sweeps(0, 0, 5)
    for i in range(len(geos)):
        assert np.array_equal(b2.m(i), b.m(i))


@pytest.mark.parametrize("precision", [1, 0], ids=["fp32", "fp64"])
def test_batch_mixed_z_padding(mmc, precision):
    """Two 3-D films sharing the planned z length 32: an open one (nz = 16,
    whose z transforms alone skip the zero half) and a periodic one (nz =
    32).  Batched, the z passes run unpruned for both; each instance must
    still match its own single context bit for bit."""
    geos = [dict(FILM3D, nx=12, ny=10, nz=16, seed=5),
            dict(FILM3D, nx=8, ny=12, nz=32, pbc_z=1, seed=6)]
    singles = []
    for kw in geos:
        c = mmc.Context(mmc.make_params(**kw), precision)
        c.build_tensor()
        c.init_random_m()
        c.set_temperature(600.0)
        c.sweeps(0, 0, 4)
        singles.append((c.m(), c.field()))
    b = mmc.Context([mmc.make_params(**kw) for kw in geos], precision)
    b.build_tensor()
    b.init_random_m()
    b.set_temperature(600.0)
    b.sweeps(0, 0, 4)
    for i, (m, h) in enumerate(singles):
        assert np.array_equal(b.m(i), m), f"instance {i}"
        assert np.array_equal(b.field(inst=i), h), f"instance {i} field"


def _run(mmc, kw, precision, n, graph=False, tyx=None, ncol=None, chunk=None):
    c = mmc.Context(mmc.make_params(**kw), precision)
    c.build_tensor()
    c.init_random_m()
    c.set_temperature(700.0)
    if tyx:
        c.set_option("tyx", tyx)
    if ncol:
        c.set_option("ncol", ncol)
    if graph:
        c.set_option("graph", 1)
        c.set_option("graph_sweeps", 4)
    out = []
    step = chunk or n
    for s0 in range(0, n, step):
        c.sweeps(1, s0, step)
        out.append(c.m())
    return out


@pytest.mark.parametrize("geo", [dict(nx=32, ny=24, nz=1, pbc_x=2),
                                 dict(nx=16, ny=16, nz=4),
                                 dict(nx=12, ny=10, nz=5, pbc_y=1),
                                 dict(nx=16, ny=16, nz=16),
                                 dict(nx=12, ny=10, nz=16, pbc_z=1)],
                         ids=["film", "3d_direct_z", "3d_lds_z", "3d_pz32", "3d_pz16"])
def test_graph_replay_matches_eager(mmc, geo):
    """hipGraph replays (4 sweeps per graph, replayed with eager work and host
    reads in between) give bit-identical states to eager launches after
    every chunk; sweeps not a multiple of the graph length run eagerly."""
    kw = film(**geo)
    eager = _run(mmc, kw, 1, 30, chunk=10)
    graph = _run(mmc, kw, 1, 30, graph=True, chunk=10)
    for a, b in zip(eager, graph):
        assert np.array_equal(a, b)


def test_tiles_do_not_change_results(mmc):
    """Row tiles and column groups only regroup independent transforms."""
    kw = film(nx=40, ny=36, nz=1, pbc_y=1)
    ref = _run(mmc, kw, 0, 6)[-1]
    for tyx, ncol in ((2, 1), (4, 3), (16, 8)):
        assert np.array_equal(_run(mmc, kw, 0, 6, tyx=tyx, ncol=ncol)[-1], ref)


@pytest.mark.parametrize("geo", [dict(nx=40, ny=36, nz=1),
                                 dict(nx=64, ny=20, nz=1, pbc_x=1, pbc_y=2),
                                 dict(nx=2, ny=6, nz=1, pbc_x=1, pbc_y=1),
                                 dict(nx=70, ny=10, nz=3, pbc_y=1),
                                 dict(nx=33, ny=17, nz=2, pbc_z=1)],
                         ids=["open", "pbc", "tiny_pbc", "3d_ragged", "3d_pbcz"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_tiled_metropolis_matches_per_colour(mmc, geo, precision):
    """Both colours in one tiled launch (halo ring recomputed per tile) give
    the bit-identical chain of one launch per colour."""
    kw = film(**geo)
    out = []
    for tiled in (0, 1):
        c = mmc.Context(mmc.make_params(**kw), precision)
        c.build_tensor()
        c.init_random_m()
        c.set_temperature(600.0)
        c.set_option("metro_tiled", tiled)
        c.sweeps(3, 0, 7)
        out.append(c.m())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("geo", [dict(nx=64, ny=40, nz=1),
                                 dict(nx=128, ny=64, nz=1, pbc_x=1, pbc_y=2),
                                 dict(nx=91, ny=37, nz=1),
                                 dict(nx=64, ny=24, nz=3),
                                 dict(nx=64, ny=16, nz=4, pbc_z=1),
                                 dict(nx=256, ny=50, nz=1, pbc_x=3),
                                 dict(nx=128, ny=37, nz=2),
                                 dict(nx=128, ny=10, nz=2, pbc_x=1, pbc_y=1),
                                 dict(nx=64, ny=16, nz=16),
                                 dict(nx=128, ny=10, nz=16, pbc_z=1),
                                 dict(nx=64, ny=64, nz=1),
                                 dict(nx=64, ny=63, nz=1),
                                 dict(nx=64, ny=2, nz=1, pbc_y=1),
                                 dict(nx=64, ny=6, nz=1, pbc_y=1)],
                         ids=["open", "pbc", "odd_open", "3d_open", "3d_pbcz", "wide_pbc",
                              "zfold_open", "zfold_pbc", "deep_z_pz32", "deep_z_pz16",
                              "planned_y_pair_major", "planned_y_odd_rows",
                              "two_periodic_rows", "six_periodic_rows"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_fused_sweeps_match_separate(mmc, geo, precision):
    """Inverse x pass fused with colour 0 and colour 1 fused with the next
    forward x pass (one row pair or full tiles per workgroup, eager and
    graph-replayed) give the bit-identical chain and energies of separate
    launches."""
    kw = film(**geo)
    out = []
    for fuse, rows, graph, zfold in ((0, 1, 0, 0), (1, 1, 0, 0), (1, 2, 0, 0), (1, 2, 1, 0),
                                     (1, 2, 1, 1)):
        c = mmc.Context(mmc.make_params(**kw), precision)
        c.build_tensor()
        c.init_random_m()
        c.set_temperature(650.0)
        c.set_option("fuse", fuse)
        c.set_option("fuse_rows", rows)
        c.set_option("zfold", zfold)
        if graph:
            c.set_option("graph", 1)
            c.set_option("graph_sweeps", 3)
        c.sweeps(2, 0, 4)
        m1 = c.m()
        c.sweeps(2, 4, 5)
        out.append((m1, c.m(), c.energy()["total"]))
    for m1, m2, e in out[1:]:
        assert np.array_equal(m1, out[0][0])
        assert np.array_equal(m2, out[0][1])
        assert e == out[0][2]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_khat_planes_stored_compactly(mmc, precision):
    """FP32 bulk: the Khat planes of an open film are real to far below FP32
    resolution and its xz, yz planes vanish, so the y pass reads four real
    planes; an arbitrary tensor keeps six complex planes.  FP64 keeps every
    plane complex (the imaginary parts left by the Newell cancellation
    noise, ~1e-10 of the largest value, matter at 1e-12)."""
    b = 8 if precision == 0 else 4
    c = mmc.Context(mmc.make_params(**film(nx=64, ny=64, nz=1)), precision)
    c.build_tensor()
    Px, Py, Pz = c.P()
    H = Px // 2 + 1
    spec = 2 * 3 * H * Pz * c.n()[1] * 2 * b
    plane = H * Pz * Py * b
    assert c.kernel_bytes("ymul") == spec + (4 if precision == 1 else 12) * plane
    K = np.random.default_rng(7).standard_normal(c.tensor().shape)
    c.set_tensor(K)
    assert c.kernel_bytes("ymul") == spec + 12 * plane


@pytest.mark.parametrize("name,geo", CASES, ids=[c[0] for c in CASES])
def test_trial_delta_e_fp64(orc, mmc, name, geo):
    """dE of every cell's trial move (the quantity the Metropolis kernels
    test) matches the oracle's within 1e-12 of its largest value."""
    kw = film(**geo) if "3d" not in name else dict(FILM3D, **geo)
    o, c = pair(orc, mmc, **kw)
    m = random_m((kw["nx"], kw["ny"], kw["nz"]), 12)
    o.set_m(m)
    c.set_m(m)
    for stage, sweep in ((0, 0), (5, 77)):
        assert rel_inf(c.trial_delta_e(stage, sweep), o.trial_delta_e(stage, sweep)) <= 1e-12


def test_fp32_delta_e_bound(orc, mmc):
    """FP32-bulk dE against the FP64 oracle on the 256x256 periodic film
    (20 images): within 2e-6 of the largest |dE| of the sweep."""
    kw = film(nx=256, ny=256, nz=1, pbc_x=20, pbc_y=20)
    c32 = mmc.Context(mmc.make_params(**kw), 1)
    c32.build_tensor()
    o = orc.Oracle(**kw)
    o.set_tensor(c32.tensor())
    m = random_m((256, 256, 1), 21)
    o.set_m(m)
    c32.set_m(m)
    err = rel_inf(c32.trial_delta_e(1, 3), o.trial_delta_e(1, 3))
    print(f"fp32 dE rel err {err:.3e}")
    assert err <= 2e-6


def _anneal(mmc, kws, precision, sweeps_per_stage, stages):
    b = mmc.Context([mmc.make_params(**kw) for kw in kws], precision)
    b.build_tensor()
    b.init_random_m()
    b.set_option("graph", 1)
    for st, T in enumerate(stages):
        b.set_temperature(T)
        b.sweeps(st, 0, sweeps_per_stage)
    return b


def _stages(T0=1400.0, T1=40.0, f=0.9):
    out, T = [], T0
    while T > T1:
        out.append(T)
        T *= f
    return out


def test_anneal_fp64_matches_oracle(orc, mmc):
    """Whole anneal (1400 K -> x0.9 -> 40 K, 34 stages) on the 32x32 open
    film: the FP64 library follows the oracle's Markov chain sweep for sweep
    (same Philox moves, same colour order), so final states agree to
    rounding."""
    stages = _stages()
    assert len(stages) == 34
    for seed in (1, 2):
        kw = film(nx=32, ny=32, nz=1, seed=seed)
        o = orc.Oracle(**kw)
        o.build_tensor()
        o.init_random_m(seed)
        c = mmc.Context(mmc.make_params(**kw), 0)
        c.set_tensor(o.tensor())
        c.init_random_m()
        for st, T in enumerate(stages):
            c.set_temperature(T)
            c.sweeps(st, 0, 12)
            for k in range(12):
                o.sweep(T, st, k)
        assert rel_inf(c.m(), o.m()) <= 1e-9
        assert abs(c.energy()["total"] - o.energy()["total"]) <= 1e-9 * abs(o.energy()["total"])


def test_anneal_statistics_16_seeds(orc, mmc):
    """16 seeds of the full anneal of the 32x32 open film (300 sweeps per
    stage): FP32-bulk and FP64 final energy density and <m_z> agree within
    four combined standard errors of the seed mean; the FP32 path's energy of
    its own final states is within 1e-6 of the FP64 oracle's (the final-energy
    error of the arithmetic, target 1e-4)."""
    stages = _stages()
    kws = [film(nx=32, ny=32, nz=1, seed=s) for s in range(1, 17)]
    V = 32 * 32 * kws[0]["dx"] * kws[0]["dy"] * kws[0]["dz"]
    res = {}
    for prec in (0, 1):
        b = _anneal(mmc, kws, prec, 300, stages)
        e = np.array([b.energy(i)["total"] / V for i in range(16)])
        mz = np.array([b.energy(i)["m_avg"][2] for i in range(16)])
        res[prec] = (e, mz, [b.m(i) for i in range(16)])
    for k, name in ((0, "energy density"), (1, "<m_z>")):
        a, f = res[0][k], res[1][k]
        se = np.sqrt(a.var(ddof=1) / 16 + f.var(ddof=1) / 16)
        print(f"{name}: fp64 {a.mean():.6e} +- {a.std(ddof=1) / 4:.2e}, "
              f"fp32 {f.mean():.6e} +- {f.std(ddof=1) / 4:.2e}")
        assert abs(a.mean() - f.mean()) <= 4 * se + 1e-12 * abs(a.mean())
    worst = 0.0
    for i in (0, 7, 15):
        o = orc.Oracle(**kws[i])
        o.build_tensor()
        o.set_m(res[1][2][i])
        e64 = o.energy()["total"] / V
        worst = max(worst, abs(res[1][0][i] - e64) / abs(e64))
    print(f"fp32 final-energy evaluation error {worst:.3e}")
    assert worst <= 1e-6


@pytest.mark.parametrize("geo", [dict(nx=128, ny=96, nz=1, pbc_x=2),
                                 dict(nx=15, ny=12, nz=1, pbc_x=1),
                                 dict(nx=64, ny=20, nz=3)],
                         ids=["fused_pbc", "three_colours", "3d"])
def test_autotune_keeps_the_chain(mmc, geo):
    """The autotuner times the admissible schedules, keeps the fastest, and
    restores the state: the chain afterwards is the untuned one."""
    kw = film(**geo)
    ref = _run(mmc, kw, 1, 6)[-1]
    c = mmc.Context(mmc.make_params(**kw), 1)
    c.build_tensor()
    c.init_random_m()
    c.set_temperature(700.0)
    c.set_option("graph", 1)
    rep = c.autotune(4)
    print(rep)
    assert "chosen:" in rep
    c.sweeps(1, 0, 3)
    # any reported schedule can be applied by name mid-run: same chain
    names = [ln.rsplit(None, 2)[0] for ln in rep.splitlines() if ln.endswith("us/sweep")]
    c.set_schedule(names[-1])
    with pytest.raises(mmc.MMCError):
        c.set_schedule("no such schedule")
    c.sweeps(1, 3, 3)
    assert np.array_equal(c.m(), ref)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_fused_tiles_fold_short_z(mmc, precision):
    """With "zfold", two-layer films take fused tiles that span both layers
    and transform z in registers: the fused passes move the xz spectrum B (Pz
    planes), and the sweep has no separate z passes."""
    b = 8 if precision == 0 else 4
    c = mmc.Context(mmc.make_params(**film(nx=128, ny=40, nz=2)), precision)
    Px, Py, Pz = c.P()
    assert Pz == 3
    H = Px // 2 + 1
    N = 128 * 40 * 2
    assert c.kernel_bytes("xinv_metro") == 3 * H * 2 * 40 * 2 * b + 6 * N * b
    c.set_option("zfold", 1)
    assert c.kernel_bytes("xinv_metro") == 3 * H * Pz * 40 * 2 * b + 6 * N * b
    c.build_tensor()
    c.init_random_m()
    c.set_temperature(500.0)
    c.set_option("timing", 1)
    c.sweeps(0, 0, 3)
    n, ms, by = c.kernel_stats("zfwd", reset=True)
    assert n == 1          # only the prologue's z pass of the xz spectrum
    n, ms, by = c.kernel_stats("zinv", reset=True)
    assert n == 0


@pytest.mark.parametrize("name,geo", CASES, ids=[c[0] for c in CASES])
def test_energy_density_fp64(orc, mmc, name, geo):
    """Per-cell energy density of each term and of all terms matches the
    oracle's within 1e-12 of its largest value, and sums to the energy."""
    kw = film(**geo) if "3d" not in name else dict(FILM3D, **geo)
    o, c = pair(orc, mmc, **kw)
    m = random_m((kw["nx"], kw["ny"], kw["nz"]), 17)
    o.set_m(m)
    c.set_m(m)
    for _, bit in TERMS:
        assert rel_inf(c.energy_density(bit), o.energy_density(bit)) <= 1e-12
    V = kw["dx"] * kw["dy"] * kw["dz"]
    E = c.energy()["total"]
    assert abs(c.energy_density(31).sum() * V - E) <= 1e-10 * abs(E)


@pytest.mark.parametrize("graph", [0, 1], ids=["eager", "graph"])
@pytest.mark.parametrize("fuse", [0, 1], ids=["separate", "fused"])
def test_concurrent_segments_match_serial(mmc, fuse, graph):
    """Per-length launches of a batch on side streams (forked from and joined
    back into the context's stream, captured as parallel graph branches)
    give the bit-identical chains of in-order launches."""
    geos = [film(nx=64, ny=64, seed=1), film(nx=96, ny=48, seed=2),
            film(nx=128, ny=16, pbc_x=1, seed=3), film(nx=64, ny=18, nz=2, seed=4)]
    out = []
    for conc in (0, 1):
        b = mmc.Context([mmc.make_params(**kw) for kw in geos], 1)
        b.build_tensor()
        b.init_random_m()
        b.set_temperature(600.0)
        b.set_option("fuse", fuse)
        b.set_option("concurrent", conc)
        if graph:
            b.set_option("graph", 1)
            b.set_option("graph_sweeps", 3)
        b.sweeps(4, 0, 7)
        out.append([b.m(i) for i in range(len(geos))])
    for i, (a, c) in enumerate(zip(*out)):
        assert np.array_equal(a, c), f"instance {i}"


@pytest.mark.parametrize("geo", [dict(nx=64, ny=64, nz=2), dict(nx=128, ny=64, nz=2, pbc_z=1),
                                 dict(nx=64, ny=48, nz=3), dict(nx=32, ny=64, nz=4, pbc_y=1,
                                                                pbc_z=2)],
                         ids=["open_pz3", "pbcz_pz2", "open_pz5", "pbc_pz4"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_few_layers_sum_z_directly(orc, mmc, geo, precision):
    """Films of 2-4 layers with a fixed-plan y length take no z transforms:
    one y pass per x frequency sums the z convolution over the x-y spectra
    of the tensor's z slices.  Fields and energies match the oracle's
    z-transform convolution (FP64 within 1e-12, FP32 bulk within 2e-6 of
    the largest field), and fused, separate, eager and graph-replayed
    sweeps give one chain."""
    kw = film(**geo)
    o, c = pair(orc, mmc, precision=precision, **kw)
    assert c.kernel_bytes("yzconv") > 0
    assert c.kernel_bytes("zfwd") == 0.0 and c.kernel_bytes("ymul") == 0.0
    m = random_m((kw["nx"], kw["ny"], kw["nz"]), 23)
    o.set_m(m)
    c.set_m(m)
    tol = 1e-12 if precision == 0 else 2e-6
    for tname, bits in (("demag", 16), ("all", 31)):
        err = rel_inf(c.field(bits), o.field(bits))
        assert err <= tol, f"{tname}: rel err {err:.3e}"
    Eo, Eg = o.energy(), c.energy()
    scale = sum(abs(Eo[k]) for k in ("exch", "dmi", "anis", "zeeman", "demag"))
    assert abs(Eg["demag"] - Eo["demag"]) <= (1e-12 if precision == 0 else 1e-6) * scale
    out = []
    for fuse, graph in ((0, 0), (1, 0), (1, 1)):
        c = mmc.Context(mmc.make_params(**kw), precision)
        c.set_tensor(o.tensor())
        c.init_random_m()
        c.set_temperature(650.0)
        try:
            c.set_option("fuse", fuse)
        except mmc.MMCError:
            assert fuse == 1        # grid without fused-kernel x lengths
        if graph:
            c.set_option("graph", 1)
            c.set_option("graph_sweeps", 3)
        h0 = c.field(31)
        c.sweeps(5, 0, 7)
        out.append((h0, c.m(), c.energy()["total"]))
    for h0, m, e in out[1:]:
        assert np.array_equal(h0, out[0][0])
        assert np.array_equal(m, out[0][1])
        assert e == out[0][2]


@pytest.mark.parametrize("geos", [[dict(nx=64, ny=40)], [dict(nx=128, ny=10, pbc_x=1, pbc_y=1)],
                                  [dict(nx=64, ny=37)], [dict(nx=128, ny=2, pbc_y=1)],
                                  [dict(nx=256, ny=1)],
                                  [dict(nx=64, ny=64, seed=1), dict(nx=96, ny=31, seed=2),
                                   dict(nx=128, ny=16, pbc_x=1, seed=3)],
                                  [dict(nx=64, ny=32, nz=2)],
                                  [dict(nx=128, ny=14, nz=2, pbc_y=1, pbc_z=1)],
                                  [dict(nx=64, ny=21, nz=2, seed=4), dict(nx=64, ny=30, seed=5)]],
                         ids=["open", "pbc", "odd_rows", "two_rows_pbc", "one_row", "batch",
                              "two_layers", "two_layers_pbc", "mixed_layers"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_row_sweeps_match_fused(mmc, geos, precision):
    """Whole-sweep row-pair tiles (halo pairs inverted and their colour-0
    updates recomputed in each tile, every layer in the tile, state in two
    alternating buffers) give the bit-identical chain and energies of the
    fused and separate schedules: eager with an odd sweep count,
    graph-replayed with an odd graph length, and after a further odd run."""
    kws = [film(**g) for g in geos]
    out = []
    for fuse, rows, graph in ((0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 1, 1)):
        c = mmc.Context([mmc.make_params(**kw) for kw in kws], precision)
        c.build_tensor()
        c.init_random_m()
        c.set_temperature(650.0)
        c.set_option("fuse", fuse)
        c.set_option("rows", rows)
        if rows:
            assert c.kernel_bytes("row_sweep") > 0
        if graph:
            c.set_option("graph", 1)
            c.set_option("graph_sweeps", 3)
        c.sweeps(3, 0, 7)
        m1 = [c.m(i) for i in range(len(kws))]
        c.sweeps(3, 7, 5)
        out.append((m1, [c.m(i) for i in range(len(kws))], c.energy()["total"]))
    for m1, m2, e in out[1:]:
        for i in range(len(kws)):
            assert np.array_equal(m1[i], out[0][0][i]), f"instance {i}"
            assert np.array_equal(m2[i], out[0][1][i]), f"instance {i}"
        assert e == out[0][2]


def test_row_sweeps_need_thin_films(mmc):
    """The row schedule is refused where an instance has more than two
    layers."""
    c = mmc.Context(mmc.make_params(**film(nx=64, ny=32, nz=3)), 1)
    with pytest.raises(mmc.MMCError):
        c.set_option("rows", 1)


def test_row_sweeps_in_anneal(orc, mmc):
    """The anneal driver under row sweeps (energies every 7 sweeps,
    snapshots every 11, graph replay of 5) records the energies, snapshots
    and final state of the fused schedule; and FP64 row sweeps follow the
    oracle's chain stage by stage."""
    kw = film(nx=64, ny=48, nz=1, seed=3)
    out = []
    for rows in (0, 1):
        c = mmc.Context(mmc.make_params(**kw), 1)
        c.build_tensor()
        c.init_random_m()
        c.set_option("rows", rows)
        c.set_option("graph", 1)
        c.set_option("graph_sweeps", 5)
        E, S = [], []
        c.anneal(1400.0, 700.0, 0.8, True, 23, energy_period=7, snapshot_period=11,
                 on_energy=lambda T, st, sw, e: E.append((st, sw, e)),
                 on_snapshot=lambda T, st, sw, m: S.append(m))
        assert len(E) > 3 and len(S) > 3
        out.append((E, S, c.m()))
    assert out[0][0] == out[1][0]
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][2], out[1][2])
    o, c = pair(orc, mmc, **kw)
    m = random_m((64, 48, 1), 9)
    o.set_m(m)
    c.set_m(m)
    c.set_option("rows", 1)
    for st, T in enumerate((1100.0, 500.0, 90.0)):
        c.set_temperature(T)
        c.sweeps(st, 0, 9)
        for k in range(9):
            o.sweep(T, st, k)
    assert rel_inf(c.m(), o.m()) <= 1e-10


def test_grid_index_limits(mmc):
    """Grids whose cell indices would not fit the kernels' 32-bit index
    arithmetic (nx or ny*nz at 2^24 or more) are refused when the context is
    created, before any device memory is allocated."""
    for geo in (dict(nx=1 << 24, ny=1, nz=1), dict(nx=1, ny=1 << 12, nz=1 << 12)):
        with pytest.raises(mmc.MMCError, match="2\\^24"):
            mmc.Context(mmc.make_params(**film(**geo)), 1)


@pytest.mark.parametrize("graph", [0, 1], ids=["eager", "graph"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_one_launch_per_pass_matches_per_length(mmc, precision, graph):
    """A batch of x lengths 128..512 and y lengths 64..512 (one periodic,
    one 3-D instance with z passes) swept with the y pass and both fused
    passes as one launch each over all instances (each item running its own
    length's transform), with one or two rows per short item, gives the
    bit-identical chains, fields and energies of per-length launches."""
    geos = [film(nx=64, ny=64, seed=1), film(nx=96, ny=48, seed=2),
            film(nx=128, ny=32, pbc_x=1, seed=3), film(nx=192, ny=32, seed=4),
            film(nx=256, ny=48, seed=5), film(nx=64, ny=32, nz=5, seed=6),
            film(nx=256, ny=256, seed=7)]
    out = []
    for sched in ("fused, minimal tiles", "one launch per pass, minimal tiles",
                  "one launch per pass, paired short rows"):
        b = mmc.Context([mmc.make_params(**kw) for kw in geos], precision)
        b.build_tensor()
        b.init_random_m()
        b.set_temperature(600.0)
        b.set_schedule(sched)
        if graph:
            b.set_option("graph", 1)
            b.set_option("graph_sweeps", 3)
        b.set_option("timing", 1)
        b.sweeps(4, 0, 7)
        n_y, _, _ = b.kernel_stats("ymul", reset=True)
        n_x, _, _ = b.kernel_stats("xinv_metro", reset=True)
        if sched.startswith("one launch") and not graph:
            assert (n_y, n_x) == (7, 7)
        out.append(([b.m(i) for i in range(len(geos))], b.field(31, 5),
                    [b.energy(i)["total"] for i in range(len(geos))]))
    for ms, h, e in out[1:]:
        for i, (a, c) in enumerate(zip(ms, out[0][0])):
            assert np.array_equal(a, c), f"instance {i}"
        assert np.array_equal(h, out[0][1])
        assert e == out[0][2]
