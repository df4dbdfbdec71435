"""Shared builders for the parity tests (configs from the anneal workloads)."""
import numpy as np

# material of the reference README example / 64x64 and 256x256 anneals
FILM = dict(dx=1e-9, dy=1e-9, dz=0.8e-9, ms=1.6e6, aex=16e-12, dind=6e-3,
            anis_ux=0.0, anis_uy=0.0, anis_uz=1.0, ku1=2e6, bext_x=0.0,
            bext_y=0.0, bext_z=0.3, cal_demag=1, seed=256)
# 3-D film material (128x128x16 workload)
FILM3D = dict(dx=1e-9, dy=1e-9, dz=1e-9, ms=1.6e6, aex=13e-12, dind=3e-3,
              anis_ux=0.0, anis_uy=0.0, anis_uz=1.0, ku1=8e5, bext_x=0.0,
              bext_y=0.0, bext_z=0.2, cal_demag=1, seed=7)


def film(**kw):
    d = dict(FILM)
    d.update(kw)
    return d


def random_m(n, seed):
    """Unit vectors, shape (3, nz, ny, nx), from a seeded normal draw."""
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((3, nz, ny, nx))
    return m / np.linalg.norm(m, axis=0)


def rel_inf(a, b):
    """max|a-b| / max|b| (norm-wise relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def pair(orc, mmc, precision=0, shared_tensor=True, **kw):
    """Oracle and library instance on the same parameters (tensor shared)."""
    o = orc.Oracle(**kw)
    p = mmc.make_params(**kw)
    c = mmc.Context(p, precision)
    if kw.get("cal_demag", 1):
        o.build_tensor()
        if shared_tensor:
            c.set_tensor(o.tensor())
        else:
            c.build_tensor()
    return o, c


def newell_noise(kw, P):
    """Rounding-noise scale of each entry of the FP64 Newell tensor, shape
    (Pz, Py, Px): every entry sums 64 corner terms per image, each of
    magnitude up to ~|r|^3 (|r| the image offset plus one cell diagonal, in
    m^3), and the sum cancels to ~V/|r|^3, so two libraries whose f/g differ
    by an ulp or two give entries that differ by about
    eps * sqrt(64 * sum_images |r|^6) / (4 pi V).  Offsets follow the
    tensor's slot map (open axes: o < n at o, o > -n at P + o; periodic
    axes: the reference's centred map, images -p..p)."""
    n = (kw["nx"], kw["ny"], kw["nz"])
    d = (kw["dx"], kw["dy"], kw["dz"])
    pbc = (kw.get("pbc_x", 0), kw.get("pbc_y", 0), kw.get("pbc_z", 0))
    axes = []
    for a in range(3):
        i = np.arange(P[a])
        if pbc[a]:
            off = (i + (P[a] + 1) // 2 - 1) % P[a] - (P[a] + 1) // 2 + 1
            imgs = np.arange(-max(pbc[a], 0), max(pbc[a], 0) + 1)
            r = (off[:, None] + imgs[None, :] * P[a]) * d[a]       # (P, images)
        else:
            off = np.where(i < n[a], i, i - P[a])
            r = (off * d[a])[:, None]
        axes.append(r)
    rx, ry, rz = axes
    diag = float(np.sqrt(sum(v * v for v in d)))
    acc = np.zeros((P[2], P[1], P[0]))
    for kz in range(rz.shape[1]):
        for ky in range(ry.shape[1]):
            r2 = (rz[:, kz][:, None, None, None] ** 2 + ry[:, ky][None, :, None, None] ** 2
                  + rx[None, None, :, :] ** 2)                     # (Pz, Py, Px, ix)
            acc += ((np.sqrt(r2) + diag) ** 6).sum(axis=-1)
    V = d[0] * d[1] * d[2]
    return np.finfo(float).eps * np.sqrt(64.0 * acc) / (4 * np.pi * V)
