"""Benchmark: Metropolis spin-update throughput of the annealing hot path.

Headline workload (one step = one sweep = one Metropolis attempt per cell of
every instance, preceded by the demag convolution of the pre-sweep state):
64 independent anneals batched in one context, nz=1 films n x n with n drawn
from {64, 96, 128, 192, 256} by a seeded generator, Bz ~ U(0, 0.6) T,
Dind ~ U(3, 6) mJ/m^2, Ms/Aex/Ku1 of the 64x64 film, open boundaries,
distinct Philox seeds, FP32-bulk (FP32 FFT/stencil/dE, FP64 tensor and
reductions), the sweep schedule chosen by the library's autotuner, sweeps
replayed from hipGraphs.  Beside it: the 64x64 film in plain FP64 (the
single-instance denominator of the batched/single ratio), the accuracy of
FP32 bulk over the 64x64 film's full anneal (16 seeds), the 256x256 film
with 20 periodic images (FP32 and FP64), the 128x128x16 film and the
1024x1024x2 film.

Per-kernel durations behind `roofline` come from eager sweeps right after
the timed ones, each launch bracketed by HIP events that its dispatch packet
records on the library's stream.

Prints one JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "micromagnetic-gpu-monte-carlo_amd")

METRIC = ("spin-update attempts/s (plain FP64 vs FP32 batched) + HBM GB/s % peak, "
          "equal accuracy")
HBM_PEAK_GBS = 8000.0


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "mmc_amd", os.path.join(PKG, "__init__.py"), submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mmc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def film256(**kw):
    d = dict(nx=256, ny=256, nz=1, dx=1e-9, dy=1e-9, dz=0.8e-9, pbc_x=20, pbc_y=20,
             pbc_z=0, ms=1.6e6, aex=16e-12, dind=6e-3, anis_ux=0.0, anis_uy=0.0,
             anis_uz=1.0, ku1=2e6, bext_x=0.0, bext_y=0.0, bext_z=0.3, cal_demag=1,
             seed=256)
    d.update(kw)
    return d


def film1024x2(**kw):
    d = dict(nx=1024, ny=1024, nz=2, dx=2e-9, dy=2e-9, dz=0.8e-9, pbc_x=10, pbc_y=10,
             pbc_z=0, ms=1.1e6, aex=15e-12, dind=4e-3, anis_ux=0.0, anis_uy=0.0,
             anis_uz=1.0, ku1=1e6, bext_x=0.0, bext_y=0.0, bext_z=0.25, cal_demag=1,
             seed=1024)
    d.update(kw)
    return d


def film128x16(**kw):
    d = dict(nx=128, ny=128, nz=16, dx=1e-9, dy=1e-9, dz=1e-9, pbc_x=0, pbc_y=0,
             pbc_z=0, ms=1.6e6, aex=13e-12, dind=3e-3, anis_ux=0.0, anis_uy=0.0,
             anis_uz=1.0, ku1=8e5, bext_x=0.0, bext_y=0.0, bext_z=0.2, cal_demag=1,
             seed=128)
    d.update(kw)
    return d


def fp32_errors(mod, kw, ctx32):
    """FP32-bulk error against the FP64 path (itself equal to the FP64 oracle
    within 1e-12) on the same states: the seeded start state and the FP32
    context's current state.  Norm-wise relative errors of H_eff (all terms),
    the total energy and every cell's trial-move dE."""
    import numpy as np
    c64 = mod.Context(mod.make_params(**kw), mod.FP64)
    c64.build_tensor()
    c32 = mod.Context(mod.make_params(**kw), mod.FP32)
    c32.build_tensor()
    out = {}
    for name, src in (("start", None), ("after_timed_sweeps", ctx32)):
        if src is None:
            c32.init_random_m()
        else:
            c32.set_m(src.m())
        m = c32.m()
        c64.set_m(m)
        c32.set_m(m)
        h32, h64 = c32.field(), c64.field()
        e32, e64 = c32.energy()["total"], c64.energy()["total"]
        d32, d64 = c32.trial_delta_e(7, 3), c64.trial_delta_e(7, 3)
        out[name] = dict(h_eff_rel=float(np.abs(h32 - h64).max() / np.abs(h64).max()),
                         energy_rel=float(abs(e32 - e64) / abs(e64)),
                         delta_e_rel=float(np.abs(d32 - d64).max() / np.abs(d64).max()))
    c64.close()
    c32.close()
    return out


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_ctx(mod, kws, precision, T=1400.0):
    c = mod.Context([mod.make_params(**kw) for kw in kws], precision)
    t0 = time.perf_counter()
    c.build_tensor()
    c.sync()
    t_tensor = time.perf_counter() - t0
    c.init_random_m()
    c.set_temperature(T)
    return c, t_tensor


def run_gpu(mod, kws, precision, steps, warmup, T=1400.0, graph=True, ksteps=None,
            default_stream=False, autotune=20):
    """Pick the sweep schedule with the library's autotuner (`autotune`
    sweeps per candidate; 0 keeps the default), warm up, time `steps` sweeps
    (hipGraph replay when `graph`), then time `ksteps` more sweeps with every
    launch bracketed by dispatch-packet HIP events for the per-kernel
    durations."""
    if isinstance(kws, dict):
        kws = [kws]
    c, t_tensor = make_ctx(mod, kws, precision, T)
    N = sum(kw["nx"] * kw["ny"] * kw["nz"] for kw in kws)
    if default_stream:
        c.set_option("default_stream", 1)
        graph = False
    if graph:
        c.set_option("graph", 1)
        c.set_option("graph_sweeps", 20)
    schedule = None
    if isinstance(autotune, str):
        c.set_schedule(autotune)
        schedule = autotune
    elif autotune:
        rep = c.autotune(autotune)
        log(rep.rstrip())
        schedule = rep.strip().splitlines()[-1].replace("chosen: ", "")
    c.sweeps(0, 0, warmup)
    c.sync()
    t0 = time.perf_counter()
    c.sweeps(0, warmup, steps)
    c.sync()
    dt = time.perf_counter() - t0
    if ksteps is None:
        ksteps = min(steps, 400)
    kern = {}
    dt_eager = None
    if ksteps <= 0:
        E = c.energy()
        return dict(attempts_per_s=N * steps / dt, ms_per_sweep=1e3 * dt / steps,
                    eager_timed_attempts_per_s=None, tensor_build_s=t_tensor, kernels=kern,
                    energy_J=E["total"], ctx=c, cells=N, schedule=schedule)
    c.set_option("timing", 1)
    for k in mod.KINDS:
        c.kernel_stats(k, reset=True)
    t1 = time.perf_counter()
    c.sweeps(0, warmup + steps, ksteps)
    c.sync()
    dt_eager = time.perf_counter() - t1
    for k in mod.KINDS:
        n, ms, by = c.kernel_stats(k, reset=True)
        if n:
            kern[k] = dict(launches=n, total_ms=ms, avg_us=1e3 * ms / n,
                           bytes_per_launch=by / n, total_bytes=by)
    c.set_option("timing", 0)
    E = c.energy()
    return dict(attempts_per_s=N * steps / dt, ms_per_sweep=1e3 * dt / steps,
                eager_timed_attempts_per_s=N * ksteps / dt_eager,
                tensor_build_s=t_tensor, kernels=kern, energy_J=E["total"], ctx=c, cells=N,
                schedule=schedule)


def batch_params(count=64, master_seed=2024):
    """Independent anneals of uneven size: nz=1 grids n x n with n drawn from
    {64,96,128,192,256}, Bz ~ U(0,0.6) T, Dind ~ U(3,6) mJ/m^2, the 64x64
    film's Ms/Aex/Ku1, distinct seeds."""
    import numpy as np
    rng = np.random.default_rng(master_seed)
    out = []
    for i in range(count):
        n = int(rng.choice([64, 96, 128, 192, 256]))
        out.append(dict(nx=n, ny=n, nz=1, dx=1e-9, dy=1e-9, dz=0.8e-9, pbc_x=0, pbc_y=0,
                        pbc_z=0, ms=1.6e6, aex=16e-12, dind=float(rng.uniform(3e-3, 6e-3)),
                        anis_ux=0.0, anis_uy=0.0, anis_uz=1.0, ku1=2e6, bext_x=0.0,
                        bext_y=0.0, bext_z=float(rng.uniform(0.0, 0.6)), cal_demag=1,
                        seed=1000 + i))
    return out


def _batch64():
    kws = batch_params()
    sizes = sorted({kw["nx"] for kw in kws})
    counts = {n: sum(kw["nx"] == n for kw in kws) for n in sizes}
    return kws, ("anneal sweep, 64 independent films batched in one context, n x n x 1 with "
                 "n drawn from {64,96,128,192,256} (master seed 2024: "
                 + ", ".join(f"{counts[n]}x{n}^2" for n in sizes)
                 + "), open boundaries, Bz ~ U(0,0.6) T, Dind ~ U(3,6) mJ/m^2, FP32-bulk, "
                   "T=1400 K")


WORKLOADS = {
    "batch64": _batch64,
    "film256": lambda: ([film256()], "anneal sweep, 256x256x1 film, PBC 20x20 images, "
                                     "FP32-bulk, T=1400 K"),
    "film128x16": lambda: ([film128x16()], "anneal sweep, 128x128x16 open film, FP32-bulk, "
                                           "T=1400 K"),
    "film1024x2": lambda: ([film1024x2()], "anneal sweep, 1024x1024x2 film, PBC 10x10 images, "
                                           "FP32-bulk, T=1400 K"),
}


def copy_calibration(nbytes, times):
    """Device-to-device copies of a known size through the HIP runtime the
    library links (hipMemcpy runs the runtime's copy kernel): the reference
    point for scaling hardware byte counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(times):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0   # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def measured_traffic(workload, kind):
    """HBM bytes per launch of `kind` on `workload` from the committed
    hardware-counter summary (FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for a device copy of known size), or
    None when no counter run covers it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["workloads"][workload][kind]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def roofline(kern, workload=None):
    """Dominant kernel (largest device time) and its achieved bandwidth:
    algorithmic bytes of its timed launches over their summed duration."""
    if not kern:
        return None, None
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    k = kern[dom]
    gbs = k["total_bytes"] / (k["total_ms"] * 1e-3) / 1e9
    return dom, dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                     frac=round(gbs / HBM_PEAK_GBS, 4),
                     traffic=measured_traffic(workload, dom) if workload else None,
                     kernel=dom, avg_us=round(k["avg_us"], 3),
                     bytes_per_launch=k["bytes_per_launch"])


def cpu_baseline(mod, ctx, kws, sweeps):
    """FP64 oracle sweeps of the same instances (each instance's tensor taken
    from the GPU build; one instance at a time), on all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    attempts, dt = 0, 0.0
    for i, kw in enumerate(kws):
        o = oracle.Oracle(**kw)
        o.set_tensor(ctx.tensor(i))
        o.init_random_m(kw["seed"])
        o.sweep(1400.0, 0, 0)       # builds the tensor spectrum outside the timed part
        t0 = time.perf_counter()
        for s in range(sweeps):
            o.sweep(1400.0, 0, 1 + s)
        dt += time.perf_counter() - t0
        attempts += kw["nx"] * kw["ny"] * kw["nz"] * sweeps
        del o
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    what = (f"{len(kws)} instances" if len(kws) > 1
            else f"the {kws[0]['nx']}x{kws[0]['ny']}x{kws[0]['nz']} film")
    return dict(value=attempts / dt, unit="spin-update attempts/s", cores=cores, kind="port",
                sample=f"{sweeps} FP64 sweeps of each of {what} ({attempts} attempts, "
                       f"{dt:.1f} s) on the C oracle (OpenMP), GPU-built tensors")


def tensor_cpu_baseline(mod):
    """The demag tensor with periodic-image sums built by the FP64 C oracle
    (OpenMP, all host cores) and by the GPU, same geometry: the 256x256 film
    with 2 periodic images per axis (5x5 image sums; the film's own 20 images
    take the oracle minutes), the one-off cost the tensor cache saves."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    kw = film256(pbc_x=2, pbc_y=2)
    o = oracle.Oracle(**kw)
    t0 = time.perf_counter()
    o.build_tensor()
    t_cpu = time.perf_counter() - t0
    c = mod.Context(mod.make_params(**kw), mod.FP64)
    c.build_tensor()          # first build: loads and warms the kernels
    c.sync()
    t0 = time.perf_counter()
    c.build_tensor()
    c.sync()
    t_gpu = time.perf_counter() - t0
    # norm-wise: single far entries carry the Newell cancellation noise of two
    # FP64 libm implementations, which grows with the image distance
    err = float(np.linalg.norm(c.tensor() - o.tensor()) / np.linalg.norm(o.tensor()))
    c.close()
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(sample="FP64 demag tensor of the 256x256x1 film with 2 periodic images per axis",
                cpu_s=t_cpu, cpu_cores=cores, kind="port", gpu_s=t_gpu, gpu_over_cpu=t_cpu / t_gpu,
                normwise_rel_diff=err)


def anneal_stages(T0=1400.0, T1=40.0, f=0.9):
    out, T = [], T0
    while T > T1:
        out.append(T)
        T *= f
    return out


def equal_accuracy(mod, seeds=16, sweeps_per_stage=1000):
    """The 64x64 film annealed on its full schedule (1400 K x0.9 to 40 K, 34
    stages) for `seeds` seeds, batched, in FP64 and in FP32 bulk: seed means
    and standard errors of the final energy density and <m_z>, and the FP32
    path's final-energy error (its energy of its final states against the FP64
    path's energy of the same states; the FP64 path equals the FP64 oracle
    within 1e-12 in the parity tests)."""
    import numpy as np
    kws = [film256(nx=64, ny=64, pbc_x=0, pbc_y=0, seed=s) for s in range(1, seeds + 1)]
    V = 64 * 64 * 1e-9 * 1e-9 * 0.8e-9
    out, finals = {}, {}
    t0 = time.perf_counter()
    for prec, name in ((mod.FP64, "fp64"), (mod.FP32, "fp32")):
        b = mod.Context([mod.make_params(**kw) for kw in kws], prec)
        b.build_tensor()
        b.init_random_m()
        b.set_option("graph", 1)
        b.set_option("graph_sweeps", 20)
        for st, T in enumerate(anneal_stages()):
            b.set_temperature(T)
            b.sweeps(st, 0, sweeps_per_stage)
        e = np.array([b.energy(i)["total"] / V for i in range(seeds)])
        mz = np.array([b.energy(i)["m_avg"][2] for i in range(seeds)])
        finals[name] = (e, [b.m(i) for i in range(seeds)])
        out[name] = dict(energy_density_mean=float(e.mean()),
                         energy_density_se=float(e.std(ddof=1) / np.sqrt(seeds)),
                         mz_mean=float(mz.mean()), mz_se=float(mz.std(ddof=1) / np.sqrt(seeds)))
        b.close()
    c64 = mod.Context([mod.make_params(**kw) for kw in kws], mod.FP64)
    c64.build_tensor()
    worst = 0.0
    for i in range(seeds):
        c64.set_m(finals["fp32"][1][i], i)
    for i in range(seeds):
        e64 = c64.energy(i)["total"] / V
        worst = max(worst, abs(finals["fp32"][0][i] - e64) / abs(e64))
    c64.close()
    d = out["fp64"]["energy_density_mean"] - out["fp32"]["energy_density_mean"]
    se = (out["fp64"]["energy_density_se"] ** 2 + out["fp32"]["energy_density_se"] ** 2) ** 0.5
    out.update(workload=f"64x64 film, {seeds} seeds, 34 stages x {sweeps_per_stage} sweeps",
               fp32_final_energy_rel_err=worst,
               mean_energy_density_diff_in_se=abs(d) / se if se > 0 else 0.0,
               seconds=time.perf_counter() - t0)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-sweeps", type=int, default=10,
                    help="oracle sweeps per instance for the CPU baseline (0: skip it)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the side measurements (other films, accuracy anneal)")
    ap.add_argument("--workload", choices=tuple(WORKLOADS), default="batch64",
                    help="workload measured as `value` (others only for profiling runs)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the timed sweeps eagerly instead of replaying hipGraphs")
    ap.add_argument("--default-stream", action="store_true",
                    help="launch on the device's default stream (hardware-counter runs; "
                         "implies --no-graph)")
    ap.add_argument("--autotune", type=int, default=20,
                    help="sweeps per candidate schedule for the library's autotuner "
                         "(0: keep the default schedule)")
    ap.add_argument("--schedule", default=None,
                    help="apply this named sweep schedule to the measured workload instead of "
                         "autotuning (profiling runs)")
    ap.add_argument("--ksteps", type=int, default=None,
                    help="event-timed eager sweeps after the timed region (default "
                         "min(steps, 400); 0 skips them, as hardware-counter runs need)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device to device three times (known-size kernel "
                         "for scaling hardware byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    mod = load_pkg()
    mod.lib()
    if a.copy_calibration:
        copy_calibration(268435456, 3)
        log("[bench] calibration: 3 device copies of 268435456 bytes")

    kws, desc = WORKLOADS[a.workload]()
    log(f"[bench] {a.workload} FP32-bulk: {a.warmup} warmup + {a.steps} timed sweeps")
    r32 = run_gpu(mod, kws, mod.FP32, a.steps, a.warmup, graph=not a.no_graph, ksteps=a.ksteps,
                  default_stream=a.default_stream,
                  autotune=a.schedule if a.schedule else a.autotune)
    dom, roof = roofline(r32["kernels"], a.workload)
    extra = {}
    if not a.no_extra:
        log("[bench] 64x64x1 open film, plain FP64 single instance")
        kw0 = film256(nx=64, ny=64, pbc_x=0, pbc_y=0)
        r0 = run_gpu(mod, kw0, mod.FP64, max(400, a.steps), a.warmup, autotune=a.autotune)
        extra["film64_fp64_attempts_per_s"] = r0["attempts_per_s"]
        extra["film64_fp64_ms_per_sweep"] = r0["ms_per_sweep"]
        extra["film64_fp64_schedule"] = r0["schedule"]
        if a.workload == "batch64":
            extra["batched_fp32_over_single_fp64"] = r32["attempts_per_s"] / r0["attempts_per_s"]
        r0["ctx"].close()
        log("[bench] equal accuracy: 16 seeds of the 64x64 film's anneal, FP64 and FP32 bulk")
        extra["equal_accuracy"] = equal_accuracy(mod)
        kw1 = film256()
        log("[bench] 256x256 PBC20 film, FP32 bulk and plain FP64")
        r1 = run_gpu(mod, kw1, mod.FP32, max(200, a.steps // 4), max(20, a.warmup // 4),
                     autotune=a.autotune)
        extra["film256_fp32_attempts_per_s"] = r1["attempts_per_s"]
        extra["film256_fp32_ms_per_sweep"] = r1["ms_per_sweep"]
        extra["film256_fp32_schedule"] = r1["schedule"]
        extra["film256_fp32_kernels_us"] = {k: round(v["avg_us"], 3) for k, v in r1["kernels"].items()}
        extra["film256_fp32_error_vs_fp64"] = fp32_errors(mod, kw1, r1["ctx"])
        r1["ctx"].close()
        r64 = run_gpu(mod, kw1, mod.FP64, max(200, a.steps // 4), max(20, a.warmup // 4),
                      autotune=a.autotune)
        extra["film256_fp64_attempts_per_s"] = r64["attempts_per_s"]
        extra["film256_fp64_ms_per_sweep"] = r64["ms_per_sweep"]
        r64["ctx"].close()
        log("[bench] 128x128x16 open 3-D film FP32-bulk")
        r2 = run_gpu(mod, film128x16(), mod.FP32, 200, 20, autotune=a.autotune)
        d2, roof2 = roofline(r2["kernels"], "film128x16")
        extra["film128x16_attempts_per_s"] = r2["attempts_per_s"]
        extra["film128x16_ms_per_sweep"] = r2["ms_per_sweep"]
        extra["film128x16_schedule"] = r2["schedule"]
        extra["film128x16_roofline"] = roof2
        extra["film128x16_kernels_us"] = {k: round(v["avg_us"], 3) for k, v in r2["kernels"].items()}
        r2["ctx"].close()
        log("[bench] 1024x1024x2 PBC10 FP32-bulk")
        r4 = run_gpu(mod, film1024x2(), mod.FP32, 200, 20, autotune=a.autotune)
        d4, roof4 = roofline(r4["kernels"], "film1024x2")
        extra["film1024x2_attempts_per_s"] = r4["attempts_per_s"]
        extra["film1024x2_ms_per_sweep"] = r4["ms_per_sweep"]
        extra["film1024x2_tensor_build_s"] = r4["tensor_build_s"]
        extra["film1024x2_roofline"] = roof4
        extra["film1024x2_schedule"] = r4["schedule"]
        extra["film1024x2_kernels_us"] = {k: round(v["avg_us"], 3) for k, v in r4["kernels"].items()}
        r4["ctx"].close()
    cpu = None
    if a.cpu_sweeps > 0:
        log("[bench] CPU baseline (oracle)")
        cpu = cpu_baseline(mod, r32["ctx"], kws, a.cpu_sweeps)
        if not a.no_extra:
            log("[bench] tensor build: oracle vs GPU")
            extra["tensor_build_cpu_baseline"] = tensor_cpu_baseline(mod)
    line = {
        "metric": METRIC,
        "value": r32["attempts_per_s"],
        "unit": "spin-update attempts/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r32["ms_per_sweep"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "cells": r32["cells"], "instances": len(kws),
                   "step": "one sweep of every instance", "precision": "fp32-bulk"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels_us": {k: round(v["avg_us"], 3) for k, v in r32["kernels"].items()},
        "schedule": r32["schedule"],
        "eager_timed_attempts_per_s": r32["eager_timed_attempts_per_s"],
        "tensor_build_s": r32["tensor_build_s"],
    }
    line.update(extra)
    r32["ctx"].close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
