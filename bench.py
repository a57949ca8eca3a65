#!/usr/bin/env python3
"""DS2D hot-path benchmark on one MI355X.

Default workload (c3, the N=1 line): a "step" is one pass of the hot path
over one batch = batched elastic forward modelling of 128 uneven-aperture
shots (windows U{400..1200} columns) on the seeded 2000x600 layered+faulted
model, dx=dz=12.5 m, 20-cell PML, dt = 1 ms, 3000 time steps, vx/vz recorded
at every window surface node.  All shots run in one launch per time step
(or per two time steps, when the autotuner picks temporal blocking), and the
whole time loop is captured in one hipGraph (--no-graph: direct launches).
The model and every wavefield are resident in HBM before timing; records
stay in HBM (fetched once, untimed).

Other workloads (--workload): c1 = single-shot elastic forward on 251x101
(2500 steps); c2 = acoustic CG LSRTM of 26 random-boundary shots on 251x101
(time to a relative data residual, residual history); c4 = the learned
preconditioner: training pairs from the plain method, U-Net training, and
plain vs learned LSRTM time to the residual target on held-out models;
c5 = PP/PS elastic Born CG LSRTM of 256 uneven windows on 4000x1000 (a step
= one CG iteration); unet = U-Net training throughput; compat = the
reference-compat ELSRTM gradient main.cpp runs (26 shots on the C1 model);
rot = the C3 windows with the P/S-decoupled theta/omega scheme; decomp = the
elastic RTM with a decomposed (Poynting / vector / scalar) imaging condition,
26 shots on the C1 model.

Prints one JSON line: value = Gcell-updates/s (one cell-update = one padded
cell advanced one step for one wavefield); roofline of the dominant kernel
(HIP-event timing inside the library on its own stream, algorithmic bytes of
the layout actually run); cpu_baseline = the FP32 C++/OpenMP oracle on a
bounded sample of the same workload, timed on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "k_fused_step_c3_traffic.json")  # rocprof FETCH/WRITE per launch


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3",
                    choices=["c1", "c2", "c3", "c4", "c5", "unet", "compat", "rot", "decomp"])
    ap.add_argument("--kind", type=int, default=0, help="decomp: image kind (0 Poynting PP, 1-3 vector, 4-6 scalar)")
    ap.add_argument("--c5-iter", type=int, default=2, help="c5: CG iterations on the full configuration")
    ap.add_argument("--train-elastic", type=int, default=0,
                    help="c5: train the U-Net on elastic PP/PS pairs of this many seeded 512x256 models first")
    ap.add_argument("--train-models", type=int, default=64, help="c4: training models (seeds 1000..)")
    ap.add_argument("--eval-models", type=int, default=4, help="c4: held-out models (seeds eval-start..)")
    ap.add_argument("--eval-start", type=int, default=2000, help="c4: seed of the first held-out model")
    ap.add_argument("--train-steps", type=int, default=1500, help="c4: Adam steps")
    ap.add_argument("--max-iter", type=int, default=400, help="c4: iteration cap of each solver")
    ap.add_argument("--ref-full", type=int, default=2,
                    help="c4: held-out models on which the reference scheme also runs in full to the target")
    ap.add_argument("--target", type=float, default=1e-2, help="c4: relative data residual target")
    ap.add_argument("--save-unet", default=None, help="c4: write the trained U-Net parameters (.npy)")
    ap.add_argument("--load-unet", default=None, help="c4: skip pair generation and training, load parameters")
    ap.add_argument("--nshots", type=int, default=None)
    ap.add_argument("--nt", type=int, default=None)
    ap.add_argument("--niter", type=int, default=10, help="c2: CG iterations")
    ap.add_argument("--mode", type=int, default=2, help="2 fused kernel (default), 1 two fast kernels, 0 reference layout")
    ap.add_argument("--graph", action="store_true",
                    help="capture each time loop in one hipGraph (always on for c3, whose configuration names it)")
    ap.add_argument("--no-graph", action="store_true", help="c3: launch the time loop's kernels directly")
    ap.add_argument("--no-fused-born", action="store_true",
                    help="c5: the elastic Born forward pass with two kernels per step instead of one fused launch")
    ap.add_argument("--no-fused-rot", action="store_true",
                    help="rot: k_rot_v + k_rot_p per step instead of the fused rotation step")
    ap.add_argument("--born-tile", default=None, help="c5: x-chunk width,lanes of the fused Born step (e.g. 64,256)")
    ap.add_argument("--no-chunk-plan", action="store_true",
                    help="two steps per launch: fixed-width chunks instead of the balanced chunk plan")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=5.0, help="CPU baseline: seconds per timed run (3 runs)")
    ap.add_argument("--traffic-json", default=TRAFFIC_JSON,
                    help="rocprof FETCH/WRITE summary (JSON) to embed as roofline.traffic when its kernel and "
                         "workload match the run")
    ap.add_argument("--random-boundary", action="store_true",
                    help="c3: the same windows with the random boundary (no PML strips)")
    ap.add_argument("--time-block", type=int, default=0, choices=[0, 1, 2],
                    help="mode 2: time steps per launch (0 = what the autotuner picked, or 1)")
    ap.add_argument("--tile", default=None,
                    help="mode 2: fixed fused tile 'width,threads,depth' instead of the autotuner")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy 5 times (calibrates rocprof FETCH/WRITE_SIZE)")
    a = ap.parse_args()
    a.graph = (a.graph or a.workload == "c3") and not a.no_graph
    return a


def _orc():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    return orc


def host_info():
    """CPU model and the logical CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count()
    return model, nproc


def _cpu_sample(run, nt, probe, seconds, repeats=3):
    """Steps that fit ~seconds (from a probe run), then the best of `repeats` runs."""
    tp = run(probe)
    steps = int(min(nt, max(probe, probe * seconds / max(tp, 1e-6))))
    return steps, min(run(steps) for _ in range(repeats))


def _cpu_record(value, threads, sample):
    model, nproc = host_info()
    return {"value": value, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "sample": sample, "cpu_model": model, "nproc": nproc, "repeats": "best of 3"}


def cpu_baseline_elastic(cfg, seconds, wav):
    """FP32 oracle elastic forward on shot 0: as many steps as fit ~seconds, best of 3."""
    orc = _orc()
    threads = orc.threads(0)
    shot = cfg["shots"][0]
    nxp, nzp = shot.nx + 2 * cfg["pml"], cfg["nz"] + 2 * cfg["pml"]

    def run(nsteps):
        o = orc.make_shot(shot, cfg["nx"], cfg["nz"], cfg["pml"], nsteps, cfg["dx"], cfg["dz"], cfg["dt"])
        t = time.perf_counter()
        orc.elastic_forward(o, cfg["vp"], cfg["vs"], cfg["rho"], wav[:nsteps], prec="f")
        return time.perf_counter() - t

    steps, t = _cpu_sample(run, cfg["nt"], min(40, cfg["nt"]), seconds)
    return _cpu_record(round(nxp * nzp * steps / t / 1e9, 4), threads,
                       f"FP32 oracle elastic forward, shot 0 only ({shot.nx}-column window, {nxp}x{nzp} padded), "
                       f"{steps} of {cfg['nt']} steps, best of 3 runs: {t:.2f} s, OpenMP threads={threads}")


def cpu_baseline_born(cfg, seconds, wav):
    """FP32 oracle plain Born forward (background + Born) on shot 0, best of 3."""
    orc = _orc()
    threads = orc.threads(0)
    shot = cfg["shots"][0]
    nxp, nzp = shot.nx + 2 * cfg["pml"], cfg["nz"] + 2 * cfg["pml"]

    def run(nsteps):
        o = orc.make_shot(shot, cfg["nx"], cfg["nz"], cfg["pml"], nsteps, cfg["dx"], cfg["dz"], cfg["dt"])
        t = time.perf_counter()
        orc.acou_forward(o, cfg["vp"], cfg["rho"], cfg["refl"], wav[:nsteps], prec="f")
        return time.perf_counter() - t

    steps, t = _cpu_sample(run, cfg["nt"], min(100, cfg["nt"]), seconds)
    return _cpu_record(round(2 * nxp * nzp * steps / t / 1e9, 4), threads,
                       f"FP32 oracle acoustic background+Born forward, shot 0 ({nxp}x{nzp} padded), {steps} of "
                       f"{cfg['nt']} steps, best of 3 runs: {t:.2f} s, OpenMP threads={threads}")


def roofline(stats):
    k_ms = float(np.mean([st["kernel_avg_ms"] for st in stats]))
    b = stats[0]["alg_bytes"]
    ach = b / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    out = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel_avg_ms": round(k_ms, 4),
           "alg_bytes_per_launch": b}
    x_ms = float(np.mean([st.get("aux_kernel_avg_ms", 0.0) for st in stats]))
    xb = stats[0].get("aux_alg_bytes", 0.0)
    if x_ms > 0 and xb > 0:
        # two kernels per step (mode 1: velocity kernel + stress kernel)
        out["second_kernel"] = {"kernel_avg_ms": round(x_ms, 4), "alg_bytes_per_launch": xb,
                                "achieved": round(xb / (x_ms * 1e-3) / 1e9, 1)}
        out["step_achieved"] = round((b + xb) / ((k_ms + x_ms) * 1e-3) / 1e9, 1)
    return out


KERNELS = {2: "fused single kernel per step (x-marching, LDS z taps, unsplit interior)",
           1: "two fast kernels per step (unsplit interior, split PML strips)",
           0: "reference split layout, two kernels per step"}


def run_forward(a, cfg):
    from ds2d_amd import Solver
    from ds2d_amd.lib import ricker
    wav = ricker(cfg["nt"], cfg["f0"], cfg["dt"])
    nx_max = max(s.nx for s in cfg["shots"])
    s = Solver(dx=cfg["dx"], dz=cfg["dz"], dt=cfg["dt"], nt=cfg["nt"], allnx=cfg["nx"], allnz=cfg["nz"],
               pml=cfg["pml"], nx_max=nx_max, wavelet=wav, pml_rr=1e-3)
    s.set_model(cfg["vp"], cfg["vs"], cfg["rho"])
    s.set_graph(a.graph)
    s.set_chunk_plan(not a.no_chunk_plan)
    chunk = threads = depth = None
    if a.mode == 2 and a.tile:
        chunk, threads, depth = (int(v) for v in a.tile.split(","))
        s.set_chunk(chunk, threads, depth)
    elif a.mode == 2:  # untimed: pick the fused kernel's tile (or the two-kernel step) for this batch
        chunk, threads, depth, _ = s.autotune(cfg["shots"], nsteps=min(120, cfg["nt"]))
    if a.mode == 2 and a.time_block:
        s.set_time_block(a.time_block)
    tblock = s.time_block() if a.mode == 2 and chunk else 1
    for _ in range(a.warmup):
        s.batch_forward(cfg["shots"], mode=a.mode, fetch=False)
    stats = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        _, _, st = s.batch_forward(cfg["shots"], mode=a.mode, fetch=False)
        stats.append(st)
    elapsed = time.perf_counter() - t0
    # the timed launches' chunk list (two steps per launch): with the balanced plan the chunk widths come from
    # the plan, so the tile's width is not reported
    plan = tblock == 2 and s.chunk_plan_used()
    rz, rx, _ = s.batch_forward(cfg["shots"][:2], mode=a.mode, fetch=True)  # untimed validation
    assert np.isfinite(rz).all() and np.isfinite(rx).all() and np.abs(rz).max() > 0
    cu = stats[0]["cell_updates"]
    out = {
        "metric": "Gcell-updates/s/GPU", "value": round(cu * a.steps / elapsed / 1e9, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(elapsed / a.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic: seeded layered+faulted elastic model (seed {cfg['seed']}), Gardner density",
        "config": {
            "workload": cfg["desc"], "shots": len(cfg["shots"]), "nt": cfg["nt"], "cell_updates_per_step": cu,
            "kernels": KERNELS[1 if (a.mode == 2 and chunk == 0) else a.mode] +
                       (", two time steps per launch (temporal blocking)" if tblock == 2 else "") +
                       (" (autotuned)" if a.mode == 2 and not a.tile else ""), "parallelism": "shots batched on one GPU",
            "autotuned_tile": None if chunk is None else {"x_width": None if plan else chunk, "threads": threads,
                                                          "prefetch_columns": depth,
                                                          "time_steps_per_launch": tblock,
                                                          "chunk_plan": "balanced" if plan else "fixed width"},
            "hipgraph": a.graph},
        "roofline": roofline(stats),
    }
    blocked = stats[0].get("blocked_bytes", 0.0)
    if blocked > 0 and out["roofline"]["kernel_avg_ms"] > 0:
        # two steps per launch: alg_bytes counts both steps' cell-updates at the one-step figure (52 B per
        # padded cell + 48 B per PML cell and step); the kernel itself moves the state once per launch
        bw = blocked / (out["roofline"]["kernel_avg_ms"] * 1e-3) / 1e9
        out["roofline"]["time_blocked"] = {"steps_per_launch": 2, "own_bytes_per_launch": blocked,
                                           "own_achieved": round(bw, 1), "own_frac": round(bw / HBM_PEAK_GBS, 4)}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_elastic(cfg, a.cpu_seconds, wav)
    return out


def run_rot(a, cfg):
    """P/S-decoupled (theta/omega) forward modelling of the C3 windows: one
    fused launch per step (velocity parts, then dilatation/rotation, in one
    x-march), or with --no-fused-rot the two per-cell kernels per step."""
    from ds2d_amd import Solver
    from ds2d_amd.lib import ricker
    wav = ricker(cfg["nt"], cfg["f0"], cfg["dt"])
    nx_max = max(s.nx for s in cfg["shots"])
    s = Solver(dx=cfg["dx"], dz=cfg["dz"], dt=cfg["dt"], nt=cfg["nt"], allnx=cfg["nx"], allnz=cfg["nz"],
               pml=cfg["pml"], nx_max=nx_max, wavelet=wav, pml_rr=1e-3)
    s.set_model(cfg["vp"], cfg["vs"], cfg["rho"])
    s.set_fused_rotation(not a.no_fused_rot)
    for _ in range(a.warmup):
        s.rot_forward(cfg["shots"], fetch=False)
    stats = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        _, _, st = s.rot_forward(cfg["shots"], fetch=False)
        stats.append(st)
    elapsed = time.perf_counter() - t0
    rz, rx, _ = s.rot_forward(cfg["shots"][:2], fetch=True)  # untimed validation
    assert np.isfinite(rz).all() and np.isfinite(rx).all() and np.abs(rz).max() > 0
    cu = stats[0]["cell_updates"]
    out = {
        "metric": "Gcell-updates/s/GPU", "value": round(cu * a.steps / elapsed / 1e9, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(elapsed / a.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic: seeded layered+faulted elastic model (seed {cfg['seed']}), vs = vp/1.73",
        "config": {"workload": cfg["desc"], "shots": len(cfg["shots"]), "nt": cfg["nt"], "cell_updates_per_step": cu,
                   "kernels": ("k_rot_v + k_rot_p per step (per-cell, reference layout of the rotation scheme)"
                               if a.no_fused_rot else "k_rot_fused: one x-marching launch per step (unsplit "
                               "theta/omega outside the damping strips)"),
                   "parallelism": "shots batched on one GPU"},
        "roofline": roofline(stats),
    }
    if not a.no_cpu_baseline:
        orc = _orc()
        threads = orc.threads(0)
        shot = cfg["shots"][0]
        nxp, nzp = shot.nx + 2 * cfg["pml"], cfg["nz"] + 2 * cfg["pml"]

        def run(nsteps):
            o = orc.make_shot(shot, cfg["nx"], cfg["nz"], cfg["pml"], nsteps, cfg["dx"], cfg["dz"], cfg["dt"])
            t = time.perf_counter()
            orc.rot_forward(o, cfg["vp"], cfg["vs"], wav[:nsteps], prec="f")
            return time.perf_counter() - t

        steps, t = _cpu_sample(run, cfg["nt"], min(40, cfg["nt"]), a.cpu_seconds)
        out["cpu_baseline"] = _cpu_record(
            round(nxp * nzp * steps / t / 1e9, 4), threads,
            f"FP32 oracle rotation forward, shot 0 only ({shot.nx}-column window, {nxp}x{nzp} padded), {steps} of "
            f"{cfg['nt']} steps, best of 3 runs: {t:.2f} s, OpenMP threads={threads}")
    return out


def run_decomp(a):
    """Elastic RTM with a decomposed imaging condition (default: the Poynting
    up/down PP image) for 26 shots 10 nodes apart on the 251x101 C1 model
    (pml 20, 2500 steps), on records of the library's own forward modelling."""
    from ds2d_amd import Solver, models
    from ds2d_amd.lib import ricker
    cfg = models.c1_elastic_forward()
    nt = a.nt or cfg["nt"]
    wav = ricker(nt, cfg["f0"], cfg["dt"])
    nx, nz, pml = cfg["nx"], cfg["nz"], cfg["pml"]
    from ds2d_amd.lib import ShotSpec
    shots = [ShotSpec(0, nx, pml + x, pml + 2, pml, pml + np.arange(nx, dtype=np.int32), shot_id=i)
             for i, x in enumerate(range(0, nx, 10))][:a.nshots or 26]
    s = Solver(dx=cfg["dx"], dz=cfg["dz"], dt=cfg["dt"], nt=nt, allnx=nx, allnz=nz, pml=pml, nx_max=nx, wavelet=wav)
    s.set_model(cfg["vp"], cfg["vs"], cfg["rho"])
    rz, _, _ = s.batch_forward(shots, mode=2)
    for _ in range(a.warmup):
        s.decomp_rtm(shots, a.kind, rz)
    times = []
    for _ in range(a.steps):
        t0 = time.perf_counter()
        img = s.decomp_rtm(shots, a.kind, rz)
        times.append(time.perf_counter() - t0)
    assert np.isfinite(img).all() and np.abs(img).max() > 0
    elapsed = float(np.mean(times))
    cu = len(shots) * (nx + 2 * pml) * (nz + 2 * pml) * nt * 2  # source and receiver wavefields
    names = {0: "PP, Poynting up/down split of vector P", 1: "PP vector", 2: "PS vector", 3: "SS vector",
             4: "PP scalar", 5: "PS scalar", 6: "SS scalar"}
    return {
        "metric": "Gcell-updates/s/GPU", "value": round(cu / elapsed / 1e9, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(elapsed * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded layered+faulted elastic model (seed 1), records of the library's forward modelling",
        "config": {"workload": f"decomp_rtm: 251x101, dx=dz=10 m, pml 20, dt 0.8 ms, {nt} steps, {len(shots)} shots "
                               f"10 nodes apart, image {names[a.kind]}",
                   "kernels": "k_ref_v/k_ref_s + k_helmholtz + k_decomp per step, source quantity resident in HBM",
                   "parallelism": "shots batched on one GPU"},
        "roofline": None,
    }


def run_lsrtm(a, cfg):
    """C2: Born data of the seeded reflectivity (untimed), then timed CG LSRTM."""
    from ds2d_amd import Solver
    from ds2d_amd.lib import laplace_filter, ricker
    wav = ricker(cfg["nt"], cfg["f0"], cfg["dt"])
    s = Solver(dx=cfg["dx"], dz=cfg["dz"], dt=cfg["dt"], nt=cfg["nt"], allnx=cfg["nx"], allnz=cfg["nz"],
               pml=cfg["pml"], nx_max=cfg["nx"], wavelet=wav, physics=1)
    s.set_model(cfg["vp"], None, cfg["rho"])
    s.set_graph(a.graph)
    _, d_obs, _ = s.born_forward(cfg["shots"], cfg["refl"])
    for _ in range(a.warmup):
        s.lsrtm_cg(cfg["shots"], d_obs, 1)
    times, res, stats = [], None, []
    for _ in range(a.steps):
        t0 = time.perf_counter()
        img, res, st = s.lsrtm_cg(cfg["shots"], d_obs, a.niter)
        times.append(time.perf_counter() - t0)
        stats.append(st)
    elapsed = float(np.mean(times))
    hit = [k for k, r in enumerate(res) if r <= 1e-2]
    # time to 1e-2: fraction of the run up to the first iteration reaching it (iterations cost the same)
    t_hit = None if not hit else round(elapsed * hit[0] / a.niter, 3)
    t_ref = None
    if hit:
        # The same CG in the reference's scheme (one shot per time loop, background rebuilt by reversal
        # through the random boundary in every adjoint pass): fixed + k * per-iteration cost from 1- and
        # 3-iteration runs, at the iteration count the batched run needed.
        sr = Solver(dx=cfg["dx"], dz=cfg["dz"], dt=cfg["dt"], nt=cfg["nt"], allnx=cfg["nx"], allnz=cfg["nz"],
                    pml=cfg["pml"], nx_max=cfg["nx"], wavelet=wav, physics=1, bg_mode=2)
        sr.set_model(cfg["vp"], None, cfg["rho"])
        sr.set_batch(1)
        tr = []
        for k in (1, 3):
            t0 = time.perf_counter()
            sr.lsrtm_cg(cfg["shots"], d_obs, k)
            tr.append(time.perf_counter() - t0)
        sr.close()
        per_iter_ref = (tr[1] - tr[0]) / 2
        t_ref = round(tr[0] - per_iter_ref + hit[0] * per_iter_ref, 2)
    img_lap = laplace_filter(img, cfg["dx"])
    cu = stats[0]["cell_updates"]
    out = {
        "metric": "Gcell-updates/s/GPU", "value": round(cu / elapsed / 1e9, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(elapsed * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic: seeded layered+faulted model (seed {cfg['seed']}), 9x9-smoothed migration model, "
                "Born data of the true reflectivity",
        "config": {"workload": cfg["desc"], "shots": len(cfg["shots"]), "nt": cfg["nt"], "cg_iterations": a.niter,
                   "residual_history": [round(float(r), 6) for r in res], "time_to_1e-2_s": t_hit,
                   "iterations_to_1e-2": hit[0] if hit else None,
                   "time_to_1e-2_reference_scheme_s": t_ref,
                   "speedup_batched_vs_reference_scheme": round(t_ref / t_hit, 2) if t_ref and t_hit else None,
                   "lsrtm_wall_s": round(elapsed, 3), "laplace_image_rms": float(np.sqrt((img_lap ** 2).mean())),
                   "parallelism": "shots batched on one GPU", "hipgraph": a.graph},
        "roofline": roofline(stats) if stats[0]["kernel_avg_ms"] > 0 else None,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_born(cfg, a.cpu_seconds, wav)
    return out


def run_learned(a):
    """C4: learned LSRTM preconditioner.  Training pairs (RTM image -> 20-iteration
    CG image) from the plain method on seeded 512x256 models, U-Net training
    with the BF16 MFMA kernels, then on each held-out model three solvers run
    to the residual target (the library's stop rule) and each call is timed
    in full, fixed costs included: learned + batched LSRTM, plain batched CG,
    and plain CG in the reference's scheme (one shot per launch, background
    rebuilt by reversal through the random boundary in every adjoint pass).
    The reference scheme's time is also extrapolated per model from 1- and
    3-iteration runs (--ref-full limits the full runs to the first models)."""
    from ds2d_amd import learned as LR
    kw = dict(nx=512, nz=256, nshots=a.nshots or 16, nt=a.nt or 1500)
    t0 = time.perf_counter()

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if a.load_unet:
        from ds2d_amd.lib import Unet
        s = LR.make_solver(LR.acoustic_problem(1000, **kw))
        u = Unet(levels=4, base=32, max_batch=8, max_h=kw["nx"], max_w=kw["nz"], seed=0)
        u.set_params(np.load(a.load_unet))
        t_gen = t_train = 0.0
        loss, tst = np.zeros(20, np.float32), {"ms_per_step": 0.0, "flops_per_step": 0.0}
        side = a.load_unet + ".json"
        if os.path.exists(side):  # the cost of the network being evaluated, recorded when it was trained
            info = json.load(open(side))
            t_gen, t_train = info["pair_generation_s"], info["training_s"]
            loss = np.array([info["train_loss_first_last"][0]] * 20 + [info["train_loss_first_last"][1]] * 20,
                            np.float32)
    else:
        X, T, info, s = LR.generate_pairs(range(1000, 1000 + a.train_models), niter=20, log=log, **kw)
        t_gen = time.perf_counter() - t0
        log(f"pairs: {len(X)} in {t_gen:.1f} s")
        u, loss, tst = LR.train_unet(X, T, steps=a.train_steps, batch=8, lr=1e-3, seed=0)
        t_train = tst["wall_s"]
        log(f"training: loss {loss[:20].mean():.4f} -> {loss[-20:].mean():.4f}, {tst['ms_per_step']:.2f} ms/step")
        if a.save_unet:
            np.save(a.save_unet, u.get_params())
            with open(a.save_unet + ".json", "w") as f:  # what the network cost, for later evaluations
                json.dump({"pair_generation_s": t_gen, "training_s": t_train, "train_models": a.train_models,
                           "train_steps": a.train_steps, "train_loss_first_last":
                           [float(loss[:20].mean()), float(loss[-20:].mean())]}, f)
    seeds = list(range(a.eval_start, a.eval_start + a.eval_models))
    prob = LR.acoustic_problem(seeds[0], **kw)
    # the reference's scheme on its own handle: one shot per launch, no resident history
    sr = LR.make_solver(prob, bg_mode=2)
    sr.set_batch(1)
    for h in (s, sr):  # progress lines from the solvers themselves
        h.set_log(25)
    # untimed warm-up of both handles (per-geometry tuning, workspace allocation)
    for h in (s, sr):
        h.set_model(prob["vp"], None, prob["rho"])
        _, d, _ = h.born_forward(prob["shots"], prob["refl"])
        h.lsrtm_cg(prob["shots"], d, 1)
    s.lsrtm_learned(u, prob["shots"], d, 1)

    def timed(h, fn):
        t1 = time.perf_counter()
        out = fn()
        return out, time.perf_counter() - t1

    rows = []
    for k, seed in enumerate(seeds):
        prob = LR.acoustic_problem(seed, **kw)
        s.set_model(prob["vp"], None, prob["rho"])
        _, d, _ = s.born_forward(prob["shots"], prob["refl"])
        s.set_stop(a.target)
        (img_l, rl, st), tl = timed(s, lambda: s.lsrtm_learned(u, prob["shots"], d, a.max_iter))
        (img_p, rp, _), tp = timed(s, lambda: s.lsrtm_cg(prob["shots"], d, a.max_iter))
        s.set_stop(0)
        kl, kp = LR.first_hit(rl, a.target), LR.first_hit(rp, a.target)
        ref = np.linalg.norm(prob["refl"])
        row = {"seed": seed, "iters_learned": kl, "iters_plain": kp,
               "time_learned_batched_s": round(tl, 2), "time_plain_batched_s": round(tp, 2),
               "final_resid_learned": float(rl[-1]), "final_resid_plain": float(rp[-1]),
               "learned_steps_accepted": st["learned_steps"],
               "image_err_learned": float(np.linalg.norm(img_l - prob["refl"]) / ref),
               "image_err_plain": float(np.linalg.norm(img_p - prob["refl"]) / ref)}
        # reference scheme: fixed + per-iteration cost measured on this model ...
        sr.set_model(prob["vp"], None, prob["rho"])
        _, t1r = timed(sr, lambda: sr.lsrtm_cg(prob["shots"], d, 1))
        _, t3r = timed(sr, lambda: sr.lsrtm_cg(prob["shots"], d, 3))
        per_it, fixed = (t3r - t1r) / 2, t1r - (t3r - t1r) / 2
        row["ref_scheme_per_iter_s"], row["ref_scheme_fixed_s"] = round(per_it, 3), round(fixed, 3)
        if kp is not None:
            row["time_plain_reference_scheme_extrapolated_s"] = round(fixed + kp * per_it, 2)
        # ... and run in full to the target on the first --ref-full models
        if k < a.ref_full:
            sr.set_stop(a.target)
            (_, rr_, _), tr = timed(sr, lambda: sr.lsrtm_cg(prob["shots"], d, a.max_iter))
            sr.set_stop(0)
            row["time_plain_reference_scheme_s"] = round(tr, 2)
            row["iters_plain_reference_scheme"] = LR.first_hit(rr_, a.target)
        rows.append(row)
        log("row " + json.dumps(row))  # each model's record survives a run that stops early
        log(f"seed {seed}: learned {kl} it {tl:.1f} s, plain {kp} it {tp:.1f} s, "
            f"reference scheme {row.get('time_plain_reference_scheme_s', '-')} s "
            f"(extrapolated {row.get('time_plain_reference_scheme_extrapolated_s', '-')} s)")
    sr.close()
    both = [r for r in rows if r["iters_learned"] is not None and r["iters_plain"] is not None]
    amort = (t_gen + t_train) / max(len(seeds), 1)

    def ref_time(r):
        return r.get("time_plain_reference_scheme_s", r.get("time_plain_reference_scheme_extrapolated_s"))

    ratios = [ref_time(r) / r["time_learned_batched_s"] for r in both]
    speed = float(np.mean(ratios)) if ratios else None
    out = {
        "metric": "LSRTM time to 1e-2 data residual, learned+batched vs plain",
        "value": None if speed is None else round(speed, 2),
        "unit": "x (plain GPU CG in the reference's scheme -- one shot per launch, background rebuilt by reversal "
                "-- over learned batched LSRTM, wall time of each solver run to the target, fixed costs included)",
        "n_gpus": 1, "steps": 1, "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 stencils (f64 background), bf16 U-Net (fp32 accumulate)",
        "data": "synthetic: seeded 512x256 layered+faulted models with lenses; Born data of the true reflectivity",
        "config": {"workload": f"c4_learned_lsrtm: {kw['nx']}x{kw['nz']}, {kw['nshots']} shots, nt {kw['nt']}, "
                               f"{a.train_models} training models, {len(seeds)} held-out (seeds {seeds[0]}..), "
                               f"U-Net 4 levels x 32",
                   "target": a.target, "max_iter": a.max_iter, "per_model": rows,
                   "models_reaching_target": {"learned": sum(r["iters_learned"] is not None for r in rows),
                                              "plain": sum(r["iters_plain"] is not None for r in rows),
                                              "of": len(rows)},
                   "speedup_models": len(both),
                   "speedup_reference_times": "measured in full where time_plain_reference_scheme_s is given, "
                                              "else fixed + iterations x per-iteration cost of that model",
                   "speedup_learned_vs_plain_both_batched": round(float(np.mean(
                       [r["time_plain_batched_s"] / r["time_learned_batched_s"] for r in both])), 3) if both else None,
                   "amortized_pairs_and_training_s_per_model": round(amort, 2),
                   "speedup_with_amortized_training": round(float(np.mean(
                       [ref_time(r) / (r["time_learned_batched_s"] + amort) for r in both])), 3) if both else None,
                   "pair_generation_s": round(t_gen, 1), "training_s": round(t_train, 1),
                   "train_steps": a.train_steps, "train_ms_per_step": round(tst["ms_per_step"], 3),
                   "train_tflops": round(tst["flops_per_step"] / tst["ms_per_step"] / 1e9, 2)
                   if tst["ms_per_step"] else None,
                   "train_loss_first_last": [float(loss[:20].mean()), float(loss[-20:].mean())],
                   "image_error": "relative L2 against the true reflectivity"},
    }
    return out


def run_c5(a):
    """C5, the full combined path: PP/PS elastic Born LSRTM of 256 uneven
    windows on the 4000x1000 model, all shots batched.  Plain CG for
    --c5-iter iterations (and the learned solver with --load-unet, the U-Net
    applied to the PP and PS gradients): residual history, seconds per
    iteration, Gcell-updates/s.  (The same solver on a 500x250 window of
    the model against CGLS on the FP64 oracle is a GPU test.)"""
    from ds2d_amd import Solver, Unet, models
    from ds2d_amd.lib import ricker

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    cfg = models.c5_combined(nshots=a.nshots or 256, nt=a.nt or 3000)
    wav = ricker(cfg["nt"], cfg["f0"], cfg["dt"])
    shots = cfg["shots"]
    s = Solver(dx=cfg["dx"], dz=cfg["dz"], dt=cfg["dt"], nt=cfg["nt"], allnx=cfg["nx"], allnz=cfg["nz"],
               pml=cfg["pml"], nx_max=max(x.nx for x in shots), wavelet=wav, physics=0)
    s.set_model(cfg["vp"], cfg["vs"], cfg["rho"])
    bw, bt = (int(v) for v in a.born_tile.split(",")) if a.born_tile else (0, 0)
    s.set_fused_born(not a.no_fused_born, bw, bt)
    s.set_log(1)  # a C5 iteration takes minutes: the solver prints each one
    t0 = time.perf_counter()
    dz, dx, stf = s.eborn_forward(shots, cfg["refl_pp"], cfg["refl_ps"])
    t_fwd = time.perf_counter() - t0
    log(f"c5 Born data: {t_fwd:.1f} s, {stf['cell_updates'] / t_fwd / 1e9:.2f} Gcell/s")
    ref = np.concatenate([cfg["refl_pp"].ravel(), cfg["refl_ps"].ravel()])
    runs = {}
    solvers = [("plain", None)]
    train = None
    params = None
    if a.train_elastic > 0:
        # elastic PP/PS pairs (RTM gradient -> 20-iteration PP/PS CG image) from the plain method on
        # seeded 512x256 elastic models with C5's physics (dx 10 m, dt 0.8 ms, random boundary 40)
        from ds2d_amd import learned as LR
        t1 = time.perf_counter()
        X, T, _, se = LR.generate_elastic_pairs(range(3000, 3000 + a.train_elastic), niter=20, log=log)
        se.close()
        t_gen = time.perf_counter() - t1
        ut, loss, tst = LR.train_unet(X, T, steps=a.train_steps, batch=8, lr=1e-3, seed=0)
        params = ut.get_params()
        ut.close()
        if a.save_unet:
            np.save(a.save_unet, params)
        train = {"elastic_models": a.train_elastic, "samples": int(len(X)), "pair_generation_s": round(t_gen, 1),
                 "train_steps": a.train_steps, "train_s": round(tst["wall_s"], 1),
                 "train_loss_first_last": [float(loss[:20].mean()), float(loss[-20:].mean())]}
        log(f"elastic U-Net: {len(X)} samples in {t_gen:.1f} s, loss {loss[:20].mean():.4f} -> "
            f"{loss[-20:].mean():.4f}")
    elif a.load_unet:
        params = np.load(a.load_unet)
    if params is not None:
        u = Unet(levels=4, base=32, max_batch=1, max_h=cfg["nx"], max_w=cfg["nz"], seed=0)
        u.set_params(params)
        solvers.append(("learned", u))
    for name, u in solvers:
        t0 = time.perf_counter()
        if u is None:
            pp, ps, res, st = s.elsrtm_cg(shots, dz, dx, a.c5_iter)
            acc = 0
        else:
            pp, ps, res, st = s.elsrtm_learned(u, shots, dz, dx, a.c5_iter)
            acc = st["learned_steps"]
        el = time.perf_counter() - t0
        img = np.concatenate([pp.ravel(), ps.ravel()])
        runs[name] = {"wall_s": round(el, 2), "gcell_per_s": round(st["cell_updates"] / el / 1e9, 3),
                      "resid": [round(float(r), 6) for r in res], "learned_steps": acc,
                      "image_rel_err_vs_true_reflectivity": round(float(np.linalg.norm(img - ref) /
                                                                        np.linalg.norm(ref)), 4)}
        log(f"c5 {name}: residual {res[-1]:.4f} after {a.c5_iter} iterations, {el:.1f} s")
    p = runs["plain"]
    return {"metric": "Gcell-updates/s/GPU", "value": p["gcell_per_s"], "unit": "Gcell-updates/s", "n_gpus": 1,
            "steps": a.c5_iter, "warmup": 0, "ms_per_step": round(p["wall_s"] * 1e3 / max(a.c5_iter, 1), 1),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic: seeded 4000x1000 layered+faulted elastic model, Born data of the true PP/PS "
                    "reflectivities",
            "config": {"shots": len(shots),
                       "workload": f"c5_combined: 4000x1000, dx=dz=10 m, {len(shots)} uneven windows (400-1200 "
                                   f"cols), random boundary {cfg['pml']}, dt {cfg['dt'] * 1e3:.1f} ms, "
                                   f"{cfg['nt']} steps, PP/PS elastic Born CG LSRTM",
                       "padded_cells_per_wavefield": int(sum((x.nx + 2 * cfg["pml"]) * (cfg["nz"] + 2 * cfg["pml"])
                                                             for x in shots)),
                       "born_data_s": round(t_fwd, 2), "solvers": runs, "elastic_unet_training": train,
                       "note": "a step = one CG iteration over all shots; 1e-2 is not reached in this many "
                               "iterations (see resid)"},
            # the fused forward step of the Born data pass (k_eb_fused, 100 B per padded cell and step)
            "roofline": roofline([stf]) if stf.get("kernel_avg_ms", 0) > 0 else None}


def run_compat(a):
    """The reference-compat per-shot gradient that main.cpp runs
    (ELSRTM_ROT_EXP_IMAGE_SINGLESHOT through ds2d_batch_gradient): 26 shots
    10 nodes apart on the C1 model (251x101, pml 20, dt 0.8 ms, 2500 steps),
    iteration 1 (Born + second-order Born, NCC residual, source-side
    reconstruction through the saved PML edge bands, receiver adjoint,
    receiver-side Born adjoint, the three imaging terms), constant density
    2000 as the reference.  A step = one gradient over all shots.
    Cell-updates per padded cell and time step: bg, b, b2 forward; bg, b
    reconstructed; bk, bb adjoint = 7 wavefields."""
    from ds2d_amd import Solver, models
    from ds2d_amd.lib import ricker
    cfg = models.c1_elastic_forward()
    nx, nz, pml, nt = cfg["nx"], cfg["nz"], cfg["pml"], a.nt or cfg["nt"]
    shots = [models.surface_shot(nx, 0, pml, min(10 * i, nx - 1), 2, 0, shot_id=i, boundary=0)
             for i in range(a.nshots or 26)]
    wav = ricker(nt, cfg["f0"], cfg["dt"])
    s = Solver(dx=cfg["dx"], dz=cfg["dz"], dt=cfg["dt"], nt=nt, allnx=nx, allnz=nz, pml=pml, nx_max=nx,
               wavelet=wav, pml_rr=1e-3, rho_const=2000.0)
    # observed data: elastic forward of a 3 % faster model (untimed)
    s.set_model(cfg["vp"] * np.float32(1.03), cfg["vs"], None)
    oz, ox, _ = s.batch_forward(shots, mode=2)
    s.set_model(cfg["vp"], cfg["vs"], None)
    rng = np.random.default_rng(4)
    s.set_image((0.02 * rng.standard_normal((nx, nz))).astype(np.float32),
                (0.02 * rng.standard_normal((nx, nz))).astype(np.float32))
    for _ in range(a.warmup):
        s.batch_gradient(shots, oz, ox, 1)
        s.take_gradients()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        s.batch_gradient(shots, oz, ox, 1)
    elapsed = time.perf_counter() - t0
    g = s.take_gradients()
    assert all(np.isfinite(x).all() for x in g) and np.abs(g[0]).max() > 0
    cells = len(shots) * (nx + 2 * pml) * (nz + 2 * pml)
    cu = 7 * cells * nt
    return {"metric": "Gcell-updates/s/GPU", "value": round(cu * a.steps / elapsed / 1e9, 3),
            "unit": "Gcell-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": round(elapsed / a.steps * 1e3, 2), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 (background and first-order Born field f64)",
            "data": "synthetic: seeded layered+faulted elastic model (seed 1), observed data from a 3 % faster model",
            "config": {"workload": f"compat_gradient: 251x101, dx=dz=10 m, pml 20, dt 0.8 ms, {nt} steps, "
                                   f"{len(shots)} shots 10 nodes apart, reference ELSRTM gradient (iteration 1)",
                       "shots": len(shots), "nt": nt, "cell_updates_per_step": cu,
                       "kernels": "reference split-PML layout, per-cell kernels (ds2d_batch_gradient)",
                       "parallelism": "shots batched on one GPU"}}


def run_unet(a):
    """U-Net training throughput: Adam steps (batch 8) of the 4-level, 32-channel
    U-Net on 512x256 images with the BF16 MFMA kernels (synthetic data)."""
    from ds2d_amd import Unet
    rng = np.random.default_rng(0)
    x = rng.standard_normal((16, 512, 256)).astype(np.float32)
    t = rng.standard_normal((16, 512, 256)).astype(np.float32)
    u = Unet(levels=4, base=32, max_batch=8, max_h=512, max_w=256, seed=1)
    u.set_data(x, t)
    u.train(max(1, a.warmup), batch=8, lr=1e-3, seed=2)
    n = max(1, a.steps) * 10
    _, st = u.train(n, batch=8, lr=1e-3, seed=3)
    tf = st["flops_per_step"] / st["ms_per_step"] / 1e9
    return {"metric": "U-Net training TFLOP/s", "value": round(tf, 2), "unit": "TFLOP/s", "n_gpus": 1, "steps": n,
            "warmup": max(1, a.warmup), "ms_per_step": round(st["ms_per_step"], 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)",
            "data": "synthetic: Gaussian 512x256 images",
            "config": {"workload": "unet_train: 4 levels x 32 channels, batch 8, 512x256, Adam",
                       "flops_per_step": st["flops_per_step"]},
            "roofline": {"bound": "mfma", "achieved": round(tf, 2), "peak": 2500.0, "unit": "TFLOP/s",
                         "frac": round(tf / 2500.0, 4), "traffic": None}}


def copy_probe():
    """A plain device copy of known size: 256 MiB read + 256 MiB written per launch."""
    import torch
    x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    print(json.dumps({"copy_probe_bytes_read": x.numel() * 4, "copy_probe_bytes_written": y.numel() * 4,
                      "ms_per_copy": round(ms, 4), "GB_per_s": round(2 * x.numel() * 4 / (ms * 1e-3) / 1e9, 1)}))


def main():
    a = parse()
    if a.copy_probe:
        copy_probe()
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {a.gpus} is not supported by this build "
              "(shots shard across GPUs by design, but no multi-GPU path is built)", file=sys.stderr)
        sys.exit(2)
    from ds2d_amd import device_info, models
    if a.workload == "c3":
        cfg = models.c3_batched_elastic(nshots=a.nshots or 128, nt=a.nt or 3000)
        cfg["seed"] = 3
        edge = f"pml {cfg['pml']}"
        if a.random_boundary:  # the same windows with the random boundary instead of the PML
            for sh in cfg["shots"]:
                sh.boundary = 1
            edge = f"random boundary {cfg['pml']}"
        cfg["desc"] = (f"c3_batched_elastic: {cfg['nx']}x{cfg['nz']} model, dx=dz={cfg['dx']} m, {len(cfg['shots'])} "
                       f"uneven windows (400-1200 cols), {edge}, dt {cfg['dt'] * 1e3:.1f} ms, {cfg['nt']} steps")
        out = run_forward(a, cfg)
    elif a.workload == "rot":
        cfg = models.c3_batched_elastic(nshots=a.nshots or 128, nt=a.nt or 3000)
        cfg["seed"] = 3
        cfg["desc"] = (f"rot_batched: the {len(cfg['shots'])} C3 windows ({cfg['nx']}x{cfg['nz']} model, dx=dz="
                       f"{cfg['dx']} m, pml {cfg['pml']}, {cfg['nt']} steps) with the P/S-decoupled theta/omega scheme")
        out = run_rot(a, cfg)
    elif a.workload == "decomp":
        out = run_decomp(a)
    elif a.workload == "c1":
        cfg = models.c1_elastic_forward()
        if a.nt:
            cfg["nt"] = a.nt
        cfg["seed"] = 1
        cfg["desc"] = (f"c1_elastic_forward: 251x101, dx=dz=10 m, pml 20, dt 0.8 ms, {cfg['nt']} steps, one shot, "
                       "vx/vz at every surface node")
        out = run_forward(a, cfg)
    elif a.workload == "c4":
        out = run_learned(a)
    elif a.workload == "c5":
        out = run_c5(a)
    elif a.workload == "unet":
        out = run_unet(a)
    elif a.workload == "compat":
        out = run_compat(a)
    else:
        cfg = models.c2_acoustic_lsrtm(nt=a.nt or 2500, nshots=a.nshots or 26)
        cfg["seed"] = 2
        cfg["desc"] = (f"c2_acoustic_lsrtm: 251x101, dx=dz=10 m, {len(cfg['shots'])} shots 10 nodes apart, random "
                       f"boundary 40 (+-30%), dt {cfg['dt'] * 1e3:.1f} ms, {cfg['nt']} steps, {a.niter} CG iterations")
        out = run_lsrtm(a, cfg)
    name, ncu, mem = device_info()
    out["device"] = {"arch": name, "cus": ncu, "hbm_gb": round(mem, 1)}
    traffic_json = a.traffic_json
    if a.workload == "rot" and traffic_json == TRAFFIC_JSON:  # the fused rotation step's own counters
        traffic_json = None if a.no_fused_rot else os.path.join(ROOT, "profiles", "k_rot_fused_c3_traffic.json")
    if a.workload == "c5" and traffic_json == TRAFFIC_JSON:  # the fused Born forward step's counters
        traffic_json = None if a.no_fused_born else os.path.join(ROOT, "profiles", "k_eb_fused_c5_traffic.json")
    if traffic_json and os.path.exists(traffic_json) and out.get("roofline"):
        tr = json.load(open(traffic_json))
        cfgd = out.get("config", {})
        tile = cfgd.get("autotuned_tile") or {}
        key = [tile.get("x_width"), tile.get("threads"), tile.get("prefetch_columns"),
               tile.get("time_steps_per_launch", 1)]
        # per-launch traffic depends on the batch and the tile, not on the number of steps
        if (tr.get("workload_prefix") and str(cfgd.get("workload", "")).startswith(tr["workload_prefix"])
                and tr.get("shots") == cfgd.get("shots")):
            balanced = tile.get("chunk_plan") == "balanced"
            rb = bool(getattr(a, "random_boundary", False))
            for e in tr.get("entries", [tr]):
                if bool(e.get("random_boundary", tr.get("random_boundary", False))) != rb:
                    continue
                if e.get("tile") is None:  # a workload without tiles: one kernel, one figure
                    if not tile:
                        out["roofline"]["traffic"] = e.get("traffic_bytes_per_launch")
                    continue
                t = list(e.get("tile") or [])
                if len(t) == 3:  # entries without a time block are one step per launch
                    t.append(1)
                if (e.get("chunk_plan") == "balanced") != balanced:
                    continue
                # the balanced plan does not depend on the width
                if (t[1:] == key[1:]) if balanced else (t == key):
                    out["roofline"]["traffic"] = e.get("traffic_bytes_per_launch")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
