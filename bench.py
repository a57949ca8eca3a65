#!/usr/bin/env python3
"""DS2D hot-path benchmark on one MI355X.

Default workload (c3, the N=1 line): a "step" is one pass of the hot path
over one batch = batched elastic forward modelling of 128 uneven-aperture
shots (windows U{400..1200} columns) on the seeded 2000x600 layered+faulted
model, dx=dz=12.5 m, 20-cell PML, dt = 1 ms, 3000 time steps, vx/vz recorded
at every window surface node.  All shots run in one launch per time step.
The model and every wavefield are resident in HBM before timing; records
stay in HBM (fetched once, untimed).

Other workloads (--workload): c1 = single-shot elastic forward on 251x101
(2500 steps); c2 = acoustic CG LSRTM of 26 random-boundary shots on 251x101
(time to a relative data residual, residual history).

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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3"])
    ap.add_argument("--nshots", type=int, default=None)
    ap.add_argument("--nt", type=int, default=None)
    ap.add_argument("--niter", type=int, default=10, help="c2: CG iterations")
    ap.add_argument("--mode", type=int, default=2, help="2 fused kernel (default), 1 two fast kernels, 0 reference layout")
    ap.add_argument("--graph", action="store_true", help="capture each time loop in one hipGraph")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--traffic-json", default=None,
                    help="rocprof FETCH/WRITE summary (JSON) to embed as roofline.traffic")
    return ap.parse_args()


def _orc():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    return orc


def cpu_baseline_elastic(cfg, seconds, wav):
    """FP32 oracle elastic forward on shot 0 for as many steps as fit ~seconds."""
    orc = _orc()
    threads = orc.threads(0)
    shot = cfg["shots"][0]
    nxp, nzp = shot.nx + 2 * cfg["pml"], cfg["nz"] + 2 * cfg["pml"]

    def run(nsteps):
        o = orc.make_shot(shot, cfg["nx"], cfg["nz"], cfg["pml"], nsteps, cfg["dx"], cfg["dz"], cfg["dt"])
        t = time.perf_counter()
        orc.elastic_forward(o, cfg["vp"], cfg["vs"], cfg["rho"], wav[:nsteps], prec="f")
        return time.perf_counter() - t

    probe = min(40, cfg["nt"])
    tp = run(probe)
    steps = int(min(cfg["nt"], max(probe, probe * seconds / max(tp, 1e-6))))
    t = run(steps)
    return {"value": round(nxp * nzp * steps / t / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"FP32 oracle elastic forward, shot 0 only ({shot.nx}-column window, {nxp}x{nzp} padded), "
                      f"{steps} of {cfg['nt']} steps, {t:.1f} s, OpenMP threads={threads}"}


def cpu_baseline_born(cfg, seconds, wav):
    """FP32 oracle plain Born forward (background + Born) on shot 0."""
    orc = _orc()
    threads = orc.threads(0)
    shot = cfg["shots"][0]
    nxp, nzp = shot.nx + 2 * cfg["pml"], cfg["nz"] + 2 * cfg["pml"]

    def run(nsteps):
        o = orc.make_shot(shot, cfg["nx"], cfg["nz"], cfg["pml"], nsteps, cfg["dx"], cfg["dz"], cfg["dt"])
        t = time.perf_counter()
        orc.acou_forward(o, cfg["vp"], cfg["rho"], cfg["refl"], wav[:nsteps], prec="f")
        return time.perf_counter() - t

    probe = min(100, cfg["nt"])
    tp = run(probe)
    steps = int(min(cfg["nt"], max(probe, probe * seconds / max(tp, 1e-6))))
    t = run(steps)
    return {"value": round(2 * nxp * nzp * steps / t / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"FP32 oracle acoustic background+Born forward, shot 0 ({nxp}x{nzp} padded), {steps} of "
                      f"{cfg['nt']} steps, {t:.1f} s, OpenMP threads={threads}"}


def roofline(stats):
    k_ms = float(np.mean([st["kernel_avg_ms"] for st in stats]))
    b = stats[0]["alg_bytes"]
    ach = b / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel_avg_ms": round(k_ms, 4),
            "alg_bytes_per_launch": b}


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
    chunk = None
    if a.mode == 2:  # untimed: pick the fused kernel's x-chunk width for this batch
        chunk, _ = s.autotune(cfg["shots"], nsteps=min(60, cfg["nt"]))
    for _ in range(a.warmup):
        s.batch_forward(cfg["shots"], mode=a.mode, fetch=False)
    stats = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        _, _, st = s.batch_forward(cfg["shots"], mode=a.mode, fetch=False)
        stats.append(st)
    elapsed = time.perf_counter() - t0
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
            "kernels": KERNELS[a.mode], "parallelism": "shots batched on one GPU",
            "autotuned_chunk_width": chunk, "hipgraph": a.graph},
        "roofline": roofline(stats),
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_elastic(cfg, a.cpu_seconds, wav)
    return out


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
                   "lsrtm_wall_s": round(elapsed, 3), "laplace_image_rms": float(np.sqrt((img_lap ** 2).mean())),
                   "parallelism": "shots batched on one GPU", "hipgraph": a.graph},
        "roofline": roofline(stats) if stats[0]["kernel_avg_ms"] > 0 else None,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_born(cfg, a.cpu_seconds, wav)
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {a.gpus} is not supported by this build "
              "(shots shard across GPUs by design, but no multi-GPU path is built)", file=sys.stderr)
        sys.exit(2)
    from ds2d_amd import device_info, models
    if a.workload == "c3":
        cfg = models.c3_batched_elastic(nshots=a.nshots or 128, nt=a.nt or 3000)
        cfg["seed"] = 3
        cfg["desc"] = (f"c3_batched_elastic: {cfg['nx']}x{cfg['nz']} model, dx=dz={cfg['dx']} m, {len(cfg['shots'])} "
                       f"uneven windows (400-1200 cols), pml {cfg['pml']}, dt {cfg['dt'] * 1e3:.1f} ms, {cfg['nt']} steps")
        out = run_forward(a, cfg)
    elif a.workload == "c1":
        cfg = models.c1_elastic_forward()
        if a.nt:
            cfg["nt"] = a.nt
        cfg["seed"] = 1
        cfg["desc"] = (f"c1_elastic_forward: 251x101, dx=dz=10 m, pml 20, dt 0.8 ms, {cfg['nt']} steps, one shot, "
                       "vx/vz at every surface node")
        out = run_forward(a, cfg)
    else:
        cfg = models.c2_acoustic_lsrtm(nt=a.nt or 2500, nshots=a.nshots or 26)
        cfg["seed"] = 2
        cfg["desc"] = (f"c2_acoustic_lsrtm: 251x101, dx=dz=10 m, {len(cfg['shots'])} shots 10 nodes apart, random "
                       f"boundary 40 (+-30%), dt {cfg['dt'] * 1e3:.1f} ms, {cfg['nt']} steps, {a.niter} CG iterations")
        out = run_lsrtm(a, cfg)
    name, ncu, mem = device_info()
    out["device"] = {"arch": name, "cus": ncu, "hbm_gb": round(mem, 1)}
    if a.traffic_json and os.path.exists(a.traffic_json) and out.get("roofline"):
        out["roofline"]["traffic"] = json.load(open(a.traffic_json)).get("traffic_bytes_per_launch")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
