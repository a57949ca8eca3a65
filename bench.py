#!/usr/bin/env python3
"""DS2D hot-path benchmark on one MI355X.

A "step" is one pass of the hot path over one batch: batched elastic forward
modelling of 128 uneven-aperture shots (windows U{400..1200} columns) on the
seeded 2000x600 layered+faulted model, dx=dz=12.5 m, 20-cell PML, dt = 1 ms,
3000 time steps, vx/vz recorded at every window surface node.  All shots run
in one launch per time-step kernel.  The model and every wavefield are
resident in HBM before timing; records stay in HBM (fetched once, untimed).

Prints one JSON line with value = Gcell-updates/s (one cell-update = one
padded cell advanced one step for one wavefield), the roofline of the
dominant kernel (HIP-event timing inside the library, algorithmic bytes of
the layout actually run) and a CPU baseline (the FP32 C++/OpenMP oracle on a
bounded sample of the same workload, timed on this host).
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
    ap.add_argument("--nshots", type=int, default=128)
    ap.add_argument("--nt", type=int, default=3000)
    ap.add_argument("--mode", type=int, default=2, help="2 fused kernel (default), 1 two fast kernels, 0 reference layout")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--traffic-json", default=None,
                    help="rocprof FETCH/WRITE summary (profiles/*.json) to embed as roofline.traffic")
    return ap.parse_args()


def cpu_baseline(cfg, seconds):
    """FP32 oracle (C++/OpenMP, test-infrastructure port of the reference
    kernels) on shot 0 of the batch for as many steps as fit ~`seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    from cases import wavelet

    threads = orc.threads(0)
    shot = cfg["shots"][0]
    nxp, nzp = shot.nx + 2 * cfg["pml"], cfg["nz"] + 2 * cfg["pml"]
    wav_full = wavelet(cfg["nt"], cfg["f0"], cfg["dt"])

    def run(nsteps):
        o = orc.make_shot(shot, cfg["nx"], cfg["nz"], cfg["pml"], nsteps, cfg["dx"], cfg["dz"], cfg["dt"])
        t = time.perf_counter()
        orc.elastic_forward(o, cfg["vp"], cfg["vs"], cfg["rho"], wav_full[:nsteps], prec="f")
        return time.perf_counter() - t

    probe = 40
    tp = run(probe)
    steps = int(min(cfg["nt"], max(probe, probe * seconds / max(tp, 1e-6))))
    t = run(steps)
    rate = nxp * nzp * steps / t / 1e9
    return {"value": round(rate, 4), "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"FP32 oracle, shot 0 only ({shot.nx}-column window, {nxp}x{nzp} padded), "
                      f"{steps} of {cfg['nt']} steps, {t:.1f} s, OpenMP threads={threads}"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {a.gpus} is not supported by this build "
              "(shots shard across GPUs by design, but no multi-GPU path is built)", file=sys.stderr)
        sys.exit(2)
    from ds2d_amd import Solver, device_info, models
    from ds2d_amd.lib import ricker

    cfg = models.c3_batched_elastic(nshots=a.nshots, nt=a.nt)
    wav = ricker(cfg["nt"], cfg["f0"], cfg["dt"])
    nx_max = max(s.nx for s in cfg["shots"])
    s = Solver(dx=cfg["dx"], dz=cfg["dz"], dt=cfg["dt"], nt=cfg["nt"], allnx=cfg["nx"], allnz=cfg["nz"],
               pml=cfg["pml"], nx_max=nx_max, wavelet=wav, pml_rr=1e-3)
    s.set_model(cfg["vp"], cfg["vs"], cfg["rho"])
    chunk, tune_ms = (None, None)
    if a.mode == 2:
        # untimed: pick the fused kernel's x-chunk width for this batch
        chunk, tune_ms = s.autotune(cfg["shots"], nsteps=min(60, cfg["nt"]))
    for _ in range(a.warmup):
        s.batch_forward(cfg["shots"], mode=a.mode, fetch=False)
    stats = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        _, _, st = s.batch_forward(cfg["shots"], mode=a.mode, fetch=False)
        stats.append(st)
    elapsed = time.perf_counter() - t0
    # validation pass (untimed): records finite and non-trivial
    rz, rx, _ = s.batch_forward(cfg["shots"][:2], mode=a.mode, fetch=True)
    assert np.isfinite(rz).all() and np.isfinite(rx).all() and np.abs(rz).max() > 0

    cu = stats[0]["cell_updates"]
    value = cu * a.steps / elapsed / 1e9
    k_ms = float(np.mean([st["kernel_avg_ms"] for st in stats]))
    bytes_launch = stats[0]["alg_bytes"]
    achieved = bytes_launch / (k_ms * 1e-3) / 1e9
    traffic = None
    if a.traffic_json and os.path.exists(a.traffic_json):
        traffic = json.load(open(a.traffic_json)).get("traffic_bytes_per_launch")
    name, ncu, mem = device_info()
    out = {
        "metric": "Gcell-updates/s/GPU",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded layered+faulted elastic model (seed 3), Gardner density",
        "config": {
            "workload": f"c3_batched_elastic: {cfg['nx']}x{cfg['nz']} model, dx=dz={cfg['dx']} m, "
                        f"{len(cfg['shots'])} uneven windows (400-1200 cols), pml {cfg['pml']}, "
                        f"dt {cfg['dt'] * 1e3:.1f} ms, {cfg['nt']} steps, one launch per kernel per step",
            "shots": len(cfg["shots"]),
            "nt": cfg["nt"],
            "cell_updates_per_step": cu,
            "kernels": {2: "fused single kernel per step (x-marching, LDS z taps, unsplit interior)",
                        1: "two fast kernels per step (unsplit interior, split PML strips)",
                        0: "reference split layout, two kernels per step"}[a.mode],
            "parallelism": "shots batched on one GPU",
            "autotuned_chunk_width": chunk,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel_avg_ms": round(k_ms, 4),
            "alg_bytes_per_launch": bytes_launch,
        },
        "device": {"arch": name, "cus": ncu, "hbm_gb": round(mem, 1)},
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
