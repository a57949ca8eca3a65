"""Learned LSRTM preconditioner workflow on the GPU library.

1. Training pairs come from the plain method itself: for a seeded acoustic
   model, Born data d = L m_true, the RTM image g0 = L^T d, and the image
   after `niter` plain CG iterations.  Both are normalised by their RMS.
2. The U-Net (ds2d_amd.lib.Unet) is trained on those pairs with Adam.
3. LSRTM on held-out models runs ds2d_lsrtm_learned (U-Net direction,
   plain CGLS fallback) beside ds2d_lsrtm_cg.
The elastic variant does the same with PP and PS images of the plain elastic
Born pair (ds2d_elsrtm_learned beside ds2d_elsrtm_cg).
"""
import time

import numpy as np

from .lib import Solver, Unet, ricker
from .models import elastic_from_vp, gaussian_smooth, layered_model, surface_shot


def acoustic_problem(seed, nx=512, nz=256, nshots=16, nt=1500, dt=0.0008, dx=10.0, pml=40, f0=15.0, lenses=3):
    """Seeded layered + faulted model with Gaussian lenses (+-15 % vp), its
    9x9-smoothed migration model, reflectivity (v - v_mig) / v_mig and
    `nshots` surface shots spread evenly (random boundary `pml` wide)."""
    vp_true = layered_model(nx, nz, seed, lenses=lenses)
    vp_mig = gaussian_smooth(vp_true)
    refl = ((vp_true - vp_mig) / vp_mig).astype(np.float32)
    _, rho = elastic_from_vp(vp_mig)
    xs = [int((i + 0.5) * nx / nshots) for i in range(nshots)]
    shots = [surface_shot(nx, 0, pml, x, 2, 0, shot_id=i, boundary=1) for i, x in enumerate(xs)]
    return dict(seed=seed, nx=nx, nz=nz, dx=dx, dz=dx, dt=dt, nt=nt, pml=pml, f0=f0, vp=vp_mig, vp_true=vp_true,
                rho=rho, refl=refl, shots=shots)


def make_solver(prob, bg_mode=1):
    wav = ricker(prob["nt"], prob["f0"], prob["dt"])
    return Solver(dx=prob["dx"], dz=prob["dz"], dt=prob["dt"], nt=prob["nt"], allnx=prob["nx"], allnz=prob["nz"],
                  pml=prob["pml"], nx_max=prob["nx"], wavelet=wav, physics=1, bg_mode=bg_mode)


def _rms(a):
    r = float(np.sqrt(np.mean(np.square(a, dtype=np.float64))))
    return r if r > 0 else 1.0


def generate_pairs(seeds, niter=20, solver=None, log=None, **kw):
    """(X, T, info): RTM images and `niter`-iteration CG images, RMS-normalised,
    shape [len(seeds)][nx][nz].  log: optional callable for progress lines."""
    X, T, info = [], [], []
    s = solver
    seeds = list(seeds)
    for i, seed in enumerate(seeds):
        prob = acoustic_problem(seed, **kw)
        if s is None:
            s = make_solver(prob)
        s.set_model(prob["vp"], None, prob["rho"])
        _, d, _ = s.born_forward(prob["shots"], prob["refl"])
        g0, _ = s.born_adjoint(prob["shots"], d)
        img, res, _ = s.lsrtm_cg(prob["shots"], d, niter)
        X.append(g0 / _rms(g0))
        T.append(img / _rms(img))
        info.append(dict(seed=seed, resid=res.tolist()))
        if log:
            log(f"pair {i + 1}/{len(seeds)} (seed {seed}): {niter}-iteration residual {res[-1]:.4f}")
    return np.asarray(X, np.float32), np.asarray(T, np.float32), info, s


def train_unet(X, T, steps=600, batch=8, lr=1e-3, seed=0, levels=4, base=32, unet=None):
    n, h, w = X.shape
    u = unet or Unet(levels=levels, base=base, max_batch=batch, max_h=h, max_w=w, seed=seed)
    u.set_data(X, T)
    t0 = time.perf_counter()
    loss, st = u.train(steps, batch=batch, lr=lr, seed=seed + 1)
    st["wall_s"] = time.perf_counter() - t0
    return u, loss, st


def first_hit(res, target=1e-2):
    k = [i for i, r in enumerate(res) if r <= target]
    return k[0] if k else None


def compare(solver, unet, prob, niter, target=1e-2, log=None):
    """Plain CG vs learned LSRTM on one problem: residual histories, wall
    times, iterations to `target`, image error against the true reflectivity."""
    solver.set_model(prob["vp"], None, prob["rho"])
    _, d, _ = solver.born_forward(prob["shots"], prob["refl"])
    out = {}
    for name in ("plain", "learned"):
        t0 = time.perf_counter()
        if name == "plain":
            img, res, st = solver.lsrtm_cg(prob["shots"], d, niter)
            acc = 0
        else:
            img, res, st = solver.lsrtm_learned(unet, prob["shots"], d, niter)
            acc = st["learned_steps"]
        wall = time.perf_counter() - t0
        err = float(np.linalg.norm(img - prob["refl"]) / np.linalg.norm(prob["refl"]))
        k = first_hit(res, target)
        if log:
            log(f"seed {prob['seed']} {name}: residual {res[-1]:.4f} after {niter} iterations, {wall:.1f} s")
        out[name] = dict(resid=[float(r) for r in res], wall_s=wall, iters_to_target=k, image_rel_err=err,
                         learned_steps=acc)
    return out


# ---------------------------------------------------------------------------
# Elastic PP/PS variant (plain-L2 elastic Born pair)

def elastic_problem(seed, nx=512, nz=256, nshots=16, nt=1500, dt=0.0008, dx=10.0, pml=40, f0=15.0, lenses=3):
    """Seeded elastic model: vp as acoustic_problem, vs = vp/1.73 scaled by
    +-15 % from a second seeded layering (so PS differs from PP), Gardner
    density of the migration vp; 9x9-smoothed migration vp and vs; PP and PS
    reflectivities (v - v_mig)/v_mig; surface shots, random boundary."""
    vp_true = layered_model(nx, nz, seed, lenses=lenses)
    r2 = layered_model(nx, nz, seed + 500).astype(np.float64)
    r2 = 2 * (r2 - r2.min()) / max(r2.max() - r2.min(), 1.0) - 1
    vs_true = (vp_true / 1.73 * (1 + 0.15 * r2)).astype(np.float32)
    vp_mig, vs_mig = gaussian_smooth(vp_true), gaussian_smooth(vs_true)
    _, rho = elastic_from_vp(vp_mig)
    refl_pp = ((vp_true - vp_mig) / vp_mig).astype(np.float32)
    refl_ps = ((vs_true - vs_mig) / vs_mig).astype(np.float32)
    xs = [int((i + 0.5) * nx / nshots) for i in range(nshots)]
    shots = [surface_shot(nx, 0, pml, x, 2, 1, shot_id=i, boundary=1) for i, x in enumerate(xs)]
    return dict(seed=seed, nx=nx, nz=nz, dx=dx, dz=dx, dt=dt, nt=nt, pml=pml, f0=f0, vp=vp_mig, vs=vs_mig, rho=rho,
                refl_pp=refl_pp, refl_ps=refl_ps, shots=shots)


def make_elastic_solver(prob, bg_mode=0):
    wav = ricker(prob["nt"], prob["f0"], prob["dt"])
    return Solver(dx=prob["dx"], dz=prob["dz"], dt=prob["dt"], nt=prob["nt"], allnx=prob["nx"], allnz=prob["nz"],
                  pml=prob["pml"], nx_max=prob["nx"], wavelet=wav, physics=0, bg_mode=bg_mode)


def generate_elastic_pairs(seeds, niter=20, solver=None, log=None, **kw):
    """(X, T, info, solver): PP and PS gradient images (RTM) and their
    `niter`-iteration CG images, RMS-normalised, two samples per model."""
    X, T, info = [], [], []
    s = solver
    for i, seed in enumerate(list(seeds)):
        prob = elastic_problem(seed, **kw)
        if s is None:
            s = make_elastic_solver(prob)
        s.set_model(prob["vp"], prob["vs"], prob["rho"])
        dz, dx, _ = s.eborn_forward(prob["shots"], prob["refl_pp"], prob["refl_ps"])
        gpp, gps, _ = s.eborn_adjoint(prob["shots"], dz, dx)
        ipp, ips, res, _ = s.elsrtm_cg(prob["shots"], dz, dx, niter)
        for g, img in ((gpp, ipp), (gps, ips)):
            X.append(g / _rms(g))
            T.append(img / _rms(img))
        info.append(dict(seed=seed, resid=res.tolist()))
        if log:
            log(f"elastic pair {i + 1}: {niter}-iteration residual {res[-1]:.4f}")
    return np.asarray(X, np.float32), np.asarray(T, np.float32), info, s


def compare_elastic(solver, unet, prob, niter, target=1e-2, log=None):
    """Plain PP/PS CG vs learned elastic LSRTM on one problem."""
    solver.set_model(prob["vp"], prob["vs"], prob["rho"])
    dz, dx, _ = solver.eborn_forward(prob["shots"], prob["refl_pp"], prob["refl_ps"])
    ref = np.concatenate([prob["refl_pp"].ravel(), prob["refl_ps"].ravel()])
    out = {}
    for name in ("plain", "learned"):
        t0 = time.perf_counter()
        if name == "plain":
            ipp, ips, res, st = solver.elsrtm_cg(prob["shots"], dz, dx, niter)
            acc = 0
        else:
            ipp, ips, res, st = solver.elsrtm_learned(unet, prob["shots"], dz, dx, niter)
            acc = st["learned_steps"]
        wall = time.perf_counter() - t0
        img = np.concatenate([ipp.ravel(), ips.ravel()])
        err = float(np.linalg.norm(img - ref) / np.linalg.norm(ref))
        k = first_hit(res, target)
        if log:
            log(f"seed {prob['seed']} elastic {name}: residual {res[-1]:.4f} after {niter} iterations, {wall:.1f} s")
        out[name] = dict(resid=[float(r) for r in res], wall_s=wall, iters_to_target=k, image_rel_err=err,
                         learned_steps=acc)
    return out
