"""The benchmark configurations C2, C4 and C5 on the GPU at their full sizes
(or, for C5, on a 500x250 window of the full model), against the FP64 oracle
and the U-Net oracle:

- C2: CG LSRTM iterations on a subset of the 26 shots at the full 251x101
  grid and 2500 steps, against CGLS driven by the FP64 oracle's Born pair,
  and the Laplacian filter (laplace.cpp:82-106) of the final image.
- C4: the U-Net loss and gradient at the training shape (batch 8 of 512x256
  images) against the U-Net oracle; learned vs plain LSRTM on a held-out
  512x256 model with a network trained on pairs the plain method made.
- C5: PP/PS elastic Born CG on a 500x250 window of the 4000x1000 model
  (2 shots, 800 steps, 2 iterations) against CGLS on the FP64 oracle pair.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as orc
from ds2d_amd import Solver, models
from ds2d_amd.lib import ricker

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import unet_ref as U  # noqa: E402

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _cgls(L, LT, d, niter, n_model):
    """CGLS in FP64 on the oracle's operators: image and residual history."""
    x = np.zeros(n_model)
    r = d.copy()
    s = LT(r)
    p, gamma, hist = s.copy(), float(s @ s), [1.0]
    for _ in range(niter):
        q = L(p)
        alpha = gamma / float(q @ q)
        x += alpha * p
        r -= alpha * q
        hist.append(float(np.linalg.norm(r) / np.linalg.norm(d)))
        s = LT(r)
        gnew = float(s @ s)
        p, gamma = s + (gnew / gamma) * p, gnew
    return x, hist


def test_c2_cg_iterations_at_full_size_and_laplacian():
    """Two CG iterations with shots 0, 13 and 25 of C2 (331x181 padded, 2500
    steps, +-30 % random boundary) match CGLS on the FP64 oracle pair: image
    and residuals within 1e-5.  The Laplacian of the GPU image is
    bit-identical to the oracle's FP32 filter of the same image."""
    cfg = models.c2_acoustic_lsrtm()
    shots = [cfg["shots"][k] for k in (0, 13, 25)]
    nx, nz, nt = cfg["nx"], cfg["nz"], cfg["nt"]
    wav = ricker(nt, cfg["f0"], cfg["dt"])
    s = Solver(dx=cfg["dx"], dz=cfg["dz"], dt=cfg["dt"], nt=nt, allnx=nx, allnz=nz, pml=cfg["pml"], nx_max=nx,
               wavelet=wav, physics=1)
    s.set_model(cfg["vp"], None, cfg["rho"])
    _, d, _ = s.born_forward(shots, cfg["refl"])
    img, res, _ = s.lsrtm_cg(shots, d, 2)
    osh = [orc.make_shot(x, nx, nz, cfg["pml"], nt, cfg["dx"], cfg["dz"], cfg["dt"]) for x in shots]
    n = [x.ntr * nt for x in shots]

    def L(m):
        return np.concatenate([orc.acou_forward(o, cfg["vp"], cfg["rho"], m.reshape(nx, nz), wav)[1] for o in osh])

    def LT(r):
        g, off = np.zeros(nx * nz), 0
        for o, k in zip(osh, n):
            g += orc.acou_adjoint(o, cfg["vp"], cfg["rho"], wav, r[off:off + k])
            off += k
        return g

    x, hist = _cgls(L, LT, d.astype(np.float64), 2, nx * nz)
    e = orc.rel_l2(img.ravel(), x)
    print(f"C2 full size, 3 shots, 2 CG iterations: image rel {e:.2e}; residuals gpu {list(res)} oracle {hist}")
    assert e <= TOL
    np.testing.assert_allclose(res, hist, rtol=TOL)
    from ds2d_amd import laplace_filter
    lap = laplace_filter(img, cfg["dx"])
    assert np.array_equal(lap, orc.laplace_order(img, cfg["dx"], prec="f"))
    lap64 = orc.laplace_order(x.reshape(nx, nz), cfg["dx"])
    print(f"Laplacian-filtered image vs the filter of the FP64 image: rel {orc.rel_l2(lap, lap64):.2e}")
    assert np.abs(lap).max() > 0


def test_c4_size_unet_loss_and_gradient_match_oracle():
    """The 4-level, 32-channel U-Net at the C4 training shape (batch 8 of
    512x256) against the U-Net oracle with the same BF16 rounding points:
    loss within 1e-3, every parameter block's gradient within 1e-2."""
    from ds2d_amd.lib import Unet
    levels, base, n, h, w = 4, 32, 8, 512, 256
    u = Unet(levels=levels, base=base, max_batch=n, max_h=h, max_w=w, seed=3)
    rng = np.random.default_rng(41)
    theta = U.init_params(levels, base, seed=42)
    theta += 0.02 * rng.standard_normal(theta.size)
    theta = theta.astype(np.float32)
    u.set_params(theta)
    x = rng.standard_normal((n, h, w)).astype(np.float32)
    t = rng.standard_normal((n, h, w)).astype(np.float32)
    loss, g = u.loss_grad(x, t)
    ref_loss, ref_g, _ = U.loss_grad(theta.astype(np.float64), x, t, levels, base, bf=True)
    eg = orc.rel_l2(g, ref_g)
    print(f"C4 U-Net loss {loss:.6e} vs {ref_loss:.6e}; gradient rel {eg:.2e}")
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    o = 0
    for name, ci, co, k in U.layers(levels, base):
        for part, cnt in (("w", (9 if k == 3 else 1) * ci * co), ("b", co)):
            e = orc.rel_l2(g[o:o + cnt], ref_g[o:o + cnt])
            assert e <= 1e-2, (name, part, e)
            o += cnt


def test_c4_held_out_model_learned_vs_plain():
    """C4 geometry (512x256, 16 shots, 1500 steps): a U-Net trained on 12
    (RTM image, 20-iteration CG image) pairs of seeds 1000-1011 preconditions
    LSRTM on held-out seed 2000.  Every accepted learned step lowers the
    residual by at least 1 %, the accepted steps come first, the residual
    never rises, and after 8 iterations the learned solver is ahead of plain
    CG.  If it fell back before the end, the tail equals plain CG started
    from its image at the switch."""
    from ds2d_amd import learned as LR
    X, T, _, s = LR.generate_pairs(range(1000, 1012), niter=20)
    u, loss, _ = LR.train_unet(X, T, steps=300, batch=8, lr=1e-3, seed=0)
    print(f"training loss {loss[:10].mean():.4f} -> {loss[-10:].mean():.4f}")
    prob = LR.acoustic_problem(2000)
    s.set_model(prob["vp"], None, prob["rho"])
    _, d, _ = s.born_forward(prob["shots"], prob["refl"])
    niter = 8
    _, rp, _ = s.lsrtm_cg(prob["shots"], d, niter)
    img, rl, st = s.lsrtm_learned(u, prob["shots"], d, niter)
    a = st["learned_steps"]
    print(f"held-out seed 2000: plain {rp[-1]:.4f}, learned {rl[-1]:.4f} ({a} learned steps)")
    assert 1 <= a <= niter
    assert all(rl[k + 1] <= 0.99 * rl[k] for k in range(a))
    assert all(rl[k + 1] <= rl[k] * (1 + 1e-6) for k in range(niter))
    assert rl[-1] < rp[-1]
    if a < niter:
        img_a, rl_a, st_a = s.lsrtm_learned(u, prob["shots"], d, a)
        assert st_a["learned_steps"] == a
        np.testing.assert_allclose(rl_a, rl[:a + 1], rtol=1e-6)
        _, rc, _ = s.lsrtm_cg(prob["shots"], d, niter - a, image0=img_a)
        print(f"fallback after {a} steps: learned tail {list(rl[a:])}, CG from its image {list(rc)}")
        np.testing.assert_allclose(rc, rl[a:], rtol=1e-4)


def test_c4_held_out_plain_cg_matches_fp64_oracle():
    """The plain solver of the C4 comparison on held-out seed 2000 at the
    full C4 geometry (512x256, random boundary 40, 1500 steps; shots 0, 5,
    10 and 15 of the 16): two CG iterations match CGLS on the FP64 oracle
    pair, image and residuals within 1e-5, so the plain GPU image the learned
    solver is compared with is the FP64 oracle's to that bar."""
    from ds2d_amd import learned as LR
    prob = LR.acoustic_problem(2000)
    shots = [prob["shots"][k] for k in (0, 5, 10, 15)]
    nx, nz, nt = prob["nx"], prob["nz"], prob["nt"]
    wav = ricker(nt, prob["f0"], prob["dt"])
    s = LR.make_solver(prob)
    s.set_model(prob["vp"], None, prob["rho"])
    _, d, _ = s.born_forward(shots, prob["refl"])
    img, res, _ = s.lsrtm_cg(shots, d, 2)
    osh = [orc.make_shot(x, nx, nz, prob["pml"], nt, prob["dx"], prob["dz"], prob["dt"]) for x in shots]
    n = [x.ntr * nt for x in shots]

    def L(m):
        return np.concatenate([orc.acou_forward(o, prob["vp"], prob["rho"], m.reshape(nx, nz), wav)[1] for o in osh])

    def LT(r):
        g, off = np.zeros(nx * nz), 0
        for o, k in zip(osh, n):
            g += orc.acou_adjoint(o, prob["vp"], prob["rho"], wav, r[off:off + k])
            off += k
        return g

    x, hist = _cgls(L, LT, d.astype(np.float64), 2, nx * nz)
    e = orc.rel_l2(img.ravel(), x)
    print(f"C4 held-out seed 2000, 4 shots, 2 CG iterations: image rel {e:.2e}; residuals gpu {list(res)} "
          f"oracle {hist}")
    assert e <= TOL
    np.testing.assert_allclose(res, hist, rtol=TOL)


def test_c5_window_pp_ps_cg_matches_fp64_oracle():
    """PP/PS CG on a 500x250 window of the C5 model (2 full-window shots,
    800 steps, random boundary 40) against CGLS on the FP64 oracle's elastic
    Born pair: image within 1e-5, residual history within 1e-5."""
    cfg = models.c5_combined(nshots=1)
    x0, z0, nx, nz, nt, pml, niter = 1750, 0, 500, 250, 800, 40, 2
    sl = (slice(x0, x0 + nx), slice(z0, z0 + nz))
    vp, vs, rho = (np.ascontiguousarray(cfg[k][sl]) for k in ("vp", "vs", "rho"))
    rpp, rps = (np.ascontiguousarray(cfg[k][sl]) for k in ("refl_pp", "refl_ps"))
    wav = ricker(nt, cfg["f0"], cfg["dt"])
    shots = [models.surface_shot(nx, 0, pml, 125 + 250 * i, 2, 1, shot_id=i, boundary=1) for i in range(2)]
    s = Solver(dx=cfg["dx"], dz=cfg["dz"], dt=cfg["dt"], nt=nt, allnx=nx, allnz=nz, pml=pml, nx_max=nx,
               wavelet=wav, physics=0)
    s.set_model(vp, vs, rho)
    dz, dx, _ = s.eborn_forward(shots, rpp, rps)
    gpp, gps, res, _ = s.elsrtm_cg(shots, dz, dx, niter)
    osh = [orc.make_shot(x, nx, nz, pml, nt, cfg["dx"], cfg["dz"], cfg["dt"]) for x in shots]
    n = [x.ntr * nt for x in shots]
    A, R = nx * nz, sum(n)

    def L(m):
        out = [orc.eborn_forward(o, vp, vs, rho, m[:A].reshape(nx, nz), m[A:].reshape(nx, nz), wav) for o in osh]
        return np.concatenate([z for z, _ in out] + [x for _, x in out])

    def LT(r):
        g, off = np.zeros(2 * A), 0
        for o, k in zip(osh, n):
            a_, b_ = orc.eborn_adjoint(o, vp, vs, rho, wav, r[off:off + k], r[R + off:R + off + k])
            g[:A] += a_
            g[A:] += b_
            off += k
        return g

    x, hist = _cgls(L, LT, np.concatenate([dz, dx]).astype(np.float64), niter, 2 * A)
    img = np.concatenate([gpp.ravel(), gps.ravel()])
    e = orc.rel_l2(img, x)
    print(f"C5 window: image rel {e:.2e}; residuals gpu {list(res)} oracle {hist}")
    assert e <= TOL
    np.testing.assert_allclose(res, hist, rtol=TOL)
