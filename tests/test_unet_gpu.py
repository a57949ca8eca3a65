"""GPU parity of the learned preconditioner's BF16 MFMA kernels and U-Net
against the FP64 oracle (oracle/unet_ref.py), training, and the gain of the
trained network inside LSRTM.

Tolerances: single conv kernels compare FP32 outputs against the oracle fed
the same BF16-rounded inputs and weights (only the FP32 accumulation order
differs): relative L2 <= 1e-5.  The whole network stores BF16 activations;
against the oracle with the same rounding points a rounding tie can flip one
BF16 ulp (2^-8), so gradients are checked at relative L2 <= 1e-2 per
parameter block and the loss at 1e-3.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import unet_ref as U  # noqa: E402

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def rnd(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


@pytest.mark.parametrize("c1,c2,up,cout", [(32, 0, 0, 32), (64, 0, 0, 64), (32, 64, 1, 32), (64, 128, 1, 64),
                                           (32, 32, 0, 96)])
def test_conv_forward_matches_oracle(c1, c2, up, cout):
    from ds2d_amd.lib import conv3x3_bf16
    n, h, w = 2, 16, 24
    x1 = U.bf16(rnd((n, h, w, c1), 1))
    x2 = U.bf16(rnd((n, h >> up, w >> up, c2), 2)) if c2 else None
    wt = U.bf16(rnd((9, cout, c1 + c2), 3) * 0.1)
    b = rnd((cout,), 4)
    y, _ = conv3x3_bf16(0, x1, wt, x2=x2, up=up, bias=b, relu=True)
    inp = x1 if not c2 else np.concatenate([x1, U.up2(x2) if up else x2], axis=-1)
    ref = np.maximum(U.conv3x3(inp, wt, b), 0)
    e = rel(y, ref)
    print(f"conv fwd c1={c1} c2={c2} up={up} cout={cout}: rel {e:.2e}")
    assert e <= 1e-5


@pytest.mark.parametrize("c1,c2,up,cout,h,w", [(32, 0, 0, 32, 16, 32), (64, 0, 0, 64, 8, 64), (32, 64, 1, 32, 16, 32),
                                               (64, 64, 0, 128, 8, 32), (128, 256, 1, 128, 16, 16),
                                               (32, 32, 0, 64, 4, 256), (32, 0, 0, 32, 8, 256),
                                               (32, 64, 1, 32, 4, 384), (64, 64, 1, 64, 8, 128), (32, 0, 0, 32, 8, 64)])
def test_conv_lds_kernel_matches_oracle(c1, c2, up, cout, h, w):
    """Shapes whose blocks are image rows (whole rows, or 4-row blocks of
    P / 4 columns on wider images) take the LDS-staged kernel (forward with
    skip/upsampled input, and the data gradient)."""
    from ds2d_amd.lib import conv3x3_bf16
    n = 2
    x1 = U.bf16(rnd((n, h, w, c1), 21))
    x2 = U.bf16(rnd((n, h >> up, w >> up, c2), 22)) if c2 else None
    wt = U.bf16(rnd((9, cout, c1 + c2), 23) * 0.1)
    b = rnd((cout,), 24)
    y, _ = conv3x3_bf16(0, x1, wt, x2=x2, up=up, bias=b, relu=True)
    inp = x1 if not c2 else np.concatenate([x1, U.up2(x2) if up else x2], axis=-1)
    e = rel(y, np.maximum(U.conv3x3(inp, wt, b), 0))
    dy = U.bf16(rnd((n, h, w, cout), 25))
    dx, _ = conv3x3_bf16(1, dy, wt)
    e2 = rel(dx, U.conv3x3_dgrad(dy, wt))
    print(f"conv (LDS) c1={c1} c2={c2} up={up} cout={cout} {h}x{w}: forward rel {e:.2e}, dgrad rel {e2:.2e}")
    assert e <= 1e-5 and e2 <= 1e-5


@pytest.mark.parametrize("cin,cout", [(32, 32), (64, 32), (96, 32), (192, 64), (64, 128)])
def test_conv_data_gradient_matches_oracle(cin, cout):
    from ds2d_amd.lib import conv3x3_bf16
    n, h, w = 2, 12, 20
    dy = U.bf16(rnd((n, h, w, cout), 5))
    wt = U.bf16(rnd((9, cout, cin), 6) * 0.1)
    dx, _ = conv3x3_bf16(1, dy, wt)
    ref = U.conv3x3_dgrad(dy, wt)
    e = rel(dx, ref)
    print(f"conv dgrad cin={cin} cout={cout}: rel {e:.2e}")
    assert e <= 1e-5


@pytest.mark.parametrize("c1,c2,up,cout,h,w", [(32, 0, 0, 32, 12, 20), (64, 0, 0, 64, 10, 70), (32, 64, 1, 32, 8, 16),
                                               (64, 64, 0, 128, 6, 8), (128, 256, 1, 128, 8, 8),
                                               (32, 0, 0, 64, 5, 33), (64, 0, 0, 32, 7, 130), (32, 0, 0, 32, 1, 150)])
def test_conv_weight_gradient_matches_oracle(c1, c2, up, cout, h, w):
    from ds2d_amd.lib import conv3x3_bf16
    n = 2
    x1 = U.bf16(rnd((n, h, w, c1), 7))
    x2 = U.bf16(rnd((n, h >> up, w >> up, c2), 8)) if c2 else None
    dy = U.bf16(rnd((n, h, w, cout), 9))
    wt = np.zeros((9, cout, c1 + c2), np.float32)
    dw, _ = conv3x3_bf16(2, x1, wt, x2=x2, up=up, dy=dy)
    inp = x1 if not c2 else np.concatenate([x1, U.up2(x2) if up else x2], axis=-1)
    ref = U.conv3x3_wgrad(inp, dy)
    e = rel(dw, ref)
    print(f"conv wgrad c1={c1} c2={c2} up={up} cout={cout} {h}x{w}: rel {e:.2e}")
    assert e <= 1e-5


@pytest.mark.parametrize("levels,base,n,h,w", [(4, 32, 2, 32, 16), (3, 32, 1, 24, 40), (2, 64, 2, 16, 16)])
def test_unet_loss_and_gradient_match_oracle(levels, base, n, h, w):
    from ds2d_amd.lib import Unet
    u = Unet(levels=levels, base=base, max_batch=n, max_h=h, max_w=w, seed=3)
    rng = np.random.default_rng(11)
    theta = U.init_params(levels, base, seed=12)
    theta += 0.02 * rng.standard_normal(theta.size)
    theta = theta.astype(np.float32)
    u.set_params(theta)
    x = rnd((n, h, w), 13)
    t = rnd((n, h, w), 14)
    loss, g = u.loss_grad(x, t)
    yg, _ = u.forward(x)
    ref_loss, ref_g, ref_y = U.loss_grad(theta.astype(np.float64), x, t, levels, base, bf=True)
    _, g64, y64 = U.loss_grad(theta.astype(np.float64), x, t, levels, base, bf=False)
    print(f"unet L{levels} C{base}: loss {loss:.6e} vs {ref_loss:.6e}; y rel {rel(yg, ref_y):.2e} "
          f"(vs FP64 net {rel(yg, y64):.2e}); grad rel {rel(g, ref_g):.2e} (vs FP64 net {rel(g, g64):.2e})")
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    assert rel(yg, ref_y) <= 1e-2
    o = 0
    for name, ci, co, k in U.layers(levels, base):
        for part, cnt in (("w", (9 if k == 3 else 1) * ci * co), ("b", co)):
            e = rel(g[o:o + cnt], ref_g[o:o + cnt])
            assert e <= 1e-2, (name, part, e)
            o += cnt


def test_unet_training_reduces_loss():
    from ds2d_amd.lib import Unet
    rng = np.random.default_rng(21)
    n, h, w = 16, 32, 32
    x = rng.standard_normal((n, h, w)).astype(np.float32)
    # target: a fixed 3x3 blur of the input (learnable by one conv)
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1)))
    t = sum(xp[:, 1 + a:1 + a + h, 1 + b:1 + b + w] for a in (-1, 0, 1) for b in (-1, 0, 1)) / 9.0
    u = Unet(levels=4, base=32, max_batch=8, max_h=h, max_w=w, seed=5)
    u.set_data(x, t.astype(np.float32))
    loss, st = u.train(200, batch=8, lr=1e-3, seed=1)
    print(f"training loss {loss[0]:.4f} -> {loss[-1]:.4f}, {st['ms_per_step']:.3f} ms/step, "
          f"{st['flops_per_step'] / st['ms_per_step'] / 1e9:.2f} TFLOP/s")
    assert np.all(np.isfinite(loss))
    assert loss[-20:].mean() < 0.2 * loss[:5].mean()


def test_learned_lsrtm_beats_plain_cg_on_held_out_models():
    """Pairs from the plain method (RTM -> 20-iteration CG image) on 24 small
    models train the U-Net; on 3 held-out models the learned solver reaches a
    lower data residual than plain CG after the same number of iterations, and
    an untrained network does not help (ablation)."""
    from ds2d_amd import learned as LR
    kw = dict(nx=96, nz=48, nshots=6, nt=700, pml=30, dt=0.0008)
    X, T, info, s = LR.generate_pairs(range(100, 124), niter=20, **kw)
    u, loss, st = LR.train_unet(X, T, steps=400, batch=8, lr=1e-3, seed=0)
    print(f"train loss {loss[:10].mean():.4f} -> {loss[-10:].mean():.4f} ({st['wall_s']:.1f} s)")
    niter = 6
    gains = []
    for seed in (900, 901, 902):
        prob = LR.acoustic_problem(seed, **kw)
        out = LR.compare(s, u, prob, niter)
        p, l = out["plain"]["resid"][-1], out["learned"]["resid"][-1]
        print(f"seed {seed}: plain {p:.4f} learned {l:.4f} (learned steps {out['learned']['learned_steps']}), "
              f"image err plain {out['plain']['image_rel_err']:.3f} learned {out['learned']['image_rel_err']:.3f}")
        gains.append(p / l)
    assert np.mean(gains) > 1.0
    # ablation: an untrained network gives no systematic gain
    from ds2d_amd.lib import Unet
    u0 = Unet(levels=4, base=32, max_batch=8, max_h=kw["nx"], max_w=kw["nz"], seed=99)
    prob = LR.acoustic_problem(900, **kw)
    out0 = LR.compare(s, u0, prob, niter)
    print(f"untrained: plain {out0['plain']['resid'][-1]:.4f} learned {out0['learned']['resid'][-1]:.4f}")
    assert out0["learned"]["resid"][-1] <= out0["plain"]["resid"][0]


def test_learned_solver_falls_back_at_once_with_a_zero_network():
    """A U-Net whose parameters are all zero returns z = 0 for every input, so
    the first learned direction has L p = 0 and cannot lower the residual:
    the solver must take the plain CGLS step from the start, count no
    learned step, and follow ds2d_lsrtm_cg's iterates."""
    from ds2d_amd import learned as LR
    from ds2d_amd.lib import Unet
    kw = dict(nx=96, nz=48, nshots=6, nt=700, pml=30, dt=0.0008)
    prob = LR.acoustic_problem(903, **kw)
    s = LR.make_solver(prob)
    s.set_model(prob["vp"], None, prob["rho"])
    _, d, _ = s.born_forward(prob["shots"], prob["refl"])
    u = Unet(levels=4, base=32, max_batch=1, max_h=kw["nx"], max_w=kw["nz"], seed=1)
    u.set_params(np.zeros(U.num_params(4, 32), np.float32))
    niter = 5
    img_p, rp, _ = s.lsrtm_cg(prob["shots"], d, niter)
    img_l, rl, st = s.lsrtm_learned(u, prob["shots"], d, niter)
    print(f"zero network: learned steps {st['learned_steps']}, residual {list(rl)} vs CG {list(rp)}")
    assert st["learned_steps"] == 0
    np.testing.assert_allclose(rl, rp, rtol=1e-5)
    assert rel(img_l, img_p) <= 1e-5
