"""Plain-L2 elastic Born pair (PP/PS) on the GPU vs the FP64 CPU oracle, and
PP/PS CG LSRTM vs CGLS driven by the oracle's pair.

Tolerance: relative L2 <= 1e-5 (FP32 GPU vs FP64 oracle) on records,
gradients (resident background history or background rebuilt by reversal)
and CG images; CG residual histories to rtol 1e-5.  The adjoint identity
<L m, r> = <m, L* r> of the GPU pair itself is checked to 1e-4 of the FP64
dot products of FP32 vectors.
"""
import numpy as np
import pytest

import oracle_lib as orc
from cases import DT, DX, DZ, elastic_model, wavelet
from ds2d_amd import Solver
from ds2d_amd.lib import ShotSpec

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _case(nt=300):
    """Two uneven windows on a 64 x 40 layered elastic model, random boundary 20."""
    allnx, allnz, pml = 64, 40, 20
    vp, vs, rho = elastic_model(allnx, allnz, 23)
    wav = wavelet(nt)
    shots = []
    for i, (x0, w) in enumerate(((0, 48), (22, 42))):
        shots.append(ShotSpec(x0, w, pml + w // 2, pml + 2, pml + 1, pml + np.arange(0, w, 2, dtype=np.int32),
                              shot_id=i, boundary=1))
    rng = np.random.default_rng(31)
    pp = (0.05 * rng.standard_normal((allnx, allnz))).astype(np.float32)
    ps = (0.03 * rng.standard_normal((allnx, allnz))).astype(np.float32)
    return allnx, allnz, pml, nt, vp, vs, rho, wav, shots, pp, ps


def _solver(allnx, allnz, pml, nt, wav, **kw):
    return Solver(dx=DX, dz=DZ, dt=DT, nt=nt, allnx=allnx, allnz=allnz, pml=pml, nx_max=allnx, wavelet=wav,
                  physics=0, **kw)


def _oshots(shots, allnx, allnz, pml, nt):
    return [orc.make_shot(sh, allnx, allnz, pml, nt, DX, DZ, DT) for sh in shots]


@pytest.mark.parametrize("bg_mode", [1, 2])
def test_eborn_forward_and_adjoint_match_oracle(bg_mode):
    allnx, allnz, pml, nt, vp, vs, rho, wav, shots, pp, ps = _case()
    s = _solver(allnx, allnz, pml, nt, wav, bg_mode=bg_mode)
    s.set_model(vp, vs, rho)
    rz, rx, _ = s.eborn_forward(shots, pp, ps)
    rng = np.random.default_rng(5)
    res_z = rng.standard_normal(rz.size).astype(np.float32)
    res_x = rng.standard_normal(rx.size).astype(np.float32)
    gpp, gps, _ = s.eborn_adjoint(shots, res_z, res_x)
    ogpp, ogps = np.zeros(allnx * allnz), np.zeros(allnx * allnz)
    off = 0
    for sh, o in zip(shots, _oshots(shots, allnx, allnz, pml, nt)):
        n = sh.ntr * nt
        oz, ox = orc.eborn_forward(o, vp, vs, rho, pp, ps, wav)
        ez, ex = orc.rel_l2(rz[off:off + n], oz), orc.rel_l2(rx[off:off + n], ox)
        print(f"shot {sh.shot_id}: records rel z {ez:.2e} x {ex:.2e}")
        assert np.abs(oz).max() > 0 and ez <= TOL and ex <= TOL
        a, b = orc.eborn_adjoint(o, vp, vs, rho, wav, res_z[off:off + n], res_x[off:off + n])
        ogpp += a
        ogps += b
        off += n
    epp, eps = orc.rel_l2(gpp.ravel(), ogpp), orc.rel_l2(gps.ravel(), ogps)
    print(f"bg_mode={bg_mode} gradient rel pp {epp:.2e} ps {eps:.2e}")
    assert epp <= TOL and eps <= TOL
    # adjoint consistency of the GPU pair itself
    lhs = float(rz.astype(np.float64) @ res_z + rx.astype(np.float64) @ res_x)
    rhs = float(pp.ravel().astype(np.float64) @ gpp.ravel() + ps.ravel().astype(np.float64) @ gps.ravel())
    assert abs(lhs - rhs) <= 1e-4 * abs(lhs)


def test_elsrtm_cg_matches_oracle_cg():
    """Three PP/PS CG iterations on the GPU vs CGLS on the FP64 oracle pair."""
    allnx, allnz, pml, nt, vp, vs, rho, wav, shots, pp, ps = _case(nt=260)
    s = _solver(allnx, allnz, pml, nt, wav, bg_mode=1)
    s.set_model(vp, vs, rho)
    dz, dx, _ = s.eborn_forward(shots, pp, ps)
    ipp, ips, res, _ = s.elsrtm_cg(shots, dz, dx, 3)
    oshots = _oshots(shots, allnx, allnz, pml, nt)
    n = [sh.ntr * nt for sh in shots]
    A = allnx * allnz

    def L(m):
        zs, xs = [], []
        for o in oshots:
            a, b = orc.eborn_forward(o, vp, vs, rho, m[:A].reshape(allnx, allnz), m[A:].reshape(allnx, allnz), wav)
            zs.append(a)
            xs.append(b)
        return np.concatenate(zs + xs)

    def LT(r):
        R = sum(n)
        out, off = np.zeros(2 * A), 0
        for o, k in zip(oshots, n):
            a, b = orc.eborn_adjoint(o, vp, vs, rho, wav, r[off:off + k], r[R + off:R + off + k])
            out[:A] += a
            out[A:] += b
            off += k
        return out

    d = np.concatenate([dz, dx]).astype(np.float64)
    x = np.zeros(2 * A)
    r = d.copy()
    sv = LT(r)
    p = sv.copy()
    gamma = float(sv @ sv)
    hist = [1.0]
    for _ in range(3):
        q = L(p)
        alpha = gamma / float(q @ q)
        x += alpha * p
        r -= alpha * q
        hist.append(float(np.linalg.norm(r) / np.linalg.norm(d)))
        sv = LT(r)
        gnew = float(sv @ sv)
        p = sv + (gnew / gamma) * p
        gamma = gnew
    print("GPU residuals", res, "oracle CGLS", hist)
    assert res[0] == pytest.approx(1.0, abs=1e-6)
    assert all(res[k + 1] < res[k] for k in range(3))
    np.testing.assert_allclose(res, hist, rtol=1e-5)
    img = np.concatenate([ipp.ravel(), ips.ravel()])
    e = orc.rel_l2(img, x)
    print(f"PP/PS CG image rel {e:.2e}")
    assert e <= TOL


def test_elsrtm_learned_trained_network_gain():
    """Pairs from the plain elastic method (PP and PS gradient -> 20-iteration
    CG image) on 12 small models train the U-Net; on 2 held-out models the
    learned PP/PS solver is compared with plain CG after the same number of
    iterations.  The CGLS fallback bounds the loss; the gain is printed."""
    from ds2d_amd import learned as LR
    kw = dict(nx=96, nz=48, nshots=6, nt=700, pml=30, dt=0.0008)
    X, T, info, s = LR.generate_elastic_pairs(range(100, 112), niter=20, **kw)
    assert X.shape == (24, 96, 48)
    u, loss, st = LR.train_unet(X, T, steps=400, batch=8, lr=1e-3, seed=0)
    print(f"elastic train loss {loss[:10].mean():.4f} -> {loss[-10:].mean():.4f} ({st['wall_s']:.1f} s)")
    assert loss[-10:].mean() < loss[:10].mean()
    niter = 6
    gains = []
    for seed in (950, 951):
        prob = LR.elastic_problem(seed, **kw)
        out = LR.compare_elastic(s, u, prob, niter)
        p, l = out["plain"]["resid"][-1], out["learned"]["resid"][-1]
        print(f"seed {seed}: plain {p:.4f} learned {l:.4f} (learned steps {out['learned']['learned_steps']}), "
              f"image err plain {out['plain']['image_rel_err']:.3f} learned {out['learned']['image_rel_err']:.3f}")
        r = out["learned"]["resid"]
        assert all(r[k + 1] <= r[k] * (1 + 1e-6) for k in range(niter))
        assert l <= 1.05 * p
        gains.append(p / l)
    print(f"elastic learned gain (plain / learned residual after {niter} iterations): {np.mean(gains):.3f}")


def test_eborn_rejects_pml_shots_and_acoustic_handles():
    """The plain elastic pair needs the random boundary and an elastic handle;
    both errors come back as status codes with a message."""
    from ds2d_amd.lib import DS2DError
    allnx, allnz, pml, nt, vp, vs, rho, wav, shots, pp, ps = _case(nt=50)
    s = _solver(allnx, allnz, pml, nt, wav)
    s.set_model(vp, vs, rho)
    bad = [ShotSpec(0, 48, pml + 24, pml + 2, pml + 1, pml + np.arange(48, dtype=np.int32), boundary=0)]
    with pytest.raises(DS2DError, match="random boundary"):
        s.eborn_forward(bad, pp, ps)
    with pytest.raises(DS2DError, match="random boundary"):
        s.elsrtm_cg(bad, np.zeros(48 * nt, np.float32), np.zeros(48 * nt, np.float32), 1)
    a = Solver(dx=DX, dz=DZ, dt=DT, nt=nt, allnx=allnx, allnz=allnz, pml=pml, nx_max=allnx, wavelet=wav, physics=1)
    a.set_model(vp, None, rho)
    with pytest.raises(DS2DError, match="elastic"):
        a.eborn_forward(shots, pp, ps)
    # zero data: CG refuses instead of dividing by zero
    with pytest.raises(DS2DError, match="all zero"):
        s.elsrtm_cg(shots, np.zeros(sum(x.ntr for x in shots) * nt, np.float32),
                    np.zeros(sum(x.ntr for x in shots) * nt, np.float32), 1)


def test_eborn_one_shot_per_launch_matches_batched():
    """ds2d_set_batch(1) runs the shots one time loop after another with the
    same kernels: records and gradients equal the batched run bit for bit."""
    allnx, allnz, pml, nt, vp, vs, rho, wav, shots, pp, ps = _case(nt=200)
    out = []
    for per_launch in (0, 1):
        s = _solver(allnx, allnz, pml, nt, wav, bg_mode=1)
        s.set_model(vp, vs, rho)
        s.set_batch(per_launch)
        rz, rx, _ = s.eborn_forward(shots, pp, ps)
        gpp, gps, _ = s.eborn_adjoint(shots, rz, rx)
        out.append((rz, rx, gpp, gps))
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def test_eborn_line_search_recovers_known_step():
    """ELSRTM_CALCULATE_ALPHA's sums: with d = L m - c L g, the ratio
    <L g, L m - d> / <L g, L g> is c."""
    allnx, allnz, pml, nt, vp, vs, rho, wav, shots, pp, ps = _case(nt=200)
    s = _solver(allnx, allnz, pml, nt, wav)
    s.set_model(vp, vs, rho)
    rng = np.random.default_rng(3)
    gpp = (0.02 * rng.standard_normal((allnx, allnz))).astype(np.float32)
    gps = (0.02 * rng.standard_normal((allnx, allnz))).astype(np.float32)
    mz, mx, _ = s.eborn_forward(shots, pp, ps)
    gz, gx, _ = s.eborn_forward(shots, gpp, gps)
    c = 0.37
    num, den = s.eborn_line_search(shots, pp, ps, gpp, gps, mz - c * gz, mx - c * gx)
    assert den > 0
    assert abs(num / den - c) <= 1e-4 * c


@pytest.mark.parametrize("bg_mode", [1, 2])
@pytest.mark.parametrize("threads,width", [(256, 64), (512, 24)])
def test_fused_born_step_matches_the_two_kernels(bg_mode, threads, width):
    """Three uneven windows (one of 12 columns) on an 80 x 300 model, so a
    256-lane block covers the 340 padded rows in two row blocks; an odd step
    count, so the fused pass ends in its second state and copies it back.
    The records of the Born forward map and the PP/PS CG images (forward
    and adjoint passes; with bg_mode = 2 the adjoint rebuilds the background
    from the forward pass's final state) of the fused steps agree with those
    of the per-cell kernels (k_eb_v + k_eb_s, k_eb_adj_v + k_eb_adj_s) to FP32
    rounding (relative L2 <= 1e-6: the expressions are the same, but the
    compiler may contract different products into FMAs in the two codes)."""
    allnx, allnz, pml, nt = 80, 300, 20, 151
    vp, vs, rho = elastic_model(allnx, allnz, 41)
    wav = wavelet(nt)
    shots = [ShotSpec(0, 60, pml + 30, pml + 2, pml + 1, pml + np.arange(0, 60, 3, dtype=np.int32), 0, 1),
             ShotSpec(15, 50, pml + 7, pml + 223, pml + 221, pml + np.arange(50, dtype=np.int32), 1, 1),
             ShotSpec(64, 12, pml + 5, pml + 3, pml + 2, pml + np.arange(12, dtype=np.int32), 2, 1)]
    rng = np.random.default_rng(8)
    pp = (0.05 * rng.standard_normal((allnx, allnz))).astype(np.float32)
    ps = (0.03 * rng.standard_normal((allnx, allnz))).astype(np.float32)
    out = {}
    for fused in (False, True):
        s = _solver(allnx, allnz, pml, nt, wav, bg_mode=bg_mode)
        s.set_model(vp, vs, rho)
        s.set_fused_born(fused, width, threads)
        rz, rx, _ = s.eborn_forward(shots, pp, ps)
        ipp, ips, hist, _ = s.elsrtm_cg(shots, rz, rx, 2)
        out[fused] = (rz, rx, ipp, ips, hist)
    for name, a, b in zip(("rec z", "rec x", "PP", "PS", "residuals"), out[False], out[True]):
        e = orc.rel_l2(np.ravel(b), np.ravel(a))
        print(f"{name}: rel {e:.2e}, identical {np.array_equal(a, b)}")
        assert e <= 1e-6, name
    assert np.abs(out[True][0]).max() > 0
