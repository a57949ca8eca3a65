"""The reference-compat entry points a main.cpp-style caller binds, through
the C-ABI on the GPU, against the FP64 oracle: the per-shot sequence
ds2d_set_shot -> ds2d_shot_gradient -> ds2d_take_gradients, ds2d_forward,
the Born records alone (ds2d_born), the line search ELSRTM_CALCULATE_ALPHA
(ds2d_calculate_alpha) and the tile the forward autotuner picks.

Tolerance: relative L2 <= 1e-5 against FP64 (FP32 records and images).
"""
import numpy as np
import pytest

import oracle_lib as orc
from cases import DT, DX, DZ, elastic_model, full_window_shot, uneven_shots, wavelet
from ds2d_amd import Solver

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _solver(allnx, allnz, pml, nt, nx_max, wav, **kw):
    return Solver(dx=DX, dz=DZ, dt=DT, nt=nt, allnx=allnx, allnz=allnz, pml=pml, nx_max=nx_max, wavelet=wav, **kw)


def _osh(spec, allnx, allnz, pml, nt, **kw):
    return orc.make_shot(spec, allnx, allnz, pml, nt, DX, DZ, DT, **kw)


def _case(boundary, nt=240):
    allnx, allnz, pml = 56, 32, 20
    vp, vs, _ = elastic_model(allnx, allnz, 21)
    wav = wavelet(nt)
    rng = np.random.default_rng(22)
    img_pp = (0.05 * rng.standard_normal((allnx, allnz))).astype(np.float32)
    img_ps = (0.05 * rng.standard_normal((allnx, allnz))).astype(np.float32)
    shots = [full_window_shot(40, pml, 10, shot_id=0, boundary=boundary, ngx_min=0),
             full_window_shot(40, pml, 27, shot_id=1, boundary=boundary, ngx_min=16)]
    return allnx, allnz, pml, nt, vp, vs, wav, img_pp, img_ps, shots


def _observed(shots, allnx, allnz, pml, nt, vp, vs, wav):
    oz, ox = [], []
    for sh in shots:
        z, x = orc.elastic_forward(_osh(sh, allnx, allnz, pml, nt, rho_const=2000.0), vp * np.float32(1.03), vs,
                                   None, wav, prec="f")
        oz.append(z.astype(np.float32))
        ox.append(x.astype(np.float32))
    return oz, ox


@pytest.mark.parametrize("boundary", [0, 1])
def test_set_shot_shot_gradient_take_gradients_sequence(boundary):
    """main.cpp:994-1035 call order, one shot at a time on one handle: the
    gradients accumulate across shots and ds2d_take_gradients returns and
    zeroes them (imagebuffer_resettozero, FD2DGPU.cu:3478-3505)."""
    allnx, allnz, pml, nt, vp, vs, wav, img_pp, img_ps, shots = _case(boundary)
    oz, ox = _observed(shots, allnx, allnz, pml, nt, vp, vs, wav)
    s = _solver(allnx, allnz, pml, nt, 40, wav, rho_const=2000.0)
    s.set_model(vp, vs, None)
    s.set_image(img_pp, img_ps)
    mf = 0.0
    for sh, z, x in zip(shots, oz, ox):
        s.set_shot(sh, z, x)
        mf += s.shot_gradient(1)
    g = s.take_gradients()
    og = np.zeros(7 * allnx * allnz)
    omf = 0.0
    for sh, z, x in zip(shots, oz, ox):
        og, m1 = orc.elsrtm_shot(_osh(sh, allnx, allnz, pml, nt, rho_const=2000.0), vp, vs, None, wav, img_pp, img_ps,
                                 z, x, 1, sh.ngx_min, grads7=og)
        omf += m1
    og = og.reshape(7, allnx, allnz)
    for k, name in enumerate(["pp", "ps", "illum", "b_pp", "b_ps", "b2_pp", "b2_ps"]):
        e = orc.rel_l2(g[k], og[k])
        print(f"boundary={boundary} {name}: rel {e:.2e}")
        assert np.abs(og[k]).max() > 0 and e <= TOL, name
    # The misfit is -1/(2 ntr) times a sum of normalised correlations, each in
    # [-1, 1]: its scale per shot is 1, whatever cancellation leaves of the sum.
    print(f"misfit {mf:.9f} vs {omf:.9f}")
    assert abs(mf - omf) <= TOL * max(1.0, abs(omf))
    # take_gradients zeroed the accumulators
    g2 = s.take_gradients()
    assert all(np.abs(a).max() == 0 for a in g2)


def test_forward_of_current_shot():
    """ds2d_forward (source, born_modeling(1,0), receivers) on the shot set by
    ds2d_set_shot."""
    allnx, allnz, pml, nt = 64, 40, 20, 350
    vp, vs, rho = elastic_model(allnx, allnz, 23)
    wav = wavelet(nt)
    sh = full_window_shot(48, pml, 20, ngx_min=9, stride=2)
    s = _solver(allnx, allnz, pml, nt, 48, wav)
    s.set_model(vp, vs, rho)
    s.set_shot(sh)
    rz, rx = s.forward()
    oz, ox = orc.elastic_forward(_osh(sh, allnx, allnz, pml, nt), vp, vs, rho, wav)
    print(f"forward rel_z={orc.rel_l2(rz, oz):.2e} rel_x={orc.rel_l2(rx, ox):.2e}")
    assert np.abs(oz).max() > 0
    assert orc.rel_l2(rz, oz) <= TOL and orc.rel_l2(rx, ox) <= TOL


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("boundary", [0, 1])
def test_born_records(order, boundary):
    """ds2d_born: Lm (born_modeling mode 1, record_scatter) and Lm + Lm2
    (modes 0/3/4/5 + scatter_add) for a two-shot batch."""
    allnx, allnz, pml, nt, vp, vs, wav, img_pp, img_ps, shots = _case(boundary)
    s = _solver(allnx, allnz, pml, nt, 40, wav, rho_const=2000.0)
    s.set_model(vp, vs, None)
    s.set_image(img_pp, img_ps)
    rz, rx = s.born(shots, order)
    off = 0
    for sh in shots:
        n = sh.ntr * nt
        oz, ox = orc.elsrtm_born(_osh(sh, allnx, allnz, pml, nt, rho_const=2000.0), vp, vs, None, wav, img_pp, img_ps,
                                 sh.ngx_min, order)
        ez, ex = orc.rel_l2(rz[off:off + n], oz), orc.rel_l2(rx[off:off + n], ox)
        print(f"order={order} boundary={boundary} shot {sh.shot_id}: rel_z={ez:.2e} rel_x={ex:.2e}")
        assert np.abs(oz).max() > 0 and ez <= TOL and ex <= TOL
        off += n


def test_calculate_alpha_matches_oracle_sums():
    """ELSRTM_CALCULATE_ALPHA (FD2DGPU.cu:4885-5035): fenzi = <Lg, Lm - d>,
    fenmu = <Lg, Lg> over both shots; the handle's image is left as it was."""
    allnx, allnz, pml, nt, vp, vs, wav, img_pp, img_ps, shots = _case(0)
    oz, ox = _observed(shots, allnx, allnz, pml, nt, vp, vs, wav)
    d_z, d_x = np.concatenate(oz), np.concatenate(ox)
    rng = np.random.default_rng(24)
    g_pp = (0.02 * rng.standard_normal((allnx, allnz))).astype(np.float32)
    g_ps = (0.02 * rng.standard_normal((allnx, allnz))).astype(np.float32)
    s = _solver(allnx, allnz, pml, nt, 40, wav, rho_const=2000.0)
    s.set_model(vp, vs, None)
    s.set_image(img_pp, img_ps)
    num, den = s.calculate_alpha(shots, g_pp, g_ps, d_z, d_x)
    lm = [orc.elsrtm_born(_osh(sh, allnx, allnz, pml, nt, rho_const=2000.0), vp, vs, None, wav, img_pp, img_ps,
                          sh.ngx_min, 1) for sh in shots]
    lg = [orc.elsrtm_born(_osh(sh, allnx, allnz, pml, nt, rho_const=2000.0), vp, vs, None, wav, g_pp, g_ps,
                          sh.ngx_min, 1) for sh in shots]
    rz = np.concatenate([a[0] for a in lm]) - d_z
    rx = np.concatenate([a[1] for a in lm]) - d_x
    qz, qx = np.concatenate([a[0] for a in lg]), np.concatenate([a[1] for a in lg])
    onum, oden = float(qz @ rz + qx @ rx), float(qz @ qz + qx @ qx)
    print(f"num {num:.6e} vs {onum:.6e}, den {den:.6e} vs {oden:.6e}")
    assert abs(num - onum) <= TOL * abs(onum) and abs(den - oden) <= TOL * oden
    # the image m is restored: Lm again equals the oracle's
    rz2, _ = s.born(shots, 1)
    assert orc.rel_l2(rz2, np.concatenate([a[0] for a in lm])) <= TOL


def test_autotuned_tile_matches_oracle():
    """ds2d_autotune_forward picks a tile (or the two-kernel step) on this
    batch; the forward run on what it picked matches the oracle."""
    allnx, allnz, pml, nt = 160, 40, 20, 300
    vp, vs, rho = elastic_model(allnx, allnz, 25)
    wav = wavelet(nt)
    shots = uneven_shots(allnx, pml, widths=[60, 100, 37, 160], starts=[0, 45, 120, 0])
    s = _solver(allnx, allnz, pml, nt, 160, wav)
    s.set_model(vp, vs, rho)
    w, t, d, ms = s.autotune(shots, nsteps=30)
    print(f"autotuner picked width={w} threads={t} depth={d} ({ms:.3f} ms / 30 steps)")
    assert ms > 0 and (w == 0 or (t in (256, 512, 704, 768) and d in (2, 4)))
    rz, rx, _ = s.batch_forward(shots, mode=2)
    off = 0
    for sh in shots:
        n = sh.ntr * nt
        oz, ox = orc.elastic_forward(_osh(sh, allnx, allnz, pml, nt), vp, vs, rho, wav)
        assert orc.rel_l2(rz[off:off + n], oz) <= TOL and orc.rel_l2(rx[off:off + n], ox) <= TOL
        off += n


def test_failed_shot_is_left_out_and_reported():
    """A NaN in shot 1's observed records makes its NCC residual and adjoint
    images non-finite.  The batched gradient leaves that shot out of the
    misfit and the images, completes shot 0 (images equal to shot 0's alone),
    returns DS2D_ESHOT and reports the shot through ds2d_shot_status.  The
    failure is forced from the host: every kernel runs to its normal end."""
    from ds2d_amd.lib import DS2DError, ESHOT
    allnx, allnz, pml, nt, vp, vs, wav, img_pp, img_ps, shots = _case(0, nt=200)
    oz, ox = _observed(shots, allnx, allnz, pml, nt, vp, vs, wav)
    bad_z = oz[1].copy()
    bad_z[5] = np.nan
    s = _solver(allnx, allnz, pml, nt, 40, wav, rho_const=2000.0)
    s.set_model(vp, vs, None)
    s.set_image(img_pp, img_ps)
    mf0 = s.batch_gradient(shots[:1], oz[0], ox[0], 1)
    ref = s.take_gradients()
    assert s.shot_status(1) == [0]
    with pytest.raises(DS2DError) as e:
        s.batch_gradient(shots, np.concatenate([oz[0], bad_z]), np.concatenate([ox[0], ox[1]]), 1)
    assert e.value.code == ESHOT
    assert s.shot_status(2) == [0, 1]
    got = s.take_gradients()
    for k in range(7):
        assert np.all(np.isfinite(got[k])) and np.array_equal(got[k], ref[k]), k
    print(f"shot 0 alone misfit {mf0:.6f}; failed shot left out, images equal shot 0's")
