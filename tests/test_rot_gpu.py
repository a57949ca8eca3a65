"""P/S-decoupled (theta/omega rotation) modelling and RTM, and the elastic
RTM with decomposed (vector P/S, Poynting up/down) imaging, on the GPU vs
the FP64 CPU oracle.

Tolerance: relative L2 <= 1e-5 on records and images.  The oracle's own FP32
restatements sit near 1e-6 on these cases; the rotation RTM runs both
wavefields in FP64 on the GPU: its image divides by the source's own energy,
which magnifies rounding wherever the source field is small (an all-FP32
restatement is 0.18 off FP64 at the C1 size).  The oracle steps with the
FP32-rounded dx, dz, dt the library holds (make_shot rounds them): at the C1
size the FP64 image moves by 1.5e-4 between dt = 0.0008 and its FP32 value."""
import numpy as np
import pytest

import oracle_lib as orc
from cases import DT, DX, DZ, elastic_model, wavelet
from ds2d_amd import Solver
from ds2d_amd.lib import DS2DError, ShotSpec

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _case(nt, boundaries=(0, 1)):
    """Two uneven windows on a 72 x 40 layered elastic model, frame 20."""
    allnx, allnz, pml = 72, 40, 20
    vp, vs, rho = elastic_model(allnx, allnz, 29)
    wav = wavelet(nt)
    shots = []
    for i, ((x0, w), bnd) in enumerate(zip(((0, 52), (24, 48)), boundaries)):
        shots.append(ShotSpec(x0, w, pml + w // 2, pml + 3, pml + 1, pml + np.arange(0, w, 2, dtype=np.int32),
                              shot_id=i, boundary=bnd))
    return allnx, allnz, pml, nt, vp, vs, rho, wav, shots


def _solver(allnx, allnz, pml, nt, wav):
    return Solver(dx=DX, dz=DZ, dt=DT, nt=nt, allnx=allnx, allnz=allnz, pml=pml, nx_max=allnx, wavelet=wav)


def _oshot(sh, allnx, allnz, pml, nt):
    return orc.make_shot(sh, allnx, allnz, pml, nt, DX, DZ, DT)


def test_rot_forward_matches_oracle():
    """A PML window and a random-boundary window in one batch: vz, vx records."""
    allnx, allnz, pml, nt, vp, vs, rho, wav, shots = _case(400)
    s = _solver(allnx, allnz, pml, nt, wav)
    s.set_model(vp, vs, rho)
    rz, rx, st = s.rot_forward(shots)
    assert st["kernel_avg_ms"] > 0
    off = 0
    for sh in shots:
        n = sh.ntr * nt
        oz, ox = orc.rot_forward(_oshot(sh, allnx, allnz, pml, nt), vp, vs, wav)
        ez, ex = orc.rel_l2(rz[off:off + n], oz), orc.rel_l2(rx[off:off + n], ox)
        print(f"shot {sh.shot_id} (boundary {sh.boundary}): records rel z {ez:.2e} x {ex:.2e}")
        assert np.abs(oz).max() > 0
        assert ez <= TOL and ex <= TOL
        off += n


def test_rot_rtm_matches_oracle():
    """Two random-boundary windows: PP and PS images of the rotation RTM on
    the oracle's own records, stacked over both shots."""
    allnx, allnz, pml, nt, vp, vs, rho, wav, shots = _case(360, boundaries=(1, 1))
    s = _solver(allnx, allnz, pml, nt, wav)
    s.set_model(vp, vs, rho)
    recs = [orc.rot_forward(_oshot(sh, allnx, allnz, pml, nt), vp, vs, wav) for sh in shots]
    rz = np.concatenate([r[0] for r in recs]).astype(np.float32)
    rx = np.concatenate([r[1] for r in recs]).astype(np.float32)
    pp, ps = s.rot_rtm(shots, rz, rx)
    opp, ops = None, None
    off = 0
    for sh in shots:
        n = sh.ntr * nt
        opp, ops = orc.rot_rtm(_oshot(sh, allnx, allnz, pml, nt), vp, vs, wav, rz[off:off + n], rx[off:off + n],
                               opp, ops)
        off += n
    epp, eps_ = orc.rel_l2(pp.ravel(), opp), orc.rel_l2(ps.ravel(), ops)
    print(f"rotation RTM: PP rel {epp:.2e}, PS rel {eps_:.2e}")
    assert np.abs(opp).max() > 0 and np.abs(ops).max() > 0
    assert epp <= TOL and eps_ <= TOL


def test_rot_rtm_rejects_damped_windows():
    allnx, allnz, pml, nt, vp, vs, rho, wav, shots = _case(50)
    s = _solver(allnx, allnz, pml, nt, wav)
    s.set_model(vp, vs, rho)
    z = np.zeros(sum(sh.ntr for sh in shots) * nt, np.float32)
    with pytest.raises(DS2DError, match="random boundary"):
        s.rot_rtm(shots, z, z)


@pytest.mark.parametrize("kind", range(7))
def test_decomp_rtm_matches_oracle(kind):
    """A PML window and a random-boundary window: the decomposed image of
    each kind (0 Poynting PP, 1-3 vector PP/PS/SS, 4-6 scalar PP/PS/SS) on
    the oracle's own elastic records."""
    allnx, allnz, pml, nt, vp, vs, rho, wav, shots = _case(320)
    s = _solver(allnx, allnz, pml, nt, wav)
    s.set_model(vp, vs, rho)
    oshots = [_oshot(sh, allnx, allnz, pml, nt) for sh in shots]
    rz = np.concatenate([orc.elastic_forward(o, vp, vs, rho, wav)[0] for o in oshots]).astype(np.float32)
    img = s.decomp_rtm(shots, kind, rz)
    oimg = None
    off = 0
    for sh, o in zip(shots, oshots):
        n = sh.ntr * nt
        oimg = orc.decomp_rtm(o, vp, vs, rho, wav, kind, rz[off:off + n], oimg)
        off += n
    e = orc.rel_l2(img.ravel(), oimg)
    print(f"decomposed RTM kind {kind}: rel {e:.2e}")
    assert np.abs(oimg).max() > 0
    assert e <= TOL


def test_rotation_and_decomposed_rtm_at_c1_size():
    """The C1 geometry at full size (251x101, 20-cell frame, 2500 steps,
    dt 0.8 ms) with the C1 model: rotation forward records of a PML shot,
    the rotation RTM of a random-boundary shot, and the Poynting (kind 0) and
    PS vector (kind 2) decomposed images of a PML shot, each within 1e-5 of
    the FP64 oracle."""
    from ds2d_amd import models
    from ds2d_amd.lib import ricker
    cfg = models.c1_elastic_forward()
    nx, nz, pml, nt = cfg["nx"], cfg["nz"], cfg["pml"], cfg["nt"]
    vp, vs, rho = cfg["vp"], cfg["vs"], cfg["rho"]
    wav = ricker(nt, cfg["f0"], cfg["dt"])
    rec = pml + np.arange(nx, dtype=np.int32)
    pml_shot = ShotSpec(0, nx, pml + 125, pml + 2, pml, rec, shot_id=0, boundary=0)
    rb_shot = ShotSpec(0, nx, pml + 60, pml + 2, pml, rec, shot_id=1, boundary=1)
    s = Solver(dx=cfg["dx"], dz=cfg["dz"], dt=cfg["dt"], nt=nt, allnx=nx, allnz=nz, pml=pml, nx_max=nx, wavelet=wav)
    s.set_model(vp, vs, rho)

    def oshot(sh):
        return orc.make_shot(sh, nx, nz, pml, nt, cfg["dx"], cfg["dz"], cfg["dt"])

    rz, rx, _ = s.rot_forward([pml_shot])
    oz, ox = orc.rot_forward(oshot(pml_shot), vp, vs, wav)
    ez, ex = orc.rel_l2(rz, oz), orc.rel_l2(rx, ox)
    print(f"C1 size rotation forward: records rel z {ez:.2e} x {ex:.2e}")
    assert ez <= TOL and ex <= TOL
    bz, bx = orc.rot_forward(oshot(rb_shot), vp, vs, wav)
    pp, ps = s.rot_rtm([rb_shot], bz.astype(np.float32), bx.astype(np.float32))
    opp, ops = orc.rot_rtm(oshot(rb_shot), vp, vs, wav, bz.astype(np.float32), bx.astype(np.float32))
    epp, eps_ = orc.rel_l2(pp.ravel(), opp), orc.rel_l2(ps.ravel(), ops)
    print(f"C1 size rotation RTM: PP rel {epp:.2e}, PS rel {eps_:.2e}")
    assert epp <= TOL and eps_ <= TOL
    ez32 = orc.elastic_forward(oshot(pml_shot), vp, vs, rho, wav)[0].astype(np.float32)
    for kind in (0, 2):
        img = s.decomp_rtm([pml_shot], kind, ez32)
        oimg = orc.decomp_rtm(oshot(pml_shot), vp, vs, rho, wav, kind, ez32)
        e = orc.rel_l2(img.ravel(), oimg)
        print(f"C1 size decomposed RTM kind {kind}: rel {e:.2e}")
        assert np.abs(oimg).max() > 0 and e <= TOL


def test_fused_rotation_step_matches_per_cell_kernels():
    """The fused rotation step (unsplit theta/omega outside the damping
    strips) against k_rot_v + k_rot_p on a PML and a random-boundary window
    plus a 12-column window, over rows that need two 256-lane row blocks:
    records agree to FP32 rounding of the two arithmetic orders."""
    allnx, allnz, pml, nt = 90, 260, 20, 300
    vp, vs, rho = elastic_model(allnx, allnz, 37)
    wav = wavelet(nt)
    shots = [ShotSpec(0, 60, pml + 30, pml + 3, pml + 1, pml + np.arange(0, 60, 2, dtype=np.int32), 0, 0),
             ShotSpec(30, 60, pml + 20, pml + 228, pml + 226, pml + np.arange(60, dtype=np.int32), 1, 1),
             ShotSpec(70, 12, pml + 6, pml + 4, pml + 2, pml + np.arange(12, dtype=np.int32), 2, 0)]
    out = {}
    for fused in (False, True):
        s = _solver(allnx, allnz, pml, nt, wav)
        s.set_model(vp, vs, rho)
        s.set_fused_rotation(fused)
        rz, rx, _ = s.rot_forward(shots)
        out[fused] = (rz, rx)
    ez, ex = orc.rel_l2(out[True][0], out[False][0]), orc.rel_l2(out[True][1], out[False][1])
    print(f"fused vs per-cell rotation forward: rel z {ez:.2e} x {ex:.2e}")
    assert np.abs(out[False][0]).max() > 0
    assert ez <= 2e-6 and ex <= 2e-6
    off = 0
    for sh in shots:
        n = sh.ntr * nt
        oz, ox = orc.rot_forward(_oshot(sh, allnx, allnz, pml, nt), vp, vs, wav)
        ez, ex = orc.rel_l2(out[True][0][off:off + n], oz), orc.rel_l2(out[True][1][off:off + n], ox)
        print(f"shot {sh.shot_id} (boundary {sh.boundary}, {sh.nx} columns): fused vs oracle rel z {ez:.2e} x {ex:.2e}")
        assert ez <= TOL and ex <= TOL
        off += n
