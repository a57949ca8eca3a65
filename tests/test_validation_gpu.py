"""Argument and state checks of the C-ABI (each returns DS2D_EINVAL or
DS2D_ESTATE with a message, before any device work), and the edge cases of
a batch: windows that touch the model's edges, a shot with no receivers, and
a window as narrow as the update region allows."""
import numpy as np
import pytest

import oracle_lib as orc
from cases import DT, DX, DZ, elastic_model, wavelet
from ds2d_amd import Solver
from ds2d_amd.lib import DS2DError, ShotSpec

pytestmark = pytest.mark.gpu
EINVAL, ESTATE = -1, -4


def _solver(allnx=60, allnz=30, pml=20, nt=40, physics=0):
    return Solver(dx=DX, dz=DZ, dt=DT, nt=nt, allnx=allnx, allnz=allnz, pml=pml, nx_max=allnx, wavelet=wavelet(nt),
                  physics=physics)


def _shot(x0=0, nx=60, pml=20, sx=30, rec=None, igz=None, boundary=0, sid=0):
    rec = pml + np.arange(nx, dtype=np.int32) if rec is None else np.asarray(rec, np.int32)
    return ShotSpec(x0, nx, pml + sx, pml + 2, pml if igz is None else igz, rec, shot_id=sid, boundary=boundary)


def _code(fn):
    with pytest.raises(DS2DError) as e:
        fn()
    return e.value.code, str(e.value)


def test_calls_out_of_order_and_bad_arguments_are_rejected():
    s = _solver()
    code, msg = _code(lambda: s.batch_forward([_shot()]))
    assert code == ESTATE and "model" in msg
    vp, vs, rho = elastic_model(60, 30, 3)
    s.set_model(vp, vs, rho)
    cases = {
        "window": [_shot(x0=10, nx=60)],                                  # past the model's last column
        "source": [_shot(sx=-18)],                                        # inside the 4-cell frame of no update
        "receiver": [_shot(rec=[20, 99, 30])],                            # column outside the window
        "receiver row": [_shot(igz=2)],                                   # row inside the no-update frame
    }
    for what, shots in cases.items():
        code, msg = _code(lambda: s.batch_forward(shots))
        print(f"{what}: {msg}")
        assert code == EINVAL
    assert _code(lambda: s.batch_forward([]))[0] == EINVAL
    assert _code(lambda: s.set_chunk(64, 300, 2))[0] == EINVAL          # lanes must be 256, 512, 704 or 768
    assert _code(lambda: s.set_chunk(64, 704, 3))[0] == EINVAL          # prefetch depth 2 or 4
    assert _code(lambda: s.set_time_block(3))[0] == EINVAL
    assert _code(lambda: s.decomp_rtm([_shot()], 7, np.zeros(60 * 40, np.float32)))[0] == EINVAL
    a = _solver(physics=1)
    a.set_model(vp, None, rho)
    assert _code(lambda: a.rot_forward([_shot(boundary=1)]))[0] == EINVAL  # the rotation scheme is elastic


def test_edge_windows_empty_receivers_and_a_narrow_window_match_oracle():
    """One batch: a window flush with the model's left edge, one flush with
    its right edge, a shot with no receivers (its records are empty and the
    others are unaffected), and a 12-column window."""
    allnx, allnz, pml, nt = 90, 36, 20, 220
    vp, vs, rho = elastic_model(allnx, allnz, 9)
    wav = wavelet(nt)
    shots = [ShotSpec(0, 40, pml + 5, pml + 2, pml, pml + np.arange(40, dtype=np.int32), shot_id=0, boundary=0),
             ShotSpec(50, 40, pml + 35, pml + 2, pml, pml + np.arange(0, 40, 3, dtype=np.int32), shot_id=1,
                      boundary=1),
             ShotSpec(20, 30, pml + 15, pml + 2, pml, np.zeros(0, np.int32), shot_id=2, boundary=0),
             ShotSpec(39, 12, pml + 6, pml + 3, pml + 1, pml + np.arange(12, dtype=np.int32), shot_id=3,
                      boundary=0)]
    s = _solver(allnx, allnz, pml, nt)
    s.set_model(vp, vs, rho)
    for mode in (2, 0):
        rz, rx, _ = s.batch_forward(shots, mode=mode)
        assert rz.size == sum(sh.ntr for sh in shots) * nt
        off = 0
        for sh in shots:
            n = sh.ntr * nt
            if n:
                oz, ox = orc.elastic_forward(orc.make_shot(sh, allnx, allnz, pml, nt, DX, DZ, DT), vp, vs, rho, wav)
                ez, ex = orc.rel_l2(rz[off:off + n], oz), orc.rel_l2(rx[off:off + n], ox)
                print(f"mode {mode} shot {sh.shot_id} ({sh.nx} columns from {sh.ngx_min}): rel z {ez:.2e} x {ex:.2e}")
                assert ez <= 1e-5 and ex <= 1e-5
            off += n
