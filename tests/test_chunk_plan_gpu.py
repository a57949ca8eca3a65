"""Two steps per launch with the balanced chunk plan on the GPU: a batch of 20
uneven windows (PML and random-boundary windows mixed) whose columns are
enough for the plan to split them into 32 chunks per XCD.  The records match
the FP64 oracle to 1e-5 and are bit-identical to the fixed-width chunk list."""
import numpy as np
import pytest

import oracle_lib as orc
from cases import DT, DX, DZ, elastic_model, wavelet
from ds2d_amd import Solver
from ds2d_amd.lib import ShotSpec

pytestmark = pytest.mark.gpu
TOL = 1e-5


def test_balanced_plan_matches_oracle_and_fixed_width():
    allnx, allnz, pml, nt = 420, 40, 20, 240
    vp, vs, rho = elastic_model(allnx, allnz, 23)
    wav = wavelet(nt)
    rng = np.random.default_rng(4)
    shots = []
    for i in range(20):
        nx = int(rng.integers(180, 421))
        x0 = int(rng.integers(0, allnx - nx + 1))
        shots.append(ShotSpec(x0, nx, pml + nx // 2, pml + 2, pml + 1, pml + np.arange(0, nx, 5, dtype=np.int32),
                              shot_id=i, boundary=i % 3 == 2))
    s = Solver(dx=DX, dz=DZ, dt=DT, nt=nt, allnx=allnx, allnz=allnz, pml=pml, nx_max=allnx, wavelet=wav)
    s.set_model(vp, vs, rho)
    s.set_chunk(240, 704, 2)
    s.set_time_block(2)
    rz, rx, _ = s.batch_forward(shots, mode=2)
    assert s.chunk_plan_used()
    s.set_chunk_plan(False)
    fz, fx, _ = s.batch_forward(shots, mode=2)
    assert not s.chunk_plan_used()
    print(f"balanced vs fixed width: identical {np.array_equal(rz, fz) and np.array_equal(rx, fx)}")
    assert np.array_equal(rz, fz) and np.array_equal(rx, fx)
    off = 0
    worst = 0.0
    for sh in shots:
        n = sh.ntr * nt
        oz, ox = orc.elastic_forward(orc.make_shot(sh, allnx, allnz, pml, nt, DX, DZ, DT), vp, vs, rho, wav)
        ez, ex = orc.rel_l2(rz[off:off + n], oz), orc.rel_l2(rx[off:off + n], ox)
        worst = max(worst, ez, ex)
        assert np.abs(oz).max() > 0 and ez <= TOL and ex <= TOL, (sh.shot_id, ez, ex)
        off += n
    print(f"20 windows, balanced plan: worst rel {worst:.2e}")


def test_balanced_plan_with_two_row_blocks():
    """24 windows on a 320 x 700 model: 740 padded rows need two 704-lane row
    blocks, so the plan splits (window, row block) units.  Records bit-identical
    to fixed-width chunks; three windows (PML and random boundary) against the
    oracle."""
    allnx, allnz, pml, nt = 320, 700, 20, 120
    vp, vs, rho = elastic_model(allnx, allnz, 29)
    wav = wavelet(nt)
    rng = np.random.default_rng(9)
    shots = []
    for i in range(24):
        nx = int(rng.integers(220, 321))
        x0 = int(rng.integers(0, allnx - nx + 1))
        shots.append(ShotSpec(x0, nx, pml + nx // 2, pml + 2, pml + 1, pml + np.arange(0, nx, 7, dtype=np.int32),
                              shot_id=i, boundary=i % 2))
    s = Solver(dx=DX, dz=DZ, dt=DT, nt=nt, allnx=allnx, allnz=allnz, pml=pml, nx_max=allnx, wavelet=wav)
    s.set_model(vp, vs, rho)
    s.set_chunk(240, 704, 2)
    s.set_time_block(2)
    rz, rx, _ = s.batch_forward(shots, mode=2)
    assert s.chunk_plan_used()
    s.set_chunk_plan(False)
    fz, fx, _ = s.batch_forward(shots, mode=2)
    assert np.array_equal(rz, fz) and np.array_equal(rx, fx)
    off = 0
    for sh in shots:
        n = sh.ntr * nt
        if sh.shot_id in (0, 1, 23):
            oz, ox = orc.elastic_forward(orc.make_shot(sh, allnx, allnz, pml, nt, DX, DZ, DT), vp, vs, rho, wav)
            ez, ex = orc.rel_l2(rz[off:off + n], oz), orc.rel_l2(rx[off:off + n], ox)
            print(f"window {sh.shot_id} (boundary {sh.boundary}): rel z {ez:.2e} x {ex:.2e}")
            assert np.abs(oz).max() > 0 and ez <= TOL and ex <= TOL
        off += n
