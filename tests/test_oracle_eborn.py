"""Pins of the oracle's plain-L2 elastic Born pair (PP/PS images -> scattered
vz, vx records, and its transpose), CPU only.

The forward map is checked against a central finite difference of the
nonlinear elastic propagator (the split-field update of born.cu:215-349 in
the oracle's elastic_forward) under the moduli perturbation that the Born
source of add_born_source (born.cu:351-405) linearises:
    d(lambda) = 2 (-(lambda + 2 mu) dmp + 2 mu dms) / dt,  d(mu) = -2 mu dms / dt,
and the adjoint by the dot-product test in FP64.
"""
import numpy as np

import oracle_lib as orc
from cases import DT, DX, DZ, full_window_shot, wavelet
from ds2d_amd import models


def _setup(nt=260, seed=21):
    allnx, allnz, pml = 56, 40, 20
    vp = models.layered_model(allnx, allnz, seed).astype(np.float64)
    vs, rho = models.elastic_from_vp(vp)
    vs, rho = vs.astype(np.float64), rho.astype(np.float64)
    wav = wavelet(nt)
    spec = full_window_shot(allnx, pml, 22, shot_id=3, boundary=1)
    return allnx, allnz, pml, nt, vp, vs, rho, wav, orc.make_shot(spec, allnx, allnz, pml, nt, DX, DZ, DT)


def _bumps(allnx, allnz):
    x, z = np.meshgrid(np.arange(allnx), np.arange(allnz), indexing="ij")
    pp = 0.05 * np.exp(-((x - 30) ** 2 + (z - 22) ** 2) / 12.0)
    ps = -0.03 * np.exp(-((x - 24) ** 2 + (z - 16) ** 2) / 10.0)
    return pp, ps


def test_eborn_linearity():
    allnx, allnz, pml, nt, vp, vs, rho, wav, shot = _setup()
    rng = np.random.default_rng(4)
    a = [rng.standard_normal((allnx, allnz)) for _ in range(4)]
    z1, x1 = orc.eborn_forward(shot, vp, vs, rho, a[0], a[1], wav)
    z2, x2 = orc.eborn_forward(shot, vp, vs, rho, a[2], a[3], wav)
    z3, x3 = orc.eborn_forward(shot, vp, vs, rho, 2 * a[0] - 3 * a[2], 2 * a[1] - 3 * a[3], wav)
    np.testing.assert_allclose(z3, 2 * z1 - 3 * z2, atol=1e-12 * np.abs(z3).max())
    np.testing.assert_allclose(x3, 2 * x1 - 3 * x2, atol=1e-12 * np.abs(x3).max())


def test_eborn_matches_finite_difference_of_the_propagator():
    allnx, allnz, pml, nt, vp, vs, rho, wav, shot = _setup()
    pp, ps = _bumps(allnx, allnz)
    bz, bx = orc.eborn_forward(shot, vp, vs, rho, pp, ps, wav)
    lam = rho * (vp ** 2 - 2 * vs ** 2)
    mu = rho * vs ** 2
    dlam = 2 * (-(lam + 2 * mu) * pp + 2 * mu * ps) / DT
    dmu = -2 * mu * ps / DT
    e = 1e-7

    def run(sign):
        lam2, mu2 = lam + sign * e * dlam, mu + sign * e * dmu
        vp2 = np.sqrt((lam2 + 2 * mu2) / rho)
        vs2 = np.sqrt(mu2 / rho)
        return orc.elastic_forward(shot, vp2, vs2, rho, wav)

    pz, px = run(+1)
    mz, mx = run(-1)
    fz, fx = (pz - mz) / (2 * e), (px - mx) / (2 * e)
    assert np.abs(bz).max() > 0 and np.abs(bx).max() > 0
    assert orc.rel_l2(bz, fz) < 1e-5
    assert orc.rel_l2(bx, fx) < 1e-5


def test_eborn_adjoint_dot_product():
    allnx, allnz, pml, nt, vp, vs, rho, wav, shot = _setup()
    rng = np.random.default_rng(12)
    pp, ps = rng.standard_normal((allnx, allnz)), rng.standard_normal((allnx, allnz))
    lz, lx = orc.eborn_forward(shot, vp, vs, rho, pp, ps, wav)
    rz, rx = rng.standard_normal(lz.size), rng.standard_normal(lx.size)
    gpp, gps = orc.eborn_adjoint(shot, vp, vs, rho, wav, rz, rx)
    lhs = float(lz @ rz + lx @ rx)
    rhs = float(pp.ravel() @ gpp + ps.ravel() @ gps)
    assert abs(gpp).max() > 0 and abs(gps).max() > 0
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_eborn_pp_and_ps_images_reach_both_components():
    """A PP-only and a PS-only perturbation each scatter into vz and vx, and the
    FP32 restatement stays within 1e-5 of FP64."""
    allnx, allnz, pml, nt, vp, vs, rho, wav, shot = _setup()
    pp, ps = _bumps(allnx, allnz)
    zero = np.zeros_like(pp)
    for a, b in ((pp, zero), (zero, ps)):
        z64, x64 = orc.eborn_forward(shot, vp, vs, rho, a, b, wav)
        z32, x32 = orc.eborn_forward(shot, vp, vs, rho, a, b, wav, prec="f")
        assert np.abs(z64).max() > 0 and np.abs(x64).max() > 0
        assert orc.rel_l2(z32, z64) < 1e-5 and orc.rel_l2(x32, x64) < 1e-5
