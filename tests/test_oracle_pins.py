"""Pins of the CPU oracle against closed forms, brute force and the compiled
reference (oracle/_ref built from the reference's own laplace.cpp).

Each check is chosen so a plausible mistake (wrong stagger, dropped term,
off-by-one on a boundary, wrong sign) fails it.
"""
import numpy as np
import pytest

import oracle_lib as orc


def _elastic_fields(nxp, nzp):
    return np.zeros((15, nxp, nzp))


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7])
def test_staggered_derivative_exact_to_degree7(k):
    """coeff2 (FD2DGPU.cuh:16) with the vx stagger of born.cu:237 differentiates
    x^k exactly at the half point for k <= 7: one velocity step from a
    polynomial txx gives vx[ix] = dt/rho * k (x_{ix+1/2})^(k-1)."""
    nxp, nzp, dx, dt = 24, 12, 1.0, 1.0
    f = _elastic_fields(nxp, nzp)
    x = np.arange(nxp, dtype=np.float64)[:, None] * dx * np.ones((1, nzp))
    x0 = 11.3  # keep magnitudes moderate
    f[2] = (x - x0) ** k  # txx
    f[10] = f[2]          # txx_x (so the split sum is consistent)
    one = np.ones((nxp, nzp))
    zero = np.zeros((nxp, nzp))
    # a velocity-only step: run dir=+1 with lam = mu = 0 so stresses stay put
    out = orc.elastic_run(zero, zero, one, zero, zero, f.reshape(-1), 1, 1, dx, dx, dt).reshape(15, nxp, nzp)
    vx = out[0]
    xs = np.arange(nxp) + 0.5 - x0
    expect = k * xs ** (k - 1)
    sl = slice(4, nxp - 4)
    np.testing.assert_allclose(vx[sl, 5], expect[sl], rtol=1e-12, atol=1e-9 * np.abs(expect[sl]).max())


def test_staggered_derivative_not_exact_degree9():
    """Control: degree 9 is not differentiated exactly (truncation error > 0)."""
    nxp, nzp = 24, 12
    f = _elastic_fields(nxp, nzp)
    x = np.arange(nxp, dtype=np.float64)[:, None] * np.ones((1, nzp)) - 11.3
    f[2] = x ** 9
    f[10] = f[2]
    one, zero = np.ones((nxp, nzp)), np.zeros((nxp, nzp))
    out = orc.elastic_run(zero, zero, one, zero, zero, f.reshape(-1), 1, 1, 1.0, 1.0, 1.0).reshape(15, nxp, nzp)
    xs = np.arange(nxp) + 0.5 - 11.3
    assert np.abs(out[0][4:-4, 5] - 9 * xs[4:-4] ** 8).max() > 1.0


def test_ricker_closed_form():
    """main.cpp:461-473: peak 1 at i = kt = int(1/(f dt)), zero crossings at
    t - kt dt = +-1/(sqrt(2) pi f)."""
    nt, f, dt = 400, 15.0, 0.0008
    s = orc.ricker(nt, f, dt)
    kt = int(1.0 / f / dt)
    assert kt == 83
    assert s[kt] == pytest.approx(1.0, abs=1e-15)
    assert np.argmax(s) == kt
    t = np.arange(nt) * dt - kt * dt
    sp = np.pi * f * t
    np.testing.assert_allclose(s, np.exp(-sp ** 2) * (1 - 2 * sp ** 2), rtol=1e-12, atol=1e-15)
    tz = 1.0 / (np.sqrt(2) * np.pi * f)
    i0 = kt + int(tz / dt)
    assert s[i0] > 0 > s[i0 + 1]


def test_pad_window_brute_force():
    rng = np.random.default_rng(0)
    g = rng.standard_normal((30, 11))
    pml, x0, nx = 5, 7, 12
    out = orc.pad_window(g, x0, nx, pml)
    for ix in range(nx + 2 * pml):
        for iz in range(11 + 2 * pml):
            sx = min(max(ix - pml, 0), nx - 1)
            sz = min(max(iz - pml, 0), 10)
            assert out[ix, iz] == g[x0 + sx, sz]


def test_damp_profile_points():
    """FD2DGPU.cu:444-494: value log10(1/rr)*5vp/(2pml) on the outermost
    near-side cell, zero at ix = pml and at ix = nxp-pml (strict '>'),
    ((pml-i)/pml)^4 decay, far side mirrored one cell inward."""
    pml, nxp, nzp, rr, v = 10, 40, 30, 1e-3, 3000.0
    vp = np.full((nxp, nzp), v)
    dxa, dza = orc.damp_profile(vp, pml, rr)
    base = np.log10(1 / rr) * 5 * v / (2 * pml)
    assert dxa[0, 15] == pytest.approx(base, rel=1e-14)
    assert dxa[pml, 15] == 0.0 and dxa[pml - 1, 15] == pytest.approx(base * (1 / pml) ** 4, rel=1e-12)
    assert dxa[nxp - pml, 15] == 0.0
    assert dxa[nxp - 1, 15] == pytest.approx(base * ((pml - 1) / pml) ** 4, rel=1e-12)
    assert np.all(dxa[pml:nxp - pml + 1] == 0)
    assert dza[20, 0] == pytest.approx(base, rel=1e-14)
    assert dza[20, nzp - pml] == 0.0 and np.all(dza[:, pml:nzp - pml + 1] == 0)
    # x profile independent of z and vice versa
    assert np.all(dxa[:, 0] == dxa[:, nzp // 2]) and np.all(dza[0, :] == dza[nxp // 2, :])


def test_random_boundary_properties():
    nxp, nzp, pml = 50, 40, 10
    rng = np.random.default_rng(1)
    vp = (2000 + 1000 * rng.random((nxp, nzp))).astype(np.float32).astype(np.float64)
    vs = (vp / 1.73).astype(np.float32).astype(np.float64)
    a, b = orc.random_boundary(vp, vs, pml, seed=77, shot=3, amp=0.3)
    a2, _ = orc.random_boundary(vp, vs, pml, seed=77, shot=3, amp=0.3)
    c, _ = orc.random_boundary(vp, vs, pml, seed=77, shot=4, amp=0.3)
    inner = (slice(pml, nxp - pml), slice(pml, nzp - pml))
    assert np.array_equal(a, a2)                     # deterministic
    assert not np.array_equal(a, c)                  # per-shot streams differ
    assert np.array_equal(a[inner], vp[inner])       # interior untouched
    frame = np.ones((nxp, nzp), bool)
    frame[inner] = False
    r = a[frame] / vp[frame]
    assert r.min() >= 0.7 - 1e-12 and r.max() <= 1.3 + 1e-12
    assert abs(r.mean() - 1.0) < 0.02 and r.std() > 0.15  # ~U(0.7,1.3): std 0.173
    np.testing.assert_allclose(b[frame] / a[frame], 1 / 1.73, rtol=3e-7)   # float quotient and product
    assert np.array_equal(a[frame], a[frame].astype(np.float32))           # the model holds floats
    assert np.array_equal(b[frame], b[frame].astype(np.float32))
    u = np.array([orc.rb_uniform(9, 0, i, j) for i in range(40) for j in range(40)])
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.06


def test_laplace_quadratic_closed_form():
    """laplace.cpp:82-106 on f = ix^2 + iz^2 gives 4/dx^2 on [2,n-2), 0 on the frame."""
    nx, nz, dx = 17, 13, 2.5
    ix, iz = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    out = orc.laplace_order((ix ** 2 + iz ** 2).astype(np.float64), dx)
    np.testing.assert_allclose(out[2:-2, 2:-2], 4 / dx ** 2, rtol=1e-13)
    fr = np.ones((nx, nz), bool)
    fr[2:-2, 2:-2] = False
    assert np.all(out[fr] == 0)


@pytest.mark.skipif(not orc.ref_laplace_available(), reason="reference tree not mounted: _ref not built")
@pytest.mark.parametrize("shape,dx", [((40, 23), 10.0), ((251, 101), 10.0), ((512, 256), 12.5), ((5, 5), 1.0)])
def test_laplace_matches_compiled_reference(shape, dx):
    rng = np.random.default_rng(2)
    img = rng.standard_normal(shape).astype(np.float32)
    ref = orc.ref_laplace_order(img, dx)
    ours = orc.laplace_order(img, dx, prec="f")
    assert np.array_equal(ours, ref)  # same arithmetic order: bit-identical
    ref_c = orc.ref_cutM(img, 3, 20, 4, 35)
    ours_c = img.copy()
    ours_c[:4] = 0
    ours_c[35:] = 0
    ours_c[:, :3] = 0
    ours_c[:, 20:] = 0
    assert np.array_equal(ref_c, ours_c)


def test_laplace_band_closed_forms():
    """Band filter (laplace.cpp:3-57, stated intent): S=(1/4,1/2,1/4) keeps
    linear functions and maps x^2 to x^2 + 1/2 (four passes add 2), so in the deep
    interior out(const) = out(ix) = out(iz) = 0 and out(ix^2) = out(iz^2) = -2."""
    nx, nz = 40, 36
    ix, iz = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(nz, dtype=np.float64), indexing="ij")
    deep = (slice(10, nx - 10), slice(10, nz - 10))
    for f, expect in ((np.full((nx, nz), 3.0), 0.0), (ix, 0.0), (iz, 0.0), (ix ** 2, -2.0), (iz ** 2, -2.0)):
        out = orc.laplace_band(f)
        np.testing.assert_allclose(out[deep], expect, atol=1e-10)


def test_ncc_closed_forms():
    rng = np.random.default_rng(3)
    ntr, nt = 5, 64
    oz, ox = rng.standard_normal(ntr * nt), rng.standard_normal(ntr * nt)
    # perfectly correlated synthetics: zero residual, misfit -1
    sz, sx, mf = orc.ncc_residual(2.5 * oz, 0.5 * ox, oz, ox, ntr, nt, 1)
    assert mf == pytest.approx(-1.0, rel=1e-7)
    assert np.abs(sz).max() < 1e-12 and np.abs(sx).max() < 1e-12
    # iteration 0: r = -D/|D|^2 per trace
    sz, sx, mf = orc.ncc_residual(np.zeros_like(oz), np.zeros_like(ox), oz, ox, ntr, nt, 0)
    assert mf == 0.0
    D = oz.reshape(ntr, nt)
    np.testing.assert_allclose(sz.reshape(ntr, nt), -D / (D ** 2).sum(1, keepdims=True), rtol=1e-12)
    # orthogonality: the NCC residual is orthogonal to the synthetic trace
    syn = rng.standard_normal(ntr * nt)
    sz, _, _ = orc.ncc_residual(syn, syn, oz, ox, ntr, nt, 1)
    assert np.abs((sz.reshape(ntr, nt) * syn.reshape(ntr, nt)).sum(1)).max() < 1e-12 * np.abs(syn).max() * 1e3


def test_adam_first_step_closed_form():
    """main.cpp:1092-1130 at iter 0: bias-corrected moments equal g and g^2,
    step = alpha * g / (|g| + eps)."""
    g = np.array([1.0, -2.0, 1e9, 0.0])
    z = np.zeros(4)
    s0, r0, img = orc.adam_update(g, z, z, z, z, z, 0)
    alpha, eps = np.float32(3e-5), np.float32(5e8)
    np.testing.assert_allclose(img, -alpha * g / (np.abs(g) + eps), rtol=1e-6)
    np.testing.assert_allclose(s0, 0.2 * g, rtol=1e-7)
