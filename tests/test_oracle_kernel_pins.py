"""Pins of the oracle's single kernels that the schedule tests only exercise
in composition: the receiver-side adjoint steps, the adjoint Born source, the
reconstruction strains, the edge bands, the imaging condition and the
reference acoustic step.

Each check relates a kernel to an already pinned one through an identity the
reference's formulas imply (moduli inside the derivative, the transpose of
add_born_source, exact reversal through the saved bands), or to a closed form
(Taylor exactness of the staggered derivatives, single-entry imaging), so a
flipped stagger, a dropped term or an off-by-one bound fails it.
"""
import numpy as np
import pytest

import oracle_lib as orc

F = {n: i for i, n in enumerate(orc.FIELDS15)}
VEL = ["vx", "vz", "vx_x", "vx_z", "vz_x", "vz_z"]
STR = ["txx", "tzz", "txz", "txx_x", "txx_z", "tzz_x", "tzz_z", "txz_x", "txz_z"]


def _media(rng, nxp, nzp, damp=True):
    lam = 1.0 + rng.random((nxp, nzp))
    mu = 0.5 + rng.random((nxp, nzp))
    rho = 1.0 + rng.random((nxp, nzp))
    dampx = 0.3 * rng.random((nxp, nzp)) if damp else np.zeros((nxp, nzp))
    dampz = 0.3 * rng.random((nxp, nzp)) if damp else np.zeros((nxp, nzp))
    return lam, mu, rho, dampx, dampz


def _close(a, b, rtol=1e-12):
    scale = max(np.abs(b).max(), 1e-300)
    np.testing.assert_allclose(a, b, rtol=0, atol=rtol * scale)


@pytest.mark.parametrize("direction", [-1, 1])
def test_adjoint_v_is_update_v_of_moduli_weighted_stress(direction):
    """FD2DGPU.cu:591-682 put the moduli inside the derivative: in a
    heterogeneous medium the receiver velocity step equals born.cu:215-279's
    velocity step applied to the cell-wise products C.sigma, i.e.
    txx' = (l+2m) txx + l tzz, tzz' = (l+2m) tzz + l txx, txz' = m txz."""
    rng = np.random.default_rng(1)
    nxp, nzp = 30, 26
    lam, mu, rho, dampx, dampz = _media(rng, nxp, nzp)
    f = rng.standard_normal((15, nxp, nzp))
    kw = dict(lam=lam, mu=mu, rho=rho, dampx=dampx, dampz=dampz, dx=1.3, dz=0.9, dt=0.7, direction=direction)
    adj, _ = orc.kernel("adjoint_v", f, **kw)
    g = f.copy()
    g[F["txx"]] = (lam + 2 * mu) * f[F["txx"]] + lam * f[F["tzz"]]
    g[F["tzz"]] = (lam + 2 * mu) * f[F["tzz"]] + lam * f[F["txx"]]
    g[F["txz"]] = mu * f[F["txz"]]
    fwd, _ = orc.kernel("update_v", g, **kw)
    for n in VEL:
        _close(adj[F[n]], fwd[F[n]])
    for n in STR:
        assert np.array_equal(adj[F[n]], f[F[n]]), n
    # the update covers [nop, n-nop)^2 only
    assert np.array_equal(adj[F["vx"]][:4], f[F["vx"]][:4]) and np.array_equal(adj[F["vx"]][:, -4:], f[F["vx"]][:, -4:])
    assert not np.allclose(adj[F["vx"]][4:-4, 4:-4], f[F["vx"]][4:-4, 4:-4])


def test_adjoint_s_is_strain_like_stress_step():
    """FD2DGPU.cu:725-761, 860-895: the receiver stress step carries no moduli
    and keeps txx = txx_x, tzz = tzz_z.  Its split updates equal
    born.cu:282-349's with (l, m) = (1, 0) for txx_x and tzz_z and with
    (l, m) = (0, 1) for txz_x and txz_z; txx_z and tzz_x are left alone."""
    rng = np.random.default_rng(2)
    nxp, nzp = 28, 30
    _, _, rho, dampx, dampz = _media(rng, nxp, nzp)
    one, zero = np.ones((nxp, nzp)), np.zeros((nxp, nzp))
    f = rng.standard_normal((15, nxp, nzp))
    kw = dict(rho=rho, dampx=dampx, dampz=dampz, dx=0.8, dz=1.4, dt=0.6, direction=-1)
    adj, _ = orc.kernel("adjoint_s", f, lam=one, mu=one, **kw)
    a, _ = orc.kernel("update_s", f, lam=one, mu=zero, **kw)
    b, _ = orc.kernel("update_s", f, lam=zero, mu=one, **kw)
    for n in ("txx_x", "tzz_z"):
        _close(adj[F[n]], a[F[n]])
    for n in ("txz_x", "txz_z", "txz"):
        _close(adj[F[n]], b[F[n]])
    assert np.array_equal(adj[F["txx"]][4:-4, 4:-4], adj[F["txx_x"]][4:-4, 4:-4])
    assert np.array_equal(adj[F["tzz"]][4:-4, 4:-4], adj[F["tzz_z"]][4:-4, 4:-4])
    for n in ("txx_z", "tzz_x") + tuple(VEL):
        assert np.array_equal(adj[F[n]], f[F[n]]), n


def _born_pair_case(rng, nxp=36, nzp=32, pml=8, dms_zero=True):
    lam, mu, rho, _, _ = _media(rng, nxp, nzp, damp=False)
    dmp = 0.1 * rng.standard_normal((nxp, nzp))
    dms = np.zeros((nxp, nzp)) if dms_zero else 0.1 * rng.standard_normal((nxp, nzp))
    v = np.zeros((15, nxp, nzp))
    v[F["vx"]] = rng.standard_normal((nxp, nzp))
    v[F["vz"]] = rng.standard_normal((nxp, nzp))
    s = np.zeros((15, nxp, nzp))
    inner = (slice(pml + 4, nxp - pml - 4), slice(pml + 4, nzp - pml - 4))
    for n in ("txx", "tzz", "txz"):
        s[F[n]][inner] = rng.standard_normal((nxp - 2 * pml - 8, nzp - 2 * pml - 8))
    kw = dict(pml=pml, lam=lam, mu=mu, rho=rho, dmp=dmp, dms=dms, dx=1.1, dz=0.7)
    return v, s, rho, kw


def _born_dot(v, s, rho, kw):
    z = np.zeros_like(v)
    bv, _ = orc.kernel("add_born_source", z, src15=v, direction=1, **kw)
    bts, _ = orc.kernel("add_born_source_adjoint", z, src15=s, direction=-1, **kw)
    lhs = sum(float((bv[F[n]] * s[F[n]]).sum()) for n in ("txx", "tzz", "txz"))
    rhs = float((rho * (bts[F["vx"]] * v[F["vx"]] + bts[F["vz"]] * v[F["vz"]])).sum())
    return lhs, rhs


def test_born_source_adjoint_is_transpose_when_dms_is_zero():
    """born.cu:408-527 with direction -1 is rho^-1 times the transpose of
    born.cu:351-405 (D-^T = -D+), exactly when dms = 0, where the
    tmp_txz_x sampling quirk of born.cu:514-515 multiplies zero."""
    rng = np.random.default_rng(3)
    v, s, rho, kw = _born_pair_case(rng)
    lhs, rhs = _born_dot(v, s, rho, kw)
    assert abs(lhs) > 1.0
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_born_source_adjoint_quirk_sits_in_the_vz_txz_term():
    """With dms != 0 the transpose still holds for every term except the one
    born.cu:514-515 computes with mu*dms sampled at ix+i-1 for both taps: the
    txz -> vz term."""
    rng = np.random.default_rng(4)
    v, s, rho, kw = _born_pair_case(rng, dms_zero=False)
    # vx outputs only: exact
    v1 = v.copy()
    v1[F["vz"]] = 0
    lhs, rhs = _born_dot(v1, s, rho, kw)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # no txz input: exact
    s1 = s.copy()
    s1[F["txz"]] = 0
    lhs, rhs = _born_dot(v, s1, rho, kw)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # vz and txz alone: the quirk breaks the transpose
    v2 = v.copy()
    v2[F["vx"]] = 0
    s2 = np.zeros_like(s)
    s2[F["txz"]] = s[F["txz"]]
    lhs, rhs = _born_dot(v2, s2, rho, kw)
    assert abs(lhs - rhs) > 1e-3 * abs(lhs)


@pytest.mark.parametrize("k,j", [(1, 0), (3, 2), (7, 1), (2, 7), (5, 5)])
def test_recon_strains_taylor_exact(k, j):
    """FD2DGPU.cu:993-1082 emit vx_gx = D-x vx / dx, vz_gz = D-z vz / dz at the
    stress node and vx_gz = D+z vx / dz, vz_gx = D+x vz / dx at the txz node.
    With vx at (x+1/2, z) and vz at (x, z+1/2) sampled from polynomials of
    degree <= 7 in the derivative direction they are the exact derivatives,
    on [pml, n-pml)^2 only."""
    nxp, nzp, pml, dx, dz = 40, 36, 9, 1.7, 1.1
    x0, z0 = 19.3 * dx, 17.6 * dz
    ix = np.arange(nxp)[:, None] * np.ones((1, nzp))
    iz = np.ones((nxp, 1)) * np.arange(nzp)[None, :]
    P = lambda x, z: (x - x0) ** k * (z - z0) ** j
    Px = lambda x, z: k * (x - x0) ** max(k - 1, 0) * (z - z0) ** j
    Pz = lambda x, z: (x - x0) ** k * j * (z - z0) ** max(j - 1, 0)
    Q = lambda x, z: (x - x0) ** j * (z - z0) ** k
    Qx = lambda x, z: j * (x - x0) ** max(j - 1, 0) * (z - z0) ** k
    Qz = lambda x, z: (x - x0) ** j * k * (z - z0) ** max(k - 1, 0)
    f = np.zeros((15, nxp, nzp))
    f[F["vx"]] = P((ix + 0.5) * dx, iz * dz)
    f[F["vz"]] = Q(ix * dx, (iz + 0.5) * dz)
    rng = np.random.default_rng(5)
    lam, mu, rho, dampx, dampz = _media(rng, nxp, nzp)
    out, st = orc.kernel("recon_s", f, pml=pml, lam=lam, mu=mu, rho=rho, dampx=dampx, dampz=dampz, dx=dx, dz=dz,
                         dt=0.5, direction=-1)
    inner = (slice(pml, nxp - pml), slice(pml, nzp - pml))
    expect = [Px(ix * dx, iz * dz), Qz(ix * dx, iz * dz), Pz((ix + 0.5) * dx, (iz + 0.5) * dz),
              Qx((ix + 0.5) * dx, (iz + 0.5) * dz)]
    for q in range(4):
        e = expect[q][inner]
        np.testing.assert_allclose(st[q][inner], e, rtol=0, atol=1e-9 * max(np.abs(e).max(), 1.0))
        frame = np.ones((nxp, nzp), bool)
        frame[inner] = False
        assert np.all(st[q][frame] == 0)
    # the stress part is born.cu:282-349's update on the interior, untouched outside
    ref, _ = orc.kernel("update_s", f, lam=lam, mu=mu, rho=rho, dampx=dampx, dampz=dampz, dx=dx, dz=dz, dt=0.5,
                        direction=-1)
    for n in STR:
        assert np.array_equal(out[F[n]][inner], ref[F[n]][inner]), n
        fr = np.ones((nxp, nzp), bool)
        fr[inner] = False
        assert np.array_equal(out[F[n]][fr], f[F[n]][fr]), n


def test_edge_bands_reconstruct_the_interior_exactly():
    """FD2DGPU.cu:3674-3731 save vx, vz in the 2N = 8 cells straddling the PML
    edge on four sides.  A forward run with damping that saves them every
    step, reversed with the interior-only reconstruction steps
    (FD2DGPU.cu:903-1082, dir -1) that load them first, recovers the initial
    stresses on [pml, n-pml)^2 and velocities on [pml+4, n-pml-4)^2."""
    rng = np.random.default_rng(6)
    nxp, nzp, pml, nt = 40, 36, 10, 25
    vp = 2.0 + 0.2 * rng.random((nxp, nzp))
    dampx, dampz = orc.damp_profile(vp, pml, 1e-3)
    lam = vp ** 2 * 0.6
    mu = vp ** 2 * 0.2
    rho = np.ones((nxp, nzp))
    kw = dict(pml=pml, lam=lam, mu=mu, rho=rho, dampx=dampx, dampz=dampz, dx=10.0, dz=10.0, dt=0.5)
    f0 = np.zeros((15, nxp, nzp))
    xs, zs = np.meshgrid(np.arange(nxp), np.arange(nzp), indexing="ij")
    bump = np.exp(-((xs - 20.0) ** 2 + (zs - 17.0) ** 2) / 12.0)
    for n, a in (("vx", 1.0), ("vz", -0.5), ("txx_x", 2.0), ("tzz_z", 1.5), ("txz_x", 0.7)):
        f0[F[n]] = a * bump
    f0[F["txx"]] = f0[F["txx_x"]]
    f0[F["tzz"]] = f0[F["tzz_z"]]
    f0[F["txz"]] = f0[F["txz_x"]]
    f0[F["vx_x"]] = f0[F["vx"]]
    f0[F["vz_x"]] = f0[F["vz"]]
    nb = orc.edge_band_size(nxp, nzp, pml, nt)
    bx, bz = np.zeros(nb), np.zeros(nb)
    f = f0.copy()
    for it in range(nt):
        f, _ = orc.kernel("update_v", f, direction=1, **kw)
        f, _ = orc.kernel("update_s", f, direction=1, **kw)
        _, bx = orc.edge_io(f[F["vx"]], bx, pml, nt, it, True)
        _, bz = orc.edge_io(f[F["vz"]], bz, pml, nt, it, True)
    assert np.abs(f[F["vx"]][:pml]).max() > 1e-3 * np.abs(f0[F["vx"]]).max()  # the wave reached the PML
    for it in range(nt - 1, -1, -1):
        f[F["vx"]], _ = orc.edge_io(f[F["vx"]], bx, pml, nt, it, False)
        f[F["vz"]], _ = orc.edge_io(f[F["vz"]], bz, pml, nt, it, False)
        f, _ = orc.kernel("recon_s", f, direction=-1, **kw)
        f, _ = orc.kernel("recon_v", f, direction=-1, **kw)
    si = (slice(pml, nxp - pml), slice(pml, nzp - pml))
    vi = (slice(pml + 4, nxp - pml - 4), slice(pml + 4, nzp - pml - 4))
    scale = np.abs(f0).max()
    for n in STR:
        np.testing.assert_allclose(f[F[n]][si], f0[F[n]][si], rtol=0, atol=1e-10 * scale, err_msg=n)
    for n in VEL:
        np.testing.assert_allclose(f[F[n]][vi], f0[F[n]][vi], rtol=0, atol=1e-10 * scale, err_msg=n)


def test_edge_band_save_load_identity():
    """One saved step loads back exactly the band cells, and only into step it."""
    rng = np.random.default_rng(7)
    nxp, nzp, pml, nt, N = 30, 24, 7, 4, 4
    a = rng.standard_normal((nxp, nzp))
    nb = orc.edge_band_size(nxp, nzp, pml, nt)
    _, bands = orc.edge_io(a, np.zeros(nb), pml, nt, 2, True)
    per_step = nb // nt
    nz_steps = [np.count_nonzero(bands.reshape(-1)) > 0]
    out, _ = orc.edge_io(np.zeros((nxp, nzp)), bands, pml, nt, 2, False)
    ix, iz = np.meshgrid(np.arange(nxp), np.arange(nzp), indexing="ij")
    inx = (ix >= pml) & (ix < nxp - pml)
    inz = (iz >= pml) & (iz < nzp - pml)
    band = (inx & (((iz >= pml - N) & (iz < pml + N)) | ((iz >= nzp - pml - N) & (iz < nzp - pml + N)))) | \
           (inz & (((ix >= pml - N) & (ix < pml + N)) | ((ix >= nxp - pml - N) & (ix < nxp - pml + N))))
    assert np.array_equal(out[band], a[band]) and np.all(out[~band] == 0)
    empty, _ = orc.edge_io(np.zeros((nxp, nzp)), bands, pml, nt, 1, False)
    assert np.all(empty == 0) and all(nz_steps)
    assert per_step == 2 * (nxp - 2 * pml) * 2 * N + 2 * 2 * N * (nzp - 2 * pml)


def test_image_gradient_closed_forms():
    """FD2DGPU.cu:3245-3263: PP += -2 vp^2 rho (exx + ezz)(sxx + szz) and
    PS += -2 vs^2 rho (-2 (exx szz + ezz sxx) + (exz + ezx) sxz), from the
    physical cell (ix+pml, iz+pml) into global column col0 + ix; illumination
    vx^2 + vz^2 (FD2DGPU.cu:3235-3243).  Constant strains against one unit
    receiver stress at a time give each product alone."""
    nxp, nzp, pml, allnx, col0 = 22, 18, 5, 30, 9
    nz = nzp - 2 * pml
    a, b, c1, c2 = 0.3, -0.7, 0.11, 0.05
    vp, vs, rho = 3.0, 1.7, 2.2
    st = np.zeros((4, nxp, nzp))
    st[0], st[1], st[2], st[3] = a, b, c1, c2
    cvp, cvs, crho = (np.full((nxp, nzp), x) for x in (vp, vs, rho))
    kp, ks = -2 * vp ** 2 * rho, -2 * vs ** 2 * rho
    for (ix0, iz0) in ((0, 0), (11, 7), (3, 4)):
        for comp, epp, eps in (("txx", kp * (a + b), ks * (-2 * b)), ("tzz", kp * (a + b), ks * (-2 * a)),
                               ("txz", 0.0, ks * (c1 + c2))):
            bk = np.zeros((15, nxp, nzp))
            bk[F[comp]][ix0 + pml, iz0 + pml] = 1.0
            pp, ps, _ = orc.image(st, bk, cvp, cvs, crho, pml, col0, allnx, nz)
            expp, exps = np.zeros((allnx, nz)), np.zeros((allnx, nz))
            expp[col0 + ix0, iz0] = epp
            exps[col0 + ix0, iz0] = eps
            np.testing.assert_allclose(pp, expp, rtol=1e-14, atol=1e-14)
            np.testing.assert_allclose(ps, exps, rtol=1e-14, atol=1e-14)
            pp2, ps2, _ = orc.image(st, bk, cvp, cvs, crho, pml, col0, allnx, nz, pp=pp, ps=ps)
            np.testing.assert_allclose(pp2, 2 * expp, rtol=1e-14, atol=1e-14)  # accumulates
    # a receiver stress in the frame contributes nothing; illumination from the velocities
    bk = np.zeros((15, nxp, nzp))
    bk[F["txx"]][pml - 1, pml + 2] = 1.0
    bk[F["vx"]][pml + 1, pml + 2] = 2.0
    bk[F["vz"]][pml + 1, pml + 2] = 3.0
    pp, ps, il = orc.image(st, bk, cvp, cvs, crho, pml, col0, allnx, nz)
    assert np.all(pp == 0) and np.all(ps == 0)
    assert il[col0 + 1, 2] == 13.0 and np.count_nonzero(il) == 1


def test_acoustic_ref_step_mirrored_stagger_taylor():
    """FD2DGPU.cu:253-325: p sits on the nodes, vx at x-1/2 and vz at z-1/2
    (mirrored against the elastic kernels); vz is scaled by dt/dx, not dt/dz
    (FD2DGPU.cu:318).  From a degree-7 polynomial p (through px, with v = 0)
    one step gives vx = -(dt/rho) dp/dx (x-1/2) and vz = -(dt/rho)(dz/dx)
    dp/dz (z-1/2); from polynomial velocities it gives px = -K dt dvx/dx and
    pz = -K dt dvz/dz at the node."""
    nxp, nzp, dx, dz, dt, K, R = 30, 28, 1.0, 1.5, 0.1, 2.5, 1.7
    x0, z0 = 14.2 * dx, 13.1 * dz
    ix = np.arange(nxp)[:, None] * np.ones((1, nzp))
    iz = np.ones((nxp, 1)) * np.arange(nzp)[None, :]
    k = np.full((nxp, nzp), K)
    rr = np.full((nxp, nzp), R)
    damp = np.zeros((nxp, nzp))
    s = np.zeros((5, nxp, nzp))
    s[0] = (ix * dx - x0) ** 7 * (iz * dz - z0) ** 3
    s[2] = s[0]  # p = px + pz inside the update, the same polynomial outside it
    out = orc.acoustic_ref_step(k, rr, damp, s, dx, dz, dt)
    xm, zm = (ix - 0.5) * dx - x0, (iz - 0.5) * dz - z0
    evx = -(dt / R) * 7 * xm ** 6 * (iz * dz - z0) ** 3
    evz = -(dt / R) * (dz / dx) * (ix * dx - x0) ** 7 * 3 * zm ** 2
    bx = (slice(4, nxp - 3), slice(4, nzp - 4))   # vx: ix in [nop, nxp-nop], iz in [nop, nzp-nop-1]
    bz = (slice(4, nxp - 4), slice(4, nzp - 3))   # vz: ix in [nop, nxp-nop-1], iz in [nop, nzp-nop]
    np.testing.assert_allclose(out[3][bx], evx[bx], rtol=0, atol=1e-9 * np.abs(evx[bx]).max())
    np.testing.assert_allclose(out[4][bz], evz[bz], rtol=0, atol=1e-9 * np.abs(evz[bz]).max())
    mx = np.ones((nxp, nzp), bool)
    mx[bx] = False
    mz = np.ones((nxp, nzp), bool)
    mz[bz] = False
    assert np.all(out[3][mx] == 0) and np.all(out[4][mz] == 0)
    # p from velocities on [nop-1, n-nop-1]^2
    s = np.zeros((5, nxp, nzp))
    s[3] = ((ix - 0.5) * dx - x0) ** 5
    s[4] = ((iz - 0.5) * dz - z0) ** 6
    out = orc.acoustic_ref_step(k, rr, damp, s, dx, dz, dt)
    bp = (slice(3, nxp - 4), slice(3, nzp - 4))   # [nop-1, n-nop-1] inclusive
    epx = -K * dt * 5 * (ix * dx - x0) ** 4
    epz = -K * dt * 6 * (iz * dz - z0) ** 5
    np.testing.assert_allclose(out[0][bp], epx[bp], rtol=0, atol=1e-9 * np.abs(epx[bp]).max())
    np.testing.assert_allclose(out[1][bp], epz[bp], rtol=0, atol=1e-9 * np.abs(epz[bp]).max())
    np.testing.assert_allclose(out[2][bp], epx[bp] + epz[bp], rtol=1e-12, atol=1e-9 * np.abs(epz[bp]).max())
    mp = np.ones((nxp, nzp), bool)
    mp[bp] = False
    assert np.all(out[0][mp] == 0) and np.all(out[2][mp] == 0)
    assert np.array_equal(out[3][mx], s[3][mx])  # velocities outside their bounds keep their values


def test_acoustic_ref_damping_one_array():
    """FD2DGPU.cu:505-519 write dampX, then dampZ, into one array: the x
    profile in the side strips, the z profile wherever z is in a strip
    (corners included), zero inside."""
    rng = np.random.default_rng(8)
    nxp, nzp, pml = 40, 34, 8
    v = 2000 + 500 * rng.random((nxp, nzp))
    d = orc.acoustic_ref_damp(v, pml, 1e-6)
    dxa, dza = orc.damp_profile(v, pml, 1e-6)
    ix, iz = np.meshgrid(np.arange(nxp), np.arange(nzp), indexing="ij")
    xs = (ix < pml) | (ix > nxp - pml)
    zs = (iz < pml) | (iz > nzp - pml)
    assert np.array_equal(d[zs], dza[zs])
    assert np.array_equal(d[xs & ~zs], dxa[xs & ~zs])
    assert np.all(d[~xs & ~zs] == 0)
    assert d[0, 3] == dza[0, 3] != dxa[0, 3]  # a corner cell keeps the z profile


def test_acoustic_ref_homogeneous_moveout():
    """Homogeneous medium, the reference acoustic step (FD2DGPU.cu:253-325)
    with its damping (FD2DGPU.cu:505-519) driven by a pressure source in the
    middle of the grid: the lag that best aligns the pressure at two offsets
    on the source row is (r2 - r1) / v within one sample."""
    nxp, nzp, pml, nt, dx, dt, v, rho = 260, 120, 20, 700, 10.0, 0.001, 2000.0, 1000.0
    vt = np.full((nxp, nzp), v)
    damp = orc.acoustic_ref_damp(vt, pml, 1e-6)
    k = vt * vt * rho
    rr = np.full((nxp, nzp), rho)
    wav = orc.ricker(nt, 20.0, dt)
    sx, sz = 40, 60
    st = np.zeros((5, nxp, nzp))
    offs = (40, 70, 100, 120)
    rec = np.zeros((len(offs), nt))
    for it in range(nt):
        st[0][sx, sz] += wav[it]
        st = orc.acoustic_ref_step(k, rr, damp, st, dx, dx, dt)
        for j, o in enumerate(offs):
            rec[j, it] = st[2][sx + o, sz]
    for (j1, j2) in ((0, 1), (1, 3), (0, 2)):
        t1, t2 = rec[j1], rec[j2]
        lags = np.arange(0, 400)
        cc = [float(np.dot(t1[: nt - L], t2[L:])) for L in lags]
        expect = (offs[j2] - offs[j1]) * dx / v / dt
        got = int(lags[int(np.argmax(cc))])
        assert abs(got - expect) <= 1, (offs[j1], offs[j2], got, expect)
