"""The parameter-file driver end to end on the GPU: synthesize SU gathers,
then run the reference's LSRTM loop (main.cpp:1046-1186) with its Adam
update and with the nonlinear CG update (sjcgsolver, main.cpp:1278), and the
plain acoustic CG mode.  Each run is a child process."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ds2d_amd", "bin", "ds2d_lsrtm")

NX, NZ, PML = 96, 64, 20


def _params(tmp, mode, extra, maxiter=2):
    vp = np.full((NX, NZ), 2500.0, np.float32)
    vp[:, NZ // 2:] = 3000.0
    vs = (vp / 1.8).astype(np.float32)
    refl = np.zeros((NX, NZ), np.float32)
    refl[:, NZ // 2] = 0.1
    for name, a in (("vp.bin", vp), ("vs.bin", vs), ("refl.bin", refl)):
        a.tofile(str(tmp / name))
    img = tmp / f"img_{mode}_{extra.get('opt', 'x')}"
    kv = {"synth_nt": "400", "synth_dt": "0.001", "synth_dshot": "32", "synth_x0": "16",
          "synth_refl": str(tmp / "refl.bin"), "mode": mode}
    kv.update(extra)
    text = (f"10 10 {PML} {NX} {NZ} 1 100\n"
            f"{tmp / 'vp.bin'} {tmp / 'vs.bin'} {tmp / 'vz.su'} {tmp / 'vx.su'} {img}\n"
            f"15 {maxiter}\n2 2\n0 0 1 1\n" + "\n".join(f"{k}={v}" for k, v in kv.items()) + "\n")
    p = tmp / f"par_{mode}_{extra.get('opt', 'x')}.txt"
    p.write_text(text)
    return p, img


def _run(*args):
    r = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=240)
    print(r.stdout[-600:], r.stderr[-600:])
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built")
@pytest.mark.parametrize("opt", ["adam", "cg"])
def test_driver_reference_mode(tmp_path, opt):
    par, img = _params(tmp_path, "ref", {"opt": opt, "batch": "2"})
    _run("synth", par)
    for name in ("vz.su", "vx.su"):
        _check_su_headers(tmp_path / name, 3, NX, 400, 0.001, [16, 48, 80], 10.0)
    _run(par)
    last = np.fromfile(str(img / f"sum_image_pp_1_{NX}_{NZ}.dat"), np.float32)
    grad = np.fromfile(str(img / f"all_grad_pp_0_{NX}_{NZ}.dat"), np.float32)
    assert last.size == NX * NZ and grad.size == NX * NZ
    assert np.all(np.isfinite(last)) and np.all(np.isfinite(grad))
    assert np.abs(grad).max() > 0 and np.abs(last).max() > 0
    # The first update moves against the gradient for both optimizers.
    first = np.fromfile(str(img / f"sum_image_pp_0_{NX}_{NZ}.dat"), np.float32)
    assert float(np.dot(first.astype(np.float64), grad)) < 0
    lines = (img / "misfit.txt").read_text().strip().split("\n")
    assert len(lines) == 2


def _check_su_headers(path, nshots, ntr, nt, dt, sxs, dx):
    """Every 240-byte header as write_shot_gather_su_3d writes it
    (read_write_su.c:43-62, one in-line index i = 0): tracl = 0,
    fldr = ep = shot + 1, tracf = cdp = receiver k, sx, sy = 0, gx, gy = 0,
    offset = gx - sx, ns, dt = (short)(dt * 1000000), trid = scalel =
    scalco = 1; every other byte zero."""
    raw = np.fromfile(str(path), np.uint8).reshape(nshots * ntr, 240 + 4 * nt)
    hdt = int(np.float32(dt) * np.float32(1000000))
    for t in range(nshots * ntr):
        j, k = divmod(t, ntr)
        sx, gx = int(round(sxs[j] * dx)), int(round(k * dx))
        want = np.zeros(240, np.uint8)

        def put(off, val, fmt):
            want[off:off + np.dtype(fmt).itemsize] = np.frombuffer(np.array([val], fmt).tobytes(), np.uint8)
        for off, val in ((0, 0), (8, j + 1), (12, k), (16, j + 1), (20, k), (36, gx - sx), (72, sx), (76, 0),
                         (80, gx), (84, 0)):
            put(off, val, "<i4")
        for off, val in ((28, 1), (68, 1), (70, 1)):
            put(off, val, "<i2")
        put(114, nt, "<u2")
        put(116, hdt, "<u2")
        assert np.array_equal(raw[t, :240], want), (t, np.nonzero(raw[t, :240] != want))


def _read_su(path, ntr, nt):
    raw = np.fromfile(str(path), np.uint8).reshape(ntr, 240 + 4 * nt)
    return raw[:, 240:].copy().view(np.float32).reshape(ntr, nt)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built")
def test_driver_plain_cg_matches_library(tmp_path):
    """Plain mode: the SU data the driver writes equal the library's Born
    records, and its 2-iteration CG image and residuals equal the library's
    on the data read back (relative L2 <= 1e-5 and 1e-4)."""
    from ds2d_amd.lib import ShotSpec, Solver, ricker
    par, img = _params(tmp_path, "plain", {})
    _run("synth", par)
    out = _run(par)
    assert "plain CG" in out
    nt, sxs = 400, [16, 48, 80]
    rec = PML + np.arange(NX)
    # synth numbers its shots from 0; the SU scan numbers them from 1 (shot id seeds the random boundary)
    synth = [ShotSpec(0, NX, PML + x, PML + 2, PML + 2, rec, shot_id=k, boundary=1) for k, x in enumerate(sxs)]
    inv = [ShotSpec(0, NX, PML + x, PML + 2, PML + 2, rec, shot_id=k + 1, boundary=1) for k, x in enumerate(sxs)]
    s = Solver(dx=10.0, dz=10.0, dt=0.001, nt=nt, allnx=NX, allnz=NZ, pml=PML, nx_max=NX, ntr_max=NX,
               wavelet=ricker(nt, 15.0, 0.001), physics=1, rho_const=1000.0)
    s.set_model(np.fromfile(str(tmp_path / "vp.bin"), np.float32).reshape(NX, NZ))
    _, d_lib, _ = s.born_forward(synth, np.fromfile(str(tmp_path / "refl.bin"), np.float32).reshape(NX, NZ))
    d_su = _read_su(tmp_path / "vz.su", 3 * NX, nt)
    d_lib = np.asarray(d_lib, np.float32).reshape(3 * NX, nt)
    assert np.linalg.norm(d_su - d_lib) <= 1e-5 * np.linalg.norm(d_lib)
    im_lib, res_lib, _ = s.lsrtm_cg(inv, d_su, 2)
    im = np.fromfile(str(img / f"cg_image_2_{NX}_{NZ}.dat"), np.float32).reshape(NX, NZ)
    res = [float(l.split("=")[-1]) for l in (img / "misfit.txt").read_text().strip().split("\n")]
    assert len(res) == 3 and res[2] < res[1] < res[0] == pytest.approx(1.0)
    np.testing.assert_allclose(res, res_lib, rtol=1e-4)
    assert np.linalg.norm(im - im_lib) <= 1e-4 * np.linalg.norm(im_lib)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built")
def test_driver_adam_matches_oracle_adam(tmp_path):
    """The driver's Adam (lsrtm_all_iteration, main.cpp:1092-1130) over three
    iterations equals the oracle's FP32 Adam applied to the summed gradients
    the driver wrote (all_grad_* = g + b_g + b_g2; adding zeros to the sum
    is exact), with the moments carried across iterations."""
    import oracle_lib as orc
    par, img = _params(tmp_path, "ref", {"opt": "adam", "batch": "3"}, maxiter=3)
    _run("synth", par)
    _run(par)
    A = NX * NZ
    z = np.zeros(A, np.float32)
    for comp in ("pp", "ps"):
        s0, r0, im = z.copy(), z.copy(), z.copy()
        for it in range(3):
            g = np.fromfile(str(img / f"all_grad_{comp}_{it}_{NX}_{NZ}.dat"), np.float32)
            got = np.fromfile(str(img / f"sum_image_{comp}_{it}_{NX}_{NZ}.dat"), np.float32)
            assert np.abs(g).max() > 0
            s0, r0, im = orc.adam_update(g, z, z, s0, r0, im, it, prec="f")
            err = float(np.linalg.norm(got - im) / np.linalg.norm(im))
            print(f"adam {comp} iter {it}: rel {err:.2e}, max |step| {np.abs(im).max():.3e}")
            assert err <= 1e-6, (comp, it)
