"""ctypes binding of the CPU oracle (oracle/libds2d_oracle.so) for tests.

Test infrastructure only: the oracle is the checker, never the product path.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "libds2d_oracle.so")
# the same source built with FMA contraction (a second valid FP32 order)
ORACLE_FMA = os.path.join(ROOT, "oracle", "libds2d_oracle_fma.so")
REF_LAPLACE = os.path.join(ROOT, "oracle", "_ref", "libref_laplace.so")

_lib = None
_lib_fma = None


class OrcShot(C.Structure):
    _fields_ = [("allnx", C.c_int), ("allnz", C.c_int), ("ngx_min", C.c_int), ("nx", C.c_int), ("pml", C.c_int),
                ("nt", C.c_int), ("dx", C.c_double), ("dz", C.c_double), ("dt", C.c_double), ("boundary", C.c_int),
                ("rr", C.c_double), ("rb_seed", C.c_ulonglong), ("rb_shot", C.c_int), ("rb_amp", C.c_double),
                ("rho_const", C.c_double), ("isx", C.c_int), ("isz", C.c_int), ("igz", C.c_int), ("ntr", C.c_int),
                ("rec_ix", C.POINTER(C.c_int))]


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE):
            raise RuntimeError("oracle not built: make -C oracle")
        _lib = C.CDLL(ORACLE)
        _lib.orc_rb_uniform.restype = C.c_double
        _lib.orc_rb_uniform.argtypes = [C.c_ulonglong, C.c_int, C.c_int, C.c_int]
    return _lib


def load_fma():
    global _lib_fma
    if _lib_fma is None:
        if not os.path.exists(ORACLE_FMA):
            raise RuntimeError("oracle not built: make -C oracle")
        _lib_fma = C.CDLL(ORACLE_FMA)
    return _lib_fma


def threads(n=0):
    return load().orc_threads(n)


def P(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


def dtype_of(prec):
    return np.float64 if prec == "d" else np.float32


def make_shot(spec, allnx, allnz, pml, nt, dx, dz, dt, *, rr=1e-3, rb_seed=1234, rb_amp=0.3, rho_const=0.0):
    """OrcShot from a ds2d_amd ShotSpec (keeps rec_ix alive on the struct).

    dx, dz and dt are rounded to FP32 first: the library's geometry holds
    them as floats, as the reference's does (FD2DGPU.cu:25 takes float dx1,
    dz1, dt1), and the FP64 oracle then steps with exactly those values.  A
    normalised image over 2500 steps moves by 1.5e-4 between dt = 0.0008 and
    its FP32 rounding (2e-11 apart), so this matters for the parity bar."""
    dx, dz, dt = (float(np.float32(v)) for v in (dx, dz, dt))
    rix = np.ascontiguousarray(spec.rec_ix, np.int32)
    s = OrcShot(allnx, allnz, spec.ngx_min, spec.nx, pml, nt, dx, dz, dt, spec.boundary, rr, rb_seed,
                spec.shot_id, rb_amp, rho_const, spec.isx, spec.isz, spec.igz, rix.size,
                rix.ctypes.data_as(C.POINTER(C.c_int)))
    s._keep = rix
    return s


def _c(a, prec):
    return None if a is None else np.ascontiguousarray(a, dtype_of(prec))


def elastic_forward(shot, vp, vs, rho, wav, prec="d", final=False):
    lib = load()
    t = dtype_of(prec)
    rz = np.zeros(shot.ntr * shot.nt, t)
    rx = np.zeros_like(rz)
    n = (shot.nx + 2 * shot.pml) * (shot.allnz + 2 * shot.pml)
    f5 = np.zeros(5 * n, t) if final else None
    vp, vs, rho, wav = _c(vp, prec), _c(vs, prec), _c(rho, prec), _c(wav, prec)
    getattr(lib, f"orc_elastic_forward_{prec}")(C.byref(shot), P(vp), P(vs), P(rho), P(wav), P(rz), P(rx), P(f5))
    return (rz, rx, f5) if final else (rz, rx)


def elsrtm_shot(shot, vp, vs, rho, wav, img_pp, img_ps, obs_z, obs_x, it, col0, grads7=None, prec="d"):
    """prec "d" (FP64), "f" (FP32, separate multiply and add) or "f_fma"
    (FP32 with multiply-adds fused)."""
    lib = load()
    if prec == "f_fma":
        lib, prec = load_fma(), "f"
    t = dtype_of(prec)
    A = shot.allnx * shot.allnz
    if grads7 is None:
        grads7 = np.zeros(7 * A, t)
    mf = C.c_double(0.0)
    args = [_c(a, prec) for a in (vp, vs, rho, wav, img_pp, img_ps, obs_z, obs_x)]
    getattr(lib, f"orc_elsrtm_shot_{prec}")(C.byref(shot), *[P(a) for a in args], it, col0, P(grads7), C.byref(mf))
    return grads7, mf.value


def acoustic_ref_forward(shot, vel, rho, wav, sxnum, prec="d"):
    lib = load()
    rec = np.zeros(shot.nx * shot.nt, dtype_of(prec))
    vel, rho, wav = _c(vel, prec), _c(rho, prec), _c(wav, prec)
    getattr(lib, f"orc_acoustic_ref_forward_{prec}")(C.byref(shot), P(vel), P(rho), P(wav), sxnum, P(rec))
    return rec


def acou_forward(shot, vp, rho, refl, wav, prec="d", final=False):
    lib = load()
    t = dtype_of(prec)
    rbg = np.zeros(shot.ntr * shot.nt, t)
    rbn = np.zeros_like(rbg)
    n = (shot.nx + 2 * shot.pml) * (shot.allnz + 2 * shot.pml)
    f3 = np.zeros(3 * n, t) if final else None
    vp, rho, refl, wav = _c(vp, prec), _c(rho, prec), _c(refl, prec), _c(wav, prec)
    getattr(lib, f"orc_acou_forward_{prec}")(C.byref(shot), P(vp), P(rho), P(refl), P(wav), P(rbg), P(rbn), P(f3))
    return (rbg, rbn, f3) if final else (rbg, rbn)


def acou_adjoint(shot, vp, rho, wav, res, prec="d"):
    lib = load()
    t = dtype_of(prec)
    g = np.zeros(shot.allnx * shot.allnz, t)
    vp, rho, wav, res = _c(vp, prec), _c(rho, prec), _c(wav, prec), _c(res, prec)
    getattr(lib, f"orc_acou_adjoint_{prec}")(C.byref(shot), P(vp), P(rho), P(wav), P(res), P(g), None)
    return g


def ricker(nt, f, dt, prec="d"):
    s = np.zeros(nt, dtype_of(prec))
    getattr(load(), f"orc_ricker_{prec}")(nt, C.c_double(f), C.c_double(dt), P(s))
    return s


def pad_window(glob, ngx_min, nx, pml, prec="d"):
    allnx, allnz = glob.shape
    out = np.zeros((nx + 2 * pml, allnz + 2 * pml), dtype_of(prec))
    g = _c(glob, prec)
    getattr(load(), f"orc_pad_window_{prec}")(P(g), allnz, ngx_min, nx, allnz, pml, P(out))
    return out


def damp_profile(vp_pad, pml, rr, prec="d"):
    nxp, nzp = vp_pad.shape
    dx = np.zeros_like(vp_pad, dtype=dtype_of(prec))
    dz = np.zeros_like(dx)
    v = _c(vp_pad, prec)
    getattr(load(), f"orc_damp_profile_{prec}")(nxp, nzp, pml, P(v), C.c_double(rr), P(dx), P(dz))
    return dx, dz


def random_boundary(vp_pad, vs_pad, pml, seed, shot, amp, prec="d"):
    nxp, nzp = vp_pad.shape
    vp = _c(vp_pad, prec).copy()
    vs = _c(vs_pad, prec).copy()
    getattr(load(), f"orc_random_boundary_{prec}")(nxp, nzp, pml, C.c_ulonglong(seed), shot, C.c_double(amp),
                                                   P(vp), P(vs))
    return vp, vs


def laplace_order(img, dx, prec="d"):
    a = _c(img, prec).copy()
    nx, nz = a.shape
    getattr(load(), f"orc_laplace_order_{prec}")(nz, nx, C.c_double(dx), P(a))
    return a


def laplace_band(img, prec="d"):
    a = _c(img, prec).copy()
    nx, nz = a.shape
    getattr(load(), f"orc_laplace_band_{prec}")(nz, nx, P(a))
    return a


def elastic_run(lam, mu, rho, dampx, dampz, fields15, nsteps, direction, dx, dz, dt, prec="d"):
    nxp, nzp = lam.shape
    f = _c(fields15, prec).copy()
    args = [_c(a, prec) for a in (lam, mu, rho, dampx, dampz)]
    getattr(load(), f"orc_elastic_run_{prec}")(nxp, nzp, C.c_double(dx), C.c_double(dz), C.c_double(dt),
                                               *[P(a) for a in args], nsteps, direction, P(f))
    return f


def ncc_residual(syn_z, syn_x, obs_z, obs_x, ntr, nt, it, prec="d"):
    sz, sx = _c(syn_z, prec).copy(), _c(syn_x, prec).copy()
    oz, ox = _c(obs_z, prec), _c(obs_x, prec)
    mf = C.c_double(0.0)
    getattr(load(), f"orc_ncc_residual_{prec}")(ntr, nt, it, P(sz), P(sx), P(oz), P(ox), C.byref(mf))
    return sz, sx, mf.value


def adam_update(g, bg, bg2, s0, r0, img, it, prec="d"):
    s0, r0, img = _c(s0, prec).copy(), _c(r0, prec).copy(), _c(img, prec).copy()
    g, bg, bg2 = _c(g, prec), _c(bg, prec), _c(bg2, prec)
    getattr(load(), f"orc_adam_update_{prec}")(g.size, it, P(g), P(bg), P(bg2), P(s0), P(r0), P(img))
    return s0, r0, img


def rb_uniform(seed, shot, ix, iz):
    return load().orc_rb_uniform(seed, shot, ix, iz)


def ref_laplace_available():
    return os.path.exists(REF_LAPLACE)


def ref_laplace_order(img, dx):
    lib = C.CDLL(REF_LAPLACE)
    a = np.ascontiguousarray(img, np.float32).copy()
    nx, nz = a.shape
    lib.ref_laplace_order(nz, nx, 2, C.c_float(dx), a.ctypes.data_as(C.c_void_p))
    return a


def ref_cutM(img, zi, ze, xi, xe):
    lib = C.CDLL(REF_LAPLACE)
    a = np.ascontiguousarray(img, np.float32).copy()
    nx, nz = a.shape
    lib.ref_cutM(a.ctypes.data_as(C.c_void_p), nz, nx, zi, ze, xi, xe)
    return a


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def eborn_forward(shot, vp, vs, rho, img_pp, img_ps, wav, prec="d"):
    """Plain-L2 elastic Born map: (img_pp, img_ps) -> scattered (vz, vx) records."""
    t = dtype_of(prec)
    rz = np.zeros(shot.ntr * shot.nt, t)
    rx = np.zeros_like(rz)
    args = [_c(a, prec) for a in (vp, vs, rho, img_pp, img_ps, wav)]
    getattr(load(), f"orc_eborn_forward_{prec}")(C.byref(shot), *[P(a) for a in args], P(rz), P(rx))
    return rz, rx


def eborn_adjoint(shot, vp, vs, rho, wav, res_z, res_x, prec="d"):
    """Its exact transpose: residual records -> (g_pp, g_ps) global images."""
    t = dtype_of(prec)
    A = shot.allnx * shot.allnz
    gpp = np.zeros(A, t)
    gps = np.zeros(A, t)
    args = [_c(a, prec) for a in (vp, vs, rho, wav, res_z, res_x)]
    getattr(load(), f"orc_eborn_adjoint_{prec}")(C.byref(shot), *[P(a) for a in args], P(gpp), P(gps))
    return gpp, gps


# --- single kernels on padded arrays (pin tests) ---------------------------
KERNELS = {"update_v": 0, "update_s": 1, "recon_v": 2, "recon_s": 3, "adjoint_v": 4, "adjoint_s": 5,
           "add_born_source": 6, "add_born_source_adjoint": 7}
FIELDS15 = ["vx", "vz", "txx", "tzz", "txz", "vx_x", "vx_z", "vz_x", "vz_z", "txx_x", "txx_z", "tzz_x", "tzz_z",
            "txz_x", "txz_z"]


def kernel(name, f15, *, pml=0, dx=1.0, dz=1.0, dt=1.0, vp=None, vs=None, rho=None, lam=None, mu=None,
           dampx=None, dampz=None, dmp=None, dms=None, direction=1, src15=None, prec="d"):
    """Apply one oracle kernel to a (15, nxp, nzp) state; returns (state, strain4)."""
    f = _c(f15, prec).copy()
    _, nxp, nzp = f.shape
    st = np.zeros((4, nxp, nzp), dtype_of(prec))
    arrs = [_c(a, prec) for a in (vp, vs, rho, lam, mu, dampx, dampz, dmp, dms)]
    src = _c(src15, prec)
    getattr(load(), f"orc_kernel_{prec}")(KERNELS[name], nxp, nzp, pml, C.c_double(dx), C.c_double(dz),
                                          C.c_double(dt), *[P(a) for a in arrs], direction, P(src), P(f), P(st))
    return f, st


def edge_io(field, bands, pml, nt, it, save, prec="d"):
    f = _c(field, prec).copy()
    b = _c(bands, prec).copy()
    nxp, nzp = f.shape
    getattr(load(), f"orc_edge_io_{prec}")(nxp, nzp, pml, nt, P(f), P(b), it, 1 if save else 0)
    return f, b


def edge_band_size(nxp, nzp, pml, nt, N=4):
    nx, nz = nxp - 2 * pml, nzp - 2 * pml
    return 2 * nt * nx * 2 * N + 2 * nt * 2 * N * nz


def image(strain4, bk15, vp, vs, rho, pml, col0, allnx, allnz, pp=None, ps=None, il=None, prec="d"):
    t = dtype_of(prec)
    pp = np.zeros((allnx, allnz), t) if pp is None else _c(pp, prec).copy()
    ps = np.zeros((allnx, allnz), t) if ps is None else _c(ps, prec).copy()
    il = np.zeros((allnx, allnz), t) if il is None else _c(il, prec).copy()
    s4, b = _c(strain4, prec), _c(bk15, prec)
    _, nxp, nzp = s4.shape
    args = [_c(a, prec) for a in (vp, vs, rho)]
    getattr(load(), f"orc_image_{prec}")(nxp, nzp, pml, P(s4), P(b), *[P(a) for a in args], col0, allnz, P(pp),
                                         P(ps), P(il))
    return pp, ps, il


def acoustic_ref_damp(vt, pml, rr, prec="d"):
    v = _c(vt, prec)
    nxp, nzp = v.shape
    d = np.zeros_like(v)
    getattr(load(), f"orc_acoustic_ref_damp_{prec}")(nxp, nzp, pml, P(v), C.c_double(rr), P(d))
    return d


def acoustic_ref_step(k, rr, damp, state5, dx, dz, dt, prec="d"):
    """state5 = (px, pz, p, vx, vz), each (nxp, nzp); one FD2DGPU.cu:253-325 step."""
    s = _c(state5, prec).copy()
    _, nxp, nzp = s.shape
    args = [_c(a, prec) for a in (k, rr, damp)]
    getattr(load(), f"orc_acoustic_ref_step_{prec}")(nxp, nzp, C.c_double(dx), C.c_double(dz), C.c_double(dt),
                                                     *[P(a) for a in args], *[P(s[i]) for i in range(5)])
    return s


def elsrtm_born(shot, vp, vs, rho, wav, img_pp, img_ps, col0, order, prec="d"):
    """Compat Born records: Lm (order 1) or Lm + Lm2 (order 2)."""
    t = dtype_of(prec)
    rz = np.zeros(shot.ntr * shot.nt, t)
    rx = np.zeros_like(rz)
    args = [_c(a, prec) for a in (vp, vs, rho, wav, img_pp, img_ps)]
    getattr(load(), f"orc_elsrtm_born_{prec}")(C.byref(shot), *[P(a) for a in args], col0, order, P(rz), P(rx))
    return rz, rx


# --- P/S-decoupled rotation propagator and decomposed imaging ----------------

ROT_FIELDS = ("th", "om", "th_x", "th_z", "om_x", "om_z", "vx", "vz", "vpx", "vpz", "vsx", "vsz")
DECOMP_FLOATS = {0: 4, 1: 2, 2: 2, 3: 2, 4: 1, 5: 1, 6: 1}


def rot_forward(shot, vp, vs, wav, prec="d", final=False):
    """Rotation forward modelling: (rec_z, rec_x[, final theta, omega, vx, vz])."""
    lib = load()
    t = dtype_of(prec)
    rz = np.zeros(shot.ntr * shot.nt, t)
    rx = np.zeros_like(rz)
    n = (shot.nx + 2 * shot.pml) * (shot.allnz + 2 * shot.pml)
    f4 = np.zeros(4 * n, t) if final else None
    vp, vs, wav = _c(vp, prec), _c(vs, prec), _c(wav, prec)
    getattr(lib, f"orc_rot_forward_{prec}")(C.byref(shot), P(vp), P(vs), P(wav), P(rz), P(rx), P(f4))
    return (rz, rx, f4) if final else (rz, rx)


def rot_reconstruct(shot, vp, vs, wav, nback, prec="d"):
    """Source side of the rotation RTM after reversing the last nback steps:
    [4][nxp*nzp] theta, omega, vx, vz."""
    lib = load()
    n = (shot.nx + 2 * shot.pml) * (shot.allnz + 2 * shot.pml)
    out = np.zeros(4 * n, dtype_of(prec))
    vp, vs, wav = _c(vp, prec), _c(vs, prec), _c(wav, prec)
    getattr(lib, f"orc_rot_reconstruct_{prec}")(C.byref(shot), P(vp), P(vs), P(wav), int(nback), P(out))
    return out


def rot_rtm(shot, vp, vs, wav, obs_z, obs_x, img_pp=None, img_ps=None, prec="d"):
    """P/S images of one rotation RTM shot, stacked into global [allnx][allnz] images."""
    lib = load()
    t = dtype_of(prec)
    A = shot.allnx * shot.allnz
    img_pp = np.zeros(A, t) if img_pp is None else img_pp
    img_ps = np.zeros(A, t) if img_ps is None else img_ps
    args = [_c(a, prec) for a in (vp, vs, wav, obs_z, obs_x)]
    getattr(lib, f"orc_rot_rtm_{prec}")(C.byref(shot), *[P(a) for a in args], P(img_pp), P(img_ps))
    return img_pp, img_ps


def decomp_rtm(shot, vp, vs, rho, wav, kind, obs_z, img=None, prec="d"):
    """Decomposed elastic RTM image of one shot (kind 0..6), stacked into a global image."""
    lib = load()
    t = dtype_of(prec)
    img = np.zeros(shot.allnx * shot.allnz, t) if img is None else img
    args = [_c(a, prec) for a in (vp, vs, rho, wav)]
    oz = _c(obs_z, prec)
    rc = getattr(lib, f"orc_decomp_rtm_{prec}")(C.byref(shot), *[P(a) for a in args], int(kind), P(oz), P(img))
    if rc:
        raise ValueError(f"bad image kind {kind}")
    return img


def rot_run(fields12, vp, vs, dampx, dampz, nsteps, direction, dx, dz, dt, prec="d"):
    """nsteps rotation steps on a bare [nxp][nzp] grid; fields12 [12][nxp][nzp] in ROT_FIELDS order."""
    lib = load()
    f = np.ascontiguousarray(fields12, dtype_of(prec)).copy()
    nxp, nzp = f.shape[1], f.shape[2]
    args = [_c(a, prec) for a in (vp, vs, dampx, dampz)]
    getattr(lib, f"orc_rot_run_{prec}")(nxp, nzp, C.c_double(dx), C.c_double(dz), C.c_double(dt),
                                        *[P(a) for a in args], int(nsteps), int(direction), P(f))
    return f


def helmholtz(vx, vz, dx, dz, prec="d"):
    lib = load()
    vx, vz = _c(vx, prec), _c(vz, prec)
    VP, VS = np.zeros_like(vx), np.zeros_like(vx)
    getattr(lib, f"orc_helmholtz_{prec}")(vx.shape[0], vx.shape[1], C.c_double(dx), C.c_double(dz), P(vx), P(vz),
                                          P(VP), P(VS))
    return VP, VS


def vector(VP, VS, dx, dz, prec="d"):
    """(vecP, vecS), each [nxp][nzp][2] (x, z components)."""
    lib = load()
    VP, VS = _c(VP, prec), _c(VS, prec)
    vp2, vs2 = np.zeros(VP.shape + (2,), VP.dtype), np.zeros(VP.shape + (2,), VP.dtype)
    getattr(lib, f"orc_vector_{prec}")(VP.shape[0], VP.shape[1], C.c_double(dx), C.c_double(dz), P(VP), P(VS),
                                       P(vp2), P(vs2))
    return vp2, vs2


def poynting_split(txz, tzz, vx, vz, vec2, prec="d"):
    lib = load()
    a = [_c(x, prec) for x in (txz, tzz, vx, vz, vec2)]
    up, down = np.zeros_like(a[4]), np.zeros_like(a[4])
    getattr(lib, f"orc_poynting_split_{prec}")(C.c_longlong(a[0].size), *[P(x) for x in a], P(up), P(down))
    return up, down


def image_cond(kind, fw, bw, prec="d"):
    lib = load()
    fw, bw = _c(fw, prec), _c(bw, prec)
    n = fw.size // DECOMP_FLOATS[kind]
    out = np.zeros(n, fw.dtype)
    rc = getattr(lib, f"orc_image_cond_{prec}")(int(kind), C.c_longlong(n), P(fw), P(bw), P(out))
    if rc:
        raise ValueError(f"bad image kind {kind}")
    return out
