"""ctypes binding of libds2d.so (include/ds2d.h).

The product path is the HIP library only: importing this module raises if the
library has not been built, and every call raises on a non-zero status with
the library's own error message.  There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# DS2D_LIB selects another build of the same library (e.g. a kernel variant under test)
LIB_PATH = os.environ.get("DS2D_LIB") or os.path.join(HERE, "libds2d.so")

_lib = None


class DS2DError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


ESHOT = -5  # some shots of a batch failed (non-finite); see Solver.shot_status


class Geom(C.Structure):
    _fields_ = [("dx", C.c_float), ("dz", C.c_float), ("dt", C.c_float),
                ("nxpml", C.c_int), ("nzpml", C.c_int), ("allnx", C.c_int), ("allnz", C.c_int),
                ("scale", C.c_int), ("pml", C.c_int), ("nt", C.c_int), ("nop", C.c_int), ("ntr", C.c_int),
                ("physics", C.c_int), ("pml_rr", C.c_float), ("rho_const", C.c_float), ("rb_amp", C.c_float),
                ("rb_seed", C.c_ulonglong), ("bg_mode", C.c_int), ("hbm_budget_gb", C.c_float)]


class Shot(C.Structure):
    _fields_ = [("ngx_min", C.c_int), ("nx", C.c_int), ("isx", C.c_int), ("isz", C.c_int), ("igz", C.c_int),
                ("ntr", C.c_int), ("rec_ix", C.POINTER(C.c_int)), ("shot_id", C.c_int), ("boundary", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("elapsed_ms", C.c_double), ("stencil_ms", C.c_double), ("cell_updates", C.c_longlong),
                ("stencil_launches", C.c_longlong), ("alg_bytes", C.c_double), ("kernel_avg_ms", C.c_double),
                ("kernel_samples", C.c_longlong), ("aux_alg_bytes", C.c_double), ("aux_kernel_avg_ms", C.c_double),
                ("blocked_bytes", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# name -> (restype, argtypes); the exported surface of include/ds2d.h
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_H = C.c_void_p
class UnetCfg(C.Structure):
    _fields_ = [("levels", C.c_int), ("base", C.c_int), ("max_batch", C.c_int), ("max_h", C.c_int),
                ("max_w", C.c_int), ("seed", C.c_ulonglong)]


class UnetStats(C.Structure):
    _fields_ = [("elapsed_ms", C.c_double), ("ms_per_step", C.c_double), ("flops_per_step", C.c_double),
                ("steps", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SIGNATURES = {
    "ds2d_last_error": (C.c_char_p, []),
    "ds2d_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), _D]),
    "ds2d_create": (C.c_int, [C.POINTER(Geom), _F, C.POINTER(_H)]),
    "ds2d_destroy": (C.c_int, [_H]),
    "ds2d_set_model": (C.c_int, [_H, _F, _F, _F]),
    "ds2d_set_image": (C.c_int, [_H, _F, _F]),
    "ds2d_set_shot": (C.c_int, [_H, C.POINTER(Shot), _F, _F]),
    "ds2d_shot_gradient": (C.c_int, [_H, C.c_int, _D]),
    "ds2d_batch_gradient": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _F, _F, C.c_int, _D]),
    "ds2d_take_gradients": (C.c_int, [_H, _F, _F, _F, _F, _F, _F, _F]),
    "ds2d_shot_status": (C.c_int, [_H, C.c_int, C.POINTER(C.c_int)]),
    "ds2d_set_time_block": (C.c_int, [_H, C.c_int]),
    "ds2d_set_chunk_plan": (C.c_int, [_H, C.c_int]),
    "ds2d_set_fused_born": (C.c_int, [_H, C.c_int, C.c_int, C.c_int]),
    "ds2d_set_fused_rotation": (C.c_int, [_H, C.c_int]),
    "ds2d_get_chunk_plan": (C.c_int, [_H, C.POINTER(C.c_int)]),
    "ds2d_set_stop": (C.c_int, [_H, C.c_double]),
    "ds2d_set_log": (C.c_int, [_H, C.c_int]),
    "ds2d_get_time_block": (C.c_int, [_H, C.POINTER(C.c_int)]),
    "ds2d_forward": (C.c_int, [_H, _F, _F]),
    "ds2d_batch_forward": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _F, _F, C.c_int, C.POINTER(Stats)]),
    "ds2d_set_chunk": (C.c_int, [_H, C.c_int, C.c_int, C.c_int]),
    "ds2d_set_graph": (C.c_int, [_H, C.c_int]),
    "ds2d_autotune_forward": (C.c_int, [_H, C.c_int, C.POINTER(Shot), C.c_int, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int), C.POINTER(C.c_int), _D]),
    "ds2d_acoustic_ref_forward": (C.c_int, [_H, C.c_int, C.c_int, _F]),
    "ds2d_born_forward": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _F, _F, _F, C.POINTER(Stats)]),
    "ds2d_born_adjoint": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _F, _F, C.POINTER(Stats)]),
    "ds2d_lsrtm_cg": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _F, C.c_int, _F, _D, C.POINTER(Stats)]),
    "ds2d_eborn_forward": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _F, _F, _F, _F, C.POINTER(Stats)]),
    "ds2d_eborn_adjoint": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _F, _F, _F, _F, C.POINTER(Stats)]),
    "ds2d_elsrtm_cg": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _F, _F, C.c_int, _F, _F, _D, C.POINTER(Stats)]),
    "ds2d_eborn_line_search": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _F, _F, _F, _F, _F, _F, _D, _D]),
    "ds2d_lsrtm_learned": (C.c_int, [_H, _H, C.c_int, C.POINTER(Shot), _F, C.c_int, _F, _D, C.POINTER(C.c_int),
                                     C.POINTER(Stats)]),
    "ds2d_elsrtm_learned": (C.c_int, [_H, _H, C.c_int, C.POINTER(Shot), _F, _F, C.c_int, _F, _F, _D,
                                      C.POINTER(C.c_int), C.POINTER(Stats)]),
    "ds2d_set_precision": (C.c_int, [_H, C.c_int]),
    "ds2d_born": (C.c_int, [_H, C.c_int, C.POINTER(Shot), C.c_int, _F, _F]),
    "ds2d_calculate_alpha": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _F, _F, _F, _F, _D, _D]),
    "ds2d_set_batch": (C.c_int, [_H, C.c_int]),
    "ds2d_set_fused": (C.c_int, [_H, C.c_int]),
    "ds2d_unet_create": (C.c_int, [C.POINTER(UnetCfg), C.POINTER(_H)]),
    "ds2d_unet_destroy": (C.c_int, [_H]),
    "ds2d_unet_num_params": (C.c_int, [_H, C.POINTER(C.c_longlong)]),
    "ds2d_unet_get_params": (C.c_int, [_H, _F]),
    "ds2d_unet_set_params": (C.c_int, [_H, _F]),
    "ds2d_unet_forward": (C.c_int, [_H, C.c_int, C.c_int, C.c_int, _F, _F, C.POINTER(UnetStats)]),
    "ds2d_unet_loss_grad": (C.c_int, [_H, C.c_int, C.c_int, C.c_int, _F, _F, _D, _F]),
    "ds2d_unet_set_data": (C.c_int, [_H, C.c_int, C.c_int, C.c_int, _F, _F]),
    "ds2d_unet_train": (C.c_int, [_H, C.c_int, C.c_int, C.c_float, C.c_ulonglong, _F, C.POINTER(UnetStats)]),
    "ds2d_conv3x3_bf16": (C.c_int, [C.c_int] * 8 + [_F, _F, _F, _F, _F, C.c_int, C.c_int, _F, _D]),
    "ds2d_sjcg_update": (C.c_int, [C.c_int, _F, _F, _F, _F, C.c_int, _F, _F]),
    "ds2d_laplace_filter": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_float, _F]),
    "ds2d_laplace_band": (C.c_int, [C.c_int, C.c_int, _F]),
    "ds2d_laplace_band_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ds2d_ricker": (C.c_int, [C.c_int, C.c_float, C.c_float, _F]),
    "ds2d_rot_forward": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _F, _F, C.POINTER(Stats)]),
    "ds2d_rot_rtm": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _F, _F, _F, _F]),
    "ds2d_decomp_rtm": (C.c_int, [_H, C.c_int, C.POINTER(Shot), C.c_int, _F, _F]),
}


def load(path=LIB_PATH):
    """Load libds2d.so; raise if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise DS2DError(f"{path} not built: run `python -m ds2d_amd.build` (the HIP library is required)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise DS2DError(f"ds2d error {rc}: {_lib.ds2d_last_error().decode()}", rc)


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "float32 C-contiguous array required"
    return a.ctypes.data_as(_F)


def ricker(nt, f0, dt):
    lib = load()
    out = np.zeros(nt, np.float32)
    _check(lib.ds2d_ricker(nt, f0, dt, fptr(out)))
    return out


def laplace_filter(img, dx, order=2):
    """laplace_filter_2d(nz, nx, order, dx, dat) on an [nx][nz] image (copy)."""
    lib = load()
    a = np.ascontiguousarray(img, np.float32).copy()
    nx, nz = a.shape
    _check(lib.ds2d_laplace_filter(nz, nx, order, dx, fptr(a)))
    return a


def laplace_band(img):
    lib = load()
    a = np.ascontiguousarray(img, np.float32).copy()
    nx, nz = a.shape
    _check(lib.ds2d_laplace_band(nz, nx, fptr(a)))
    return a


def device_info():
    lib = load()
    name = C.create_string_buffer(64)
    ncu = C.c_int(0)
    mem = C.c_double(0)
    _check(lib.ds2d_device_info(name, 64, C.byref(ncu), C.byref(mem)))
    return name.value.decode(), ncu.value, mem.value


class ShotSpec:
    """One shot window in padded-window coordinates (see ds2d_shot)."""

    def __init__(self, ngx_min, nx, isx, isz, igz, rec_ix, shot_id=0, boundary=0):
        self.ngx_min, self.nx, self.isx, self.isz, self.igz = int(ngx_min), int(nx), int(isx), int(isz), int(igz)
        self.rec_ix = np.ascontiguousarray(rec_ix, np.int32)
        self.shot_id, self.boundary = int(shot_id), int(boundary)

    @property
    def ntr(self):
        return int(self.rec_ix.size)


def _shots_array(shots):
    arr = (Shot * len(shots))()
    for i, s in enumerate(shots):
        arr[i] = Shot(s.ngx_min, s.nx, s.isx, s.isz, s.igz, s.ntr, s.rec_ix.ctypes.data_as(C.POINTER(C.c_int)),
                      s.shot_id, s.boundary)
    return arr


class Solver:
    """Owning wrapper of a ds2d handle (FD2DGPU_ELASTIC counterpart)."""

    def __init__(self, *, dx, dz, dt, nt, allnx, allnz, pml, nx_max, ntr_max=0, wavelet, physics=0,
                 pml_rr=1e-3, rho_const=0.0, rb_amp=0.3, rb_seed=1234, bg_mode=0, hbm_budget_gb=64.0):
        self.lib = load()
        g = Geom(dx, dz, dt, nx_max + 2 * pml, allnz + 2 * pml, allnx, allnz, 1, pml, nt, 4, ntr_max, physics,
                 pml_rr, rho_const, rb_amp, rb_seed, bg_mode, hbm_budget_gb)
        self.geom = g
        self.nt, self.allnx, self.allnz, self.pml = nt, allnx, allnz, pml
        self.wavelet = np.ascontiguousarray(wavelet, np.float32)
        h = C.c_void_p()
        _check(self.lib.ds2d_create(C.byref(g), fptr(self.wavelet), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.ds2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_model(self, vp, vs=None, rho=None):
        self._vp = np.ascontiguousarray(vp, np.float32)
        self._vs = None if vs is None else np.ascontiguousarray(vs, np.float32)
        self._rho = None if rho is None else np.ascontiguousarray(rho, np.float32)
        _check(self.lib.ds2d_set_model(self.h, fptr(self._vp), fptr(self._vs), fptr(self._rho)))

    def set_image(self, pp, ps=None):
        pp = np.ascontiguousarray(pp, np.float32)
        ps = None if ps is None else np.ascontiguousarray(ps, np.float32)
        _check(self.lib.ds2d_set_image(self.h, fptr(pp), fptr(ps)))

    def _rec(self, shots):
        return np.zeros(sum(s.ntr for s in shots) * self.nt, np.float32)

    def set_chunk(self, width, threads=0, depth=0):
        """Fused-kernel tile: x-chunk width, workgroup lanes, prefetch depth (0 = defaults)."""
        _check(self.lib.ds2d_set_chunk(self.h, int(width), int(threads), int(depth)))

    def set_time_block(self, steps):
        """Fused elastic forward: 1 or 2 time steps per launch (temporal blocking)."""
        _check(self.lib.ds2d_set_time_block(self.h, int(steps)))

    def set_chunk_plan(self, on):
        """Two steps per launch: balanced chunk plan across XCDs/CUs (True, default) or fixed-width chunks."""
        _check(self.lib.ds2d_set_chunk_plan(self.h, 1 if on else 0))

    def set_fused_born(self, on, width=0, threads=0):
        """Elastic Born forward pass: one fused launch per step (True, default) or the two per-step kernels."""
        _check(self.lib.ds2d_set_fused_born(self.h, 1 if on else 0, int(width), int(threads)))

    def set_fused_rotation(self, on):
        """Rotation forward: one fused launch per step (True, default) or k_rot_v + k_rot_p."""
        _check(self.lib.ds2d_set_fused_rotation(self.h, 1 if on else 0))

    def chunk_plan_used(self):
        """Whether the last two-step launch list was the balanced plan."""
        v = C.c_int(0)
        _check(self.lib.ds2d_get_chunk_plan(self.h, C.byref(v)))
        return bool(v.value)

    def time_block(self):
        v = C.c_int(0)
        _check(self.lib.ds2d_get_time_block(self.h, C.byref(v)))
        return v.value

    def set_graph(self, on):
        _check(self.lib.ds2d_set_graph(self.h, 1 if on else 0))

    def autotune(self, shots, nsteps=100):
        """Pick the forward step's tile for this batch; returns (width, threads, depth, ms);
        width 0 means the two-kernel step won."""
        w, t, d = C.c_int(0), C.c_int(0), C.c_int(0)
        ms = C.c_double(0)
        _check(self.lib.ds2d_autotune_forward(self.h, len(shots), _shots_array(shots), nsteps, C.byref(w),
                                              C.byref(t), C.byref(d), C.byref(ms)))
        return w.value, t.value, d.value, ms.value

    def batch_forward(self, shots, mode=2, fetch=True):
        """Batched elastic forward; fetch=False leaves the records in HBM."""
        rz, rx = (self._rec(shots), self._rec(shots)) if fetch else (None, None)
        st = Stats()
        arr = _shots_array(shots)
        _check(self.lib.ds2d_batch_forward(self.h, len(shots), arr, fptr(rz), fptr(rx), mode, C.byref(st)))
        return rz, rx, st.as_dict()

    def rot_forward(self, shots, fetch=True):
        """P/S-decoupled (theta/omega) forward modelling: (rec_z, rec_x, stats)."""
        rz, rx = (self._rec(shots), self._rec(shots)) if fetch else (None, None)
        st = Stats()
        _check(self.lib.ds2d_rot_forward(self.h, len(shots), _shots_array(shots), fptr(rz), fptr(rx), C.byref(st)))
        return rz, rx, st.as_dict()

    def rot_rtm(self, shots, rec_z, rec_x):
        """P/S-decomposed rotation RTM (random-boundary shots): (PP, PS) images [allnx][allnz]."""
        rz = np.ascontiguousarray(rec_z, np.float32)
        rx = np.ascontiguousarray(rec_x, np.float32)
        pp = np.zeros((self.allnx, self.allnz), np.float32)
        ps = np.zeros_like(pp)
        _check(self.lib.ds2d_rot_rtm(self.h, len(shots), _shots_array(shots), fptr(rz), fptr(rx), fptr(pp),
                                     fptr(ps)))
        return pp, ps

    def decomp_rtm(self, shots, kind, rec_z):
        """Elastic RTM with a decomposed imaging condition (kind 0 Poynting PP; 1-3 vector PP/PS/SS;
        4-6 scalar PP/PS/SS): image [allnx][allnz]."""
        rz = np.ascontiguousarray(rec_z, np.float32)
        img = np.zeros((self.allnx, self.allnz), np.float32)
        _check(self.lib.ds2d_decomp_rtm(self.h, len(shots), _shots_array(shots), int(kind), fptr(rz), fptr(img)))
        return img

    def batch_gradient(self, shots, obs_z, obs_x, it):
        mf = C.c_double(0.0)
        arr = _shots_array(shots)
        oz = np.ascontiguousarray(obs_z, np.float32)
        ox = np.ascontiguousarray(obs_x, np.float32)
        _check(self.lib.ds2d_batch_gradient(self.h, len(shots), arr, fptr(oz), fptr(ox), it, C.byref(mf)))
        return mf.value

    def shot_status(self, nshots):
        """Per shot of the last reference-compat gradient: 0 included, 1 left out (non-finite)."""
        st = (C.c_int * nshots)()
        _check(self.lib.ds2d_shot_status(self.h, nshots, st))
        return list(st)

    def set_precision(self, source_fp64):
        """Reference-compat gradient: background and first-order Born field in FP64 (1) or FP32 (0)."""
        _check(self.lib.ds2d_set_precision(self.h, 1 if source_fp64 else 0))

    def born(self, shots, order=2):
        """Reference Born records of the handle's image: Lm (order 1) or Lm + Lm2 (order 2)."""
        rz, rx = self._rec(shots), self._rec(shots)
        _check(self.lib.ds2d_born(self.h, len(shots), _shots_array(shots), int(order), fptr(rz), fptr(rx)))
        return rz, rx

    def calculate_alpha(self, shots, g_pp, g_ps, obs_z, obs_x):
        """ELSRTM_CALCULATE_ALPHA: (num, den) = (<Lg, Lm - d>, <Lg, Lg>) with m the handle's image."""
        num, den = C.c_double(0.0), C.c_double(0.0)
        arrs = [np.ascontiguousarray(a, np.float32) for a in (g_pp, g_ps, obs_z, obs_x)]
        _check(self.lib.ds2d_calculate_alpha(self.h, len(shots), _shots_array(shots), *[fptr(a) for a in arrs],
                                             C.byref(num), C.byref(den)))
        return num.value, den.value

    def set_shot(self, shot, rec_z=None, rec_x=None):
        """ds2d_set_shot: the current shot (main.cpp:994-1025) and its observed records."""
        self._shot_keep = _shots_array([shot])
        rz = None if rec_z is None else np.ascontiguousarray(rec_z, np.float32)
        rx = None if rec_x is None else np.ascontiguousarray(rec_x, np.float32)
        _check(self.lib.ds2d_set_shot(self.h, self._shot_keep, fptr(rz), fptr(rx)))
        self._cur_ntr = shot.ntr

    def shot_gradient(self, it):
        """ds2d_shot_gradient on the current shot; returns the NCC misfit increment."""
        mf = C.c_double(0.0)
        _check(self.lib.ds2d_shot_gradient(self.h, int(it), C.byref(mf)))
        return mf.value

    def forward(self):
        """ds2d_forward: elastic forward records (vz, vx) of the current shot."""
        rz = np.zeros(self._cur_ntr * self.nt, np.float32)
        rx = np.zeros_like(rz)
        _check(self.lib.ds2d_forward(self.h, fptr(rz), fptr(rx)))
        return rz, rx

    def take_gradients(self):
        A = self.allnx * self.allnz
        outs = [np.zeros(A, np.float32) for _ in range(7)]
        _check(self.lib.ds2d_take_gradients(self.h, *[fptr(o) for o in outs]))
        return [o.reshape(self.allnx, self.allnz) for o in outs]

    def acoustic_ref_forward(self, sxnum, ngx_min):
        nx = self.geom.nxpml - 2 * self.pml
        rec = np.zeros(nx * self.nt, np.float32)
        _check(self.lib.ds2d_acoustic_ref_forward(self.h, sxnum, ngx_min, fptr(rec)))
        return rec.reshape(nx, self.nt)

    def born_forward(self, shots, refl):
        refl = np.ascontiguousarray(refl, np.float32)
        rbg, rbn = self._rec(shots), self._rec(shots)
        st = Stats()
        _check(self.lib.ds2d_born_forward(self.h, len(shots), _shots_array(shots), fptr(refl), fptr(rbg),
                                          fptr(rbn), C.byref(st)))
        return rbg, rbn, st.as_dict()

    def born_adjoint(self, shots, res):
        res = np.ascontiguousarray(res, np.float32)
        g = np.zeros(self.allnx * self.allnz, np.float32)
        st = Stats()
        _check(self.lib.ds2d_born_adjoint(self.h, len(shots), _shots_array(shots), fptr(res), fptr(g),
                                          C.byref(st)))
        return g.reshape(self.allnx, self.allnz), st.as_dict()

    def eborn_forward(self, shots, img_pp, img_ps):
        """Plain-L2 elastic Born map: (PP, PS) images -> scattered (vz, vx) records."""
        pp = np.ascontiguousarray(img_pp, np.float32)
        ps = np.ascontiguousarray(img_ps, np.float32)
        rz, rx = self._rec(shots), self._rec(shots)
        st = Stats()
        _check(self.lib.ds2d_eborn_forward(self.h, len(shots), _shots_array(shots), fptr(pp), fptr(ps), fptr(rz),
                                           fptr(rx), C.byref(st)))
        return rz, rx, st.as_dict()

    def eborn_adjoint(self, shots, res_z, res_x):
        """Exact adjoint: residual records -> (PP, PS) gradients (allnx x allnz)."""
        rz = np.ascontiguousarray(res_z, np.float32)
        rx = np.ascontiguousarray(res_x, np.float32)
        gpp = np.zeros(self.allnx * self.allnz, np.float32)
        gps = np.zeros_like(gpp)
        st = Stats()
        _check(self.lib.ds2d_eborn_adjoint(self.h, len(shots), _shots_array(shots), fptr(rz), fptr(rx), fptr(gpp),
                                           fptr(gps), C.byref(st)))
        return gpp.reshape(self.allnx, self.allnz), gps.reshape(self.allnx, self.allnz), st.as_dict()

    def eborn_line_search(self, shots, img_pp, img_ps, g_pp, g_ps, d_z, d_x):
        """(num, den): <L g, L m - d> and <L g, L g> (the reference's line-search sums)."""
        a = [np.ascontiguousarray(x, np.float32) for x in (img_pp, img_ps, g_pp, g_ps, d_z, d_x)]
        num, den = C.c_double(0), C.c_double(0)
        _check(self.lib.ds2d_eborn_line_search(self.h, len(shots), _shots_array(shots), *[fptr(x) for x in a],
                                               C.byref(num), C.byref(den)))
        return num.value, den.value

    def elsrtm_cg(self, shots, d_z, d_x, niter, pp0=None, ps0=None):
        """CGLS PP/PS LSRTM on the elastic Born pair: (pp, ps, residual history, stats)."""
        A = self.allnx * self.allnz
        pp = np.zeros(A, np.float32) if pp0 is None else np.ascontiguousarray(pp0, np.float32).ravel().copy()
        ps = np.zeros(A, np.float32) if ps0 is None else np.ascontiguousarray(ps0, np.float32).ravel().copy()
        res = np.zeros(niter + 1, np.float64)
        st = Stats()
        dz = np.ascontiguousarray(d_z, np.float32)
        dx = np.ascontiguousarray(d_x, np.float32)
        _check(self.lib.ds2d_elsrtm_cg(self.h, len(shots), _shots_array(shots), fptr(dz), fptr(dx), niter,
                                       fptr(pp), fptr(ps), res.ctypes.data_as(_D), C.byref(st)))
        return pp.reshape(self.allnx, self.allnz), ps.reshape(self.allnx, self.allnz), res, st.as_dict()

    def set_fused(self, mode):
        """Fused Born/adjoint steps of the resident-history passes: -1 auto (timed), 0 never,
        1 always (x width 64), k in [8, 1024] always with x width k."""
        _check(self.lib.ds2d_set_fused(self.h, int(mode)))

    def set_stop(self, target):
        """CG solvers stop once the relative data residual is at most target (0: never)."""
        _check(self.lib.ds2d_set_stop(self.h, float(target)))

    def set_log(self, every):
        """CG solvers print their residual to stderr every `every` iterations (0: silent)."""
        _check(self.lib.ds2d_set_log(self.h, int(every)))

    def set_batch(self, shots_per_launch):
        """Shots per time loop for the plain Born/adjoint/CG entries (0 = all)."""
        _check(self.lib.ds2d_set_batch(self.h, int(shots_per_launch)))

    def lsrtm_learned(self, unet, shots, d_obs, niter, image0=None):
        """LSRTM with the U-Net preconditioner (plain CGLS fallback)."""
        img = np.zeros(self.allnx * self.allnz, np.float32) if image0 is None else \
            np.ascontiguousarray(image0, np.float32).ravel().copy()
        res = np.zeros(niter + 1, np.float64)
        st = Stats()
        acc = C.c_int(0)
        d = np.ascontiguousarray(d_obs, np.float32)
        _check(self.lib.ds2d_lsrtm_learned(self.h, unet.h, len(shots), _shots_array(shots), fptr(d), niter,
                                           fptr(img), res.ctypes.data_as(_D), C.byref(acc), C.byref(st)))
        out = st.as_dict()
        out["learned_steps"] = acc.value
        return img.reshape(self.allnx, self.allnz), res, out

    def elsrtm_learned(self, unet, shots, d_z, d_x, niter, pp0=None, ps0=None):
        """PP/PS LSRTM with the U-Net preconditioner (plain CGLS fallback)."""
        A = self.allnx * self.allnz
        pp = np.zeros(A, np.float32) if pp0 is None else np.ascontiguousarray(pp0, np.float32).ravel().copy()
        ps = np.zeros(A, np.float32) if ps0 is None else np.ascontiguousarray(ps0, np.float32).ravel().copy()
        res = np.zeros(niter + 1, np.float64)
        st = Stats()
        acc = C.c_int(0)
        dz = np.ascontiguousarray(d_z, np.float32)
        dx = np.ascontiguousarray(d_x, np.float32)
        _check(self.lib.ds2d_elsrtm_learned(self.h, unet.h, len(shots), _shots_array(shots), fptr(dz), fptr(dx),
                                            niter, fptr(pp), fptr(ps), res.ctypes.data_as(_D), C.byref(acc),
                                            C.byref(st)))
        out = st.as_dict()
        out["learned_steps"] = acc.value
        return pp.reshape(self.allnx, self.allnz), ps.reshape(self.allnx, self.allnz), res, out

    def lsrtm_cg(self, shots, d_obs, niter, image0=None):
        img = np.zeros(self.allnx * self.allnz, np.float32) if image0 is None else \
            np.ascontiguousarray(image0, np.float32).ravel().copy()
        res = np.zeros(niter + 1, np.float64)
        st = Stats()
        d = np.ascontiguousarray(d_obs, np.float32)
        _check(self.lib.ds2d_lsrtm_cg(self.h, len(shots), _shots_array(shots), fptr(d), niter, fptr(img),
                                      res.ctypes.data_as(_D), C.byref(st)))
        return img.reshape(self.allnx, self.allnz), res, st.as_dict()


class Unet:
    """The learned LSRTM preconditioner (BF16 MFMA U-Net) of libds2d.so."""

    def __init__(self, *, levels=4, base=32, max_batch=8, max_h=512, max_w=256, seed=0):
        self.lib = load()
        self.levels, self.base = levels, base
        cfg = UnetCfg(levels, base, max_batch, max_h, max_w, seed)
        self.h = _H()
        _check(self.lib.ds2d_unet_create(C.byref(cfg), C.byref(self.h)))
        n = C.c_longlong(0)
        _check(self.lib.ds2d_unet_num_params(self.h, C.byref(n)))
        self.num_params = n.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.ds2d_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        th = np.zeros(self.num_params, np.float32)
        _check(self.lib.ds2d_unet_get_params(self.h, fptr(th)))
        return th

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        if th.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {th.size}")
        _check(self.lib.ds2d_unet_set_params(self.h, fptr(th)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        n, hh, ww = x.shape
        y = np.zeros_like(x)
        st = UnetStats()
        _check(self.lib.ds2d_unet_forward(self.h, n, hh, ww, fptr(x), fptr(y), C.byref(st)))
        return y, st.as_dict()

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        n, hh, ww = x.shape
        g = np.zeros(self.num_params, np.float32)
        loss = C.c_double(0)
        _check(self.lib.ds2d_unet_loss_grad(self.h, n, hh, ww, fptr(x), fptr(t), C.byref(loss), fptr(g)))
        return loss.value, g

    def set_data(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        n, hh, ww = x.shape
        _check(self.lib.ds2d_unet_set_data(self.h, n, hh, ww, fptr(x), fptr(t)))

    def train(self, nsteps, batch=8, lr=1e-3, seed=0):
        lh = np.zeros(nsteps, np.float32)
        st = UnetStats()
        _check(self.lib.ds2d_unet_train(self.h, nsteps, batch, lr, seed, fptr(lh), C.byref(st)))
        return lh, st.as_dict()


def conv3x3_bf16(op, x1, wt, *, x2=None, up=0, dy=None, bias=None, relu=False, repeat=1):
    """One BF16 MFMA conv kernel on NHWC float data (op 0 fwd, 1 data grad, 2 weight grad).
    Returns (out, average kernel ms)."""
    lib = load()
    x1 = np.ascontiguousarray(x1, np.float32)
    wt = np.ascontiguousarray(wt, np.float32)
    n, hh, ww, c = x1.shape
    cout = wt.shape[1]
    cin = wt.shape[2]
    c1 = c if op != 1 else cin
    c2 = 0 if x2 is None else x2.shape[-1]
    if op == 0:
        out = np.zeros((n, hh, ww, cout), np.float32)
    elif op == 1:
        out = np.zeros((n, hh, ww, cin), np.float32)
    else:
        out = np.zeros((9, cout, cin), np.float32)
    ms = C.c_double(0)
    x2a = None if x2 is None else np.ascontiguousarray(x2, np.float32)
    dya = None if dy is None else np.ascontiguousarray(dy, np.float32)
    ba = None if bias is None else np.ascontiguousarray(bias, np.float32)
    _check(lib.ds2d_conv3x3_bf16(op, n, hh, ww, c1 - c2 if op == 1 else c1, c2, up, cout, fptr(x1),
                                 fptr(x2a) if x2a is not None else None, fptr(dya) if dya is not None else None,
                                 fptr(wt), fptr(ba) if ba is not None else None, int(relu), repeat, fptr(out),
                                 C.byref(ms)))
    return out, ms.value


def sjcg_update(m, cg, g1, g0, it):
    """The reference's nonlinear CG update (in place on m, cg, g0); returns (alpha, beta)."""
    lib = load()
    a, b = C.c_float(0), C.c_float(0)
    for v in (m, cg, g1, g0):
        assert v.dtype == np.float32 and v.flags["C_CONTIGUOUS"]
    _check(lib.ds2d_sjcg_update(m.size, fptr(m), fptr(cg), fptr(g1), fptr(g0), int(it), C.byref(a), C.byref(b)))
    return a.value, b.value
