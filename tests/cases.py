"""Small seeded cases shared by the oracle, golden and GPU parity tests."""
import numpy as np

from ds2d_amd import models
from ds2d_amd.lib import ShotSpec

DX, DZ, DT = 10.0, 10.0, 0.0008


def wavelet(nt, f0=15.0, dt=DT):
    """Ricker (main.cpp:461-473) in float32, the input handed to both sides."""
    pi = np.float32(3.14159265358979)
    kt = int(np.float32(1.0) / np.float32(f0) / np.float32(dt))
    i = np.arange(nt, dtype=np.float32)
    tt = i * np.float32(dt) - np.float32(kt) * np.float32(dt)
    sp = pi * np.float32(f0) * tt
    return (np.exp(-sp.astype(np.float64) ** 2) * (1.0 - 2.0 * sp.astype(np.float64) ** 2)).astype(np.float32)


def elastic_model(nx, nz, seed):
    vp = models.layered_model(nx, nz, seed)
    vs, rho = models.elastic_from_vp(vp)
    return vp, vs, rho


def full_window_shot(nx, pml, sx, sz=2, gz=0, shot_id=0, boundary=0, ngx_min=0, stride=1):
    rec = pml + np.arange(0, nx, stride, dtype=np.int32)
    return ShotSpec(ngx_min, nx, pml + sx, pml + sz, pml + gz, rec, shot_id=shot_id, boundary=boundary)


def uneven_shots(allnx, pml, widths, starts, boundary=0):
    shots = []
    for i, (w, x0) in enumerate(zip(widths, starts)):
        shots.append(ShotSpec(x0, w, pml + w // 2, pml + 2, pml, pml + np.arange(w, dtype=np.int32), shot_id=i,
                              boundary=boundary))
    return shots
