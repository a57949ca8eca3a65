"""Seeded synthetic models and shot layouts for the DS2D workloads.

All randomness comes from numpy's PCG64 generator seeded with the workload
seed, so every caller (GPU path, oracle checks, bench) sees identical inputs.
Arrays are [ix][iz] (x-major, z contiguous), float32, like the reference's
model files (main.cpp:1464-1468).
"""
import numpy as np

from .lib import ShotSpec


def layered_model(nx, nz, seed, n_layers=(8, 15), vmin=2000.0, vmax=5500.0, fault=True, lenses=0):
    """Layered vp with one dipping fault (and optional Gaussian lenses).

    L ~ U{8..15} interfaces at sorted U(0.05, 0.95)*nz; layer velocities
    2000 + 3500*sorted U(0,1); a fault through x0 ~ U(0.3,0.7)*nx dipping
    60-80 degrees shifts the hanging-wall interfaces down by U{3..10} cells;
    lenses are Gaussian anomalies of +-15 % vp.
    """
    rng = np.random.default_rng(seed)
    L = int(rng.integers(n_layers[0], n_layers[1] + 1))
    zi = np.sort(rng.uniform(0.05, 0.95, L)) * nz
    vel = vmin + (vmax - vmin) * np.sort(rng.uniform(0.0, 1.0, L + 1))
    x0 = rng.uniform(0.3, 0.7) * nx
    dip = np.deg2rad(rng.uniform(60.0, 80.0))
    throw = int(rng.integers(3, 11)) if fault else 0
    ix = np.arange(nx)[:, None].astype(np.float64)
    iz = np.arange(nz)[None, :].astype(np.float64)
    xf = x0 + iz / np.tan(dip)
    shift = np.where(ix > xf, throw, 0)
    zeff = iz - shift
    layer = np.searchsorted(zi, zeff, side="right")
    vp = vel[layer]
    for _ in range(lenses):
        cx, cz = rng.uniform(0.1, 0.9) * nx, rng.uniform(0.2, 0.9) * nz
        r = rng.uniform(0.03, 0.08) * min(nx, nz) * 2
        amp = rng.uniform(-0.15, 0.15)
        vp = vp * (1.0 + amp * np.exp(-((ix - cx) ** 2 + (iz - cz) ** 2) / (2 * r * r)))
    return np.ascontiguousarray(vp, np.float32)


def elastic_from_vp(vp):
    """vs = vp/1.73, rho = 310 vp^0.25 (Gardner)."""
    vp = np.asarray(vp, np.float32)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    rho = (np.float32(310.0) * np.power(vp.astype(np.float64), 0.25)).astype(np.float32)
    return vs, rho


def gaussian_smooth(v, size=9, sigma=2.0):
    """Normalised size x size Gaussian, edge-replicated borders."""
    h = size // 2
    k1 = np.exp(-0.5 * (np.arange(-h, h + 1) / sigma) ** 2)
    k1 /= k1.sum()
    p = np.pad(np.asarray(v, np.float64), h, mode="edge")
    a = sum(k1[i] * p[i:i + v.shape[0], :] for i in range(size))
    b = sum(k1[j] * a[:, j:j + v.shape[1]] for j in range(size))
    return b.astype(np.float32)


def surface_shot(nx_win, ngx_min, pml, sx_local, sz=2, gz=0, shot_id=0, boundary=0):
    """Source at window column sx_local, depth sz cells; receivers at every
    surface node of the window at depth gz (padded coordinates)."""
    return ShotSpec(ngx_min, nx_win, pml + sx_local, pml + sz, pml + gz, pml + np.arange(nx_win, dtype=np.int32),
                    shot_id=shot_id, boundary=boundary)


# ---------------------------------------------------------------------------
# Workloads (the benchmark configurations C1-C5)
# ---------------------------------------------------------------------------

def c1_elastic_forward(seed=1):
    """Plain elastic forward modelling, 251x101, dx=dz=10 m, pml 20, dt 0.8 ms,
    2500 steps, 15 Hz Ricker at ix=125 iz=2, vx/vz at every surface node."""
    nx, nz = 251, 101
    vp = layered_model(nx, nz, seed)
    vs, rho = elastic_from_vp(vp)
    return dict(name="c1_elastic_forward", nx=nx, nz=nz, dx=10.0, dz=10.0, dt=0.0008, nt=2500, pml=20, f0=15.0,
                vp=vp, vs=vs, rho=rho, shots=[surface_shot(nx, 0, 20, 125, 2, 0)])


def c2_acoustic_lsrtm(seed=2, nt=2500, dt=0.0006, nshots=26, spacing=10):
    """Acoustic RTM + Born LSRTM on 251x101: 26 shots 10 nodes apart, random
    boundary 40 cells (+-30 %), migration model = true vp smoothed 9x9."""
    nx, nz, pml = 251, 101, 40
    vp_true = layered_model(nx, nz, seed)
    vp_mig = gaussian_smooth(vp_true)
    refl = ((vp_true - vp_mig) / vp_mig).astype(np.float32)
    _, rho = elastic_from_vp(vp_mig)
    shots = [surface_shot(nx, 0, pml, min(i * spacing, nx - 1), 2, 0, shot_id=i, boundary=1)
             for i in range(nshots)]
    return dict(name="c2_acoustic_lsrtm", nx=nx, nz=nz, dx=10.0, dz=10.0, dt=dt, nt=nt, pml=pml, f0=15.0,
                vp=vp_mig, vp_true=vp_true, rho=rho, refl=refl, shots=shots)


def c3_batched_elastic(seed=3, nshots=128, nt=3000, wmin=400, wmax=1200):
    """128 uneven-aperture elastic shots on 2000x600, dx=dz=12.5 m, dt 1 ms,
    3000 steps, pml 20; window width U{400..1200}, start uniform in the model,
    source at the window centre, receivers at every window surface node."""
    nx, nz, pml = 2000, 600, 20
    vp = layered_model(nx, nz, seed)
    vs, rho = elastic_from_vp(vp)
    rng = np.random.default_rng(seed + 1000)
    shots = []
    for i in range(nshots):
        w = int(rng.integers(wmin, wmax + 1))
        x0 = int(rng.integers(0, nx - w + 1))
        shots.append(surface_shot(w, x0, pml, w // 2, 2, 0, shot_id=i, boundary=0))
    return dict(name="c3_batched_elastic", nx=nx, nz=nz, dx=12.5, dz=12.5, dt=0.001, nt=nt, pml=pml, f0=15.0,
                vp=vp, vs=vs, rho=rho, shots=shots)


def c5_combined(seed=7, nshots=256, nt=3000, dt=0.0008, wmin=400, wmax=1200, pml=40):
    """Full combined path: 4000x1000 seeded elastic model, dx=dz=10 m, 256
    uneven-aperture shots (window width U{400..1200}, start uniform), random
    boundary `pml` cells; vs from vp/1.73 scaled +-15 % by a second layering;
    9x9-smoothed migration vp/vs; PP and PS reflectivities (v - v_mig)/v_mig.
    dt 0.8 ms (CFL 0.80 at 5500 m/s), `nt` steps; source at the window
    centre at depth 2, receivers at depth 1 on every window node."""
    nx, nz = 4000, 1000
    vp_true = layered_model(nx, nz, seed)
    r2 = layered_model(nx, nz, seed + 500).astype(np.float64)
    r2 = 2 * (r2 - r2.min()) / max(r2.max() - r2.min(), 1.0) - 1
    vs_true = (vp_true / 1.73 * (1 + 0.15 * r2)).astype(np.float32)
    vp_mig, vs_mig = gaussian_smooth(vp_true), gaussian_smooth(vs_true)
    _, rho = elastic_from_vp(vp_mig)
    rng = np.random.default_rng(seed + 1000)
    shots = []
    for i in range(nshots):
        w = int(rng.integers(wmin, wmax + 1))
        x0 = int(rng.integers(0, nx - w + 1))
        shots.append(surface_shot(w, x0, pml, w // 2, 2, 1, shot_id=i, boundary=1))
    return dict(name="c5_combined", nx=nx, nz=nz, dx=10.0, dz=10.0, dt=dt, nt=nt, pml=pml, f0=15.0, vp=vp_mig,
                vs=vs_mig, rho=rho, refl_pp=((vp_true - vp_mig) / vp_mig).astype(np.float32),
                refl_ps=((vs_true - vs_mig) / vs_mig).astype(np.float32), shots=shots)

