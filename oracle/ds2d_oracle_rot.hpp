// DS2D CPU oracle — TEST INFRASTRUCTURE ONLY (see ds2d_oracle.hpp header).
//
// The reference's alternative imaging paths, restated on the CPU:
//   * the P/S-decoupled (theta/omega rotation) elastic propagator,
//   * the P/S-decomposed RTM image of its random-boundary RTM,
//   * the Helmholtz (divergence / curl) and vector P/S decompositions,
//   * the Poynting up/down split and the illumination-normalised imaging
//     conditions of the elastic RTM that images decomposed wavefields.
//
// Parity pins (tests/test_oracle_rot.py):
//   rot_v / rot_p      single plane-wave mode against the closed-form 2x2
//                      recurrence of the discrete scheme (weights, stagger,
//                      vp^2 / vs^2, dt/dx vs dt/dz); curl-free evolution of a
//                      dilatation pulse and divergence-free evolution of a
//                      rotation pulse with dx != dz; exact time reversal
//   helmholtz, vector  Taylor exactness on polynomials of degree <= 7
//   poynting_split     partition identity; a down-going plane P wave is "down"
//   image conditions   closed forms on constant fields
//   rot_rtm, decomp_rtm  composition of the pinned parts.  The reference's
//                      drivers of these paths are commented out
//                      (ROT_EXP_IMAGE_SINGLESHOT, FD2DGPU.cu:3505-3585) or do
//                      not compile (FD_ELASTIC calls calculateVx() with no
//                      arguments, FD2DGPU.cu:2283 vs FD2DGPU.cuh:268), so their
//                      schedules are restated: parity unpinned beyond the parts.
#pragma once
#include "ds2d_oracle.hpp"

namespace orc {

// The rotation wavefield.  The reference aliases it onto the split elastic
// arrays (FD2DGPU.cu:1402-1412: theta = txx, omega = txz, vpx = vx_x, ...).
template <typename R>
struct Rot {
    std::vector<R> th, om, th_x, th_z, om_x, om_z, vx, vz, vpx, vpz, vsx, vsz;
    std::vector<std::vector<R>*> all() {
        return {&th, &om, &th_x, &th_z, &om_x, &om_z, &vx, &vz, &vpx, &vpz, &vsx, &vsz};
    }
    void init(size_t n) {
        for (auto* v : all()) v->assign(n, R(0));
    }
};

// FD2DGPU.cu:3005-3041 GPUcalculate_rotation_v (damped form; with zero damping
// d1 = d2 = 1 exactly, so it equals the undamped form of FD2DGPU.cu:2545-2575).
// Bounds [nop, n-nop): the reference's inclusive last row and column read
// omega one cell past the window (omega[(ix+4)*nzpml+iz] at ix = nxpml-4).
template <typename R>
void rot_v(const Grid<R>& g, Rot<R>& f, const R* vp, const R* vs, const R* dampx, const R* dampz, int dir) {
    const R dt = g.dt, dx = g.dx, dz = g.dz;
    const int nzp = g.nzp;
#pragma omp parallel for schedule(static)
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            const size_t c = g.at(ix, iz);
            const R d1 = 1 - dt * dampx[c] / 2, d2 = 1 + dt * dampx[c] / 2;
            const R d3 = 1 - dt * dampz[c] / 2, d4 = 1 + dt * dampz[c] / 2;
            R tx = 0, tz = 0, ox = 0, oz = 0;
            for (int i = 1; i <= NOP; ++i) {
                tx += (R)C8[i] * (f.th[c + (size_t)(i - 1) * nzp] - f.th[c - (size_t)i * nzp]);
                tz += (R)C8[i] * (f.th[c + i - 1] - f.th[c - i]);
                ox += (R)C8[i] * (f.om[c + (size_t)i * nzp] - f.om[c - (size_t)(i - 1) * nzp]);
                oz += (R)C8[i] * (f.om[c + i] - f.om[c - i + 1]);
            }
            const R p2 = vp[c] * vp[c], s2 = vs[c] * vs[c];
            f.vpx[c] = (d1 * f.vpx[c] + dir * p2 * (dt / dx) * tx) / d2;
            f.vpz[c] = (d3 * f.vpz[c] + dir * p2 * (dt / dz) * tz) / d4;
            f.vsx[c] = (d1 * f.vsx[c] + dir * s2 * (dt / dz) * oz) / d2;
            f.vsz[c] = (d3 * f.vsz[c] - dir * s2 * (dt / dx) * ox) / d4;
            f.vx[c] = f.vpx[c] + f.vsx[c];
            f.vz[c] = f.vpz[c] + f.vsz[c];
        }
}

// FD2DGPU.cu:3043-3080 GPUcalculate_rotation_p (damped form of 2577-2609),
// bounds as rot_v.
template <typename R>
void rot_p(const Grid<R>& g, Rot<R>& f, const R* dampx, const R* dampz, int dir) {
    const R dt = g.dt, dx = g.dx, dz = g.dz;
    const int nzp = g.nzp;
#pragma omp parallel for schedule(static)
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            const size_t c = g.at(ix, iz);
            const R d1 = 1 - dt * dampx[c] / 2, d2 = 1 + dt * dampx[c] / 2;
            const R d3 = 1 - dt * dampz[c] / 2, d4 = 1 + dt * dampz[c] / 2;
            R xx = 0, xz = 0, zx = 0, zz = 0;
            for (int i = 1; i <= NOP; ++i) {
                xx += (R)C8[i] * (f.vx[c + (size_t)i * nzp] - f.vx[c - (size_t)(i - 1) * nzp]);
                xz += (R)C8[i] * (f.vx[c + i - 1] - f.vx[c - i]);
                zx += (R)C8[i] * (f.vz[c + (size_t)(i - 1) * nzp] - f.vz[c - (size_t)i * nzp]);
                zz += (R)C8[i] * (f.vz[c + i] - f.vz[c - i + 1]);
            }
            f.th_x[c] = (d1 * f.th_x[c] + dir * (dt / dx) * xx) / d2;
            f.th_z[c] = (d3 * f.th_z[c] + dir * (dt / dz) * zz) / d4;
            f.om_x[c] = (d1 * f.om_x[c] - dir * (dt / dx) * zx) / d2;
            f.om_z[c] = (d3 * f.om_z[c] + dir * (dt / dz) * xz) / d4;
            f.th[c] = f.th_x[c] + f.th_z[c];
            f.om[c] = f.om_x[c] + f.om_z[c];
        }
}

// FD2DGPU.cu:1963-1986 GPUcalculate_helmholtz on [nop, n-nop] (inclusive, as
// the reference: its taps reach ix+3 at most, inside the window):
// VP = div v, VS = d(vx)/dz - d(vz)/dx.  Cells outside keep their value.
template <typename R>
void helmholtz(const Grid<R>& g, const R* vx, const R* vz, R* VP, R* VS) {
    const int nzp = g.nzp;
#pragma omp parallel for schedule(static)
    for (int ix = NOP; ix <= g.nxp - NOP; ++ix)
        for (int iz = NOP; iz <= g.nzp - NOP; ++iz) {
            const size_t c = g.at(ix, iz);
            R xx = 0, zz = 0, xz = 0, zx = 0;
            for (int i = 1; i <= NOP; ++i) {
                xx += (R)C8[i] * (vx[c + (size_t)(i - 1) * nzp] - vx[c - (size_t)i * nzp]);
                zz += (R)C8[i] * (vz[c + i - 1] - vz[c - i]);
                xz += (R)C8[i] * (vx[c + i - 1] - vx[c - i]);
                zx += (R)C8[i] * (vz[c + (size_t)(i - 1) * nzp] - vz[c - (size_t)i * nzp]);
            }
            VP[c] = xx / g.dx + zz / g.dz;
            VS[c] = xz / g.dz - zx / g.dx;
        }
}

// FD2DGPU.cu:1988-2017 GPUcalculate_vector_VPVS at one cell:
// vecP = (dVP/dx, dVP/dz), vecS = (-dVS/dz, dVS/dx).
template <typename R>
inline void vector_at(const Grid<R>& g, size_t c, const R* VP, const R* VS, R* pxz, R* sxz) {
    const int nzp = g.nzp;
    R px = 0, pz = 0, sz = 0, sx = 0;
    for (int i = 1; i <= NOP; ++i) {
        px += (R)C8[i] * (VP[c + (size_t)(i - 1) * nzp] - VP[c - (size_t)i * nzp]);
        pz += (R)C8[i] * (VP[c + i - 1] - VP[c - i]);
        sz += (R)C8[i] * (VS[c + i - 1] - VS[c - i]);
        sx += (R)C8[i] * (VS[c + (size_t)(i - 1) * nzp] - VS[c - (size_t)i * nzp]);
    }
    if (pxz) { pxz[0] = px / g.dx; pxz[1] = pz / g.dz; }
    if (sxz) { sxz[0] = -sz / g.dz; sxz[1] = sx / g.dx; }
}

// FD2DGPU.cu:2223-2236 poynting_both_up_down at one cell: the z energy flux
// p = -txz vx - tzz vz sends the vector to "down" when p >= 0, else to "up".
template <typename R>
inline void poynting_split_at(R txz, R tzz, R vx, R vz, const R* vec, R* up, R* down) {
    const R p = -txz * vx - tzz * vz;
    const bool dn = p >= 0;
    down[0] = dn ? vec[0] : R(0);
    down[1] = dn ? vec[1] : R(0);
    up[0] = dn ? R(0) : vec[0];
    up[1] = dn ? R(0) : vec[1];
}

// Imaging conditions, one time step at one cell.  EPS is the reference's
// float epsilon = 1e-6 (FD2DGPU.cu:2147, 2170, 3223).
template <typename R>
inline R image_eps() { return (R)(float)1e-6; }
// FD2DGPU.cu:2164-2196 doimage_poynting: (fu.bd + fd.bu) / (fu.fu + fd.fd + eps).
template <typename R>
inline R image_poynting(const R* fu, const R* fd, const R* bu, const R* bd) {
    const R il = (fu[0] * fu[0] + fu[1] * fu[1]) + (fd[0] * fd[0] + fd[1] * fd[1]) + image_eps<R>();
    const R cc = (fu[0] * bd[0] + fu[1] * bd[1]) + (fd[0] * bu[0] + fd[1] * bu[1]);
    return cc / il;
}
// FD2DGPU.cu:2141-2162 doimage, VECTORIMG form: f.b / (f.f + eps).
template <typename R>
inline R image_vector(const R* f, const R* b) {
    const R il = (f[0] * f[0] + f[1] * f[1]) + image_eps<R>();
    const R cc = f[0] * b[0] + f[1] * b[1];
    return cc / il;
}
// FD2DGPU.cu:2141-2162 doimage, scalar form: f b / (f^2 + eps).
template <typename R>
inline R image_scalar(R f, R b) {
    return f * b / (f * f + image_eps<R>());
}

}  // namespace orc
