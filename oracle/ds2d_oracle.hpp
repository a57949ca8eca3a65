// DS2D CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// A plain C++17/OpenMP restatement of the reference's time-domain hot path,
// templated on the arithmetic type R (double = the parity truth, float = an
// FP32 restatement).  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as a checker or a timed
// CPU baseline; the product library (ds2d_amd/libds2d.so) never links it.
//
// Every sweep keeps the reference's loop bounds and per-cell arithmetic order.
// One `omp parallel for` over ix per sweep reproduces the GPU semantics
// exactly because no live kernel reads an array it writes at a neighbour.
//
// Layout: every padded field is [ix][iz], z contiguous, nxp*nzp cells.
// Reference files are cited as file:line of the upstream DS2D tree.
//
// Parity pins (what checks each function, see tests/):
//   coeff           Taylor exactness x^k, k<=7                  (test_oracle_pins)
//   ricker          closed form                                 (test_oracle_pins)
//   damp_profile    closed form pointwise                       (test_oracle_pins)
//   pad_window      brute-force index check                     (test_oracle_pins)
//   elastic step    homogeneous arrival time, symmetry, mu=0 invariant,
//                   time reversal (no damping)                  (test_oracle_physics)
//   born source     linearity + finite difference of the forward map
//   plain adjoint   dot-product test <Lm,d> = <m,L^T d> in FP64 (test_oracle_adjoint)
//   laplace         quadratic closed form + compiled reference laplace.cpp
//   ref ELSRTM      composition of the pinned kernels; the reference's own
//                   program cannot be built here (CUDA/MPI), so the schedule
//                   itself is "parity unpinned" beyond its pinned parts.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace orc {

// FD2DGPU.cuh:16 — coeff2 = exact 8th-order staggered first-derivative weights.
static const double C8[5] = {0.0, 1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0};
static const int NOP = 4;

template <typename R>
struct Grid {
    int nxp, nzp, pml;
    R dx, dz, dt;
    size_t n() const { return (size_t)nxp * nzp; }
    size_t at(int ix, int iz) const { return (size_t)ix * nzp + iz; }
};

// One elastic wavefield in the reference's split-PML layout (15 arrays).
template <typename R>
struct Elastic {
    std::vector<R> vx, vz, txx, tzz, txz;
    std::vector<R> vx_x, vx_z, vz_x, vz_z;
    std::vector<R> txx_x, txx_z, tzz_x, tzz_z, txz_x, txz_z;
    void init(size_t n) {
        for (auto* v : all()) v->assign(n, R(0));
    }
    std::vector<std::vector<R>*> all() {
        return {&vx, &vz, &txx, &tzz, &txz, &vx_x, &vx_z, &vz_x, &vz_z,
                &txx_x, &txx_z, &tzz_x, &tzz_z, &txz_x, &txz_z};
    }
};

// Medium on the padded window.
template <typename R>
struct Medium {
    std::vector<R> vp, vs, rho, lam, mu, dampx, dampz;
};

// ---------------------------------------------------------------------------
// Setup helpers
// ---------------------------------------------------------------------------

// main.cpp:461-473 ricker1: kt=(int)(1/(f*dt)); s[i]=exp(-sp^2)(1-2sp^2), sp=pi f (i-kt) dt.
template <typename R>
void ricker(int nt, R f, R dt, R* s) {
    const R pi = (R)3.14159265358979;
    const R t0 = (R)1 / f;
    const int kt = (int)(t0 / dt);
    for (int i = 0; i < nt; ++i) {
        R tt = (R)i * dt - (R)kt * dt;
        R sp = pi * f * tt;
        s[i] = (R)(std::exp(-sp * sp) * (1.0 - 2.0 * sp * sp));
    }
}

// FD2DGPU.cu:122-138 copyvelocityHtoD, then FD2DGPU.cu:234-250 copypmlVz and
// copypmlVx (separate launches, so top/bottom replication happens first).
template <typename R>
void pad_window(const R* glob, int allnz, int ngx_min, int nx, int nz, int pml, R* out) {
    const int nzp = nz + 2 * pml, nxp = nx + 2 * pml;
    for (int ix = 0; ix < nx; ++ix)
        for (int iz = 0; iz < nz; ++iz)
            out[(size_t)(ix + pml) * nzp + iz + pml] = glob[(size_t)(ngx_min + ix) * allnz + iz];
    for (int ix = 0; ix < nxp; ++ix)
        for (int iz = 0; iz < pml; ++iz) {
            out[(size_t)ix * nzp + iz] = out[(size_t)ix * nzp + pml];
            out[(size_t)ix * nzp + nzp - pml + iz] = out[(size_t)ix * nzp + nzp - pml - 1];
        }
    for (int ix = 0; ix < pml; ++ix)
        for (int iz = 0; iz < nzp; ++iz) {
            out[(size_t)ix * nzp + iz] = out[(size_t)pml * nzp + iz];
            out[(size_t)(nxp - pml + ix) * nzp + iz] = out[(size_t)(nxp - pml - 1) * nzp + iz];
        }
}

// FD2DGPU.cu:444-494 dampX/dampZ: d = log10(1/rr)*(5 vp/(2 pml))*((pml-i)/pml)^4 on
// the near side (i<pml) and ((i-(n-pml))/pml)^4 on the far side with a STRICT
// i > n-pml; every other cell keeps its zero initialisation (FD2DGPU.cu:1394-1395).
template <typename R>
void damp_profile(const Grid<R>& g, const R* vp, R rr, R* dampx, R* dampz) {
    const int pml = g.pml;
    const double a = std::log10(1.0 / (double)rr);
    for (int ix = 0; ix < g.nxp; ++ix)
        for (int iz = 0; iz < g.nzp; ++iz) {
            size_t c = g.at(ix, iz);
            double base = a * (5.0 * (double)vp[c] / (2.0 * pml));
            dampx[c] = 0;
            dampz[c] = 0;
            if (ix < pml) dampx[c] = (R)(base * std::pow((double)(pml - ix) / pml, 4.0));
            if (ix > g.nxp - pml) dampx[c] = (R)(base * std::pow((double)(ix - (g.nxp - pml)) / pml, 4.0));
            if (iz < pml) dampz[c] = (R)(base * std::pow((double)(pml - iz) / pml, 4.0));
            if (iz > g.nzp - pml) dampz[c] = (R)(base * std::pow((double)(iz - (g.nzp - pml)) / pml, 4.0));
        }
}

// FD2DGPU.cu:1117-1125 GPUcalculate_elastic_rho_C: lam = rho(vp^2-2vs^2), mu = rho vs^2.
// rho_const > 0 reproduces the reference's fixed rho=2000; otherwise rho_in is used.
template <typename R>
void elastic_moduli(size_t n, const R* vp, const R* vs, const R* rho_in, R rho_const,
                    R* rho, R* lam, R* mu) {
    for (size_t c = 0; c < n; ++c) {
        R r = rho_const > 0 ? rho_const : rho_in[c];
        rho[c] = r;
        lam[c] = r * (vp[c] * vp[c] - 2 * vs[c] * vs[c]);
        mu[c] = r * vs[c] * vs[c];
    }
}

// Counter-based hash used for the seeded random boundary.  The product library
// implements the same published SplitMix64 finaliser (Steele, Lea, Flood 2014).
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
// Uniform in [-1, 1) with 24 random bits, exact in float.
inline double rb_uniform(uint64_t seed, int shot, int ix, int iz) {
    uint64_t key = seed * 0x100000001B3ull ^ ((uint64_t)(uint32_t)shot << 42) ^
                   ((uint64_t)(uint32_t)ix << 21) ^ (uint64_t)(uint32_t)iz;
    uint64_t h = splitmix64(key);
    double u = (double)(h >> 40) * (1.0 / 16777216.0);  // [0,1)
    return 2.0 * u - 1.0;
}

// Random boundary (replaces FD2DGPU.cu:154-230, 3156-3216, whose cuRAND seed is
// time(0) and whose index_opposite is out of bounds; see Appendix fix):
// every cell of the pml-wide frame gets v = v_edge*(1 + amp*u), u ~ U[-1,1)
// from the counter hash; v_edge is the edge-replicated value already in the
// padded array.  vs keeps vs/vp of the replicated cell; rho is left replicated.
// The reference writes the frame into its float vp / vs arrays (random_vel,
// FD2DGPU.cu:3181), so the new velocities are float products and quotients
// whatever R is: the FP64 oracle then steps with the values the model holds.
template <typename R>
void random_boundary(const Grid<R>& g, uint64_t seed, int shot, R amp, R* vp, R* vs) {
    const int pml = g.pml;
    for (int ix = 0; ix < g.nxp; ++ix)
        for (int iz = 0; iz < g.nzp; ++iz) {
            bool frame = ix < pml || ix >= g.nxp - pml || iz < pml || iz >= g.nzp - pml;
            if (!frame) continue;
            size_t c = g.at(ix, iz);
            double u = rb_uniform(seed, shot, ix, iz);
            float fac = (float)(1.0 + (double)(float)amp * u);
            const float v0 = (float)vp[c];
            const float ratio = v0 > 0.f ? (float)vs[c] / v0 : 0.f;
            const float vnew = v0 * fac;
            vp[c] = (R)vnew;
            vs[c] = (R)(ratio * vnew);
        }
}

// ---------------------------------------------------------------------------
// Elastic stencils (reference layout)
// ---------------------------------------------------------------------------

// born.cu:215-279 update_vx + update_vz on [nop, n-nop)^2 (two launches that
// touch disjoint outputs, so one sweep is identical).
template <typename R>
void update_v(const Grid<R>& g, Elastic<R>& f, const Medium<R>& m, int dir) {
    const R dt = g.dt, dx = g.dx, dz = g.dz;
    const int nzp = g.nzp;
#pragma omp parallel for schedule(static)
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            const size_t c = g.at(ix, iz);
            R d1 = 1 - dt * m.dampx[c] / 2, d2 = 1 + dt * m.dampx[c] / 2;
            R d3 = 1 - dt * m.dampz[c] / 2, d4 = 1 + dt * m.dampz[c] / 2;
            R a = 0, b = 0, cz = 0, e = 0;
            for (int i = 1; i <= NOP; ++i) {
                a += (R)C8[i] * (f.txx[c + (size_t)i * nzp] - f.txx[c - (size_t)(i - 1) * nzp]);
                b += (R)C8[i] * (f.txz[c + i - 1] - f.txz[c - i]);
                cz += (R)C8[i] * (f.tzz[c + i] - f.tzz[c - i + 1]);
                e += (R)C8[i] * (f.txz[c + (size_t)(i - 1) * nzp] - f.txz[c - (size_t)i * nzp]);
            }
            R br = (R)1 / m.rho[c];
            f.vx_x[c] = (d1 * f.vx_x[c] + dir * br * (dt / dx) * a) / d2;
            f.vx_z[c] = (d3 * f.vx_z[c] + dir * br * (dt / dz) * b) / d4;
            f.vx[c] = f.vx_x[c] + f.vx_z[c];
            f.vz_x[c] = (d1 * f.vz_x[c] + dir * br * (dt / dx) * e) / d2;
            f.vz_z[c] = (d3 * f.vz_z[c] + dir * br * (dt / dz) * cz) / d4;
            f.vz[c] = f.vz_x[c] + f.vz_z[c];
        }
}

// born.cu:282-349 update_txx_tzz + update_txz on [nop, n-nop)^2.
template <typename R>
void update_s(const Grid<R>& g, Elastic<R>& f, const Medium<R>& m, int dir) {
    const R dt = g.dt, dx = g.dx, dz = g.dz;
    const int nzp = g.nzp;
#pragma omp parallel for schedule(static)
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            const size_t c = g.at(ix, iz);
            R d1 = 1 - dt * m.dampx[c] / 2, d2 = 1 + dt * m.dampx[c] / 2;
            R d3 = 1 - dt * m.dampz[c] / 2, d4 = 1 + dt * m.dampz[c] / 2;
            R dvx = 0, dvz = 0, dvxz = 0, dvzx = 0;
            for (int i = 1; i <= NOP; ++i) {
                dvx += (R)C8[i] * (f.vx[c + (size_t)(i - 1) * nzp] - f.vx[c - (size_t)i * nzp]);
                dvz += (R)C8[i] * (f.vz[c + i - 1] - f.vz[c - i]);
                dvxz += (R)C8[i] * (f.vx[c + i] - f.vx[c - i + 1]);
                dvzx += (R)C8[i] * (f.vz[c + (size_t)i * nzp] - f.vz[c - (size_t)(i - 1) * nzp]);
            }
            const R l = m.lam[c], mu = m.mu[c], l2m = l + 2 * mu;
            f.txx_x[c] = (d1 * f.txx_x[c] + dir * l2m * (dt / dx) * dvx) / d2;
            f.txx_z[c] = (d3 * f.txx_z[c] + dir * l * (dt / dz) * dvz) / d4;
            f.tzz_x[c] = (d1 * f.tzz_x[c] + dir * l * (dt / dx) * dvx) / d2;
            f.tzz_z[c] = (d3 * f.tzz_z[c] + dir * l2m * (dt / dz) * dvz) / d4;
            f.txx[c] = f.txx_x[c] + f.txx_z[c];
            f.tzz[c] = f.tzz_x[c] + f.tzz_z[c];
            f.txz_x[c] = (d1 * f.txz_x[c] + dir * (dt / dx) * mu * dvzx) / d2;
            f.txz_z[c] = (d3 * f.txz_z[c] + dir * (dt / dz) * mu * dvxz) / d4;
            f.txz[c] = f.txz_x[c] + f.txz_z[c];
        }
}

// FD2DGPU.cu:903-989 GPUcalculate_elastic_vx/vz: same update as update_v but on
// the interior [pml, n-pml)^2 only (reconstruction).
template <typename R>
void recon_v(const Grid<R>& g, Elastic<R>& f, const Medium<R>& m, int dir) {
    const R dt = g.dt, dx = g.dx, dz = g.dz;
    const int nzp = g.nzp, p = g.pml;
#pragma omp parallel for schedule(static)
    for (int ix = p; ix < g.nxp - p; ++ix)
        for (int iz = p; iz < g.nzp - p; ++iz) {
            const size_t c = g.at(ix, iz);
            R d1 = 1 - dt * m.dampx[c] / 2, d2 = 1 + dt * m.dampx[c] / 2;
            R d3 = 1 - dt * m.dampz[c] / 2, d4 = 1 + dt * m.dampz[c] / 2;
            R a = 0, b = 0, cz = 0, e = 0;
            for (int i = 1; i <= NOP; ++i) {
                a += (R)C8[i] * (f.txx[c + (size_t)i * nzp] - f.txx[c - (size_t)(i - 1) * nzp]);
                b += (R)C8[i] * (f.txz[c + i - 1] - f.txz[c - i]);
                cz += (R)C8[i] * (f.tzz[c + i] - f.tzz[c - i + 1]);
                e += (R)C8[i] * (f.txz[c + (size_t)(i - 1) * nzp] - f.txz[c - (size_t)i * nzp]);
            }
            R br = (R)1 / m.rho[c];
            f.vx_x[c] = (d1 * f.vx_x[c] + dir * br * (dt / dx) * a) / d2;
            f.vx_z[c] = (d3 * f.vx_z[c] + dir * br * (dt / dz) * b) / d4;
            f.vx[c] = f.vx_x[c] + f.vx_z[c];
            f.vz_x[c] = (d1 * f.vz_x[c] + dir * br * (dt / dx) * e) / d2;
            f.vz_z[c] = (d3 * f.vz_z[c] + dir * br * (dt / dz) * cz) / d4;
            f.vz[c] = f.vz_x[c] + f.vz_z[c];
        }
}

// Strains emitted by the reconstruction stress kernels.
template <typename R>
struct Strain {
    std::vector<R> gx_vx, gz_vz, gz_vx, gx_vz;  // vx_gx, vz_gz, vx_gz, vz_gx
    void init(size_t n) {
        gx_vx.assign(n, 0); gz_vz.assign(n, 0); gz_vx.assign(n, 0); gx_vz.assign(n, 0);
    }
};

// FD2DGPU.cu:993-1082 GPUcalculate_elastic_txx_tzz_grad + _txz_grad on [pml,n-pml)^2;
// strains are the same tap sums divided by dx or dz.  The launch order in the
// reference is txz_grad first (FD2DGPU.cu:4735-4736); both read only v, so one
// sweep is identical.
template <typename R>
void recon_s(const Grid<R>& g, Elastic<R>& f, const Medium<R>& m, int dir, Strain<R>& s) {
    const R dt = g.dt, dx = g.dx, dz = g.dz;
    const int nzp = g.nzp, p = g.pml;
#pragma omp parallel for schedule(static)
    for (int ix = p; ix < g.nxp - p; ++ix)
        for (int iz = p; iz < g.nzp - p; ++iz) {
            const size_t c = g.at(ix, iz);
            R d1 = 1 - dt * m.dampx[c] / 2, d2 = 1 + dt * m.dampx[c] / 2;
            R d3 = 1 - dt * m.dampz[c] / 2, d4 = 1 + dt * m.dampz[c] / 2;
            R dvx = 0, dvz = 0, dvxz = 0, dvzx = 0;
            for (int i = 1; i <= NOP; ++i) {
                dvx += (R)C8[i] * (f.vx[c + (size_t)(i - 1) * nzp] - f.vx[c - (size_t)i * nzp]);
                dvz += (R)C8[i] * (f.vz[c + i - 1] - f.vz[c - i]);
                dvxz += (R)C8[i] * (f.vx[c + i] - f.vx[c - i + 1]);
                dvzx += (R)C8[i] * (f.vz[c + (size_t)i * nzp] - f.vz[c - (size_t)(i - 1) * nzp]);
            }
            const R l = m.lam[c], mu = m.mu[c], l2m = l + 2 * mu;
            f.txx_x[c] = (d1 * f.txx_x[c] + dir * l2m * (dt / dx) * dvx) / d2;
            f.txx_z[c] = (d3 * f.txx_z[c] + dir * l * (dt / dz) * dvz) / d4;
            f.tzz_x[c] = (d1 * f.tzz_x[c] + dir * l * (dt / dx) * dvx) / d2;
            f.tzz_z[c] = (d3 * f.tzz_z[c] + dir * l2m * (dt / dz) * dvz) / d4;
            f.txx[c] = f.txx_x[c] + f.txx_z[c];
            f.tzz[c] = f.tzz_x[c] + f.tzz_z[c];
            s.gx_vx[c] = dvx / dx;
            s.gz_vz[c] = dvz / dz;
            f.txz_x[c] = (d1 * f.txz_x[c] + dir * mu * (dt / dx) * dvzx) / d2;
            f.txz_z[c] = (d3 * f.txz_z[c] + dir * mu * (dt / dz) * dvxz) / d4;
            f.txz[c] = f.txz_x[c] + f.txz_z[c];
            s.gz_vx[c] = dvxz / dz;
            s.gx_vz[c] = dvzx / dx;
        }
}

// FD2DGPU.cu:591-682 GPUcalculate_elastic_vx_back/_vz_back: receiver-side
// velocity with the moduli inside the derivative, on [nop, n-nop)^2.
template <typename R>
void adjoint_v(const Grid<R>& g, Elastic<R>& f, const Medium<R>& m, int dir) {
    const R dt = g.dt, dx = g.dx, dz = g.dz;
    const int nzp = g.nzp;
    const std::vector<R>& L = m.lam;
    const std::vector<R>& M = m.mu;
#pragma omp parallel for schedule(static)
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            const size_t c = g.at(ix, iz);
            R d1 = 1 - dt * m.dampx[c] / 2, d2 = 1 + dt * m.dampx[c] / 2;
            R d3 = 1 - dt * m.dampz[c] / 2, d4 = 1 + dt * m.dampz[c] / 2;
            // vx_back
            R ttxx = 0, ttxz = 0, ttzz = 0;
            for (int i = 1; i <= NOP; ++i) {
                size_t p1 = c + (size_t)i * nzp, p0 = c - (size_t)(i - 1) * nzp;
                ttxx += (R)C8[i] * ((L[p1] + 2 * M[p1]) * f.txx[p1] - (L[p0] + 2 * M[p0]) * f.txx[p0]);
                ttxz += (R)C8[i] * (M[c + i - 1] * f.txz[c + i - 1] - M[c - i] * f.txz[c - i]);
                ttzz += (R)C8[i] * (L[p1] * f.tzz[p1] - L[p0] * f.tzz[p0]);
            }
            R br = (R)1 / m.rho[c];
            R nvx_x = (d1 * f.vx_x[c] + dir * br * (dt / dx) * ttxx + dir * br * (dt / dx) * ttzz) / d2;
            R nvx_z = (d3 * f.vx_z[c] + dir * br * (dt / dz) * ttxz) / d4;
            // vz_back
            R utzz = 0, utxz = 0, utxx = 0;
            for (int i = 1; i <= NOP; ++i) {
                utzz += (R)C8[i] * ((L[c + i] + 2 * M[c + i]) * f.tzz[c + i] -
                                    (L[c - i + 1] + 2 * M[c - i + 1]) * f.tzz[c - i + 1]);
                size_t q1 = c + (size_t)(i - 1) * nzp, q0 = c - (size_t)i * nzp;
                utxz += (R)C8[i] * (M[q1] * f.txz[q1] - M[q0] * f.txz[q0]);
                utxx += (R)C8[i] * (L[c + i] * f.txx[c + i] - L[c - i + 1] * f.txx[c - i + 1]);
            }
            R nvz_x = (d1 * f.vz_x[c] + dir * br * (dt / dx) * utxz) / d2;
            R nvz_z = (d3 * f.vz_z[c] + dir * br * (dt / dz) * utzz + dir * br * (dt / dz) * utxx) / d4;
            f.vx_x[c] = nvx_x; f.vx_z[c] = nvx_z; f.vx[c] = nvx_x + nvx_z;
            f.vz_x[c] = nvz_x; f.vz_z[c] = nvz_z; f.vz[c] = nvz_x + nvz_z;
        }
}

// FD2DGPU.cu:725-761 GPUcalculate_elastic_txx_tzz (txx=txx_x, tzz=tzz_z, no
// moduli) + FD2DGPU.cu:860-895 GPUcalculate_elastic_txz (no mu), [nop,n-nop)^2.
template <typename R>
void adjoint_s(const Grid<R>& g, Elastic<R>& f, const Medium<R>& m, int dir) {
    const R dt = g.dt, dx = g.dx, dz = g.dz;
    const int nzp = g.nzp;
#pragma omp parallel for schedule(static)
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            const size_t c = g.at(ix, iz);
            R d1 = 1 - dt * m.dampx[c] / 2, d2 = 1 + dt * m.dampx[c] / 2;
            R d3 = 1 - dt * m.dampz[c] / 2, d4 = 1 + dt * m.dampz[c] / 2;
            R dvx = 0, dvz = 0, dvxz = 0, dvzx = 0;
            for (int i = 1; i <= NOP; ++i) {
                dvx += (R)C8[i] * (f.vx[c + (size_t)(i - 1) * nzp] - f.vx[c - (size_t)i * nzp]);
                dvz += (R)C8[i] * (f.vz[c + i - 1] - f.vz[c - i]);
                dvxz += (R)C8[i] * (f.vx[c + i] - f.vx[c - i + 1]);
                dvzx += (R)C8[i] * (f.vz[c + (size_t)i * nzp] - f.vz[c - (size_t)(i - 1) * nzp]);
            }
            f.txx_x[c] = (d1 * f.txx_x[c] + dir * (dt / dx) * dvx) / d2;
            f.tzz_z[c] = (d3 * f.tzz_z[c] + dir * (dt / dz) * dvz) / d4;
            f.txx[c] = f.txx_x[c];
            f.tzz[c] = f.tzz_z[c];
            f.txz_x[c] = (d1 * f.txz_x[c] + dir * (dt / dx) * dvzx) / d2;
            f.txz_z[c] = (d3 * f.txz_z[c] + dir * (dt / dz) * dvxz) / d4;
            f.txz[c] = f.txz_x[c] + f.txz_z[c];
        }
}

// born.cu:351-405 add_born_source on [pml, n-pml)^2: adds to the stress SUM
// fields of the target wavefield, using velocities of the source wavefield.
// Operator precedence makes `direction` scale only the first term; every live
// call passes direction = 1.
template <typename R>
void add_born_source(const Grid<R>& g, const std::vector<R>& vx, const std::vector<R>& vz,
                     Elastic<R>& b, const std::vector<R>& dmp, const std::vector<R>& dms,
                     const Medium<R>& m, int dir) {
    const R dx = g.dx, dz = g.dz;
    const int nzp = g.nzp, p = g.pml;
#pragma omp parallel for schedule(static)
    for (int ix = p; ix < g.nxp - p; ++ix)
        for (int iz = p; iz < g.nzp - p; ++iz) {
            const size_t c = g.at(ix, iz);
            R vxx = 0, vzz = 0, vxz = 0, vzx = 0;
            for (int i = 1; i <= NOP; ++i) {
                vxx += (R)C8[i] * (vx[c + (size_t)(i - 1) * nzp] - vx[c - (size_t)i * nzp]);
                vzz += (R)C8[i] * (vz[c + i - 1] - vz[c - i]);
                vxz += (R)C8[i] * (vx[c + i] - vx[c - i + 1]);
                vzx += (R)C8[i] * (vz[c + (size_t)i * nzp] - vz[c - (size_t)(i - 1) * nzp]);
            }
            const R l = m.lam[c], mu = m.mu[c], l2m = l + 2 * mu;
            const R mp = -dmp[c], ms = -dms[c];
            b.txx[c] = b.txx[c] + dir * 2 * l2m * mp * vxx / dx + 2 * (l2m * mp - 2 * mu * ms) * vzz / dz;
            b.tzz[c] = b.tzz[c] + dir * 2 * (l2m * mp - 2 * mu * ms) * vxx / dx + 2 * l2m * mp * vzz / dz;
            b.txz[c] = b.txz[c] + dir * 2 * mu * ms * (vzx / dx + vxz / dz);
        }
}

// born.cu:408-527 add_born_source_adjoint on [pml, n-pml)^2: adds to the
// velocity SUM fields.  tmp_txz_x samples mu*dms at ix+i-1 for both taps
// (born.cu:514-515) — kept as in the reference (compat).
template <typename R>
void add_born_source_adjoint(const Grid<R>& g, const Elastic<R>& s, std::vector<R>& bvx,
                             std::vector<R>& bvz, const std::vector<R>& dmp,
                             const std::vector<R>& dms, const Medium<R>& m, int dir) {
    const R dx = g.dx, dz = g.dz;
    const int nzp = g.nzp, p = g.pml;
    const std::vector<R>& L = m.lam;
    const std::vector<R>& M = m.mu;
    auto CP = [&](size_t q) { return 2 * (L[q] + 2 * M[q]) * (-dmp[q]); };
    auto CX = [&](size_t q) { return 2 * ((L[q] + 2 * M[q]) * (-dmp[q]) - 2 * M[q] * (-dms[q])); };
    auto CS = [&](size_t q) { return 2 * M[q] * (-dms[q]); };
#pragma omp parallel for schedule(static)
    for (int ix = p; ix < g.nxp - p; ++ix)
        for (int iz = p; iz < g.nzp - p; ++iz) {
            const size_t c = g.at(ix, iz);
            R txx_x = 0, txx_z = 0, tzz_x = 0, tzz_z = 0, txz_x = 0, txz_z = 0;
            for (int i = 1; i <= NOP; ++i) {
                size_t xp = c + (size_t)i * nzp, xm = c - (size_t)(i - 1) * nzp;
                size_t zp = c + i, zm = c - i + 1;
                txx_x += (R)C8[i] * (CP(xp) * s.txx[xp] - CP(xm) * s.txx[xm]);
                txx_z += (R)C8[i] * (CX(zp) * s.txx[zp] - CX(zm) * s.txx[zm]);
                tzz_x += (R)C8[i] * (CX(xp) * s.tzz[xp] - CX(xm) * s.tzz[xm]);
                tzz_z += (R)C8[i] * (CP(zp) * s.tzz[zp] - CP(zm) * s.tzz[zm]);
                size_t a1 = c + (size_t)(i - 1) * nzp, a0 = c - (size_t)i * nzp;
                txz_x += (R)C8[i] * (CS(a1) * s.txz[a1] - CS(a1) * s.txz[a0]);
                txz_z += (R)C8[i] * (CS(c + i - 1) * s.txz[c + i - 1] - CS(c - i) * s.txz[c - i]);
            }
            R br = (R)1 / m.rho[c];
            bvx[c] = bvx[c] + dir * br * (txx_x / dx + tzz_x / dx + txz_z / dz);
            bvz[c] = bvz[c] + dir * br * (tzz_z / dz + txx_z / dz + txz_x / dx);
        }
}

// FD2DGPU.cu:3674-3731 record_edge_wavefield / load_edge_wavefield (10-arg):
// 2N = 8 bands straddling the PML edge on four sides.
template <typename R>
struct EdgeBands {
    int nx, nz, N;
    std::vector<R> up, down, left, right;  // for one field, [nt][...]
    void init(int nt, int nx_, int nz_, int N_) {
        nx = nx_; nz = nz_; N = N_;
        up.assign((size_t)nt * nx * 2 * N, 0); down.assign(up.size(), 0);
        left.assign((size_t)nt * 2 * N * nz, 0); right.assign(left.size(), 0);
    }
};
template <typename R>
void edge_io(const Grid<R>& g, std::vector<R>& f, EdgeBands<R>& e, int it, bool save) {
    const int pml = g.pml, nzp = g.nzp, nxp = g.nxp, N = e.N, nx = e.nx, nz = e.nz;
    for (int ix = 0; ix < nxp; ++ix)
        for (int iz = 0; iz < nzp; ++iz) {
            size_t c = g.at(ix, iz);
            auto io = [&](std::vector<R>& buf, size_t k) {
                if (save) buf[k] = f[c]; else f[c] = buf[k];
            };
            if (ix >= pml && ix < nxp - pml && iz >= pml - N && iz < pml + N)
                io(e.up, (size_t)it * nx * 2 * N + (size_t)(ix - pml) * 2 * N + (iz - (pml - N)));
            if (ix >= pml && ix < nxp - pml && iz >= nzp - pml - N && iz < nzp - pml + N)
                io(e.down, (size_t)it * nx * 2 * N + (size_t)(ix - pml) * 2 * N + (iz - (nzp - pml - N)));
            if (ix >= pml - N && ix < pml + N && iz >= pml && iz < nzp - pml)
                io(e.left, (size_t)it * 2 * N * nz + (size_t)(ix - (pml - N)) * nz + (iz - pml));
            if (ix >= nxp - pml - N && ix < nxp - pml + N && iz >= pml && iz < nzp - pml)
                io(e.right, (size_t)it * 2 * N * nz + (size_t)(ix - (nxp - pml - N)) * nz + (iz - pml));
        }
}

// FD2DGPU.cu:3245-3263 doimage_gradient (first overload): PP += -2 vp^2 rho lam_g,
// PS += -2 vs^2 rho mu_g at global column col0+ix (image stride allnz = nz).
template <typename R>
void image_gradient(const Grid<R>& g, const Strain<R>& s, const Elastic<R>& bk, const Medium<R>& m,
                    int col0, int allnz, R* pp, R* ps) {
    const int pml = g.pml, nx = g.nxp - 2 * pml, nz = g.nzp - 2 * pml;
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nx; ++ix)
        for (int iz = 0; iz < nz; ++iz) {
            size_t c = g.at(ix + pml, iz + pml);
            R lg = (s.gx_vx[c] + s.gz_vz[c]) * (bk.txx[c] + bk.tzz[c]);
            R mg = -2 * (s.gx_vx[c] * bk.tzz[c] + s.gz_vz[c] * bk.txx[c]) + (s.gz_vx[c] + s.gx_vz[c]) * bk.txz[c];
            R cpp = -2 * m.vp[c] * m.vp[c] * m.rho[c] * lg;
            R cps = -2 * m.vs[c] * m.vs[c] * m.rho[c] * mg;
            size_t o = (size_t)(col0 + ix) * allnz + iz;
            pp[o] += cpp;
            ps[o] += cps;
        }
}
// FD2DGPU.cu:3235-3243 doimage_illum.
template <typename R>
void image_illum(const Grid<R>& g, const Elastic<R>& f, int col0, int allnz, R* il) {
    const int pml = g.pml, nx = g.nxp - 2 * pml, nz = g.nzp - 2 * pml;
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nx; ++ix)
        for (int iz = 0; iz < nz; ++iz) {
            size_t c = g.at(ix + pml, iz + pml);
            il[(size_t)(col0 + ix) * allnz + iz] += f.vx[c] * f.vx[c] + f.vz[c] * f.vz[c];
        }
}

}  // namespace orc
