// Acoustic kernels (gfx950), batched over shot slots.
//
// Plain-L2 Born / adjoint pair: the mu = 0 specialisation of the elastic live
// stencils (born.cu:215-349: same stagger and [nop, n-nop) bounds, with
// txx = tzz = p and txz = 0), random boundary, no damping.  The adjoint
// kernels are the exact transposes of the forward ones (masked to the update
// region), run backward together with the exact reversal of the background.
#include "common.hpp"
#include "kernels.hpp"

namespace ds2d {

namespace {

__device__ __forceinline__ bool in_u(const SlotDesc& S, int ix, int iz) {
    return ix >= NOP && ix < S.nxp - NOP && iz >= NOP && iz < S.nzp - NOP;
}

struct Tile {
    int ix, iz;
    size_t c;
};
__device__ __forceinline__ bool tile_cell(const SlotDesc& S, Tile& t) {
    t.iz = blockIdx.x * BZ + threadIdx.x;
    t.ix = blockIdx.y * BX + threadIdx.y;
    if (!in_u(S, t.ix, t.iz)) return false;
    t.c = (size_t)S.off + (size_t)t.ix * S.nzp + t.iz;
    return true;
}

template <typename R>
__device__ __forceinline__ R dpx(const R* f, size_t c, int s) {
    return (R)DS2D_C1 * (f[c + s] - f[c]) + (R)DS2D_C2 * (f[c + 2 * s] - f[c - s]) +
           (R)DS2D_C3 * (f[c + 3 * s] - f[c - 2 * s]) + (R)DS2D_C4 * (f[c + 4 * s] - f[c - 3 * s]);
}
template <typename R>
__device__ __forceinline__ R dmx(const R* f, size_t c, int s) {
    return (R)DS2D_C1 * (f[c] - f[c - s]) + (R)DS2D_C2 * (f[c + s] - f[c - 2 * s]) +
           (R)DS2D_C3 * (f[c + 2 * s] - f[c - 3 * s]) + (R)DS2D_C4 * (f[c + 3 * s] - f[c - 4 * s]);
}
// v += sign (dt/d) b D+ (p + w delta_src) of the FP64 background at cell c.
__device__ __forceinline__ void bg_v_step(const SlotDesc& S, AcouFieldD bg, size_t c, size_t s, double w, double bu,
                                          double ax, double az, double sign) {
    const size_t src = (size_t)S.off + (size_t)S.isx * S.nzp + S.isz;
    auto P = [&](size_t j) { return j == src ? bg.p[j] + w : bg.p[j]; };
    const double a = (double)DS2D_C1 * (P(c + s) - P(c)) + (double)DS2D_C2 * (P(c + 2 * s) - P(c - s)) +
                     (double)DS2D_C3 * (P(c + 3 * s) - P(c - 2 * s)) + (double)DS2D_C4 * (P(c + 4 * s) - P(c - 3 * s));
    const double b = (double)DS2D_C1 * (P(c + 1) - P(c)) + (double)DS2D_C2 * (P(c + 2) - P(c - 1)) +
                     (double)DS2D_C3 * (P(c + 3) - P(c - 2)) + (double)DS2D_C4 * (P(c + 4) - P(c - 3));
    bg.vx[c] += sign * (ax * bu * a);
    bg.vz[c] += sign * (az * bu * b);
}

}  // namespace

// Physical-window cell test and its index in the per-step history of the
// background increment dp (hist[hist_off + (it*nx + ix-pml)*nz + iz-pml]).
__device__ __forceinline__ bool hist_cell(const SlotDesc& S, int ix, int iz, int it, size_t& h) {
    const int lx = ix - S.pml, lz = iz - S.pml;
    if (lx < 0 || lx >= S.nx || lz < 0 || lz >= S.nz) return false;
    h = (size_t)S.hist_off + ((size_t)it * S.nx + lx) * S.nz + lz;
    return true;
}

// bg (FP64): v += sign * (dt/d) b D+ (p + w delta_src);  born: dv += (dt/d) b D+ dp.
__global__ __launch_bounds__(256) void k_acou_v(AcouFieldD bg, AcouField bn, AcouMed m, const float* __restrict__ wav,
                                                const SlotDesc* __restrict__ slots, float ax, float az, double axd,
                                                double azd, int it, float sign, int do_bg, int do_born) {
    const SlotDesc S = slots[blockIdx.z];
    Tile t;
    if (!tile_cell(S, t)) return;
    const size_t c = t.c, s = S.nzp;
    const float bu = m.b[c];
    if (do_bg) bg_v_step(S, bg, c, s, (double)wav[it], (double)bu, axd, azd, (double)sign);
    if (do_born) {
        bn.vx[c] += ax * bu * dpx(bn.p, c, (int)s);
        bn.vz[c] += az * bu * dpx(bn.p, c, 1);
    }
}

// bg: dp = k (dt/dx D- vx + dt/dz D- vz); p += dp (hist_mode 1 also stores dp
// on the physical window); without bg, hist_mode 2 reads dp back.
// born: dpb += k D- dv + 2 m dp.
__global__ __launch_bounds__(256) void k_acou_p(AcouFieldD bg, AcouField bn, AcouMed m, float* hist,
                                                const SlotDesc* __restrict__ slots, float ax, float az, double axd,
                                                double azd, int it, int do_bg, int do_born, int hist_mode,
                                                const int* __restrict__ rmap, float* rbg, float* rbn, int nt) {
    const SlotDesc S = slots[blockIdx.z];
    Tile t;
    if (!tile_cell(S, t)) return;
    const size_t c = t.c;
    const int s = S.nzp;
    const float kk = m.k[c];
    size_t h = 0;
    const bool inside = hist_cell(S, t.ix, t.iz, it, h);
    float dp = 0.f;
    if (do_bg) {
        const double d = (double)kk * (axd * dmx(bg.vx, c, s) + azd * dmx(bg.vz, c, 1));
        bg.p[c] += d;
        dp = (float)d;
        if (hist_mode == 1 && inside) hist[h] = dp;
    } else if (hist_mode == 2 && inside) {
        dp = hist[h];
    }
    if (do_born) {
        const float bd = kk * (ax * dmx(bn.vx, c, s) + az * dmx(bn.vz, c, 1));
        bn.p[c] += bd + 2 * m.m[c] * dp;
    }
    // surface recorder (FD2DGPU.cu:3093-3101): p after this step at (rec_ix, igz)
    if (t.iz == S.igz && (rbg || rbn)) {
        const int r = rmap[S.rmap_off + t.ix];
        if (r >= 0) {
            const size_t o = (size_t)S.rec_off + (size_t)r * nt + it;
            if (rbg) rbg[o] = (float)bg.p[c];
            if (rbn) rbn[o] = bn.p[c];
        }
    }
}

// Backward kernel 1 (lam holds mu_n after the receiver injection):
//   d = k D- v^{n+1} (reconstruct) or the stored increment (hist != null);
//   grad += 2 d lp;  reconstruct: p^n = p^{n+1} - d;
//   lv += -(dt/d) D+ (k lp) with k lp masked to the update region.
__global__ __launch_bounds__(256) void k_acou_adj_p(AcouFieldD bg, AcouField lm, AcouMed m, float* grad,
                                                    const float* __restrict__ hist,
                                                    const SlotDesc* __restrict__ slots, float ax, float az,
                                                    double axd, double azd, int it) {
    const SlotDesc S = slots[blockIdx.z];
    Tile t;
    if (!tile_cell(S, t)) return;
    const size_t c = t.c;
    const int s = S.nzp;
    if (hist) {
        size_t h = 0;
        if (hist_cell(S, t.ix, t.iz, it, h)) grad[c] += 2 * hist[h] * lm.p[c];
    } else {
        const double d = (double)m.k[c] * (axd * dmx(bg.vx, c, s) + azd * dmx(bg.vz, c, 1));
        grad[c] += 2 * (float)d * lm.p[c];
        bg.p[c] -= d;
    }
    auto W = [&](int jx, int jz) {
        return in_u(S, jx, jz) ? m.k[(size_t)S.off + (size_t)jx * s + jz] * lm.p[(size_t)S.off + (size_t)jx * s + jz]
                               : 0.f;
    };
    const int ix = t.ix, iz = t.iz;
    const float gx = DS2D_C1 * (W(ix + 1, iz) - W(ix, iz)) + DS2D_C2 * (W(ix + 2, iz) - W(ix - 1, iz)) +
                     DS2D_C3 * (W(ix + 3, iz) - W(ix - 2, iz)) + DS2D_C4 * (W(ix + 4, iz) - W(ix - 3, iz));
    const float gz = DS2D_C1 * (W(ix, iz + 1) - W(ix, iz)) + DS2D_C2 * (W(ix, iz + 2) - W(ix, iz - 1)) +
                     DS2D_C3 * (W(ix, iz + 3) - W(ix, iz - 2)) + DS2D_C4 * (W(ix, iz + 4) - W(ix, iz - 3));
    lm.vx[c] += -(ax * gx);
    lm.vz[c] += -(az * gz);
}

// Backward kernel 2: (reconstruct only) background velocity reversal with the
// transient source, v^n = v^{n+1} - (dt/d) b D+ (p^n + w_n delta_src);
//   lp += -(dt/d) D- (b lv) with b lv masked to the update region.
__global__ __launch_bounds__(256) void k_acou_adj_v(AcouFieldD bg, AcouField lm, AcouMed m,
                                                    const float* __restrict__ wav, const SlotDesc* __restrict__ slots,
                                                    float ax, float az, double axd, double azd, int it, int do_bg,
                                                    const int* __restrict__ rmap, const float* __restrict__ res,
                                                    int nt) {
    const SlotDesc S = slots[blockIdx.z];
    Tile t;
    if (!tile_cell(S, t)) return;
    const size_t c = t.c, s = S.nzp;
    if (do_bg) bg_v_step(S, bg, c, s, (double)wav[it], (double)m.b[c], axd, azd, -1.0);
    const int ix = t.ix, iz = t.iz;
    auto UX = [&](int jx, int jz) {
        const size_t j = (size_t)S.off + (size_t)jx * s + jz;
        return in_u(S, jx, jz) ? m.b[j] * lm.vx[j] : 0.f;
    };
    auto UZ = [&](int jx, int jz) {
        const size_t j = (size_t)S.off + (size_t)jx * s + jz;
        return in_u(S, jx, jz) ? m.b[j] * lm.vz[j] : 0.f;
    };
    const float gx = DS2D_C1 * (UX(ix, iz) - UX(ix - 1, iz)) + DS2D_C2 * (UX(ix + 1, iz) - UX(ix - 2, iz)) +
                     DS2D_C3 * (UX(ix + 2, iz) - UX(ix - 3, iz)) + DS2D_C4 * (UX(ix + 3, iz) - UX(ix - 4, iz));
    const float gz = DS2D_C1 * (UZ(ix, iz) - UZ(ix, iz - 1)) + DS2D_C2 * (UZ(ix, iz + 1) - UZ(ix, iz - 2)) +
                     DS2D_C3 * (UZ(ix, iz + 2) - UZ(ix, iz - 3)) + DS2D_C4 * (UZ(ix, iz + 3) - UZ(ix, iz - 4));
    lm.p[c] += -(ax * gx + az * gz);
    // receiver injection of the next backward step (it - 1), fused here: this
    // cell's lp is final for step it and no other cell reads it in this kernel
    if (res && it > 0 && iz == S.igz) {
        const int r = rmap[S.rmap_off + ix];
        if (r >= 0) lm.p[c] += res[(size_t)S.rec_off + (size_t)r * nt + (it - 1)];
    }
}

// lp[rec] += res[r][it] (first backward step; later steps inject inside k_acou_adj_v).
__global__ void k_acou_inj(float* lp, const float* __restrict__ res, const int* __restrict__ rix,
                           const SlotDesc* __restrict__ slots, int nt, int it) {
    const SlotDesc S = slots[blockIdx.y];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.ntr) return;
    const size_t c = (size_t)S.off + (size_t)rix[S.rix_off + r] * S.nzp + S.igz;
    lp[c] += res[(size_t)S.rec_off + (size_t)r * nt + it];
}

static dim3 cgrid(const BatchShape& b) { return dim3(ceil_div(b.nzp, BZ), ceil_div(b.max_nxp, BX), b.nslots); }

// FP64 background coefficients dt/dx, dt/dz from the FP32 grid constants
static double dxd(const BatchShape& b) { return (double)b.dt / (double)b.dx; }
static double dzd(const BatchShape& b) { return (double)b.dt / (double)b.dz; }

void launch_acou_v(hipStream_t st, const BatchShape& b, AcouFieldD bg, AcouField born, AcouMed m, const float* wav,
                   int it, float sign, int do_bg, int do_born) {
    hipLaunchKernelGGL(k_acou_v, cgrid(b), dim3(BZ, BX), 0, st, bg, born, m, wav, b.slots, b.dt / b.dx, b.dt / b.dz,
                       dxd(b), dzd(b), it, sign, do_bg, do_born);
}
void launch_acou_p(hipStream_t st, const BatchShape& b, AcouFieldD bg, AcouField born, AcouMed m, float* hist,
                   float* rec_bg, float* rec_born, int it, int do_bg, int do_born, int hist_mode, const int* rmap) {
    hipLaunchKernelGGL(k_acou_p, cgrid(b), dim3(BZ, BX), 0, st, bg, born, m, hist, b.slots, b.dt / b.dx, b.dt / b.dz,
                       dxd(b), dzd(b), it, do_bg, do_born, hist_mode, rmap, rec_bg, rec_born, b.nt);
}
void launch_acou_inj(hipStream_t st, const BatchShape& b, float* lp, const float* res, int it) {
    hipLaunchKernelGGL(k_acou_inj, dim3(ceil_div(b.max_ntr, 64), b.nslots), dim3(64), 0, st, lp, res, b.rix,
                       b.slots, b.nt, it);
}
void launch_acou_adj_p(hipStream_t st, const BatchShape& b, AcouFieldD bg, AcouField lam, AcouMed m, float* grad,
                       const float* hist, int it) {
    hipLaunchKernelGGL(k_acou_adj_p, cgrid(b), dim3(BZ, BX), 0, st, bg, lam, m, grad, hist, b.slots, b.dt / b.dx,
                       b.dt / b.dz, dxd(b), dzd(b), it);
}
void launch_acou_adj_v(hipStream_t st, const BatchShape& b, AcouFieldD bg, AcouField lam, AcouMed m,
                       const float* wav, int it, int do_bg, const int* rmap, const float* res) {
    hipLaunchKernelGGL(k_acou_adj_v, cgrid(b), dim3(BZ, BX), 0, st, bg, lam, m, wav, b.slots, b.dt / b.dx,
                       b.dt / b.dz, dxd(b), dzd(b), it, do_bg, rmap, res, b.nt);
}

// ---------------------------------------------------------------------------
// Reference acoustic layout (FD2DGPU.cu:253-340): split p with one damping
// array, mirrored velocity stagger, vz uses dt/dx, source on vz, p recorded.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_aref_p(float* px, float* pz, float* p, const float* __restrict__ vx,
                                                const float* __restrict__ vz, const float* __restrict__ k,
                                                const float* __restrict__ damp, const SlotDesc* __restrict__ slots,
                                                float dt, float dx, float dz) {
    const SlotDesc S = slots[blockIdx.z];
    const int iz = blockIdx.x * BZ + threadIdx.x, ix = blockIdx.y * BX + threadIdx.y;
    if (iz > S.nzp - NOP - 1 || ix > S.nxp - NOP - 1 || iz < NOP - 1 || ix < NOP - 1) return;
    const size_t c = (size_t)S.off + (size_t)ix * S.nzp + iz;
    const float d1 = 1.f - dt * damp[c] / 2, d2 = 1.f + dt * damp[c] / 2;
    const float tvx = dpx(vx, c, S.nzp), tvz = dpx(vz, c, 1);
    const float a = (d1 * px[c] - k[c] * (dt / dx) * tvx) / d2;
    const float b = (d1 * pz[c] - k[c] * (dt / dz) * tvz) / d2;
    px[c] = a; pz[c] = b; p[c] = a + b;
}
__global__ __launch_bounds__(256) void k_aref_v(const float* __restrict__ p, float* vx, float* vz,
                                                const float* __restrict__ rho, const float* __restrict__ damp,
                                                const SlotDesc* __restrict__ slots, float dt, float dx) {
    const SlotDesc S = slots[blockIdx.z];
    const int iz = blockIdx.x * BZ + threadIdx.x, ix = blockIdx.y * BX + threadIdx.y;
    const size_t c = (size_t)S.off + (size_t)ix * S.nzp + iz;
    if (ix >= NOP && iz >= NOP && ix <= S.nxp - NOP && iz <= S.nzp - NOP - 1) {  // GPUcalculate_vx
        const float d1 = 1.f - dt * damp[c] / 2, d2 = 1.f + dt * damp[c] / 2;
        vx[c] = (d1 * vx[c] - (1.f / rho[c]) * (dt / dx) * dmx(p, c, S.nzp)) / d2;
    }
    if (ix >= NOP && iz >= NOP && ix <= S.nxp - NOP - 1 && iz <= S.nzp - NOP) {  // GPUcalculate_vz
        const float d1 = 1.f - dt * damp[c] / 2, d2 = 1.f + dt * damp[c] / 2;
        vz[c] = (d1 * vz[c] - (1.f / rho[c]) * (dt / dx) * dmx(p, c, 1)) / d2;
    }
}
// FD2DGPU.cu:327-340 addS on vz at (sxnum+pml, pml), then recordD (342-353).
__global__ void k_aref_src_rec(float* vz, const float* __restrict__ p, const float* __restrict__ wav, float* rec,
                               const SlotDesc* __restrict__ slots, int sxnum, int nt, int it) {
    const SlotDesc S = slots[0];
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix == 0) vz[(size_t)S.off + (size_t)(sxnum + S.pml) * S.nzp + S.pml] += wav[it];
    if (ix < S.nx) rec[(size_t)ix * nt + it] = p[(size_t)S.off + (size_t)(ix + S.pml) * S.nzp + S.pml];
}

void launch_acou_ref_step(hipStream_t st, const BatchShape& b, float* px, float* pz, float* p, float* vx, float* vz,
                          const float* k, const float* rho, const float* damp, const float* wav, int sxnum,
                          float* rec, int it) {
    hipLaunchKernelGGL(k_aref_p, cgrid(b), dim3(BZ, BX), 0, st, px, pz, p, vx, vz, k, damp, b.slots, b.dt, b.dx,
                       b.dz);
    hipLaunchKernelGGL(k_aref_v, cgrid(b), dim3(BZ, BX), 0, st, p, vx, vz, rho, damp, b.slots, b.dt, b.dx);
    hipLaunchKernelGGL(k_aref_src_rec, dim3(ceil_div(b.max_nxp, 64)), dim3(64), 0, st, vz, p, wav, rec, b.slots,
                       sxnum, b.nt, it);
}

}  // namespace ds2d

namespace ds2d {
// FD2DGPU.cu:505-519: dampX then dampZ write the same array, so the z profile
// wins where both strips overlap.
__global__ void k_combine_damp(const float* __restrict__ dampx, const float* __restrict__ dampz, float* damp,
                               const SlotDesc* __restrict__ slots) {
    const SlotDesc S = slots[blockIdx.z];
    const int iz = blockIdx.x * BZ + threadIdx.x, ix = blockIdx.y * BX + threadIdx.y;
    if (ix >= S.nxp || iz >= S.nzp) return;
    const size_t c = (size_t)S.off + (size_t)ix * S.nzp + iz;
    damp[c] = (iz < S.pml || iz > S.nzp - S.pml) ? dampz[c] : dampx[c];
}
void launch_combine_damp(hipStream_t st, const BatchShape& b, const float* dampx, const float* dampz, float* damp) {
    hipLaunchKernelGGL(k_combine_damp, dim3(ceil_div(b.nzp, BZ), ceil_div(b.max_nxp, BX), b.nslots), dim3(BZ, BX), 0,
                       st, dampx, dampz, damp, b.slots);
}
}  // namespace ds2d
