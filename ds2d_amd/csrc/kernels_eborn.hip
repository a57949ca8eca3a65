// Plain-L2 elastic Born pair for gfx950, batched over shot slots: PP/PS
// image -> scattered (vz, vx) records, and the exact transpose.
//
// Per step of the forward map: background velocity (transient source on txx,
// tzz at the source cell, FD2DGPU.cu:2039-2049) and scattered velocity
// (update_v, born.cu:215-279) in one kernel; background stress, scattered
// stress and the Born source of add_born_source (born.cu:351-405) on
// [pml, n-pml)^2 from the new background velocity in the second kernel, which
// also records the scattered velocity.  Random boundary, no damping, unsplit
// fields, update region [nop, n-nop)^2.
//
// The adjoint runs the transposed steps backward: v* += -(dt/d) D (C s*) and
// s* += -(dt/d) D ((1/rho) v*), each operand masked to the update region (the
// products C s* and (1/rho) v* are written by the kernel that updates s* or
// v*, so cells outside the region keep their zero), then the transpose of the
// Born source (the image_gradient products of FD2DGPU.cu:3336-3339) from the
// background strains of that step, read from the resident history or rebuilt
// by reversing the background.
#include "common.hpp"
#include "kernels.hpp"

namespace ds2d {

namespace {

__device__ __forceinline__ bool in_u(const SlotDesc& S, int ix, int iz) {
    return ix >= NOP && ix < S.nxp - NOP && iz >= NOP && iz < S.nzp - NOP;
}

struct Cell {
    int ix, iz;
    size_t c;
};
__device__ __forceinline__ bool cell_u(const SlotDesc& S, Cell& t) {
    t.iz = blockIdx.x * BZ + threadIdx.x;
    t.ix = blockIdx.y * BX + threadIdx.y;
    if (!in_u(S, t.ix, t.iz)) return false;
    t.c = (size_t)S.off + (size_t)t.ix * S.nzp + t.iz;
    return true;
}

// Staggered 8th-order differences; f(k) is the operand k cells away along the
// difference axis.  plus: sum c_i (f(i) - f(1-i)); minus: sum c_i (f(i-1) - f(-i)).
template <typename F>
__device__ __forceinline__ float dplus(F f) {
    return DS2D_C1 * (f(1) - f(0)) + DS2D_C2 * (f(2) - f(-1)) + DS2D_C3 * (f(3) - f(-2)) + DS2D_C4 * (f(4) - f(-3));
}
template <typename F>
__device__ __forceinline__ float dminus(F f) {
    return DS2D_C1 * (f(0) - f(-1)) + DS2D_C2 * (f(1) - f(-2)) + DS2D_C3 * (f(2) - f(-3)) + DS2D_C4 * (f(3) - f(-4));
}
// operand k cells away along x (stride s) or z (stride 1) of array a at cell c
#define EB_X(a) [&](int k) { return (a)[c + (long long)k * s]; }
#define EB_Z(a) [&](int k) { return (a)[c + k]; }

// Physical-window cell (the Born source region [pml, n-pml)^2) and its index
// in the per-step strain history hist[hist_off + (it*nx + ix-pml)*nz + iz-pml].
__device__ __forceinline__ bool hist_at(const SlotDesc& S, int ix, int iz, int it, size_t& h) {
    const int lx = ix - S.pml, lz = iz - S.pml;
    if (lx < 0 || lx >= S.nx || lz < 0 || lz >= S.nz) return false;
    h = (size_t)S.hist_off + ((size_t)it * S.nx + lx) * S.nz + lz;
    return true;
}

// Background strains of the Born source at cell c (add_born_source's
// vxx/dx, vzz/dz and vzx/dx + vxz/dz).
__device__ __forceinline__ void strains(const EbField& f, long long c, long long s, float idx, float idz, float& exx,
                                        float& ezz, float& exz) {
    exx = dminus(EB_X(f.vx)) * idx;
    ezz = dminus(EB_Z(f.vz)) * idz;
    exz = dplus(EB_X(f.vz)) * idx + dplus(EB_Z(f.vx)) * idz;
}

}  // namespace

// Velocity of the background (sign -1 undoes a step; w = source sample) and of
// the scattered field.
__global__ __launch_bounds__(256) void k_eb_v(EbField bg, EbField b, const float* __restrict__ rho,
                                              const SlotDesc* __restrict__ slots, float ax, float az, float w,
                                              float sign, int do_bg, int do_b) {
    const SlotDesc S = slots[blockIdx.z];
    Cell t;
    if (!cell_u(S, t)) return;
    const long long c = (long long)t.c, s = S.nzp;
    const float br = 1.f / rho[c];
    if (do_bg) {
        const long long src = S.off + (long long)S.isx * S.nzp + S.isz;
        auto txx = [&](int k) {
            const long long j = c + (long long)k * s;
            return j == src ? bg.txx[j] + w : bg.txx[j];
        };
        auto tzz = [&](int k) {
            const long long j = c + k;
            return j == src ? bg.tzz[j] + w : bg.tzz[j];
        };
        bg.vx[c] += sign * (br * ax * dplus(txx) + br * az * dminus(EB_Z(bg.txz)));
        bg.vz[c] += sign * (br * ax * dminus(EB_X(bg.txz)) + br * az * dplus(tzz));
    }
    if (do_b) {
        b.vx[c] += br * ax * dplus(EB_X(b.txx)) + br * az * dminus(EB_Z(b.txz));
        b.vz[c] += br * ax * dminus(EB_X(b.txz)) + br * az * dplus(EB_Z(b.tzz));
    }
}

// Stress of the background (sign -1 undoes a step; hist_mode 1 stores its
// strains) and of the scattered field with the Born source (strains from the
// background, or hist_mode 2: from the history); records the scattered
// velocity after the step.
__global__ __launch_bounds__(256) void k_eb_s(EbField bg, EbField b, EbMed m, float* hist, long long hstride,
                                              const SlotDesc* __restrict__ slots, float ax, float az, float idx,
                                              float idz, int it, float sign, int do_bg, int do_b, int hist_mode,
                                              const int* __restrict__ rmap, float* rec_z, float* rec_x, int nt) {
    const SlotDesc S = slots[blockIdx.z];
    Cell t;
    if (!cell_u(S, t)) return;
    const long long c = (long long)t.c, s = S.nzp;
    const float l = m.lam[c], mu = m.mu[c], l2m = l + 2 * mu;
    size_t h = 0;
    const bool inside = hist_at(S, t.ix, t.iz, it, h);
    float exx = 0.f, ezz = 0.f, exz = 0.f;
    if (do_bg) {
        const float dvx = dminus(EB_X(bg.vx)), dvz = dminus(EB_Z(bg.vz));
        const float dvxz = dplus(EB_Z(bg.vx)), dvzx = dplus(EB_X(bg.vz));
        bg.txx[c] += sign * (ax * l2m * dvx + az * l * dvz);
        bg.tzz[c] += sign * (ax * l * dvx + az * l2m * dvz);
        bg.txz[c] += sign * (mu * (ax * dvzx + az * dvxz));
        exx = dvx * idx;
        ezz = dvz * idz;
        exz = dvzx * idx + dvxz * idz;
        if (hist_mode == 1 && inside) {
            hist[h] = exx;
            hist[h + hstride] = ezz;
            hist[h + 2 * hstride] = exz;
        }
    } else if (hist_mode == 2 && inside) {
        exx = hist[h];
        ezz = hist[h + hstride];
        exz = hist[h + 2 * hstride];
    }
    if (do_b) {
        const float dvx = dminus(EB_X(b.vx)), dvz = dminus(EB_Z(b.vz));
        const float dvxz = dplus(EB_Z(b.vx)), dvzx = dplus(EB_X(b.vz));
        float txx = b.txx[c] + ax * l2m * dvx + az * l * dvz;
        float tzz = b.tzz[c] + ax * l * dvx + az * l2m * dvz;
        float txz = b.txz[c] + mu * (ax * dvzx + az * dvxz);
        if (inside) {
            const float mp = -m.dmp[c], ms = -m.dms[c];
            txx += 2 * l2m * mp * exx + 2 * (l2m * mp - 2 * mu * ms) * ezz;
            tzz += 2 * (l2m * mp - 2 * mu * ms) * exx + 2 * l2m * mp * ezz;
            txz += 2 * mu * ms * exz;
        }
        b.txx[c] = txx;
        b.tzz[c] = tzz;
        b.txz[c] = txz;
        if (t.iz == S.igz && rec_z) {
            const int r = rmap[S.rmap_off + t.ix];
            if (r >= 0) {
                const size_t o = (size_t)S.rec_off + (size_t)r * nt + it;
                rec_z[o] = b.vz[c];
                rec_x[o] = b.vx[c];
            }
        }
    }
}

// Adjoint, part 1: v* += inj(step it_inj) - (dt/d) D (C s*), reading the
// products C s* that part 2 left (zero outside the update region); leaves
// u = (1/rho) v* for part 2.
// undo = 1 also takes one step off the background velocity, v -= (dt/d) D sigma
// with the source sample w (the background stress is already one step back).
__global__ __launch_bounds__(256) void k_eb_adj_v(EbField lm, EbAux x, EbField bg, EbMed m,
                                                  const SlotDesc* __restrict__ slots, float ax, float az,
                                                  const int* __restrict__ rmap, const float* res_z,
                                                  const float* res_x, int nt, int it_inj, int undo, float w) {
    const SlotDesc S = slots[blockIdx.z];
    Cell t;
    if (!cell_u(S, t)) return;
    const long long c = (long long)t.c, s = S.nzp;
    const float br = 1.f / m.rho[c];
    if (undo) {
        const long long src = S.off + (long long)S.isx * S.nzp + S.isz;
        auto txx = [&](int k) {
            const long long j = c + (long long)k * s;
            return j == src ? bg.txx[j] + w : bg.txx[j];
        };
        auto tzz = [&](int k) {
            const long long j = c + k;
            return j == src ? bg.tzz[j] + w : bg.tzz[j];
        };
        bg.vx[c] += -(br * ax * dplus(txx) + br * az * dminus(EB_Z(bg.txz)));
        bg.vz[c] += -(br * ax * dminus(EB_X(bg.txz)) + br * az * dplus(tzz));
    }
    float ix_ = 0.f, iz_ = 0.f;
    if (it_inj >= 0 && t.iz == S.igz) {
        const int r = rmap[S.rmap_off + t.ix];
        if (r >= 0) {
            const size_t o = (size_t)S.rec_off + (size_t)r * nt + it_inj;
            ix_ = res_x[o];
            iz_ = res_z[o];
        }
    }
    const float vx = lm.vx[c] + ix_ - (ax * dplus(EB_X(x.wa)) + az * dminus(EB_Z(x.wc)));
    const float vz = lm.vz[c] + iz_ - (az * dplus(EB_Z(x.wb)) + ax * dminus(EB_X(x.wc)));
    lm.vx[c] = vx;
    lm.vz[c] = vz;
    x.ux[c] = br * vx;
    x.uz[c] = br * vz;
}

// Adjoint, part 2: s* += -(dt/d) D u (u = (1/rho) v* from part 1, zero outside
// the update region); leaves C s* for the next part 1; then the gradient of
// step it at the physical window: the strains come from the history
// (hist != null) or from the background velocity.
// undo = 1 (no history) also takes the background stress one step back,
// sigma -= (dt/d) C D v, from the same velocity differences as the strains.
__global__ __launch_bounds__(256) void k_eb_adj_s(EbField lm, EbAux x, EbField bg, EbMed m,
                                                  const float* __restrict__ hist, long long hstride, float* gpp,
                                                  float* gps, const SlotDesc* __restrict__ slots, float ax, float az,
                                                  float idx, float idz, int it, int undo) {
    const SlotDesc S = slots[blockIdx.z];
    Cell t;
    if (!cell_u(S, t)) return;
    const long long c = (long long)t.c, s = S.nzp;
    const float txx = lm.txx[c] - ax * dminus(EB_X(x.ux));
    const float tzz = lm.tzz[c] - az * dminus(EB_Z(x.uz));
    const float txz = lm.txz[c] - (az * dplus(EB_Z(x.ux)) + ax * dplus(EB_X(x.uz)));
    lm.txx[c] = txx;
    lm.tzz[c] = tzz;
    lm.txz[c] = txz;
    const float l = m.lam[c], mu = m.mu[c], l2m = l + 2 * mu;
    x.wa[c] = l2m * txx + l * tzz;
    x.wb[c] = l * txx + l2m * tzz;
    x.wc[c] = mu * txz;
    size_t h;
    const bool inside = hist_at(S, t.ix, t.iz, it, h);
    float exx = 0.f, ezz = 0.f, exz = 0.f;
    if (hist) {
        if (!inside) return;
        exx = hist[h];
        ezz = hist[h + hstride];
        exz = hist[h + 2 * hstride];
    } else {
        if (!inside && !undo) return;
        const float dvx = dminus(EB_X(bg.vx)), dvz = dminus(EB_Z(bg.vz));
        const float dvxz = dplus(EB_Z(bg.vx)), dvzx = dplus(EB_X(bg.vz));
        exx = dvx * idx;
        ezz = dvz * idz;
        exz = dvzx * idx + dvxz * idz;
        if (undo) {
            bg.txx[c] += -(ax * l2m * dvx + az * l * dvz);
            bg.tzz[c] += -(ax * l * dvx + az * l2m * dvz);
            bg.txz[c] += -(mu * (ax * dvzx + az * dvxz));
        }
        if (!inside) return;
    }
    gpp[c] += -2 * l2m * (exx + ezz) * (txx + tzz);
    gps[c] += 4 * mu * (ezz * txx + exx * tzz) - 2 * mu * exz * txz;
}

namespace {
dim3 egrid(const BatchShape& b) { return dim3(ceil_div(b.nzp, BZ), ceil_div(b.max_nxp, BX), b.nslots); }
}  // namespace

void launch_eb_v(hipStream_t st, const BatchShape& b, EbField bg, EbField born, const float* rho, float w,
                 float sign, int do_bg, int do_born) {
    hipLaunchKernelGGL(k_eb_v, egrid(b), dim3(BZ, BX), 0, st, bg, born, rho, b.slots, b.dt / b.dx, b.dt / b.dz, w,
                       sign, do_bg, do_born);
}
void launch_eb_s(hipStream_t st, const BatchShape& b, EbField bg, EbField born, EbMed m, float* hist,
                 long long hstride, int it, float sign, int do_bg, int do_born, int hist_mode, const int* rmap,
                 float* rec_z, float* rec_x) {
    hipLaunchKernelGGL(k_eb_s, egrid(b), dim3(BZ, BX), 0, st, bg, born, m, hist, hstride, b.slots, b.dt / b.dx,
                       b.dt / b.dz, 1.f / b.dx, 1.f / b.dz, it, sign, do_bg, do_born, hist_mode, rmap, rec_z, rec_x,
                       b.nt);
}
void launch_eb_adj_v(hipStream_t st, const BatchShape& b, EbField lm, EbAux x, EbField bg, EbMed m,
                     const int* rmap, const float* res_z, const float* res_x, int it_inj, int undo, float w) {
    hipLaunchKernelGGL(k_eb_adj_v, egrid(b), dim3(BZ, BX), 0, st, lm, x, bg, m, b.slots, b.dt / b.dx, b.dt / b.dz,
                       rmap, res_z, res_x, b.nt, it_inj, undo, w);
}
void launch_eb_adj_s(hipStream_t st, const BatchShape& b, EbField lm, EbAux x, EbField bg, EbMed m,
                     const float* hist, long long hstride, float* gpp, float* gps, int it, int undo) {
    hipLaunchKernelGGL(k_eb_adj_s, egrid(b), dim3(BZ, BX), 0, st, lm, x, bg, m, hist, hstride, gpp, gps, b.slots,
                       b.dt / b.dx, b.dt / b.dz, 1.f / b.dx, 1.f / b.dz, it, undo);
}

}  // namespace ds2d
