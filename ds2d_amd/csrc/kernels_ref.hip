// Reference-compatible elastic kernels for gfx950, batched over shot slots.
//
// Each kernel keeps the reference kernel's loop bounds and per-cell update
// (cited below) but runs every shot of a batch in one launch: blockIdx.z is
// the slot, the block is 64 lanes along contiguous z x 4 rows along x, and
// the two reference launches that write disjoint outputs (vx/vz, txx_tzz/txz)
// are fused into one.
#include "common.hpp"
#include "kernels.hpp"

namespace ds2d {

namespace {

struct Cell {
    int ix, iz;
    size_t c;
};

__device__ __forceinline__ bool cell_in(const SlotDesc& S, int lo, Cell& k) {
    k.iz = blockIdx.x * BZ + threadIdx.x;
    k.ix = blockIdx.y * BX + threadIdx.y;
    if (k.ix < lo || k.ix >= S.nxp - lo || k.iz < lo || k.iz >= S.nzp - lo) return false;
    k.c = (size_t)S.off + (size_t)k.ix * S.nzp + k.iz;
    return true;
}

// Staggered-derivative weights in the wavefield's precision: FP32 fields use
// coeff2 rounded to float (as the reference's __constant__ array), FP64 fields
// the exact quotients.
template <typename T>
struct Cf;
template <>
struct Cf<float> {
    static constexpr float c1 = DS2D_C1, c2 = DS2D_C2, c3 = DS2D_C3, c4 = DS2D_C4;
};
template <>
struct Cf<double> {
    static constexpr double c1 = 1225.0 / 1024.0, c2 = -245.0 / 3072.0, c3 = 49.0 / 5120.0, c4 = -5.0 / 7168.0;
};

// Tap sums of the eight-point staggered derivatives.
//  dpx: sum c_i (f[ix+i] - f[ix-i+1])   (forward half-point in x)
//  dmx: sum c_i (f[ix+i-1] - f[ix-i])   (backward half-point in x)
//  dpz / dmz: same along z.
template <typename T>
__device__ __forceinline__ T dpx(const T* f, size_t c, int s) {
    return Cf<T>::c1 * (f[c + s] - f[c]) + Cf<T>::c2 * (f[c + 2 * s] - f[c - s]) +
           Cf<T>::c3 * (f[c + 3 * s] - f[c - 2 * s]) + Cf<T>::c4 * (f[c + 4 * s] - f[c - 3 * s]);
}
template <typename T>
__device__ __forceinline__ T dmx(const T* f, size_t c, int s) {
    return Cf<T>::c1 * (f[c] - f[c - s]) + Cf<T>::c2 * (f[c + s] - f[c - 2 * s]) +
           Cf<T>::c3 * (f[c + 2 * s] - f[c - 3 * s]) + Cf<T>::c4 * (f[c + 3 * s] - f[c - 4 * s]);
}

// Velocity split update of born.cu:215-279 (update_vx/update_vz) and
// FD2DGPU.cu:903-989 (GPUcalculate_elastic_vx/vz): LO selects the bounds.
template <typename T>
__device__ __forceinline__ void vel_update(EFieldT<T> f, const Med& m, size_t c, int nzp, T dt, T dx, T dz, T dir) {
    const T d1 = T(1) - dt * (T)m.dampx[c] / 2, d2 = T(1) + dt * (T)m.dampx[c] / 2;
    const T d3 = T(1) - dt * (T)m.dampz[c] / 2, d4 = T(1) + dt * (T)m.dampz[c] / 2;
    const T a = dpx(f.txx, c, nzp);
    const T b = dmx(f.txz, c, 1);
    const T cz = dpx(f.tzz, c, 1);
    const T e = dmx(f.txz, c, nzp);
    const T br = T(1) / (T)m.rho[c];
    const T nvx_x = (d1 * f.vx_x[c] + dir * br * (dt / dx) * a) / d2;
    const T nvx_z = (d3 * f.vx_z[c] + dir * br * (dt / dz) * b) / d4;
    const T nvz_x = (d1 * f.vz_x[c] + dir * br * (dt / dx) * e) / d2;
    const T nvz_z = (d3 * f.vz_z[c] + dir * br * (dt / dz) * cz) / d4;
    f.vx_x[c] = nvx_x; f.vx_z[c] = nvx_z; f.vx[c] = nvx_x + nvx_z;
    f.vz_x[c] = nvz_x; f.vz_z[c] = nvz_z; f.vz[c] = nvz_x + nvz_z;
}

}  // namespace

// born.cu:215-279 (LO = nop) or FD2DGPU.cu:903-989 (LO = pml).
template <typename T>
__global__ __launch_bounds__(256) void k_ref_v(EFieldT<T> f, Med m, const SlotDesc* __restrict__ slots, float dt,
                                               float dx, float dz, float dir, int recon) {
    const SlotDesc S = slots[blockIdx.z];
    Cell k;
    if (!cell_in(S, recon ? S.pml : NOP, k)) return;
    vel_update<T>(f, m, k.c, S.nzp, (T)dt, (T)dx, (T)dz, (T)dir);
}

// born.cu:282-349 update_txx_tzz + update_txz (recon = 0), or
// FD2DGPU.cu:993-1082 *_grad on [pml, n-pml) with strain outputs (recon = 1;
// the strains are stored in FP32 whatever the wavefield's precision).
template <typename T>
__global__ __launch_bounds__(256) void k_ref_s(EFieldT<T> f, Med m, Strain st, const SlotDesc* __restrict__ slots,
                                               float dt_, float dx_, float dz_, float dir_, int recon) {
    const SlotDesc S = slots[blockIdx.z];
    Cell k;
    if (!cell_in(S, recon ? S.pml : NOP, k)) return;
    const T dt = dt_, dx = dx_, dz = dz_, dir = dir_;
    const size_t c = k.c;
    const int nzp = S.nzp;
    const T d1 = T(1) - dt * (T)m.dampx[c] / 2, d2 = T(1) + dt * (T)m.dampx[c] / 2;
    const T d3 = T(1) - dt * (T)m.dampz[c] / 2, d4 = T(1) + dt * (T)m.dampz[c] / 2;
    const T dvx = dmx(f.vx, c, nzp);
    const T dvz = dmx(f.vz, c, 1);
    const T dvxz = dpx(f.vx, c, 1);
    const T dvzx = dpx(f.vz, c, nzp);
    const T l = m.lam[c], mu = m.mu[c], l2m = l + 2 * mu;
    const T a = f.txx_x[c], b = f.txx_z[c], cc = f.tzz_x[c], d = f.tzz_z[c];
    const T ntxx_x = (d1 * a + dir * l2m * (dt / dx) * dvx) / d2;
    const T ntxx_z = (d3 * b + dir * l * (dt / dz) * dvz) / d4;
    const T ntzz_x = (d1 * cc + dir * l * (dt / dx) * dvx) / d2;
    const T ntzz_z = (d3 * d + dir * l2m * (dt / dz) * dvz) / d4;
    f.txx_x[c] = ntxx_x; f.txx_z[c] = ntxx_z; f.tzz_x[c] = ntzz_x; f.tzz_z[c] = ntzz_z;
    f.txx[c] = ntxx_x + ntxx_z;
    f.tzz[c] = ntzz_x + ntzz_z;
    const T ntxz_x = (d1 * f.txz_x[c] + dir * (dt / dx) * mu * dvzx) / d2;
    const T ntxz_z = (d3 * f.txz_z[c] + dir * (dt / dz) * mu * dvxz) / d4;
    f.txz_x[c] = ntxz_x; f.txz_z[c] = ntxz_z;
    f.txz[c] = ntxz_x + ntxz_z;
    if (recon) {
        st.gx_vx[c] = (float)(dvx / dx);
        st.gz_vz[c] = (float)(dvz / dz);
        st.gz_vx[c] = (float)(dvxz / dz);
        st.gx_vz[c] = (float)(dvzx / dx);
    }
}

// FD2DGPU.cu:591-682 GPUcalculate_elastic_vx_back + _vz_back: receiver-side
// velocity with the moduli inside the derivative, [nop, n-nop).
__global__ __launch_bounds__(256) void k_adj_v(EField f, Med m, const SlotDesc* __restrict__ slots, float dt,
                                               float dx, float dz, float dir) {
    const SlotDesc S = slots[blockIdx.z];
    Cell k;
    if (!cell_in(S, NOP, k)) return;
    const size_t c = k.c;
    const int s = S.nzp;
    const float* L = m.lam;
    const float* M = m.mu;
    const float d1 = 1.f - dt * m.dampx[c] / 2, d2 = 1.f + dt * m.dampx[c] / 2;
    const float d3 = 1.f - dt * m.dampz[c] / 2, d4 = 1.f + dt * m.dampz[c] / 2;
    const float cf[5] = {0.f, DS2D_C1, DS2D_C2, DS2D_C3, DS2D_C4};
    float ttxx = 0, ttxz = 0, ttzz = 0, utzz = 0, utxz = 0, utxx = 0;
#pragma unroll
    for (int i = 1; i <= NOP; ++i) {
        const size_t p1 = c + (size_t)i * s, p0 = c - (size_t)(i - 1) * s;
        ttxx += cf[i] * ((L[p1] + 2 * M[p1]) * f.txx[p1] - (L[p0] + 2 * M[p0]) * f.txx[p0]);
        ttxz += cf[i] * (M[c + i - 1] * f.txz[c + i - 1] - M[c - i] * f.txz[c - i]);
        ttzz += cf[i] * (L[p1] * f.tzz[p1] - L[p0] * f.tzz[p0]);
        utzz += cf[i] * ((L[c + i] + 2 * M[c + i]) * f.tzz[c + i] - (L[c - i + 1] + 2 * M[c - i + 1]) * f.tzz[c - i + 1]);
        const size_t q1 = c + (size_t)(i - 1) * s, q0 = c - (size_t)i * s;
        utxz += cf[i] * (M[q1] * f.txz[q1] - M[q0] * f.txz[q0]);
        utxx += cf[i] * (L[c + i] * f.txx[c + i] - L[c - i + 1] * f.txx[c - i + 1]);
    }
    const float br = 1.f / m.rho[c];
    const float nvx_x = (d1 * f.vx_x[c] + dir * br * (dt / dx) * ttxx + dir * br * (dt / dx) * ttzz) / d2;
    const float nvx_z = (d3 * f.vx_z[c] + dir * br * (dt / dz) * ttxz) / d4;
    const float nvz_x = (d1 * f.vz_x[c] + dir * br * (dt / dx) * utxz) / d2;
    const float nvz_z = (d3 * f.vz_z[c] + dir * br * (dt / dz) * utzz + dir * br * (dt / dz) * utxx) / d4;
    f.vx_x[c] = nvx_x; f.vx_z[c] = nvx_z; f.vx[c] = nvx_x + nvx_z;
    f.vz_x[c] = nvz_x; f.vz_z[c] = nvz_z; f.vz[c] = nvz_x + nvz_z;
}

// FD2DGPU.cu:725-761 GPUcalculate_elastic_txx_tzz (txx = txx_x, tzz = tzz_z,
// no moduli) + FD2DGPU.cu:860-895 GPUcalculate_elastic_txz, [nop, n-nop).
__global__ __launch_bounds__(256) void k_adj_s(EField f, Med m, const SlotDesc* __restrict__ slots, float dt,
                                               float dx, float dz, float dir) {
    const SlotDesc S = slots[blockIdx.z];
    Cell k;
    if (!cell_in(S, NOP, k)) return;
    const size_t c = k.c;
    const int nzp = S.nzp;
    const float d1 = 1.f - dt * m.dampx[c] / 2, d2 = 1.f + dt * m.dampx[c] / 2;
    const float d3 = 1.f - dt * m.dampz[c] / 2, d4 = 1.f + dt * m.dampz[c] / 2;
    const float dvx = dmx(f.vx, c, nzp);
    const float dvz = dmx(f.vz, c, 1);
    const float dvxz = dpx(f.vx, c, 1);
    const float dvzx = dpx(f.vz, c, nzp);
    const float a = (d1 * f.txx_x[c] + dir * (dt / dx) * dvx) / d2;
    const float b = (d3 * f.tzz_z[c] + dir * (dt / dz) * dvz) / d4;
    f.txx_x[c] = a; f.tzz_z[c] = b; f.txx[c] = a; f.tzz[c] = b;
    const float e = (d1 * f.txz_x[c] + dir * (dt / dx) * dvzx) / d2;
    const float g = (d3 * f.txz_z[c] + dir * (dt / dz) * dvxz) / d4;
    f.txz_x[c] = e; f.txz_z[c] = g; f.txz[c] = e + g;
}

// born.cu:351-405 add_born_source on [pml, n-pml): adds to the stress SUM
// fields of the target wavefield from the velocities of the source wavefield.
// The strain taps run in the source field's precision TS, the update in the
// target's TT.
template <typename TS, typename TT>
__global__ __launch_bounds__(256) void k_born_src(const TS* __restrict__ vx, const TS* __restrict__ vz, TT* btxx,
                                                  TT* btzz, TT* btxz, Med m, const SlotDesc* __restrict__ slots,
                                                  float dx_, float dz_, float dir_) {
    const SlotDesc S = slots[blockIdx.z];
    Cell k;
    if (!cell_in(S, S.pml, k)) return;
    const size_t c = k.c;
    const int nzp = S.nzp;
    const TT dx = dx_, dz = dz_, dir = dir_;
    const TT vxx = (TT)dmx(vx, c, nzp);
    const TT vzz = (TT)dmx(vz, c, 1);
    const TT vxz = (TT)dpx(vx, c, 1);
    const TT vzx = (TT)dpx(vz, c, nzp);
    const TT l = m.lam[c], mu = m.mu[c], l2m = l + 2 * mu;
    const TT mp = -(TT)m.dmp[c], ms = -(TT)m.dms[c];
    btxx[c] = btxx[c] + dir * 2 * l2m * mp * vxx / dx + 2 * (l2m * mp - 2 * mu * ms) * vzz / dz;
    btzz[c] = btzz[c] + dir * 2 * (l2m * mp - 2 * mu * ms) * vxx / dx + 2 * l2m * mp * vzz / dz;
    btxz[c] = btxz[c] + dir * 2 * mu * ms * (vzx / dx + vxz / dz);
}

// born.cu:408-527 add_born_source_adjoint on [pml, n-pml).  tmp_txz_x samples
// mu*dms at ix+i-1 for both taps (born.cu:514-515), as in the reference.
__global__ __launch_bounds__(256) void k_born_adj(const float* __restrict__ txx, const float* __restrict__ tzz,
                                                  const float* __restrict__ txz, float* bvx, float* bvz, Med m,
                                                  const SlotDesc* __restrict__ slots, float dx, float dz,
                                                  float dir) {
    const SlotDesc S = slots[blockIdx.z];
    Cell k;
    if (!cell_in(S, S.pml, k)) return;
    const size_t c = k.c;
    const int s = S.nzp;
    const float* L = m.lam;
    const float* M = m.mu;
    const float* P = m.dmp;
    const float* Q = m.dms;
    auto CP = [&](size_t q) { return 2 * (L[q] + 2 * M[q]) * (-P[q]); };
    auto CX = [&](size_t q) { return 2 * ((L[q] + 2 * M[q]) * (-P[q]) - 2 * M[q] * (-Q[q])); };
    auto CS = [&](size_t q) { return 2 * M[q] * (-Q[q]); };
    const float cf[5] = {0.f, DS2D_C1, DS2D_C2, DS2D_C3, DS2D_C4};
    float txx_x = 0, txx_z = 0, tzz_x = 0, tzz_z = 0, txz_x = 0, txz_z = 0;
#pragma unroll
    for (int i = 1; i <= NOP; ++i) {
        const size_t xp = c + (size_t)i * s, xm = c - (size_t)(i - 1) * s;
        const size_t zp = c + i, zm = c - i + 1;
        txx_x += cf[i] * (CP(xp) * txx[xp] - CP(xm) * txx[xm]);
        txx_z += cf[i] * (CX(zp) * txx[zp] - CX(zm) * txx[zm]);
        tzz_x += cf[i] * (CX(xp) * tzz[xp] - CX(xm) * tzz[xm]);
        tzz_z += cf[i] * (CP(zp) * tzz[zp] - CP(zm) * tzz[zm]);
        const size_t a1 = c + (size_t)(i - 1) * s, a0 = c - (size_t)i * s;
        txz_x += cf[i] * (CS(a1) * txz[a1] - CS(a1) * txz[a0]);
        txz_z += cf[i] * (CS(c + i - 1) * txz[c + i - 1] - CS(c - i) * txz[c - i]);
    }
    const float br = 1.f / m.rho[c];
    bvx[c] = bvx[c] + dir * br * (txx_x / dx + tzz_x / dx + txz_z / dz);
    bvz[c] = bvz[c] + dir * br * (tzz_z / dz + txx_z / dz + txz_x / dx);
}

// FD2DGPU.cu:2039-2049 GPUapplysource on txx and tzz (one thread per slot).
template <typename T>
__global__ void k_inject_src(T* txx, T* tzz, const float* __restrict__ wav, const SlotDesc* __restrict__ slots,
                             int nslots, int it) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const SlotDesc S = slots[s];
    const size_t c = (size_t)S.off + (size_t)S.isx * S.nzp + S.isz;
    const T w = wav[it];
    txx[c] += w;
    tzz[c] += w;
}

// FD2DGPU.cu:3093-3101 GPUrecord_scatterwavefield for one or two wavefields
// (FP32 records of wavefields of either precision).
template <typename T1, typename T2>
__global__ void k_record(const T1* __restrict__ vz, const T1* __restrict__ vx, float* rz, float* rx,
                         const T2* __restrict__ vz2, const T2* __restrict__ vx2, float* rz2, float* rx2,
                         const int* __restrict__ rix, const SlotDesc* __restrict__ slots, int nt, int it) {
    const SlotDesc S = slots[blockIdx.y];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.ntr) return;
    const size_t c = (size_t)S.off + (size_t)rix[S.rix_off + r] * S.nzp + S.igz;
    const size_t o = (size_t)S.rec_off + (size_t)r * nt + it;
    rz[o] = (float)vz[c];
    rx[o] = (float)vx[c];
    if (vz2) {
        rz2[o] = (float)vz2[c];
        rx2[o] = (float)vx2[c];
    }
}

// FD2DGPU.cu:3306-3327 GPUapplysource_backward_from_record with stack = true.
__global__ void k_inject_res(float* vz, float* vx, const float* __restrict__ rz, const float* __restrict__ rx,
                             const int* __restrict__ rix, const SlotDesc* __restrict__ slots, int nt, int it) {
    const SlotDesc S = slots[blockIdx.y];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.ntr) return;
    const size_t c = (size_t)S.off + (size_t)rix[S.rix_off + r] * S.nzp + S.igz;
    const size_t o = (size_t)S.rec_off + (size_t)r * nt + it;
    vz[c] += rz[o];
    vx[c] += rx[o];
}

// FD2DGPU.cu:3674-3731 record_edge_wavefield / load_edge_wavefield (10-arg)
// for two fields; bands are stored per slot as [nt][up|down|left|right].
template <typename T>
__global__ __launch_bounds__(256) void k_edge_io(T* fa, T* fb, T* ea, T* eb, const SlotDesc* __restrict__ slots,
                                                 int it, int save) {
    const SlotDesc S = slots[blockIdx.z];
    const int iz = blockIdx.x * BZ + threadIdx.x;
    const int ix = blockIdx.y * BX + threadIdx.y;
    // random-boundary slots own no band storage
    if (S.boundary != 0 || ix >= S.nxp || iz >= S.nzp) return;
    const int N = NOP, pml = S.pml, nx = S.nx, nz = S.nz, nxp = S.nxp, nzp = S.nzp;
    const size_t per_step = (size_t)2 * (2 * N) * nx + (size_t)2 * (2 * N) * nz;
    const size_t base = (size_t)S.edge_off + (size_t)it * per_step;
    const size_t c = (size_t)S.off + (size_t)ix * nzp + iz;
    auto io = [&](size_t k) {
        if (save) { ea[base + k] = fa[c]; eb[base + k] = fb[c]; }
        else { fa[c] = ea[base + k]; fb[c] = eb[base + k]; }
    };
    const size_t up = 0, down = (size_t)nx * 2 * N, left = 2 * down, right = left + (size_t)2 * N * nz;
    if (ix >= pml && ix < nxp - pml && iz >= pml - N && iz < pml + N)
        io(up + (size_t)(ix - pml) * 2 * N + (iz - (pml - N)));
    if (ix >= pml && ix < nxp - pml && iz >= nzp - pml - N && iz < nzp - pml + N)
        io(down + (size_t)(ix - pml) * 2 * N + (iz - (nzp - pml - N)));
    if (ix >= pml - N && ix < pml + N && iz >= pml && iz < nzp - pml)
        io(left + (size_t)(ix - (pml - N)) * nz + (iz - pml));
    if (ix >= nxp - pml - N && ix < nxp - pml + N && iz >= pml && iz < nzp - pml)
        io(right + (size_t)(ix - (nxp - pml - N)) * nz + (iz - pml));
}

// FD2DGPU.cu:3333-3349 image_gradient: three doimage_gradient terms
// (FD2DGPU.cu:3245-3263) and doimage_illum (3235-3243), accumulated in the
// slot's window images w7 = [pp, ps, illum, b_pp, b_ps, b2_pp, b2_ps].
template <typename TV>
__global__ __launch_bounds__(256) void k_image(Strain sg, Strain sb, const float* __restrict__ bk_txx,
                                               const float* __restrict__ bk_tzz, const float* __restrict__ bk_txz,
                                               const float* __restrict__ bb_txx, const float* __restrict__ bb_tzz,
                                               const float* __restrict__ bb_txz, const TV* __restrict__ vx,
                                               const TV* __restrict__ vz, Med m, float* w7, long long wstride,
                                               const SlotDesc* __restrict__ slots) {
    const SlotDesc S = slots[blockIdx.z];
    const int iz = blockIdx.x * BZ + threadIdx.x;
    const int ix = blockIdx.y * BX + threadIdx.y;
    if (ix >= S.nx || iz >= S.nz) return;
    const size_t c = (size_t)S.off + (size_t)(ix + S.pml) * S.nzp + iz + S.pml;
    const size_t o = (size_t)S.win_off + (size_t)ix * S.nz + iz;
    const float gxx = sg.gx_vx[c], gzz = sg.gz_vz[c], gxz = sg.gz_vx[c] + sg.gx_vz[c];
    const float hxx = sb.gx_vx[c], hzz = sb.gz_vz[c], hxz = sb.gz_vx[c] + sb.gx_vz[c];
    const float kp = -2 * m.vp[c] * m.vp[c] * m.rho[c];
    const float ks = -2 * m.vs[c] * m.vs[c] * m.rho[c];
    const float axx = bk_txx[c], azz = bk_tzz[c], axz = bk_txz[c];
    const float bxx = bb_txx[c], bzz = bb_tzz[c], bxz = bb_txz[c];
    // term 1: G0 strain x receiver G0 stress
    w7[o] += kp * ((gxx + gzz) * (axx + azz));
    w7[o + wstride] += ks * (-2 * (gxx * azz + gzz * axx) + gxz * axz);
    const float ux = (float)vx[c], uz = (float)vz[c];
    w7[o + 2 * wstride] += ux * ux + uz * uz;
    // term 2: source-side scatter strain x receiver G0 stress
    w7[o + 3 * wstride] += kp * ((hxx + hzz) * (axx + azz));
    w7[o + 4 * wstride] += ks * (-2 * (hxx * azz + hzz * axx) + hxz * axz);
    // term 3: G0 strain x receiver-side scatter stress
    w7[o + 5 * wstride] += kp * ((gxx + gzz) * (bxx + bzz));
    w7[o + 6 * wstride] += ks * (-2 * (gxx * bzz + gzz * bxx) + gxz * bxz);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static dim3 cell_grid(const BatchShape& b) { return dim3(ceil_div(b.nzp, BZ), ceil_div(b.max_nxp, BX), b.nslots); }

template <typename T>
static void ref_v(hipStream_t st, const BatchShape& b, EFieldT<T> f, Med m, float dir, int recon) {
    hipLaunchKernelGGL(k_ref_v<T>, cell_grid(b), dim3(BZ, BX), 0, st, f, m, b.slots, b.dt, b.dx, b.dz, dir, recon);
}
template <typename T>
static void ref_s(hipStream_t st, const BatchShape& b, EFieldT<T> f, Med m, Strain s, float dir, int recon) {
    hipLaunchKernelGGL(k_ref_s<T>, cell_grid(b), dim3(BZ, BX), 0, st, f, m, s, b.slots, b.dt, b.dx, b.dz, dir,
                       recon);
}
void launch_ref_v(hipStream_t st, const BatchShape& b, EField f, Med m, float dir, int recon) {
    ref_v<float>(st, b, f, m, dir, recon);
}
void launch_ref_v(hipStream_t st, const BatchShape& b, EFieldD f, Med m, float dir, int recon) {
    ref_v<double>(st, b, f, m, dir, recon);
}
void launch_ref_s(hipStream_t st, const BatchShape& b, EField f, Med m, Strain s, float dir, int recon) {
    ref_s<float>(st, b, f, m, s, dir, recon);
}
void launch_ref_s(hipStream_t st, const BatchShape& b, EFieldD f, Med m, Strain s, float dir, int recon) {
    ref_s<double>(st, b, f, m, s, dir, recon);
}
void launch_adj_v(hipStream_t st, const BatchShape& b, EField f, Med m, float dir) {
    hipLaunchKernelGGL(k_adj_v, cell_grid(b), dim3(BZ, BX), 0, st, f, m, b.slots, b.dt, b.dx, b.dz, dir);
}
void launch_adj_s(hipStream_t st, const BatchShape& b, EField f, Med m, float dir) {
    hipLaunchKernelGGL(k_adj_s, cell_grid(b), dim3(BZ, BX), 0, st, f, m, b.slots, b.dt, b.dx, b.dz, dir);
}
template <typename TS, typename TT>
static void born_src(hipStream_t st, const BatchShape& b, const TS* vx, const TS* vz, EFieldT<TT> tgt, Med m,
                     float dir) {
    hipLaunchKernelGGL((k_born_src<TS, TT>), cell_grid(b), dim3(BZ, BX), 0, st, vx, vz, tgt.txx, tgt.tzz, tgt.txz, m,
                       b.slots, b.dx, b.dz, dir);
}
void launch_born_src(hipStream_t st, const BatchShape& b, const float* vx, const float* vz, EField tgt, Med m,
                     float dir) {
    born_src<float, float>(st, b, vx, vz, tgt, m, dir);
}
void launch_born_src(hipStream_t st, const BatchShape& b, const double* vx, const double* vz, EFieldD tgt, Med m,
                     float dir) {
    born_src<double, double>(st, b, vx, vz, tgt, m, dir);
}
void launch_born_src(hipStream_t st, const BatchShape& b, const double* vx, const double* vz, EField tgt, Med m,
                     float dir) {
    born_src<double, float>(st, b, vx, vz, tgt, m, dir);
}
void launch_born_adj(hipStream_t st, const BatchShape& b, EField src, EField tgt, Med m, float dir) {
    hipLaunchKernelGGL(k_born_adj, cell_grid(b), dim3(BZ, BX), 0, st, src.txx, src.tzz, src.txz, tgt.vx, tgt.vz, m,
                       b.slots, b.dx, b.dz, dir);
}
void launch_inject_src(hipStream_t st, const BatchShape& b, float* txx, float* tzz, const float* wav, int it) {
    hipLaunchKernelGGL(k_inject_src<float>, dim3(ceil_div(b.nslots, 64)), dim3(64), 0, st, txx, tzz, wav, b.slots,
                       b.nslots, it);
}
void launch_inject_src(hipStream_t st, const BatchShape& b, double* txx, double* tzz, const float* wav, int it) {
    hipLaunchKernelGGL(k_inject_src<double>, dim3(ceil_div(b.nslots, 64)), dim3(64), 0, st, txx, tzz, wav, b.slots,
                       b.nslots, it);
}
template <typename T1, typename T2>
static void record(hipStream_t st, const BatchShape& b, const T1* vz, const T1* vx, float* rz, float* rx,
                   const T2* vz2, const T2* vx2, float* rz2, float* rx2, int it) {
    hipLaunchKernelGGL((k_record<T1, T2>), dim3(ceil_div(b.max_ntr, 64), b.nslots), dim3(64), 0, st, vz, vx, rz, rx,
                       vz2, vx2, rz2, rx2, b.rix, b.slots, b.nt, it);
}
void launch_record(hipStream_t st, const BatchShape& b, const float* vz, const float* vx, float* rz, float* rx,
                   const float* vz2, const float* vx2, float* rz2, float* rx2, int it) {
    record<float, float>(st, b, vz, vx, rz, rx, vz2, vx2, rz2, rx2, it);
}
void launch_record(hipStream_t st, const BatchShape& b, const double* vz, const double* vx, float* rz, float* rx,
                   const float* vz2, const float* vx2, float* rz2, float* rx2, int it) {
    record<double, float>(st, b, vz, vx, rz, rx, vz2, vx2, rz2, rx2, it);
}
void launch_inject_res(hipStream_t st, const BatchShape& b, float* vz, float* vx, const float* rz, const float* rx,
                       int it) {
    hipLaunchKernelGGL(k_inject_res, dim3(ceil_div(b.max_ntr, 64), b.nslots), dim3(64), 0, st, vz, vx, rz, rx, b.rix,
                       b.slots, b.nt, it);
}
void launch_edge_io(hipStream_t st, const BatchShape& b, float* fa, float* fb, float* ea, float* eb, int it,
                    int save) {
    hipLaunchKernelGGL(k_edge_io<float>, cell_grid(b), dim3(BZ, BX), 0, st, fa, fb, ea, eb, b.slots, it, save);
}
void launch_edge_io(hipStream_t st, const BatchShape& b, double* fa, double* fb, double* ea, double* eb, int it,
                    int save) {
    hipLaunchKernelGGL(k_edge_io<double>, cell_grid(b), dim3(BZ, BX), 0, st, fa, fb, ea, eb, b.slots, it, save);
}
template <typename TV>
static void image(hipStream_t st, const BatchShape& b, Strain sg, Strain sb, EField bk, EField bb, const TV* vx,
                  const TV* vz, Med m, float* w7, long long wstride) {
    dim3 grid(ceil_div(b.nz, BZ), ceil_div(b.max_nxp, BX), b.nslots);
    hipLaunchKernelGGL(k_image<TV>, grid, dim3(BZ, BX), 0, st, sg, sb, bk.txx, bk.tzz, bk.txz, bb.txx, bb.tzz, bb.txz,
                       vx, vz, m, w7, wstride, b.slots);
}
void launch_image(hipStream_t st, const BatchShape& b, Strain sg, Strain sb, EField bk, EField bb, EField bg, Med m,
                  float* w7, long long wstride) {
    image<float>(st, b, sg, sb, bk, bb, bg.vx, bg.vz, m, w7, wstride);
}
void launch_image(hipStream_t st, const BatchShape& b, Strain sg, Strain sb, EField bk, EField bb, EFieldD bg, Med m,
                  float* w7, long long wstride) {
    image<double>(st, b, sg, sb, bk, bb, bg.vx, bg.vz, m, w7, wstride);
}

}  // namespace ds2d
