// Fast elastic forward kernels (gfx950), batched over uneven shot windows.
//
// Same scheme as born.cu:215-349 (update_vx/vz, update_txx_tzz/txz) with the
// source of FD2DGPU.cu:2039-2049 and the recorder of FD2DGPU.cu:355-365 fused
// in, but the split PML memory (vx_x, vx_z, ...) is kept only where a damping
// profile can be non-zero (ix < pml, ix > nxp-pml, iz < pml, iz > nzp-pml).
// Elsewhere dampx = dampz = 0 and the split update reduces exactly (in exact
// arithmetic) to the unsplit one, so interior cells carry only the five
// physical fields: v reads 6 floats and writes 2, s reads 7 and writes 3.
#include "common.hpp"
#include "kernels.hpp"

namespace ds2d {

namespace {
__device__ __forceinline__ float dpx(const float* f, size_t c, int s) {
    return DS2D_C1 * (f[c + s] - f[c]) + DS2D_C2 * (f[c + 2 * s] - f[c - s]) +
           DS2D_C3 * (f[c + 3 * s] - f[c - 2 * s]) + DS2D_C4 * (f[c + 4 * s] - f[c - 3 * s]);
}
__device__ __forceinline__ float dmx(const float* f, size_t c, int s) {
    return DS2D_C1 * (f[c] - f[c - s]) + DS2D_C2 * (f[c + s] - f[c - 2 * s]) +
           DS2D_C3 * (f[c + 2 * s] - f[c - 3 * s]) + DS2D_C4 * (f[c + 3 * s] - f[c - 4 * s]);
}
__device__ __forceinline__ bool damp_cell(const SlotDesc& S, int ix, int iz) {
    return S.boundary == 0 &&
           (ix < S.pml || ix > S.nxp - S.pml || iz < S.pml || iz > S.nzp - S.pml);
}
}  // namespace

// Velocity half-step.  bx = dt/rho.  Source w added to txx/tzz at the source
// cell for this step only (it is not part of the stress state, exactly as the
// reference's split update rebuilds txx/tzz from their splits).
__global__ __launch_bounds__(256) void k_fast_v(FastArgs a, const int* __restrict__ rmap,
                                                const SlotDesc* __restrict__ slots, float idx, float idz,
                                                float dt, int nt, int it) {
    const SlotDesc S = slots[blockIdx.z];
    const int iz = blockIdx.x * BZ + threadIdx.x, ix = blockIdx.y * BX + threadIdx.y;
    if (ix < NOP || ix >= S.nxp - NOP || iz < NOP || iz >= S.nzp - NOP) return;
    const int s = S.nzp;
    const size_t c = (size_t)S.off + (size_t)ix * s + iz;
    float ax = dpx(a.txx, c, s);
    float bz = dmx(a.txz, c, 1);
    float cz = dpx(a.tzz, c, 1);
    float ex = dmx(a.txz, c, s);
    // transient explosive source (only cells whose stencil reaches it)
    const int dxs = S.isx - ix, dzs = S.isz - iz;
    if (iz == S.isz && dxs >= -3 && dxs <= 4) {
        const float w = a.wav[it];
        const float cf[9] = {-DS2D_C4, -DS2D_C3, -DS2D_C2, -DS2D_C1, DS2D_C1, DS2D_C2, DS2D_C3, DS2D_C4, 0.f};
        // dpx taps: +c_i at ix+i (i=1..4), -c_i at ix-i+1 (offsets 0..-3)
        ax += w * cf[dxs + 3];
    }
    if (ix == S.isx && dzs >= -3 && dzs <= 4) {
        const float w = a.wav[it];
        const float cf[9] = {-DS2D_C4, -DS2D_C3, -DS2D_C2, -DS2D_C1, DS2D_C1, DS2D_C2, DS2D_C3, DS2D_C4, 0.f};
        cz += w * cf[dzs + 3];
    }
    const float bx = a.bx[c];
    float nvx, nvz;
    if (damp_cell(S, ix, iz)) {
        const float dpx_ = a.dampx[c], dpz_ = a.dampz[c];
        // one hardware reciprocal (1 ulp) per direction instead of an IEEE division per component
        const float d1 = 1.f - dt * dpx_ / 2, r2 = __builtin_amdgcn_rcpf(1.f + dt * dpx_ / 2);
        const float d3 = 1.f - dt * dpz_ / 2, r4 = __builtin_amdgcn_rcpf(1.f + dt * dpz_ / 2);
        const float vxx = (d1 * a.vx_x[c] + bx * idx * ax) * r2;
        const float vxz = (d3 * a.vx_z[c] + bx * idz * bz) * r4;
        const float vzx = (d1 * a.vz_x[c] + bx * idx * ex) * r2;
        const float vzz = (d3 * a.vz_z[c] + bx * idz * cz) * r4;
        a.vx_x[c] = vxx; a.vx_z[c] = vxz; a.vz_x[c] = vzx; a.vz_z[c] = vzz;
        nvx = vxx + vxz;
        nvz = vzx + vzz;
    } else {
        nvx = a.vx[c] + bx * (ax * idx + bz * idz);
        nvz = a.vz[c] + bx * (ex * idx + cz * idz);
    }
    a.vx[c] = nvx;
    a.vz[c] = nvz;
    if (iz == S.igz) {
        const int r = rmap[S.rmap_off + ix];
        if (r >= 0) {
            const size_t o = (size_t)S.rec_off + (size_t)r * nt + it;
            a.rec_z[o] = nvz;
            a.rec_x[o] = nvx;
        }
    }
}

// Stress half-step.  lam/mu are pre-multiplied by dt.
__global__ __launch_bounds__(256) void k_fast_s(FastArgs a, const SlotDesc* __restrict__ slots, float idx,
                                                float idz, float dt) {
    const SlotDesc S = slots[blockIdx.z];
    const int iz = blockIdx.x * BZ + threadIdx.x, ix = blockIdx.y * BX + threadIdx.y;
    if (ix < NOP || ix >= S.nxp - NOP || iz < NOP || iz >= S.nzp - NOP) return;
    const int s = S.nzp;
    const size_t c = (size_t)S.off + (size_t)ix * s + iz;
    const float dvx = dmx(a.vx, c, s) * idx;
    const float dvz = dmx(a.vz, c, 1) * idz;
    const float dvxz = dpx(a.vx, c, 1) * idz;
    const float dvzx = dpx(a.vz, c, s) * idx;
    const float l = a.lam[c], mu = a.mu[c], l2m = l + 2 * mu;
    if (damp_cell(S, ix, iz)) {
        const float dpx_ = a.dampx[c], dpz_ = a.dampz[c];
        const float d1 = 1.f - dt * dpx_ / 2, r2 = __builtin_amdgcn_rcpf(1.f + dt * dpx_ / 2);
        const float d3 = 1.f - dt * dpz_ / 2, r4 = __builtin_amdgcn_rcpf(1.f + dt * dpz_ / 2);
        const float a1 = (d1 * a.txx_x[c] + l2m * dvx) * r2;
        const float a2 = (d3 * a.txx_z[c] + l * dvz) * r4;
        const float b1 = (d1 * a.tzz_x[c] + l * dvx) * r2;
        const float b2 = (d3 * a.tzz_z[c] + l2m * dvz) * r4;
        const float c1 = (d1 * a.txz_x[c] + mu * dvzx) * r2;
        const float c2 = (d3 * a.txz_z[c] + mu * dvxz) * r4;
        a.txx_x[c] = a1; a.txx_z[c] = a2; a.tzz_x[c] = b1; a.tzz_z[c] = b2; a.txz_x[c] = c1; a.txz_z[c] = c2;
        a.txx[c] = a1 + a2;
        a.tzz[c] = b1 + b2;
        a.txz[c] = c1 + c2;
    } else {
        a.txx[c] = a.txx[c] + (l2m * dvx + l * dvz);
        a.tzz[c] = a.tzz[c] + (l * dvx + l2m * dvz);
        a.txz[c] = a.txz[c] + mu * (dvzx + dvxz);
    }
}

// bx = dt/rho, lam*dt, mu*dt from the batch medium.
__global__ void k_fast_params(const float* __restrict__ rho, const float* __restrict__ lam,
                              const float* __restrict__ mu, float* bx, float* ldt, float* mdt, long long n,
                              float dt) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bx[i] = dt / rho[i];
    ldt[i] = lam[i] * dt;
    mdt[i] = mu[i] * dt;
}

// Receiver map: rmap[pad0 + ix] = trace index recording column ix, else -1.
__global__ void k_fast_rmap(int* rmap, const int* __restrict__ rix, const SlotDesc* __restrict__ slots) {
    const SlotDesc S = slots[blockIdx.y];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.ntr) return;
    rmap[S.rmap_off + rix[S.rix_off + r]] = r;
}

static dim3 fgrid(const BatchShape& b) { return dim3(ceil_div(b.nzp, BZ), ceil_div(b.max_nxp, BX), b.nslots); }

void launch_fast_v(hipStream_t st, const BatchShape& b, const FastArgs& a, const int* rmap, int it) {
    hipLaunchKernelGGL(k_fast_v, fgrid(b), dim3(BZ, BX), 0, st, a, rmap, b.slots, 1.f / b.dx, 1.f / b.dz, b.dt, b.nt,
                       it);
}
void launch_fast_s(hipStream_t st, const BatchShape& b, const FastArgs& a) {
    hipLaunchKernelGGL(k_fast_s, fgrid(b), dim3(BZ, BX), 0, st, a, b.slots, 1.f / b.dx, 1.f / b.dz, b.dt);
}
void launch_fast_params(hipStream_t st, const float* rho, const float* lam, const float* mu, float* bx, float* ldt,
                        float* mdt, long long n, float dt) {
    hipLaunchKernelGGL(k_fast_params, dim3(ceil_div(n, 256)), dim3(256), 0, st, rho, lam, mu, bx, ldt, mdt, n, dt);
}
void launch_fast_rmap(hipStream_t st, const BatchShape& b, int* rmap) {
    hipLaunchKernelGGL(k_fast_rmap, dim3(ceil_div(b.max_ntr, 64), b.nslots), dim3(64), 0, st, rmap, b.rix, b.slots);
}

}  // namespace ds2d
