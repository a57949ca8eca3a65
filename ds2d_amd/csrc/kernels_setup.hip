// Per-batch setup, image gather and reduction kernels (gfx950).
#include "common.hpp"
#include "kernels.hpp"

namespace ds2d {

// SplitMix64 finaliser (Steele, Lea, Flood 2014): the counter hash of the
// seeded random boundary.  Same constants and key mixing as the oracle.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ double rb_uniform(unsigned long long seed, int shot, int ix, int iz) {
    unsigned long long key = seed * 0x100000001B3ull ^ ((unsigned long long)(unsigned)shot << 42) ^
                             ((unsigned long long)(unsigned)ix << 21) ^ (unsigned long long)(unsigned)iz;
    unsigned long long h = splitmix64(key);
    double u = (double)(h >> 40) * (1.0 / 16777216.0);
    return 2.0 * u - 1.0;
}

// One pass over every padded cell of every slot:
//  * window + edge replication (FD2DGPU.cu:122-138 copyvelocityHtoD followed by
//    copypmlVz/copypmlVx 234-250): value = global[col0 + clamp(ix-pml)][clamp(iz-pml)];
//  * random boundary frame (replaces FD2DGPU.cu:154-230 / 3156-3216) or PML
//    profiles (FD2DGPU.cu:444-494, strict '>' on the far side);
//  * moduli (FD2DGPU.cu:1117-1125) with rho constant or from the model;
//  * image window (buffer_image_extrapolation, FD2DGPU.cu:1324-1333).
__global__ __launch_bounds__(256) void k_setup(SetupArgs a, Med m, float* bu, float* kk, float* refl,
                                               const SlotDesc* __restrict__ slots) {
    const SlotDesc S = slots[blockIdx.z];
    const int iz = blockIdx.x * BZ + threadIdx.x;
    const int ix = blockIdx.y * BX + threadIdx.y;
    if (ix >= S.nxp || iz >= S.nzp) return;
    const int pml = S.pml;
    const int lx = min(max(ix - pml, 0), S.nx - 1);
    const int lz = min(max(iz - pml, 0), S.nz - 1);
    const size_t gidx = (size_t)(S.col0 + lx) * a.allnz + lz;
    const size_t c = (size_t)S.off + (size_t)ix * S.nzp + iz;
    float vp = a.vp_all[gidx];
    float vs = a.vs_all ? a.vs_all[gidx] : 0.f;
    const float rin = a.rho_all ? a.rho_all[gidx] : 0.f;
    const bool frame = ix < pml || ix >= S.nxp - pml || iz < pml || iz >= S.nzp - pml;
    float dxv = 0.f, dzv = 0.f;
    if (S.boundary == 1) {
        if (frame) {
            const double u = rb_uniform(a.rb_seed, S.shot_id, ix, iz);
            const float fac = (float)(1.0 + (double)a.rb_amp * u);
            const float ratio = vp > 0.f ? vs / vp : 0.f;
            vp = vp * fac;
            vs = ratio * vp;
        }
    } else {
        const double base = log10(1.0 / (double)a.pml_rr) * (5.0 * (double)vp / (2.0 * pml));
        auto prof = [&](int d) {
            const double t = (double)d / pml;
            return (float)(base * (t * t) * (t * t));
        };
        if (ix < pml) dxv = prof(pml - ix);
        if (ix > S.nxp - pml) dxv = prof(ix - (S.nxp - pml));
        if (iz < pml) dzv = prof(pml - iz);
        if (iz > S.nzp - pml) dzv = prof(iz - (S.nzp - pml));
    }
    const float rho = a.rho_const > 0.f ? a.rho_const : rin;
    m.vp[c] = vp;
    m.vs[c] = vs;
    m.rho[c] = rho;
    m.lam[c] = rho * (vp * vp - 2 * vs * vs);
    m.mu[c] = rho * vs * vs;
    m.dampx[c] = dxv;
    m.dampz[c] = dzv;
    m.dmp[c] = a.img_pp ? a.img_pp[gidx] : 0.f;
    m.dms[c] = a.img_ps ? a.img_ps[gidx] : 0.f;
    if (bu) {
        bu[c] = 1.f / rho;
        kk[c] = rho * vp * vp;
        refl[c] = (!frame && a.img_pp) ? a.img_pp[gidx] : 0.f;
    }
}

struct SetupExtra {
    float *bu, *kk, *refl;
};

void launch_setup(hipStream_t st, const BatchShape& b, const SetupArgs& a, Med m) {
    dim3 grid(ceil_div(b.nzp, BZ), ceil_div(b.max_nxp, BX), b.nslots);
    hipLaunchKernelGGL(k_setup, grid, dim3(BZ, BX), 0, st, a, m, (float*)nullptr, (float*)nullptr, (float*)nullptr,
                       b.slots);
}

void launch_setup_acoustic(hipStream_t st, const BatchShape& b, const SetupArgs& a, Med m, float* bu, float* kk,
                           float* refl) {
    dim3 grid(ceil_div(b.nzp, BZ), ceil_div(b.max_nxp, BX), b.nslots);
    hipLaunchKernelGGL(k_setup, grid, dim3(BZ, BX), 0, st, a, m, bu, kk, refl, b.slots);
}

// glob[k][gx][z] += sum_{slots s covering gx, in slot order} win[k][s][gx-col0_s][z];
// slots flagged in skip (may be null) are left out.
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ win, long long wstride, float* glob,
                                                long long gstride, int nimg, int allnx, int allnz,
                                                const SlotDesc* __restrict__ slots, int nslots,
                                                const int* __restrict__ skip) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)allnx * allnz) return;
    const int gx = (int)(t / allnz), z = (int)(t % allnz);
    for (int k = 0; k < nimg; ++k) {
        float acc = 0.f;
        bool any = false;
        for (int s = 0; s < nslots; ++s) {
            const int c0 = slots[s].col0, nx = slots[s].nx;
            if (gx < c0 || gx >= c0 + nx || (skip && skip[s])) continue;
            const float v = win[k * wstride + slots[s].win_off + (long long)(gx - c0) * allnz + z];
            acc = any ? acc + v : v;
            any = true;
        }
        if (any) glob[k * gstride + t] += acc;
    }
}

void launch_gather(hipStream_t st, const BatchShape& b, const float* win, long long wstride, float* glob,
                   long long gstride, int nimg, int allnx, int allnz, const int* skip) {
    const long long n = (long long)allnx * allnz;
    hipLaunchKernelGGL(k_gather, dim3(ceil_div(n, 256)), dim3(256), 0, st, win, wstride, glob, gstride, nimg, allnx,
                       allnz, b.slots, b.nslots, skip);
}

// flags[s] = 1 when any of the narr arrays (stride astride) holds a NaN or an
// infinity in slot s's records (windows = 0: ntr*nt samples from rec_off) or
// image window (windows = 1: nx*nz cells from win_off); flags are only set.
__global__ __launch_bounds__(256) void k_slot_nonfinite(const float* __restrict__ a, long long astride, int narr,
                                                        const SlotDesc* __restrict__ slots, int windows, int nt,
                                                        int* flags) {
    const SlotDesc S = slots[blockIdx.y];
    const long long len = windows ? (long long)S.nx * S.nz : (long long)S.ntr * nt;
    const long long base = windows ? S.win_off : S.rec_off;
    bool bad = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < len;
         i += (long long)gridDim.x * blockDim.x)
        for (int k = 0; k < narr; ++k) bad |= !(fabsf(a[k * astride + base + i]) <= 3.402823466e38f);
    if (__any(bad) && (threadIdx.x & 63) == 0) flags[blockIdx.y] = 1;
}

void launch_slot_nonfinite(hipStream_t st, const BatchShape& b, const float* a, long long astride, int narr,
                           int windows, int* flags) {
    hipLaunchKernelGGL(k_slot_nonfinite, dim3(64, b.nslots), dim3(256), 0, st, a, astride, narr, b.slots, windows,
                       b.nt, flags);
}

// One workgroup per (trace, slot): FP64 sums over time of the NCC terms
// (FD2DGPU.cu:4590-4604).
__global__ __launch_bounds__(256) void k_ncc_sums(const float* __restrict__ sz, const float* __restrict__ sx,
                                                  const float* __restrict__ oz, const float* __restrict__ ox,
                                                  double* sums, const int* __restrict__ tbase,
                                                  const SlotDesc* __restrict__ slots, int nt) {
    const SlotDesc S = slots[blockIdx.y];
    const int r = blockIdx.x;
    if (r >= S.ntr) return;
    const size_t o = (size_t)S.rec_off + (size_t)r * nt;
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int j = threadIdx.x; j < nt; j += blockDim.x) {
        const double dx = sx[o + j], Dx = ox[o + j], dz = sz[o + j], Dz = oz[o + j];
        a[0] += dx * dx; a[1] += Dx * Dx; a[2] += dx * Dx;
        a[3] += dz * dz; a[4] += Dz * Dz; a[5] += dz * Dz;
    }
    __shared__ double red[6][256];
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* out = sums + (size_t)(tbase[blockIdx.y] + r) * 6;
        for (int k = 0; k < 6; ++k) out[k] = red[k][0];
    }
}

void launch_ncc_sums(hipStream_t st, const BatchShape& b, const float* sz, const float* sx, const float* oz,
                     const float* ox, double* sums, const int* trace_base) {
    hipLaunchKernelGGL(k_ncc_sums, dim3(b.max_ntr, b.nslots), dim3(256), 0, st, sz, sx, oz, ox, sums, trace_base,
                       b.slots, b.nt);
}

// FD2DGPU.cu:4566-4664 residual: iter 0: r = -D/|D|^2; else
// r = ((dD/|d|^2) d - D)/(|d||D|).  A zero-norm trace gets a zero residual.
__global__ __launch_bounds__(256) void k_ncc_apply(float* sz, float* sx, const float* __restrict__ oz,
                                                   const float* __restrict__ ox, const double* __restrict__ sums,
                                                   const int* __restrict__ tbase, const SlotDesc* __restrict__ slots,
                                                   int nt, int iter) {
    const SlotDesc S = slots[blockIdx.z];
    const int r = blockIdx.y;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.ntr || j >= nt) return;
    const double* q = sums + (size_t)(tbase[blockIdx.z] + r) * 6;
    const size_t o = (size_t)S.rec_off + (size_t)r * nt + j;
    if (iter == 0) {
        const double ax = q[1] > 0 ? 1.0 / (sqrt(q[1]) * sqrt(q[1])) : 0.0;
        const double az = q[4] > 0 ? 1.0 / (sqrt(q[4]) * sqrt(q[4])) : 0.0;
        sx[o] = (float)(ax * (0.0 - (double)ox[o]));
        sz[o] = (float)(az * (0.0 - (double)oz[o]));
    } else {
        const bool okx = q[0] > 0 && q[1] > 0, okz = q[3] > 0 && q[4] > 0;
        const double ax = okx ? 1.0 / (sqrt(q[0]) * sqrt(q[1])) : 0.0;
        const double az = okz ? 1.0 / (sqrt(q[3]) * sqrt(q[4])) : 0.0;
        const double gx = okx ? q[2] / q[0] : 0.0, gz = okz ? q[5] / q[3] : 0.0;
        sx[o] = (float)(ax * (gx * (double)sx[o] - (double)ox[o]));
        sz[o] = (float)(az * (gz * (double)sz[o] - (double)oz[o]));
    }
}

void launch_ncc_apply(hipStream_t st, const BatchShape& b, float* sz, float* sx, const float* oz, const float* ox,
                      const double* sums, const int* trace_base, int iter) {
    hipLaunchKernelGGL(k_ncc_apply, dim3(ceil_div(b.nt, 256), b.max_ntr, b.nslots), dim3(256), 0, st, sz, sx, oz, ox,
                       sums, trace_base, b.slots, b.nt, iter);
}

__global__ void k_fill(float* p, long long n, float v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill(hipStream_t st, float* p, long long n, float v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(ceil_div(n, 256)), dim3(256), 0, st, p, n, v);
}

__global__ void k_axpby(float* y, const float* __restrict__ x, long long n, float a, float b) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}
void launch_axpby(hipStream_t st, float* y, const float* x, long long n, float a, float b) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_axpby, dim3(ceil_div(n, 256)), dim3(256), 0, st, y, x, n, a, b);
}

// Block-partial FP64 dot products (grid-stride), reduced on the host in block order.
__global__ __launch_bounds__(256) void k_dot(const float* __restrict__ x, const float* __restrict__ y, long long n,
                                             double* partial) {
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        acc += (double)x[i] * (double)y[i];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
void launch_dot(hipStream_t st, const float* x, const float* y, long long n, double* partial, int nblocks) {
    hipLaunchKernelGGL(k_dot, dim3(nblocks), dim3(256), 0, st, x, y, n, partial);
}

}  // namespace ds2d
