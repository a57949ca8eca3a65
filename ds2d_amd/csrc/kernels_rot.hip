// P/S-decoupled (theta/omega rotation) elastic kernels and the decomposed
// imaging kernels for gfx950, batched over shot slots (blockIdx.z is the
// slot; blocks of 64 lanes along contiguous z x 4 rows along x).
//
//   k_rot_v / k_rot_p   GPUcalculate_rotation_v / _p (FD2DGPU.cu:3005-3080,
//                       undamped forms 2545-2609), on [nop, n-nop): the
//                       reference's inclusive last row and column read one
//                       cell past the window.
//   k_rot_image         doimage of the rotation RTM (FD2DGPU.cu:3218-3233):
//                       PP and PS normalised by the source P energy.
//   k_helmholtz         GPUcalculate_helmholtz (FD2DGPU.cu:1963-1986), on the
//                       reference's inclusive [nop, n-nop].
//   k_decomp            per physical cell: GPUcalculate_vector_VPVS
//                       (FD2DGPU.cu:1988-2017), the Poynting up/down split
//                       (poynting_both_up_down, FD2DGPU.cu:2223-2236) and the
//                       imaging conditions doimage / doimage_poynting
//                       (FD2DGPU.cu:2141-2196).  The source side stores its
//                       imaging quantity per step in HBM (the reference
//                       writes 100-step blocks to disk); the receiver side
//                       images against it at once.
//
// Wavefields of the rotation kernels are FP32 or FP64 (T): the rotation RTM
// images the source field against the receiver field normalised by the
// source's own energy, which amplifies the rounding of either wavefield
// wherever the source field is small, so the RTM runs both in FP64 (the
// forward modelling runs in FP32).
#include "common.hpp"
#include "kernels.hpp"

namespace ds2d {

namespace {

template <typename T>
struct Cw;
template <>
struct Cw<float> {
    static constexpr float c1 = DS2D_C1, c2 = DS2D_C2, c3 = DS2D_C3, c4 = DS2D_C4;
};
template <>
struct Cw<double> {
    static constexpr double c1 = 1225.0 / 1024.0, c2 = -245.0 / 3072.0, c3 = 49.0 / 5120.0, c4 = -5.0 / 7168.0;
};

// Staggered tap sums: fwd = sum c_i (f[c+i s] - f[c-(i-1) s]),
// bwd = sum c_i (f[c+(i-1) s] - f[c-i s]).
template <typename T>
__device__ __forceinline__ T fwd(const T* f, size_t c, int s) {
    return Cw<T>::c1 * (f[c + s] - f[c]) + Cw<T>::c2 * (f[c + 2 * s] - f[c - s]) +
           Cw<T>::c3 * (f[c + 3 * s] - f[c - 2 * s]) + Cw<T>::c4 * (f[c + 4 * s] - f[c - 3 * s]);
}
template <typename T>
__device__ __forceinline__ T bwd(const T* f, size_t c, int s) {
    return Cw<T>::c1 * (f[c] - f[c - s]) + Cw<T>::c2 * (f[c + s] - f[c - 2 * s]) +
           Cw<T>::c3 * (f[c + 2 * s] - f[c - 3 * s]) + Cw<T>::c4 * (f[c + 3 * s] - f[c - 4 * s]);
}

__device__ __forceinline__ bool update_cell(const SlotDesc& S, int& ix, int& iz, size_t& c) {
    iz = blockIdx.x * BZ + threadIdx.x;
    ix = blockIdx.y * BX + threadIdx.y;
    if (ix < NOP || ix >= S.nxp - NOP || iz < NOP || iz >= S.nzp - NOP) return false;
    c = (size_t)S.off + (size_t)ix * S.nzp + iz;
    return true;
}

}  // namespace

template <typename T>
__global__ __launch_bounds__(256) void k_rot_v(RotFieldT<T> f, const float* __restrict__ vp,
                                               const float* __restrict__ vs, const float* __restrict__ dampx,
                                               const float* __restrict__ dampz, const SlotDesc* __restrict__ slots,
                                               float dt_, float dx_, float dz_, float dir_) {
    const SlotDesc S = slots[blockIdx.z];
    int ix, iz;
    size_t c;
    if (!update_cell(S, ix, iz, c)) return;
    const T dt = dt_, dx = dx_, dz = dz_, dir = dir_;
    const int nzp = S.nzp;
    const T d1 = T(1) - dt * (T)dampx[c] / 2, d2 = T(1) + dt * (T)dampx[c] / 2;
    const T d3 = T(1) - dt * (T)dampz[c] / 2, d4 = T(1) + dt * (T)dampz[c] / 2;
    const T tx = bwd(f.th, c, nzp), tz = bwd(f.th, c, 1);
    const T ox = fwd(f.om, c, nzp), oz = fwd(f.om, c, 1);
    const T p = vp[c], s = vs[c];
    const T p2 = p * p, s2 = s * s;
    const T npx = (d1 * f.vpx[c] + dir * p2 * (dt / dx) * tx) / d2;
    const T npz = (d3 * f.vpz[c] + dir * p2 * (dt / dz) * tz) / d4;
    const T nsx = (d1 * f.vsx[c] + dir * s2 * (dt / dz) * oz) / d2;
    const T nsz = (d3 * f.vsz[c] - dir * s2 * (dt / dx) * ox) / d4;
    f.vpx[c] = npx; f.vpz[c] = npz; f.vsx[c] = nsx; f.vsz[c] = nsz;
    f.vx[c] = npx + nsx;
    f.vz[c] = npz + nsz;
}

template <typename T>
__global__ __launch_bounds__(256) void k_rot_p(RotFieldT<T> f, const float* __restrict__ dampx,
                                               const float* __restrict__ dampz, const SlotDesc* __restrict__ slots,
                                               float dt_, float dx_, float dz_, float dir_) {
    const SlotDesc S = slots[blockIdx.z];
    int ix, iz;
    size_t c;
    if (!update_cell(S, ix, iz, c)) return;
    const T dt = dt_, dx = dx_, dz = dz_, dir = dir_;
    const int nzp = S.nzp;
    const T d1 = T(1) - dt * (T)dampx[c] / 2, d2 = T(1) + dt * (T)dampx[c] / 2;
    const T d3 = T(1) - dt * (T)dampz[c] / 2, d4 = T(1) + dt * (T)dampz[c] / 2;
    const T xx = fwd(f.vx, c, nzp), xz = bwd(f.vx, c, 1);
    const T zx = bwd(f.vz, c, nzp), zz = fwd(f.vz, c, 1);
    const T ntx = (d1 * f.th_x[c] + dir * (dt / dx) * xx) / d2;
    const T ntz = (d3 * f.th_z[c] + dir * (dt / dz) * zz) / d4;
    const T nox = (d1 * f.om_x[c] - dir * (dt / dx) * zx) / d2;
    const T noz = (d3 * f.om_z[c] + dir * (dt / dz) * xz) / d4;
    f.th_x[c] = ntx; f.th_z[c] = ntz; f.om_x[c] = nox; f.om_z[c] = noz;
    f.th[c] = ntx + ntz;
    f.om[c] = nox + noz;
}

// theta += s[it] at each slot's source (elastic_rotation_addS_for injects into
// txx, which the reference aliases to theta, FD2DGPU.cu:2046-2049, 1407).
template <typename T>
__global__ void k_rot_src(T* th, const float* __restrict__ wav, const SlotDesc* __restrict__ slots, int nslots,
                          int it) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const SlotDesc S = slots[s];
    th[(size_t)S.off + (size_t)S.isx * S.nzp + S.isz] += (T)wav[it];
}

// FD2DGPU.cu:3218-3233 doimage on the physical window: the source P
// velocity against the receiver's P (PP) and S (PS) velocity parts,
// normalised by the source P energy + eps (float eps = 1e-6), accumulated in
// the slot's window images w2 = [pp, ps].
template <typename TS, typename TR>
__global__ __launch_bounds__(256) void k_rot_image(const TS* __restrict__ spx, const TS* __restrict__ spz,
                                                   const TR* __restrict__ rpx, const TR* __restrict__ rpz,
                                                   const TR* __restrict__ rsx, const TR* __restrict__ rsz,
                                                   float* w2, long long wstride, const SlotDesc* __restrict__ slots) {
    const SlotDesc S = slots[blockIdx.z];
    const int iz = blockIdx.x * BZ + threadIdx.x;
    const int ix = blockIdx.y * BX + threadIdx.y;
    if (ix >= S.nx || iz >= S.nz) return;
    const size_t c = (size_t)S.off + (size_t)(ix + S.pml) * S.nzp + iz + S.pml;
    const size_t o = (size_t)S.win_off + (size_t)ix * S.nz + iz;
    const TS ax = spx[c], az = spz[c];
    const TS e = (TS)(float)1e-6;
    const TS den = ax * ax + az * az + e;
    const TS pp = ax * (TS)rpx[c] + az * (TS)rpz[c];
    const TS ps = ax * (TS)rsx[c] + az * (TS)rsz[c];
    w2[o] += (float)(pp / den);
    w2[o + wstride] += (float)(ps / den);
}

// GPUapplysource_backward_from_record with stack (FD2DGPU.cu:3306-3327):
// vz += rec_z[tr][it], vx += rec_x[tr][it] at each receiver.
template <typename T>
__global__ void k_rot_inject(T* vz, T* vx, const float* __restrict__ rz, const float* __restrict__ rx,
                             const int* __restrict__ rix, const SlotDesc* __restrict__ slots, int nt, int it) {
    const SlotDesc S = slots[blockIdx.y];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.ntr) return;
    const size_t c = (size_t)S.off + (size_t)rix[S.rix_off + r] * S.nzp + S.igz;
    const size_t o = (size_t)S.rec_off + (size_t)r * nt + it;
    vz[c] += (T)rz[o];
    vx[c] += (T)rx[o];
}

// FD2DGPU.cu:1963-1986: VP = div v, VS = d(vx)/dz - d(vz)/dx on the
// reference's inclusive [nop, n-nop] (its taps stay inside the window).
__global__ __launch_bounds__(256) void k_helmholtz(const float* __restrict__ vx, const float* __restrict__ vz,
                                                   float* __restrict__ VP, float* __restrict__ VS,
                                                   const SlotDesc* __restrict__ slots, float dx, float dz) {
    const SlotDesc S = slots[blockIdx.z];
    const int iz = blockIdx.x * BZ + threadIdx.x;
    const int ix = blockIdx.y * BX + threadIdx.y;
    if (ix < NOP || ix > S.nxp - NOP || iz < NOP || iz > S.nzp - NOP) return;
    const size_t c = (size_t)S.off + (size_t)ix * S.nzp + iz;
    const int nzp = S.nzp;
    const float xx = bwd(vx, c, nzp), zz = bwd(vz, c, 1);
    const float xz = bwd(vx, c, 1), zx = bwd(vz, c, nzp);
    VP[c] = xx / dx + zz / dz;
    VS[c] = xz / dz - zx / dx;
}

namespace {

// Floats per cell and step of each image kind's quantity: 0 PP with the
// Poynting split of vector P (up x, up z, down x, down z); 1 PP, 2 PS, 3 SS
// vector images (x, z); 4 PP, 5 PS, 6 SS scalar images.
__host__ __device__ constexpr int decomp_k(int kind) { return kind == 0 ? 4 : kind <= 3 ? 2 : 1; }

// The imaging quantity of one side at padded cell c (side 1 = receiver: PS
// images the source's P against the receiver's S).
__device__ __forceinline__ void quantity(int kind, int side, const float* VP, const float* VS, size_t c, int nzp,
                                         float dx, float dz, float txz, float tzz, float vx, float vz, float* q) {
    const bool use_s = kind == 3 || kind == 6 || ((kind == 2 || kind == 5) && side == 1);
    if (kind >= 4) {
        q[0] = use_s ? VS[c] : VP[c];
        return;
    }
    float a, b;
    if (use_s) {  // vecS = (-dVS/dz, dVS/dx)
        a = -bwd(VS, c, 1) / dz;
        b = bwd(VS, c, nzp) / dx;
    } else {      // vecP = (dVP/dx, dVP/dz)
        a = bwd(VP, c, nzp) / dx;
        b = bwd(VP, c, 1) / dz;
    }
    if (kind != 0) {
        q[0] = a;
        q[1] = b;
        return;
    }
    // z energy flux -txz vx - tzz vz >= 0: down-going
    const float p = -txz * vx - tzz * vz;
    const bool dn = p >= 0;
    q[0] = dn ? 0.f : a;
    q[1] = dn ? 0.f : b;
    q[2] = dn ? a : 0.f;
    q[3] = dn ? b : 0.f;
}

}  // namespace

// store = 1: the source quantity of step it into hist; store = 0: the
// receiver quantity imaged against hist of step it into the window image.
template <int KIND>
__global__ __launch_bounds__(256) void k_decomp(const float* __restrict__ VP, const float* __restrict__ VS,
                                                const float* __restrict__ txz, const float* __restrict__ tzz,
                                                const float* __restrict__ vx, const float* __restrict__ vz,
                                                float* __restrict__ hist, float* __restrict__ win,
                                                const SlotDesc* __restrict__ slots, float dx, float dz, int store) {
    constexpr int K = decomp_k(KIND);
    const SlotDesc S = slots[blockIdx.z];
    const int iz = blockIdx.x * BZ + threadIdx.x;
    const int ix = blockIdx.y * BX + threadIdx.y;
    if (ix >= S.nx || iz >= S.nz) return;
    const size_t c = (size_t)S.off + (size_t)(ix + S.pml) * S.nzp + iz + S.pml;
    const size_t o = (size_t)S.win_off + (size_t)ix * S.nz + iz;
    float q[4];
    quantity(KIND, store ? 0 : 1, VP, VS, c, S.nzp, dx, dz, txz[c], tzz[c], vx[c], vz[c], q);
    float* h = hist + o * K;
    if (store) {
        if constexpr (K == 4) {
            *reinterpret_cast<float4*>(h) = make_float4(q[0], q[1], q[2], q[3]);
        } else if constexpr (K == 2) {
            *reinterpret_cast<float2*>(h) = make_float2(q[0], q[1]);
        } else {
            h[0] = q[0];
        }
        return;
    }
    const float e = 1e-6f;
    float r;
    if constexpr (K == 4) {  // doimage_poynting: (fu.bd + fd.bu) / (fu.fu + fd.fd + eps)
        const float4 f = *reinterpret_cast<const float4*>(h);
        const float il = (f.x * f.x + f.y * f.y) + (f.z * f.z + f.w * f.w) + e;
        const float cc = (f.x * q[2] + f.y * q[3]) + (f.z * q[0] + f.w * q[1]);
        r = cc / il;
    } else if constexpr (K == 2) {  // doimage, vector: f.b / (f.f + eps)
        const float2 f = *reinterpret_cast<const float2*>(h);
        const float il = (f.x * f.x + f.y * f.y) + e;
        r = (f.x * q[0] + f.y * q[1]) / il;
    } else {  // doimage, scalar: f b / (f^2 + eps)
        const float f = h[0];
        r = f * q[0] / (f * f + e);
    }
    win[o] += r;
}

// FD2DGPU.cu:2128-2134 GPUapplysource_backward: the receiver records are
// assigned (not added) to txx and tzz at each receiver.
__global__ void k_assign_rec(float* txx, float* tzz, const float* __restrict__ rz, const int* __restrict__ rix,
                             const SlotDesc* __restrict__ slots, int nt, int it) {
    const SlotDesc S = slots[blockIdx.y];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S.ntr) return;
    const size_t c = (size_t)S.off + (size_t)rix[S.rix_off + r] * S.nzp + S.igz;
    const float v = rz[(size_t)S.rec_off + (size_t)r * nt + it];
    txx[c] = v;
    tzz[c] = v;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static dim3 rot_grid(const BatchShape& b) { return dim3(ceil_div(b.nzp, BZ), ceil_div(b.max_nxp, BX), b.nslots); }
static dim3 win_grid(const BatchShape& b) { return dim3(ceil_div(b.nz, BZ), ceil_div(b.max_nxp, BX), b.nslots); }

template <typename T>
static void rot_v(hipStream_t st, const BatchShape& b, RotFieldT<T> f, Med m, float dir) {
    hipLaunchKernelGGL(k_rot_v<T>, rot_grid(b), dim3(BZ, BX), 0, st, f, m.vp, m.vs, m.dampx, m.dampz, b.slots, b.dt,
                       b.dx, b.dz, dir);
}
template <typename T>
static void rot_p(hipStream_t st, const BatchShape& b, RotFieldT<T> f, Med m, float dir) {
    hipLaunchKernelGGL(k_rot_p<T>, rot_grid(b), dim3(BZ, BX), 0, st, f, m.dampx, m.dampz, b.slots, b.dt, b.dx, b.dz,
                       dir);
}
void launch_rot_v(hipStream_t st, const BatchShape& b, RotField f, Med m, float dir) { rot_v<float>(st, b, f, m, dir); }
void launch_rot_v(hipStream_t st, const BatchShape& b, RotFieldD f, Med m, float dir) {
    rot_v<double>(st, b, f, m, dir);
}
void launch_rot_p(hipStream_t st, const BatchShape& b, RotField f, Med m, float dir) { rot_p<float>(st, b, f, m, dir); }
void launch_rot_p(hipStream_t st, const BatchShape& b, RotFieldD f, Med m, float dir) {
    rot_p<double>(st, b, f, m, dir);
}
void launch_rot_src(hipStream_t st, const BatchShape& b, float* th, const float* wav, int it) {
    hipLaunchKernelGGL(k_rot_src<float>, dim3(ceil_div(b.nslots, 64)), dim3(64), 0, st, th, wav, b.slots, b.nslots, it);
}
void launch_rot_src(hipStream_t st, const BatchShape& b, double* th, const float* wav, int it) {
    hipLaunchKernelGGL(k_rot_src<double>, dim3(ceil_div(b.nslots, 64)), dim3(64), 0, st, th, wav, b.slots, b.nslots,
                       it);
}
void launch_rot_image(hipStream_t st, const BatchShape& b, RotFieldD src, RotFieldD rcv, float* w2,
                      long long wstride) {
    hipLaunchKernelGGL((k_rot_image<double, double>), win_grid(b), dim3(BZ, BX), 0, st, src.vpx, src.vpz, rcv.vpx,
                       rcv.vpz, rcv.vsx, rcv.vsz, w2, wstride, b.slots);
}
void launch_rot_inject(hipStream_t st, const BatchShape& b, double* vz, double* vx, const float* rz, const float* rx,
                       int it) {
    hipLaunchKernelGGL(k_rot_inject<double>, dim3(ceil_div(b.max_ntr, 64), b.nslots), dim3(64), 0, st, vz, vx, rz, rx,
                       b.rix, b.slots, b.nt, it);
}
void launch_helmholtz(hipStream_t st, const BatchShape& b, const float* vx, const float* vz, float* VP, float* VS) {
    hipLaunchKernelGGL(k_helmholtz, rot_grid(b), dim3(BZ, BX), 0, st, vx, vz, VP, VS, b.slots, b.dx, b.dz);
}
int decomp_floats(int kind) { return decomp_k(kind); }
void launch_decomp(hipStream_t st, const BatchShape& b, int kind, const float* VP, const float* VS, EField f,
                   float* hist, float* win, int store) {
#define DS2D_DECOMP(KD)                                                                                        \
    hipLaunchKernelGGL(k_decomp<KD>, win_grid(b), dim3(BZ, BX), 0, st, VP, VS, f.txz, f.tzz, f.vx, f.vz, hist, \
                       win, b.slots, b.dx, b.dz, store)
    switch (kind) {
        case 0: DS2D_DECOMP(0); break;
        case 1: DS2D_DECOMP(1); break;
        case 2: DS2D_DECOMP(2); break;
        case 3: DS2D_DECOMP(3); break;
        case 4: DS2D_DECOMP(4); break;
        case 5: DS2D_DECOMP(5); break;
        default: DS2D_DECOMP(6); break;
    }
#undef DS2D_DECOMP
}
void launch_assign_rec(hipStream_t st, const BatchShape& b, float* txx, float* tzz, const float* rz, int it) {
    hipLaunchKernelGGL(k_assign_rec, dim3(ceil_div(b.max_ntr, 64), b.nslots), dim3(64), 0, st, txx, tzz, rz, b.rix,
                       b.slots, b.nt, it);
}

}  // namespace ds2d
