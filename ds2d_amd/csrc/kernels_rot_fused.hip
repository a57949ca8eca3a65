// One time step of the P/S-decoupled (theta/omega rotation) scheme in a
// single launch for gfx950, batched over shot windows: the velocity parts from
// the theta / omega gradients (GPUcalculate_rotation_v, FD2DGPU.cu:3005-3040)
// and theta / omega from the new velocity (GPUcalculate_rotation_p,
// FD2DGPU.cu:3042-3080), with the source sample on theta for this step's
// velocity update (FD2DGPU.cu:2046-2049, 1407) and the recorder of vz, vx.
//
// Layout (as k_fused_step): a workgroup owns columns [x0, x1) and a block of
// rows of one window, one z row per lane, and marches along x; at column j it
// updates the velocity parts of column j and theta / omega of column j-4.
// x taps in register queues, z taps in a double-buffered LDS column, state
// ping-ponged (A read, B written).
//
// State per cell: theta, omega, the four velocity parts and, in the damping
// strips, one split of each of theta and omega (theta_x, omega_x; the other
// split is the total minus it).  Outside the strips the split update equals
// the unsplit one (theta += dt/dx d(vx)/dx + dt/dz d(vz)/dz) in exact
// arithmetic, so only the totals are kept there: 8 floats in (6 fields, vp,
// vs) and 6 out per cell and step, 56 B, against 120 B for k_rot_v + k_rot_p.
#include "common.hpp"
#include "kernels.hpp"

namespace ds2d {

namespace {

struct RotCol {
    float th4, om4;              // theta, omega of column j+4 (x-tap queues)
    float px, pz, sx, sz;        // velocity parts of column j
    float vp, vs;                // of column j
    float dxv, dzv;              // damping of column j (damping lanes)
    float dxs, dzs, thx, omx;    // damping and theta_x, omega_x of column j-4 (damping lanes)
    int rec;                     // trace recorded at column j (receiver-row lane; -1 none)
};

__device__ __forceinline__ float fresh_rot(float x) {
    float r;
    asm("v_mov_b32 %0, %1" : "=&v"(r) : "v"(x));
    return r;
}

__device__ __forceinline__ bool damp_rot(const SlotDesc& S, int ix, int iz) {
    return S.boundary == 0 && (ix < S.pml || ix > S.nxp - S.pml || iz < S.pml || iz > S.nzp - S.pml);
}

}  // namespace

template <int T, bool PML>
__global__ __launch_bounds__(T) void k_rot_fused(RotFusedArgs a, const ChunkDesc* __restrict__ chunks, int nchunks,
                                                 int per_xcd, const int* __restrict__ rmap,
                                                 const SlotDesc* __restrict__ slots, float dt, float dx, float dz,
                                                 int nt, int it) {
    const int b = blockIdx.x;
    const int ch = (b % 8) * per_xcd + b / 8;
    if (ch >= nchunks) return;
    const ChunkDesc C = chunks[ch];
    const SlotDesc S = slots[C.slot];
    const int L = threadIdx.x;
    const int z = C.z0 - FUSED_HALO + L;
    const int nzp = S.nzp, nxp = S.nxp;
    const bool zin = z >= 0 && z < nzp;
    const bool zown = zin && L >= FUSED_HALO && L < T - FUSED_HALO;
    const bool vlane = L >= 4 && L < T - 4;
    const bool zupd = z >= NOP && z < nzp - NOP;
    const int x0 = C.x0, x1 = C.x1;
    const float c1 = DS2D_C1, c2 = DS2D_C2, c3 = DS2D_C3, c4 = DS2D_C4;
    const float rx = dt / dx, rz = dt / dz;

    // z-tap columns: theta, omega (column j), new vx, vz (column j-4)
    __shared__ float lds[2][4][T];

    const long long n = a.n;
    const float* A = a.A;  // theta omega vpx vpz vsx vsz theta_x omega_x
    const uint32_t colb = (uint32_t)nzp * 4u;
    const uint32_t zo4 = (uint32_t)(S.off + z) * 4u;
    const uint32_t zc4 = (uint32_t)(S.off + min(max(z, 0), nzp - 1)) * 4u;
    auto at = [](const float* p, uint32_t boff) -> float {
        return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + boff);
    };
    auto st = [](float* p, uint32_t boff, float v) {
        *reinterpret_cast<float*>(reinterpret_cast<char*>(p) + boff) = v;
    };
    auto ld = [&](const float* p, int col) -> float {
        return (zin && col >= 0 && col < nxp) ? at(p, zo4 + (uint32_t)col * colb) : 0.f;
    };
    auto ldc = [&](const float* p, int col) -> float {
        return at(p, zc4 + (uint32_t)min(max(col, 0), nxp - 1) * colb);
    };
    auto inu = [&](int col) { return zupd && col >= NOP && col < nxp - NOP; };

    auto fetch = [&](int j) {
        RotCol o;
        o.th4 = ldc(A, j + 4);
        o.om4 = ldc(A + n, j + 4);
        o.px = ldc(A + 2 * n, j);
        o.pz = ldc(A + 3 * n, j);
        o.sx = ldc(A + 4 * n, j);
        o.sz = ldc(A + 5 * n, j);
        o.vp = ldc(a.vp, j);
        o.vs = ldc(a.vs, j);
        o.dxv = o.dzv = o.dxs = o.dzs = o.thx = o.omx = 0.f;
        if (PML && inu(j) && damp_rot(S, j, z)) {
            const uint32_t c = zo4 + (uint32_t)j * colb;
            o.dxv = at(a.dampx, c);
            o.dzv = at(a.dampz, c);
        }
        const int cs = j - 4;
        if (PML && inu(cs) && damp_rot(S, cs, z)) {
            const uint32_t c = zo4 + (uint32_t)cs * colb;
            o.dxs = at(a.dampx, c);
            o.dzs = at(a.dampz, c);
            o.thx = at(A + 6 * n, c);
            o.omx = at(A + 7 * n, c);
        }
        o.rec = -1;
        if (z == S.igz && j >= x0 && j < x1) o.rec = rmap[S.rmap_off + j];
        return o;
    };

    // queues: theta, omega of columns j-4+k (old); vx, vz of columns j-8+k (new)
    float qth[9], qom[9], qvx[9], qvz[9];
    const int j0 = x0 - 4;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        qth[k + 1] = ld(A, j0 - 4 + k);
        qom[k + 1] = ld(A + n, j0 - 4 + k);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) qvx[k] = qvz[k] = 0.f;

    const float w = a.wav[it];
    const int lsrc = S.isz - (C.z0 - FUSED_HALO);
    const bool srow = z == S.isz;

    auto column = [&](const int j, const RotCol& o, const int p) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            qth[k] = qth[k + 1];
            qom[k] = qom[k + 1];
            qvx[k] = qvx[k + 1];
            qvz[k] = qvz[k + 1];
        }
        qth[8] = fresh_rot(o.th4);
        qom[8] = fresh_rot(o.om4);
        lds[p][0][L] = qth[4];
        lds[p][1][L] = qom[4];
        lds[p][2][L] = qvx[4];  // new vx, vz of column j-4
        lds[p][3][L] = qvz[4];
        __syncthreads();

        // ---- velocity parts of column j (theta carries the source sample here)
        float px = o.px, pz = o.pz, sx = o.sx, sz = o.sz;
        if (vlane && inu(j)) {
            const float* L0 = lds[p][0];
            const float* L1 = lds[p][1];
            const bool sxw = srow && (unsigned)(S.isx - j + 4) < 8u;  // source column in j-4 .. j+3
            auto th = [&](int k) { return (sxw && j + k == S.isx) ? qth[4 + k] + w : qth[4 + k]; };
            auto thz = [&](int k) { return (j == S.isx && L + k == lsrc) ? L0[L + k] + w : L0[L + k]; };
            const float tx = c1 * (th(0) - th(-1)) + c2 * (th(1) - th(-2)) + c3 * (th(2) - th(-3)) + c4 * (th(3) - th(-4));
            const float tz = c1 * (thz(0) - thz(-1)) + c2 * (thz(1) - thz(-2)) + c3 * (thz(2) - thz(-3)) +
                             c4 * (thz(3) - thz(-4));
            const float ox = c1 * (qom[5] - qom[4]) + c2 * (qom[6] - qom[3]) + c3 * (qom[7] - qom[2]) + c4 * (qom[8] - qom[1]);
            const float oz = c1 * (L1[L + 1] - L1[L]) + c2 * (L1[L + 2] - L1[L - 1]) + c3 * (L1[L + 3] - L1[L - 2]) +
                             c4 * (L1[L + 4] - L1[L - 3]);
            const float p2 = o.vp * o.vp, s2 = o.vs * o.vs;
            const bool dv = PML && damp_rot(S, j, z);
            if (PML && __any(dv)) {
                const float d1 = 1.f - dt * o.dxv / 2, r2 = 1.f / (1.f + dt * o.dxv / 2);
                const float d3 = 1.f - dt * o.dzv / 2, r4 = 1.f / (1.f + dt * o.dzv / 2);
                px = (d1 * o.px + p2 * rx * tx) * r2;
                pz = (d3 * o.pz + p2 * rz * tz) * r4;
                sx = (d1 * o.sx + s2 * rz * oz) * r2;
                sz = (d3 * o.sz - s2 * rx * ox) * r4;
            } else {
                px = o.px + p2 * rx * tx;
                pz = o.pz + p2 * rz * tz;
                sx = o.sx + s2 * rz * oz;
                sz = o.sz - s2 * rx * ox;
            }
        }
        const float vx = px + sx, vz = pz + sz;
        qvx[8] = vx;
        qvz[8] = vz;
        if (zown && j >= x0 && j < x1) {
            const uint32_t c = zo4 + (uint32_t)j * colb;
            st(a.B + 2 * n, c, px);
            st(a.B + 3 * n, c, pz);
            st(a.B + 4 * n, c, sx);
            st(a.B + 5 * n, c, sz);
            if (o.rec >= 0) {
                const size_t q = (size_t)S.rec_off + (size_t)o.rec * nt + it;
                a.rec_z[q] = vz;
                a.rec_x[q] = vx;
            }
        }

        // ---- theta, omega of column cs = j - 4
        const int cs = j - 4;
        if (zown && cs >= x0 && cs < x1) {
            const uint32_t c = zo4 + (uint32_t)cs * colb;
            float th = qth[0], om = qom[0], thx = o.thx, omx = o.omx;
            if (inu(cs)) {
                const float* L2 = lds[p][2];
                const float* L3 = lds[p][3];
                const float xx = c1 * (qvx[5] - qvx[4]) + c2 * (qvx[6] - qvx[3]) + c3 * (qvx[7] - qvx[2]) +
                                 c4 * (qvx[8] - qvx[1]);
                const float zx = c1 * (qvz[4] - qvz[3]) + c2 * (qvz[5] - qvz[2]) + c3 * (qvz[6] - qvz[1]) +
                                 c4 * (qvz[7] - qvz[0]);
                const float xz = c1 * (L2[L] - L2[L - 1]) + c2 * (L2[L + 1] - L2[L - 2]) + c3 * (L2[L + 2] - L2[L - 3]) +
                                 c4 * (L2[L + 3] - L2[L - 4]);
                const float zz = c1 * (L3[L + 1] - L3[L]) + c2 * (L3[L + 2] - L3[L - 1]) + c3 * (L3[L + 3] - L3[L - 2]) +
                                 c4 * (L3[L + 4] - L3[L - 3]);
                const bool ds = PML && damp_rot(S, cs, z);
                if (PML && __any(ds)) {
                    const float d1 = 1.f - dt * o.dxs / 2, r2 = 1.f / (1.f + dt * o.dxs / 2);
                    const float d3 = 1.f - dt * o.dzs / 2, r4 = 1.f / (1.f + dt * o.dzs / 2);
                    const float ntx = (d1 * o.thx + rx * xx) * r2;
                    const float ntz = (d3 * (th - o.thx) + rz * zz) * r4;
                    const float nox = (d1 * o.omx - rx * zx) * r2;
                    const float noz = (d3 * (om - o.omx) + rz * xz) * r4;
                    th = ntx + ntz;
                    om = nox + noz;
                    thx = ntx;
                    omx = nox;
                } else {
                    th = th + (rx * xx + rz * zz);
                    om = om + (rz * xz - rx * zx);
                }
            }
            st(a.B, c, th);
            st(a.B + n, c, om);
            if (PML && damp_rot(S, cs, z)) {
                st(a.B + 6 * n, c, thx);
                st(a.B + 7 * n, c, omx);
            }
        }
    };

    const int jend = j0 + (x1 + 4 - j0 + 1) / 2 * 2;
    RotCol cA = fetch(j0), cB = fetch(j0 + 1);
    for (int j = j0; j < jend; j += 2) {
        const bool more = j + 2 < jend;
        column(j, cA, 0);
        if (more) cA = fetch(j + 2);
        column(j + 1, cB, 1);
        if (more) cB = fetch(j + 3);
    }
}

void launch_rot_fused(hipStream_t st, const BatchShape& b, const RotFusedArgs& a, const ChunkDesc* chunks,
                      int nchunks, const int* rmap, int it, int threads, bool pml) {
    const int per_xcd = (nchunks + 7) / 8;
    const dim3 grid(per_xcd * 8);
#define DS2D_ROTF(TT, MM)                                                                                        \
    hipLaunchKernelGGL((k_rot_fused<TT, MM>), grid, dim3(TT), 0, st, a, chunks, nchunks, per_xcd, rmap, b.slots, \
                       b.dt, b.dx, b.dz, b.nt, it)
    if (threads == 512) {
        if (pml)
            DS2D_ROTF(512, true);
        else
            DS2D_ROTF(512, false);
    } else {
        if (pml)
            DS2D_ROTF(256, true);
        else
            DS2D_ROTF(256, false);
    }
#undef DS2D_ROTF
}

}  // namespace ds2d
