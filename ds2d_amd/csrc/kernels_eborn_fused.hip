// One forward step of the plain-L2 elastic Born pair in a single launch for
// gfx950: background velocity and stress (transient source on txx, tzz) and
// scattered velocity and stress with the Born source of add_born_source
// (born.cu:351-405, direction 1), plus the recorder of the scattered
// velocity, batched over shot windows with the random boundary (no damping,
// unsplit fields, update region [nop, n-nop)^2).  It replaces the pair
// k_eb_v + k_eb_s of one step (kernels_eborn.hip) with the same arithmetic,
// expression for expression (the compiler may still contract other products
// into FMAs, so results agree to FP32 rounding rather than bit for bit).
//
// Work layout (as k_fused_step): a workgroup owns columns [x0, x1) and a block
// of rows of one window, one z row per lane, and marches along x; at column j
// it computes the new velocity of both fields at column j and their new stress
// at column j-4.  x taps live in per-lane register queues, z taps in a
// double-buffered LDS column (one barrier per column).  State is ping-ponged:
// A (10 fields: background vx vz txx tzz txz, then the scattered five) is read
// and B written, so the halo columns each workgroup recomputes never race.
// Per cell and step it moves 15 floats in (both states, 1/rho's rho, lambda,
// mu, the two image samples) and 10 out: 100 B for two cell-updates.
#include "common.hpp"
#include "kernels.hpp"

namespace ds2d {

namespace {

// Operands of march column j for one lane.
struct EbCol {
    float gxx4, gxz4, gzz, gvx, gvz;  // background: txx, txz of column j+4; tzz, vx, vz of column j
    float sxx4, sxz4, szz, svx, svz;  // scattered field, the same
    float rho;                        // rho of column j
    float lam, mu, dmp, dms;          // moduli and image samples of column j-4
    int rec;                          // trace recorded at column j-4 (receiver-row lane only; -1 none)
};

__device__ __forceinline__ float fresh_eb(float x) {
    float r;
    asm("v_mov_b32 %0, %1" : "=&v"(r) : "v"(x));
    return r;
}

// Staggered 8th-order differences of f(k), the operand k cells away:
// plus: sum c_i (f(i) - f(1-i)); minus: sum c_i (f(i-1) - f(-i)), as in the
// per-cell kernels.
template <typename F>
__device__ __forceinline__ float ebp(F f) {
    return DS2D_C1 * (f(1) - f(0)) + DS2D_C2 * (f(2) - f(-1)) + DS2D_C3 * (f(3) - f(-2)) + DS2D_C4 * (f(4) - f(-3));
}
template <typename F>
__device__ __forceinline__ float ebm(F f) {
    return DS2D_C1 * (f(0) - f(-1)) + DS2D_C2 * (f(1) - f(-2)) + DS2D_C3 * (f(2) - f(-3)) + DS2D_C4 * (f(3) - f(-4));
}

}  // namespace

template <int T>
__global__ __launch_bounds__(T) void k_eb_fused(EbFusedArgs a, const ChunkDesc* __restrict__ chunks, int nchunks,
                                                int per_xcd, const int* __restrict__ rmap,
                                                const SlotDesc* __restrict__ slots, float cax, float caz, float idx,
                                                float idz, float w, float sign, int it, int nt, int hist_mode) {
    const int b = blockIdx.x;
    const int ch = (b % 8) * per_xcd + b / 8;  // a contiguous run of chunks per XCD
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
    const bool zphys = z >= S.pml && z < S.pml + S.nz;
    const int x0 = C.x0, x1 = C.x1;

    // z-tap columns: background txz, tzz (column j), scattered txz, tzz (j),
    // new background vx, vz (j-4), new scattered vx, vz (j-4)
    __shared__ float lds[2][8][T];

    const long long n = a.n;
    const float* A = a.A;
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
    // clamped, unconditional load (the value of a clamped cell is never used)
    auto ldc = [&](const float* p, int col) -> float {
        return at(p, zc4 + (uint32_t)min(max(col, 0), nxp - 1) * colb);
    };
    auto fld = [&](int k) { return A + k * n; };  // field k of A (0-4 background, 5-9 scattered)

    auto fetch = [&](int j) {
        EbCol o;
        o.gxx4 = ldc(fld(2), j + 4);
        o.gxz4 = ldc(fld(4), j + 4);
        o.gzz = ldc(fld(3), j);
        o.gvx = ldc(fld(0), j);
        o.gvz = ldc(fld(1), j);
        o.sxx4 = ldc(fld(7), j + 4);
        o.sxz4 = ldc(fld(9), j + 4);
        o.szz = ldc(fld(8), j);
        o.svx = ldc(fld(5), j);
        o.svz = ldc(fld(6), j);
        o.rho = ldc(a.m.rho, j);
        const int cs = j - 4;
        o.lam = ldc(a.m.lam, cs);
        o.mu = ldc(a.m.mu, cs);
        o.dmp = ldc(a.m.dmp, cs);
        o.dms = ldc(a.m.dms, cs);
        o.rec = -1;
        if (z == S.igz && cs >= x0 && cs < x1 && cs >= NOP && cs < nxp - NOP) o.rec = rmap[S.rmap_off + cs];
        return o;
    };

    // queues: q*[k] is column j-4+k (old stress), v*[k] column j-8+k (new vx),
    // u*[k] column j-7+k (new vz); g = background, s = scattered
    float gxx[9], gxz[9], gzz[5], gvx[9], gvz[8];
    float sxx[9], sxz[9], szz[5], svx[9], svz[8];
    const int j0 = x0 - 4;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        gxx[k + 1] = ld(fld(2), j0 - 4 + k);
        gxz[k + 1] = ld(fld(4), j0 - 4 + k);
        sxx[k + 1] = ld(fld(7), j0 - 4 + k);
        sxz[k + 1] = ld(fld(9), j0 - 4 + k);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        gzz[k + 1] = ld(fld(3), j0 - 4 + k);
        szz[k + 1] = ld(fld(8), j0 - 4 + k);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) gvx[k] = svx[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) gvz[k] = svz[k] = 0.f;

    const int lsrc = S.isz - (C.z0 - FUSED_HALO);  // LDS row of the source
    const bool srow = z == S.isz;

    auto column = [&](const int j, const EbCol& o, const int p) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            gxx[k] = gxx[k + 1];
            gxz[k] = gxz[k + 1];
            gvx[k] = gvx[k + 1];
            sxx[k] = sxx[k + 1];
            sxz[k] = sxz[k + 1];
            svx[k] = svx[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gzz[k] = gzz[k + 1];
            szz[k] = szz[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            gvz[k] = gvz[k + 1];
            svz[k] = svz[k + 1];
        }
        gxx[8] = fresh_eb(o.gxx4);
        gxz[8] = fresh_eb(o.gxz4);
        gzz[4] = fresh_eb(o.gzz);
        sxx[8] = fresh_eb(o.sxx4);
        sxz[8] = fresh_eb(o.sxz4);
        szz[4] = fresh_eb(o.szz);
        lds[p][0][L] = gxz[4];
        lds[p][1][L] = gzz[4];
        lds[p][2][L] = sxz[4];
        lds[p][3][L] = szz[4];
        lds[p][4][L] = gvx[4];  // new velocity of column j-4
        lds[p][5][L] = gvz[3];
        lds[p][6][L] = svx[4];
        lds[p][7][L] = svz[3];
        __syncthreads();

        // ---- velocity of column j (k_eb_v)
        float ngvx = o.gvx, ngvz = o.gvz, nsvx = o.svx, nsvz = o.svz;
        if (vlane && zupd && j >= NOP && j < nxp - NOP) {
            const float* L0 = lds[p][0];
            const float* L1 = lds[p][1];
            const float* L2 = lds[p][2];
            const float* L3 = lds[p][3];
            const float br = 1.f / o.rho;
            // the source sample sits on txx, tzz of the source cell for these taps only
            const bool sx = srow && (unsigned)(S.isx - j + 3) < 8u;
            auto txx = [&](int k) { return (sx && j + k == S.isx) ? gxx[4 + k] + w : gxx[4 + k]; };
            auto tzz = [&](int k) { return (j == S.isx && L + k == lsrc) ? L1[L + k] + w : L1[L + k]; };
            ngvx += sign * (br * cax * ebp(txx) + br * caz * ebm([&](int k) { return L0[L + k]; }));
            ngvz += sign * (br * cax * ebm([&](int k) { return gxz[4 + k]; }) + br * caz * ebp(tzz));
            nsvx += br * cax * ebp([&](int k) { return sxx[4 + k]; }) + br * caz * ebm([&](int k) { return L2[L + k]; });
            nsvz += br * cax * ebm([&](int k) { return sxz[4 + k]; }) + br * caz * ebp([&](int k) { return L3[L + k]; });
        }
        gvx[8] = ngvx;
        gvz[7] = ngvz;
        svx[8] = nsvx;
        svz[7] = nsvz;
        if (zown && j >= x0 && j < x1) {
            const uint32_t c = zo4 + (uint32_t)j * colb;
            st(a.B, c, ngvx);
            st(a.B + n, c, ngvz);
            st(a.B + 5 * n, c, nsvx);
            st(a.B + 6 * n, c, nsvz);
        }

        // ---- stress of column cs = j - 4 (k_eb_s)
        const int cs = j - 4;
        if (zown && cs >= x0 && cs < x1) {
            float gtxx = gxx[0], gtzz = gzz[0], gtxz = gxz[0];
            float stxx = sxx[0], stzz = szz[0], stxz = sxz[0];
            if (zupd && cs >= NOP && cs < nxp - NOP) {
                const float* L4 = lds[p][4];
                const float* L5 = lds[p][5];
                const float* L6 = lds[p][6];
                const float* L7 = lds[p][7];
                const float l = o.lam, mu = o.mu, l2m = l + 2 * mu;
                const bool inside = zphys && cs >= S.pml && cs < S.pml + S.nx;
                // background: q[k] of vx is column cs+k = v*[4+k]; of vz, u*[3+k]
                const float dvx = ebm([&](int k) { return gvx[4 + k]; }), dvz = ebm([&](int k) { return L5[L + k]; });
                const float dvxz = ebp([&](int k) { return L4[L + k]; }), dvzx = ebp([&](int k) { return gvz[3 + k]; });
                gtxx += sign * (cax * l2m * dvx + caz * l * dvz);
                gtzz += sign * (cax * l * dvx + caz * l2m * dvz);
                gtxz += sign * (mu * (cax * dvzx + caz * dvxz));
                const float exx = dvx * idx, ezz = dvz * idz, exz = dvzx * idx + dvxz * idz;
                if (hist_mode == 1 && inside) {
                    const size_t h =
                        (size_t)S.hist_off + ((size_t)it * S.nx + (cs - S.pml)) * S.nz + (size_t)(z - S.pml);
                    a.hist[h] = exx;
                    a.hist[h + a.hstride] = ezz;
                    a.hist[h + 2 * a.hstride] = exz;
                }
                // scattered field with the Born source
                const float bvx = ebm([&](int k) { return svx[4 + k]; }), bvz = ebm([&](int k) { return L7[L + k]; });
                const float bvxz = ebp([&](int k) { return L6[L + k]; }), bvzx = ebp([&](int k) { return svz[3 + k]; });
                float txx = stxx + cax * l2m * bvx + caz * l * bvz;
                float tzz = stzz + cax * l * bvx + caz * l2m * bvz;
                float txz = stxz + mu * (cax * bvzx + caz * bvxz);
                if (inside) {
                    const float mp = -o.dmp, ms = -o.dms;
                    txx += 2 * l2m * mp * exx + 2 * (l2m * mp - 2 * mu * ms) * ezz;
                    tzz += 2 * (l2m * mp - 2 * mu * ms) * exx + 2 * l2m * mp * ezz;
                    txz += 2 * mu * ms * exz;
                }
                stxx = txx;
                stzz = tzz;
                stxz = txz;
                if (o.rec >= 0 && a.rec_z) {
                    const size_t q = (size_t)S.rec_off + (size_t)o.rec * nt + it;
                    a.rec_z[q] = svz[3];
                    a.rec_x[q] = svx[4];
                }
            }
            const uint32_t c = zo4 + (uint32_t)cs * colb;
            st(a.B + 2 * n, c, gtxx);
            st(a.B + 3 * n, c, gtzz);
            st(a.B + 4 * n, c, gtxz);
            st(a.B + 7 * n, c, stxx);
            st(a.B + 8 * n, c, stzz);
            st(a.B + 9 * n, c, stxz);
        }
    };

    // two columns in flight; the trailing column of an odd count stores nothing
    const int jend = j0 + (x1 + 4 - j0 + 1) / 2 * 2;
    EbCol cA = fetch(j0), cB = fetch(j0 + 1);
    for (int j = j0; j < jend; j += 2) {
        const bool more = j + 2 < jend;
        column(j, cA, 0);
        if (more) cA = fetch(j + 2);
        column(j + 1, cB, 1);
        if (more) cB = fetch(j + 3);
    }
}

// One adjoint step of the pair (k_eb_adj_v + k_eb_adj_s of step it, with the
// receiver residual of step it+1) in a single x-marching launch.  The products
// C s* (wa, wb, wc) and (1/rho) v* (ux, uz) that the two per-cell kernels hand
// each other through memory are formed in registers here: wa, wb, wc of the
// adjoint stress of column j+4 as it is loaded (zero outside the update
// region, as the per-cell kernels leave them), ux, uz of column j as its
// velocity is updated.  UNDO (no strain history): the background is taken one
// step back alongside, velocity at column j from its stress (source sample w),
// stress at column j-4 from the new velocity, whose differences are also the
// strains of the gradient; otherwise the strains come from the history.  The
// adjoint state (lm) and the background are ping-ponged (A read, B written);
// the gradients accumulate in place (each cell has one owner).
template <int T, bool UNDO>
__global__ __launch_bounds__(T) void k_eb_adj_fused(EbAdjArgs a, const ChunkDesc* __restrict__ chunks,
                                                    int nchunks, int per_xcd, const int* __restrict__ rmap,
                                                    const SlotDesc* __restrict__ slots, float cax, float caz,
                                                    float idx, float idz, float w, int it, int it_inj, int nt) {
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
    const bool zphys = z >= S.pml && z < S.pml + S.nz;
    const int x0 = C.x0, x1 = C.x1;

    // z-tap columns: background txz, tzz (column j), wc, wb (j), ux, uz (j-4),
    // new background vx, vz (j-4)
    __shared__ float lds[2][8][T];

    const long long n = a.n;
    const uint32_t colb = (uint32_t)nzp * 4u;
    const uint32_t zo4 = (uint32_t)(S.off + z) * 4u;
    const uint32_t zc4 = (uint32_t)(S.off + min(max(z, 0), nzp - 1)) * 4u;
    auto at = [](const float* p, uint32_t boff) -> float {
        return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + boff);
    };
    auto st = [](float* p, uint32_t boff, float v) {
        *reinterpret_cast<float*>(reinterpret_cast<char*>(p) + boff) = v;
    };
    auto ldc = [&](const float* p, int col) -> float {
        return at(p, zc4 + (uint32_t)min(max(col, 0), nxp - 1) * colb);
    };
    auto inu = [&](int col) { return zupd && col >= NOP && col < nxp - NOP; };
    const float* Alm = a.Alm;  // vx vz txx tzz txz
    const float* Abg = a.Abg;

    struct AdjCol {
        float wa4, wb4, wc4;            // products C s* of column j+4
        float lvx, lvz, rho;            // adjoint velocity and rho of column j
        float ix, iz;                   // residual injected at column j (receiver-row lane)
        float gxx4, gxz4, gzz, gvx, gvz;  // background: txx, txz of column j+4; tzz, vx, vz of column j
    };
    auto fetch = [&](int j) {
        AdjCol o;
        const int c4 = j + 4;
        const float txx = ldc(Alm + 2 * n, c4), tzz = ldc(Alm + 3 * n, c4), txz = ldc(Alm + 4 * n, c4);
        const float l = ldc(a.m.lam, c4), mu = ldc(a.m.mu, c4), l2m = l + 2 * mu;
        const bool u4 = inu(c4);
        o.wa4 = u4 ? l2m * txx + l * tzz : 0.f;
        o.wb4 = u4 ? l * txx + l2m * tzz : 0.f;
        o.wc4 = u4 ? mu * txz : 0.f;
        o.lvx = ldc(Alm, j);
        o.lvz = ldc(Alm + n, j);
        o.rho = ldc(a.m.rho, j);
        o.ix = o.iz = 0.f;
        // every computed column, halo included: the stress taps of owned columns read them
        if (it_inj >= 0 && z == S.igz && inu(j)) {
            const int r = rmap[S.rmap_off + j];
            if (r >= 0) {
                const size_t q = (size_t)S.rec_off + (size_t)r * nt + it_inj;
                o.ix = a.res_x[q];
                o.iz = a.res_z[q];
            }
        }
        if (UNDO) {
            o.gxx4 = ldc(Abg + 2 * n, c4);
            o.gxz4 = ldc(Abg + 4 * n, c4);
            o.gzz = ldc(Abg + 3 * n, j);
            o.gvx = ldc(Abg, j);
            o.gvz = ldc(Abg + n, j);
        } else {
            o.gxx4 = o.gxz4 = o.gzz = o.gvx = o.gvz = 0.f;
        }
        return o;
    };

    // queues: wa, wc (x taps, column j-4+k), wb (column j+k: z taps only),
    // ux, uz (columns j-8+k, j-7+k); background as in the forward step
    float qwa[9], qwc[9], qwb[5], qux[9], quz[8];
    float gxx[9], gxz[9], gzz[5], gvx[9], gvz[8];
    const int j0 = x0 - 4;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int c = j0 - 4 + k;
        const bool u = zin && c >= 0 && c < nxp && inu(c);
        const float txx = u ? ldc(Alm + 2 * n, c) : 0.f, tzz = u ? ldc(Alm + 3 * n, c) : 0.f;
        const float txz = u ? ldc(Alm + 4 * n, c) : 0.f;
        const float l = ldc(a.m.lam, c), mu = ldc(a.m.mu, c), l2m = l + 2 * mu;
        qwa[k + 1] = u ? l2m * txx + l * tzz : 0.f;
        qwc[k + 1] = u ? mu * txz : 0.f;
        if (k >= 4) qwb[k - 3] = u ? l * txx + l2m * tzz : 0.f;
        gxx[k + 1] = UNDO && zin && c >= 0 && c < nxp ? ldc(Abg + 2 * n, c) : 0.f;
        gxz[k + 1] = UNDO && zin && c >= 0 && c < nxp ? ldc(Abg + 4 * n, c) : 0.f;
        if (k < 4) gzz[k + 1] = UNDO && zin && c >= 0 && c < nxp ? ldc(Abg + 3 * n, c) : 0.f;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) qux[k] = gvx[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) quz[k] = gvz[k] = 0.f;

    const int lsrc = S.isz - (C.z0 - FUSED_HALO);
    const bool srow = z == S.isz;

    auto column = [&](const int j, const AdjCol& o, const int p) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            qwa[k] = qwa[k + 1];
            qwc[k] = qwc[k + 1];
            qux[k] = qux[k + 1];
            gxx[k] = gxx[k + 1];
            gxz[k] = gxz[k + 1];
            gvx[k] = gvx[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            qwb[k] = qwb[k + 1];
            gzz[k] = gzz[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            quz[k] = quz[k + 1];
            gvz[k] = gvz[k + 1];
        }
        // operands of the stress part of column j-4, issued before the barrier
        const int cs = j - 4;
        const bool own_s = zown && cs >= x0 && cs < x1;
        const uint32_t cS = zo4 + (uint32_t)cs * colb;
        const bool ins = own_s && inu(cs) && zphys && cs >= S.pml && cs < S.pml + S.nx;
        float e_ltxx = 0.f, e_ltzz = 0.f, e_ltxz = 0.f, e_l = 0.f, e_mu = 0.f, e_gpp = 0.f, e_gps = 0.f;
        if (own_s) {
            e_ltxx = at(Alm + 2 * n, cS);
            e_ltzz = at(Alm + 3 * n, cS);
            e_ltxz = at(Alm + 4 * n, cS);
            e_l = at(a.m.lam, cS);
            e_mu = at(a.m.mu, cS);
        }
        if (ins) {
            e_gpp = at(a.gpp, cS);
            e_gps = at(a.gps, cS);
        }
        qwa[8] = fresh_eb(o.wa4);
        qwc[8] = fresh_eb(o.wc4);
        qwb[4] = fresh_eb(o.wb4);
        gxx[8] = fresh_eb(o.gxx4);
        gxz[8] = fresh_eb(o.gxz4);
        gzz[4] = fresh_eb(o.gzz);
        lds[p][0][L] = gxz[4];
        lds[p][1][L] = gzz[4];
        lds[p][2][L] = qwc[4];
        lds[p][3][L] = qwb[0];  // wb of column j (qwb[k]: column j+k)
        lds[p][4][L] = qux[4];  // column j-4
        lds[p][5][L] = quz[3];
        lds[p][6][L] = gvx[4];
        lds[p][7][L] = gvz[3];
        __syncthreads();

        // ---- velocity of column j (k_eb_adj_v)
        float ngvx = o.gvx, ngvz = o.gvz, nlvx = o.lvx, nlvz = o.lvz, ux = 0.f, uz = 0.f;
        if (vlane && inu(j)) {
            const float* L0 = lds[p][0];
            const float* L1 = lds[p][1];
            const float* L2 = lds[p][2];
            const float* L3 = lds[p][3];
            const float br = 1.f / o.rho;
            if (UNDO) {
                const bool sx = srow && (unsigned)(S.isx - j + 3) < 8u;
                auto txx = [&](int k) { return (sx && j + k == S.isx) ? gxx[4 + k] + w : gxx[4 + k]; };
                auto tzz = [&](int k) { return (j == S.isx && L + k == lsrc) ? L1[L + k] + w : L1[L + k]; };
                ngvx += -(br * cax * ebp(txx) + br * caz * ebm([&](int k) { return L0[L + k]; }));
                ngvz += -(br * cax * ebm([&](int k) { return gxz[4 + k]; }) + br * caz * ebp(tzz));
            }
            nlvx = o.lvx + o.ix - (cax * ebp([&](int k) { return qwa[4 + k]; }) +
                                   caz * ebm([&](int k) { return L2[L + k]; }));
            nlvz = o.lvz + o.iz - (caz * ebp([&](int k) { return L3[L + k]; }) +
                                   cax * ebm([&](int k) { return qwc[4 + k]; }));
            ux = br * nlvx;
            uz = br * nlvz;
        }
        qux[8] = ux;
        quz[7] = uz;
        gvx[8] = ngvx;
        gvz[7] = ngvz;
        if (zown && j >= x0 && j < x1) {
            const uint32_t c = zo4 + (uint32_t)j * colb;
            st(a.Blm, c, nlvx);
            st(a.Blm + n, c, nlvz);
            if (UNDO) {
                st(a.Bbg, c, ngvx);
                st(a.Bbg + n, c, ngvz);
            }
        }

        // ---- stress of column cs = j - 4 (k_eb_adj_s)
        if (own_s) {
            const uint32_t c = cS;
            float ltxx = e_ltxx, ltzz = e_ltzz, ltxz = e_ltxz;
            float btxx = gxx[0], btzz = gzz[0], btxz = gxz[0];
            if (inu(cs)) {
                const float* L4 = lds[p][4];
                const float* L5 = lds[p][5];
                const float* L6 = lds[p][6];
                const float* L7 = lds[p][7];
                const float txx = ltxx - cax * ebm([&](int k) { return qux[4 + k]; });
                const float tzz = ltzz - caz * ebm([&](int k) { return L5[L + k]; });
                const float txz = ltxz - (caz * ebp([&](int k) { return L4[L + k]; }) +
                                          cax * ebp([&](int k) { return quz[3 + k]; }));
                ltxx = txx;
                ltzz = tzz;
                ltxz = txz;
                const float l = e_l, mu = e_mu, l2m = l + 2 * mu;
                const bool inside = zphys && cs >= S.pml && cs < S.pml + S.nx;
                float exx = 0.f, ezz = 0.f, exz = 0.f;
                if (UNDO) {
                    const float dvx = ebm([&](int k) { return gvx[4 + k]; }), dvz = ebm([&](int k) { return L7[L + k]; });
                    const float dvxz = ebp([&](int k) { return L6[L + k]; }), dvzx = ebp([&](int k) { return gvz[3 + k]; });
                    exx = dvx * idx;
                    ezz = dvz * idz;
                    exz = dvzx * idx + dvxz * idz;
                    btxx += -(cax * l2m * dvx + caz * l * dvz);
                    btzz += -(cax * l * dvx + caz * l2m * dvz);
                    btxz += -(mu * (cax * dvzx + caz * dvxz));
                } else if (inside) {
                    const size_t h =
                        (size_t)S.hist_off + ((size_t)it * S.nx + (cs - S.pml)) * S.nz + (size_t)(z - S.pml);
                    exx = a.hist[h];
                    ezz = a.hist[h + a.hstride];
                    exz = a.hist[h + 2 * a.hstride];
                }
                if (inside) {
                    float gpp = e_gpp, gps = e_gps;
                    gpp += -2 * l2m * (exx + ezz) * (txx + tzz);
                    gps += 4 * mu * (ezz * txx + exx * tzz) - 2 * mu * exz * txz;
                    st(a.gpp, c, gpp);
                    st(a.gps, c, gps);
                }
            }
            st(a.Blm + 2 * n, c, ltxx);
            st(a.Blm + 3 * n, c, ltzz);
            st(a.Blm + 4 * n, c, ltxz);
            if (UNDO) {
                st(a.Bbg + 2 * n, c, btxx);
                st(a.Bbg + 3 * n, c, btzz);
                st(a.Bbg + 4 * n, c, btxz);
            }
        }
    };

    const int jend = j0 + (x1 + 4 - j0 + 1) / 2 * 2;
    AdjCol cA = fetch(j0), cB = fetch(j0 + 1);
    for (int j = j0; j < jend; j += 2) {
        const bool more = j + 2 < jend;
        column(j, cA, 0);
        if (more) cA = fetch(j + 2);
        column(j + 1, cB, 1);
        if (more) cB = fetch(j + 3);
    }
}

bool eb_fused_threads_ok(int threads) { return threads == 256 || threads == 512; }

void launch_eb_fused(hipStream_t st, const BatchShape& b, const EbFusedArgs& a, const ChunkDesc* chunks,
                     int nchunks, const int* rmap, float w, float sign, int it, int hist_mode, int threads) {
    const int per_xcd = (nchunks + 7) / 8;
    const dim3 grid(per_xcd * 8);
    const float cax = b.dt / b.dx, caz = b.dt / b.dz, idx = 1.f / b.dx, idz = 1.f / b.dz;
    if (threads == 512)
        hipLaunchKernelGGL(k_eb_fused<512>, grid, dim3(512), 0, st, a, chunks, nchunks, per_xcd, rmap, b.slots, cax,
                           caz, idx, idz, w, sign, it, b.nt, hist_mode);
    else
        hipLaunchKernelGGL(k_eb_fused<256>, grid, dim3(256), 0, st, a, chunks, nchunks, per_xcd, rmap, b.slots, cax,
                           caz, idx, idz, w, sign, it, b.nt, hist_mode);
}

void launch_eb_adj_fused(hipStream_t st, const BatchShape& b, const EbAdjArgs& a, const ChunkDesc* chunks,
                         int nchunks, const int* rmap, float w, int it, int it_inj, bool undo, int threads) {
    const int per_xcd = (nchunks + 7) / 8;
    const dim3 grid(per_xcd * 8);
    const float cax = b.dt / b.dx, caz = b.dt / b.dz, idx = 1.f / b.dx, idz = 1.f / b.dz;
#define DS2D_EBA_LAUNCH(TT, UU)                                                                                \
    hipLaunchKernelGGL((k_eb_adj_fused<TT, UU>), grid, dim3(TT), 0, st, a, chunks, nchunks, per_xcd, rmap, b.slots, \
                       cax, caz, idx, idz, w, it, it_inj, b.nt)
    if (threads == 512) {
        if (undo)
            DS2D_EBA_LAUNCH(512, true);
        else
            DS2D_EBA_LAUNCH(512, false);
    } else {
        if (undo)
            DS2D_EBA_LAUNCH(256, true);
        else
            DS2D_EBA_LAUNCH(256, false);
    }
#undef DS2D_EBA_LAUNCH
}

}  // namespace ds2d
