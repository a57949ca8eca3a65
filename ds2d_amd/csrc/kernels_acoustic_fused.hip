// Fused plain-L2 acoustic Born and adjoint steps (gfx950), batched over shot
// windows, for the resident-history (store) mode of LSRTM.
//
// Same scheme and arithmetic as k_acou_v/k_acou_p (Born field only) and
// k_acou_inj/k_acou_adj_p/k_acou_adj_v (store mode) in kernels_acoustic.hip,
// i.e. the mu = 0 specialisation of born.cu:215-349 with the Born source
// 2 m dp_bg, but one launch per time step: a workgroup owns an x-chunk and a
// block of rows of one window (one row per lane, 8 halo rows each side) and
// marches along x, computing the velocity of column j and the pressure of
// column j-4 from register queues (x taps) and a double-buffered LDS column
// (z taps).  The march is unrolled by two so operands loaded two columns
// ahead are not copied before use.  State is ping-ponged (read A, write B).
#include "common.hpp"
#include "kernels.hpp"

namespace ds2d {

namespace {

constexpr int AHALO = 8;

__device__ __forceinline__ bool in_u_f(const SlotDesc& S, int ix, int iz) {
    return ix >= NOP && ix < S.nxp - NOP && iz >= NOP && iz < S.nzp - NOP;
}

// Background-increment history index of cell (ix, iz) at step it, or -1.
__device__ __forceinline__ long long hist_index(const SlotDesc& S, int ix, int iz, int it) {
    const int lx = ix - S.pml, lz = iz - S.pml;
    if (lx < 0 || lx >= S.nx || lz < 0 || lz >= S.nz) return -1;
    return S.hist_off + ((long long)it * S.nx + lx) * S.nz + lz;
}

// Lane geometry shared by both kernels.
struct Lanes {
    int L, z;
    bool zin, zown, vlane;
    uint32_t zo4, zc4, colb;
};
template <int T>
__device__ __forceinline__ Lanes lanes(const SlotDesc& S, const ChunkDesc& C) {
    Lanes g;
    g.L = threadIdx.x;
    g.z = C.z0 - AHALO + g.L;
    g.zin = g.z >= 0 && g.z < S.nzp;
    g.zown = g.zin && g.L >= AHALO && g.L < T - AHALO;
    g.vlane = g.L >= 4 && g.L < T - 4;
    g.colb = (uint32_t)S.nzp * 4u;
    g.zo4 = (uint32_t)(S.off + g.z) * 4u;
    g.zc4 = (uint32_t)(S.off + min(max(g.z, 0), S.nzp - 1)) * 4u;
    return g;
}

__device__ __forceinline__ float at(const float* p, uint32_t boff) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + boff);
}
__device__ __forceinline__ void put(float* p, uint32_t boff, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(p) + boff) = v;
}

// A copy in a register distinct from the source (early clobber).  A loaded
// operand that enters an x-tap queue stays live there for eight columns;
// copying it lets the next prefetch reuse the load register, so the loop
// latch never moves (and waits for) freshly loaded registers.
__device__ __forceinline__ float fresh(float x) {
    float r;
    asm("v_mov_b32 %0, %1" : "=&v"(r) : "v"(x));
    return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// Born step: v += (dt/d) b D+ p;  p += k (dt/dx D- vx + dt/dz D- vz) + 2 m dp_bg
// ---------------------------------------------------------------------------
template <int T>
__global__ __launch_bounds__(T) void k_acou_born_fused(AcouFusedArgs a, const ChunkDesc* __restrict__ chunks,
                                                       int nchunks, int per_xcd, const int* __restrict__ rmap,
                                                       const SlotDesc* __restrict__ slots, int nt, int it) {
    const int bid = blockIdx.x;
    const int ch = (bid % 8) * per_xcd + bid / 8;
    if (ch >= nchunks) return;
    const ChunkDesc C = chunks[ch];
    const SlotDesc S = slots[C.slot];
    const Lanes g = lanes<T>(S, C);
    const int L = g.L, z = g.z, nxp = S.nxp;
    const long long n = a.n;
    const float *Avx = a.A, *Avz = a.A + n, *Ap = a.A + 2 * n;
    float *Bvx = a.B, *Bvz = a.B + n, *Bp = a.B + 2 * n;
    const float ax = a.ax, az = a.az;
    __shared__ float lds[2][2][T];

    struct Col {
        float p4, vx, vz, bu, kk, mm, dp;
        int rec;  // trace recorded at column j - 4 (receiver row lane only; -1 none)
    };
    const int x0 = C.x0, x1 = C.x1, j0 = x0 - 4;
    auto ldc = [&](const float* p, int col) { return at(p, g.zc4 + (uint32_t)min(max(col, 0), nxp - 1) * g.colb); };
    auto fetch = [&](int j) {
        Col o;
        o.p4 = ldc(Ap, j + 4);
        o.vx = ldc(Avx, j);
        o.vz = ldc(Avz, j);
        o.bu = ldc(a.b, j);
        const int cs = j - 4;
        o.kk = ldc(a.k, cs);
        o.mm = ldc(a.m, cs);
        const long long h = (g.zin && cs >= 0 && cs < nxp) ? hist_index(S, cs, z, it) : -1;
        o.dp = h >= 0 ? a.hist[h] : 0.f;
        // prefetched with the operands: a load at the store would wait for
        // every prefetch issued before it
        o.rec = -1;
        if (a.rec && z == S.igz && cs >= x0 && cs < x1) o.rec = rmap[S.rmap_off + cs];
        return o;
    };
    const float c1 = DS2D_C1, c2 = DS2D_C2, c3 = DS2D_C3, c4 = DS2D_C4;
    float qp[9], qvx[9], qvz[5];
#pragma unroll
    for (int k = 0; k < 8; ++k) qp[k + 1] = ldc(Ap, j0 - 4 + k);
#pragma unroll
    for (int k = 0; k < 9; ++k) qvx[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 5; ++k) qvz[k] = 0.f;

    auto column = [&](const int j, const Col& o, const int p) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            qp[k] = qp[k + 1];
            qvx[k] = qvx[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) qvz[k] = qvz[k + 1];
        qp[8] = fresh(o.p4);
        lds[p][0][L] = qp[4];   // p of column j
        lds[p][1][L] = qvz[0];  // new vz of column j - 4
        __syncthreads();
        float nvx = o.vx, nvz = o.vz;
        if (g.vlane && in_u_f(S, j, z)) {
            const float* L0 = lds[p][0];
            const float dx = c1 * (qp[5] - qp[4]) + c2 * (qp[6] - qp[3]) + c3 * (qp[7] - qp[2]) + c4 * (qp[8] - qp[1]);
            const float dz = c1 * (L0[L + 1] - L0[L]) + c2 * (L0[L + 2] - L0[L - 1]) + c3 * (L0[L + 3] - L0[L - 2]) +
                             c4 * (L0[L + 4] - L0[L - 3]);
            nvx = o.vx + ax * o.bu * dx;
            nvz = o.vz + az * o.bu * dz;
        }
        qvx[8] = nvx;
        qvz[4] = nvz;
        if (g.zown && j >= x0 && j < x1) {
            const uint32_t c = g.zo4 + (uint32_t)j * g.colb;
            put(Bvx, c, nvx);
            put(Bvz, c, nvz);
        }
        const int cs = j - 4;
        if (g.zown && cs >= x0 && cs < x1) {
            float pn = qp[0];
            if (in_u_f(S, cs, z)) {
                const float* L1 = lds[p][1];
                const float dvx = c1 * (qvx[4] - qvx[3]) + c2 * (qvx[5] - qvx[2]) + c3 * (qvx[6] - qvx[1]) +
                                  c4 * (qvx[7] - qvx[0]);
                const float dvz = c1 * (L1[L] - L1[L - 1]) + c2 * (L1[L + 1] - L1[L - 2]) +
                                  c3 * (L1[L + 2] - L1[L - 3]) + c4 * (L1[L + 3] - L1[L - 4]);
                const float bd = o.kk * (ax * dvx + az * dvz);
                pn = pn + (bd + 2 * o.mm * o.dp);
            }
            put(Bp, g.zo4 + (uint32_t)cs * g.colb, pn);
            if (o.rec >= 0) a.rec[S.rec_off + (long long)o.rec * nt + it] = pn;
        }
    };
    const int jend = j0 + ((x1 + 4 - j0 + 1) & ~1);
    Col cA = fetch(j0), cB = fetch(j0 + 1);
    for (int j = j0; j < jend; j += 2) {
        column(j, cA, 0);
        cA = fetch(j + 2);
        column(j + 1, cB, 1);
        cB = fetch(j + 3);
    }
}

// ---------------------------------------------------------------------------
// Adjoint step (store mode), time reversed:
//   lp += res at the receivers;  grad += 2 dp_bg lp;
//   lv += -(dt/d) D+ (k lp);  lp += -(dt/d) D- (b lv)   (k lp, b lv masked to the update region)
// ---------------------------------------------------------------------------
template <int T>
__global__ __launch_bounds__(T) void k_acou_adj_fused(AcouFusedArgs a, const ChunkDesc* __restrict__ chunks,
                                                      int nchunks, int per_xcd, const int* __restrict__ rmap,
                                                      const SlotDesc* __restrict__ slots, int nt, int it) {
    const int bid = blockIdx.x;
    const int ch = (bid % 8) * per_xcd + bid / 8;
    if (ch >= nchunks) return;
    const ChunkDesc C = chunks[ch];
    const SlotDesc S = slots[C.slot];
    const Lanes g = lanes<T>(S, C);
    const int L = g.L, z = g.z, nxp = S.nxp;
    const long long n = a.n;
    const float *Avx = a.A, *Avz = a.A + n, *Ap = a.A + 2 * n;
    float *Bvx = a.B, *Bvz = a.B + n, *Bp = a.B + 2 * n;
    const float ax = a.ax, az = a.az;
    __shared__ float lds[2][2][T];

    struct Col {
        float lp4, k4, lvx, lvz, bu, dp, gr;
    };
    auto ldc = [&](const float* p, int col) { return at(p, g.zc4 + (uint32_t)min(max(col, 0), nxp - 1) * g.colb); };
    auto fetch = [&](int j) {
        Col o;
        const int jl = j + 4;
        o.lp4 = ldc(Ap, jl);
        o.k4 = ldc(a.k, jl);
        o.lvx = ldc(Avx, j);
        o.lvz = ldc(Avz, j);
        o.bu = ldc(a.b, j);
        const int cs = j - 4;
        const long long h = (g.zin && cs >= 0 && cs < nxp) ? hist_index(S, cs, z, it) : -1;
        o.dp = h >= 0 ? a.hist[h] : 0.f;
        o.gr = ldc(a.grad, cs);
        return o;
    };
    const float c1 = DS2D_C1, c2 = DS2D_C2, c3 = DS2D_C3, c4 = DS2D_C4;
    float qw[9], qlp[9], qux[9], quz[5];
    const int x0 = C.x0, x1 = C.x1, j0 = x0 - 4;
    const int jend = j0 + ((x1 + 4 - j0 + 1) & ~1);
    // Residual of this step at the receivers of columns [j0 - 4, jend + 4),
    // staged once per launch (the host keeps chunks at most ACOU_FUSED_MAX_W
    // columns wide): the receiver row then needs no dependent load per column.
    __shared__ float resv[ACOU_FUSED_MAX_W + 24];
    __shared__ int resk[ACOU_FUSED_MAX_W + 24];
    const int cbase = j0 - 4, ncols = jend + 4 - cbase;
    for (int q = L; q < ncols; q += T) {
        const int col = cbase + q;
        int r = -1;
        float v = 0.f;
        if (col >= 0 && col < nxp) {
            r = rmap[S.rmap_off + col];
            if (r >= 0) v = a.res[S.rec_off + (long long)r * nt + it];
        }
        resk[q] = r;
        resv[q] = v;
    }
    __syncthreads();
    // lp of column col after the receiver injection of this step
    const bool rrow = z == S.igz;
    auto inject = [&](float v, int col) { return (rrow && resk[col - cbase] >= 0) ? v + resv[col - cbase] : v; };
    // prefill: issue the 16 loads first (clamped, unconditional), then combine
    float plp[8], pk[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        plp[k] = ldc(Ap, j0 - 4 + k);
        pk[k] = ldc(a.k, j0 - 4 + k);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int col = j0 - 4 + k;
        const float lp = inject(plp[k], col);
        qw[k + 1] = in_u_f(S, col, z) ? pk[k] * lp : 0.f;
        qlp[k + 1] = lp;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) qux[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 5; ++k) quz[k] = 0.f;

    auto column = [&](const int j, const Col& o, const int p) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            qw[k] = qw[k + 1];
            qlp[k] = qlp[k + 1];
            qux[k] = qux[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) quz[k] = quz[k + 1];
        const int jl = j + 4;
        const float lp4 = fresh(inject(o.lp4, jl));
        qw[8] = in_u_f(S, jl, z) ? o.k4 * lp4 : 0.f;
        qlp[8] = lp4;
        lds[p][0][L] = qw[4];   // k lp of column j
        lds[p][1][L] = quz[0];  // b lvz of column j - 4
        __syncthreads();
        float nvx = o.lvx, nvz = o.lvz;
        const bool uj = in_u_f(S, j, z);
        if (g.vlane && uj) {
            const float* L0 = lds[p][0];
            const float gx = c1 * (qw[5] - qw[4]) + c2 * (qw[6] - qw[3]) + c3 * (qw[7] - qw[2]) + c4 * (qw[8] - qw[1]);
            const float gz = c1 * (L0[L + 1] - L0[L]) + c2 * (L0[L + 2] - L0[L - 1]) + c3 * (L0[L + 3] - L0[L - 2]) +
                             c4 * (L0[L + 4] - L0[L - 3]);
            nvx = o.lvx + -(ax * gx);
            nvz = o.lvz + -(az * gz);
        }
        qux[8] = uj ? o.bu * nvx : 0.f;
        quz[4] = uj ? o.bu * nvz : 0.f;
        if (g.zown && j >= x0 && j < x1) {
            const uint32_t c = g.zo4 + (uint32_t)j * g.colb;
            put(Bvx, c, nvx);
            put(Bvz, c, nvz);
        }
        const int cs = j - 4;
        if (g.zown && cs >= x0 && cs < x1) {
            const float lp = qlp[0];
            float np = lp;
            float gr = o.gr;
            if (in_u_f(S, cs, z)) {
                const float* L1 = lds[p][1];
                const float gx = c1 * (qux[4] - qux[3]) + c2 * (qux[5] - qux[2]) + c3 * (qux[6] - qux[1]) +
                                 c4 * (qux[7] - qux[0]);
                const float gz = c1 * (L1[L] - L1[L - 1]) + c2 * (L1[L + 1] - L1[L - 2]) +
                                 c3 * (L1[L + 2] - L1[L - 3]) + c4 * (L1[L + 3] - L1[L - 4]);
                gr = gr + 2 * o.dp * lp;
                np = lp + -(ax * gx + az * gz);
            }
            const uint32_t c = g.zo4 + (uint32_t)cs * g.colb;
            put(Bp, c, np);
            put(a.grad, c, gr);
        }
    };
    Col cA = fetch(j0), cB = fetch(j0 + 1);
    for (int j = j0; j < jend; j += 2) {
        column(j, cA, 0);
        cA = fetch(j + 2);
        column(j + 1, cB, 1);
        cB = fetch(j + 3);
    }
}

bool acou_fused_threads_ok(int threads) { return threads == 256 || threads == 512; }

void launch_acou_born_fused(hipStream_t st, const BatchShape& b, const AcouFusedArgs& a, const ChunkDesc* chunks,
                            int nchunks, const int* rmap, int it, int threads) {
    const int per_xcd = (nchunks + 7) / 8;
    const dim3 grid(per_xcd * 8);
    if (threads == 512)
        hipLaunchKernelGGL(k_acou_born_fused<512>, grid, dim3(512), 0, st, a, chunks, nchunks, per_xcd, rmap,
                           b.slots, b.nt, it);
    else
        hipLaunchKernelGGL(k_acou_born_fused<256>, grid, dim3(256), 0, st, a, chunks, nchunks, per_xcd, rmap,
                           b.slots, b.nt, it);
}

void launch_acou_adj_fused(hipStream_t st, const BatchShape& b, const AcouFusedArgs& a, const ChunkDesc* chunks,
                           int nchunks, const int* rmap, int it, int threads) {
    const int per_xcd = (nchunks + 7) / 8;
    const dim3 grid(per_xcd * 8);
    if (threads == 512)
        hipLaunchKernelGGL(k_acou_adj_fused<512>, grid, dim3(512), 0, st, a, chunks, nchunks, per_xcd, rmap,
                           b.slots, b.nt, it);
    else
        hipLaunchKernelGGL(k_acou_adj_fused<256>, grid, dim3(256), 0, st, a, chunks, nchunks, per_xcd, rmap,
                           b.slots, b.nt, it);
}

}  // namespace ds2d
