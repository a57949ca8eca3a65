// Fused elastic time step for gfx950: velocity and stress of one time step in
// a single launch, batched over uneven shot windows.
//
// Scheme: born.cu:215-349 (update_vx/vz then update_txx_tzz/txz) with the
// transient explosive source of FD2DGPU.cu:2039-2049 and the surface
// recorder of FD2DGPU.cu:355-365; split PML memory only in the damping
// strips (elsewhere the split update equals the unsplit one), one split per
// component (the other is the unsplit value minus it).
//
// Layout of the work: one workgroup owns columns [x0, x1) and a block of
// rows of one shot window, one z row per lane (rows contiguous in memory, so
// every column load is a coalesced 64-lane row).  Eight halo rows above and
// below the block are loaded too: velocity is computed on the block plus
// four rows each side, which is what the stress z taps of the block read.
// Several small workgroups per CU keep loads in flight while others wait at
// their barrier.  The workgroup marches along x:
// at column j it computes the new velocity of column j and the new stress of
// column j-4.  The eight x taps of every derivative live in per-lane
// register queues; the eight z taps are read from a double-buffered LDS copy
// of the current column (one barrier per column).  The march is unrolled by
// the prefetch depth PF (2 or 4) with one operand set per phase: the operands
// of column j + PF are loaded right after column j is consumed, into the
// registers column j just freed, so no loaded value is copied before use and
// HBM latency overlaps PF column iterations.  The eight operands every cell needs are loaded
// unconditionally from clamped addresses (values of clamped cells are never
// used), which keeps those loads free of branches.  The state is ping-ponged
// (read A, write B), so the four halo columns of velocity on each side are
// recomputed without races.
#include "common.hpp"
#include "kernels.hpp"

namespace ds2d {

namespace {

__device__ __forceinline__ bool damp_cell_f(const SlotDesc& S, int ix, int iz) {
    return S.boundary == 0 && (ix < S.pml || ix > S.nxp - S.pml || iz < S.pml || iz > S.nzp - S.pml);
}

// Operands of one column for one lane.
struct Col {
    float txx4, txz4;            // stress of column j+4 (enters the x-tap queues)
    float tzz, vx, vz, bx;       // stress tzz, old velocity, dt/rho of column j
    float lam, mu;               // moduli of column j-4
    float dxv, dzv;              // damping of column j (velocity, PML lanes only)
    float vxx, vzz;              // old vx_x, vz_z of column j (PML lanes only)
    float dxs, dzs;              // damping of column j-4 (stress, PML lanes only)
    float sxx, szz, sxz;         // old txx_x, tzz_z, txz_x of column j-4 (PML lanes only)
    int rec;                     // trace recorded at column j (receiver row lane only; -1 none)
};

// A copy in a register distinct from the source (early clobber).  Operands
// that enter the x-tap queues stay live there for eight more columns; when
// the queue held the loaded register itself, the next prefetch could not
// reuse it and the loop latch had to move freshly loaded registers, which
// waited for every load of the pass (the prefetch depth collapsed to zero).
__device__ __forceinline__ float fresh(float x) {
    float r;
    asm("v_mov_b32 %0, %1" : "=&v"(r) : "v"(x));
    return r;
}
}  // namespace

template <int T, int PF, bool PML>
__global__ __launch_bounds__(T) void k_fused_step(FusedArgs a, const ChunkDesc* __restrict__ chunks, int nchunks,
                                                  int per_xcd, const int* __restrict__ rmap,
                                                  const SlotDesc* __restrict__ slots, float idx, float idz, float dt,
                                                  int nt, int it) {
    // XCD-aware chunk order: consecutive workgroups land on different XCDs,
    // so give each XCD a contiguous run of chunks (neighbours share halos in L2).
    const int b = blockIdx.x;
    const int ch = (b % 8) * per_xcd + b / 8;
    if (ch >= nchunks) return;
    const ChunkDesc C = chunks[ch];
    const SlotDesc S = slots[C.slot];
    // Lane -> local row of the block (LDS index).  With damping strips the
    // rows of the top and bottom strips (and those outside the window) go to
    // the first lanes, so one wave instead of several takes the damped-row
    // branch at every column; the other waves hold contiguous interior rows.
    // The map is a bijection on [0, T): every row still has exactly one lane.
    const int zb = C.z0 - FUSED_HALO;
    int rr = threadIdx.x;
    if (PML && S.boundary == 0) {
        const int a0 = min(max(S.pml - zb, 0), T);
        const int b0 = min(max(S.nzp - S.pml + 1 - zb, a0), T);
        const int nb = T - b0;
        const int l = threadIdx.x;
        rr = l < a0 ? l : (l < a0 + nb ? b0 + (l - a0) : a0 + (l - a0 - nb));
    }
    const int L = rr;                          // local row = LDS index
    const int z = zb + L;                      // padded row of this lane
    const int nzp = S.nzp, nxp = S.nxp;
    const bool zin = z >= 0 && z < nzp;
    const bool zown = zin && L >= FUSED_HALO && L < T - FUSED_HALO;  // rows this workgroup stores
    const bool vlane = L >= 4 && L < T - 4;    // velocity needed (z taps of owned stress rows)
    const bool zupd = z >= NOP && z < nzp - NOP;
    const int x0 = C.x0, x1 = C.x1;

    __shared__ float lds[2][4][T];

    // Every field of the batch is below 4 GiB, so addresses are a uniform
    // field base plus a 32-bit per-lane byte offset (one VALU add per access).
    const long long n = a.n;
    const float* Avx = a.A;
    const float* Avz = a.A + n;
    const float* Atxx = a.A + 2 * n;
    const float* Atzz = a.A + 3 * n;
    const float* Atxz = a.A + 4 * n;
    const float* Pbx = a.par;
    const float* Plam = a.par + n;
    const float* Pmu = a.par + 2 * n;
    const uint32_t colb = (uint32_t)nzp * 4u;
    const uint32_t zo4 = (uint32_t)(S.off + z) * 4u;                           // lanes with zin only
    const uint32_t zc4 = (uint32_t)(S.off + min(max(z, 0), nzp - 1)) * 4u;   // clamped row
    auto at = [](const float* p, uint32_t boff) -> float {
        return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + boff);
    };
    auto ld = [&](const float* p, int col) -> float {
        return (zin && col >= 0 && col < nxp) ? at(p, zo4 + (uint32_t)col * colb) : 0.f;
    };
    // clamped, unconditional load (the value of a clamped cell is never used)
    auto ldc = [&](const float* p, int col) -> float {
        return at(p, zc4 + (uint32_t)min(max(col, 0), nxp - 1) * colb);
    };
    // Split-field bases are rebuilt inside the damping-strip branches from an
    // opaque copy of n, so the compiler does not keep twenty loop-invariant
    // pointers live in scalar registers for the whole march.
    auto opaque_n = [&]() {
        long long v = n;
        asm volatile("" : "+s"(v));
        return v;
    };
    auto split = [&](const float* base, int k) { return base + (5 + k) * opaque_n(); };
    auto splitB = [&](int k) { return a.B + (5 + k) * opaque_n(); };
    // operands of column j (velocity side) and j-4 (stress side)
    auto fetch = [&](int j) {
        Col o;
        o.txx4 = ldc(Atxx, j + 4);
        o.txz4 = ldc(Atxz, j + 4);
        o.tzz = ldc(Atzz, j);
        o.vx = ldc(Avx, j);
        o.vz = ldc(Avz, j);
        const bool vupd = zupd && j >= NOP && j < nxp - NOP;
        o.bx = ldc(Pbx, j);
        const int cs = j - 4;
        const bool supd = zupd && cs >= NOP && cs < nxp - NOP;
        o.lam = ldc(Plam, cs);
        o.mu = ldc(Pmu, cs);
        // damping-strip operands: one branch per side (vupd/supd imply an in-window cell)
        // One split of each component is kept (vx_x, vz_z, txx_x, tzz_z, txz_x);
        // its partner is the unsplit value minus it.
        o.dxv = o.dzv = o.vxx = o.vzz = 0.f;
        if (PML && vupd && damp_cell_f(S, j, z)) {
            const uint32_t c = zo4 + (uint32_t)j * colb;
            o.dxv = at(a.dampx, c);
            o.dzv = at(a.dampz, c);
            o.vxx = at(split(a.A, 0), c);
            o.vzz = at(split(a.A, 3), c);
        }
        // receiver map of column j, prefetched like the operands (a load at
        // the store would wait for every prefetch issued before it)
        o.rec = -1;
        if (z == S.igz && j >= x0 && j < x1) o.rec = rmap[S.rmap_off + j];
        o.dxs = o.dzs = o.sxx = o.szz = o.sxz = 0.f;
        if (PML && supd && damp_cell_f(S, cs, z)) {
            const uint32_t c = zo4 + (uint32_t)cs * colb;
            o.dxs = at(a.dampx, c);
            o.dzs = at(a.dampz, c);
            o.sxx = at(split(a.A, 4), c);
            o.szz = at(split(a.A, 7), c);
            o.sxz = at(split(a.A, 8), c);
        }
        return o;
    };
    auto st = [](float* p, uint32_t boff, float v) {
        *reinterpret_cast<float*>(reinterpret_cast<char*>(p) + boff) = v;
    };

    const float c1 = DS2D_C1, c2 = DS2D_C2, c3 = DS2D_C3, c4 = DS2D_C4;
    float qxx[9], qxz[9], qzz[5], qvx[9], qvz[8];
    const int j0 = x0 - 4;
    // prefill: qxx/qxz[1..8] = old stress of columns j0-4 .. j0+3; qzz[1..4] = j0-4 .. j0-1
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        qxx[k + 1] = ld(Atxx, j0 - 4 + k);
        qxz[k + 1] = ld(Atxz, j0 - 4 + k);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) qzz[k + 1] = ld(Atzz, j0 - 4 + k);
#pragma unroll
    for (int k = 0; k < 9; ++k) qvx[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) qvz[k] = 0.f;

    const float w = a.wav[it];
    // coefficient of the tap at offset d = (source - cell) of a forward
    // half-point difference: +c_d for d = 1..4, -c_(1-d) for d = 0..-3
    auto tap = [&](int d) -> float {
        return d == 1 ? c1 : d == 2 ? c2 : d == 3 ? c3 : d == 4 ? c4 :
               d == 0 ? -c1 : d == -1 ? -c2 : d == -2 ? -c3 : d == -3 ? -c4 : 0.f;
    };
    // z-tap weight of this lane's row in the source column (once per lane)
    const float wz = w * tap(S.isz - z);

    // One column: velocity of column j and stress of column j - 4, from the
    // operands o of column j; p selects the LDS buffer of this phase.
    auto column = [&](const int j, const Col& o, const int p) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            qxx[k] = qxx[k + 1];
            qxz[k] = qxz[k + 1];
            qvx[k] = qvx[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) qzz[k] = qzz[k + 1];
#pragma unroll
        for (int k = 0; k < 7; ++k) qvz[k] = qvz[k + 1];
        qxx[8] = fresh(o.txx4);
        qxz[8] = fresh(o.txz4);
        qzz[4] = fresh(o.tzz);
        // ---- z taps through LDS: txz, tzz of column j; new vx, vz of column j-4
        lds[p][0][L] = qxz[4];
        lds[p][1][L] = qzz[4];
        lds[p][2][L] = qvx[4];
        lds[p][3][L] = qvz[3];
        __syncthreads();
        // ---- velocity of column j (halo columns and rows included)
        float nvx = o.vx, nvz = o.vz;
        if (vlane && zupd && j >= NOP && j < nxp - NOP) {
            const float* L0 = lds[p][0];
            const float* L1 = lds[p][1];
            float ax = c1 * (qxx[5] - qxx[4]) + c2 * (qxx[6] - qxx[3]) + c3 * (qxx[7] - qxx[2]) + c4 * (qxx[8] - qxx[1]);
            float ex = c1 * (qxz[4] - qxz[3]) + c2 * (qxz[5] - qxz[2]) + c3 * (qxz[6] - qxz[1]) + c4 * (qxz[7] - qxz[0]);
            float bz = c1 * (L0[L] - L0[L - 1]) + c2 * (L0[L + 1] - L0[L - 2]) + c3 * (L0[L + 2] - L0[L - 3]) +
                       c4 * (L0[L + 3] - L0[L - 4]);
            float cz = c1 * (L1[L + 1] - L1[L]) + c2 * (L1[L + 2] - L1[L - 1]) + c3 * (L1[L + 3] - L1[L - 2]) +
                       c4 * (L1[L + 4] - L1[L - 3]);
            // the source reaches the x taps of columns isx-4 .. isx+3 (a uniform test)
            if ((unsigned)(S.isx - j + 3) < 8u && z == S.isz) ax += w * tap(S.isx - j);
            if (j == S.isx) cz += wz;
            const float bx = o.bx;
            // A wave holding a damped cell runs the split update on all its lanes (in the
            // undamped ones the damping and the loaded split are zero, so it is the unsplit
            // update): the damping waves set the pace at every column's barrier.
            const bool dv = PML && damp_cell_f(S, j, z);
            if (PML && __any(dv)) {
                // one hardware reciprocal (1 ulp) per direction instead of four IEEE divisions
                const float d1 = 1.f - dt * o.dxv / 2, r2 = __builtin_amdgcn_rcpf(1.f + dt * o.dxv / 2);
                const float d3 = 1.f - dt * o.dzv / 2, r4 = __builtin_amdgcn_rcpf(1.f + dt * o.dzv / 2);
                const float vxx = (d1 * o.vxx + bx * idx * ax) * r2;
                const float vxz = (d3 * (o.vx - o.vxx) + bx * idz * bz) * r4;
                const float vzx = (d1 * (o.vz - o.vzz) + bx * idx * ex) * r2;
                const float vzz = (d3 * o.vzz + bx * idz * cz) * r4;
                if (dv && zown && j >= x0 && j < x1) {
                    const uint32_t c = zo4 + (uint32_t)j * colb;
                    st(splitB(0), c, vxx);
                    st(splitB(3), c, vzz);
                }
                nvx = vxx + vxz;
                nvz = vzx + vzz;
            } else {
                nvx = o.vx + bx * (ax * idx + bz * idz);
                nvz = o.vz + bx * (ex * idx + cz * idz);
            }
        }
        qvx[8] = nvx;
        qvz[7] = nvz;
        if (zown && j >= x0 && j < x1) {
            const uint32_t c = zo4 + (uint32_t)j * colb;
            st(a.B, c, nvx);
            st(a.B + n, c, nvz);
            if (o.rec >= 0) {
                const size_t q = (size_t)S.rec_off + (size_t)o.rec * nt + it;
                a.rec_z[q] = nvz;
                a.rec_x[q] = nvx;
            }
        }
        // ---- stress of column cs = j - 4 (owned columns only)
        const int cs = j - 4;
        if (zown && cs >= x0 && cs < x1) {
            const uint32_t c = zo4 + (uint32_t)cs * colb;
            float txx = qxx[0], tzz = qzz[0], txz = qxz[0];
            if (zupd && cs >= NOP && cs < nxp - NOP) {
                const float* L2 = lds[p][2];
                const float* L3 = lds[p][3];
                const float dvx = (c1 * (qvx[4] - qvx[3]) + c2 * (qvx[5] - qvx[2]) + c3 * (qvx[6] - qvx[1]) +
                                   c4 * (qvx[7] - qvx[0])) * idx;
                const float dvzx = (c1 * (qvz[4] - qvz[3]) + c2 * (qvz[5] - qvz[2]) + c3 * (qvz[6] - qvz[1]) +
                                    c4 * (qvz[7] - qvz[0])) * idx;
                const float dvz = (c1 * (L3[L] - L3[L - 1]) + c2 * (L3[L + 1] - L3[L - 2]) +
                                   c3 * (L3[L + 2] - L3[L - 3]) + c4 * (L3[L + 3] - L3[L - 4])) * idz;
                const float dvxz = (c1 * (L2[L + 1] - L2[L]) + c2 * (L2[L + 2] - L2[L - 1]) +
                                    c3 * (L2[L + 3] - L2[L - 2]) + c4 * (L2[L + 4] - L2[L - 3])) * idz;
                const float l = o.lam, mu = o.mu, l2m = l + 2 * mu;
                const bool ds = PML && damp_cell_f(S, cs, z);
                if (PML && __any(ds)) {
                    const float d1 = 1.f - dt * o.dxs / 2, r2 = __builtin_amdgcn_rcpf(1.f + dt * o.dxs / 2);
                    const float d3 = 1.f - dt * o.dzs / 2, r4 = __builtin_amdgcn_rcpf(1.f + dt * o.dzs / 2);
                    const float a1 = (d1 * o.sxx + l2m * dvx) * r2;
                    const float a2 = (d3 * (txx - o.sxx) + l * dvz) * r4;
                    const float b1 = (d1 * (tzz - o.szz) + l * dvx) * r2;
                    const float b2 = (d3 * o.szz + l2m * dvz) * r4;
                    const float e1 = (d1 * o.sxz + mu * dvzx) * r2;
                    const float e2 = (d3 * (txz - o.sxz) + mu * dvxz) * r4;
                    if (ds) {
                        st(splitB(4), c, a1);
                        st(splitB(7), c, b2);
                        st(splitB(8), c, e1);
                    }
                    txx = a1 + a2;
                    tzz = b1 + b2;
                    txz = e1 + e2;
                } else {
                    txx = txx + (l2m * dvx + l * dvz);
                    tzz = tzz + (l * dvx + l2m * dvz);
                    txz = txz + mu * (dvzx + dvxz);
                }
            }
            st(a.B + 2 * n, c, txx);
            st(a.B + 3 * n, c, tzz);
            st(a.B + 4 * n, c, txz);
        }
    };

    // a multiple of PF columns: up to PF-1 extra trailing columns store nothing
    const int jend = j0 + (x1 + 4 - j0 + PF - 1) / PF * PF;
    // Without damping strips the last pass prefetches nothing (a uniform
    // branch; with them, the unguarded march measured faster).
    auto more = [&](int j) { return PML || j + PF < jend; };
    if (PF == 2) {
        Col cA = fetch(j0), cB = fetch(j0 + 1);
        for (int j = j0; j < jend; j += 2) {
            const bool m = more(j);
            column(j, cA, 0);
            if (m) cA = fetch(j + 2);
            column(j + 1, cB, 1);
            if (m) cB = fetch(j + 3);
        }
    } else {
        Col cA = fetch(j0), cB = fetch(j0 + 1), cC = fetch(j0 + 2), cD = fetch(j0 + 3);
        for (int j = j0; j < jend; j += 4) {
            const bool m = more(j);
            column(j, cA, 0);
            if (m) cA = fetch(j + 4);
            column(j + 1, cB, 1);
            if (m) cB = fetch(j + 5);
            column(j + 2, cC, 0);
            if (m) cC = fetch(j + 6);
            column(j + 3, cD, 1);
            if (m) cD = fetch(j + 7);
        }
    }
}

// 704 lanes own 688 rows: one block covers a 600-row model with a 20-cell
// frame (640 padded rows) with one wave fewer than 768.
bool fused_threads_ok(int threads) { return threads == 256 || threads == 512 || threads == 704 || threads == 768; }

void launch_fused_step(hipStream_t st, const BatchShape& b, const FusedArgs& a, const ChunkDesc* chunks,
                       int nchunks, const int* rmap, int it, int threads, int depth, bool pml) {
    const int per_xcd = (nchunks + 7) / 8;
    const dim3 grid(per_xcd * 8);
    const float idx = 1.f / b.dx, idz = 1.f / b.dz;
#define DS2D_FUSED_LAUNCH(TT, PP, MM)                                                                             \
    hipLaunchKernelGGL((k_fused_step<TT, PP, MM>), grid, dim3(TT), 0, st, a, chunks, nchunks, per_xcd, rmap, b.slots, \
                       idx, idz, b.dt, b.nt, it)
#define DS2D_FUSED_THREADS(PP, MM)       \
    if (threads == 256)                  \
        DS2D_FUSED_LAUNCH(256, PP, MM);  \
    else if (threads == 512)             \
        DS2D_FUSED_LAUNCH(512, PP, MM);  \
    else if (threads == 704)             \
        DS2D_FUSED_LAUNCH(704, PP, MM);  \
    else                                 \
        DS2D_FUSED_LAUNCH(768, PP, MM);
    if (pml) {
        if (depth == 4) {
            DS2D_FUSED_THREADS(4, true)
        } else {
            DS2D_FUSED_THREADS(2, true)
        }
    } else {
        if (depth == 4) {
            DS2D_FUSED_THREADS(4, false)
        } else {
            DS2D_FUSED_THREADS(2, false)
        }
    }
#undef DS2D_FUSED_THREADS
#undef DS2D_FUSED_LAUNCH
}

}  // namespace ds2d
