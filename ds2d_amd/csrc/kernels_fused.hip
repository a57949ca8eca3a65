// Fused elastic time step for gfx950: velocity and stress of one time step in
// a single launch, batched over uneven shot windows.
//
// Scheme: born.cu:215-349 (update_vx/vz then update_txx_tzz/txz) with the
// transient explosive source of FD2DGPU.cu:2039-2049 and the surface
// recorder of FD2DGPU.cu:355-365; split PML memory only in the damping
// strips (elsewhere the split update equals the unsplit one).
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
// two with one operand set per phase: the operands of column j + 2 are
// loaded right after column j is consumed, into the registers column j just
// freed, so no loaded value is copied before use and HBM latency overlaps
// two column iterations.  The eight operands every cell needs are loaded
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
    float vxx, vxz, vzx, vzz;    // old velocity splits of column j (PML lanes only)
    float dxs, dzs;              // damping of column j-4 (stress, PML lanes only)
    float sxx, sxz, szx, szz, sex, sez;  // old stress splits of column j-4 (PML lanes only)
};
}  // namespace

template <int T>
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
    const int L = threadIdx.x;                 // lane = LDS index
    const int z = C.z0 - FUSED_HALO + L;       // padded row of this lane
    const int nzp = S.nzp, nxp = S.nxp;
    const bool zin = z >= 0 && z < nzp;
    const bool zown = zin && L >= FUSED_HALO && L < T - FUSED_HALO;  // rows this workgroup stores
    const bool vlane = L >= 4 && L < T - 4;    // velocity needed (z taps of owned stress rows)
    const bool zupd = z >= NOP && z < nzp - NOP;
    const size_t base = (size_t)S.off + z;

    __shared__ float lds[2][4][T];

    auto ld = [&](const float* p, int col) -> float {
        return (zin && col >= 0 && col < nxp) ? p[base + (size_t)col * nzp] : 0.f;
    };
    // clamped, unconditional load (the value of a clamped cell is never used)
    const size_t basec = (size_t)S.off + (size_t)min(max(z, 0), nzp - 1);
    auto ldc = [&](const float* p, int col) -> float {
        return p[basec + (size_t)min(max(col, 0), nxp - 1) * nzp];
    };
    // operands of column j (velocity side) and j-4 (stress side)
    auto fetch = [&](int j) {
        Col o;
        o.txx4 = ldc(a.txx_a, j + 4);
        o.txz4 = ldc(a.txz_a, j + 4);
        o.tzz = ldc(a.tzz_a, j);
        o.vx = ldc(a.vx_a, j);
        o.vz = ldc(a.vz_a, j);
        const bool vupd = zupd && j >= NOP && j < nxp - NOP;
        o.bx = ldc(a.bx, j);
        const int cs = j - 4;
        const bool supd = zupd && cs >= NOP && cs < nxp - NOP;
        o.lam = ldc(a.lam, cs);
        o.mu = ldc(a.mu, cs);
        const bool vpml = vupd && damp_cell_f(S, j, z);
        o.dxv = vpml ? ld(a.dampx, j) : 0.f;
        o.dzv = vpml ? ld(a.dampz, j) : 0.f;
        o.vxx = vpml ? ld(a.vx_x_a, j) : 0.f;
        o.vxz = vpml ? ld(a.vx_z_a, j) : 0.f;
        o.vzx = vpml ? ld(a.vz_x_a, j) : 0.f;
        o.vzz = vpml ? ld(a.vz_z_a, j) : 0.f;
        const bool spml = supd && damp_cell_f(S, cs, z);
        o.dxs = spml ? ld(a.dampx, cs) : 0.f;
        o.dzs = spml ? ld(a.dampz, cs) : 0.f;
        o.sxx = spml ? ld(a.txx_x_a, cs) : 0.f;
        o.sxz = spml ? ld(a.txx_z_a, cs) : 0.f;
        o.szx = spml ? ld(a.tzz_x_a, cs) : 0.f;
        o.szz = spml ? ld(a.tzz_z_a, cs) : 0.f;
        o.sex = spml ? ld(a.txz_x_a, cs) : 0.f;
        o.sez = spml ? ld(a.txz_z_a, cs) : 0.f;
        return o;
    };

    const float c1 = DS2D_C1, c2 = DS2D_C2, c3 = DS2D_C3, c4 = DS2D_C4;
    float qxx[9], qxz[9], qzz[5], qvx[9], qvz[8];
    const int x0 = C.x0, x1 = C.x1;
    const int j0 = x0 - 4;
    // prefill: qxx/qxz[1..8] = old stress of columns j0-4 .. j0+3; qzz[1..4] = j0-4 .. j0-1
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        qxx[k + 1] = ld(a.txx_a, j0 - 4 + k);
        qxz[k + 1] = ld(a.txz_a, j0 - 4 + k);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) qzz[k + 1] = ld(a.tzz_a, j0 - 4 + k);
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
        qxx[8] = o.txx4;
        qxz[8] = o.txz4;
        qzz[4] = o.tzz;
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
            if (z == S.isz) ax += w * tap(S.isx - j);
            if (j == S.isx) cz += w * tap(S.isz - z);
            const float bx = o.bx;
            if (damp_cell_f(S, j, z)) {
                const float d1 = 1.f - dt * o.dxv / 2, d2 = 1.f + dt * o.dxv / 2;
                const float d3 = 1.f - dt * o.dzv / 2, d4 = 1.f + dt * o.dzv / 2;
                const float vxx = (d1 * o.vxx + bx * idx * ax) / d2;
                const float vxz = (d3 * o.vxz + bx * idz * bz) / d4;
                const float vzx = (d1 * o.vzx + bx * idx * ex) / d2;
                const float vzz = (d3 * o.vzz + bx * idz * cz) / d4;
                if (zown && j >= x0 && j < x1) {
                    const size_t c = base + (size_t)j * nzp;
                    a.vx_x_b[c] = vxx; a.vx_z_b[c] = vxz; a.vz_x_b[c] = vzx; a.vz_z_b[c] = vzz;
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
            const size_t c = base + (size_t)j * nzp;
            a.vx_b[c] = nvx;
            a.vz_b[c] = nvz;
            if (z == S.igz) {
                const int r = rmap[S.rmap_off + j];
                if (r >= 0) {
                    const size_t q = (size_t)S.rec_off + (size_t)r * nt + it;
                    a.rec_z[q] = nvz;
                    a.rec_x[q] = nvx;
                }
            }
        }
        // ---- stress of column cs = j - 4 (owned columns only)
        const int cs = j - 4;
        if (zown && cs >= x0 && cs < x1) {
            const size_t c = base + (size_t)cs * nzp;
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
                if (damp_cell_f(S, cs, z)) {
                    const float d1 = 1.f - dt * o.dxs / 2, d2 = 1.f + dt * o.dxs / 2;
                    const float d3 = 1.f - dt * o.dzs / 2, d4 = 1.f + dt * o.dzs / 2;
                    const float a1 = (d1 * o.sxx + l2m * dvx) / d2;
                    const float a2 = (d3 * o.sxz + l * dvz) / d4;
                    const float b1 = (d1 * o.szx + l * dvx) / d2;
                    const float b2 = (d3 * o.szz + l2m * dvz) / d4;
                    const float e1 = (d1 * o.sex + mu * dvzx) / d2;
                    const float e2 = (d3 * o.sez + mu * dvxz) / d4;
                    a.txx_x_b[c] = a1; a.txx_z_b[c] = a2; a.tzz_x_b[c] = b1; a.tzz_z_b[c] = b2;
                    a.txz_x_b[c] = e1; a.txz_z_b[c] = e2;
                    txx = a1 + a2;
                    tzz = b1 + b2;
                    txz = e1 + e2;
                } else {
                    txx = txx + (l2m * dvx + l * dvz);
                    tzz = tzz + (l * dvx + l2m * dvz);
                    txz = txz + mu * (dvzx + dvxz);
                }
            }
            a.txx_b[c] = txx;
            a.tzz_b[c] = tzz;
            a.txz_b[c] = txz;
        }
    };

    // even number of columns: an extra trailing column stores nothing
    const int jend = j0 + ((x1 + 4 - j0 + 1) & ~1);
    Col cA = fetch(j0), cB = fetch(j0 + 1);
    for (int j = j0; j < jend; j += 2) {
        column(j, cA, 0);
        cA = fetch(j + 2);
        column(j + 1, cB, 1);
        cB = fetch(j + 3);
    }
}

bool fused_threads_ok(int threads) { return threads == 256 || threads == 512 || threads == 768; }

void launch_fused_step(hipStream_t st, const BatchShape& b, const FusedArgs& a, const ChunkDesc* chunks,
                       int nchunks, const int* rmap, int it, int threads) {
    const int per_xcd = (nchunks + 7) / 8;
    const dim3 grid(per_xcd * 8);
    const float idx = 1.f / b.dx, idz = 1.f / b.dz;
    if (threads == 256)
        hipLaunchKernelGGL(k_fused_step<256>, grid, dim3(256), 0, st, a, chunks, nchunks, per_xcd, rmap, b.slots, idx,
                           idz, b.dt, b.nt, it);
    else if (threads == 512)
        hipLaunchKernelGGL(k_fused_step<512>, grid, dim3(512), 0, st, a, chunks, nchunks, per_xcd, rmap, b.slots, idx,
                           idz, b.dt, b.nt, it);
    else
        hipLaunchKernelGGL(k_fused_step<768>, grid, dim3(768), 0, st, a, chunks, nchunks, per_xcd, rmap, b.slots, idx,
                           idz, b.dt, b.nt, it);
}

}  // namespace ds2d
