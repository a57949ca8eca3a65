// Two elastic time steps per launch for gfx950 (temporal blocking of depth
// 2), batched over uneven shot windows.
//
// Scheme: born.cu:215-349 (update_vx/vz then update_txx_tzz/txz) applied for
// steps it and it+1, each with the transient explosive source of
// FD2DGPU.cu:2039-2049 and followed by the surface recorder of
// FD2DGPU.cu:355-365.  Arithmetic per cell is that of k_fused_step; the
// state read (A) is step it and the state written (B) is step it+2, so every
// field crosses HBM once per two steps.
//
// Work layout: as k_fused_step, one workgroup owns columns [x0, x1) and a
// block of rows of one window, one z row per lane, and marches along x.  A
// second stage trails the first by eight columns:
//   at column j   stage 1: velocity of column j, stress of column j-4 (step it)
//                 stage 2: velocity of column j-8, stress of column j-12 (step it+1)
// Stage 2 reads the step-it+1 values stage 1 left in register queues (x taps)
// and in an LDS column (z taps), so the intermediate step never leaves the
// CU; its parameters (dt/rho, lambda dt, mu dt) come from a per-lane LDS ring
// of what stage 1 loaded eight columns earlier, not from memory again.  Stage 1 is therefore computed on 16 halo rows and columns each side
// of the owned block (velocity on 12, stress on 8): the owned block needs
// nothing of its neighbours' intermediate step.
//
// Damping strips: stage 1 writes the split fields of the damped cells (one
// split per component, as k_fused_step) to a scratch arena and stage 2 reads
// them back eight columns later.  The reading lane wrote the value itself
// (program order), or a neighbouring workgroup whose halo overlaps wrote the
// same bits (both compute it from the same inputs with the same code).
#include "common.hpp"
#include "kernels.hpp"

namespace ds2d {

namespace {

__device__ __forceinline__ bool damp_col2(const SlotDesc& S, int ix) {
    return S.boundary == 0 && (ix < S.pml || ix > S.nxp - S.pml);
}

// A copy in a register distinct from the source (early clobber): loaded
// operands that enter an x-tap queue stay live there for eight columns, and
// the copy lets the next prefetch reuse the load register (no moves of
// in-flight loads at the loop latch).
__device__ __forceinline__ float fresh2(float x) {
    float r;
    asm("v_mov_b32 %0, %1" : "=&v"(r) : "v"(x));
    return r;
}

// Operands of one march column j for one lane.
struct Col2 {
    float txx4, txz4;        // step-it stress of column j+4 (x-tap queues)
    float tzz, vx, vz, bx;   // step-it tzz, velocity and dt/rho of column j
    float lam, mu;           // moduli (x dt) of column j-4
    int rec, rec2;           // traces recorded at column j (step it) and j-8 (step it+1)
    // damping lanes only
    float dxv, dzv, vxx, vzz;        // stage 1 velocity: damping, old vx_x, vz_z (column j)
    float dxs, dzs, sxx, szz, sxz;   // stage 1 stress: damping, old txx_x, tzz_z, txz_x (column j-4)
};
// Damping operands of stage 2 (loaded at the end of the previous column, in
// flight across the barrier and stage 1: prefetching them two columns ahead
// with the other operands would not fit the register budget of 704-lane
// workgroups).
struct Late2 {
    float dxv2, dzv2, vxx2, vzz2;        // stage 2 velocity (column j-8), splits from the scratch
    float dxs2, dzs2, sxx2, szz2, sxz2;  // stage 2 stress (column j-12), splits from the scratch
};

}  // namespace

template <int T, int PF, bool PML>
__global__ __launch_bounds__(T) void k_fused2_step(FusedArgs a, float* __restrict__ sc,
                                                   const ChunkDesc* __restrict__ chunks, int nchunks, int per_xcd,
                                                   const int* __restrict__ rmap, const SlotDesc* __restrict__ slots,
                                                   float idx, float idz, float dt, int nt, int it) {
    // XCD-aware chunk order (neighbouring chunks share halo lines in one L2)
    const int b = blockIdx.x;
    const int ch = (b % 8) * per_xcd + b / 8;
    if (ch >= nchunks) return;
    const ChunkDesc C = chunks[ch];
    const SlotDesc S = slots[C.slot];
    const int nzp = S.nzp, nxp = S.nxp;
    // Lane -> local row r of the block (LDS index).  With damping strips the
    // rows of the z strips (and those outside the window) are gathered into
    // consecutive ranks, so one wave instead of three carries the damped rows
    // of the top and bottom strips; the other waves hold contiguous interior
    // rows.  Rank = lane - 192 (mod T): the damping wave is the block's fourth
    // wave.  Its place matters, with the results bit-identical: on 128 uneven
    // windows of a 2000x600 model with a 20-cell PML (704 lanes) a launch
    // takes 1.485 ms with it in the fourth wave, 1.519 ms in the first,
    // 1.52-1.55 ms in waves 2-8 and 1.66 ms in the last three.
    const int zb = C.z0 - FUSED2_HALO;
    int r = threadIdx.x;
    if (PML && S.boundary == 0) {
        const int a0 = min(max(S.pml - zb, 0), T);                // rows with z < pml
        const int b0 = min(max(nzp - S.pml + 1 - zb, a0), T);      // first row with z > nzp - pml
        const int nb = T - b0;
        int l = (int)threadIdx.x - 192;                             // T >= 512
        if (l < 0) l += T;
        r = l < a0 ? l : (l < a0 + nb ? b0 + (l - a0) : a0 + (l - a0 - nb));
    }
    const int L = r;
    const int z = zb + r;
    const bool zin = z >= 0 && z < nzp;
    const bool zupd = z >= NOP && z < nzp - NOP;
    const bool l1v = L >= 4 && L < T - 4;     // stage 1 velocity (its z taps read rows L-4..L+4)
    const bool l1s = L >= 8 && L < T - 8;     // stage 1 stress
    const bool l2v = L >= 12 && L < T - 12;   // stage 2 velocity
    const bool zown = zin && L >= FUSED2_HALO && L < T - FUSED2_HALO;  // stage 2 stress: rows stored
    const bool zdamp = PML && S.boundary == 0 && (z < S.pml || z > nzp - S.pml);
    const int x0 = C.x0, x1 = C.x1;

    __shared__ float lds[2][8][T];
    // per-lane ring of stage 1's parameters (dt/rho of column j, moduli of
    // column j-4): stage 2 reads them eight columns later instead of loading
    // them again from memory
    __shared__ float pring[8][3][T];

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
    // split-field bases rebuilt from an opaque copy of n inside the damping
    // branches (no long-lived scalar pointers)
    auto opaque_n = [&]() {
        long long v = n;
        asm volatile("" : "+s"(v));
        return v;
    };
    auto splitA = [&](int k) { return a.A + (5 + k) * opaque_n(); };
    auto splitB = [&](int k) { return a.B + (5 + k) * opaque_n(); };
    auto scr = [&](int k) { return sc + k * opaque_n(); };  // vx_x vz_z txx_x tzz_z txz_x of step it+1
    auto inx = [&](int c) { return c >= NOP && c < nxp - NOP; };

    auto fetch = [&](int j) {
        Col2 o;
        o.txx4 = ldc(Atxx, j + 4);
        o.txz4 = ldc(Atxz, j + 4);
        o.tzz = ldc(Atzz, j);
        o.vx = ldc(Avx, j);
        o.vz = ldc(Avz, j);
        o.bx = ldc(Pbx, j);
        o.lam = ldc(Plam, j - 4);
        o.mu = ldc(Pmu, j - 4);
        o.rec = o.rec2 = -1;
        if (z == S.igz) {
            if (j >= x0 && j < x1) o.rec = rmap[S.rmap_off + j];
            if (j - 8 >= x0 && j - 8 < x1) o.rec2 = rmap[S.rmap_off + j - 8];
        }
        o.dxv = o.dzv = o.vxx = o.vzz = 0.f;
        o.dxs = o.dzs = o.sxx = o.szz = o.sxz = 0.f;
        if (PML && zupd) {
            if (inx(j) && (zdamp || damp_col2(S, j))) {
                const uint32_t c = zo4 + (uint32_t)j * colb;
                o.dxv = at(a.dampx, c);
                o.dzv = at(a.dampz, c);
                o.vxx = at(splitA(0), c);
                o.vzz = at(splitA(3), c);
            }
            const int cs = j - 4;
            if (inx(cs) && (zdamp || damp_col2(S, cs))) {
                const uint32_t c = zo4 + (uint32_t)cs * colb;
                o.dxs = at(a.dampx, c);
                o.dzs = at(a.dampz, c);
                o.sxx = at(splitA(4), c);
                o.szz = at(splitA(7), c);
                o.sxz = at(splitA(8), c);
            }
        }
        return o;
    };
    // stage 2 damping operands of column j: the scratch holds what stage 1 of
    // this lane wrote eight (velocity) and eight (stress) columns earlier
    auto fetch_late = [&](int j) {
        Late2 q;
        q.dxv2 = q.dzv2 = q.vxx2 = q.vzz2 = 0.f;
        q.dxs2 = q.dzs2 = q.sxx2 = q.szz2 = q.sxz2 = 0.f;
        if (PML && zupd) {
            const int c2 = j - 8;
            if (l2v && inx(c2) && c2 >= x0 - 4 && c2 < x1 + 4 && (zdamp || damp_col2(S, c2))) {
                const uint32_t c = zo4 + (uint32_t)c2 * colb;
                q.dxv2 = at(a.dampx, c);
                q.dzv2 = at(a.dampz, c);
                q.vxx2 = at(scr(0), c);
                q.vzz2 = at(scr(1), c);
            }
            const int c3 = j - 12;
            if (zown && inx(c3) && c3 >= x0 && c3 < x1 && (zdamp || damp_col2(S, c3))) {
                const uint32_t c = zo4 + (uint32_t)c3 * colb;
                q.dxs2 = at(a.dampx, c);
                q.dzs2 = at(a.dampz, c);
                q.sxx2 = at(scr(2), c);
                q.szz2 = at(scr(3), c);
                q.sxz2 = at(scr(4), c);
            }
        }
        return q;
    };

    const float c1 = DS2D_C1, c2c = DS2D_C2, c3c = DS2D_C3, c4c = DS2D_C4;
    // stage 1 queues (step it values; velocity of step it+1): q*[k] is column j-4+k
    // (stress) or j-8+k (velocity) after the shift at column j
    float qxx[9], qxz[9], qzz[5], qvx[9], qvz[9];
    // stage 2 queues: step-it+1 stress s*[k] = column j-12+k, step-it+2 velocity w*[k] = column j-16+k
    float sxx[9], sxz[9], szz[9], wvx[9], wvz[9];
    const int j0 = x0 - 12;
#pragma unroll
    for (int k = 0; k < 8; ++k) pring[k][0][L] = pring[k][1][L] = pring[k][2][L] = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        qxx[k + 1] = ld(Atxx, j0 - 4 + k);
        qxz[k + 1] = ld(Atxz, j0 - 4 + k);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) qzz[k + 1] = ld(Atzz, j0 - 4 + k);
#pragma unroll
    for (int k = 0; k < 9; ++k) qvx[k] = qvz[k] = sxx[k] = sxz[k] = szz[k] = wvx[k] = wvz[k] = 0.f;

    const float w1 = a.wav[it], w2 = a.wav[it + 1];
    auto tap = [&](int d) -> float {
        return d == 1 ? c1 : d == 2 ? c2c : d == 3 ? c3c : d == 4 ? c4c :
               d == 0 ? -c1 : d == -1 ? -c2c : d == -2 ? -c3c : d == -3 ? -c4c : 0.f;
    };
    const float tz = tap(S.isz - z);
    auto dxp = [&](const float* q) {  // forward difference of a 9-queue centred at [4]/[5]
        return c1 * (q[5] - q[4]) + c2c * (q[6] - q[3]) + c3c * (q[7] - q[2]) + c4c * (q[8] - q[1]);
    };
    auto dxm = [&](const float* q) {  // backward difference of a queue, taps [0..7]
        return c1 * (q[4] - q[3]) + c2c * (q[5] - q[2]) + c3c * (q[6] - q[1]) + c4c * (q[7] - q[0]);
    };
    auto dzp = [&](const float* Lc) {
        return c1 * (Lc[L + 1] - Lc[L]) + c2c * (Lc[L + 2] - Lc[L - 1]) + c3c * (Lc[L + 3] - Lc[L - 2]) +
               c4c * (Lc[L + 4] - Lc[L - 3]);
    };
    auto dzm = [&](const float* Lc) {
        return c1 * (Lc[L] - Lc[L - 1]) + c2c * (Lc[L + 1] - Lc[L - 2]) + c3c * (Lc[L + 2] - Lc[L - 3]) +
               c4c * (Lc[L + 3] - Lc[L - 4]);
    };

    // velocity update of one stage: old velocity (ovx, ovz), x taps (ax, ex),
    // z taps (bz, cz), dt/rho; damping operands when the wave holds a damped cell
    struct V2 {
        float vx, vz, sx, sz;  // new velocity, new splits vx_x, vz_z
    };
    auto vel = [&](float ovx, float ovz, float ax, float ex, float bz, float cz, float bx, float dxv, float dzv,
                   float pvxx, float pvzz, bool damped) -> V2 {
        V2 r;
        r.sx = r.sz = 0.f;
        if (PML && __any(damped)) {
            const float d1 = 1.f - dt * dxv / 2, r2 = __builtin_amdgcn_rcpf(1.f + dt * dxv / 2);
            const float d3 = 1.f - dt * dzv / 2, r4 = __builtin_amdgcn_rcpf(1.f + dt * dzv / 2);
            const float vxx = (d1 * pvxx + bx * idx * ax) * r2;
            const float vxz = (d3 * (ovx - pvxx) + bx * idz * bz) * r4;
            const float vzx = (d1 * (ovz - pvzz) + bx * idx * ex) * r2;
            const float vzz = (d3 * pvzz + bx * idz * cz) * r4;
            r.vx = vxx + vxz;
            r.vz = vzx + vzz;
            r.sx = vxx;
            r.sz = vzz;
        } else {
            r.vx = ovx + bx * (ax * idx + bz * idz);
            r.vz = ovz + bx * (ex * idx + cz * idz);
        }
        return r;
    };
    struct S3 {
        float xx, zz, xz, sxx, szz, sxz;  // new stress and splits txx_x, tzz_z, txz_x
    };
    auto str = [&](float txx, float tzz, float txz, float dvx, float dvzx, float dvz, float dvxz, float l, float mu,
                   float dxs, float dzs, float psxx, float pszz, float psxz, bool damped) -> S3 {
        S3 r;
        r.sxx = r.szz = r.sxz = 0.f;
        const float l2m = l + 2 * mu;
        if (PML && __any(damped)) {
            const float d1 = 1.f - dt * dxs / 2, r2 = __builtin_amdgcn_rcpf(1.f + dt * dxs / 2);
            const float d3 = 1.f - dt * dzs / 2, r4 = __builtin_amdgcn_rcpf(1.f + dt * dzs / 2);
            const float a1 = (d1 * psxx + l2m * dvx) * r2;
            const float a2 = (d3 * (txx - psxx) + l * dvz) * r4;
            const float b1 = (d1 * (tzz - pszz) + l * dvx) * r2;
            const float b2 = (d3 * pszz + l2m * dvz) * r4;
            const float e1 = (d1 * psxz + mu * dvzx) * r2;
            const float e2 = (d3 * (txz - psxz) + mu * dvxz) * r4;
            r.xx = a1 + a2;
            r.zz = b1 + b2;
            r.xz = e1 + e2;
            r.sxx = a1;
            r.szz = b2;
            r.sxz = e1;
        } else {
            r.xx = txx + (l2m * dvx + l * dvz);
            r.zz = tzz + (l * dvx + l2m * dvz);
            r.xz = txz + mu * (dvzx + dvxz);
        }
        return r;
    };

    auto column = [&](const int j, const Col2& o, const int p, Late2& q) {
        const int slot = j & 7;  // (j - 8) & 7 as well
        const float bx2 = pring[slot][0][L], lam2 = pring[slot][1][L], mu2 = pring[slot][2][L];
        pring[slot][0][L] = o.bx;
        pring[slot][1][L] = o.lam;
        pring[slot][2][L] = o.mu;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            qxx[k] = qxx[k + 1];
            qxz[k] = qxz[k + 1];
            qvx[k] = qvx[k + 1];
            qvz[k] = qvz[k + 1];
            sxx[k] = sxx[k + 1];
            sxz[k] = sxz[k + 1];
            szz[k] = szz[k + 1];
            wvx[k] = wvx[k + 1];
            wvz[k] = wvz[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) qzz[k] = qzz[k + 1];
        qxx[8] = fresh2(o.txx4);
        qxz[8] = fresh2(o.txz4);
        qzz[4] = fresh2(o.tzz);
        // z-tap columns of all four updates, one barrier
        lds[p][0][L] = qxz[4];  // txz, step it, column j
        lds[p][1][L] = qzz[4];  // tzz, step it, column j
        lds[p][2][L] = qvx[4];  // vx, step it+1, column j-4
        lds[p][3][L] = qvz[4];  // vz, step it+1, column j-4
        lds[p][4][L] = sxz[4];  // txz, step it+1, column j-8
        lds[p][5][L] = szz[4];  // tzz, step it+1, column j-8
        lds[p][6][L] = wvx[4];  // vx, step it+2, column j-12
        lds[p][7][L] = wvz[4];  // vz, step it+2, column j-12
        __syncthreads();

        // ---- stage 1 velocity, column j (step it -> it+1)
        float nvx = o.vx, nvz = o.vz;
        if (l1v && zupd && inx(j)) {
            float ax = dxp(qxx), ex = dxm(qxz), bz = dzm(lds[p][0]), cz = dzp(lds[p][1]);
            if ((unsigned)(S.isx - j + 3) < 8u && z == S.isz) ax += w1 * tap(S.isx - j);
            if (j == S.isx) cz += w1 * tz;
            const bool dv = PML && (zdamp || damp_col2(S, j));
            const V2 r = vel(o.vx, o.vz, ax, ex, bz, cz, o.bx, o.dxv, o.dzv, o.vxx, o.vzz, dv);
            nvx = r.vx;
            nvz = r.vz;
            if (PML && dv && l2v && j >= x0 - 4 && j < x1 + 4) {
                const uint32_t c = zo4 + (uint32_t)j * colb;
                st(scr(0), c, r.sx);
                st(scr(1), c, r.sz);
            }
        }
        qvx[8] = nvx;
        qvz[8] = nvz;
        if (zown && o.rec >= 0) {
            const size_t q = (size_t)S.rec_off + (size_t)o.rec * nt + it;
            a.rec_z[q] = nvz;
            a.rec_x[q] = nvx;
        }

        // ---- stage 1 stress, column j-4
        {
            const int cs = j - 4;
            float txx = qxx[0], tzz = qzz[0], txz = qxz[0];
            if (l1s && zupd && inx(cs)) {
                const float dvx = dxm(qvx) * idx;
                const float dvzx = dxp(qvz) * idx;
                const float dvz = dzm(lds[p][3]) * idz;
                const float dvxz = dzp(lds[p][2]) * idz;
                const bool ds = PML && (zdamp || damp_col2(S, cs));
                const S3 r = str(txx, tzz, txz, dvx, dvzx, dvz, dvxz, o.lam, o.mu, o.dxs, o.dzs, o.sxx, o.szz, o.sxz,
                                 ds);
                txx = r.xx;
                tzz = r.zz;
                txz = r.xz;
                if (PML && ds && zown && cs >= x0 && cs < x1) {
                    const uint32_t c = zo4 + (uint32_t)cs * colb;
                    st(scr(2), c, r.sxx);
                    st(scr(3), c, r.szz);
                    st(scr(4), c, r.sxz);
                }
            }
            sxx[8] = txx;
            szz[8] = tzz;
            sxz[8] = txz;
        }

        // ---- stage 2 velocity, column j-8 (step it+1 -> it+2)
        {
            const int c2 = j - 8;
            float mvx = qvx[0], mvz = qvz[0];
            if (l2v && zupd && inx(c2)) {
                float ax = dxp(sxx), ex = dxm(sxz), bz = dzm(lds[p][4]), cz = dzp(lds[p][5]);
                if ((unsigned)(S.isx - c2 + 3) < 8u && z == S.isz) ax += w2 * tap(S.isx - c2);
                if (c2 == S.isx) cz += w2 * tz;
                const bool dv = PML && (zdamp || damp_col2(S, c2));
                const V2 r = vel(mvx, mvz, ax, ex, bz, cz, bx2, q.dxv2, q.dzv2, q.vxx2, q.vzz2, dv);
                mvx = r.vx;
                mvz = r.vz;
                if (PML && dv && zown && c2 >= x0 && c2 < x1) {
                    const uint32_t c = zo4 + (uint32_t)c2 * colb;
                    st(splitB(0), c, r.sx);
                    st(splitB(3), c, r.sz);
                }
            }
            wvx[8] = mvx;
            wvz[8] = mvz;
            if (zown && c2 >= x0 && c2 < x1) {
                const uint32_t c = zo4 + (uint32_t)c2 * colb;
                st(a.B, c, mvx);
                st(a.B + n, c, mvz);
                if (o.rec2 >= 0) {
                    const size_t q = (size_t)S.rec_off + (size_t)o.rec2 * nt + it + 1;
                    a.rec_z[q] = mvz;
                    a.rec_x[q] = mvx;
                }
            }
        }

        // ---- stage 2 stress, column j-12 (owned rows and columns only)
        {
            const int c3 = j - 12;
            if (zown && c3 >= x0 && c3 < x1) {
                float txx = sxx[0], tzz = szz[0], txz = sxz[0];
                if (zupd && inx(c3)) {
                    const float dvx = dxm(wvx) * idx;
                    const float dvzx = dxp(wvz) * idx;
                    const float dvz = dzm(lds[p][7]) * idz;
                    const float dvxz = dzp(lds[p][6]) * idz;
                    const bool ds = PML && (zdamp || damp_col2(S, c3));
                    const S3 r = str(txx, tzz, txz, dvx, dvzx, dvz, dvxz, lam2, mu2, q.dxs2, q.dzs2, q.sxx2,
                                     q.szz2, q.sxz2, ds);
                    txx = r.xx;
                    tzz = r.zz;
                    txz = r.xz;
                    if (PML && ds) {
                        const uint32_t c = zo4 + (uint32_t)c3 * colb;
                        st(splitB(4), c, r.sxx);
                        st(splitB(7), c, r.szz);
                        st(splitB(8), c, r.sxz);
                    }
                }
                const uint32_t c = zo4 + (uint32_t)c3 * colb;
                st(a.B + 2 * n, c, txx);
                st(a.B + 3 * n, c, tzz);
                st(a.B + 4 * n, c, txz);
            }
        }
        // stage 2 damping operands of the next column, in flight across the barrier
        q = fetch_late(j + 1);
    };

    // march columns [x0-12, x1+12) in passes of two columns, the operands of
    // the next two in flight
    static_assert(PF == 2, "two steps per launch march with two columns in flight");
    const int jend = j0 + (x1 + 12 - j0 + 1) / 2 * 2;
    Col2 cA = fetch(j0), cB = fetch(j0 + 1);
    Late2 q = fetch_late(j0);
    for (int j = j0; j < jend; j += 2) {
        column(j, cA, 0, q);
        cA = fetch(j + 2);
        column(j + 1, cB, 1, q);
        cB = fetch(j + 3);
    }
}

bool fused2_threads_ok(int threads) { return threads == 512 || threads == 704 || threads == 768; }

void launch_fused2_step(hipStream_t st, const BatchShape& b, const FusedArgs& a, float* scratch,
                        const ChunkDesc* chunks, int nchunks, const int* rmap, int it, int threads, int depth,
                        bool pml) {
    const int per_xcd = (nchunks + 7) / 8;
    const dim3 grid(per_xcd * 8);
    const float idx = 1.f / b.dx, idz = 1.f / b.dz;
#define DS2D_F2_LAUNCH(TT, PP, MM)                                                                                  \
    hipLaunchKernelGGL((k_fused2_step<TT, PP, MM>), grid, dim3(TT), 0, st, a, scratch, chunks, nchunks, per_xcd, rmap, \
                       b.slots, idx, idz, b.dt, b.nt, it)
#define DS2D_F2_THREADS(PP, MM)         \
    if (threads == 512)                 \
        DS2D_F2_LAUNCH(512, PP, MM);    \
    else if (threads == 704)            \
        DS2D_F2_LAUNCH(704, PP, MM);    \
    else                                \
        DS2D_F2_LAUNCH(768, PP, MM);
    (void)depth;  // two columns in flight (the only depth of this kernel)
    if (pml) {
        DS2D_F2_THREADS(2, true)
    } else {
        DS2D_F2_THREADS(2, false)
    }
#undef DS2D_F2_THREADS
#undef DS2D_F2_LAUNCH
}

}  // namespace ds2d
