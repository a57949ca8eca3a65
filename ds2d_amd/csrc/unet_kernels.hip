// BF16 MFMA kernels of the learned LSRTM preconditioner (U-Net), gfx950.
//
// The 3x3 convolutions are implicit GEMMs on v_mfma_f32_32x32x16_bf16:
//   forward / data gradient: rows = output pixels, K = (tap, input channel),
//     columns = output channels.  NHWC makes both MFMA operands one 16-byte
//     load per lane (8 consecutive channels of one pixel; 8 consecutive input
//     channels of one weight row), so no LDS is needed.
//   weight gradient: rows = output channels, columns = input channels,
//     K = pixels.  Both operands are transposed through LDS (channel-major
//     rows of a 64-pixel z segment); the three z shifts of a tap row come from
//     one 16-byte + one 4-byte LDS read with v_alignbyte, and a 3-row ring
//     keeps each input row in LDS for the three output rows that use it.
// MFMA lane layouts (32x32x16, 64 lanes, lane = l, li = l % 32, kh = l / 32):
//   A[i][k]: lane i + 32 * (k / 8), element k % 8
//   B[k][j]: lane j + 32 * (k / 8), element k % 8
//   D[i][j]: lane j + 32 * ((i / 4) % 2), element (i % 4) + 4 * (i / 8)
#include <hip/hip_runtime.h>

#include "unet.hpp"

namespace ds2d {
namespace unet {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16_t f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ u32x4 ld16(const bf16_t* p) { return *reinterpret_cast<const u32x4*>(p); }
__device__ __forceinline__ f32x16 mfma(u32x4 a, u32x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// ---------------------------------------------------------------------------
// 3x3 convolution (forward and data gradient)
// ---------------------------------------------------------------------------
// Block = 4 waves stacked along pixels; a wave owns WM x 32 pixels times
// WN x 32 output channels (WM * WN accumulator tiles of 16 floats per lane).
template <int WM, int WN>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 31, kh = lane >> 5;
    const long long npix = (long long)a.n * a.h * a.w;
    const long long pbase = ((long long)blockIdx.x * 4 + wave) * (WM * 32);
    if (pbase >= npix) return;  // no barriers in this kernel
    const int co0 = blockIdx.y * (WN * 32);
    const int ctot = a.c1 + a.c2;
    const int h2 = a.h >> a.up, w2 = a.w >> a.up;
    int pn[WM], px[WM], pz[WM];
    bool pv[WM];
#pragma unroll
    for (int m = 0; m < WM; ++m) {
        const long long p = pbase + m * 32 + li;
        pv[m] = p < npix;
        long long q = pv[m] ? p : 0;
        pz[m] = (int)(q % a.w);
        q /= a.w;
        px[m] = (int)(q % a.h);
        pn[m] = (int)(q / a.h);
    }
    f32x16 acc[WM][WN];
#pragma unroll
    for (int m = 0; m < WM; ++m)
#pragma unroll
        for (int j = 0; j < WN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[m][j][r] = 0.f;
    const u32x4 zero = {0u, 0u, 0u, 0u};
    for (int tap = 0; tap < 9; ++tap) {
        const int dx = tap / 3 - 1, dz = tap % 3 - 1;
        const bf16_t* s1[WM];
        const bf16_t* s2[WM];
#pragma unroll
        for (int m = 0; m < WM; ++m) {
            const int xs = px[m] + dx, zs = pz[m] + dz;
            const bool ok = pv[m] && xs >= 0 && xs < a.h && zs >= 0 && zs < a.w;
            s1[m] = ok ? a.x1 + (((long long)pn[m] * a.h + xs) * a.w + zs) * a.c1 + 8 * kh : nullptr;
            s2[m] = (ok && a.c2) ? a.x2 + (((long long)pn[m] * h2 + (xs >> a.up)) * w2 + (zs >> a.up)) * a.c2 + 8 * kh
                                 : nullptr;
        }
        const bf16_t* wt = a.wt + ((long long)tap * a.cout + co0 + li) * ctot + 8 * kh;
#pragma unroll 2
        for (int c = 0; c < ctot; c += 16) {
            u32x4 A[WM], B[WN];
#pragma unroll
            for (int m = 0; m < WM; ++m)
                A[m] = !s1[m] ? zero : (c < a.c1 ? ld16(s1[m] + c) : ld16(s2[m] + (c - a.c1)));
#pragma unroll
            for (int j = 0; j < WN; ++j) B[j] = ld16(wt + (long long)j * 32 * ctot + c);
#pragma unroll
            for (int m = 0; m < WM; ++m)
#pragma unroll
                for (int j = 0; j < WN; ++j) acc[m][j] = mfma(A[m], B[j], acc[m][j]);
        }
    }
    const int c2o = a.cout - a.split;
#pragma unroll
    for (int m = 0; m < WM; ++m)
#pragma unroll
        for (int j = 0; j < WN; ++j) {
            const int co = co0 + j * 32 + li;
            const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = (r & 3) + 4 * kh + 8 * (r >> 2);
                const long long p = pbase + m * 32 + i;
                if (p >= npix) continue;
                float v = acc[m][j][r] + b;
                if (a.relu) v = fmaxf(v, 0.f);
                if (a.mask && !(bf2f(a.mask[p * a.cout + co]) > 0.f)) v = 0.f;
                if (co < a.split) {
                    if (a.out_bf16)
                        ((bf16_t*)a.y1)[p * a.split + co] = f2bf(v);
                    else
                        ((float*)a.y1)[p * a.split + co] = v;
                } else {
                    if (a.out_bf16)
                        ((bf16_t*)a.y2)[p * c2o + co - a.split] = f2bf(v);
                    else
                        ((float*)a.y2)[p * c2o + co - a.split] = v;
                }
            }
        }
}

// Columns of a block of P pixels in the LDS kernel: P / 4 (blocks of 4 rows,
// a smaller frame per pixel than whole rows) when that divides a wider image,
// else whole rows.
__host__ __device__ inline int conv_lds_cols(int w, int P) { return w > P / 4 && w % (P / 4) == 0 ? P / 4 : w; }

// Same GEMM with the input staged through LDS: a block's P = 128 * WM pixels
// are R = P / CB image rows of CB = conv_lds_cols(w, P) columns; per K step of 16
// channels the block loads them plus a one-pixel frame ((R + 2) x (CB + 2)
// pixels, 32 B each, zero
// outside the image, skip concatenation and upsampling resolved in the
// address) and the step's weights of all nine taps for the block's output
// channels once; all nine taps then read both MFMA operands from LDS.
// The next K step is staged in registers while the MFMAs of the current one run
// (3 waves per SIMD).  CB = conv_lds_cols(w, P) is a template argument (P / 4,
// P / 8 or P / 16), so the pixel <-> (row, column) arithmetic of the tile is by
// constants.
template <int WM, int WN, int CB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_conv_lds(ConvArgs a) {
    extern __shared__ u32x4 tile[];  // [(R + 2) * (CB + 2)][2]: 16 channels of one pixel; then the weights
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 31, kh = lane >> 5;
    constexpr int P = 4 * WM * 32, NC = WN * 32, R = P / CB, wp = CB + 2;
    const int w = a.w, h = a.h;
    u32x4* wl = tile + 2 * (R + 2) * wp;  // [9][NC][2]: 16 input channels of each tap and output channel
    const int ncb = w / CB, bpi = (h / R) * ncb;  // column blocks per row block, blocks per image
    const int nimg = blockIdx.x / bpi, rem = blockIdx.x % bpi;
    const int x0 = (rem / ncb) * R, zb0 = (rem % ncb) * CB;
    const int co0 = blockIdx.y * (WN * 32);
    const int ctot = a.c1 + a.c2;
    const int h2 = h >> a.up, w2 = w >> a.up;
    constexpr int npx = (R + 2) * wp;
    int lp[WM];  // LDS pixel of each tile row's pixel at tap (0, 0)
#pragma unroll
    for (int m = 0; m < WM; ++m) {
        const int i = wave * WM * 32 + m * 32 + li;
        lp[m] = (i / CB + 1) * wp + (i % CB) + 1;
    }
    f32x16 acc[WM][WN];
#pragma unroll
    for (int m = 0; m < WM; ++m)
#pragma unroll
        for (int j = 0; j < WN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[m][j][r] = 0.f;
    const u32x4 zero = {0u, 0u, 0u, 0u};
    // The items of K step c: the framed tile, then wl[(tap * NC + col) * 2 + half] =
    // wt[tap][co0 + col][c + 8 * half .. + 8].  They are loaded into registers
    // while the MFMAs of step c - 16 run, and stored to LDS after them.
    constexpr int ntot = 2 * npx + 18 * NC;
    constexpr int NST = (ntot + 255) / 256;
    u32x4 stg[NST];
    auto fetch = [&](int c) {
#pragma unroll
        for (int k = 0; k < NST; ++k) {
            const int q = tid + 256 * k;
            u32x4 v = zero;
            if (q < 2 * npx) {
                const int px = q >> 1, half = q & 1;
                const int x = x0 - 1 + px / wp, z = zb0 + px % wp - 1;
                if (x >= 0 && x < h && z >= 0 && z < w)
                    v = c < a.c1 ? ld16(a.x1 + (((long long)nimg * h + x) * w + z) * a.c1 + c + 8 * half)
                                 : ld16(a.x2 + (((long long)nimg * h2 + (x >> a.up)) * w2 + (z >> a.up)) * a.c2 +
                                        (c - a.c1) + 8 * half);
            } else if (q < ntot) {
                const int qw = q - 2 * npx, half = qw & 1, r = qw >> 1, tap = r / NC, col = r % NC;
                v = ld16(a.wt + ((long long)tap * a.cout + co0 + col) * ctot + c + 8 * half);
            }
            stg[k] = v;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int k = 0; k < NST; ++k) {
            const int q = tid + 256 * k;
            if (q < ntot) tile[q] = stg[k];
        }
    };
    fetch(0);
    stash();
    for (int c = 0; c < ctot; c += 16) {
        const bool more = c + 16 < ctot;
        __syncthreads();
        if (more) fetch(c + 16);
#pragma unroll 3
        for (int tap = 0; tap < 9; ++tap) {
            const int off = (tap / 3 - 1) * wp + (tap % 3 - 1);
            u32x4 A[WM], B[WN];
#pragma unroll
            for (int j = 0; j < WN; ++j) B[j] = wl[2 * (tap * NC + j * 32 + li) + kh];
#pragma unroll
            for (int m = 0; m < WM; ++m) A[m] = tile[2 * (lp[m] + off) + kh];
#pragma unroll
            for (int m = 0; m < WM; ++m)
#pragma unroll
                for (int j = 0; j < WN; ++j) acc[m][j] = mfma(A[m], B[j], acc[m][j]);
        }
        __syncthreads();
        if (more) stash();
    }
    const long long img0 = (long long)nimg * h * w;
    const int c2o = a.cout - a.split;
#pragma unroll
    for (int m = 0; m < WM; ++m)
#pragma unroll
        for (int j = 0; j < WN; ++j) {
            const int co = co0 + j * 32 + li;
            const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = wave * WM * 32 + m * 32 + (r & 3) + 4 * kh + 8 * (r >> 2);  // pixel of the block
                const long long p = img0 + (long long)(x0 + i / CB) * w + zb0 + i % CB;
                float v = acc[m][j][r] + b;
                if (a.relu) v = fmaxf(v, 0.f);
                if (a.mask && !(bf2f(a.mask[p * a.cout + co]) > 0.f)) v = 0.f;
                if (co < a.split) {
                    if (a.out_bf16)
                        ((bf16_t*)a.y1)[p * a.split + co] = f2bf(v);
                    else
                        ((float*)a.y1)[p * a.split + co] = v;
                } else {
                    if (a.out_bf16)
                        ((bf16_t*)a.y2)[p * c2o + co - a.split] = f2bf(v);
                    else
                        ((float*)a.y2)[p * c2o + co - a.split] = v;
                }
            }
        }
}

// The LDS kernel takes blocks of R = P / CB rows of CB = conv_lds_cols(w, P)
// columns: CB divides P, R divides h, channels come in 16s.
bool conv_lds_ok(const ConvArgs& a, int P) {
    if (a.w < 16 || a.h <= 0) return false;
    const int cb = conv_lds_cols(a.w, P);
    return (cb == P / 4 || cb == P / 8 || cb == P / 16) && a.h % (P / cb) == 0 && a.c1 % 16 == 0 &&
           a.c2 % 16 == 0;
}

// Dynamic LDS of the kernel: the framed pixel tile and the weights of nine taps.
static size_t conv_lds_bytes(int w, int P, int nc) {
    const int cb = conv_lds_cols(w, P);
    return (size_t)(P / cb + 2) * (cb + 2) * 32 + 9 * (size_t)nc * 32;
}

template <int WM, int WN>
static void launch_conv_lds(hipStream_t st, const ConvArgs& a, long long npix) {
    constexpr int P = 4 * WM * 32;
    const dim3 grid((unsigned)(npix / P), a.cout / (WN * 32));
    const size_t lds = conv_lds_bytes(a.w, P, WN * 32);
    switch (conv_lds_cols(a.w, P)) {
        case P / 4: hipLaunchKernelGGL((k_conv_lds<WM, WN, P / 4>), grid, dim3(256), lds, st, a); break;
        case P / 8: hipLaunchKernelGGL((k_conv_lds<WM, WN, P / 8>), grid, dim3(256), lds, st, a); break;
        default: hipLaunchKernelGGL((k_conv_lds<WM, WN, P / 16>), grid, dim3(256), lds, st, a); break;
    }
}

void launch_conv(hipStream_t st, const ConvArgs& a) {
    const long long npix = (long long)a.n * a.h * a.w;
    if (a.cout % 64 == 0 && conv_lds_ok(a, 256)) {
        launch_conv_lds<2, 2>(st, a, npix);
        return;
    }
    if (a.cout % 64 != 0 && conv_lds_ok(a, 256)) {
        launch_conv_lds<2, 1>(st, a, npix);
        return;
    }
    if (a.cout % 64 == 0) {
        dim3 grid((unsigned)((npix + 255) / 256), a.cout / 64);
        hipLaunchKernelGGL((k_conv<2, 2>), grid, dim3(256), 0, st, a);
    } else {
        dim3 grid((unsigned)((npix + 511) / 512), a.cout / 32);
        hipLaunchKernelGGL((k_conv<4, 1>), grid, dim3(256), 0, st, a);
    }
}

// ---------------------------------------------------------------------------
// 3x3 weight gradient
// ---------------------------------------------------------------------------
constexpr int ZS = 64;  // z segment (pixels of K per row)
constexpr int LP = 72;  // LDS row pitch in elements (144 B: 16-byte aligned rows)

// LDS rows of the weight gradient are swizzled in 16-byte chunks: logical chunk
// j < 8 (pixels 8j .. 8j + 7) of channel row c sits at chunk j ^ ((c >> 3) & 7);
// chunk 8 (pixels 64, 65 of an input row) stays in place.  The transposing
// writes of 8 channels x 2 pixels then fall in distinct banks.
__device__ __forceinline__ int wg_chunk(int c, int j) { return j < 8 ? j ^ ((c >> 3) & 7) : j; }

// Block = (TCO/32) x (TCI/32) x DS waves; wave (cot, cit, dsel) accumulates the
// 32x32 (co, ci) tile of the taps dx = d - 1 with d = dsel, dsel + DS, ... < 3.
// Input rows live in a 4-slot LDS ring (row r in slot r & 3, r >= -1) and dy
// rows in two buffers, so the global loads of rows x + 2 (input) and x + 1
// (dy) are issued before the MFMAs of row x and land in LDS after them, with
// one barrier per row.  Each thread moves two pixels x 8 channels per item
// (two 16-byte loads, eight 4-byte LDS writes of a channel's pixel pair).
template <int TCO, int TCI, int DS>
__global__ __launch_bounds__(64 * (TCO / 32) * (TCI / 32) * DS) __attribute__((amdgpu_waves_per_eu(4))) void k_wgrad(WgradArgs a) {
    constexpr int NT = 64 * (TCO / 32) * (TCI / 32) * DS;
    constexpr int GI = TCI / 8, GO = TCO / 8;
    constexpr int QI = (ZS + 2) / 2 * GI, QO = ZS / 2 * GO;  // items: (pixel pair, channel group)
    constexpr int NI = (QI + NT - 1) / NT, NO = (QO + NT - 1) / NT;
    constexpr int ND = 3 / DS;
    static_assert(DS == 1 || DS == 3, "taps split over 1 or 3 waves");
    __shared__ __attribute__((aligned(16))) bf16_t As[2][TCO][LP];
    __shared__ __attribute__((aligned(16))) bf16_t Bs[4][TCI][LP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 31, kh = lane >> 5;
    const int cot = wave % (TCO / 32), cit = (wave / (TCO / 32)) % (TCI / 32);
    const int dsel = wave / ((TCO / 32) * (TCI / 32));
    const int cin = a.c1 + a.c2;
    const int ntci = cin / TCI;
    const int co0 = (blockIdx.y / ntci) * TCO, ci0 = (blockIdx.y % ntci) * TCI;
    const int chunk = blockIdx.x;
    const int xb = chunk % a.nxblk;
    const int seg = (chunk / a.nxblk) % a.nseg;
    const int n = chunk / (a.nxblk * a.nseg);
    const int zs = seg * ZS;
    const int xa = xb * a.rows, xe = min(a.h, xa + a.rows);
    const int h2 = a.h >> a.up, w2 = a.w >> a.up;
    const int nks = (min(ZS, a.w - zs) + 15) / 16;
    const u32x4 zero = {0u, 0u, 0u, 0u};

    // Four 4-byte words per 16-byte pixel load: word k holds channels 2k, 2k + 1.
    // (v0, v1) = pixels (zz, zz + 1) of channels 8g .. 8g + 7 -> rows 8g + 2k, 8g + 2k + 1.
    auto put_pair = [](bf16_t (*rows)[LP], int g, int zz, const u32x4& v0, const u32x4& v1) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int c = 8 * g + 2 * k, e = wg_chunk(c, zz >> 3) * 8 + (zz & 7);
            *reinterpret_cast<unsigned*>(&rows[c][e]) = (v0[k] & 0xFFFFu) | (v1[k] << 16);
            *reinterpret_cast<unsigned*>(&rows[c + 1][e]) = (v0[k] >> 16) | (v1[k] & 0xFFFF0000u);
        }
    };
    // input row x: item q = (pair p, g) holds in[n][x][zs + 2p - 1 .. + 1][ci0 + 8g .. + 8], zero outside the image
    u32x4 ri[NI][2], ro[NO][2];
    auto in_px = [&](int x, int z, int c) -> u32x4 {
        if (x < 0 || x >= a.h || z < 0 || z >= a.w) return zero;
        return c < a.c1 ? ld16(a.x1 + (((long long)n * a.h + x) * a.w + z) * a.c1 + c)
                        : ld16(a.x2 + (((long long)n * h2 + (x >> a.up)) * w2 + (z >> a.up)) * a.c2 + (c - a.c1));
    };
    auto fetch_in = [&](int x) {
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int q = tid + k * NT, p = q / GI, g = q % GI;
            const int z = zs + 2 * p - 1, c = ci0 + g * 8;
            ri[k][0] = ri[k][1] = zero;
            if (q < QI) {
                ri[k][0] = in_px(x, z, c);
                ri[k][1] = in_px(x, z + 1, c);
            }
        }
    };
    auto store_in = [&](int slot) {
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int q = tid + k * NT, p = q / GI, g = q % GI;
            if (q < QI) put_pair(Bs[slot], g, 2 * p, ri[k][0], ri[k][1]);
        }
    };
    // dy row x: item q = (pair p, g) holds dy[n][x][zs + 2p .. + 1][co0 + 8g .. + 8], zero beyond the image
    auto fetch_dy = [&](int x) {
#pragma unroll
        for (int k = 0; k < NO; ++k) {
            const int q = tid + k * NT, p = q / GO, g = q % GO;
            const int z = zs + 2 * p;
            const bf16_t* src = a.dy + (((long long)n * a.h + x) * a.w + z) * a.cout + co0 + g * 8;
            ro[k][0] = ro[k][1] = zero;
            if (q < QO && z < a.w) ro[k][0] = ld16(src);
            if (q < QO && z + 1 < a.w) ro[k][1] = ld16(src + a.cout);
        }
    };
    auto store_dy = [&](int buf) {
#pragma unroll
        for (int k = 0; k < NO; ++k) {
            const int q = tid + k * NT, p = q / GO, g = q % GO;
            if (q < QO) put_pair(As[buf], g, 2 * p, ro[k][0], ro[k][1]);
        }
    };

    f32x16 acc[3 * ND];
#pragma unroll
    for (int t = 0; t < 3 * ND; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.f;
    // bias partials (blocks of the first cin tile): NB threads per output
    // channel, each summing ZS / NB pixels of every dy row
    constexpr int NB = NT / TCO >= 8 ? 8 : NT / TCO >= 4 ? 4 : NT / TCO >= 2 ? 2 : 1;
    const bool do_bias = a.bias && blockIdx.y % ntci == 0;
    const int bco = tid % TCO, bpart = tid / TCO;
    float bacc = 0.f;

    for (int r = xa - 1; r <= xa + 1; ++r) {
        fetch_in(r);
        store_in((r + 4) & 3);
    }
    fetch_dy(xa);
    store_dy(xa & 1);
    __syncthreads();
    for (int x = xa; x < xe; ++x) {
        const bool more = x + 1 < xe;
        if (more) {
            fetch_in(x + 2);
            fetch_dy(x + 1);
        }
        for (int s = 0; s < nks; ++s) {
            const int base = 16 * s + 8 * kh;
            const int ca = cot * 32 + li, cb = cit * 32 + li, jb = base >> 3;
            const u32x4 A = *reinterpret_cast<const u32x4*>(&As[x & 1][ca][wg_chunk(ca, jb) * 8]);
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                const int d = dsel + j * DS;  // dx = d - 1: input row x + d - 1
                const bf16_t* row = Bs[(x + d + 3) & 3][cb];
                const u32x4 q = *reinterpret_cast<const u32x4*>(row + wg_chunk(cb, jb) * 8);
                const unsigned e = *reinterpret_cast<const unsigned*>(row + wg_chunk(cb, jb + 1) * 8);
                const u32x4 bm = q;                                                   // dz = -1
                const u32x4 b0 = {__builtin_amdgcn_alignbyte(q[1], q[0], 2), __builtin_amdgcn_alignbyte(q[2], q[1], 2),
                                  __builtin_amdgcn_alignbyte(q[3], q[2], 2), __builtin_amdgcn_alignbyte(e, q[3], 2)};
                const u32x4 bp = {q[1], q[2], q[3], e};                               // dz = +1
                acc[j * 3 + 0] = mfma(A, bm, acc[j * 3 + 0]);
                acc[j * 3 + 1] = mfma(A, b0, acc[j * 3 + 1]);
                acc[j * 3 + 2] = mfma(A, bp, acc[j * 3 + 2]);
            }
        }
        if (do_bias && bpart < NB) {
            const bf16_t* rp = As[x & 1][bco];
#pragma unroll
            for (int k = 0; k < ZS / NB / 8; ++k) {
                const u32x4 v = *reinterpret_cast<const u32x4*>(rp + wg_chunk(bco, bpart * (ZS / NB / 8) + k) * 8);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    bacc += __uint_as_float(v[e] << 16);
                    bacc += __uint_as_float(v[e] & 0xFFFF0000u);
                }
            }
        }
        if (more) {
            store_in((x + 2) & 3);
            store_dy((x + 1) & 1);
        }
        __syncthreads();
    }
    const long long rec = 9LL * a.cout * cin;
    float* out = a.part + (long long)chunk * (rec + (a.bias ? a.cout : 0));
    if (do_bias) {  // the ring is free after the last barrier
        float* red = reinterpret_cast<float*>(&Bs[0][0][0]);
        if (bpart < NB) red[bpart * TCO + bco] = bacc;
        __syncthreads();
        if (tid < TCO) {
            float sb = 0.f;
            for (int k = 0; k < NB; ++k) sb += red[k * TCO + tid];
            out[rec + co0 + tid] = sb;
        }
    }
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int t = (dsel + j * DS) * 3 + dz;
                const int i = (r & 3) + 4 * kh + 8 * (r >> 2);
                const int co = co0 + cot * 32 + i, ci = ci0 + cit * 32 + li;
                out[((long long)t * a.cout + co) * cin + ci] = acc[j * 3 + dz][r];
            }
}

void wgrad_plan(int n, int h, int w, int cout, int cin, long long max_part, int* rows, int* nchunk) {
    const int tco = cout % 64 == 0 ? 64 : 32, tci = cin % 64 == 0 ? 64 : 32;
    const long long tiles = (long long)(cout / tco) * (cin / tci);
    const int nseg = (w + ZS - 1) / ZS;
    // workgroups for one full round of the GPU: 256 CUs x the workgroups of this tile shape that fit on a CU
    // (64x64: 12 waves, one per CU; 64x32 and 32x64: 6 waves, two; 32x32: 3 waves, five)
    const long long target = tco == 64 && tci == 64 ? 256 : tco == 32 && tci == 32 ? 1280 : 512;
    const long long want = std::max<long long>(1, (target + tiles - 1) / tiles);
    long long per = std::max<long long>(1, (want + (long long)n * nseg - 1) / ((long long)n * nseg));
    int r = (int)std::max<long long>(1, (h + per - 1) / per);
    r = std::max(r, std::min(4, h));
    const long long per_chunk = 9LL * cout * cin + cout;
    for (;;) {
        const long long nc = (long long)n * nseg * ((h + r - 1) / r);
        if (nc * per_chunk <= max_part || r >= h) {
            *rows = r;
            *nchunk = (int)nc;
            return;
        }
        r = std::min(h, r * 2);
    }
}

void launch_wgrad(hipStream_t st, const WgradArgs& a) {
    const int cin = a.c1 + a.c2;
    const int tco = a.cout % 64 == 0 ? 64 : 32, tci = cin % 64 == 0 ? 64 : 32;
    dim3 grid((unsigned)(a.n * a.nseg * a.nxblk), (a.cout / tco) * (cin / tci));
    if (tco == 64 && tci == 64)
        hipLaunchKernelGGL((k_wgrad<64, 64, 3>), grid, dim3(768), 0, st, a);
    else if (tco == 64)
        hipLaunchKernelGGL((k_wgrad<64, 32, 3>), grid, dim3(384), 0, st, a);
    else if (tci == 64)
        hipLaunchKernelGGL((k_wgrad<32, 64, 3>), grid, dim3(384), 0, st, a);
    else
        hipLaunchKernelGGL((k_wgrad<32, 32, 3>), grid, dim3(192), 0, st, a);
}

__global__ void k_reduce(const float* part, int nchunk, long long len, float* out, int accumulate) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(long long)c * len + e];
    out[e] = accumulate ? out[e] + s : s;
}

// Level 1 of a two-level reduction, in place: part[g][e] = sum over chunks
// c = g, g + G, g + 2G, ... of part[c][e] (set g is read only by block row g).
__global__ void k_reduce_strided(float* part, int nchunk, int G, long long len) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;
    if (e >= len) return;
    float s = 0.f;
    for (int c = g; c < nchunk; c += G) s += part[(long long)c * len + e];
    part[(long long)g * len + e] = s;
}

// Few elements, many chunks: one block per element, fixed-order tree sum.
__global__ __launch_bounds__(256) void k_reduce_cols(const float* part, int nchunk, long long len, float* out,
                                                     int accumulate) {
    __shared__ float red[256];
    const long long e = blockIdx.x;
    float s = 0.f;
    for (int c = threadIdx.x; c < nchunk; c += 256) s += part[(long long)c * len + e];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[e] = accumulate ? out[e] + red[0] : red[0];
}

// The same two levels on 4 consecutive elements per thread (len a multiple of 4,
// 16-byte aligned rows and output), with the chunk grouping of the scalar form,
// so every element is summed in the same order.
__global__ void k_reduce4(const float4* part, int nchunk, long long len4, float4* out, int accumulate) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len4) return;
    float4 s = {0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < nchunk; ++c) {
        const float4 v = part[(long long)c * len4 + e];
        s.x += v.x, s.y += v.y, s.z += v.z, s.w += v.w;
    }
    if (accumulate) {
        const float4 o = out[e];
        s = float4{o.x + s.x, o.y + s.y, o.z + s.z, o.w + s.w};
    }
    out[e] = s;
}

__global__ void k_reduce_strided4(float4* part, int nchunk, int G, long long len4) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;
    if (e >= len4) return;
    float4 s = {0.f, 0.f, 0.f, 0.f};
    for (int c = g; c < nchunk; c += G) {
        const float4 v = part[(long long)c * len4 + e];
        s.x += v.x, s.y += v.y, s.z += v.z, s.w += v.w;
    }
    part[(long long)g * len4 + e] = s;
}

void launch_reduce(hipStream_t st, float* part, int nchunk, long long len, float* out, int accumulate) {
    if (len % 4 == 0 && len > 4096 && ((uintptr_t)part & 15) == 0 && ((uintptr_t)out & 15) == 0) {
        const long long eb = (len + 255) / 256, len4 = len / 4, eb4 = (len4 + 255) / 256;
        const int G = (int)std::min<long long>(nchunk, std::max<long long>(1, (2048 + eb - 1) / eb));
        if (G > 1 && nchunk >= 4 * G) {
            hipLaunchKernelGGL(k_reduce_strided4, dim3((unsigned)eb4, G), dim3(256), 0, st,
                               reinterpret_cast<float4*>(part), nchunk, G, len4);
            nchunk = G;
        }
        hipLaunchKernelGGL(k_reduce4, dim3((unsigned)eb4), dim3(256), 0, st, reinterpret_cast<const float4*>(part),
                           nchunk, len4, reinterpret_cast<float4*>(out), accumulate);
        return;
    }
    if (len <= 4096 && nchunk >= 64) {
        hipLaunchKernelGGL(k_reduce_cols, dim3((unsigned)len), dim3(256), 0, st, part, nchunk, len, out, accumulate);
        return;
    }
    const long long eb = (len + 255) / 256;
    // enough blocks to keep every CU busy: split the chunk loop when the element count is small
    const int G = (int)std::min<long long>(nchunk, std::max<long long>(1, (2048 + eb - 1) / eb));
    if (G > 1 && nchunk >= 4 * G) {
        hipLaunchKernelGGL(k_reduce_strided, dim3((unsigned)eb, G), dim3(256), 0, st, part, nchunk,
                           G, len);
        nchunk = G;
    }
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)eb), dim3(256), 0, st, part, nchunk, len, out, accumulate);
}

// ---------------------------------------------------------------------------
// First layer (one fp32 input channel), biases, head
// ---------------------------------------------------------------------------
__device__ __forceinline__ float tap_in(const float* x, int n, int h, int w, int ix, int iz, int t) {
    const int xs = ix + t / 3 - 1, zs = iz + t % 3 - 1;
    return (xs >= 0 && xs < h && zs >= 0 && zs < w) ? x[((long long)n * h + xs) * w + zs] : 0.f;
}

// y[p][co] = relu(sum_t x[p + off(t)] wt[t][co] + b[co]); one thread per
// (pixel, 8 output channels), so consecutive lanes store consecutive 16 bytes.
__global__ __launch_bounds__(256) void k_conv1(const float* x, const float* wt, const float* b, bf16_t* y, int n,
                                               int h, int w, int cout) {
    __shared__ float ws[10][256];  // taps 0..8, then the bias
    for (int i = threadIdx.x; i < 10 * cout; i += 256) ws[i / cout][i % cout] = i < 9 * cout ? wt[i] : b[i - 9 * cout];
    __syncthreads();
    const int G = cout / 8;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long npix = (long long)n * h * w;
    if (e >= npix * G) return;
    const long long p = e / G;
    const int c0 = 8 * (int)(e % G);
    const int iz = (int)(p % w), ix = (int)((p / w) % h), in = (int)(p / ((long long)w * h));
    float v[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) v[t] = tap_in(x, in, h, w, ix, iz, t);
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float s2[2];
#pragma unroll
        for (int e2 = 0; e2 < 2; ++e2) {
            const int co = c0 + 2 * k + e2;
            float acc = 0.f;
#pragma unroll
            for (int t = 0; t < 9; ++t) acc += v[t] * ws[t][co];
            s2[e2] = fmaxf(acc + ws[9][co], 0.f);
        }
        o[k] = (unsigned)f2bf(s2[0]) | ((unsigned)f2bf(s2[1]) << 16);
    }
    *reinterpret_cast<u32x4*>(y + p * cout + c0) = o;
}

void launch_conv1(hipStream_t st, const float* x, const float* w, const float* b, bf16_t* y, int n, int h, int wd,
                  int cout) {
    const long long m = (long long)n * h * wd * (cout / 8);
    hipLaunchKernelGGL(k_conv1, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, x, w, b, y, n, h, wd, cout);
}

static int reduce_blocks(long long npix) { return (int)std::min<long long>(1024, (npix + 255) / 256); }
int conv1_wgrad_blocks(long long npix) { return reduce_blocks(npix); }
int head_blocks(long long npix) { return reduce_blocks(npix); }

// part[blk][t * cout + co] = sum_{p in blk} dy[p, co] * x[p + off(t)] for the
// taps t < 9 and sum_{p in blk} dy[p, co] at t = 9; cout a multiple of 8
// dividing 2048.  Each thread reads 8 channels of a pixel (one
// 16-byte load) per pass; the per-thread sums are combined in a fixed order.
__global__ __launch_bounds__(256) void k_conv1_wgrad(const float* x, const bf16_t* dy, float* part, int n, int h,
                                                     int w, int cout) {
    __shared__ float red[256 * 8];
    const long long npix = (long long)n * h * w;
    const long long per = (npix + gridDim.x - 1) / gridDim.x;
    const long long p0 = blockIdx.x * per, p1 = std::min(npix, p0 + per);
    const int ng = cout / 8, g = threadIdx.x % ng, sub = threadIdx.x / ng, nsub = 256 / ng;
    float acc[10][8];  // taps 0..8, then the bias
#pragma unroll
    for (int t = 0; t < 10; ++t)
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[t][k] = 0.f;
    // pixel coordinates of p, advanced by nsub per pass without divisions
    const long long q0 = p0 + sub;
    int iz = (int)(q0 % w), ix = (int)((q0 / w) % h), in = (int)(q0 / ((long long)w * h));
    for (long long p = q0; p < p1; p += nsub) {
        const u32x4 v = ld16(dy + p * cout + 8 * g);
        float gv[8];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gv[2 * k] = __uint_as_float(v[k] << 16);
            gv[2 * k + 1] = __uint_as_float(v[k] & 0xFFFF0000u);
        }
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const float xt = tap_in(x, in, h, w, ix, iz, t);
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[t][k] += gv[k] * xt;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[9][k] += gv[k];
        iz += nsub;
        while (iz >= w) {
            iz -= w;
            if (++ix == h) ix = 0, ++in;
        }
    }
    float* out = part + (long long)blockIdx.x * 10 * cout;
#pragma unroll
    for (int t = 0; t < 10; ++t) {
#pragma unroll
        for (int k = 0; k < 8; ++k) red[threadIdx.x * 8 + k] = acc[t][k];
        __syncthreads();
        for (int c = threadIdx.x; c < cout; c += 256) {
            float s = 0.f;
            for (int j = 0; j < nsub; ++j) s += red[(j * ng + c / 8) * 8 + c % 8];
            out[t * cout + c] = s;
        }
        __syncthreads();
    }
}

void launch_conv1_wgrad(hipStream_t st, const float* x, const bf16_t* dy, float* part, int n, int h, int wd,
                        int cout) {
    const long long npix = (long long)n * h * wd;
    hipLaunchKernelGGL(k_conv1_wgrad, dim3(conv1_wgrad_blocks(npix)), dim3(256), 0, st, x, dy, part, n, h, wd, cout);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// C0 channels (multiple of 8); 256 threads, one pixel per thread per pass
template <int c0>
__global__ __launch_bounds__(256) void k_head(const bf16_t* a, const float* wh, const float* bh, float* y,
                                              const float* t, bf16_t* dz, float* part, double* part_loss,
                                              long long npix) {
    __shared__ float red[4][c0 + 1];
    __shared__ double redl[4];
    const long long per = (npix + gridDim.x - 1) / gridDim.x;
    const long long p0 = blockIdx.x * per, p1 = std::min(npix, p0 + per);
    float gw[c0 + 1];
#pragma unroll
    for (int c = 0; c <= c0; ++c) gw[c] = 0.f;
    double ls = 0.0;
    for (long long p = p0 + threadIdx.x; p < p1; p += 256) {
        float av[c0];
#pragma unroll
        for (int c = 0; c < c0; c += 8) {
            const u32x4 v = ld16(a + p * c0 + c);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                av[c + 2 * k] = bf2f((bf16_t)(v[k] & 0xFFFFu));
                av[c + 2 * k + 1] = bf2f((bf16_t)(v[k] >> 16));
            }
        }
        float s = 0.f;
#pragma unroll
        for (int c = 0; c < c0; ++c) s += av[c] * wh[c];
        s += bh[0];
        y[p] = s;
        if (t) {
            const float diff = s - t[p];
            ls += (double)diff * diff;
            const float g = 2.f * diff / (float)npix;
#pragma unroll
            for (int c = 0; c < c0; c += 8) {
                u32x4 o;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int ca = c + 2 * k, cb = ca + 1;
                    const float za = av[ca] > 0.f ? g * wh[ca] : 0.f;
                    const float zb = av[cb] > 0.f ? g * wh[cb] : 0.f;
                    o[k] = (unsigned)f2bf(za) | ((unsigned)f2bf(zb) << 16);
                }
                *reinterpret_cast<u32x4*>(dz + p * c0 + c) = o;
            }
#pragma unroll
            for (int c = 0; c < c0; ++c) gw[c] += g * av[c];
            gw[c0] += g;
        }
    }
    if (!t) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c <= c0; ++c) {
        const float v = wave_sum(gw[c]);
        if (lane == 0) red[wave][c] = v;
    }
    const double lv = wave_sum_d(ls);
    if (lane == 0) redl[wave] = lv;
    __syncthreads();
    if (threadIdx.x <= c0)
        part[(long long)blockIdx.x * (c0 + 1) + threadIdx.x] =
            red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
    if (threadIdx.x == 0) part_loss[blockIdx.x] = redl[0] + redl[1] + redl[2] + redl[3];
}

void launch_head(hipStream_t st, const bf16_t* a, const float* wh, const float* bh, float* y, const float* t,
                 bf16_t* dz, float* part, double* part_loss, long long npix, int c0) {
    const dim3 g(head_blocks(npix));
    switch (c0) {
        case 8: hipLaunchKernelGGL(k_head<8>, g, dim3(256), 0, st, a, wh, bh, y, t, dz, part, part_loss, npix); break;
        case 16: hipLaunchKernelGGL(k_head<16>, g, dim3(256), 0, st, a, wh, bh, y, t, dz, part, part_loss, npix); break;
        case 32: hipLaunchKernelGGL(k_head<32>, g, dim3(256), 0, st, a, wh, bh, y, t, dz, part, part_loss, npix); break;
        default: hipLaunchKernelGGL(k_head<64>, g, dim3(256), 0, st, a, wh, bh, y, t, dz, part, part_loss, npix); break;
    }
}

// ---------------------------------------------------------------------------
// Pooling, upsampling gradients, weight conversion, Adam
// ---------------------------------------------------------------------------
// y[n][x][z][c] = bf16(mean of the 2x2 block), 8 channels per thread
__global__ void k_pool(const bf16_t* x, bf16_t* y, int n, int h, int w, int c) {
    const int ho = h / 2, wo = w / 2, G = c / 8;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * ho * wo * G) return;
    const int g = (int)(e % G);
    const long long q = e / G;
    const int oz = (int)(q % wo), ox = (int)((q / wo) % ho), in = (int)(q / ((long long)wo * ho));
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int da = 0; da < 2; ++da)
        for (int db = 0; db < 2; ++db) {
            const u32x4 v = ld16(x + (((long long)in * h + 2 * ox + da) * w + 2 * oz + db) * c + g * 8);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                s[2 * k] += bf2f((bf16_t)(v[k] & 0xFFFFu));
                s[2 * k + 1] += bf2f((bf16_t)(v[k] >> 16));
            }
        }
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = (unsigned)f2bf(0.25f * s[2 * k]) | ((unsigned)f2bf(0.25f * s[2 * k + 1]) << 16);
    *reinterpret_cast<u32x4*>(y + q * c + g * 8) = o;
}

void launch_pool(hipStream_t st, const bf16_t* x, bf16_t* y, int n, int h, int w, int c) {
    const long long m = (long long)n * (h / 2) * (w / 2) * (c / 8);
    hipLaunchKernelGGL(k_pool, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, x, y, n, h, w, c);
}

// 8 consecutive floats (two 16-byte accesses)
__device__ __forceinline__ void ld8f(const float* p, float v[8]) {
    const float4 a = reinterpret_cast<const float4*>(p)[0], b = reinterpret_cast<const float4*>(p)[1];
    v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
}
// dz = bf16(act > 0 ? v : 0) for 8 channels, one 16-byte load of act and one store of dz
__device__ __forceinline__ void masked_store8(const bf16_t* act, bf16_t* dz, const float v[8]) {
    const u32x4 m = ld16(act);
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float za = __uint_as_float(m[k] << 16) > 0.f ? v[2 * k] : 0.f;
        const float zb = __uint_as_float(m[k] & 0xFFFF0000u) > 0.f ? v[2 * k + 1] : 0.f;
        o[k] = (unsigned)f2bf(za) | ((unsigned)f2bf(zb) << 16);
    }
    *reinterpret_cast<u32x4*>(dz) = o;
}

// dz[p][ch] = bf16(act > 0 ? dp[pool(p)][ch] / 4 + dskip[p][ch] : 0); 8 channels per thread
__global__ void k_combine_pool(const float* dp, const float* dskip, const bf16_t* act, bf16_t* dz, int n, int h,
                               int w, int c) {
    const int G = c / 8;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * h * w * G) return;
    const int ch = 8 * (int)(e % G);
    const long long p = e / G;
    const int iz = (int)(p % w), ix = (int)((p / w) % h), in = (int)(p / ((long long)w * h));
    const long long q = ((long long)in * (h / 2) + ix / 2) * (w / 2) + iz / 2;
    float v[8], sk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    ld8f(dp + q * c + ch, v);
    if (dskip) ld8f(dskip + p * c + ch, sk);
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = v[k] * 0.25f + sk[k];
    masked_store8(act + p * c + ch, dz + p * c + ch, v);
}

void launch_combine_pool(hipStream_t st, const float* dp, const float* dskip, const bf16_t* act, bf16_t* dz, int n,
                         int h, int w, int c) {
    const long long m = (long long)n * h * w * (c / 8);
    hipLaunchKernelGGL(k_combine_pool, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, dp, dskip, act, dz, n, h,
                       w, c);
}

// dz[p][ch] = bf16(act > 0 ? sum of dup over the 2x2 block above p : 0); 8 channels per thread
__global__ void k_combine_up(const float* dup, const bf16_t* act, bf16_t* dz, int n, int h, int w, int c) {
    const int G = c / 8;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * h * w * G) return;
    const int ch = 8 * (int)(e % G);
    const long long p = e / G;
    const int iz = (int)(p % w), ix = (int)((p / w) % h), in = (int)(p / ((long long)w * h));
    const int H = 2 * h, W = 2 * w;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int da = 0; da < 2; ++da)
        for (int db = 0; db < 2; ++db) {
            float u[8];
            ld8f(dup + (((long long)in * H + 2 * ix + da) * W + 2 * iz + db) * c + ch, u);
#pragma unroll
            for (int k = 0; k < 8; ++k) s[k] += u[k];
        }
    masked_store8(act + p * c + ch, dz + p * c + ch, s);
}

void launch_combine_up(hipStream_t st, const float* dup, const bf16_t* act, bf16_t* dz, int n, int h, int w, int c) {
    const long long m = (long long)n * h * w * (c / 8);
    hipLaunchKernelGGL(k_combine_up, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, dup, act, dz, n, h, w, c);
}

__global__ void k_prep_weights(const float* w, bf16_t* wf, bf16_t* wb, int cout, int cin) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)cout * cin;
    if (e >= 9 * per) return;
    const int t = (int)(e / per);
    const int co = (int)((e % per) / cin), ci = (int)(e % cin);
    const bf16_t v = f2bf(w[e]);
    wf[e] = v;
    wb[((long long)(8 - t) * cin + ci) * cout + co] = v;
}

void launch_prep_weights(hipStream_t st, const float* w, bf16_t* wf, bf16_t* wb, int cout, int cin) {
    const long long m = 9LL * cout * cin;
    hipLaunchKernelGGL(k_prep_weights, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, w, wf, wb, cout, cin);
}

__global__ void k_prep_weights_list(PrepList pl) {
    const PrepList::Job& J = pl.job[blockIdx.y];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)J.cout * J.cin;
    if (e >= 9 * per) return;
    const int t = (int)(e / per);
    const int co = (int)((e % per) / J.cin), ci = (int)(e % J.cin);
    const bf16_t v = f2bf(J.w[e]);
    J.wf[e] = v;
    J.wb[((long long)(8 - t) * J.cin + ci) * J.cout + co] = v;
}

void launch_prep_weights_list(hipStream_t st, const PrepList& pl) {
    long long m = 0;
    for (int i = 0; i < pl.n; ++i) m = std::max(m, 9LL * pl.job[i].cout * pl.job[i].cin);
    if (pl.n > 0)
        hipLaunchKernelGGL(k_prep_weights_list, dim3((unsigned)((m + 255) / 256), pl.n), dim3(256), 0, st, pl);
}

__global__ void k_to_bf16(const float* x, bf16_t* y, long long n) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) y[e] = f2bf(x[e]);
}
__global__ void k_from_bf16(const bf16_t* x, float* y, long long n) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) y[e] = bf2f(x[e]);
}
void launch_to_bf16(hipStream_t st, const float* x, bf16_t* y, long long n) {
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, y, n);
}
void launch_from_bf16(hipStream_t st, const bf16_t* x, float* y, long long n) {
    hipLaunchKernelGGL(k_from_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, y, n);
}

__global__ void k_gather_images(const float* sx, const float* stg, float* x, float* t, GatherIdx idx, long long img) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (e >= img) return;
    const long long src = (long long)idx.j[b] * img + e, dst = (long long)b * img + e;
    x[dst] = sx[src];
    t[dst] = stg[src];
}

void launch_gather_images(hipStream_t st, const float* src_x, const float* src_t, float* x, float* t,
                          const GatherIdx& idx, int batch, long long img) {
    hipLaunchKernelGGL(k_gather_images, dim3((unsigned)((img + 255) / 256), batch), dim3(256), 0, st, src_x, src_t, x,
                       t, idx, img);
}

__global__ void k_adam(float* th, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    th[e] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

void launch_adam(hipStream_t st, float* theta, const float* g, float* m, float* v, long long n, float lr, float b1,
                 float b2, float eps, float bc1, float bc2) {
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, theta, g, m, v, n, lr, b1, b2, eps,
                       bc1, bc2);
}

__global__ void k_sum_double(const double* part, int n, double* out, double scale) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = scale * red[0];
}

void launch_sum_double(hipStream_t st, const double* part, int n, double* out, double scale) {
    hipLaunchKernelGGL(k_sum_double, dim3(1), dim3(256), 0, st, part, n, out, scale);
}

}  // namespace unet
}  // namespace ds2d
