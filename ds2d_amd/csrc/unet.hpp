// Learned LSRTM preconditioner: BF16 MFMA convolution kernels and the U-Net
// that stands in for the inverse-Hessian deblur of the LSRTM gradient.
//
// Tensors are NHWC (H = x, W = z of an [ix][iz] image, C innermost) in
// bfloat16 (raw uint16 storage); accumulation is FP32 on the MFMA units.
// 3x3 weights are [tap][Cout][Cin], tap = (dx + 1) * 3 + (dz + 1), padding 1.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ds2d {
namespace unet {

typedef uint16_t bf16_t;

// Implicit-GEMM 3x3 convolution, one output pixel per GEMM row:
//   y[p, co] = act(sum_{tap, ci} in[p + off(tap), ci] * w[tap, co, ci] + bias[co]) * mask
// The input is concat(x1 [C1], up2^u(x2) [C2]) along channels, where x2 is
// read at (x >> u, z >> u) (nearest upsampling fused into the load).  The
// same kernel is the data gradient when fed dy and the flipped-transposed
// weights.  Output channels [0, split) go to y1, [split, cout) to y2.
struct ConvArgs {
    const bf16_t* x1;
    const bf16_t* x2;
    int c1, c2, up;     // up: x2 is at (h >> up, w >> up)
    const bf16_t* wt;   // [9][cout][c1 + c2]
    const float* bias;  // [cout] or null
    const bf16_t* mask; // [npix][cout] or null: output *= (mask > 0)
    void* y1;
    void* y2;
    int split;          // channels of y1 (= cout when y2 is unused)
    int out_bf16;       // 1: bf16 outputs, 0: fp32 outputs
    int relu;
    int n, h, w, cout;
};

// Weight gradient: part[chunk][tap][co][ci] = sum over the chunk's pixels of
// dy[p, co] * in[p + off(tap), ci]  (in as in ConvArgs).  Chunks are
// (image, 64-wide z segment, block of `rows` x rows).  With bias set, each
// chunk's record is followed by the bias partials part[chunk][9 cout cin + co]
// = sum over the chunk's pixels of dy[p, co] (the layout of a layer's weights
// followed by its biases, so one reduction yields both gradients).
struct WgradArgs {
    const bf16_t* x1;
    const bf16_t* x2;
    int c1, c2, up;
    const bf16_t* dy;   // [npix][cout]
    float* part;        // [nchunk][9 cout (c1 + c2) (+ cout with bias)]
    int n, h, w, cout;
    int rows, nseg, nxblk;
    int bias;
};

void launch_conv(hipStream_t st, const ConvArgs& a);
// Number of chunks and rows per chunk for a weight-gradient launch whose
// partial buffer may hold at most max_part floats.
void wgrad_plan(int n, int h, int w, int cout, int cin, long long max_part, int* rows, int* nchunk);
void launch_wgrad(hipStream_t st, const WgradArgs& a);
// out[e] = sum_c part[c * len + e] in a fixed order (two levels when the
// element count is small; the first level works in place, so part is
// clobbered), optionally += into out.
void launch_reduce(hipStream_t st, float* part, int nchunk, long long len, float* out, int accumulate);

// First layer (one input channel, fp32 image): y = relu(conv(x) + b) -> bf16 [npix][cout]
void launch_conv1(hipStream_t st, const float* x, const float* w, const float* b, bf16_t* y, int n, int h, int wd,
                  int cout);
// First-layer weight and bias gradient partials: part[blk][9 * cout + cout]
int conv1_wgrad_blocks(long long npix);
void launch_conv1_wgrad(hipStream_t st, const float* x, const bf16_t* dy, float* part, int n, int h, int wd,
                        int cout);

// Head: y[p] = sum_c a[p, c] * wh[c] + bh.  With a target t: per-block partials
// of the loss sum (double, part_loss[blk]) and of the head gradients
// (part[blk][c0 + 1]), and dz[p, c] = bf16(dy_p * wh[c] * (a[p, c] > 0)) with
// dy_p = 2 (y_p - t_p) / npix.
int head_blocks(long long npix);
void launch_head(hipStream_t st, const bf16_t* a, const float* wh, const float* bh, float* y, const float* t,
                 bf16_t* dz, float* part, double* part_loss, long long npix, int c0);

void launch_pool(hipStream_t st, const bf16_t* x, bf16_t* y, int n, int h, int w, int c);
// dz[p, c] = bf16((dp[p/2, c] / 4 + dskip[p, c]) * (act[p, c] > 0))  at (n, h, w)
void launch_combine_pool(hipStream_t st, const float* dp, const float* dskip, const bf16_t* act, bf16_t* dz, int n,
                         int h, int w, int c);
// dz[q, c] = bf16(sum_{2x2 children} dup[., c] * (act[q, c] > 0)); q at (n, h, w), dup at (2h, 2w)
void launch_combine_up(hipStream_t st, const float* dup, const bf16_t* act, bf16_t* dz, int n, int h, int w, int c);
// bf16 forward weights [9][co][ci] and flipped-transposed data-gradient weights [9][ci][co]
void launch_prep_weights(hipStream_t st, const float* w, bf16_t* wf, bf16_t* wb, int cout, int cin);
// The same for up to PREP_MAX layers in one launch (grid row = layer).
constexpr int PREP_MAX = 24;
struct PrepList {
    struct Job {
        const float* w;
        bf16_t* wf;
        bf16_t* wb;
        int cout, cin;
    } job[PREP_MAX];
    int n;
};
void launch_prep_weights_list(hipStream_t st, const PrepList& pl);
void launch_to_bf16(hipStream_t st, const float* x, bf16_t* y, long long n);
void launch_from_bf16(hipStream_t st, const bf16_t* x, float* y, long long n);
// Training batch assembly: x[b] = src_x[idx.j[b]], t[b] = src_t[idx.j[b]] for
// b < batch (images of img floats), in one launch.
constexpr int GATHER_MAX = 64;
struct GatherIdx {
    int j[GATHER_MAX];
};
void launch_gather_images(hipStream_t st, const float* src_x, const float* src_t, float* x, float* t,
                          const GatherIdx& idx, int batch, long long img);
void launch_adam(hipStream_t st, float* theta, const float* g, float* m, float* v, long long n, float lr, float b1,
                 float b2, float eps, float bc1, float bc2);
// out[0] = scale * sum_i part[i] (one block, fixed order)
void launch_sum_double(hipStream_t st, const double* part, int n, double* out, double scale);

}  // namespace unet
}  // namespace ds2d

// Host-side hooks of the U-Net handle used by the learned LSRTM solver.
struct ds2d_unet;
int ds2d_internal_fail(int code, const char* msg);
// y = U(x) for n images of h x w (device pointers, FP32), on stream st.
int unet_infer_dev(ds2d_unet* u, int n, int h, int w, const float* x, float* y, hipStream_t st);
void unet_limits(const ds2d_unet* u, int* levels, int* max_n, int* max_h, int* max_w);
