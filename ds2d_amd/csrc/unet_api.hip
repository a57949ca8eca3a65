// Host side of the learned LSRTM preconditioner: U-Net handle, forward,
// backward (hand-written BF16 MFMA kernels of unet_kernels.hip), Adam
// training on a resident data set, and a single-kernel test hook.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ds2d.h"
#include "unet.hpp"

using namespace ds2d::unet;

namespace {

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return ds2d_internal_fail(code, buf);
}

#define HIP_TRY(x)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(DS2D_EDEVICE, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                       \
    } while (0)

constexpr long long MAX_PART = 32LL << 20;  // floats of split-K partial sums (128 MiB)

struct Layer {
    int cin, cout, k;         // k = 3 (3x3) or 1 (head)
    long long w_off, b_off;   // into theta
    long long wf_off, wb_off; // into the bf16 weight buffer (3x3 layers with cin > 1)
};

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct ds2d_unet {
    ds2d_unet_cfg cfg{};
    int L = 0, C0 = 0;
    hipStream_t st = nullptr;
    std::vector<Layer> layers;
    std::vector<int> ch;  // channels per level
    long long np = 0, nwbf = 0;
    float *theta = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    bf16_t* wbf = nullptr;
    bool wdirty = true;
    std::vector<bf16_t*> ea, eb, pool, da, db;
    std::vector<float*> dskip, dup, dp;
    bf16_t *dzA = nullptr, *dzB = nullptr;
    float *x = nullptr, *t = nullptr, *y = nullptr;
    float* part = nullptr;
    double* part_loss = nullptr;
    double* d_loss = nullptr;  // per-step losses of a training run
    int loss_cap = 0;
    float *dset_x = nullptr, *dset_t = nullptr;
    int ds_n = 0, ds_h = 0, ds_w = 0;
    long long adam_t = 0;
    hipEvent_t ev[2]{};
    std::vector<void*> allocs;

    ~ds2d_unet() {
        for (void* p : allocs)
            if (p) (void)hipFree(p);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
    template <typename T>
    int alloc(T** p, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) return fail(DS2D_ENOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
        allocs.push_back(q);
        *p = (T*)q;
        return 0;
    }
    int enc(int l, int b) const { return 2 * l + b; }
    int dec(int l, int b) const { return 2 * L + 2 * (L - 2 - l) + b; }
    int head() const { return 2 * L + 2 * (L - 1); }
};

namespace {

double layer_flops(const ds2d_unet* u, int n, int h, int w) {
    double f = 0;
    for (int l = 0; l < u->L; ++l) {
        const double np = (double)n * (h >> l) * (w >> l);
        for (int b = 0; b < 2; ++b) {
            const Layer& q = u->layers[u->enc(l, b)];
            f += 2.0 * 9 * q.cin * q.cout * np;
        }
        if (l < u->L - 1)
            for (int b = 0; b < 2; ++b) {
                const Layer& q = u->layers[u->dec(l, b)];
                f += 2.0 * 9 * q.cin * q.cout * np;
            }
    }
    f += 2.0 * u->C0 * n * h * w;
    return f;
}

int check_dims(const ds2d_unet* u, int n, int h, int w) {
    const int q = 1 << (u->L - 1);
    if (n <= 0 || n > u->cfg.max_batch || h <= 0 || h > u->cfg.max_h || w <= 0 || w > u->cfg.max_w)
        return fail(DS2D_EINVAL, "images %dx%dx%d exceed the handle's limits %dx%dx%d", n, h, w, u->cfg.max_batch,
                    u->cfg.max_h, u->cfg.max_w);
    if (h % q || w % q) return fail(DS2D_EINVAL, "image %dx%d is not a multiple of %d", h, w, q);
    return 0;
}

void prep_weights(ds2d_unet* u, hipStream_t st) {
    if (!u->wdirty) return;
    PrepList pl{};
    for (const Layer& q : u->layers) {
        if (q.k != 3 || q.cin <= 1) continue;
        pl.job[pl.n++] = {u->theta + q.w_off, u->wbf + q.wf_off, u->wbf + q.wb_off, q.cout, q.cin};
        if (pl.n == PREP_MAX) {
            launch_prep_weights_list(st, pl);
            pl.n = 0;
        }
    }
    launch_prep_weights_list(st, pl);
    u->wdirty = false;
}

ConvArgs conv_args(const ds2d_unet* u, const Layer& q, int n, int h, int w) {
    ConvArgs a{};
    a.wt = u->wbf + q.wf_off;
    a.bias = u->theta + q.b_off;
    a.relu = 1;
    a.out_bf16 = 1;
    a.n = n;
    a.h = h;
    a.w = w;
    a.cout = q.cout;
    a.split = q.cout;
    return a;
}

// Forward pass from u->x (n x h x w fp32) to u->y; activations stay resident.
void forward(ds2d_unet* u, int n, int h, int w, hipStream_t st) {
    prep_weights(u, st);
    const int L = u->L;
    for (int l = 0; l < L; ++l) {
        const int hl = h >> l, wl = w >> l;
        const Layer& qa = u->layers[u->enc(l, 0)];
        if (l == 0) {
            launch_conv1(st, u->x, u->theta + qa.w_off, u->theta + qa.b_off, u->ea[0], n, hl, wl, qa.cout);
        } else {
            ConvArgs a = conv_args(u, qa, n, hl, wl);
            a.x1 = u->pool[l - 1];
            a.c1 = qa.cin;
            a.y1 = u->ea[l];
            launch_conv(st, a);
        }
        const Layer& qb = u->layers[u->enc(l, 1)];
        ConvArgs b = conv_args(u, qb, n, hl, wl);
        b.x1 = u->ea[l];
        b.c1 = qb.cin;
        b.y1 = u->eb[l];
        launch_conv(st, b);
        if (l < L - 1) launch_pool(st, u->eb[l], u->pool[l], n, hl, wl, u->ch[l]);
    }
    const bf16_t* low = u->eb[L - 1];
    for (int l = L - 2; l >= 0; --l) {
        const int hl = h >> l, wl = w >> l;
        const Layer& qa = u->layers[u->dec(l, 0)];
        ConvArgs a = conv_args(u, qa, n, hl, wl);
        a.x1 = u->eb[l];
        a.c1 = u->ch[l];
        a.x2 = low;
        a.c2 = u->ch[l + 1];
        a.up = 1;
        a.y1 = u->da[l];
        launch_conv(st, a);
        const Layer& qb = u->layers[u->dec(l, 1)];
        ConvArgs b = conv_args(u, qb, n, hl, wl);
        b.x1 = u->da[l];
        b.c1 = qb.cin;
        b.y1 = u->db[l];
        launch_conv(st, b);
        low = u->db[l];
    }
    const Layer& hq = u->layers[u->head()];
    launch_head(st, low, u->theta + hq.w_off, u->theta + hq.b_off, u->y, nullptr, nullptr, nullptr, nullptr,
                (long long)n * h * w, u->C0);
}

// Weight and bias gradients of one 3x3 layer (bf16 operands), into u->grad.
void layer_grads(ds2d_unet* u, const Layer& q, const bf16_t* x1, int c1, const bf16_t* x2, int c2, const bf16_t* dz,
                 int n, int h, int w, hipStream_t st) {
    int rows, nchunk;
    wgrad_plan(n, h, w, q.cout, q.cin, MAX_PART, &rows, &nchunk);
    WgradArgs a{};
    a.x1 = x1;
    a.c1 = c1;
    a.x2 = x2;
    a.c2 = c2;
    a.up = x2 ? 1 : 0;
    a.dy = dz;
    a.part = u->part;
    a.n = n;
    a.h = h;
    a.w = w;
    a.cout = q.cout;
    a.rows = rows;
    a.nseg = (w + 63) / 64;
    a.nxblk = (h + rows - 1) / rows;
    a.bias = 1;
    launch_wgrad(st, a);
    // weights then biases: b_off follows w_off in the parameter vector
    launch_reduce(st, u->part, nchunk, 9LL * q.cout * q.cin + q.cout, u->grad + q.w_off, 0);
}

// Data gradient of one 3x3 layer: conv of dz with the flipped weights.
ConvArgs dgrad_args(const ds2d_unet* u, const Layer& q, const bf16_t* dz, int n, int h, int w) {
    ConvArgs a{};
    a.x1 = dz;
    a.c1 = q.cout;
    a.wt = u->wbf + q.wb_off;
    a.n = n;
    a.h = h;
    a.w = w;
    a.cout = q.cin;
    a.split = q.cin;
    return a;
}

// Backward pass after forward(); target in u->t.  Loss sum (double) goes to
// *loss_dev (scaled by 1/npix), gradients to u->grad.
void backward(ds2d_unet* u, int n, int h, int w, double* loss_dev, hipStream_t st) {
    const int L = u->L;
    const long long npix = (long long)n * h * w;
    const Layer& hq = u->layers[u->head()];
    const bf16_t* top = L > 1 ? u->db[0] : u->eb[0];
    bf16_t *cur = u->dzA, *nxt = u->dzB;
    launch_head(st, top, u->theta + hq.w_off, u->theta + hq.b_off, u->y, u->t, cur, u->part, u->part_loss, npix,
                u->C0);
    launch_reduce(st, u->part, head_blocks(npix), u->C0 + 1, u->grad + hq.w_off, 0);
    launch_sum_double(st, u->part_loss, head_blocks(npix), loss_dev, 1.0 / (double)npix);
    for (int l = 0; l <= L - 2; ++l) {
        const int hl = h >> l, wl = w >> l;
        const Layer& qb = u->layers[u->dec(l, 1)];
        layer_grads(u, qb, u->da[l], qb.cin, nullptr, 0, cur, n, hl, wl, st);
        ConvArgs d = dgrad_args(u, qb, cur, n, hl, wl);
        d.mask = u->da[l];
        d.out_bf16 = 1;
        d.y1 = nxt;
        launch_conv(st, d);
        std::swap(cur, nxt);
        const Layer& qa = u->layers[u->dec(l, 0)];
        const bf16_t* low = l + 1 <= L - 2 ? u->db[l + 1] : u->eb[L - 1];
        layer_grads(u, qa, u->eb[l], u->ch[l], low, u->ch[l + 1], cur, n, hl, wl, st);
        ConvArgs e = dgrad_args(u, qa, cur, n, hl, wl);
        e.out_bf16 = 0;
        e.split = u->ch[l];
        e.y1 = u->dskip[l];
        e.y2 = u->dup[l];
        launch_conv(st, e);
        launch_combine_up(st, u->dup[l], low, nxt, n, hl / 2, wl / 2, u->ch[l + 1]);
        std::swap(cur, nxt);
    }
    for (int l = L - 1; l >= 0; --l) {
        const int hl = h >> l, wl = w >> l;
        const Layer& qb = u->layers[u->enc(l, 1)];
        layer_grads(u, qb, u->ea[l], qb.cin, nullptr, 0, cur, n, hl, wl, st);
        ConvArgs d = dgrad_args(u, qb, cur, n, hl, wl);
        d.mask = u->ea[l];
        d.out_bf16 = 1;
        d.y1 = nxt;
        launch_conv(st, d);
        std::swap(cur, nxt);
        const Layer& qa = u->layers[u->enc(l, 0)];
        if (l == 0) {
            launch_conv1_wgrad(st, u->x, cur, u->part, n, hl, wl, qa.cout);
            launch_reduce(st, u->part, conv1_wgrad_blocks(npix), 10LL * qa.cout, u->grad + qa.w_off, 0);
            break;
        }
        layer_grads(u, qa, u->pool[l - 1], qa.cin, nullptr, 0, cur, n, hl, wl, st);
        ConvArgs e = dgrad_args(u, qa, cur, n, hl, wl);
        e.out_bf16 = 0;
        e.y1 = u->dp[l];
        launch_conv(st, e);
        launch_combine_pool(st, u->dp[l], u->dskip[l - 1], u->eb[l - 1], nxt, n, h >> (l - 1), w >> (l - 1),
                            u->ch[l - 1]);
        std::swap(cur, nxt);
    }
}

int sync_check(ds2d_unet* u) {
    HIP_TRY(hipStreamSynchronize(u->st));
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

int unet_infer_dev(ds2d_unet* u, int n, int h, int w, const float* x, float* y, hipStream_t st) {
    int rc;
    if ((rc = check_dims(u, n, h, w))) return rc;
    const size_t bytes = (size_t)n * h * w * 4;
    HIP_TRY(hipMemcpyAsync(u->x, x, bytes, hipMemcpyDeviceToDevice, st));
    forward(u, n, h, w, st);
    HIP_TRY(hipMemcpyAsync(y, u->y, bytes, hipMemcpyDeviceToDevice, st));
    return 0;
}

void unet_limits(const ds2d_unet* u, int* levels, int* max_n, int* max_h, int* max_w) {
    *levels = u->L;
    *max_n = u->cfg.max_batch;
    *max_h = u->cfg.max_h;
    *max_w = u->cfg.max_w;
}

extern "C" {

int ds2d_unet_create(const ds2d_unet_cfg* cfg, ds2d_unet_t** out) {
    if (!cfg || !out) return fail(DS2D_EINVAL, "null argument");
    *out = nullptr;
    const int L = cfg->levels, C0 = cfg->base;
    if (L < 1 || L > 6 || C0 <= 0 || C0 % 32 || C0 > 64 || cfg->max_batch <= 0)
        return fail(DS2D_EINVAL, "unsupported U-Net shape (levels %d, base %d)", L, C0);
    if (C0 << (L - 1) > 256) return fail(DS2D_EINVAL, "deepest level would have more than 256 channels");
    const int q = 1 << (L - 1);
    if (cfg->max_h <= 0 || cfg->max_w <= 0 || cfg->max_h % q || cfg->max_w % q)
        return fail(DS2D_EINVAL, "max image %dx%d must be a positive multiple of %d", cfg->max_h, cfg->max_w, q);
    HIP_TRY(hipSetDevice(0));
    auto* u = new ds2d_unet();
    u->cfg = *cfg;
    u->L = L;
    u->C0 = C0;
    for (int l = 0; l < L; ++l) u->ch.push_back(C0 << l);
    auto add = [&](int cin, int cout, int k) {
        Layer q{cin, cout, k, 0, 0, -1, -1};
        q.w_off = u->np;
        u->np += (long long)(k == 3 ? 9 : 1) * cin * cout;
        q.b_off = u->np;
        u->np += cout;
        if (k == 3 && cin > 1) {
            q.wf_off = u->nwbf;
            u->nwbf += 9LL * cin * cout;
            q.wb_off = u->nwbf;
            u->nwbf += 9LL * cin * cout;
        }
        u->layers.push_back(q);
    };
    for (int l = 0; l < L; ++l) {
        add(l == 0 ? 1 : u->ch[l - 1], u->ch[l], 3);
        add(u->ch[l], u->ch[l], 3);
    }
    for (int l = L - 2; l >= 0; --l) {
        add(u->ch[l] + u->ch[l + 1], u->ch[l], 3);
        add(u->ch[l], u->ch[l], 3);
    }
    add(C0, 1, 1);
    int rc = 0;
    auto bail = [&](int r) {
        delete u;
        return r;
    };
    if (hipStreamCreateWithFlags(&u->st, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(DS2D_EDEVICE, "stream creation failed"));
    for (auto& e : u->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(DS2D_EDEVICE, "event creation failed"));
    const long long N = cfg->max_batch;
    auto npix = [&](int l) { return N * (cfg->max_h >> l) * (cfg->max_w >> l); };
    if ((rc = u->alloc(&u->theta, u->np)) || (rc = u->alloc(&u->grad, u->np)) || (rc = u->alloc(&u->m, u->np)) ||
        (rc = u->alloc(&u->v, u->np)) || (rc = u->alloc(&u->wbf, u->nwbf)))
        return bail(rc);
    u->ea.resize(L);
    u->eb.resize(L);
    u->pool.resize(L);
    u->da.resize(L);
    u->db.resize(L);
    u->dskip.resize(L);
    u->dup.resize(L);
    u->dp.resize(L);
    for (int l = 0; l < L; ++l) {
        const long long e = npix(l) * u->ch[l];
        if ((rc = u->alloc(&u->ea[l], e)) || (rc = u->alloc(&u->eb[l], e))) return bail(rc);
        if (l < L - 1) {
            if ((rc = u->alloc(&u->pool[l], npix(l + 1) * u->ch[l])) || (rc = u->alloc(&u->da[l], e)) ||
                (rc = u->alloc(&u->db[l], e)) || (rc = u->alloc(&u->dskip[l], e)) ||
                (rc = u->alloc(&u->dup[l], npix(l) * u->ch[l + 1])))
                return bail(rc);
        }
        if (l >= 1 && (rc = u->alloc(&u->dp[l], npix(l) * u->ch[l - 1]))) return bail(rc);
    }
    long long dzmax = 0;
    for (int l = 0; l < L; ++l) dzmax = std::max(dzmax, npix(l) * u->ch[l]);
    if ((rc = u->alloc(&u->dzA, dzmax)) || (rc = u->alloc(&u->dzB, dzmax))) return bail(rc);
    if ((rc = u->alloc(&u->x, npix(0))) || (rc = u->alloc(&u->t, npix(0))) || (rc = u->alloc(&u->y, npix(0))))
        return bail(rc);
    // split-K partial buffer: the largest of any weight-gradient plan and the small reductions
    long long part = (long long)head_blocks(npix(0)) * (C0 + 1);
    part = std::max(part, (long long)conv1_wgrad_blocks(npix(0)) * 10 * C0);
    // a weight-gradient plan stays within MAX_PART unless it is down to one
    // chunk per (image, z segment); bound that case over every layer and level
    part = std::max(part, MAX_PART);
    for (int l = 0; l < L; ++l) {
        const long long nseg = ((cfg->max_w >> l) + 63) / 64;
        for (const Layer& q : u->layers)
            if (q.k == 3 && q.cin > 1) part = std::max(part, N * nseg * (9 * q.cout * q.cin + q.cout));
    }
    if ((rc = u->alloc(&u->part, part)) || (rc = u->alloc(&u->part_loss, head_blocks(npix(0)))) ||
        (rc = u->alloc(&u->d_loss, 1)))
        return bail(rc);
    u->loss_cap = 1;
    // He-uniform weights U(-sqrt(6/fan_in), sqrt(6/fan_in)), zero biases (SplitMix64 stream)
    std::vector<float> th(u->np, 0.f);
    uint64_t s = cfg->seed;
    for (const Layer& q : u->layers) {
        const long long nw = (long long)(q.k == 3 ? 9 : 1) * q.cin * q.cout;
        const double lim = std::sqrt(6.0 / (q.cin * (q.k == 3 ? 9 : 1)));
        for (long long i = 0; i < nw; ++i) th[q.w_off + i] = (float)((2 * uniform01(s) - 1) * lim);
    }
    if (hipMemcpy(u->theta, th.data(), u->np * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(u->m, 0, u->np * 4) != hipSuccess || hipMemset(u->v, 0, u->np * 4) != hipSuccess)
        return bail(fail(DS2D_EDEVICE, "parameter upload failed"));
    *out = u;
    return 0;
}

int ds2d_unet_destroy(ds2d_unet_t* u) {
    delete u;
    return 0;
}

int ds2d_unet_num_params(const ds2d_unet_t* u, long long* n) {
    if (!u || !n) return fail(DS2D_EINVAL, "null argument");
    *n = u->np;
    return 0;
}

int ds2d_unet_get_params(ds2d_unet_t* u, float* theta) {
    if (!u || !theta) return fail(DS2D_EINVAL, "null argument");
    HIP_TRY(hipMemcpyAsync(theta, u->theta, u->np * 4, hipMemcpyDeviceToHost, u->st));
    return sync_check(u);
}

int ds2d_unet_set_params(ds2d_unet_t* u, const float* theta) {
    if (!u || !theta) return fail(DS2D_EINVAL, "null argument");
    HIP_TRY(hipMemcpyAsync(u->theta, theta, u->np * 4, hipMemcpyHostToDevice, u->st));
    HIP_TRY(hipMemsetAsync(u->m, 0, u->np * 4, u->st));
    HIP_TRY(hipMemsetAsync(u->v, 0, u->np * 4, u->st));
    u->adam_t = 0;
    u->wdirty = true;
    return sync_check(u);
}

int ds2d_unet_forward(ds2d_unet_t* u, int n, int h, int w, const float* x, float* y, ds2d_unet_stats* stats) {
    if (!u || !x || !y) return fail(DS2D_EINVAL, "null argument");
    int rc;
    if ((rc = check_dims(u, n, h, w))) return rc;
    const size_t bytes = (size_t)n * h * w * 4;
    HIP_TRY(hipMemcpyAsync(u->x, x, bytes, hipMemcpyHostToDevice, u->st));
    HIP_TRY(hipEventRecord(u->ev[0], u->st));
    forward(u, n, h, w, u->st);
    HIP_TRY(hipEventRecord(u->ev[1], u->st));
    HIP_TRY(hipMemcpyAsync(y, u->y, bytes, hipMemcpyDeviceToHost, u->st));
    if ((rc = sync_check(u))) return rc;
    if (stats) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, u->ev[0], u->ev[1]));
        stats->elapsed_ms = ms;
        stats->ms_per_step = ms;
        stats->flops_per_step = layer_flops(u, n, h, w);
        stats->steps = 1;
    }
    return 0;
}

int ds2d_unet_loss_grad(ds2d_unet_t* u, int n, int h, int w, const float* x, const float* t, double* loss,
                        float* grad) {
    if (!u || !x || !t) return fail(DS2D_EINVAL, "null argument");
    int rc;
    if ((rc = check_dims(u, n, h, w))) return rc;
    const size_t bytes = (size_t)n * h * w * 4;
    HIP_TRY(hipMemcpyAsync(u->x, x, bytes, hipMemcpyHostToDevice, u->st));
    HIP_TRY(hipMemcpyAsync(u->t, t, bytes, hipMemcpyHostToDevice, u->st));
    forward(u, n, h, w, u->st);
    backward(u, n, h, w, u->d_loss, u->st);
    double l = 0;
    HIP_TRY(hipMemcpyAsync(&l, u->d_loss, sizeof l, hipMemcpyDeviceToHost, u->st));
    if (grad) HIP_TRY(hipMemcpyAsync(grad, u->grad, u->np * 4, hipMemcpyDeviceToHost, u->st));
    if ((rc = sync_check(u))) return rc;
    if (loss) *loss = l;
    return 0;
}

int ds2d_unet_set_data(ds2d_unet_t* u, int nsamples, int h, int w, const float* x, const float* t) {
    if (!u || !x || !t || nsamples <= 0) return fail(DS2D_EINVAL, "bad training set");
    int rc;
    if ((rc = check_dims(u, 1, h, w))) return rc;
    const size_t n = (size_t)nsamples * h * w;
    if ((rc = u->alloc(&u->dset_x, n)) || (rc = u->alloc(&u->dset_t, n))) return rc;
    HIP_TRY(hipMemcpyAsync(u->dset_x, x, n * 4, hipMemcpyHostToDevice, u->st));
    HIP_TRY(hipMemcpyAsync(u->dset_t, t, n * 4, hipMemcpyHostToDevice, u->st));
    u->ds_n = nsamples;
    u->ds_h = h;
    u->ds_w = w;
    return sync_check(u);
}

int ds2d_unet_train(ds2d_unet_t* u, int nsteps, int batch, float lr, unsigned long long seed, float* loss_hist,
                    ds2d_unet_stats* stats) {
    if (!u) return fail(DS2D_EINVAL, "null handle");
    if (!u->dset_x) return fail(DS2D_ESTATE, "no training set");
    if (nsteps <= 0 || batch <= 0 || batch > u->ds_n || !(lr > 0)) return fail(DS2D_EINVAL, "bad training arguments");
    int rc;
    const int h = u->ds_h, w = u->ds_w;
    if ((rc = check_dims(u, batch, h, w))) return rc;
    if (u->loss_cap < nsteps) {
        if ((rc = u->alloc(&u->d_loss, nsteps))) return rc;
        u->loss_cap = nsteps;
    }
    const size_t img = (size_t)h * w;
    std::vector<int> perm(u->ds_n);
    std::iota(perm.begin(), perm.end(), 0);
    uint64_t s = seed;
    int pos = u->ds_n;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    HIP_TRY(hipEventRecord(u->ev[0], u->st));
    GatherIdx gi{};
    for (int k = 0; k < nsteps; ++k) {
        for (int b = 0; b < batch; ++b) {
            if (pos >= u->ds_n) {  // new epoch: Fisher-Yates shuffle
                for (int i = u->ds_n - 1; i > 0; --i) std::swap(perm[i], perm[splitmix(s) % (uint64_t)(i + 1)]);
                pos = 0;
            }
            const int j = perm[pos++];
            if (batch <= GATHER_MAX) {
                gi.j[b] = j;
            } else {
                HIP_TRY(hipMemcpyAsync(u->x + b * img, u->dset_x + j * img, img * 4, hipMemcpyDeviceToDevice, u->st));
                HIP_TRY(hipMemcpyAsync(u->t + b * img, u->dset_t + j * img, img * 4, hipMemcpyDeviceToDevice, u->st));
            }
        }
        if (batch <= GATHER_MAX) launch_gather_images(u->st, u->dset_x, u->dset_t, u->x, u->t, gi, batch, (long long)img);
        forward(u, batch, h, w, u->st);
        backward(u, batch, h, w, u->d_loss + k, u->st);
        ++u->adam_t;
        const float bc1 = 1.f - (float)std::pow((double)b1, (double)u->adam_t);
        const float bc2 = 1.f - (float)std::pow((double)b2, (double)u->adam_t);
        launch_adam(u->st, u->theta, u->grad, u->m, u->v, u->np, lr, b1, b2, eps, bc1, bc2);
        u->wdirty = true;
    }
    HIP_TRY(hipEventRecord(u->ev[1], u->st));
    std::vector<double> lh(nsteps);
    HIP_TRY(hipMemcpyAsync(lh.data(), u->d_loss, nsteps * sizeof(double), hipMemcpyDeviceToHost, u->st));
    if ((rc = sync_check(u))) return rc;
    if (loss_hist)
        for (int k = 0; k < nsteps; ++k) loss_hist[k] = (float)lh[k];
    if (stats) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, u->ev[0], u->ev[1]));
        stats->elapsed_ms = ms;
        stats->ms_per_step = ms / nsteps;
        stats->flops_per_step = 3.0 * layer_flops(u, batch, h, w);
        stats->steps = nsteps;
    }
    return 0;
}

// Test / measurement hook: one conv kernel on host data.
int ds2d_conv3x3_bf16(int op, int n, int h, int w, int c1, int c2, int up, int cout, const float* x1,
                      const float* x2, const float* dy, const float* wt, const float* bias, int relu, int repeat,
                      float* out, double* kernel_ms) {
    if (op < 0 || op > 2 || n <= 0 || h <= 0 || w <= 0 || c1 <= 0 || c1 % 16 || c2 < 0 || c2 % 16 || cout <= 0 ||
        cout % 32 || cout > 4096 || (c1 + c2) % 32 || up < 0 || up > 1 || !x1 || !wt || !out || (c2 && !x2) ||
        (op == 2 && !dy))
        return fail(DS2D_EINVAL, "bad conv arguments");
    if (c2 && up && (h % 2 || w % 2)) return fail(DS2D_EINVAL, "upsampled input needs even h, w");
    if (op == 1 && c2 && up) return fail(DS2D_EINVAL, "data gradient hook takes one full-resolution input");
    HIP_TRY(hipSetDevice(0));
    const int cin = c1 + c2;
    const long long npix = (long long)n * h * w, npix2 = (long long)n * (h >> up) * (w >> up);
    const long long nx1 = npix * (op == 1 ? cout : c1), nx2 = c2 ? npix2 * c2 : 0, nw = 9LL * cout * cin;
    const long long nout = op == 0 ? npix * cout : (op == 1 ? npix * cin : nw);
    std::vector<void*> mem;
    auto cleanup = [&]() {
        for (void* p : mem) (void)hipFree(p);
    };
    auto dal = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 4)) != hipSuccess) return nullptr;
        mem.push_back(p);
        return p;
    };
    float* f_in = (float*)dal(std::max({nx1, nx2, nw, op == 2 ? npix * cout : 0LL}) * 4);
    bf16_t* b_x1 = (bf16_t*)dal(nx1 * 2);
    bf16_t* b_x2 = (bf16_t*)dal(nx2 * 2);
    bf16_t* b_dy = (bf16_t*)dal(op == 2 ? npix * cout * 2 : 4);
    float* f_w = (float*)dal(nw * 4);
    bf16_t* b_wf = (bf16_t*)dal(nw * 2);
    bf16_t* b_wb = (bf16_t*)dal(nw * 2);
    float* f_b = (float*)dal(cout * 4);
    float* f_out = (float*)dal(nout * 4);
    int rows = 1, nchunk = 1;
    if (op == 2) wgrad_plan(n, h, w, cout, cin, MAX_PART, &rows, &nchunk);
    float* f_part = (float*)dal(op == 2 ? (size_t)nchunk * nw * 4 : 4);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!f_in || !b_x1 || !b_x2 || !b_dy || !f_w || !b_wf || !b_wb || !f_b || !f_out || !f_part ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return fail(DS2D_ENOMEM, "conv hook allocation failed");
    }
    hipStream_t st = nullptr;
    bool ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
    auto up16 = [&](const float* h_src, bf16_t* dst, long long cnt) {
        ok = ok && hipMemcpyAsync(f_in, h_src, cnt * 4, hipMemcpyHostToDevice, st) == hipSuccess;
        if (ok) launch_to_bf16(st, f_in, dst, cnt);
        ok = ok && hipStreamSynchronize(st) == hipSuccess;
    };
    up16(x1, b_x1, nx1);
    if (c2) up16(x2, b_x2, nx2);
    if (op == 2) up16(dy, b_dy, npix * cout);
    ok = ok && hipMemcpyAsync(f_w, wt, nw * 4, hipMemcpyHostToDevice, st) == hipSuccess;
    if (ok) launch_prep_weights(st, f_w, b_wf, b_wb, cout, cin);
    ok = ok && hipMemsetAsync(f_b, 0, cout * 4, st) == hipSuccess;
    if (bias) ok = ok && hipMemcpyAsync(f_b, bias, cout * 4, hipMemcpyHostToDevice, st) == hipSuccess;
    const int reps = std::max(1, repeat);
    ConvArgs a{};
    WgradArgs g{};
    if (op == 0) {
        a.x1 = b_x1;
        a.c1 = c1;
        a.x2 = c2 ? b_x2 : nullptr;
        a.c2 = c2;
        a.up = up;
        a.wt = b_wf;
        a.bias = bias ? f_b : nullptr;
        a.relu = relu;
        a.y1 = f_out;
        a.split = cout;
        a.n = n;
        a.h = h;
        a.w = w;
        a.cout = cout;
    } else if (op == 1) {
        a.x1 = b_x1;
        a.c1 = cout;
        a.wt = b_wb;
        a.y1 = f_out;
        a.split = cin;
        a.n = n;
        a.h = h;
        a.w = w;
        a.cout = cin;
    } else {
        g.x1 = b_x1;
        g.c1 = c1;
        g.x2 = c2 ? b_x2 : nullptr;
        g.c2 = c2;
        g.up = up;
        g.dy = b_dy;
        g.part = f_part;
        g.n = n;
        g.h = h;
        g.w = w;
        g.cout = cout;
        g.rows = rows;
        g.nseg = (w + 63) / 64;
        g.nxblk = (h + rows - 1) / rows;
    }
    auto launch = [&]() {
        if (op == 2)
            launch_wgrad(st, g);
        else
            launch_conv(st, a);
    };
    if (ok) {
        launch();  // warm-up
        ok = hipEventRecord(e0, st) == hipSuccess;
        for (int r = 0; r < reps && ok; ++r) launch();
        ok = ok && hipEventRecord(e1, st) == hipSuccess;
        if (ok && op == 2) launch_reduce(st, f_part, nchunk, nw, f_out, 0);
        ok = ok && hipMemcpyAsync(out, f_out, nout * 4, hipMemcpyDeviceToHost, st) == hipSuccess;
        ok = ok && hipStreamSynchronize(st) == hipSuccess && hipGetLastError() == hipSuccess;
    }
    float ms = 0;
    if (ok && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) ok = false;
    if (kernel_ms) *kernel_ms = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    cleanup();
    return ok ? 0 : fail(DS2D_EDEVICE, "conv hook failed");
}

}  // extern "C"
