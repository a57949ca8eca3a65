// Laplacian image filters of laplace.cpp on gfx950, and their C-ABI.
//
//   ds2d_laplace_filter  laplace_filter_2d(nz, nx, order, dx, dat), laplace.cpp:82-106:
//                        (1/dx^2)(5-point Laplacian) on [2, n-2)^2, zero frame.
//   ds2d_laplace_band    laplace_filter_2d(nz, nx, dat), laplace.cpp:3-57, with its
//                        stated intent (zero frame around the image):
//                        out = Sx(in - (Sx o Sz)^4 in), S = (1/4, 1/2, 1/4).
//
// Images are [ix][iz], z contiguous.  The arithmetic order of every output is
// the reference's (left-to-right sums, no fused multiply-add, the 1/dx^2
// factor in FP64), so the filters are bit-identical to laplace.cpp compiled
// for x86-64 without FMA.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "../../include/ds2d.h"

int ds2d_internal_fail(int code, const char* msg);

namespace {

constexpr int LAP_BLOCK = 256;

// laplace.cpp:95: the float sum, then the FP64 factor, rounded to float.
__global__ __launch_bounds__(LAP_BLOCK) void k_laplace5(const float* __restrict__ in, float* __restrict__ out,
                                                         int nx, int nz, double scale) {
#pragma clang fp contract(off)
    const long long c = (long long)blockIdx.x * LAP_BLOCK + threadIdx.x;
    if (c >= (long long)nx * nz) return;
    const int ix = (int)(c / nz), iz = (int)(c - (long long)ix * nz);
    float r = 0.f;
    if (ix >= 2 && ix < nx - 2 && iz >= 2 && iz < nz - 2) {
        const float s = in[c + nz] + in[c - nz] - 2 * in[c] + in[c + 1] + in[c - 1] - 2 * in[c];
        r = (float)(scale * (double)s);
    }
    out[c] = r;
}

// One 3-tap smoothing pass S = (1/2 centre, 1/4 neighbours) along z (axis 0)
// or x (axis 1), with zero outside the image (laplace.cpp:19-29).
__global__ __launch_bounds__(LAP_BLOCK) void k_smooth(const float* __restrict__ in, float* __restrict__ out, int nx,
                                                       int nz, int axis) {
#pragma clang fp contract(off)
    const long long c = (long long)blockIdx.x * LAP_BLOCK + threadIdx.x;
    if (c >= (long long)nx * nz) return;
    const int ix = (int)(c / nz), iz = (int)(c - (long long)ix * nz);
    const float a = 0.5f, b = 0.25f;
    float lo, hi;
    if (axis == 0) {
        lo = iz > 0 ? in[c - 1] : 0.f;
        hi = iz < nz - 1 ? in[c + 1] : 0.f;
    } else {
        lo = ix > 0 ? in[c - nz] : 0.f;
        hi = ix < nx - 1 ? in[c + nz] : 0.f;
    }
    out[c] = in[c] * a + lo * b + hi * b;
}

// laplace.cpp:34-38: out = a - b.
__global__ __launch_bounds__(LAP_BLOCK) void k_diff(const float* __restrict__ a, const float* __restrict__ b,
                                                     float* __restrict__ out, long long n) {
    const long long c = (long long)blockIdx.x * LAP_BLOCK + threadIdx.x;
    if (c < n) out[c] = a[c] - b[c];
}

int grid_of(long long n) { return (int)((n + LAP_BLOCK - 1) / LAP_BLOCK); }

// The band filter's launch sequence: in -> out through the two intermediate
// images t1, t2 (each nx*nz floats).  Every launch writes all cells of its
// output before the next one reads them, so t1 and t2 need no clearing.
// `out` may equal `in`: in is last read by the difference pass, before the
// final pass writes out.
void launch_band(hipStream_t st, const float* in, float* out, float* t1, float* t2, int nx, int nz) {
    const long long n = (long long)nx * nz;
    const int g = grid_of(n);
    const float* src = in;
    for (int k = 0; k < 4; ++k) {  // laplace.cpp:19-31, n1 = 4
        hipLaunchKernelGGL(k_smooth, dim3(g), dim3(LAP_BLOCK), 0, st, src, t2, nx, nz, 0);
        hipLaunchKernelGGL(k_smooth, dim3(g), dim3(LAP_BLOCK), 0, st, t2, t1, nx, nz, 1);
        src = t1;
    }
    hipLaunchKernelGGL(k_diff, dim3(g), dim3(LAP_BLOCK), 0, st, in, t1, t2, n);  // laplace.cpp:34-38
    // laplace.cpp:41-48: the n3 = 2 loop re-reads the same tmp2, so one x pass
    hipLaunchKernelGGL(k_smooth, dim3(g), dim3(LAP_BLOCK), 0, st, t2, out, nx, nz, 1);
}

int hip_fail(hipError_t e, const char* what) {
    static thread_local char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    return ds2d_internal_fail(DS2D_EDEVICE, buf);
}

#define LAP_TRY(x)                                  \
    do {                                            \
        hipError_t e_ = (x);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #x); \
    } while (0)

// Device buffers of one allocating call, freed on every exit path.
struct Scoped {
    std::vector<void*> p;
    ~Scoped() {
        for (void* q : p)
            if (q) (void)hipFree(q);
    }
    int alloc(size_t bytes, float** out) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes ? bytes : 4);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc");
        p.push_back(q);
        *out = (float*)q;
        return 0;
    }
};

}  // namespace

// Scratch of the allocation-free band filter: the two intermediate images,
// at most LAP_SCRATCH_CELLS cells each (2 x 16 MiB).  Every GPU the library
// runs on holds its own copy until the process ends.
constexpr long long LAP_SCRATCH_CELLS = 1LL << 22;
static __device__ __align__(256) float scratch[2 * LAP_SCRATCH_CELLS];

namespace {

// Device address of `scratch` on the current device, looked up once per
// device (the array's host-side name is not a device address).
constexpr int LAP_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[LAP_MAX_DEVICES] = {};

int scratch_on_current_device(float** out) {
    int dev = -1;
    LAP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= LAP_MAX_DEVICES) return ds2d_internal_fail(DS2D_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        LAP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_dev[dev];
    return 0;
}

}  // namespace

extern "C" {

int ds2d_laplace_band_noalloc(int nz, int nx, const float* in, float* out, void* stream) {
    if (!in || !out || nz <= 0 || nx <= 0) return ds2d_internal_fail(DS2D_EINVAL, "bad laplace arguments");
    const long long n = (long long)nx * nz;
    if (n > LAP_SCRATCH_CELLS) {
        char msg[160];
        snprintf(msg, sizeof msg, "image of %lld cells exceeds the %lld-cell limit of the allocation-free filter", n,
                 LAP_SCRATCH_CELLS);
        return ds2d_internal_fail(DS2D_EINVAL, msg);
    }
    float* s = nullptr;
    int rc;
    if ((rc = scratch_on_current_device(&s))) return rc;
    // every pass writes all cells of its output, so the scratch needs no clearing
    launch_band(static_cast<hipStream_t>(stream), in, out, s, s + LAP_SCRATCH_CELLS, nx, nz);
    LAP_TRY(hipGetLastError());
    return 0;
}

int ds2d_laplace_filter(int nz, int nx, int order, float dx, float* dat) {
    (void)order;  // laplace.cpp:82 takes it and never reads it
    if (!dat || nz <= 0 || nx <= 0 || dx == 0) return ds2d_internal_fail(DS2D_EINVAL, "bad laplace arguments");
    const long long n = (long long)nx * nz;
    Scoped s;
    float *din, *dout;
    int rc;
    if ((rc = s.alloc(n * 4, &din)) || (rc = s.alloc(n * 4, &dout))) return rc;
    LAP_TRY(hipMemcpy(din, dat, n * 4, hipMemcpyHostToDevice));
    const double scale = 1.0 / (double)dx / (double)dx;
    hipLaunchKernelGGL(k_laplace5, dim3(grid_of(n)), dim3(LAP_BLOCK), 0, 0, din, dout, nx, nz, scale);
    LAP_TRY(hipGetLastError());
    LAP_TRY(hipMemcpy(dat, dout, n * 4, hipMemcpyDeviceToHost));
    return 0;
}

int ds2d_laplace_band(int nz, int nx, float* dat) {
    if (!dat || nz <= 0 || nx <= 0) return ds2d_internal_fail(DS2D_EINVAL, "bad laplace arguments");
    const long long n = (long long)nx * nz;
    Scoped s;
    float *din, *t1, *t2;
    int rc;
    if ((rc = s.alloc(n * 4, &din)) || (rc = s.alloc(n * 4, &t1)) || (rc = s.alloc(n * 4, &t2))) return rc;
    LAP_TRY(hipMemcpy(din, dat, n * 4, hipMemcpyHostToDevice));
    launch_band(0, din, din, t1, t2, nx, nz);
    LAP_TRY(hipGetLastError());
    LAP_TRY(hipMemcpy(dat, din, n * 4, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
