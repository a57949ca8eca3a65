// Host side of libds2d.so: handle, batch layout and the per-step schedules.
//
// The handle owns every device buffer (as FD2DGPU_ELASTIC does, FD2DGPU.cu:
// 1350-1801) and one HIP stream.  A batch packs many shot windows ("slots")
// into shared field arenas so each reference kernel runs once per step for
// all shots (the MPI master/worker shot loop of main.cpp:896-1036 becomes
// one launch per step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/ds2d.h"
#include "common.hpp"
#include "kernels.hpp"
#include "chunk_plan.hpp"
#include "unet.hpp"

using namespace ds2d;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(x)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(DS2D_EDEVICE, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                       \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

// Error state shared with the other host files of the library.
int ds2d_internal_fail(int code, const char* msg) {
    g_err = msg;
    return code;
}

struct ds2d_ctx {
    ds2d_geom g{};
    hipStream_t st = nullptr;
    std::vector<float> wav;
    float* d_wav = nullptr;
    float *d_vp = nullptr, *d_vs = nullptr, *d_rho = nullptr;  // global model
    float *d_ipp = nullptr, *d_ips = nullptr;                  // global image
    float* d_grad7 = nullptr;                                  // accumulated gradients
    bool have_model = false, have_image = false, have_shot = false;
    // current single shot (ds2d_set_shot)
    ds2d_shot cur{};
    std::vector<int> cur_rix;
    std::vector<float> cur_oz, cur_ox;
    std::map<std::string, DevBuf> ws;  // grow-only workspace
    hipEvent_t ev[4]{};
    ds2d_stats last{};
    int chunk_w = 0;    // fused-kernel x-chunk width (0: default)
    int chunk_t = 0;    // fused-kernel workgroup size (256/512/704/768; 0: default)
    int chunk_d = 0;    // fused-kernel prefetch depth in columns (2 or 4; 0: default 2)
    int tblock = 1;     // fused elastic forward: time steps per launch (1 or 2)
    int chunk_plan = 1; // two steps per launch: balanced chunk plan when it simulates faster (0: fixed width)
    int eb_fused = 1;   // elastic Born forward pass: one fused launch per step (0: k_eb_v + k_eb_s)
    int eb_w = 0, eb_t = 0;  // its x-chunk width and workgroup lanes (0: defaults 64, 256)
    int rot_fused = 1;  // rotation forward: one fused launch per step (0: k_rot_v + k_rot_p)
    int last_plan = 0;  // the last two-step launch list: 1 balanced, 0 fixed width
    std::vector<int> plan_key;         // batch shape of the cached two-step chunk list
    std::vector<ChunkDesc> plan_list;  // its chunk list
    int plan_list_bal = 0;             // whether that list is the balanced plan
    double stop_resid = 0;  // CG solvers stop once the relative data residual is at most this (0: never)
    int log_every = 0;      // CG solvers print the residual to stderr every this many iterations (0: never)
    int use_graph = 0;  // capture time loops into one hipGraph per loop
    int per_launch = 0; // plain Born/adjoint: shots per time loop (0: all)
    int fwd_two = 0;    // autotuned: mode 2 runs the two-kernel step instead of the fused one
    int acou_fused = -1;  // fused Born/adjoint steps: -1 auto, 0 never, 1 whenever eligible
    int src_fp64 = 1;     // reference-compat gradient: background and first-order Born field in FP64
    std::vector<int> shot_status;  // per shot of the last reference-compat gradient: 1 = left out (non-finite)
    // (cells, slots, shots per launch) -> 0 per-cell kernels, else fused x width
    std::map<std::tuple<long long, int, int>, int> acou_tuned;

    ~ds2d_ctx() {
        for (auto& kv : ws)
            if (kv.second.p) (void)hipFree(kv.second.p);
        void* ps[] = {d_wav, d_vp, d_vs, d_rho, d_ipp, d_ips, d_grad7};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }

    // Grow-only named device buffer.
    int buf(const char* name, size_t bytes, void** out) {
        DevBuf& b = ws[name];
        if (b.bytes < bytes) {
            if (b.p) (void)hipFree(b.p);
            b.p = nullptr;
            b.bytes = 0;
            hipError_t e = hipMalloc(&b.p, bytes ? bytes : 256);
            if (e != hipSuccess) {
                b.p = nullptr;
                return fail(DS2D_ENOMEM, "hipMalloc(%s, %zu bytes): %s", name, bytes, hipGetErrorString(e));
            }
            b.bytes = bytes;
        }
        *out = b.p;
        return 0;
    }
    template <typename T>
    int fbuf(const char* name, size_t n, T** out) {
        return buf(name, n * sizeof(T), (void**)out);
    }
};

namespace {

// Host image of a batch.
struct Batch {
    std::vector<SlotDesc> slots;
    std::vector<int> rix, tbase;
    BatchShape shape{};
    long long total_cells = 0, total_rec = 0, total_win = 0, total_edge = 0, total_rmap = 0, total_hist = 0;
    int total_traces = 0;
    bool any_pml = false;
};

int build_batch(ds2d_ctx* h, int nshots, const ds2d_shot* shots, Batch& B) {
    const ds2d_geom& g = h->g;
    if (nshots <= 0 || !shots) return fail(DS2D_EINVAL, "nshots must be > 0");
    const int pml = g.pml, nzp = g.allnz + 2 * pml;
    B.slots.resize(nshots);
    long long off = 0, roff = 0, woff = 0, eoff = 0, moff = 0;
    int rixo = 0, max_nxp = 0, max_ntr = 0;
    for (int s = 0; s < nshots; ++s) {
        const ds2d_shot& q = shots[s];
        if (q.nx <= 0 || q.ngx_min < 0 || q.ngx_min + q.nx > g.allnx)
            return fail(DS2D_EINVAL, "shot %d: window [%d, %d) outside model of %d columns", s, q.ngx_min,
                        q.ngx_min + q.nx, g.allnx);
        const int nxp = q.nx + 2 * pml;
        if (q.isx < NOP || q.isx >= nxp - NOP || q.isz < NOP || q.isz >= nzp - NOP)
            return fail(DS2D_EINVAL, "shot %d: source (%d,%d) outside the update region", s, q.isx, q.isz);
        if (q.ntr < 0 || (q.ntr > 0 && !q.rec_ix)) return fail(DS2D_EINVAL, "shot %d: bad receiver list", s);
        if (q.igz < NOP || q.igz >= nzp - NOP) return fail(DS2D_EINVAL, "shot %d: receiver row %d", s, q.igz);
        for (int r = 0; r < q.ntr; ++r)
            if (q.rec_ix[r] < NOP || q.rec_ix[r] >= nxp - NOP)
                return fail(DS2D_EINVAL, "shot %d: receiver %d column %d outside [%d,%d)", s, r, q.rec_ix[r], NOP,
                            nxp - NOP);
        SlotDesc& S = B.slots[s];
        std::memset(&S, 0, sizeof S);
        S.off = off;
        S.rec_off = roff;
        S.win_off = woff;
        S.edge_off = eoff;
        S.nxp = nxp;
        S.nzp = nzp;
        S.nx = q.nx;
        S.nz = g.allnz;
        S.pml = pml;
        S.col0 = q.ngx_min;
        S.isx = q.isx;
        S.isz = q.isz;
        S.igz = q.igz;
        S.ntr = q.ntr;
        S.rix_off = rixo;
        S.shot_id = q.shot_id;
        S.boundary = q.boundary;
        S.rmap_off = (int)moff;
        const long long n = (long long)nxp * nzp;
        off += (n + 63) / 64 * 64;
        roff += (long long)q.ntr * g.nt;
        woff += (long long)q.nx * g.allnz;
        if (q.boundary == 0) {
            B.any_pml = true;
            eoff += (long long)g.nt * (2LL * (2 * NOP) * q.nx + 2LL * (2 * NOP) * g.allnz);
        }
        moff += nxp;
        S.hist_off = B.total_hist;
        B.total_hist += (long long)g.nt * q.nx * g.allnz;
        B.tbase.push_back(B.total_traces);
        B.total_traces += q.ntr;
        for (int r = 0; r < q.ntr; ++r) B.rix.push_back(q.rec_ix[r]);
        rixo += q.ntr;
        max_nxp = std::max(max_nxp, nxp);
        max_ntr = std::max(max_ntr, q.ntr);
    }
    B.total_cells = off;
    B.total_rec = roff;
    B.total_win = woff;
    B.total_edge = eoff;
    B.total_rmap = moff;
    SlotDesc* d_slots;
    int* d_rix;
    int* d_tb;
    int rc;
    if ((rc = h->fbuf("slots", B.slots.size(), &d_slots))) return rc;
    if ((rc = h->fbuf("rix", std::max<size_t>(B.rix.size(), 1), &d_rix))) return rc;
    if ((rc = h->fbuf("tbase", B.tbase.size(), &d_tb))) return rc;
    HIP_TRY(hipMemcpyAsync(d_slots, B.slots.data(), B.slots.size() * sizeof(SlotDesc), hipMemcpyHostToDevice, h->st));
    if (!B.rix.empty())
        HIP_TRY(hipMemcpyAsync(d_rix, B.rix.data(), B.rix.size() * sizeof(int), hipMemcpyHostToDevice, h->st));
    HIP_TRY(hipMemcpyAsync(d_tb, B.tbase.data(), B.tbase.size() * sizeof(int), hipMemcpyHostToDevice, h->st));
    BatchShape& sh = B.shape;
    sh.slots = d_slots;
    sh.rix = d_rix;
    sh.nslots = nshots;
    sh.max_nxp = max_nxp;
    sh.nzp = nzp;
    sh.nz = g.allnz;
    sh.max_ntr = std::max(max_ntr, 1);
    sh.nt = g.nt;
    sh.total_cells = off;
    sh.dt = g.dt;
    sh.dx = g.dx;
    sh.dz = g.dz;
    return 0;
}

// Allocate (grow-only) and zero one 15-array elastic wavefield.
template <typename T>
int alloc_efield(ds2d_ctx* h, const char* tag, long long n, EFieldT<T>& f, bool zero = true) {
    T** ptrs[15] = {&f.vx, &f.vz, &f.txx, &f.tzz, &f.txz, &f.vx_x, &f.vx_z, &f.vz_x, &f.vz_z,
                    &f.txx_x, &f.txx_z, &f.tzz_x, &f.tzz_z, &f.txz_x, &f.txz_z};
    T* base;
    std::string name = std::string("ef_") + tag;
    int rc = h->fbuf(name.c_str(), (size_t)n * 15, &base);
    if (rc) return rc;
    for (int k = 0; k < 15; ++k) *ptrs[k] = base + (size_t)k * n;
    if (zero) HIP_TRY(hipMemsetAsync(base, 0, (size_t)n * 15 * sizeof(T), h->st));
    return 0;
}

int alloc_med(ds2d_ctx* h, long long n, Med& m) {
    float* base;
    int rc = h->fbuf("med", (size_t)n * 9, &base);
    if (rc) return rc;
    float** ptrs[9] = {&m.vp, &m.vs, &m.rho, &m.lam, &m.mu, &m.dampx, &m.dampz, &m.dmp, &m.dms};
    for (int k = 0; k < 9; ++k) *ptrs[k] = base + (size_t)k * n;
    return 0;
}

SetupArgs setup_args(ds2d_ctx* h, bool with_image) {
    SetupArgs a{};
    a.vp_all = h->d_vp;
    a.vs_all = h->d_vs;
    a.rho_all = h->d_rho;
    a.img_pp = with_image && h->have_image ? h->d_ipp : nullptr;
    a.img_ps = with_image && h->have_image ? h->d_ips : nullptr;
    a.allnx = h->g.allnx;
    a.allnz = h->g.allnz;
    a.pml_rr = h->g.pml_rr > 0 ? h->g.pml_rr : 1e-3f;
    a.rho_const = h->g.rho_const;
    a.rb_amp = h->g.rb_amp;
    a.rb_seed = h->g.rb_seed;
    return a;
}

double host_dot(ds2d_ctx* h, const float* x, const float* y, long long n, int* rc) {
    const int nb = 1024;
    double* d_part;
    if ((*rc = h->fbuf("dotpart", nb, &d_part))) return 0;
    launch_dot(h->st, x, y, n, d_part, nb);
    std::vector<double> part(nb);
    if (hipMemcpyAsync(part.data(), d_part, nb * sizeof(double), hipMemcpyDeviceToHost, h->st) != hipSuccess ||
        hipStreamSynchronize(h->st) != hipSuccess) {
        *rc = fail(DS2D_EDEVICE, "dot product copy failed");
        return 0;
    }
    double s = 0;
    for (double v : part) s += v;
    return s;
}

void stats_begin(ds2d_ctx* h) { (void)hipEventRecord(h->ev[0], h->st); }
void stats_loop_begin(ds2d_ctx* h) { (void)hipEventRecord(h->ev[1], h->st); }
void stats_loop_end(ds2d_ctx* h) { (void)hipEventRecord(h->ev[2], h->st); }
int stats_end(ds2d_ctx* h, ds2d_stats* out, long long cell_updates, long long launches, double alg_bytes) {
    HIP_TRY(hipEventRecord(h->ev[3], h->st));
    HIP_TRY(hipEventSynchronize(h->ev[3]));
    float a = 0, b = 0;
    HIP_TRY(hipEventElapsedTime(&a, h->ev[0], h->ev[3]));
    HIP_TRY(hipEventElapsedTime(&b, h->ev[1], h->ev[2]));
    h->last.elapsed_ms = a;
    h->last.stencil_ms = b;
    h->last.cell_updates = cell_updates;
    h->last.stencil_launches = launches;
    h->last.alg_bytes = alg_bytes;
    h->last.aux_alg_bytes = h->last.aux_kernel_avg_ms = 0;
    h->last.blocked_bytes = 0;
    if (out) *out = h->last;
    return 0;
}

int check_h(ds2d_ctx* h) {
    if (!h) return fail(DS2D_EINVAL, "null handle");
    HIP_TRY(hipSetDevice(0));
    return 0;
}

// Run body(it) for it = first, first+step, ... (n iterations) either as
// direct launches or, with use_graph, captured once into a hipGraph and
// launched as one graph.  Bodies only launch kernels (no memsets, copies,
// events or host syncs), so the capture holds kernel nodes only.
template <typename F>
int run_steps(ds2d_ctx* h, int first, int n, int step, F&& body) {
    if (!h->use_graph) {
        for (int k = 0, it = first; k < n; ++k, it += step) body(it);
        return 0;
    }
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    HIP_TRY(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
    for (int k = 0, it = first; k < n; ++k, it += step) body(it);
    HIP_TRY(hipStreamEndCapture(h->st, &g));
    hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (err == hipSuccess) err = hipGraphLaunch(e, h->st);
    if (err == hipSuccess) err = hipStreamSynchronize(h->st);
    if (e) (void)hipGraphExecDestroy(e);
    (void)hipGraphDestroy(g);
    if (err != hipSuccess) return fail(DS2D_EDEVICE, "graph launch: %s", hipGetErrorString(err));
    return 0;
}

// Event triplets around sampled steps of a two-kernel time loop: [a] k0 [b] k1 [c].
struct StepSampler {
    std::vector<hipEvent_t> ev;
    int every = 1, n = 0;
    int init(int nt, int max_samples = 48) {
        every = std::max(1, nt / max_samples);
        n = 0;
        ev.resize(3 * (nt / every + 1));
        for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
        return 0;
    }
    bool take(int it) const { return it % every == 0; }
    void mark(hipStream_t st, int k) { (void)hipEventRecord(ev[3 * n + k], st); if (k == 2) ++n; }
    // average durations of kernel 0 and kernel 1 (ms); call after a stream sync
    void averages(double* k0, double* k1) {
        double a = 0, b = 0;
        for (int i = 0; i < n; ++i) {
            float x = 0, y = 0;
            (void)hipEventElapsedTime(&x, ev[3 * i], ev[3 * i + 1]);
            (void)hipEventElapsedTime(&y, ev[3 * i + 1], ev[3 * i + 2]);
            a += x;
            b += y;
        }
        *k0 = n ? a / n : 0;
        *k1 = n ? b / n : 0;
    }
    ~StepSampler() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char* ds2d_last_error(void) { return g_err.c_str(); }

int ds2d_device_info(char* name, int len, int* n_cu, double* hbm_gb) {
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (n < 1) return fail(DS2D_EDEVICE, "no HIP device");
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, 0));
    if (name && len > 0) {
        snprintf(name, len, "%s", p.gcnArchName);
    }
    if (n_cu) *n_cu = p.multiProcessorCount;
    if (hbm_gb) *hbm_gb = (double)p.totalGlobalMem / 1e9;
    return 0;
}

int ds2d_create(const ds2d_geom* g, const float* wavelet, ds2d_t** out) {
    if (!g || !wavelet || !out) return fail(DS2D_EINVAL, "null argument");
    if (g->nop != NOP) return fail(DS2D_EINVAL, "nop must be %d (8th-order coefficients), got %d", NOP, g->nop);
    if (g->nt <= 0 || g->allnx <= 0 || g->allnz <= 0 || g->pml < NOP || g->dx <= 0 || g->dz <= 0 || g->dt <= 0)
        return fail(DS2D_EINVAL, "bad geometry");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (ndev < 1) return fail(DS2D_EDEVICE, "no HIP device");
    HIP_TRY(hipSetDevice(0));
    ds2d_ctx* h = new ds2d_ctx();
    h->g = *g;
    if (h->g.scale <= 0) h->g.scale = 1;
    h->wav.assign(wavelet, wavelet + g->nt);
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(DS2D_EDEVICE, "stream creation failed");
    }
    for (auto& e : h->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete h;
            return fail(DS2D_EDEVICE, "event creation failed");
        }
    const size_t A = (size_t)g->allnx * g->allnz;
    if (hipMalloc(&h->d_wav, g->nt * sizeof(float)) != hipSuccess || hipMalloc(&h->d_vp, A * 4) != hipSuccess ||
        hipMalloc(&h->d_vs, A * 4) != hipSuccess || hipMalloc(&h->d_rho, A * 4) != hipSuccess ||
        hipMalloc(&h->d_ipp, A * 4) != hipSuccess || hipMalloc(&h->d_ips, A * 4) != hipSuccess ||
        hipMalloc(&h->d_grad7, 7 * A * 4) != hipSuccess) {
        delete h;
        return fail(DS2D_ENOMEM, "model allocation failed");
    }
    if (hipMemcpy(h->d_wav, wavelet, g->nt * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(h->d_grad7, 0, 7 * A * 4) != hipSuccess || hipMemset(h->d_ipp, 0, A * 4) != hipSuccess ||
        hipMemset(h->d_ips, 0, A * 4) != hipSuccess) {
        delete h;
        return fail(DS2D_EDEVICE, "initialising the model buffers failed");
    }
    *out = h;
    return 0;
}

int ds2d_destroy(ds2d_t* h) {
    if (!h) return 0;
    (void)hipStreamSynchronize(h->st);
    delete h;
    return 0;
}

int ds2d_set_model(ds2d_t* h, const float* vp, const float* vs, const float* rho) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!vp) return fail(DS2D_EINVAL, "vp is required");
    const size_t A = (size_t)h->g.allnx * h->g.allnz * 4;
    HIP_TRY(hipMemcpyAsync(h->d_vp, vp, A, hipMemcpyHostToDevice, h->st));
    if (vs) HIP_TRY(hipMemcpyAsync(h->d_vs, vs, A, hipMemcpyHostToDevice, h->st));
    else HIP_TRY(hipMemsetAsync(h->d_vs, 0, A, h->st));
    if (rho) HIP_TRY(hipMemcpyAsync(h->d_rho, rho, A, hipMemcpyHostToDevice, h->st));
    else if (h->g.rho_const <= 0) return fail(DS2D_EINVAL, "rho array required when rho_const <= 0");
    HIP_TRY(hipStreamSynchronize(h->st));
    h->have_model = true;
    return 0;
}

int ds2d_set_image(ds2d_t* h, const float* pp, const float* ps) {
    int rc;
    if ((rc = check_h(h))) return rc;
    const size_t A = (size_t)h->g.allnx * h->g.allnz * 4;
    if (pp) HIP_TRY(hipMemcpyAsync(h->d_ipp, pp, A, hipMemcpyHostToDevice, h->st));
    else HIP_TRY(hipMemsetAsync(h->d_ipp, 0, A, h->st));
    if (ps) HIP_TRY(hipMemcpyAsync(h->d_ips, ps, A, hipMemcpyHostToDevice, h->st));
    else HIP_TRY(hipMemsetAsync(h->d_ips, 0, A, h->st));
    HIP_TRY(hipStreamSynchronize(h->st));
    h->have_image = true;
    return 0;
}

int ds2d_set_shot(ds2d_t* h, const ds2d_shot* s, const float* rec_z, const float* rec_x) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!s) return fail(DS2D_EINVAL, "null shot");
    h->cur = *s;
    h->cur_rix.assign(s->rec_ix, s->rec_ix + s->ntr);
    h->cur.rec_ix = h->cur_rix.data();
    const size_t nr = (size_t)s->ntr * h->g.nt;
    if (rec_z) h->cur_oz.assign(rec_z, rec_z + nr); else h->cur_oz.assign(nr, 0.f);
    if (rec_x) h->cur_ox.assign(rec_x, rec_x + nr); else h->cur_ox.assign(nr, 0.f);
    h->have_shot = true;
    return 0;
}

// ---------------------------------------------------------------------------
// Reference ELSRTM gradient, batched (FD2DGPU.cu:4184-4879 per slot)
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

// TB is the precision of the two source-side wavefields (the background and
// the first-order Born field), whose velocities drive a Born source: in FP32
// their rounding noise is differentiated by add_born_source and amplified in
// the Born fields and images (relative L2 up to 3e-4 against FP64 after a few
// hundred steps), in FP64 the images stay within 1e-5 of FP64.  The
// second-order Born field, the receiver-side fields, the strains, records and
// images are FP32.
//
// born_order > 0 runs the forward part only (born_modeling mode 1, plus mode 5
// when born_order = 2) and returns the scatter records: the first-order Born
// records Lm (order 1, FD2DGPU.cu:3460-3467 without the second field) or
// Lm + Lm2 (order 2, scatter_add of FD2DGPU.cu:3147-3152) in out_z / out_x.
template <typename TB>
int batch_gradient_impl(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* obs_z, const float* obs_x,
                        int iter, double* misfit, int born_order = 0, float* out_z = nullptr,
                        float* out_x = nullptr) {
    int rc;
    Batch B;
    if ((rc = build_batch(h, nshots, shots, B))) return rc;
    const BatchShape& sh = B.shape;
    const long long n = B.total_cells;
    hipStream_t st = h->st;
    EFieldT<TB> bg, b;
    EField b2, bk, bb;
    Med m;
    Strain sg, sb;
    const char* tg = sizeof(TB) == 8 ? "bg64" : "bg";
    const char* tb = sizeof(TB) == 8 ? "b64" : "b";
    if ((rc = alloc_efield(h, tg, n, bg)) || (rc = alloc_efield(h, tb, n, b)) ||
        (rc = alloc_efield(h, "b2", n, b2)) || (rc = alloc_efield(h, "bk", n, bk)) ||
        (rc = alloc_efield(h, "bb", n, bb)) || (rc = alloc_med(h, n, m)))
        return rc;
    float* strain;
    if ((rc = h->fbuf("strain", (size_t)n * 8, &strain))) return rc;
    HIP_TRY(hipMemsetAsync(strain, 0, (size_t)n * 8 * 4, st));
    sg = Strain{strain, strain + n, strain + 2 * n, strain + 3 * n};
    sb = Strain{strain + 4 * n, strain + 5 * n, strain + 6 * n, strain + 7 * n};
    const size_t R = std::max<long long>(B.total_rec, 1);
    float *rs, *obs;
    if ((rc = h->fbuf("recs", R * 4, &rs)) || (rc = h->fbuf("obs", R * 2, &obs))) return rc;
    float *rs_z = rs, *rs_x = rs + R, *rs2_z = rs + 2 * R, *rs2_x = rs + 3 * R, *ob_z = obs, *ob_x = obs + R;
    HIP_TRY(hipMemsetAsync(rs, 0, R * 4 * 4, st));
    if (obs_z) HIP_TRY(hipMemcpyAsync(ob_z, obs_z, B.total_rec * 4, hipMemcpyHostToDevice, st));
    if (obs_x) HIP_TRY(hipMemcpyAsync(ob_x, obs_x, B.total_rec * 4, hipMemcpyHostToDevice, st));
    const bool full = born_order == 0, second = born_order != 1, edges = full && B.any_pml;
    const long long W = std::max<long long>(B.total_win, 1);
    float* win;
    if ((rc = h->fbuf("win7", (size_t)W * 7, &win))) return rc;
    HIP_TRY(hipMemsetAsync(win, 0, (size_t)W * 7 * 4, st));
    TB *e_bvx = nullptr, *e_bvz = nullptr, *e_cvx = nullptr, *e_cvz = nullptr;
    if (edges) {
        TB* e;
        if ((rc = h->fbuf(sizeof(TB) == 8 ? "edges64" : "edges", (size_t)B.total_edge * 4, &e))) return rc;
        e_bvx = e; e_bvz = e + B.total_edge; e_cvx = e + 2 * B.total_edge; e_cvz = e + 3 * B.total_edge;
    }
    launch_setup(st, sh, setup_args(h, true), m);
    const int nt = h->g.nt;
    for (int it = 0; it < nt; ++it) {
        launch_inject_src(st, sh, bg.txx, bg.tzz, h->d_wav, it);
        launch_ref_v(st, sh, bg, m, 1.f, 0);
        launch_ref_s(st, sh, bg, m, sg, 1.f, 0);
        launch_born_src(st, sh, bg.vx, bg.vz, b, m, 1.f);
        launch_ref_v(st, sh, b, m, 1.f, 0);
        launch_ref_s(st, sh, b, m, sg, 1.f, 0);
        if (second) {
            launch_born_src(st, sh, b.vx, b.vz, b2, m, 1.f);
            launch_ref_v(st, sh, b2, m, 1.f, 0);
            launch_ref_s(st, sh, b2, m, sg, 1.f, 0);
            launch_record(st, sh, b.vz, b.vx, rs_z, rs_x, b2.vz, b2.vx, rs2_z, rs2_x, it);
        } else {
            launch_record(st, sh, b.vz, b.vx, rs_z, rs_x, (const float*)nullptr, nullptr, nullptr, nullptr, it);
        }
        if (edges) {
            launch_edge_io(st, sh, bg.vx, bg.vz, e_bvx, e_bvz, it, 1);
            launch_edge_io(st, sh, b.vx, b.vz, e_cvx, e_cvz, it, 1);
        }
    }
    // scatter_add (FD2DGPU.cu:3143-3149): s = 1*s + s2
    if (second) {
        launch_axpby(st, rs_z, rs2_z, B.total_rec, 1.f, 1.f);
        launch_axpby(st, rs_x, rs2_x, B.total_rec, 1.f, 1.f);
    }
    if (!full) {
        if (B.total_rec > 0) {
            HIP_TRY(hipMemcpyAsync(out_z, rs_z, B.total_rec * 4, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(out_x, rs_x, B.total_rec * 4, hipMemcpyDeviceToHost, st));
        }
        HIP_TRY(hipStreamSynchronize(st));
        HIP_TRY(hipGetLastError());
        return 0;
    }
    // NCC (FD2DGPU.cu:4532-4673) on the device, misfit finished on the host
    double* sums;
    if ((rc = h->fbuf("nccsum", (size_t)std::max(B.total_traces, 1) * 6, &sums))) return rc;
    int* d_tb;
    if ((rc = h->fbuf("tbase", B.tbase.size(), &d_tb))) return rc;
    launch_ncc_sums(st, sh, rs_z, rs_x, ob_z, ob_x, sums, d_tb);
    std::vector<double> hs((size_t)B.total_traces * 6);
    if (!hs.empty()) HIP_TRY(hipMemcpyAsync(hs.data(), sums, hs.size() * 8, hipMemcpyDeviceToHost, st));
    launch_ncc_apply(st, sh, rs_z, rs_x, ob_z, ob_x, sums, d_tb, iter);
    // per-shot failure detection: a shot whose residual (and so its adjoint
    // images) is not finite is left out of the misfit and the gathered images
    int* d_bad;
    if ((rc = h->fbuf("shotbad", (size_t)nshots, &d_bad))) return rc;
    HIP_TRY(hipMemsetAsync(d_bad, 0, (size_t)nshots * 4, st));
    launch_slot_nonfinite(st, sh, rs_z, (long long)R, 2, 0, d_bad);
    std::vector<int> bad(nshots, 0);
    HIP_TRY(hipMemcpyAsync(bad.data(), d_bad, (size_t)nshots * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (misfit && iter != 0) {
        for (int s = 0; s < nshots; ++s) {
            if (bad[s]) continue;
            const int ntr = shots[s].ntr;
            for (int r = 0; r < ntr; ++r) {
                const double* q = &hs[(size_t)(B.tbase[s] + r) * 6];
                double tx = (q[0] > 0 && q[1] > 0) ? q[2] / (std::sqrt(q[0]) * std::sqrt(q[1])) : 0.0;
                double tz = (q[3] > 0 && q[4] > 0) ? q[5] / (std::sqrt(q[3]) * std::sqrt(q[4])) : 0.0;
                *misfit -= (double)(1.0f / (2 * ntr)) * (tx + tz);
            }
        }
    }
    for (int it = nt - 1; it >= 0; --it) {
        if (B.any_pml) {
            launch_edge_io(st, sh, bg.vx, bg.vz, e_bvx, e_bvz, it, 0);
            launch_edge_io(st, sh, b.vx, b.vz, e_cvx, e_cvz, it, 0);
        }
        launch_ref_s(st, sh, bg, m, sg, -1.f, 1);
        launch_ref_v(st, sh, bg, m, -1.f, 1);
        launch_ref_s(st, sh, b, m, sb, -1.f, 1);
        launch_born_src(st, sh, bg.vx, bg.vz, b, m, 1.f);
        launch_ref_v(st, sh, b, m, -1.f, 1);
        launch_inject_res(st, sh, bk.vz, bk.vx, rs_z, rs_x, it);
        launch_adj_s(st, sh, bk, m, -1.f);
        launch_adj_v(st, sh, bk, m, -1.f);
        launch_adj_v(st, sh, bb, m, -1.f);
        launch_born_adj(st, sh, bk, bb, m, -1.f);
        launch_adj_s(st, sh, bb, m, -1.f);
        launch_image(st, sh, sg, sb, bk, bb, bg, m, win, W);
    }
    const long long A = (long long)h->g.allnx * h->g.allnz;
    launch_slot_nonfinite(st, sh, win, W, 7, 1, d_bad);
    launch_gather(st, sh, win, W, h->d_grad7, A, 7, h->g.allnx, h->g.allnz, d_bad);
    HIP_TRY(hipMemcpyAsync(bad.data(), d_bad, (size_t)nshots * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipGetLastError());
    h->shot_status.assign(bad.begin(), bad.end());
    int nbad = 0;
    for (int s = 0; s < nshots; ++s) nbad += bad[s] != 0;
    if (nbad)
        return fail(DS2D_ESHOT, "%d of %d shots gave non-finite residuals or images and were left out of the misfit "
                    "and the gradients (ds2d_shot_status lists them)", nbad, nshots);
    return 0;
}

}  // namespace

extern "C" {

int ds2d_batch_gradient(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* obs_z, const float* obs_x,
                        int iter, double* misfit) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!h->have_model) return fail(DS2D_ESTATE, "no model");
    if (h->g.physics != 0) return fail(DS2D_EINVAL, "the reference gradient is elastic");
    if (!obs_z || !obs_x) return fail(DS2D_EINVAL, "observed records required");
    return h->src_fp64 ? batch_gradient_impl<double>(h, nshots, shots, obs_z, obs_x, iter, misfit)
                       : batch_gradient_impl<float>(h, nshots, shots, obs_z, obs_x, iter, misfit);
}

int ds2d_born(ds2d_t* h, int nshots, const ds2d_shot* shots, int order, float* rec_z, float* rec_x) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!h->have_model) return fail(DS2D_ESTATE, "no model");
    if (h->g.physics != 0) return fail(DS2D_EINVAL, "the reference Born modelling is elastic");
    if (order != 1 && order != 2) return fail(DS2D_EINVAL, "order must be 1 (Lm) or 2 (Lm + Lm2)");
    if (!rec_z || !rec_x) return fail(DS2D_EINVAL, "output records required");
    return h->src_fp64 ? batch_gradient_impl<double>(h, nshots, shots, nullptr, nullptr, 0, nullptr, order, rec_z, rec_x)
                       : batch_gradient_impl<float>(h, nshots, shots, nullptr, nullptr, 0, nullptr, order, rec_z, rec_x);
}

// ELSRTM_CALCULATE_ALPHA (FD2DGPU.cu:4885-5035): r = Lm - d with the handle's
// image m (first-order Born, born_modeling mode 1, then GPU_subtract with
// alpha 1), q = Lg with the gradient images g; num += sum q r, den += sum q q
// over every sample of both components (FP64 sums, as fenzi / fenmu).
int ds2d_calculate_alpha(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* g_pp, const float* g_ps,
                         const float* obs_z, const float* obs_x, double* num, double* den) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!g_pp || !g_ps || !obs_z || !obs_x || !num || !den) return fail(DS2D_EINVAL, "null argument");
    long long R = 0;
    for (int i = 0; i < nshots; ++i) R += (long long)shots[i].ntr * h->g.nt;
    std::vector<float> rz(R), rx(R), qz(R), qx(R);
    const size_t A = (size_t)h->g.allnx * h->g.allnz;
    std::vector<float> keep_pp(A, 0.f), keep_ps(A, 0.f);
    const bool had_image = h->have_image;
    if (had_image) {
        HIP_TRY(hipMemcpyAsync(keep_pp.data(), h->d_ipp, A * 4, hipMemcpyDeviceToHost, h->st));
        HIP_TRY(hipMemcpyAsync(keep_ps.data(), h->d_ips, A * 4, hipMemcpyDeviceToHost, h->st));
        HIP_TRY(hipStreamSynchronize(h->st));
    }
    if ((rc = ds2d_born(h, nshots, shots, 1, rz.data(), rx.data()))) return rc;
    if ((rc = ds2d_set_image(h, g_pp, g_ps))) return rc;
    rc = ds2d_born(h, nshots, shots, 1, qz.data(), qx.data());
    // restore the handle's image (m) whatever the outcome
    int rc2 = had_image ? ds2d_set_image(h, keep_pp.data(), keep_ps.data()) : ds2d_set_image(h, nullptr, nullptr);
    h->have_image = had_image;
    if (rc) return rc;
    if (rc2) return rc2;
    double a = 0, b = 0;
    for (long long k = 0; k < R; ++k) {
        const double ez = (double)rz[k] - obs_z[k], ex = (double)rx[k] - obs_x[k];
        a += (double)qz[k] * ez + (double)qx[k] * ex;
        b += (double)qz[k] * qz[k] + (double)qx[k] * qx[k];
    }
    *num += a;
    *den += b;
    return 0;
}

int ds2d_set_precision(ds2d_t* h, int source_fp64) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (source_fp64 != 0 && source_fp64 != 1) return fail(DS2D_EINVAL, "source_fp64 must be 0 or 1");
    h->src_fp64 = source_fp64;
    return 0;
}

int ds2d_shot_gradient(ds2d_t* h, int iter, double* misfit) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!h->have_shot) return fail(DS2D_ESTATE, "no shot set");
    return ds2d_batch_gradient(h, 1, &h->cur, h->cur_oz.data(), h->cur_ox.data(), iter, misfit);
}

int ds2d_shot_status(ds2d_t* h, int nshots, int* status) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!status) return fail(DS2D_EINVAL, "null status");
    if (nshots != (int)h->shot_status.size())
        return fail(DS2D_EINVAL, "the last gradient had %zu shots, not %d", h->shot_status.size(), nshots);
    std::copy(h->shot_status.begin(), h->shot_status.end(), status);
    return 0;
}

int ds2d_take_gradients(ds2d_t* h, float* pp, float* ps, float* il, float* bpp, float* bps, float* b2pp,
                        float* b2ps) {
    int rc;
    if ((rc = check_h(h))) return rc;
    const size_t A = (size_t)h->g.allnx * h->g.allnz;
    float* outs[7] = {pp, ps, il, bpp, bps, b2pp, b2ps};
    for (int k = 0; k < 7; ++k)
        if (outs[k]) HIP_TRY(hipMemcpyAsync(outs[k], h->d_grad7 + k * A, A * 4, hipMemcpyDeviceToHost, h->st));
    HIP_TRY(hipMemsetAsync(h->d_grad7, 0, 7 * A * 4, h->st));
    HIP_TRY(hipStreamSynchronize(h->st));
    return 0;
}

// ---------------------------------------------------------------------------
// Elastic forward modelling
// ---------------------------------------------------------------------------
// Forward modelling for the first nt_run steps (nt_run <= nt; the autotuner
// times short runs).
static int batch_forward_n(ds2d_t* h, int nshots, const ds2d_shot* shots, float* rec_z, float* rec_x, int mode,
                           ds2d_stats* stats, int nt_run) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!h->have_model) return fail(DS2D_ESTATE, "no model");
    if (h->g.physics != 0) return fail(DS2D_EINVAL, "elastic forward needs physics = 0");
    if (mode < 0 || mode > 2) return fail(DS2D_EINVAL, "mode must be 0, 1 or 2");
    Batch B;
    if ((rc = build_batch(h, nshots, shots, B))) return rc;
    const BatchShape& sh = B.shape;
    const long long n = B.total_cells;
    hipStream_t st = h->st;
    Med m;
    if ((rc = alloc_med(h, n, m))) return rc;
    const size_t R = std::max<long long>(B.total_rec, 1);
    float* rs;
    if ((rc = h->fbuf("frec", R * 2, &rs))) return rc;
    float *rz = rs, *rx = rs + R;
    EField f;
    if ((rc = alloc_efield(h, "bg", n, f, false))) return rc;
    const int nt = std::max(1, std::min(nt_run, h->g.nt));
    long long cu = 0;
    for (auto& S : B.slots) cu += (long long)S.nxp * S.nzp;
    cu *= nt;
    stats_begin(h);
    launch_setup(st, sh, setup_args(h, false), m);
    HIP_TRY(hipMemsetAsync(f.vx, 0, (size_t)n * 15 * 4, st));
    double bytes_v = 0, bytes_s = 0, blocked = 0;
    // cells of the update region that carry split PML memory (PML windows only)
    double pml_cells = 0;
    for (auto& S : B.slots)
        if (S.boundary == 0 && S.pml >= NOP)
            pml_cells += (double)(S.nxp - 2 * NOP) * (S.nzp - 2 * NOP) -
                         (double)(S.nxp - 2 * S.pml + 1) * (S.nzp - 2 * S.pml + 1);
    long long launches = 0;
    StepSampler smp;
    if ((rc = smp.init(nt))) return rc;
    if (mode == 0) {
        Strain none{nullptr, nullptr, nullptr, nullptr};
        stats_loop_begin(h);
        for (int it = 0; it < nt; ++it) {
            const bool tk = smp.take(it);
            launch_inject_src(st, sh, f.txx, f.tzz, h->d_wav, it);
            if (tk) smp.mark(st, 0);
            launch_ref_v(st, sh, f, m, 1.f, 0);
            if (tk) smp.mark(st, 1);
            launch_ref_s(st, sh, f, m, none, 1.f, 0);
            if (tk) smp.mark(st, 2);
            launch_record(st, sh, f.vz, f.vx, rz, rx, nullptr, nullptr, nullptr, nullptr, it);
        }
        stats_loop_end(h);
        launches = (long long)nt;
        // split layout, per padded cell: v reads 10 writes 6; s reads 12 writes 9
        bytes_v = (double)B.total_cells * 16 * 4;
        bytes_s = (double)B.total_cells * 21 * 4;
    } else if (mode == 2 && !h->fwd_two && n <= FUSED_MAX_CELLS) {
        // single fused kernel per step, ping-pong state
        float* par;
        int* rmap;
        if ((rc = h->fbuf("fpar", (size_t)n * 3, &par)) || (rc = h->fbuf("rmap", B.total_rmap, &rmap))) return rc;
        launch_fast_params(st, m.rho, m.lam, m.mu, par, par + n, par + 2 * n, n, h->g.dt);
        HIP_TRY(hipMemsetAsync(rmap, 0xff, B.total_rmap * 4, st));
        launch_fast_rmap(st, sh, rmap);
        EField g2;
        if ((rc = alloc_efield(h, "b", n, g2, true))) return rc;
        std::vector<ChunkDesc> ch;
        const int W = h->chunk_w > 0 ? h->chunk_w : 64;
        const int T = h->chunk_t > 0 ? h->chunk_t : 256;
        const int D = h->chunk_d == 4 ? 4 : 2;
        const int rows = T - 2 * FUSED_HALO;
        for (int s = 0; s < nshots; ++s)
            for (int x0 = 0; x0 < B.slots[s].nxp; x0 += W)
                for (int z0 = 0; z0 < B.slots[s].nzp; z0 += rows)
                    ch.push_back(ChunkDesc{s, x0, std::min(x0 + W, B.slots[s].nxp), z0});
        ChunkDesc* d_ch;
        if ((rc = h->fbuf("chunks", ch.size(), &d_ch))) return rc;
        HIP_TRY(hipMemcpyAsync(d_ch, ch.data(), ch.size() * sizeof(ChunkDesc), hipMemcpyHostToDevice, st));
        auto mk = [&](const EField& A, const EField& Bf) {
            FusedArgs q{};
            q.A = A.vx;  // alloc_efield: 15 fields of n floats from vx on
            q.B = Bf.vx;
            q.n = n;
            q.par = par;
            q.dampx = m.dampx;
            q.dampz = m.dampz;
            q.wav = h->d_wav;
            q.rec_z = rz;
            q.rec_x = rx;
            return q;
        };
        const FusedArgs even = mk(f, g2), odd = mk(g2, f);
        if (h->tblock == 2 && nt >= 2) {
            // temporal blocking of depth 2: steps (it, it+1) per launch; an odd
            // last step runs the single-step kernel on its own chunk list
            const int T2 = fused2_threads_ok(T) ? T : 704;
            const int D2 = 2;  // two columns in flight
            std::vector<ChunkDesc> ch2;
            const int rows2 = T2 - 2 * FUSED2_HALO;
            {
                // fixed-width chunks, or the balanced plan when its simulated
                // launch time is shorter (chunk_plan.hpp)
                std::vector<int> nxv(nshots), nzv(nshots);
                for (int s = 0; s < nshots; ++s) {
                    nxv[s] = B.slots[s].nxp;
                    nzv[s] = B.slots[s].nzp;
                }
                int ncu = 0;
                if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess ||
                    ncu < kPlanXcds)
                    ncu = 256;
                const int cpx = ncu / kPlanXcds;
                std::vector<int> key{W, rows2, cpx, h->chunk_plan};
                key.insert(key.end(), nxv.begin(), nxv.end());
                key.insert(key.end(), nzv.begin(), nzv.end());
                if (key != h->plan_key) {  // plan once per batch shape (the search takes milliseconds)
                    std::vector<PlanChunk> fixed, bal;
                    const long long span_f = plan_fixed(nxv, nzv, rows2, W, cpx, fixed);
                    const long long span_b = h->chunk_plan ? plan_balanced(nxv, nzv, rows2, cpx, bal) : -1;
                    h->plan_list_bal = span_b >= 0 && span_b < span_f;
                    h->plan_list.clear();
                    for (const PlanChunk& c : h->plan_list_bal ? bal : fixed)
                        h->plan_list.push_back(ChunkDesc{c.slot, c.x0, c.x1, c.z0});
                    h->plan_key = key;
                }
                h->last_plan = h->plan_list_bal;
                ch2 = h->plan_list;
            }
            ChunkDesc* d_ch2;
            if ((rc = h->fbuf("chunks2", ch2.size(), &d_ch2))) return rc;
            HIP_TRY(hipMemcpyAsync(d_ch2, ch2.data(), ch2.size() * sizeof(ChunkDesc), hipMemcpyHostToDevice, st));
            float* scratch = nullptr;
            if (B.any_pml && (rc = h->fbuf("f2split", (size_t)n * 5, &scratch))) return rc;
            int k = 0;  // launch parity: even launches read f, odd ones read g2
            auto two = [&](int it) {
                launch_fused2_step(st, sh, (k & 1) ? odd : even, scratch, d_ch2, (int)ch2.size(), rmap, it, T2, D2,
                                   B.any_pml);
                ++k;
            };
            stats_loop_begin(h);
            int it = 0;
            if (h->use_graph) {
                // the paired steps as one graph; per-launch time = loop time / launches
                if ((rc = run_steps(h, 0, nt / 2, 2, two))) return rc;
                it = nt / 2 * 2;
            } else {
                for (; it + 1 < nt; it += 2) {
                    const bool tk = smp.take(it);
                    if (tk) smp.mark(st, 0);
                    two(it);
                    if (tk) { smp.mark(st, 1); smp.mark(st, 2); }
                }
            }
            if (it < nt) {
                launch_fused_step(st, sh, (k & 1) ? odd : even, d_ch, (int)ch.size(), rmap, it, T, D, B.any_pml);
                ++k;
            }
            stats_loop_end(h);
            launches = (long long)(nt / 2);
            // algorithmic bytes of the cell-updates one launch performs: two steps at the one-step figure
            // (52 B per padded cell, +48 B per PML cell).  What this kernel itself must move is less: the
            // state and parameters once per launch (13 floats per cell) and per PML cell dampx, dampz twice
            // and 4 x 5 splits (A read, scratch write + read, B write) = 24 floats.
            bytes_v = 2.0 * ((double)B.total_cells * 13 * 4 + pml_cells * 12 * 4);
            blocked = (double)B.total_cells * 13 * 4 + pml_cells * 24 * 4;
            bytes_s = 0;
        } else {
        stats_loop_begin(h);
        if (h->use_graph) {
            // whole time loop as one graph; per-launch time = loop time / nt
            if ((rc = run_steps(h, 0, nt, 1, [&](int it) {
                     launch_fused_step(st, sh, (it & 1) ? odd : even, d_ch, (int)ch.size(), rmap, it, T, D, B.any_pml);
                 })))
                return rc;
        } else {
            for (int it = 0; it < nt; ++it) {
                const bool tk = smp.take(it);
                if (tk) smp.mark(st, 0);
                launch_fused_step(st, sh, (it & 1) ? odd : even, d_ch, (int)ch.size(), rmap, it, T, D, B.any_pml);
                if (tk) { smp.mark(st, 1); smp.mark(st, 2); }
            }
        }
        stats_loop_end(h);
        launches = (long long)nt;
        // unsplit interior, per padded cell: reads stress 3 + velocity 2 + dt/rho, lambda, mu; writes 5.
        // Each PML cell also reads dampx, dampz and reads + writes one split of each of the 5 components.
        bytes_v = (double)B.total_cells * 13 * 4 + pml_cells * 12 * 4;
        bytes_s = 0;
        }
    } else {
        float* par;
        int* rmap;
        if ((rc = h->fbuf("fpar", (size_t)n * 3, &par)) || (rc = h->fbuf("rmap", B.total_rmap, &rmap))) return rc;
        launch_fast_params(st, m.rho, m.lam, m.mu, par, par + n, par + 2 * n, n, h->g.dt);
        HIP_TRY(hipMemsetAsync(rmap, 0xff, B.total_rmap * 4, st));
        launch_fast_rmap(st, sh, rmap);
        FastArgs a{};
        a.vx = f.vx; a.vz = f.vz; a.txx = f.txx; a.tzz = f.tzz; a.txz = f.txz;
        a.vx_x = f.vx_x; a.vx_z = f.vx_z; a.vz_x = f.vz_x; a.vz_z = f.vz_z;
        a.txx_x = f.txx_x; a.txx_z = f.txx_z; a.tzz_x = f.tzz_x; a.tzz_z = f.tzz_z; a.txz_x = f.txz_x;
        a.txz_z = f.txz_z;
        a.bx = par; a.lam = par + n; a.mu = par + 2 * n;
        a.dampx = m.dampx; a.dampz = m.dampz;
        a.wav = h->d_wav;
        a.rec_z = rz; a.rec_x = rx;
        stats_loop_begin(h);
        for (int it = 0; it < nt; ++it) {
            const bool tk = smp.take(it);
            if (tk) smp.mark(st, 0);
            launch_fast_v(st, sh, a, rmap, it);
            if (tk) smp.mark(st, 1);
            launch_fast_s(st, sh, a);
            if (tk) smp.mark(st, 2);
        }
        stats_loop_end(h);
        launches = (long long)nt;
        // unsplit layout, per padded cell: v reads 6 writes 2; s reads 7 writes 3.
        // PML cells: v also reads dampx, dampz, 4 splits and writes 4; s reads dampx, dampz, 6 and writes 6.
        bytes_v = (double)B.total_cells * 8 * 4 + pml_cells * 10 * 4;
        bytes_s = (double)B.total_cells * 10 * 4 + pml_cells * 14 * 4;
    }
    if ((rc = stats_end(h, stats, cu, launches, 0))) return rc;
    {
        double kv, ks;
        smp.averages(&kv, &ks);
        // report the kernel with the larger average duration as the dominant one
        const bool s_dom = ks >= kv && bytes_s > 0;
        h->last.kernel_avg_ms = s_dom ? ks : kv;
        h->last.alg_bytes = s_dom ? bytes_s : bytes_v;
        // the other kernel of a two-kernel step
        h->last.aux_kernel_avg_ms = bytes_s > 0 && bytes_v > 0 ? (s_dom ? kv : ks) : 0;
        h->last.aux_alg_bytes = bytes_s > 0 && bytes_v > 0 ? (s_dom ? bytes_v : bytes_s) : 0;
        h->last.kernel_samples = smp.n;
        h->last.blocked_bytes = blocked;
        if (smp.n == 0 && launches > 0) {  // graph mode: one kernel per launch, no gaps
            h->last.kernel_avg_ms = h->last.stencil_ms / (double)launches;
            h->last.aux_kernel_avg_ms = h->last.aux_alg_bytes = 0;
            h->last.kernel_samples = launches;
        }
        if (stats) *stats = h->last;
    }
    if (rec_z) HIP_TRY(hipMemcpyAsync(rec_z, rz, B.total_rec * 4, hipMemcpyDeviceToHost, st));
    if (rec_x) HIP_TRY(hipMemcpyAsync(rec_x, rx, B.total_rec * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipGetLastError());
    return 0;
}

int ds2d_batch_forward(ds2d_t* h, int nshots, const ds2d_shot* shots, float* rec_z, float* rec_x, int mode,
                       ds2d_stats* stats) {
    return batch_forward_n(h, nshots, shots, rec_z, rec_x, mode, stats, h ? h->g.nt : 0);
}

int ds2d_set_time_block(ds2d_t* h, int steps) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (steps != 1 && steps != 2) return fail(DS2D_EINVAL, "time steps per launch must be 1 or 2, got %d", steps);
    h->tblock = steps;
    return 0;
}

int ds2d_set_chunk_plan(ds2d_t* h, int on) {
    int rc;
    if ((rc = check_h(h))) return rc;
    h->chunk_plan = on ? 1 : 0;
    return 0;
}

int ds2d_set_fused_born(ds2d_t* h, int on, int width, int threads) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (width < 0 || (width > 0 && width < 8)) return fail(DS2D_EINVAL, "chunk width must be 0 or >= 8");
    if (threads != 0 && !eb_fused_threads_ok(threads)) return fail(DS2D_EINVAL, "threads must be 0, 256 or 512");
    h->eb_fused = on ? 1 : 0;
    h->eb_w = width;
    h->eb_t = threads;
    return 0;
}

int ds2d_set_fused_rotation(ds2d_t* h, int on) {
    int rc;
    if ((rc = check_h(h))) return rc;
    h->rot_fused = on ? 1 : 0;
    return 0;
}

int ds2d_get_chunk_plan(ds2d_t* h, int* used) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!used) return fail(DS2D_EINVAL, "null argument");
    *used = h->last_plan;
    return 0;
}

int ds2d_get_time_block(ds2d_t* h, int* steps) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!steps) return fail(DS2D_EINVAL, "null argument");
    *steps = h->tblock;
    return 0;
}

int ds2d_set_graph(ds2d_t* h, int on) {
    int rc;
    if ((rc = check_h(h))) return rc;
    h->use_graph = on ? 1 : 0;
    return 0;
}

int ds2d_set_chunk(ds2d_t* h, int width, int threads, int depth) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (width < 0 || (width > 0 && width < 8)) return fail(DS2D_EINVAL, "chunk width must be 0 or >= 8");
    if (threads != 0 && !fused_threads_ok(threads))
        return fail(DS2D_EINVAL, "threads must be 0, 256, 512, 704 or 768");
    if (depth != 0 && depth != 2 && depth != 4) return fail(DS2D_EINVAL, "depth must be 0, 2 or 4");
    h->chunk_w = width;
    h->chunk_t = threads;
    h->chunk_d = depth;
    h->fwd_two = 0;
    return 0;
}

// Time nsteps of the fused kernel for each candidate tile on this batch, and
// of the two-kernel step (width 0): best of two runs after one untimed warm
// run; keep the fastest for mode 2.
int ds2d_autotune_forward(ds2d_t* h, int nshots, const ds2d_shot* shots, int nsteps, int* best_width,
                          int* best_threads, int* best_depth, double* best_ms) {
    int rc;
    if ((rc = check_h(h))) return rc;
    struct Cand {
        int w, t, d, tb;  // x width (0: two-kernel step), lanes, prefetch columns, time steps per launch
    };
    static const Cand cand[] = {{0, 0, 0, 1},     {32, 256, 2, 1},  {64, 256, 2, 1},  {64, 256, 4, 1},
                                {128, 256, 4, 1}, {64, 512, 4, 1},  {64, 768, 2, 1},  {64, 768, 4, 1},
                                {128, 768, 4, 1}, {64, 704, 4, 1},  {112, 704, 4, 1}, {160, 704, 4, 1},
                                {64, 704, 2, 2},  {128, 704, 2, 2}};
    std::vector<Cand> cands(std::begin(cand), std::end(cand));
    // 704-lane tiles whose chunk count just fills R rounds of one workgroup per CU
    // (R = 2, 3, 4): the step time moves with the number of rounds, not with the width alone
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0) == hipSuccess && ncu > 0) {
        const int pml = h->g.pml, nzp = h->g.allnz + 2 * pml, rb = (nzp + 687) / 688;
        int max_nxp = 1;
        for (int i = 0; i < nshots; ++i) max_nxp = std::max(max_nxp, shots[i].nx + 2 * pml);
        auto nchunks = [&](int W) {
            long long k = 0;
            for (int i = 0; i < nshots; ++i) k += (shots[i].nx + 2 * pml + W - 1) / W;
            return k * rb;
        };
        for (int R = 2; R <= 4; ++R) {
            int lo = 16, hi = std::max(16, max_nxp);  // smallest W with nchunks(W) <= R * ncu
            if (nchunks(hi) > (long long)R * ncu) continue;
            while (lo < hi) {
                const int mid = (lo + hi) / 2;
                if (nchunks(mid) <= (long long)R * ncu)
                    hi = mid;
                else
                    lo = mid + 1;
            }
            bool have = false;
            for (const Cand& c : cands) have = have || (c.w == lo && c.t == 704 && c.d == 4);
            if (!have) cands.push_back(Cand{lo, 704, 4, 1});
            // the same widths two steps per launch (704 lanes own 672 rows; 640-row windows stay one block)
            cands.push_back(Cand{lo, 704, 2, 2});
        }
    }
    Cand best{64, 256, 2, 1};
    double bt = 1e300;
    const int keep_w = h->chunk_w, keep_t = h->chunk_t, keep_d = h->chunk_d, keep_two = h->fwd_two,
              keep_tb = h->tblock, keep_graph = h->use_graph;
    h->use_graph = 0;  // candidates are timed as direct launches (a short capture would time its instantiation)
    for (const Cand& c : cands) {
        if (c.tb == 2 && nsteps < 2) continue;
        h->fwd_two = c.w == 0;
        h->chunk_w = c.w;
        h->chunk_t = c.t;
        h->chunk_d = c.d;
        h->tblock = c.tb;
        ds2d_stats s{};
        double t = 1e300;  // best of two timed runs after one warm run
        for (int rep = 0; rep < 3; ++rep) {
            if ((rc = batch_forward_n(h, nshots, shots, nullptr, nullptr, 2, &s, nsteps))) {
                h->chunk_w = keep_w;
                h->chunk_t = keep_t;
                h->chunk_d = keep_d;
                h->fwd_two = keep_two;
                h->tblock = keep_tb;
                h->use_graph = keep_graph;
                return rc;
            }
            if (rep > 0) t = std::min(t, s.stencil_ms);
        }
        if (t < bt) {
            bt = t;
            best = c;
        }
    }
    h->use_graph = keep_graph;
    h->fwd_two = best.w == 0;
    h->chunk_w = best.w;
    h->chunk_t = best.t;
    h->chunk_d = best.d;
    h->tblock = best.tb;
    if (best_width) *best_width = best.w;
    if (best_threads) *best_threads = best.t;
    if (best_depth) *best_depth = best.d;
    if (best_ms) *best_ms = bt;
    return 0;
}

int ds2d_forward(ds2d_t* h, float* rec_z, float* rec_x) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!h->have_shot) return fail(DS2D_ESTATE, "no shot set");
    return ds2d_batch_forward(h, 1, &h->cur, rec_z, rec_x, 2, nullptr);
}

// ---------------------------------------------------------------------------
// Reference acoustic forward (FD() of FD2DGPU.cu:554-584)
// ---------------------------------------------------------------------------
int ds2d_acoustic_ref_forward(ds2d_t* h, int sxnum, int ngx_min, float* rec_p) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!h->have_model) return fail(DS2D_ESTATE, "no model");
    const ds2d_geom& g = h->g;
    const int nx = g.nxpml - 2 * g.pml;
    if (sxnum < 0 || sxnum >= nx) return fail(DS2D_EINVAL, "sxnum %d outside window of %d", sxnum, nx);
    int rix = NOP;
    ds2d_shot s{};
    s.ngx_min = ngx_min;
    s.nx = nx;
    s.isx = g.pml + sxnum;
    s.isz = g.pml;
    s.igz = g.pml;
    s.ntr = 0;
    s.rec_ix = &rix;
    s.boundary = 0;
    Batch B;
    if ((rc = build_batch(h, 1, &s, B))) return rc;
    const BatchShape& sh = B.shape;
    const long long n = B.total_cells;
    Med m;
    if ((rc = alloc_med(h, n, m))) return rc;
    float* ab;
    if ((rc = h->fbuf("aref", (size_t)n * 9, &ab))) return rc;
    HIP_TRY(hipMemsetAsync(ab, 0, (size_t)n * 9 * 4, h->st));
    float *px = ab, *pz = ab + n, *p = ab + 2 * n, *vx = ab + 3 * n, *vz = ab + 4 * n, *bu = ab + 5 * n,
          *kk = ab + 6 * n, *refl = ab + 7 * n, *damp = ab + 8 * n;
    SetupArgs a = setup_args(h, false);
    a.pml_rr = g.pml_rr > 0 ? g.pml_rr : 1e-6f;
    launch_setup_acoustic(h->st, sh, a, m, bu, kk, refl);
    launch_combine_damp(h->st, sh, m.dampx, m.dampz, damp);
    float* drec;
    if ((rc = h->fbuf("arec", (size_t)nx * g.nt, &drec))) return rc;
    for (int it = 0; it < g.nt; ++it)
        launch_acou_ref_step(h->st, sh, px, pz, p, vx, vz, kk, m.rho, damp, h->d_wav, sxnum, drec, it);
    HIP_TRY(hipMemcpyAsync(rec_p, drec, (size_t)nx * g.nt * 4, hipMemcpyDeviceToHost, h->st));
    HIP_TRY(hipStreamSynchronize(h->st));
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Plain-L2 acoustic Born operator, adjoint and CG
// ---------------------------------------------------------------------------
namespace {

struct AcouWS {
    AcouFieldD bg;              // FP64 background
    AcouField bn, lm;
    AcouField bn2, lm2;         // second state of the fused (ping-pong) Born / adjoint steps
    AcouMed m;
    float *grad_pad = nullptr;  // per-slot padded gradient (n)
    float *win = nullptr;       // per-slot window images
    float *rec = nullptr;       // R floats: bg records
    float *rec2 = nullptr;      // R floats: born records / residual
    float *hist = nullptr;      // resident background increments (store mode)
    bool store = false;
    bool fused = false;         // store mode with the fused one-launch-per-step kernels
    int* rmap = nullptr;        // column -> trace of every slot (-1: none)
    ChunkDesc* chunks = nullptr;
    std::vector<std::pair<int, int>> gchunks;  // per launch group: first chunk, count
    bool tune = false;          // choose per-cell vs fused (and the x width) at the first Born pass
    long long n = 0, R = 0, W = 0;
};

constexpr int ACOU_CHUNK_W = 64, ACOU_THREADS = 256, ACOU_TUNE_STEPS = 48;

// Chunk list of the fused acoustic steps for x width W (per launch group).
int acou_build_chunks(ds2d_ctx* h, const Batch& B, AcouWS& w, int W) {
    if (W < 1 || W > ACOU_FUSED_MAX_W) return fail(DS2D_EINVAL, "fused acoustic x width %d outside [1, %d]", W,
                                                   ACOU_FUSED_MAX_W);
    std::vector<ChunkDesc> ch;
    const int ns = (int)B.slots.size();
    const int gs = h->per_launch > 0 ? std::min(h->per_launch, ns) : ns;
    const int rows = ACOU_THREADS - 16;
    w.gchunks.clear();
    for (int s0 = 0; s0 < ns; s0 += gs) {
        const int first = (int)ch.size();
        for (int s = s0; s < std::min(ns, s0 + gs); ++s)
            for (int x0 = 0; x0 < B.slots[s].nxp; x0 += W)
                for (int z0 = 0; z0 < B.slots[s].nzp; z0 += rows)
                    ch.push_back(ChunkDesc{s, x0, std::min(x0 + W, B.slots[s].nxp), z0});
        w.gchunks.push_back({first, (int)ch.size() - first});
    }
    int rc;
    if ((rc = h->fbuf("achunks", ch.size(), &w.chunks))) return rc;
    HIP_TRY(hipMemcpyAsync(w.chunks, ch.data(), ch.size() * sizeof(ChunkDesc), hipMemcpyHostToDevice, h->st));
    return 0;
}

int acou_prepare(ds2d_ctx* h, const Batch& B, const float* refl_glob_dev, AcouWS& w) {
    int rc;
    const long long n = B.total_cells;
    w.n = n;
    w.R = std::max<long long>(B.total_rec, 1);
    w.W = std::max<long long>(B.total_win, 1);
    float* base;
    double* bgd;
    if ((rc = h->fbuf("acou", (size_t)n * 16, &base)) || (rc = h->fbuf("acoubg", (size_t)n * 3, &bgd))) return rc;
    w.bg = AcouFieldD{bgd, bgd + n, bgd + 2 * n};
    w.bn = AcouField{base, base + n, base + 2 * n};
    w.lm = AcouField{base + 3 * n, base + 4 * n, base + 5 * n};
    w.grad_pad = base + 6 * n;
    float *bu = base + 7 * n, *kk = base + 8 * n, *refl = base + 9 * n;
    w.bn2 = AcouField{base + 10 * n, base + 11 * n, base + 12 * n};
    w.lm2 = AcouField{base + 13 * n, base + 14 * n, base + 15 * n};
    Med m;
    if ((rc = alloc_med(h, n, m))) return rc;
    SetupArgs a = setup_args(h, false);
    a.img_pp = refl_glob_dev;
    launch_setup_acoustic(h->st, B.shape, a, m, bu, kk, refl);
    w.m = AcouMed{bu, kk, refl, m.dampx, m.rho};
    if ((rc = h->fbuf("acrec", (size_t)w.R * 2, &w.rec))) return rc;
    w.rec2 = w.rec + w.R;
    if ((rc = h->fbuf("acwin", (size_t)w.W, &w.win))) return rc;
    // background mode: keep the per-step increment resident when it fits
    const double hist_gb = (double)B.total_hist * 4 / 1e9;
    const double budget = h->g.hbm_budget_gb > 0 ? h->g.hbm_budget_gb : 64.0;
    w.store = h->g.bg_mode == 1 || (h->g.bg_mode == 0 && hist_gb <= budget);
    if (w.store && (rc = h->fbuf("hist", (size_t)B.total_hist, &w.hist))) return rc;
    // receiver map (column -> trace) of every slot, for the in-kernel recorder and injection
    if ((rc = h->fbuf("armap", std::max<long long>(B.total_rmap, 1), &w.rmap))) return rc;
    HIP_TRY(hipMemsetAsync(w.rmap, 0xff, B.total_rmap * 4, h->st));
    launch_fast_rmap(h->st, B.shape, w.rmap);
    // fused Born / adjoint steps: store mode, 32-bit byte offsets within a field; the choice
    // (per-cell kernels or fused at some x width) is timed once per batch geometry
    w.fused = false;
    w.tune = false;
    if (w.store && n <= FUSED_MAX_CELLS && h->acou_fused != 0) {
        int width = h->acou_fused > 1 ? h->acou_fused : ACOU_CHUNK_W;
        if (h->acou_fused < 0) {
            auto it = h->acou_tuned.find({n, (int)B.slots.size(), h->per_launch});
            if (it == h->acou_tuned.end()) {
                w.tune = true;
                width = 0;
            } else {
                width = it->second;
            }
        }
        if (width > 0) {
            w.fused = true;
            if ((rc = acou_build_chunks(h, B, w, width))) return rc;
        }
    }
    return 0;
}
AcouFusedArgs acou_fused_args(ds2d_ctx* h, AcouWS& w, const AcouField& A, const AcouField& Bf) {
    AcouFusedArgs q{};
    q.A = A.vx;
    q.B = Bf.vx;
    q.n = w.n;
    q.b = w.m.b;
    q.k = w.m.k;
    q.m = w.m.m;
    q.hist = w.hist;
    q.ax = h->g.dt / h->g.dx;
    q.az = h->g.dt / h->g.dz;
    return q;
}

// Re-window a new reflectivity into the existing workspace (zero frame).
int acou_set_refl(ds2d_ctx* h, const Batch& B, const float* refl_glob_dev, AcouWS& w) {
    Med m;
    int rc;
    if ((rc = alloc_med(h, w.n, m))) return rc;
    SetupArgs a = setup_args(h, false);
    a.img_pp = refl_glob_dev;
    launch_setup_acoustic(h->st, B.shape, a, m, (float*)w.m.b, (float*)w.m.k, (float*)w.m.m);
    return 0;
}

// Shot groups launched as one time loop each: the whole batch, or runs of
// per_launch consecutive slots (per_launch = 1 is the reference's one shot
// after another).  Slots keep their arena offsets, so a group is a view.
std::vector<BatchShape> launch_groups(const ds2d_ctx* h, const Batch& B) {
    const int ns = (int)B.slots.size();
    const int gs = h->per_launch > 0 ? std::min(h->per_launch, ns) : ns;
    std::vector<BatchShape> out;
    for (int s0 = 0; s0 < ns; s0 += gs) {
        BatchShape sh = B.shape;
        const int s1 = std::min(ns, s0 + gs);
        sh.slots = B.shape.slots + s0;
        sh.nslots = s1 - s0;
        sh.max_nxp = 0;
        sh.max_ntr = 1;
        for (int k = s0; k < s1; ++k) {
            sh.max_nxp = std::max(sh.max_nxp, B.slots[k].nxp);
            sh.max_ntr = std::max(sh.max_ntr, B.slots[k].ntr);
        }
        out.push_back(sh);
    }
    return out;
}

// Forward pass from rest.  do_bg: propagate the background (storing its
// increment when w.store); otherwise the Born field reads the stored
// increment.  Records: w.rec (bg), w.rec2 (born).
void acou_fwd(ds2d_ctx* h, const Batch& B, AcouWS& w, bool do_bg, bool do_born, bool rec_bg);

// Autotuner of the resident-history Born pass: time ACOU_TUNE_STEPS steps of
// the per-cell kernels and of the fused step at several x widths, keep the
// fastest for this batch geometry (cached in the handle).
void acou_tune(ds2d_ctx* h, const Batch& B, AcouWS& w) {
    hipStream_t st = h->st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return;
    const int steps = std::min(h->g.nt, ACOU_TUNE_STEPS);
    auto time_pass = [&](int width) -> float {
        (void)hipMemsetAsync(w.bn.vx, 0, (size_t)w.n * 3 * 4, st);
        if (width > 0 && acou_build_chunks(h, B, w, width)) return 1e30f;
        (void)hipEventRecord(e0, st);
        if (width == 0) {
            for (const BatchShape& sh : launch_groups(h, B))
                for (int it = 0; it < steps; ++it) {
                    launch_acou_v(st, sh, w.bg, w.bn, w.m, h->d_wav, it, 1.f, 0, 1);
                    launch_acou_p(st, sh, w.bg, w.bn, w.m, w.hist, nullptr, w.rec2, it, 0, 1, 2, w.rmap);
                }
        } else {
            AcouFusedArgs even = acou_fused_args(h, w, w.bn, w.bn2), odd = acou_fused_args(h, w, w.bn2, w.bn);
            even.rec = odd.rec = w.rec2;
            for (size_t gi = 0; gi < w.gchunks.size(); ++gi)
                for (int it = 0; it < steps; ++it)
                    launch_acou_born_fused(st, B.shape, (it & 1) ? odd : even, w.chunks + w.gchunks[gi].first,
                                           w.gchunks[gi].second, w.rmap, it, ACOU_THREADS);
        }
        (void)hipEventRecord(e1, st);
        float ms = 1e30f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1e30f;
        return ms;
    };
    // widths 16, 32, 64 and those whose chunk count just fills 1, 2 or 3 rounds of one workgroup per CU
    std::vector<int> widths = {16, 32, 64};
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0) == hipSuccess && ncu > 0) {
        const int rows = ACOU_THREADS - 16;
        int max_nxp = 16;
        for (const SlotDesc& S : B.slots) max_nxp = std::max(max_nxp, S.nxp);
        auto nchunks = [&](int W) {
            long long k = 0;
            for (const SlotDesc& S : B.slots) k += (long long)((S.nxp + W - 1) / W) * ((S.nzp + rows - 1) / rows);
            return k;
        };
        for (int R = 1; R <= 3; ++R) {
            int lo = 8, hi = std::min(max_nxp, ACOU_FUSED_MAX_W);  // smallest W with nchunks(W) <= R * ncu
            if (nchunks(hi) > (long long)R * ncu) continue;
            while (lo < hi) {
                const int mid = (lo + hi) / 2;
                if (nchunks(mid) <= (long long)R * ncu)
                    hi = mid;
                else
                    lo = mid + 1;
            }
            if (std::find(widths.begin(), widths.end(), lo) == widths.end()) widths.push_back(lo);
        }
    }
    int best = 0;
    float bt = time_pass(0);
    for (int width : widths) {
        const float t = time_pass(width);
        if (t < bt) {
            bt = t;
            best = width;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->acou_tuned[{w.n, (int)B.slots.size(), h->per_launch}] = best;
    w.fused = best > 0 && acou_build_chunks(h, B, w, best) == 0;
}

void acou_fwd(ds2d_ctx* h, const Batch& B, AcouWS& w, bool do_bg, bool do_born, bool rec_bg) {
    hipStream_t st = h->st;
    if (w.tune && !do_bg && do_born && !rec_bg) {
        w.tune = false;
        acou_tune(h, B, w);
    }
    (void)hipMemsetAsync(w.bg.vx, 0, (size_t)w.n * 3 * 8, st);
    (void)hipMemsetAsync(w.bn.vx, 0, (size_t)w.n * 3 * 4, st);
    if (w.fused && !do_bg && do_born && !rec_bg) {
        // Born field only, from the resident history: one fused launch per step
        AcouFusedArgs even = acou_fused_args(h, w, w.bn, w.bn2), odd = acou_fused_args(h, w, w.bn2, w.bn);
        even.rec = odd.rec = w.rec2;
        // chunk lists hold absolute slot indices: launch with the whole slot table
        for (size_t gi = 0; gi < w.gchunks.size(); ++gi)
            for (int it = 0; it < h->g.nt; ++it)
                launch_acou_born_fused(st, B.shape, (it & 1) ? odd : even, w.chunks + w.gchunks[gi].first,
                                       w.gchunks[gi].second, w.rmap, it, ACOU_THREADS);
        return;
    }
    const int hmode = do_bg ? (w.store ? 1 : 0) : 2;
    for (const BatchShape& sh : launch_groups(h, B))
        for (int it = 0; it < h->g.nt; ++it) {
            launch_acou_v(st, sh, w.bg, w.bn, w.m, h->d_wav, it, 1.f, do_bg, do_born);
            launch_acou_p(st, sh, w.bg, w.bn, w.m, w.hist, rec_bg ? w.rec : nullptr, do_born ? w.rec2 : nullptr, it,
                          do_bg, do_born, hmode, w.rmap);
        }
}

// Backward pass (residual in w.rec2): the gradient is accumulated in
// w.grad_pad from the stored increment, or from the background rebuilt by
// reversal from the state acou_fwd left at t = nt.
void acou_bwd(ds2d_ctx* h, const Batch& B, AcouWS& w) {
    hipStream_t st = h->st;
    (void)hipMemsetAsync(w.lm.vx, 0, (size_t)w.n * 3 * 4, st);
    (void)hipMemsetAsync(w.grad_pad, 0, (size_t)w.n * 4, st);
    if (w.fused) {
        AcouFusedArgs even = acou_fused_args(h, w, w.lm, w.lm2), odd = acou_fused_args(h, w, w.lm2, w.lm);
        even.grad = odd.grad = w.grad_pad;
        even.res = odd.res = w.rec2;
        for (size_t gi = 0; gi < w.gchunks.size(); ++gi)
            for (int it = h->g.nt - 1, k = 0; it >= 0; --it, ++k)
                launch_acou_adj_fused(st, B.shape, (k & 1) ? odd : even, w.chunks + w.gchunks[gi].first,
                                      w.gchunks[gi].second, w.rmap, it, ACOU_THREADS);
        return;
    }
    for (const BatchShape& sh : launch_groups(h, B)) {
        launch_acou_inj(st, sh, w.lm.p, w.rec2, h->g.nt - 1);
        for (int it = h->g.nt - 1; it >= 0; --it) {
            launch_acou_adj_p(st, sh, w.bg, w.lm, w.m, w.grad_pad, w.store ? w.hist : nullptr, it);
            launch_acou_adj_v(st, sh, w.bg, w.lm, w.m, h->d_wav, it, w.store ? 0 : 1, w.rmap, w.rec2);
        }
    }
}

// Padded per-slot gradient -> global (deterministic slot order), overwrite.
int acou_grad_to_global(ds2d_ctx* h, const Batch& B, AcouWS& w, float* d_glob) {
    for (size_t s = 0; s < B.slots.size(); ++s) {
        const SlotDesc& S = B.slots[s];
        HIP_TRY(hipMemcpy2DAsync(w.win + S.win_off, (size_t)S.nz * 4,
                                 w.grad_pad + S.off + (size_t)S.pml * S.nzp + S.pml, (size_t)S.nzp * 4,
                                 (size_t)S.nz * 4, S.nx, hipMemcpyDeviceToDevice, h->st));
    }
    const long long A = (long long)h->g.allnx * h->g.allnz;
    HIP_TRY(hipMemsetAsync(d_glob, 0, A * 4, h->st));
    launch_gather(h->st, B.shape, w.win, w.W, d_glob, A, 1, h->g.allnx, h->g.allnz);
    return 0;
}

int acou_begin(ds2d_ctx* h, int nshots, const ds2d_shot* shots, Batch& B) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!h->have_model) return fail(DS2D_ESTATE, "no model");
    if (!shots) return fail(DS2D_EINVAL, "null shots");
    for (int s = 0; s < nshots; ++s)
        if (shots[s].boundary != 1)
            return fail(DS2D_EINVAL, "plain Born/adjoint needs the random boundary (shot %d)", s);
    return build_batch(h, nshots, shots, B);
}

long long padded_cells(const Batch& B) {
    long long cu = 0;
    for (auto& S : B.slots) cu += (long long)S.nxp * S.nzp;
    return cu;
}

// Stop rule of the CG solvers (ds2d_set_stop): once the relative residual
// after iteration k is at most the target, the remaining entries of resid
// repeat it and the solver returns.
bool reached(const ds2d_ctx* h, double rr, double dd, double* resid, int k, int niter) {
    if (h->log_every > 0 && ((k + 1) % h->log_every == 0 || k + 1 == niter))
        std::fprintf(stderr, "ds2d: iteration %d, relative data residual %.6e\n", k + 1, std::sqrt(rr / dd));
    if (!(h->stop_resid > 0) || std::sqrt(rr / dd) > h->stop_resid) return false;
    if (resid)
        for (int j = k + 1; j < niter; ++j) resid[j + 1] = resid[k + 1];
    return true;
}

// CGLS iterations on ||L x - d||^2 (image of N floats, records of Rn).  On
// entry: x, r = d - L x, rr = |r|^2, dd = |d|^2 and s = L^T r.  born(p)
// leaves L p in q; adjoint(r) leaves L^T r in s.  resid[k] = |r_k| / |d|.
template <class BornF, class AdjF>
int cgls_loop(ds2d_ctx* h, long long N, long long Rn, float* x, float* r, float* p, float* s, const float* q,
              double rr, double dd, BornF born, AdjF adjoint, int niter, double* resid) {
    hipStream_t st = h->st;
    int rc = 0;
    HIP_TRY(hipMemcpyAsync(p, s, N * 4, hipMemcpyDeviceToDevice, st));
    double gamma = host_dot(h, s, s, N, &rc);
    if (rc) return rc;
    for (int k = 0; k < niter; ++k) {
        if (gamma <= 0) {
            if (resid)
                for (int j = k; j < niter; ++j) resid[j + 1] = std::sqrt(rr / dd);
            break;
        }
        if ((rc = born(p))) return rc;  // q = L p
        const double qq = host_dot(h, q, q, Rn, &rc);
        if (rc) return rc;
        const double alpha = qq > 0 ? gamma / qq : 0.0;
        launch_axpby(st, x, p, N, (float)alpha, 1.f);
        launch_axpby(st, r, q, Rn, (float)-alpha, 1.f);
        rr = host_dot(h, r, r, Rn, &rc);
        if (rc) return rc;
        if (resid) resid[k + 1] = std::sqrt(rr / dd);
        if (reached(h, rr, dd, resid, k, niter)) break;
        if ((rc = adjoint(r))) return rc;  // s = L^T r
        const double gnew = host_dot(h, s, s, N, &rc);
        if (rc) return rc;
        const double beta = gamma > 0 ? gnew / gamma : 0.0;
        launch_axpby(st, p, s, N, 1.f, (float)beta);  // p = s + beta p
        gamma = gnew;
    }
    return 0;
}

}  // namespace

extern "C" {

int ds2d_born_forward(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* refl, float* rec_bg,
                      float* rec_born, ds2d_stats* stats) {
    int rc;
    Batch B;
    if ((rc = acou_begin(h, nshots, shots, B))) return rc;
    if (!refl) return fail(DS2D_EINVAL, "null reflectivity");
    const size_t A = (size_t)h->g.allnx * h->g.allnz;
    float* d_refl;
    if ((rc = h->fbuf("refl", A, &d_refl))) return rc;
    HIP_TRY(hipMemcpyAsync(d_refl, refl, A * 4, hipMemcpyHostToDevice, h->st));
    AcouWS w;
    if ((rc = acou_prepare(h, B, d_refl, w))) return rc;
    w.store = false;  // a single pass needs no history
    stats_begin(h);
    stats_loop_begin(h);
    acou_fwd(h, B, w, true, true, true);
    stats_loop_end(h);
    if ((rc = stats_end(h, stats, padded_cells(B) * h->g.nt * 2, h->g.nt, 0))) return rc;
    if (rec_bg) HIP_TRY(hipMemcpyAsync(rec_bg, w.rec, B.total_rec * 4, hipMemcpyDeviceToHost, h->st));
    if (rec_born) HIP_TRY(hipMemcpyAsync(rec_born, w.rec2, B.total_rec * 4, hipMemcpyDeviceToHost, h->st));
    HIP_TRY(hipStreamSynchronize(h->st));
    HIP_TRY(hipGetLastError());
    return 0;
}

int ds2d_born_adjoint(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* res, float* grad,
                      ds2d_stats* stats) {
    int rc;
    Batch B;
    if ((rc = acou_begin(h, nshots, shots, B))) return rc;
    if (!res || !grad) return fail(DS2D_EINVAL, "null residual or gradient");
    AcouWS w;
    if ((rc = acou_prepare(h, B, nullptr, w))) return rc;
    stats_begin(h);
    stats_loop_begin(h);
    acou_fwd(h, B, w, true, false, false);
    HIP_TRY(hipMemcpyAsync(w.rec2, res, B.total_rec * 4, hipMemcpyHostToDevice, h->st));
    acou_bwd(h, B, w);
    stats_loop_end(h);
    const size_t A = (size_t)h->g.allnx * h->g.allnz;
    float* d_g;
    if ((rc = h->fbuf("gglob", A, &d_g))) return rc;
    if ((rc = acou_grad_to_global(h, B, w, d_g))) return rc;
    if ((rc = stats_end(h, stats, padded_cells(B) * h->g.nt * (w.store ? 2 : 3), 2LL * h->g.nt, 0))) return rc;
    HIP_TRY(hipMemcpyAsync(grad, d_g, A * 4, hipMemcpyDeviceToHost, h->st));
    HIP_TRY(hipStreamSynchronize(h->st));
    HIP_TRY(hipGetLastError());
    return 0;
}

// CGLS on ||L m - d||^2.  Store mode: one background pass fills the resident
// increment history, then every iteration is one Born pass (q = L p) and one
// adjoint pass (s = L^T r).  Reconstruct mode: every Born pass carries the
// background and every adjoint pass reverses it.
int ds2d_lsrtm_cg(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* d_obs, int niter, float* image,
                  double* resid, ds2d_stats* stats) {
    int rc;
    Batch B;
    if ((rc = acou_begin(h, nshots, shots, B))) return rc;
    if (niter < 0 || !d_obs || !image) return fail(DS2D_EINVAL, "bad CG arguments");
    const long long A = (long long)h->g.allnx * h->g.allnz;
    hipStream_t st = h->st;
    float *x, *p, *s, *d, *r, *q;
    if ((rc = h->fbuf("cg_x", A, &x)) || (rc = h->fbuf("cg_p", A, &p)) || (rc = h->fbuf("cg_s", A, &s)))
        return rc;
    const long long R = std::max<long long>(B.total_rec, 1);
    if ((rc = h->fbuf("cg_d", R, &d)) || (rc = h->fbuf("cg_r", R, &r)) || (rc = h->fbuf("cg_q", R, &q))) return rc;
    HIP_TRY(hipMemcpyAsync(x, image, A * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d, d_obs, B.total_rec * 4, hipMemcpyHostToDevice, st));
    AcouWS w;
    if ((rc = acou_prepare(h, B, x, w))) return rc;
    const long long cells = padded_cells(B);
    long long props = 0;
    stats_begin(h);
    stats_loop_begin(h);
    // r = d - L x (this pass also fills the history in store mode)
    acou_fwd(h, B, w, true, true, false);
    props += 2;
    HIP_TRY(hipMemcpyAsync(r, d, B.total_rec * 4, hipMemcpyDeviceToDevice, st));
    launch_axpby(st, r, w.rec2, B.total_rec, -1.f, 1.f);
    const double dd = host_dot(h, d, d, B.total_rec, &rc);
    if (rc) return rc;
    if (dd <= 0) return fail(DS2D_EINVAL, "observed data are all zero");
    const double rr = host_dot(h, r, r, B.total_rec, &rc);
    if (rc) return rc;
    if (resid) resid[0] = std::sqrt(rr / dd);
    auto born = [&](const float* m) -> int {  // q = L m
        int e;
        if ((e = acou_set_refl(h, B, m, w))) return e;
        acou_fwd(h, B, w, !w.store, true, false);
        props += w.store ? 1 : 2;
        HIP_TRY(hipMemcpyAsync(q, w.rec2, B.total_rec * 4, hipMemcpyDeviceToDevice, st));
        return 0;
    };
    auto adjoint = [&](const float* res) -> int {  // s = L^T res
        HIP_TRY(hipMemcpyAsync(w.rec2, res, B.total_rec * 4, hipMemcpyDeviceToDevice, st));
        acou_bwd(h, B, w);
        props += w.store ? 1 : 2;
        return acou_grad_to_global(h, B, w, s);
    };
    if ((rc = adjoint(r))) return rc;
    if ((rc = cgls_loop(h, A, B.total_rec, x, r, p, s, q, rr, dd, born, adjoint, niter, resid))) return rc;
    stats_loop_end(h);
    if ((rc = stats_end(h, stats, cells * h->g.nt * props, 0, 0))) return rc;
    HIP_TRY(hipMemcpyAsync(image, x, A * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipGetLastError());
    return 0;
}

int ds2d_set_fused(ds2d_t* h, int mode) {
    if (!h) return fail(DS2D_EINVAL, "null handle");
    if (mode < -1 || (mode > 1 && mode < 8) || mode > ACOU_FUSED_MAX_W)
        return fail(DS2D_EINVAL, "mode must be -1, 0, 1 or an x width in [8, %d], got %d", ACOU_FUSED_MAX_W, mode);
    h->acou_fused = mode;
    return 0;
}

int ds2d_set_stop(ds2d_t* h, double target) {
    if (!h) return fail(DS2D_EINVAL, "null handle");
    if (!(target >= 0) || target >= 1) return fail(DS2D_EINVAL, "stop target must be in [0, 1), got %g", target);
    h->stop_resid = target;
    return 0;
}

int ds2d_set_log(ds2d_t* h, int every) {
    if (!h) return fail(DS2D_EINVAL, "null handle");
    if (every < 0) return fail(DS2D_EINVAL, "log interval must be >= 0");
    h->log_every = every;
    return 0;
}

int ds2d_set_batch(ds2d_t* h, int shots_per_launch) {
    if (!h) return fail(DS2D_EINVAL, "null handle");
    if (shots_per_launch < 0) return fail(DS2D_EINVAL, "shots per launch must be >= 0");
    h->per_launch = shots_per_launch;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Plain-L2 elastic Born pair (PP/PS) and its CG LSRTM
// ---------------------------------------------------------------------------
namespace {

struct EbWS {
    EbField bg{}, bn{}, lm{};
    EbAux aux{};
    EbMed m{};
    float *gpp = nullptr, *gps = nullptr;  // per-slot padded gradients (n each)
    float* win = nullptr;                  // per-slot window images, PP then PS (2 W)
    float* rec = nullptr;                  // scattered records: z then x (2 R)
    float* res = nullptr;                  // residual records: z then x (2 R)
    float* hist = nullptr;                 // background strains exx | ezz | exz (3 H)
    bool store = false;
    int* rmap = nullptr;
    long long n = 0, R = 0, W = 0, H = 0;
    // fused forward step: second state (10 n) and chunk lists per launch group
    float* alt = nullptr;
    ChunkDesc* chunks = nullptr;
    std::vector<std::pair<int, int>> gchunks;
    int threads = 256;
    StepSampler* smp = nullptr;  // when set, the fused forward launches are timed (sampled steps)
};

EbField eb_field(float* base, long long n) {
    return EbField{base, base + n, base + 2 * n, base + 3 * n, base + 4 * n};
}

int eb_begin(ds2d_ctx* h, int nshots, const ds2d_shot* shots, Batch& B) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!h->have_model || !h->d_vs) return fail(DS2D_ESTATE, "no elastic model (vp, vs)");
    if (h->g.physics != 0) return fail(DS2D_EINVAL, "the elastic Born pair needs physics = 0");
    if (!shots) return fail(DS2D_EINVAL, "null shots");
    for (int s = 0; s < nshots; ++s)
        if (shots[s].boundary != 1)
            return fail(DS2D_EINVAL, "plain elastic Born/adjoint needs the random boundary (shot %d)", s);
    return build_batch(h, nshots, shots, B);
}

// Medium (with the PP/PS image windows when given) and workspace.
int eb_set_image(ds2d_ctx* h, const Batch& B, const float* pp_dev, const float* ps_dev, EbWS& w) {
    Med m;
    int rc;
    if ((rc = alloc_med(h, w.n, m))) return rc;
    SetupArgs a = setup_args(h, false);
    a.img_pp = pp_dev;
    a.img_ps = ps_dev;
    launch_setup(h->st, B.shape, a, m);
    if (!pp_dev) HIP_TRY(hipMemsetAsync(m.dmp, 0, (size_t)w.n * 4, h->st));
    if (!ps_dev) HIP_TRY(hipMemsetAsync(m.dms, 0, (size_t)w.n * 4, h->st));
    w.m = EbMed{m.rho, m.lam, m.mu, m.dmp, m.dms};
    return 0;
}

int eb_prepare(ds2d_ctx* h, const Batch& B, const float* pp_dev, const float* ps_dev, bool want_store, EbWS& w) {
    int rc;
    const long long n = B.total_cells;
    w.n = n;
    w.R = std::max<long long>(B.total_rec, 1);
    w.W = std::max<long long>(B.total_win, 1);
    w.H = std::max<long long>(B.total_hist, 1);
    float* base;
    if ((rc = h->fbuf("eborn", (size_t)n * 22, &base))) return rc;
    w.bg = eb_field(base, n);
    w.bn = eb_field(base + 5 * n, n);
    w.lm = eb_field(base + 10 * n, n);
    w.gpp = base + 15 * n;
    w.gps = base + 16 * n;
    w.aux = EbAux{base + 17 * n, base + 18 * n, base + 19 * n, base + 20 * n, base + 21 * n};
    if ((rc = eb_set_image(h, B, pp_dev, ps_dev, w))) return rc;
    if ((rc = h->fbuf("ebrec", (size_t)w.R * 4, &w.rec))) return rc;
    w.res = w.rec + 2 * w.R;
    if ((rc = h->fbuf("ebwin", (size_t)w.W * 2, &w.win))) return rc;
    // resident background strains when they fit (3 floats per physical cell and step)
    const double hist_gb = (double)B.total_hist * 3 * 4 / 1e9;
    const double budget = h->g.hbm_budget_gb > 0 ? h->g.hbm_budget_gb : 64.0;
    w.store = want_store && (h->g.bg_mode == 1 || (h->g.bg_mode == 0 && hist_gb <= budget));
    if (w.store && (rc = h->fbuf("ehist", (size_t)w.H * 3, &w.hist))) return rc;
    if ((rc = h->fbuf("ebrmap", std::max<long long>(B.total_rmap, 1), &w.rmap))) return rc;
    HIP_TRY(hipMemsetAsync(w.rmap, 0xff, B.total_rmap * 4, h->st));
    launch_fast_rmap(h->st, B.shape, w.rmap);
    // fused forward step (32-bit byte offsets within a field: n < 2^30 cells)
    w.alt = nullptr;
    if (h->eb_fused && n < (1LL << 30)) {
        const int W = h->eb_w > 0 ? h->eb_w : 64;
        w.threads = eb_fused_threads_ok(h->eb_t) ? h->eb_t : 256;
        const int rows = w.threads - 2 * FUSED_HALO;
        std::vector<ChunkDesc> ch;
        w.gchunks.clear();
        const int ns = (int)B.slots.size();
        const int gs = h->per_launch > 0 ? std::min(h->per_launch, ns) : ns;  // the groups of launch_groups
        for (int s0 = 0; s0 < ns; s0 += gs) {
            const int first = (int)ch.size();
            for (int s = s0; s < std::min(ns, s0 + gs); ++s)
                for (int x0 = 0; x0 < B.slots[s].nxp; x0 += W)
                    for (int z0 = 0; z0 < B.slots[s].nzp; z0 += rows)
                        ch.push_back(ChunkDesc{s - s0, x0, std::min(x0 + W, B.slots[s].nxp), z0});
            w.gchunks.push_back({first, (int)ch.size() - first});
        }
        if ((rc = h->fbuf("ebchunks", ch.size(), &w.chunks))) return rc;
        HIP_TRY(hipMemcpyAsync(w.chunks, ch.data(), ch.size() * sizeof(ChunkDesc), hipMemcpyHostToDevice, h->st));
        // the second state; when it does not fit, the per-cell kernels (one state) run instead
        if (h->fbuf("eborn2", (size_t)n * 10, &w.alt)) {
            w.alt = nullptr;
            (void)hipGetLastError();
        }
    }
    return 0;
}

// Forward pass from rest into w.rec (z then x).  do_bg: propagate the
// background (storing its strains when w.store); otherwise the scattered
// field reads the stored strains.
void eb_fwd(ds2d_ctx* h, const Batch& B, EbWS& w, bool do_bg) {
    hipStream_t st = h->st;
    (void)hipMemsetAsync(w.bg.vx, 0, (size_t)w.n * 10 * 4, st);  // bg + bn
    const int hmode = do_bg ? (w.store ? 1 : 0) : 2;
    if (do_bg && w.alt) {
        // one fused launch per step, state ping-ponged between the workspace
        // (bg then bn, 10 n floats) and w.alt; the final state goes back to
        // the workspace, where the adjoint pass expects it
        const int nt = h->g.nt;
        int g = 0;
        for (const BatchShape& sh : launch_groups(h, B)) {
            const std::pair<int, int> gc = w.gchunks[g++];
            float *a = w.bg.vx, *b = w.alt;
            for (int it = 0; it < nt; ++it) {
                EbFusedArgs q{a, b, w.n, w.m, w.hist, w.H, w.rec, w.rec + w.R};
                const bool tk = w.smp && g == 1 && w.smp->take(it);
                if (tk) w.smp->mark(st, 0);
                launch_eb_fused(st, sh, q, w.chunks + gc.first, gc.second, w.rmap, h->wav[it], 1.f, it, hmode,
                                w.threads);
                if (tk) {
                    w.smp->mark(st, 1);
                    w.smp->mark(st, 2);
                }
                std::swap(a, b);
            }
        }
        if (nt & 1) (void)hipMemcpyAsync(w.bg.vx, w.alt, (size_t)w.n * 10 * 4, hipMemcpyDeviceToDevice, st);
        return;
    }
    for (const BatchShape& sh : launch_groups(h, B))
        for (int it = 0; it < h->g.nt; ++it) {
            launch_eb_v(st, sh, w.bg, w.bn, w.m.rho, h->wav[it], 1.f, do_bg, 1);
            launch_eb_s(st, sh, w.bg, w.bn, w.m, w.hist, w.H, it, 1.f, do_bg, 1, hmode, w.rmap, w.rec, w.rec + w.R);
        }
}

// Background only (fills the strain history), from rest to t = nt.
void eb_bg(ds2d_ctx* h, const Batch& B, EbWS& w) {
    hipStream_t st = h->st;
    (void)hipMemsetAsync(w.bg.vx, 0, (size_t)w.n * 5 * 4, st);
    for (const BatchShape& sh : launch_groups(h, B))
        for (int it = 0; it < h->g.nt; ++it) {
            launch_eb_v(st, sh, w.bg, w.bn, w.m.rho, h->wav[it], 1.f, 1, 0);
            launch_eb_s(st, sh, w.bg, w.bn, w.m, w.hist, w.H, it, 1.f, 1, 0, w.store ? 1 : 0, w.rmap, nullptr,
                        nullptr);
        }
}

// Backward pass (residual in w.res): gradients in w.gpp / w.gps.  Without the
// history the background, left at t = nt by the forward pass, is reversed.
void eb_bwd(ds2d_ctx* h, const Batch& B, EbWS& w) {
    hipStream_t st = h->st;
    (void)hipMemsetAsync(w.lm.vx, 0, (size_t)w.n * 12 * 4, st);  // lm + gpp + gps + adjoint products
    const int nt = h->g.nt, undo = w.store ? 0 : 1;
    int g = 0;
    for (const BatchShape& sh : launch_groups(h, B)) {
        // Without the history the background is rebuilt alongside: the stress of
        // step nt-1 comes off here, then each iteration takes the velocity of
        // step it+1 (with k_eb_adj_v) and the stress of step it (with k_eb_adj_s).
        if (undo && nt >= 2)
            launch_eb_s(st, sh, w.bg, w.bn, w.m, nullptr, w.H, nt - 1, -1.f, 1, 0, 0, w.rmap, nullptr, nullptr);
        if (w.alt) {
            // one fused launch per step: adjoint state (and the background when
            // it is reversed) ping-ponged between the workspace and w.alt
            const std::pair<int, int> gc = w.gchunks[g++];
            float *alm = w.lm.vx, *abg = w.bg.vx, *blm = w.alt, *bbg = w.alt + 5 * w.n;
            for (int it = nt - 2; it >= 0; --it) {
                EbAdjArgs q{alm, abg, blm, bbg, w.n, w.m, w.store ? w.hist : nullptr, w.H, w.gpp, w.gps,
                            w.res, w.res + w.R};
                launch_eb_adj_fused(st, sh, q, w.chunks + gc.first, gc.second, w.rmap, h->wav[it + 1], it, it + 1,
                                    undo != 0, w.threads);
                std::swap(alm, blm);
                if (undo) std::swap(abg, bbg);
            }
            continue;
        }
        for (int it = nt - 2; it >= 0; --it) {
            launch_eb_adj_v(st, sh, w.lm, w.aux, w.bg, w.m, w.rmap, w.res, w.res + w.R, it + 1, undo,
                            h->wav[it + 1]);
            launch_eb_adj_s(st, sh, w.lm, w.aux, w.bg, w.m, w.store ? w.hist : nullptr, w.H, w.gpp, w.gps, it, undo);
        }
    }
}

// Padded per-slot gradients -> global PP then PS (deterministic slot order), overwrite.
int eb_grad_to_global(ds2d_ctx* h, const Batch& B, EbWS& w, float* d_glob) {
    for (size_t s = 0; s < B.slots.size(); ++s) {
        const SlotDesc& S = B.slots[s];
        const size_t o = S.off + (size_t)S.pml * S.nzp + S.pml;
        HIP_TRY(hipMemcpy2DAsync(w.win + S.win_off, (size_t)S.nz * 4, w.gpp + o, (size_t)S.nzp * 4, (size_t)S.nz * 4,
                                 S.nx, hipMemcpyDeviceToDevice, h->st));
        HIP_TRY(hipMemcpy2DAsync(w.win + w.W + S.win_off, (size_t)S.nz * 4, w.gps + o, (size_t)S.nzp * 4,
                                 (size_t)S.nz * 4, S.nx, hipMemcpyDeviceToDevice, h->st));
    }
    const long long A = (long long)h->g.allnx * h->g.allnz;
    HIP_TRY(hipMemsetAsync(d_glob, 0, 2 * A * 4, h->st));
    launch_gather(h->st, B.shape, w.win, w.W, d_glob, A, 2, h->g.allnx, h->g.allnz);
    return 0;
}

// Padded cell-updates of one pass over the batch per propagated wavefield.
long long eb_cells(const Batch& B, int nt) {
    long long cu = 0;
    for (auto& S : B.slots) cu += (long long)S.nxp * S.nzp;
    return cu * nt;
}

}  // namespace

extern "C" {

int ds2d_eborn_forward(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* img_pp, const float* img_ps,
                       float* rec_z, float* rec_x, ds2d_stats* stats) {
    int rc;
    Batch B;
    if ((rc = eb_begin(h, nshots, shots, B))) return rc;
    if (!img_pp || !img_ps || !rec_z || !rec_x) return fail(DS2D_EINVAL, "null image or record pointer");
    const size_t A = (size_t)h->g.allnx * h->g.allnz;
    float* d_img;
    if ((rc = h->fbuf("ebimg", 2 * A, &d_img))) return rc;
    HIP_TRY(hipMemcpyAsync(d_img, img_pp, A * 4, hipMemcpyHostToDevice, h->st));
    HIP_TRY(hipMemcpyAsync(d_img + A, img_ps, A * 4, hipMemcpyHostToDevice, h->st));
    EbWS w;
    if ((rc = eb_prepare(h, B, d_img, d_img + A, false, w))) return rc;
    StepSampler smp;
    if ((rc = smp.init(h->g.nt))) return rc;
    w.smp = &smp;
    stats_begin(h);
    stats_loop_begin(h);
    eb_fwd(h, B, w, true);
    stats_loop_end(h);
    if ((rc = stats_end(h, stats, 2 * eb_cells(B, h->g.nt), 2LL * h->g.nt, 0))) return rc;
    // the fused forward step (first launch group): 100 B per padded cell (both states in and out, rho,
    // lambda, mu, the two image samples)
    double kf = 0, unused = 0;
    smp.averages(&kf, &unused);
    h->last.kernel_avg_ms = w.alt ? kf : 0;
    h->last.kernel_samples = w.alt ? smp.n : 0;
    long long cells0 = 0;  // padded cells of the first launch group
    {
        const std::vector<BatchShape> gr = launch_groups(h, B);
        for (int s = 0; s < gr[0].nslots; ++s) cells0 += (long long)B.slots[s].nxp * B.slots[s].nzp;
    }
    h->last.alg_bytes = w.alt ? (double)cells0 * 25 * 4 : 0;
    if (stats) *stats = h->last;
    HIP_TRY(hipMemcpyAsync(rec_z, w.rec, B.total_rec * 4, hipMemcpyDeviceToHost, h->st));
    HIP_TRY(hipMemcpyAsync(rec_x, w.rec + w.R, B.total_rec * 4, hipMemcpyDeviceToHost, h->st));
    HIP_TRY(hipStreamSynchronize(h->st));
    HIP_TRY(hipGetLastError());
    return 0;
}

int ds2d_eborn_adjoint(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* res_z, const float* res_x,
                       float* g_pp, float* g_ps, ds2d_stats* stats) {
    int rc;
    Batch B;
    if ((rc = eb_begin(h, nshots, shots, B))) return rc;
    if (!res_z || !res_x || !g_pp || !g_ps) return fail(DS2D_EINVAL, "null residual or gradient pointer");
    EbWS w;
    if ((rc = eb_prepare(h, B, nullptr, nullptr, true, w))) return rc;
    stats_begin(h);
    stats_loop_begin(h);
    eb_bg(h, B, w);
    HIP_TRY(hipMemcpyAsync(w.res, res_z, B.total_rec * 4, hipMemcpyHostToDevice, h->st));
    HIP_TRY(hipMemcpyAsync(w.res + w.R, res_x, B.total_rec * 4, hipMemcpyHostToDevice, h->st));
    eb_bwd(h, B, w);
    stats_loop_end(h);
    const size_t A = (size_t)h->g.allnx * h->g.allnz;
    float* d_g;
    if ((rc = h->fbuf("ebglob", 2 * A, &d_g))) return rc;
    if ((rc = eb_grad_to_global(h, B, w, d_g))) return rc;
    if ((rc = stats_end(h, stats, eb_cells(B, h->g.nt) * (w.store ? 2 : 3), 2LL * h->g.nt, 0))) return rc;
    HIP_TRY(hipMemcpyAsync(g_pp, d_g, A * 4, hipMemcpyDeviceToHost, h->st));
    HIP_TRY(hipMemcpyAsync(g_ps, d_g + A, A * 4, hipMemcpyDeviceToHost, h->st));
    HIP_TRY(hipStreamSynchronize(h->st));
    HIP_TRY(hipGetLastError());
    return 0;
}

// ELSRTM_CALCULATE_ALPHA (FD2DGPU.cu:4885-5035) on the plain elastic pair:
// r = L m - d, q = L g; num = <q, r>, den = <q, q> summed in FP64 over all
// shots (the step length of m - alpha g is num / den).
int ds2d_eborn_line_search(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* img_pp, const float* img_ps,
                           const float* g_pp, const float* g_ps, const float* d_z, const float* d_x, double* num,
                           double* den) {
    int rc;
    Batch B;
    if ((rc = eb_begin(h, nshots, shots, B))) return rc;
    if (!img_pp || !img_ps || !g_pp || !g_ps || !d_z || !d_x || !num || !den)
        return fail(DS2D_EINVAL, "null line-search argument");
    const long long A = (long long)h->g.allnx * h->g.allnz;
    const long long R = std::max<long long>(B.total_rec, 1);
    hipStream_t st = h->st;
    float *m, *r;
    if ((rc = h->fbuf("els_m", 4 * A, &m)) || (rc = h->fbuf("els_r", 2 * R, &r))) return rc;
    HIP_TRY(hipMemcpyAsync(m, img_pp, A * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(m + A, img_ps, A * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(m + 2 * A, g_pp, A * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(m + 3 * A, g_ps, A * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(r, 0, 2 * R * 4, st));
    HIP_TRY(hipMemcpyAsync(r, d_z, B.total_rec * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(r + R, d_x, B.total_rec * 4, hipMemcpyHostToDevice, st));
    EbWS w;
    if ((rc = eb_prepare(h, B, m, m + A, false, w))) return rc;
    eb_fwd(h, B, w, true);
    launch_axpby(st, r, w.rec, 2 * R, 1.f, -1.f);  // r = L m - d
    if ((rc = eb_set_image(h, B, m + 2 * A, m + 3 * A, w))) return rc;
    eb_fwd(h, B, w, true);  // q = L g in w.rec
    *num = host_dot(h, w.rec, r, 2 * R, &rc);
    if (rc) return rc;
    *den = host_dot(h, w.rec, w.rec, 2 * R, &rc);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    return 0;
}

// CGLS on ||L [pp; ps] - [d_z; d_x]||^2 with the plain elastic Born pair.
// Store mode: one background pass fills the strain history, then every
// iteration is one scattered-field pass and one adjoint pass; otherwise each
// pass carries the background (forward) or reverses it (adjoint).
int ds2d_elsrtm_cg(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* d_z, const float* d_x, int niter,
                   float* img_pp, float* img_ps, double* resid, ds2d_stats* stats) {
    int rc;
    Batch B;
    if ((rc = eb_begin(h, nshots, shots, B))) return rc;
    if (niter < 0 || !d_z || !d_x || !img_pp || !img_ps) return fail(DS2D_EINVAL, "bad CG arguments");
    const long long A = (long long)h->g.allnx * h->g.allnz;
    hipStream_t st = h->st;
    EbWS w;
    float *x, *p, *s, *d, *r;
    if ((rc = h->fbuf("ecg_x", 2 * A, &x)) || (rc = h->fbuf("ecg_p", 2 * A, &p)) ||
        (rc = h->fbuf("ecg_s", 2 * A, &s)))
        return rc;
    const long long R = std::max<long long>(B.total_rec, 1);
    if ((rc = h->fbuf("ecg_d", 2 * R, &d)) || (rc = h->fbuf("ecg_r", 2 * R, &r))) return rc;
    HIP_TRY(hipMemcpyAsync(x, img_pp, A * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(x + A, img_ps, A * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(d, 0, 2 * R * 4, st));
    HIP_TRY(hipMemcpyAsync(d, d_z, B.total_rec * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d + R, d_x, B.total_rec * 4, hipMemcpyHostToDevice, st));
    if ((rc = eb_prepare(h, B, x, x + A, true, w))) return rc;
    long long props = 0;
    stats_begin(h);
    stats_loop_begin(h);
    // r = d - L x (this pass also fills the history in store mode)
    eb_fwd(h, B, w, true);
    props += 2;
    HIP_TRY(hipMemcpyAsync(r, d, 2 * R * 4, hipMemcpyDeviceToDevice, st));
    launch_axpby(st, r, w.rec, 2 * R, -1.f, 1.f);
    const double dd = host_dot(h, d, d, 2 * R, &rc);
    if (rc) return rc;
    if (dd <= 0) return fail(DS2D_EINVAL, "observed data are all zero");
    const double rr = host_dot(h, r, r, 2 * R, &rc);
    if (rc) return rc;
    if (resid) resid[0] = std::sqrt(rr / dd);
    auto born = [&](const float* m) -> int {  // L m in w.rec
        int e;
        if ((e = eb_set_image(h, B, m, m + A, w))) return e;
        eb_fwd(h, B, w, !w.store);
        props += w.store ? 1 : 2;
        return 0;
    };
    auto adjoint = [&](const float* res) -> int {  // s = L^T res
        HIP_TRY(hipMemcpyAsync(w.res, res, 2 * R * 4, hipMemcpyDeviceToDevice, st));
        eb_bwd(h, B, w);
        props += w.store ? 1 : 2;
        return eb_grad_to_global(h, B, w, s);
    };
    if ((rc = adjoint(r))) return rc;
    if ((rc = cgls_loop(h, 2 * A, 2 * R, x, r, p, s, w.rec, rr, dd, born, adjoint, niter, resid))) return rc;
    stats_loop_end(h);
    if ((rc = stats_end(h, stats, eb_cells(B, h->g.nt) * props, 0, 0))) return rc;
    HIP_TRY(hipMemcpyAsync(img_pp, x, A * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(img_ps, x + A, A * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"

namespace {

// Learned preconditioner z = rms(g) * U(g / rms(g)): positively homogeneous,
// so the conjugacy recurrences see consistent scales across iterations.  The
// image (allnx x allnz) is zero-padded to the U-Net's multiple-of-2^(L-1) grid.
struct Precond {
    ds2d_unet* u;
    int H, W;
    float *in, *out;
    int apply(ds2d_ctx* h, const float* g, float* z) {
        hipStream_t st = h->st;
        const int nx = h->g.allnx, nz = h->g.allnz;
        const long long A = (long long)nx * nz;
        int rc;
        const double gg = host_dot(h, g, g, A, &rc);
        if (rc) return rc;
        if (!(gg > 0)) {
            HIP_TRY(hipMemsetAsync(z, 0, A * 4, st));
            return 0;
        }
        const double rms = std::sqrt(gg / A);
        HIP_TRY(hipMemsetAsync(in, 0, (size_t)H * W * 4, st));
        HIP_TRY(hipMemcpy2DAsync(in, (size_t)W * 4, g, (size_t)nz * 4, (size_t)nz * 4, nx, hipMemcpyDeviceToDevice, st));
        launch_axpby(st, in, in, (long long)H * W, 0.f, (float)(1.0 / rms));
        if ((rc = unet_infer_dev(u, 1, H, W, in, out, st))) return rc;
        HIP_TRY(hipMemcpy2DAsync(z, (size_t)nz * 4, out, (size_t)W * 4, (size_t)nz * 4, nx, hipMemcpyDeviceToDevice, st));
        launch_axpby(st, z, z, A, 0.f, (float)rms);
        return 0;
    }
    // nch images of allnx x allnz one after another, each preconditioned alone
    int apply_n(ds2d_ctx* h, int nch, const float* g, float* z) {
        const long long A = (long long)h->g.allnx * h->g.allnz;
        for (int k = 0; k < nch; ++k) {
            const int rc = apply(h, g + k * A, z + k * A);
            if (rc) return rc;
        }
        return 0;
    }
};

// Flexible preconditioned CG (Polak-Ribiere+, exact line search on the
// quadratic) with the learned direction, falling back to plain CGLS from the
// current image when a learned step would lower |r| by less than 1%.
// On entry: x the image (nch * A), r = d - L x (Rn samples), rr = |r|^2,
// dd = |d|^2, g = L^T r.  born(m, q): q = L m; adjoint(r, g): g = L^T r.
template <class BornF, class AdjF>
int learned_cg(ds2d_ctx* h, Precond& P, int nch, long long A, long long Rn, float* x, float* r, float* q, float* p,
               float* g, float* z, float* gn, float* zn, double rr, double dd, BornF born, AdjF adjoint, int niter,
               double* resid, int* accepted_out) {
    hipStream_t st = h->st;
    int rc = 0;
    const long long N = nch * A;
    auto dot = [&](const float* a, const float* b, long long n) { return host_dot(h, a, b, n, &rc); };
    bool learned = true;
    int accepted = 0;
    double gz = 0, gamma = 0;
    if ((rc = P.apply_n(h, nch, g, z))) return rc;
    HIP_TRY(hipMemcpyAsync(p, z, N * 4, hipMemcpyDeviceToDevice, st));
    gz = dot(g, z, N);
    if (rc) return rc;
    for (int k = 0; k < niter; ++k) {
        if (learned) {
            if ((rc = born(p, q))) return rc;
            const double qr = dot(q, r, Rn), qq = dot(q, q, Rn);
            if (rc) return rc;
            const double rr_new = qq > 0 ? rr - qr * qr / qq : rr;
            if (qq > 0 && qr > 0 && rr_new < 0.99 * 0.99 * rr) {
                const double alpha = qr / qq;
                launch_axpby(st, x, p, N, (float)alpha, 1.f);
                launch_axpby(st, r, q, Rn, (float)-alpha, 1.f);
                rr = dot(r, r, Rn);
                if (rc) return rc;
                if (resid) resid[k + 1] = std::sqrt(rr / dd);
                ++accepted;
                if (reached(h, rr, dd, resid, k, niter)) break;
                if ((rc = adjoint(r, gn))) return rc;
                if ((rc = P.apply_n(h, nch, gn, zn))) return rc;
                const double zg_new = dot(zn, gn, N), zg_old = dot(zn, g, N);
                if (rc) return rc;
                const double beta = gz > 0 ? std::max(0.0, (zg_new - zg_old) / gz) : 0.0;
                launch_axpby(st, p, zn, N, 1.f, (float)beta);  // p = zn + beta p
                std::swap(g, gn);
                std::swap(z, zn);
                gz = zg_new;
                continue;
            }
            // fall back: plain CGLS from the current image, s = g
            learned = false;
            HIP_TRY(hipMemcpyAsync(p, g, N * 4, hipMemcpyDeviceToDevice, st));
            gamma = dot(g, g, N);
            if (rc) return rc;
        }
        if (!(gamma > 0)) {
            if (resid)
                for (int j = k; j < niter; ++j) resid[j + 1] = std::sqrt(rr / dd);
            break;
        }
        if ((rc = born(p, q))) return rc;
        const double qq = dot(q, q, Rn);
        if (rc) return rc;
        const double alpha = qq > 0 ? gamma / qq : 0.0;
        launch_axpby(st, x, p, N, (float)alpha, 1.f);
        launch_axpby(st, r, q, Rn, (float)-alpha, 1.f);
        rr = dot(r, r, Rn);
        if (rc) return rc;
        if (resid) resid[k + 1] = std::sqrt(rr / dd);
        if (reached(h, rr, dd, resid, k, niter)) break;
        if ((rc = adjoint(r, g))) return rc;
        const double gnew = dot(g, g, N);
        if (rc) return rc;
        launch_axpby(st, p, g, N, 1.f, (float)(gnew / gamma));  // p = s + beta p
        gamma = gnew;
    }
    if (accepted_out) *accepted_out = accepted;
    return 0;
}

// The U-Net's padded grid for this model, checked against its limits.
int precond_for(ds2d_ctx* h, ds2d_unet* u, Precond& P) {
    int L, mn, mh, mw;
    unet_limits(u, &L, &mn, &mh, &mw);
    const int qd = 1 << (L - 1);
    P = Precond{u, (h->g.allnx + qd - 1) / qd * qd, (h->g.allnz + qd - 1) / qd * qd, nullptr, nullptr};
    if (P.H > mh || P.W > mw)
        return fail(DS2D_EINVAL, "model %dx%d (padded %dx%d) exceeds the U-Net's %dx%d", h->g.allnx, h->g.allnz, P.H,
                    P.W, mh, mw);
    int rc;
    if ((rc = h->fbuf("lr_in", (size_t)P.H * P.W, &P.in)) || (rc = h->fbuf("lr_out", (size_t)P.H * P.W, &P.out)))
        return rc;
    return 0;
}

}  // namespace

extern "C" {

// Learned LSRTM on the plain acoustic Born pair (learned_cg above).
int ds2d_lsrtm_learned(ds2d_t* h, ds2d_unet_t* u, int nshots, const ds2d_shot* shots, const float* d_obs,
                       int niter, float* image, double* resid, int* learned_steps, ds2d_stats* stats) {
    int rc;
    Batch B;
    if ((rc = acou_begin(h, nshots, shots, B))) return rc;
    if (!u) return fail(DS2D_EINVAL, "null U-Net handle");
    if (niter < 0 || !d_obs || !image) return fail(DS2D_EINVAL, "bad LSRTM arguments");
    Precond P{};
    if ((rc = precond_for(h, u, P))) return rc;
    const long long A = (long long)h->g.allnx * h->g.allnz;
    hipStream_t st = h->st;
    float *x, *p, *g, *z, *gn, *zn;
    if ((rc = h->fbuf("cg_x", A, &x)) || (rc = h->fbuf("cg_p", A, &p)) || (rc = h->fbuf("cg_s", A, &g)) ||
        (rc = h->fbuf("lr_z", A, &z)) || (rc = h->fbuf("lr_gn", A, &gn)) || (rc = h->fbuf("lr_zn", A, &zn)))
        return rc;
    float *d, *r, *q;
    const long long R = std::max<long long>(B.total_rec, 1);
    if ((rc = h->fbuf("cg_d", R, &d)) || (rc = h->fbuf("cg_r", R, &r)) || (rc = h->fbuf("cg_q", R, &q))) return rc;
    HIP_TRY(hipMemcpyAsync(x, image, A * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d, d_obs, B.total_rec * 4, hipMemcpyHostToDevice, st));
    AcouWS w;
    if ((rc = acou_prepare(h, B, x, w))) return rc;
    const long long cells = padded_cells(B);
    long long props = 0;
    auto born = [&](const float* m, float* out) -> int {  // out = L m (records)
        int e;
        if ((e = acou_set_refl(h, B, m, w))) return e;
        acou_fwd(h, B, w, !w.store, true, false);
        props += w.store ? 1 : 2;
        HIP_TRY(hipMemcpyAsync(out, w.rec2, B.total_rec * 4, hipMemcpyDeviceToDevice, st));
        return 0;
    };
    auto adjoint = [&](const float* res, float* out) -> int {  // out = L^T res (image)
        HIP_TRY(hipMemcpyAsync(w.rec2, res, B.total_rec * 4, hipMemcpyDeviceToDevice, st));
        acou_bwd(h, B, w);
        props += w.store ? 1 : 2;
        return acou_grad_to_global(h, B, w, out);
    };
    stats_begin(h);
    stats_loop_begin(h);
    // r = d - L x (fills the background history in store mode)
    acou_fwd(h, B, w, true, true, false);
    props += 2;
    HIP_TRY(hipMemcpyAsync(r, d, B.total_rec * 4, hipMemcpyDeviceToDevice, st));
    launch_axpby(st, r, w.rec2, B.total_rec, -1.f, 1.f);
    const double dd = host_dot(h, d, d, B.total_rec, &rc);
    if (rc) return rc;
    if (!(dd > 0)) return fail(DS2D_EINVAL, "observed data are all zero");
    const double rr = host_dot(h, r, r, B.total_rec, &rc);
    if (rc) return rc;
    if (resid) resid[0] = std::sqrt(rr / dd);
    if ((rc = adjoint(r, g))) return rc;
    if ((rc = learned_cg(h, P, 1, A, B.total_rec, x, r, q, p, g, z, gn, zn, rr, dd, born, adjoint, niter, resid,
                         learned_steps)))
        return rc;
    stats_loop_end(h);
    if ((rc = stats_end(h, stats, cells * h->g.nt * props, 0, 0))) return rc;
    HIP_TRY(hipMemcpyAsync(image, x, A * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipGetLastError());
    return 0;
}

// Learned PP/PS LSRTM on the plain elastic Born pair: the U-Net preconditions
// the PP and the PS gradient images one at a time.
int ds2d_elsrtm_learned(ds2d_t* h, ds2d_unet_t* u, int nshots, const ds2d_shot* shots, const float* d_z,
                        const float* d_x, int niter, float* img_pp, float* img_ps, double* resid, int* learned_steps,
                        ds2d_stats* stats) {
    int rc;
    Batch B;
    if ((rc = eb_begin(h, nshots, shots, B))) return rc;
    if (!u) return fail(DS2D_EINVAL, "null U-Net handle");
    if (niter < 0 || !d_z || !d_x || !img_pp || !img_ps) return fail(DS2D_EINVAL, "bad LSRTM arguments");
    Precond P{};
    if ((rc = precond_for(h, u, P))) return rc;
    const long long A = (long long)h->g.allnx * h->g.allnz;
    hipStream_t st = h->st;
    float *x, *p, *g, *z, *gn, *zn;
    if ((rc = h->fbuf("ecg_x", 2 * A, &x)) || (rc = h->fbuf("ecg_p", 2 * A, &p)) ||
        (rc = h->fbuf("ecg_s", 2 * A, &g)) || (rc = h->fbuf("elr_z", 2 * A, &z)) ||
        (rc = h->fbuf("elr_gn", 2 * A, &gn)) || (rc = h->fbuf("elr_zn", 2 * A, &zn)))
        return rc;
    const long long R = std::max<long long>(B.total_rec, 1);
    float *d, *r, *q;
    if ((rc = h->fbuf("ecg_d", 2 * R, &d)) || (rc = h->fbuf("ecg_r", 2 * R, &r)) || (rc = h->fbuf("ecg_q", 2 * R, &q)))
        return rc;
    HIP_TRY(hipMemcpyAsync(x, img_pp, A * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(x + A, img_ps, A * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(d, 0, 2 * R * 4, st));
    HIP_TRY(hipMemcpyAsync(d, d_z, B.total_rec * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d + R, d_x, B.total_rec * 4, hipMemcpyHostToDevice, st));
    EbWS w;
    if ((rc = eb_prepare(h, B, x, x + A, true, w))) return rc;
    long long props = 0;
    auto born = [&](const float* m, float* out) -> int {
        int e;
        if ((e = eb_set_image(h, B, m, m + A, w))) return e;
        eb_fwd(h, B, w, !w.store);
        props += w.store ? 1 : 2;
        HIP_TRY(hipMemcpyAsync(out, w.rec, 2 * R * 4, hipMemcpyDeviceToDevice, st));
        return 0;
    };
    auto adjoint = [&](const float* res, float* out) -> int {
        HIP_TRY(hipMemcpyAsync(w.res, res, 2 * R * 4, hipMemcpyDeviceToDevice, st));
        eb_bwd(h, B, w);
        props += w.store ? 1 : 2;
        return eb_grad_to_global(h, B, w, out);
    };
    stats_begin(h);
    stats_loop_begin(h);
    eb_fwd(h, B, w, true);
    props += 2;
    HIP_TRY(hipMemcpyAsync(r, d, 2 * R * 4, hipMemcpyDeviceToDevice, st));
    launch_axpby(st, r, w.rec, 2 * R, -1.f, 1.f);
    const double dd = host_dot(h, d, d, 2 * R, &rc);
    if (rc) return rc;
    if (!(dd > 0)) return fail(DS2D_EINVAL, "observed data are all zero");
    const double rr = host_dot(h, r, r, 2 * R, &rc);
    if (rc) return rc;
    if (resid) resid[0] = std::sqrt(rr / dd);
    if ((rc = adjoint(r, g))) return rc;
    if ((rc = learned_cg(h, P, 2, A, 2 * R, x, r, q, p, g, z, gn, zn, rr, dd, born, adjoint, niter, resid,
                         learned_steps)))
        return rc;
    stats_loop_end(h);
    if ((rc = stats_end(h, stats, eb_cells(B, h->g.nt) * props, 0, 0))) return rc;
    HIP_TRY(hipMemcpyAsync(img_pp, x, A * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(img_ps, x + A, A * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipGetLastError());
    return 0;
}

// main.cpp:1232-1303 sjcgbeta / sjcglength / sjcgsolver: conjugate direction
// cg = -g1 + beta cg with beta = max(0, min(beta_HS, beta_DY)) (0 at iter 0),
// g0 <- g1, step alpha = 1e-4 at iter 0, else 0.1 max|m| / max|cg| (first
// index of the largest magnitude), m += alpha cg.  FP32 sums, as there.
int ds2d_sjcg_update(int n, float* m, float* cg, const float* g1, float* g0, int iter, float* alpha_out,
                     float* beta_out) {
    if (n <= 0 || !m || !cg || !g1 || !g0 || iter < 0) return fail(DS2D_EINVAL, "bad CG update arguments");
    float beta = 0.f;
    if (iter > 0) {
        float a = 0.f, b = 0.f, c = 0.f;
        for (int i = 0; i < n; ++i) {
            a += g1[i] * (g1[i] - g0[i]);
            b += cg[i] * (g1[i] - g0[i]);
            c += g1[i] * g1[i];
        }
        const float hs = std::fabs(b) > 0.f ? a / b : 0.f;
        const float dy = std::fabs(b) > 0.f ? c / b : 0.f;
        beta = std::max(0.f, std::min(hs, dy));
    }
    for (int i = 0; i < n; ++i) cg[i] = -1.f * g1[i] + beta * cg[i];
    std::memcpy(g0, g1, (size_t)n * sizeof(float));
    float alpha = 1e-4f;
    if (iter > 0) {
        int is = 0, ix = 0;
        for (int i = 0; i < n; ++i) {
            if (std::fabs(cg[i]) > std::fabs(cg[is])) is = i;
            if (std::fabs(m[i]) > std::fabs(m[ix])) ix = i;
        }
        alpha = 0.1f * std::fabs(m[ix]) / std::fabs(cg[is]);
    }
    for (int i = 0; i < n; ++i) m[i] = 1.f * m[i] + alpha * cg[i];
    if (alpha_out) *alpha_out = alpha;
    if (beta_out) *beta_out = beta;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// P/S-decoupled (theta/omega rotation) modelling and RTM; elastic RTM with
// decomposed (vector P/S, Poynting up/down) imaging
// ---------------------------------------------------------------------------
namespace {

template <typename T>
int alloc_rot(ds2d_ctx* h, const char* tag, long long n, RotFieldT<T>& f) {
    T** ptrs[12] = {&f.th, &f.om, &f.th_x, &f.th_z, &f.om_x, &f.om_z, &f.vx, &f.vz, &f.vpx, &f.vpz, &f.vsx, &f.vsz};
    T* base;
    std::string name = std::string("rot_") + tag;
    int rc = h->fbuf(name.c_str(), (size_t)n * 12, &base);
    if (rc) return rc;
    for (int k = 0; k < 12; ++k) *ptrs[k] = base + (size_t)k * n;
    HIP_TRY(hipMemsetAsync(base, 0, (size_t)n * 12 * sizeof(T), h->st));
    return 0;
}

// Entry checks, batch and padded medium of the rotation and decomposed paths.
int rot_setup(ds2d_ctx* h, int nshots, const ds2d_shot* shots, Batch& B, Med& m) {
    int rc;
    if ((rc = check_h(h))) return rc;
    if (!h->have_model) return fail(DS2D_ESTATE, "no model");
    if (h->g.physics != 0) return fail(DS2D_EINVAL, "needs physics = 0 (elastic)");
    if (h->g.pml < NOP) return fail(DS2D_EINVAL, "needs a frame of at least %d cells", NOP);
    if ((rc = build_batch(h, nshots, shots, B))) return rc;
    if ((rc = alloc_med(h, B.total_cells, m))) return rc;
    launch_setup(h->st, B.shape, setup_args(h, false), m);
    return 0;
}

// Records concatenated in shot order ([ntr_s][nt] each) into a device arena.
int upload_records(ds2d_ctx* h, const Batch& B, const char* name, const float* host, float** dev) {
    int rc;
    if ((rc = h->fbuf(name, std::max<long long>(B.total_rec, 1), dev))) return rc;
    if (B.total_rec > 0) HIP_TRY(hipMemcpyAsync(*dev, host, B.total_rec * 4, hipMemcpyHostToDevice, h->st));
    return 0;
}

// Window images (nimg x total_win) -> global images (nimg x allnx*allnz) -> host.
int gather_out(ds2d_ctx* h, const Batch& B, const float* win, int nimg, float* const* out) {
    int rc;
    const long long A = (long long)h->g.allnx * h->g.allnz;
    float* g;
    if ((rc = h->fbuf("rot_glob", (size_t)A * nimg, &g))) return rc;
    HIP_TRY(hipMemsetAsync(g, 0, (size_t)A * nimg * 4, h->st));
    launch_gather(h->st, B.shape, win, B.total_win, g, A, nimg, h->g.allnx, h->g.allnz);
    for (int k = 0; k < nimg; ++k)
        if (out[k]) HIP_TRY(hipMemcpyAsync(out[k], g + (size_t)k * A, A * 4, hipMemcpyDeviceToHost, h->st));
    HIP_TRY(hipStreamSynchronize(h->st));
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int ds2d_rot_forward(ds2d_t* h, int nshots, const ds2d_shot* shots, float* rec_z, float* rec_x,
                     ds2d_stats* stats) {
    Batch B;
    Med m;
    int rc;
    if ((rc = rot_setup(h, nshots, shots, B, m))) return rc;
    const BatchShape& sh = B.shape;
    hipStream_t st = h->st;
    const int nt = h->g.nt;
    const size_t R = std::max<long long>(B.total_rec, 1);
    float* rs;
    if ((rc = h->fbuf("frec", R * 2, &rs))) return rc;
    StepSampler smp;
    if ((rc = smp.init(nt))) return rc;
    const long long n = B.total_cells;
    float* fa = nullptr;
    // one fused launch per step needs two sets of 8 fields; when they do not
    // fit, the per-cell kernels (12 fields) run instead
    if (h->rot_fused && n < (1LL << 30) && h->fbuf("rotf", (size_t)n * 16, &fa)) {
        fa = nullptr;
        (void)hipGetLastError();
    }
    if (fa) {
        // one fused launch per step (64-column chunks, 256 lanes), state
        // ping-ponged between two sets of 8 fields
        int* rmap;
        if ((rc = h->fbuf("rmap", B.total_rmap, &rmap))) return rc;
        HIP_TRY(hipMemsetAsync(fa, 0, (size_t)n * 16 * 4, st));
        HIP_TRY(hipMemsetAsync(rmap, 0xff, B.total_rmap * 4, st));
        HIP_TRY(hipMemsetAsync(rs, 0, R * 2 * 4, st));
        launch_fast_rmap(st, sh, rmap);
        const int W = 64, T = 256, rows = T - 2 * FUSED_HALO;
        std::vector<ChunkDesc> ch;
        for (int s = 0; s < nshots; ++s)
            for (int x0 = 0; x0 < B.slots[s].nxp; x0 += W)
                for (int z0 = 0; z0 < B.slots[s].nzp; z0 += rows)
                    ch.push_back(ChunkDesc{s, x0, std::min(x0 + W, B.slots[s].nxp), z0});
        ChunkDesc* d_ch;
        if ((rc = h->fbuf("rotchunks", ch.size(), &d_ch))) return rc;
        HIP_TRY(hipMemcpyAsync(d_ch, ch.data(), ch.size() * sizeof(ChunkDesc), hipMemcpyHostToDevice, st));
        stats_begin(h);
        stats_loop_begin(h);
        float *a = fa, *b = fa + 8 * n;
        for (int it = 0; it < nt; ++it) {
            const bool tk = smp.take(it);
            if (tk) smp.mark(st, 0);
            RotFusedArgs q{a, b, n, m.vp, m.vs, m.dampx, m.dampz, h->d_wav, rs, rs + R};
            launch_rot_fused(st, sh, q, d_ch, (int)ch.size(), rmap, it, T, B.any_pml);
            if (tk) {
                smp.mark(st, 1);
                smp.mark(st, 2);
            }
            std::swap(a, b);
        }
        stats_loop_end(h);
        if ((rc = stats_end(h, stats, padded_cells(B) * nt, nt, 0))) return rc;
        double kf, unused;
        smp.averages(&kf, &unused);
        // per padded cell: theta, omega, 4 velocity parts, vp, vs in, 6 fields out (56 B); the damping cells
        // add dampx, dampz and theta_x, omega_x read and written (24 B)
        h->last.kernel_avg_ms = kf;
        double pml_cells = 0;  // update-region cells of the damping strips (PML windows)
        for (auto& S : B.slots)
            if (S.boundary == 0 && S.pml >= NOP)
                pml_cells += (double)(S.nxp - 2 * NOP) * (S.nzp - 2 * NOP) -
                             (double)(S.nxp - 2 * S.pml + 1) * (S.nzp - 2 * S.pml + 1);
        h->last.alg_bytes = (double)n * 14 * 4 + pml_cells * 6 * 4;
        h->last.aux_kernel_avg_ms = 0;
        h->last.aux_alg_bytes = 0;
        h->last.kernel_samples = smp.n;
        if (stats) *stats = h->last;
        if (rec_z) HIP_TRY(hipMemcpyAsync(rec_z, rs, B.total_rec * 4, hipMemcpyDeviceToHost, st));
        if (rec_x) HIP_TRY(hipMemcpyAsync(rec_x, rs + R, B.total_rec * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        HIP_TRY(hipGetLastError());
        return 0;
    }
    RotField f;
    if ((rc = alloc_rot(h, "f", B.total_cells, f))) return rc;
    stats_begin(h);
    stats_loop_begin(h);
    for (int it = 0; it < nt; ++it) {
        const bool tk = smp.take(it);
        launch_rot_src(st, sh, f.th, h->d_wav, it);
        if (tk) smp.mark(st, 0);
        launch_rot_v(st, sh, f, m, 1.f);
        if (tk) smp.mark(st, 1);
        launch_rot_p(st, sh, f, m, 1.f);
        if (tk) smp.mark(st, 2);
        launch_record(st, sh, f.vz, f.vx, rs, rs + R, nullptr, nullptr, nullptr, nullptr, it);
    }
    stats_loop_end(h);
    if ((rc = stats_end(h, stats, padded_cells(B) * nt, nt, 0))) return rc;
    double kv, kp;
    smp.averages(&kv, &kp);
    // per padded cell: rot_v reads theta, omega, 4 velocity parts, vp, vs, dampx, dampz and writes
    // 4 parts + vx, vz (16 floats); rot_p reads vx, vz, 4 splits, dampx, dampz and writes 4 splits
    // + theta, omega (14 floats)
    const double bv = (double)B.total_cells * 16 * 4, bp = (double)B.total_cells * 14 * 4;
    const bool p_dom = kp >= kv;
    h->last.kernel_avg_ms = p_dom ? kp : kv;
    h->last.alg_bytes = p_dom ? bp : bv;
    h->last.aux_kernel_avg_ms = p_dom ? kv : kp;
    h->last.aux_alg_bytes = p_dom ? bv : bp;
    h->last.kernel_samples = smp.n;
    if (stats) *stats = h->last;
    if (rec_z) HIP_TRY(hipMemcpyAsync(rec_z, rs, B.total_rec * 4, hipMemcpyDeviceToHost, st));
    if (rec_x) HIP_TRY(hipMemcpyAsync(rec_x, rs + R, B.total_rec * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipGetLastError());
    return 0;
}

int ds2d_rot_rtm(ds2d_t* h, int nshots, const ds2d_shot* shots, const float* rec_z, const float* rec_x,
                 float* img_pp, float* img_ps) {
    if (!rec_z || !rec_x) return fail(DS2D_EINVAL, "records required");
    for (int s = 0; s < nshots && shots; ++s)
        if (shots[s].boundary != 1)
            return fail(DS2D_EINVAL,
                        "shot %d: the rotation RTM rebuilds its source wavefield by reversal, which needs the "
                        "random boundary (boundary = 1)", s);
    Batch B;
    Med m;
    int rc;
    if ((rc = rot_setup(h, nshots, shots, B, m))) return rc;
    const BatchShape& sh = B.shape;
    hipStream_t st = h->st;
    const int nt = h->g.nt;
    // both sides in FP64: the image divides by the source's own P energy, which magnifies the rounding of
    // either wavefield wherever the source field is small
    RotFieldD f, b;
    if ((rc = alloc_rot(h, "src", B.total_cells, f)) || (rc = alloc_rot(h, "rcv", B.total_cells, b))) return rc;
    float *dz, *dx, *w2;
    if ((rc = upload_records(h, B, "obs_z", rec_z, &dz)) || (rc = upload_records(h, B, "obs_x", rec_x, &dx)))
        return rc;
    if ((rc = h->fbuf("rot_win", (size_t)B.total_win * 2, &w2))) return rc;
    HIP_TRY(hipMemsetAsync(w2, 0, (size_t)B.total_win * 2 * 4, st));
    stats_begin(h);
    stats_loop_begin(h);
    for (int it = 0; it < nt; ++it) {
        launch_rot_src(st, sh, f.th, h->d_wav, it);
        launch_rot_v(st, sh, f, m, 1.f);
        launch_rot_p(st, sh, f, m, 1.f);
    }
    for (int it = nt - 1; it >= 0; --it) {
        // reversal of step it: the sample is back in theta before rot_v(-1) reads it
        launch_rot_p(st, sh, f, m, -1.f);
        launch_rot_src(st, sh, f.th, h->d_wav, it);
        launch_rot_v(st, sh, f, m, -1.f);
        launch_rot_v(st, sh, b, m, 1.f);
        launch_rot_inject(st, sh, b.vz, b.vx, dz, dx, it);
        launch_rot_p(st, sh, b, m, 1.f);
        launch_rot_image(st, sh, f, b, w2, B.total_win);
    }
    stats_loop_end(h);
    if ((rc = stats_end(h, nullptr, padded_cells(B) * nt * 3, 0, 0))) return rc;
    float* out[2] = {img_pp, img_ps};
    return gather_out(h, B, w2, 2, out);
}

int ds2d_decomp_rtm(ds2d_t* h, int nshots, const ds2d_shot* shots, int kind, const float* rec_z, float* image) {
    if (kind < 0 || kind > 6) return fail(DS2D_EINVAL, "image kind must be 0..6");
    if (!rec_z) return fail(DS2D_EINVAL, "records required");
    Batch B;
    Med m;
    int rc;
    if ((rc = rot_setup(h, nshots, shots, B, m))) return rc;
    const BatchShape& sh = B.shape;
    hipStream_t st = h->st;
    const int nt = h->g.nt;
    const long long n = B.total_cells, W = B.total_win, K = decomp_floats(kind);
    EField f;
    if ((rc = alloc_efield(h, "bg", n, f))) return rc;
    float *vps, *hist, *win, *dz;
    if ((rc = h->fbuf("helm", (size_t)n * 2, &vps))) return rc;
    if ((rc = h->fbuf("decomp_hist", (size_t)nt * W * K, &hist))) return rc;
    if ((rc = h->fbuf("rot_win", (size_t)W, &win))) return rc;
    if ((rc = upload_records(h, B, "obs_z", rec_z, &dz))) return rc;
    float *VP = vps, *VS = vps + n;
    HIP_TRY(hipMemsetAsync(vps, 0, (size_t)n * 2 * 4, st));
    HIP_TRY(hipMemsetAsync(win, 0, (size_t)W * 4, st));
    const Strain none{nullptr, nullptr, nullptr, nullptr};
    stats_begin(h);
    stats_loop_begin(h);
    for (int it = 0; it < nt; ++it) {
        launch_inject_src(st, sh, f.txx, f.tzz, h->d_wav, it);
        launch_ref_v(st, sh, f, m, 1.f, 0);
        launch_helmholtz(st, sh, f.vx, f.vz, VP, VS);
        launch_decomp(st, sh, kind, VP, VS, f, hist + (size_t)it * W * K, nullptr, 1);
        launch_ref_s(st, sh, f, m, none, 1.f, 0);
    }
    // receiver side from rest; step nt-1 is not imaged (its backward slot is never written)
    HIP_TRY(hipMemsetAsync(f.vx, 0, (size_t)n * 15 * 4, st));
    HIP_TRY(hipMemsetAsync(vps, 0, (size_t)n * 2 * 4, st));
    for (int it = nt - 1; it >= 0; --it) {
        launch_assign_rec(st, sh, f.txx, f.tzz, dz, it);
        launch_ref_v(st, sh, f, m, 1.f, 0);
        launch_helmholtz(st, sh, f.vx, f.vz, VP, VS);
        if (it < nt - 1) launch_decomp(st, sh, kind, VP, VS, f, hist + (size_t)it * W * K, win, 0);
        launch_ref_s(st, sh, f, m, none, 1.f, 0);
    }
    stats_loop_end(h);
    if ((rc = stats_end(h, nullptr, padded_cells(B) * nt * 2, 0, 0))) return rc;
    float* out[1] = {image};
    return gather_out(h, B, win, 1, out);
}

// main.cpp:461-473 ricker1.
int ds2d_ricker(int nt, float f, float dt, float* s) {
    if (!s || nt <= 0 || f <= 0 || dt <= 0) return fail(DS2D_EINVAL, "bad ricker arguments");
    const float pi = 3.14159265358979f;
    const float t0 = 1 / f;
    const int kt = (int)(t0 / dt);
    for (int i = 0; i < nt; ++i) {
        const float tt = i * dt - kt * dt;
        const float sp = pi * f * tt;
        s[i] = (float)(1.0 * std::exp(-sp * sp) * (1. - 2. * sp * sp));
    }
    return 0;
}

}  // extern "C"
