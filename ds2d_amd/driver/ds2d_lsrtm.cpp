// ds2d_lsrtm — parameter-file driver of the DS2D hot path (main.cpp counterpart).
//
//   ds2d_lsrtm <parameters>          LSRTM over the shots of the vz/vx SU files
//   ds2d_lsrtm synth <parameters>    write synthetic vz/vx SU files for the same file
//
// Reads the reference's positional parameter file (main.cpp:1349-1359):
//   dx dz pml allnx allnz minshot maxshot
//   <vp .bin> <vs .bin> <vz .su> <vx .su> <image dir>
//   f0 maxiter
//   sy gy
//   light rbc cpu_mem iointerval
// followed by optional key=value lines (extension):
//   mode=ref|plain   ref (default): elastic Born + 2nd-order Born, NCC misfit and
//                    the Adam update of lsrtm_all_iteration (main.cpp:788-1217);
//                    plain: acoustic Born CG (L2), vz file holds pressure data
//   rho=<.bin>       density model (default: the reference's constant 2000)
//   batch=<n>        shots per batched launch (default: all selected shots)
//   seed=<n>         random-boundary seed
//   synth_vp=<.bin> synth_vs=<.bin>  models used by `synth` (default: vp/vs above)
//   synth_refl=<.bin>                reflectivity for `synth` in plain mode
//   synth_nt=<n> synth_dt=<s> synth_dshot=<cells> synth_x0=<cells>
// Models are x-major FP32 (allnx columns of allnz samples, main.cpp:1464-1468).
// Only the C-ABI of libds2d.so is used; the GPU does all wavefield work.
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ds2d.h"

namespace {

[[noreturn]] void die(const char* msg, const char* extra = "") {
    std::fprintf(stderr, "ds2d_lsrtm: %s%s\n", msg, extra);
    std::exit(1);
}

void check(int rc, const char* what) {
    if (rc != 0) {
        std::fprintf(stderr, "ds2d_lsrtm: %s failed (%d): %s\n", what, rc, ds2d_last_error());
        std::exit(1);
    }
}

struct Params {
    float dx = 0, dz = 0;
    int pml = 0, allnx = 0, allnz = 0, minshot = 1, maxshot = 1;
    std::string fvp, fvs, fn1, fn2, imagedir;
    float f0 = 15;
    int maxiter = 1, sy = 0, gy = 0, light = 0, rbc = 0, cpu_mem = 0, iointerval = 1;
    std::map<std::string, std::string> kv;
    std::string get(const std::string& k, const std::string& d) const {
        auto it = kv.find(k);
        return it == kv.end() ? d : it->second;
    }
};

// main.cpp:1349-1359 (parsed with the intended types; the reference's format
// string has nine conversions for seven arguments).
Params read_params(const char* fn) {
    Params p;
    FILE* fp = std::fopen(fn, "r");
    if (!fp) die("cannot open parameter file ", fn);
    char a[1024], b[1024], c[1024], d[1024], e[1024];
    if (std::fscanf(fp, "%f %f %d %d %d %d %d", &p.dx, &p.dz, &p.pml, &p.allnx, &p.allnz, &p.minshot, &p.maxshot) !=
        7)
        die("bad first line in ", fn);
    if (std::fscanf(fp, "%1023s %1023s %1023s %1023s %1023s", a, b, c, d, e) != 5) die("bad file names in ", fn);
    p.fvp = a; p.fvs = b; p.fn1 = c; p.fn2 = d; p.imagedir = e;
    if (std::fscanf(fp, "%f %d", &p.f0, &p.maxiter) != 2) die("bad f0/maxiter line in ", fn);
    if (std::fscanf(fp, "%d %d", &p.sy, &p.gy) != 2) die("bad sy/gy line in ", fn);
    if (std::fscanf(fp, "%d %d %d %d", &p.light, &p.rbc, &p.cpu_mem, &p.iointerval) != 4)
        die("bad light/rbc/cpu_mem/iointerval line in ", fn);
    char tok[2048];
    while (std::fscanf(fp, "%2047s", tok) == 1) {
        char* eq = std::strchr(tok, '=');
        if (!eq) continue;
        *eq = 0;
        p.kv[tok] = eq + 1;
    }
    std::fclose(fp);
    return p;
}

std::vector<float> read_bin(const std::string& fn, size_t n) {
    std::vector<float> v(n);
    FILE* fp = std::fopen(fn.c_str(), "rb");
    if (!fp) die("cannot open model file ", fn.c_str());
    if (std::fread(v.data(), sizeof(float), n, fp) != n) die("short model file ", fn.c_str());
    std::fclose(fp);
    return v;
}

void write_bin(const std::string& fn, const float* v, size_t n) {
    FILE* fp = std::fopen(fn.c_str(), "wb");
    if (!fp) die("cannot write ", fn.c_str());
    std::fwrite(v, sizeof(float), n, fp);
    std::fclose(fp);
}

// ---- SU (SEG-Y trace header, subfun/segy.h:8-309): 240 bytes, native endian
constexpr int SU_HDR = 240;
constexpr int OFF_TRACL = 0, OFF_FLDR = 8, OFF_TRACF = 12, OFF_EP = 16, OFF_CDP = 20, OFF_TRID = 28, OFF_OFFSET = 36,
              OFF_SCALEL = 68, OFF_SCALCO = 70, OFF_SX = 72, OFF_SY = 76, OFF_GX = 80, OFF_GY = 84, OFF_NS = 114,
              OFF_DT = 116;

template <typename T>
T hget(const unsigned char* h, int off) {
    T v;
    std::memcpy(&v, h + off, sizeof v);
    return v;
}
template <typename T>
void hset(unsigned char* h, int off, T v) {
    std::memcpy(h + off, &v, sizeof v);
}

struct ShotEntry {
    int sno, ntr, sx, cx_min, cx_max;
    long long pos;  // first trace index
};

// main.cpp:52-155 index_shot: scan headers, group consecutive traces by sx.
std::vector<ShotEntry> index_shot(const std::string& fn, int* nt, float* dt) {
    FILE* fp = std::fopen(fn.c_str(), "rb");
    if (!fp) die("cannot open seismic file ", fn.c_str());
    unsigned char h[SU_HDR];
    if (std::fread(h, 1, SU_HDR, fp) != SU_HDR) die("empty seismic file ", fn.c_str());
    *nt = hget<unsigned short>(h, OFF_NS);
    *dt = hget<unsigned short>(h, OFF_DT) / 1000000.0f;
    std::fseek(fp, 0, SEEK_SET);
    std::vector<ShotEntry> t;
    long long pos = 0;
    const long tl = (long)*nt * 4;
    int sx0 = INT32_MIN;
    while (std::fread(h, 1, SU_HDR, fp) == SU_HDR) {
        const int sx = hget<int>(h, OFF_SX), gx = hget<int>(h, OFF_GX);
        if (t.empty() || sx != sx0) {
            t.push_back(ShotEntry{(int)t.size() + 1, 0, sx, INT32_MAX, INT32_MIN, pos});
            sx0 = sx;
        }
        ShotEntry& e = t.back();
        e.ntr++;
        e.cx_min = std::min(e.cx_min, std::min(sx, gx));
        e.cx_max = std::max(e.cx_max, std::max(sx, gx));
        ++pos;
        if (std::fseek(fp, tl, SEEK_CUR) != 0) break;
    }
    std::fclose(fp);
    return t;
}

// main.cpp:254-276 read_shot_gather_su: ntr traces from trace index pos.
void read_gather(const std::string& fn, long long pos, int ntr, int nt, float* dat, int* gc) {
    FILE* fp = std::fopen(fn.c_str(), "rb");
    if (!fp) die("cannot open seismic file ", fn.c_str());
    const long long tl = SU_HDR + (long long)nt * 4;
    std::fseek(fp, (long)(tl * pos), SEEK_SET);
    unsigned char h[SU_HDR];
    for (int i = 0; i < ntr; ++i) {
        if (std::fread(h, 1, SU_HDR, fp) != SU_HDR) die("short seismic file ", fn.c_str());
        gc[i] = hget<int>(h, OFF_GX);
        if (std::fread(dat + (size_t)i * nt, 4, nt, fp) != (size_t)nt) die("short seismic file ", fn.c_str());
    }
    std::fclose(fp);
}

// One trace of shot j (0-based), receiver k of the shot, with the header
// fields write_shot_gather_su_3d sets (read_write_su.c:43-62, ny = 1 so the
// in-line index i is 0): tracl = i, fldr = ep = j + 1, tracf = cdp = k,
// offset = gx - sx, trid = scalel = scalco = 1, sy = gy = 0, and the sample
// interval truncated to microseconds in float arithmetic (hdt = dt*1000000).
void write_su(FILE* fp, int j, int k, int sx, int gx, int nt, float dt, const float* tr) {
    unsigned char h[SU_HDR];
    std::memset(h, 0, sizeof h);
    hset<int>(h, OFF_TRACL, 0);
    hset<int>(h, OFF_FLDR, j + 1);
    hset<int>(h, OFF_TRACF, k);
    hset<int>(h, OFF_EP, j + 1);
    hset<int>(h, OFF_CDP, k);
    hset<short>(h, OFF_TRID, 1);
    hset<int>(h, OFF_OFFSET, gx - sx);
    hset<short>(h, OFF_SCALEL, 1);
    hset<short>(h, OFF_SCALCO, 1);
    hset<int>(h, OFF_SX, sx);
    hset<int>(h, OFF_SY, 0);
    hset<int>(h, OFF_GX, gx);
    hset<int>(h, OFF_GY, 0);
    hset<unsigned short>(h, OFF_NS, (unsigned short)nt);
    hset<unsigned short>(h, OFF_DT, (unsigned short)(short)(dt * 1000000));
    std::fwrite(h, 1, SU_HDR, fp);
    std::fwrite(tr, 4, nt, fp);
}

ds2d_geom make_geom(const Params& p, int nt, float dt, int nx_max, int ntr_max, int physics) {
    ds2d_geom g{};
    g.dx = p.dx; g.dz = p.dz; g.dt = dt;
    g.nxpml = nx_max + 2 * p.pml; g.nzpml = p.allnz + 2 * p.pml;
    g.allnx = p.allnx; g.allnz = p.allnz;
    g.scale = 1; g.pml = p.pml; g.nt = nt; g.nop = 4; g.ntr = ntr_max;
    g.physics = physics;
    g.pml_rr = 1e-3f;
    g.rho_const = p.kv.count("rho") ? 0.f : (physics ? 1000.f : 2000.f);
    g.rb_amp = 0.3f;
    g.rb_seed = std::strtoull(p.get("seed", "1234").c_str(), nullptr, 10);
    return g;
}

// ---------------------------------------------------------------------------
// synth: write vz/vx SU files (elastic forward, or acoustic Born pressure in
// plain mode) for shots every synth_dshot cells, receivers on every node.
int run_synth(const Params& p) {
    const int nt = std::atoi(p.get("synth_nt", "2000").c_str());
    const float dt = std::atof(p.get("synth_dt", "0.0008").c_str());
    const int dshot = std::atoi(p.get("synth_dshot", "10").c_str());
    const int x0 = std::atoi(p.get("synth_x0", "0").c_str());
    const bool plain = p.get("mode", "ref") == "plain";
    const size_t A = (size_t)p.allnx * p.allnz;
    std::vector<float> vp = read_bin(p.get("synth_vp", p.fvp), A);
    std::vector<float> vs = plain ? std::vector<float>() : read_bin(p.get("synth_vs", p.fvs), A);
    std::vector<float> rho = p.kv.count("rho") ? read_bin(p.get("rho", ""), A) : std::vector<float>();
    std::vector<float> wav(nt);
    check(ds2d_ricker(nt, p.f0, dt, wav.data()), "ricker");
    std::vector<int> rix(p.allnx);
    for (int i = 0; i < p.allnx; ++i) rix[i] = p.pml + i;
    std::vector<ds2d_shot> shots;
    for (int sxc = x0; sxc < p.allnx; sxc += dshot) {
        ds2d_shot s{};
        s.ngx_min = 0; s.nx = p.allnx; s.isx = p.pml + sxc; s.isz = p.pml + p.sy; s.igz = p.pml + p.gy;
        s.ntr = p.allnx; s.rec_ix = rix.data(); s.shot_id = (int)shots.size(); s.boundary = plain ? 1 : 0;
        shots.push_back(s);
    }
    ds2d_geom g = make_geom(p, nt, dt, p.allnx, p.allnx, plain ? 1 : 0);
    ds2d_t* h = nullptr;
    check(ds2d_create(&g, wav.data(), &h), "create");
    check(ds2d_set_model(h, vp.data(), vs.empty() ? nullptr : vs.data(), rho.empty() ? nullptr : rho.data()),
          "set_model");
    const size_t R = (size_t)shots.size() * p.allnx * nt;
    std::vector<float> rz(R), rx(R, 0.f);
    if (plain) {
        std::vector<float> refl = read_bin(p.get("synth_refl", ""), A);
        check(ds2d_born_forward(h, (int)shots.size(), shots.data(), refl.data(), nullptr, rz.data(), nullptr),
              "born_forward");
    } else {
        check(ds2d_batch_forward(h, (int)shots.size(), shots.data(), rz.data(), rx.data(), 2, nullptr),
              "batch_forward");
    }
    FILE* fz = std::fopen(p.fn1.c_str(), "wb");
    FILE* fx = std::fopen(p.fn2.c_str(), "wb");
    if (!fz || !fx) die("cannot write SU files");
    for (size_t s = 0; s < shots.size(); ++s) {
        const int sx = (int)std::lround((shots[s].isx - p.pml) * p.dx);
        for (int r = 0; r < p.allnx; ++r) {
            const size_t o = (s * p.allnx + r) * (size_t)nt;
            const int gx = (int)std::lround(r * p.dx);
            write_su(fz, (int)s, r, sx, gx, nt, dt, &rz[o]);
            write_su(fx, (int)s, r, sx, gx, nt, dt, &rx[o]);
        }
    }
    std::fclose(fz);
    std::fclose(fx);
    std::printf("synth: %zu shots x %d traces x %d samples -> %s, %s\n", shots.size(), p.allnx, nt, p.fn1.c_str(),
                p.fn2.c_str());
    ds2d_destroy(h);
    return 0;
}

// ---------------------------------------------------------------------------
// LSRTM
struct Loaded {
    std::vector<ds2d_shot> shots;
    std::vector<std::vector<int>> rix;
    std::vector<float> oz, ox;
};

Loaded load_shots(const Params& p, const std::vector<ShotEntry>& table, int nt, int all_left, bool plain) {
    Loaded L;
    for (const ShotEntry& e : table) {
        if (e.sno < p.minshot || e.sno > p.maxshot) continue;  // main.cpp:967-970
        std::vector<float> z((size_t)e.ntr * nt), x((size_t)e.ntr * nt);
        std::vector<int> gc(e.ntr);
        read_gather(p.fn1, e.pos, e.ntr, nt, z.data(), gc.data());
        if (!plain) read_gather(p.fn2, e.pos, e.ntr, nt, x.data(), gc.data());
        // main.cpp:998-1025: window and source/receiver cells
        const int ngx_left = e.cx_min;
        const int nx = (int)((e.cx_max - e.cx_min) / p.dx) + 1;
        const int delta_left = (int)((ngx_left - all_left) / p.dx);
        std::vector<int> r(e.ntr);
        for (int i = 0; i < e.ntr; ++i) r[i] = p.pml + (int)((float)(gc[i] - ngx_left) / p.dx);
        ds2d_shot s{};
        s.ngx_min = delta_left;
        s.nx = nx;
        s.isx = p.pml + (int)((e.sx - ngx_left) / p.dx);
        s.isz = p.pml + p.sy;
        s.igz = p.pml + p.gy;
        s.ntr = e.ntr;
        s.shot_id = e.sno;
        s.boundary = (plain || p.rbc) ? 1 : 0;
        L.rix.push_back(std::move(r));
        L.shots.push_back(s);
        L.oz.insert(L.oz.end(), z.begin(), z.end());
        L.ox.insert(L.ox.end(), x.begin(), x.end());
    }
    for (size_t i = 0; i < L.shots.size(); ++i) L.rix[i].shrink_to_fit(), L.shots[i].rec_ix = L.rix[i].data();
    return L;
}

// Adam update of lsrtm_all_iteration (main.cpp:1092-1130) on g + b_g + b_g2.
void adam(int n, int iter, const float* g, const float* bg, const float* bg2, float* s0, float* r0, float* img,
          float* sum) {
    const float beta1 = 0.8f, beta2 = 0.999f, alpha = 3e-5f, eps = 5e8f, gama = 0.99f;
    const float alpha_k = iter == 0 ? alpha : (float)((alpha / 2) * std::pow((double)gama, (double)iter));
    const int k = iter + 1;
    for (int i = 0; i < n; ++i) {
        sum[i] = g[i] + bg[i] + bg2[i];
        const float sk = beta1 * s0[i] + (1 - beta1) * sum[i];
        const float rk = beta2 * r0[i] + (1 - beta2) * (sum[i] * sum[i]);
        const float sm = sk / (float)(1 - std::pow((double)beta1, (double)k));
        const float rm = rk / (float)(1 - std::pow((double)beta2, (double)k));
        img[i] -= alpha_k * (sm / (std::sqrt(rm) + eps));
        s0[i] = sk;
        r0[i] = rk;
    }
}

int run_lsrtm(const Params& p) {
    const bool plain = p.get("mode", "ref") == "plain";
    int nt = 0;
    float dt = 0;
    std::vector<ShotEntry> table = index_shot(p.fn1, &nt, &dt);
    if (table.empty()) die("no shots in ", p.fn1.c_str());
    int all_left = INT32_MAX;
    for (auto& e : table) all_left = std::min(all_left, e.cx_min);
    std::printf("shots %zu, nt %d, dt %g s\n", table.size(), nt, dt);
    Loaded L = load_shots(p, table, nt, all_left, plain);
    if (L.shots.empty()) die("no shot in [minshot, maxshot]");
    int nx_max = 0, ntr_max = 0;
    for (auto& s : L.shots) nx_max = std::max(nx_max, s.nx), ntr_max = std::max(ntr_max, s.ntr);
    const size_t A = (size_t)p.allnx * p.allnz;
    std::vector<float> vp = read_bin(p.fvp, A);
    std::vector<float> vs = plain ? std::vector<float>() : read_bin(p.fvs, A);
    std::vector<float> rho = p.kv.count("rho") ? read_bin(p.get("rho", ""), A) : std::vector<float>();
    std::vector<float> wav(nt);
    check(ds2d_ricker(nt, p.f0, dt, wav.data()), "ricker");  // main.cpp:1476
    ds2d_geom g = make_geom(p, nt, dt, nx_max, ntr_max, plain ? 1 : 0);
    ds2d_t* h = nullptr;
    check(ds2d_create(&g, wav.data(), &h), "create");
    check(ds2d_set_model(h, vp.data(), vs.empty() ? nullptr : vs.data(), rho.empty() ? nullptr : rho.data()),
          "set_model");
    mkdir(p.imagedir.c_str(), 0755);
    char path[2048];
    if (plain) {
        std::vector<float> img(A, 0.f);
        std::vector<double> res(p.maxiter + 1);
        ds2d_stats st{};
        check(ds2d_lsrtm_cg(h, (int)L.shots.size(), L.shots.data(), L.oz.data(), p.maxiter, img.data(), res.data(),
                            &st),
              "lsrtm_cg");
        std::snprintf(path, sizeof path, "%s/cg_image_%d_%d_%d.dat", p.imagedir.c_str(), p.maxiter, p.allnx, p.allnz);
        write_bin(path, img.data(), A);
        check(ds2d_laplace_filter(p.allnz, p.allnx, 2, p.dx, img.data()), "laplace");
        std::snprintf(path, sizeof path, "%s/cg_image_laplace_%d_%d_%d.dat", p.imagedir.c_str(), p.maxiter, p.allnx,
                      p.allnz);
        write_bin(path, img.data(), A);
        std::snprintf(path, sizeof path, "%s/misfit.txt", p.imagedir.c_str());
        FILE* fp = std::fopen(path, "w");
        for (int i = 0; i <= p.maxiter; ++i) std::fprintf(fp, "i = %d, relative residual = %e\n", i, res[i]);
        std::fclose(fp);
        std::printf("plain CG: residual %g -> %g in %.1f ms\n", res[0], res[p.maxiter], st.elapsed_ms);
        ds2d_destroy(h);
        return 0;
    }
    const int batch = std::max(1, std::atoi(p.get("batch", std::to_string(L.shots.size())).c_str()));
    std::vector<float> ipp(A, 0.f), ips(A, 0.f), s0p(A, 0.f), s0s(A, 0.f), r0p(A, 0.f), r0s(A, 0.f);
    std::vector<float> gpp(A), gps(A), il(A), bpp(A), bps(A), b2pp(A), b2ps(A), sump(A), sums(A);
    std::vector<float> mis(p.maxiter);
    // opt = adam (main.cpp:1092-1130, the reference's active update) or
    // cg (sjcgsolver, main.cpp:1278, present in the reference but not called)
    const bool use_cg = p.get("opt", "adam") == "cg";
    if (!use_cg && p.get("opt", "adam") != "adam") die("opt must be adam or cg");
    std::vector<float> cgp(use_cg ? A : 0), cgs(use_cg ? A : 0), g0p(use_cg ? A : 0), g0s(use_cg ? A : 0);
    const size_t R = (size_t)nt;
    for (int iter = 0; iter < p.maxiter; ++iter) {
        check(ds2d_set_image(h, ipp.data(), ips.data()), "set_image");
        double misfit = 0;
        size_t roff = 0;
        for (size_t s0 = 0; s0 < L.shots.size(); s0 += batch) {
            const int nb = (int)std::min<size_t>(batch, L.shots.size() - s0);
            const int grc = ds2d_batch_gradient(h, nb, &L.shots[s0], &L.oz[roff], &L.ox[roff], iter, &misfit);
            if (grc == DS2D_ESHOT) {  // the batch completed without the failed shots: report them, go on
                std::vector<int> st(nb);
                check(ds2d_shot_status(h, nb, st.data()), "shot status");
                for (int k = 0; k < nb; ++k)
                    if (st[k])
                        std::fprintf(stderr, "ds2d_lsrtm: iteration %d: shot %d left out (non-finite residual or "
                                     "gradient)\n", iter, L.shots[s0 + k].shot_id);
            } else {
                check(grc, "gradient");
            }
            for (int k = 0; k < nb; ++k) roff += (size_t)L.shots[s0 + k].ntr * R;
        }
        check(ds2d_take_gradients(h, gpp.data(), gps.data(), il.data(), bpp.data(), bps.data(), b2pp.data(),
                                  b2ps.data()),
              "take_gradients");
        if (iter == 0) misfit = 0;  // main.cpp:1065-1067
        mis[iter] = 1.0f - (float)std::fabs((1.0 / table.size()) * misfit);
        std::printf("iter = %d  misfit = %g  rmsmisfit = %g\n", iter, misfit, mis[iter]);
        if (use_cg) {  // the reference's nonlinear CG (sjcgsolver) on g + b_g + b_g2
            for (size_t i = 0; i < A; ++i) {
                sump[i] = gpp[i] + bpp[i] + b2pp[i];
                sums[i] = gps[i] + bps[i] + b2ps[i];
            }
            check(ds2d_sjcg_update((int)A, ipp.data(), cgp.data(), sump.data(), g0p.data(), iter, nullptr, nullptr),
                  "cg pp");
            check(ds2d_sjcg_update((int)A, ips.data(), cgs.data(), sums.data(), g0s.data(), iter, nullptr, nullptr),
                  "cg ps");
        } else {
            adam((int)A, iter, gpp.data(), bpp.data(), b2pp.data(), s0p.data(), r0p.data(), ipp.data(), sump.data());
            adam((int)A, iter, gps.data(), bps.data(), b2ps.data(), s0s.data(), r0s.data(), ips.data(), sums.data());
        }
        const struct { const char* tag; const float* v; } outs[] = {
            {"sum_bgrad2_pp", b2pp.data()}, {"sum_bgrad2_ps", b2ps.data()}, {"all_grad_pp", sump.data()},
            {"all_grad_ps", sums.data()},   {"sum_image_pp", ipp.data()},   {"sum_image_ps", ips.data()}};
        for (auto& o : outs) {  // main.cpp:1134-1165 file names
            std::snprintf(path, sizeof path, "%s/%s_%d_%d_%d.dat", p.imagedir.c_str(), o.tag, iter, p.allnx, p.allnz);
            write_bin(path, o.v, A);
        }
    }
    std::snprintf(path, sizeof path, "%s/misfit.txt", p.imagedir.c_str());  // main.cpp:1178-1186
    FILE* fp = std::fopen(path, "w");
    for (int i = 0; i < p.maxiter; ++i) std::fprintf(fp, "i = %d,misfit = %e,rms_misfit = %f\n", i, mis[i], mis[i]);
    std::fclose(fp);
    ds2d_destroy(h);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 3 && std::strcmp(argv[1], "synth") == 0) return run_synth(read_params(argv[2]));
    if (argc == 3 && std::strcmp(argv[1], "index") == 0) {
        // shot table of an SU file (no device work): sno ntr sx cx_min cx_max pos
        int nt = 0;
        float dt = 0;
        std::vector<ShotEntry> t = index_shot(argv[2], &nt, &dt);
        std::printf("nt %d dt %.6f shots %zu\n", nt, dt, t.size());
        for (auto& e : t) std::printf("%d %d %d %d %d %lld\n", e.sno, e.ntr, e.sx, e.cx_min, e.cx_max, e.pos);
        return 0;
    }
    if (argc != 2) {
        std::fprintf(stderr, "usage: ds2d_lsrtm [synth|index] <parameter file | su file>\n");
        return 2;
    }
    return run_lsrtm(read_params(argv[1]));
}
