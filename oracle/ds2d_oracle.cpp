// DS2D CPU oracle — TEST INFRASTRUCTURE ONLY (see ds2d_oracle.hpp header).
// extern "C" surface used by tests/ (ctypes) and bench.py's cpu_baseline leg.
#include "ds2d_oracle.hpp"
#include "ds2d_oracle_rot.hpp"

#include <cstdio>
#include <omp.h>

using namespace orc;

extern "C" {

// Shot description shared by the oracle drivers (all indices are padded-window
// coordinates unless stated otherwise).
typedef struct {
    int allnx, allnz;    // global model size (x-major, z contiguous)
    int ngx_min, nx;     // window start column in the global model, window width
    int pml, nt;         // frame width, time steps
    double dx, dz, dt;
    int boundary;        // 0 = PML damping, 1 = random boundary (no damping)
    double rr;           // PML reflection coefficient (elastic 1e-3, acoustic 1e-6)
    unsigned long long rb_seed;
    int rb_shot;
    double rb_amp;       // random boundary relative amplitude (0.3)
    double rho_const;    // > 0: constant density (reference compat), else rho array
    int isx, isz;        // source cell
    int igz, ntr;        // receiver row, trace count
    const int* rec_ix;   // receiver columns [ntr]
} orc_shot;

}  // extern "C"

namespace {

template <typename R>
struct Setup {
    Grid<R> g;
    Medium<R> m;
    int nx, nz;
};

// Window, pad and build the elastic medium exactly as FD2DGPU_ELASTIC does per
// shot: bufferVpVsHtoD (FD2DGPU.cu:1336-1348), then either set_random_boundary
// (replaced by the seeded frame, FD2DGPU.cu:3156-3216) or
// calculate_elastic_damp_C (FD2DGPU.cu:1292-1309).
template <typename R>
Setup<R> make_setup(const orc_shot* s, const R* vp, const R* vs, const R* rho) {
    Setup<R> S;
    S.nx = s->nx;
    S.nz = s->allnz;
    S.g = Grid<R>{s->nx + 2 * s->pml, s->allnz + 2 * s->pml, s->pml, (R)s->dx, (R)s->dz, (R)s->dt};
    const size_t n = S.g.n();
    Medium<R>& m = S.m;
    m.vp.assign(n, 0); m.vs.assign(n, 0); m.rho.assign(n, 0); m.lam.assign(n, 0); m.mu.assign(n, 0);
    m.dampx.assign(n, 0); m.dampz.assign(n, 0);
    pad_window(vp, s->allnz, s->ngx_min, s->nx, S.nz, s->pml, m.vp.data());
    if (vs) pad_window(vs, s->allnz, s->ngx_min, s->nx, S.nz, s->pml, m.vs.data());
    std::vector<R> rhop(n, 0);
    if (rho) pad_window(rho, s->allnz, s->ngx_min, s->nx, S.nz, s->pml, rhop.data());
    if (s->boundary == 1) {
        random_boundary(S.g, (uint64_t)s->rb_seed, s->rb_shot, (R)s->rb_amp, m.vp.data(), m.vs.data());
    } else {
        damp_profile(S.g, m.vp.data(), (R)s->rr, m.dampx.data(), m.dampz.data());
    }
    elastic_moduli(n, m.vp.data(), m.vs.data(), rhop.data(), (R)s->rho_const, m.rho.data(), m.lam.data(),
                   m.mu.data());
    return S;
}

template <typename R>
void record_pair(const Grid<R>& g, const orc_shot* s, const std::vector<R>& vz, const std::vector<R>& vx,
                 int it, R* rz, R* rxx) {
    for (int r = 0; r < s->ntr; ++r) {
        size_t c = g.at(s->rec_ix[r], s->igz);
        if (rz) rz[(size_t)r * s->nt + it] = vz[c];
        if (rxx) rxx[(size_t)r * s->nt + it] = vx[c];
    }
}

// Elastic forward modelling: FD2DGPU.cu:2046 source before the step,
// born_modeling(1,0) (FD2DGPU.cu:1834-1839), then the surface recorder
// GPUseisrecord (FD2DGPU.cu:355-365) generalised to a receiver list.
template <typename R>
void elastic_forward(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* wav, R* rec_z,
                     R* rec_x, R* final5) {
    Setup<R> S = make_setup<R>(s, vp, vs, rho);
    Elastic<R> f;
    f.init(S.g.n());
    const size_t src = S.g.at(s->isx, s->isz);
    for (int it = 0; it < s->nt; ++it) {
        f.txx[src] += wav[it];
        f.tzz[src] += wav[it];
        update_v(S.g, f, S.m, 1);
        update_s(S.g, f, S.m, 1);
        record_pair(S.g, s, f.vz, f.vx, it, rec_z, rec_x);
    }
    if (final5) {
        const size_t n = S.g.n();
        std::copy(f.vx.begin(), f.vx.end(), final5);
        std::copy(f.vz.begin(), f.vz.end(), final5 + n);
        std::copy(f.txx.begin(), f.txx.end(), final5 + 2 * n);
        std::copy(f.tzz.begin(), f.tzz.end(), final5 + 3 * n);
        std::copy(f.txz.begin(), f.txz.end(), final5 + 4 * n);
    }
}

// FD2DGPU.cu:4532-4673 normalised cross-correlation residual (host, double
// sums).  iter==0: r = -D/|D|^2; else r = ((dD/|d|^2) d - D)/(|d||D|) and
// misfit -= (1/(2 ntr)) sum_tr (dD_x/(|d_x||D_x|) + dD_z/(|d_z||D_z|)).
// A zero-norm trace yields a zero residual here (the reference divides by 0).
template <typename R>
void ncc_residual(int ntr, int nt, int iter, R* syn_z, R* syn_x, const R* obs_z, const R* obs_x,
                  double* misfit) {
    std::vector<double> dx(ntr, 0), Dx(ntr, 0), dDx(ntr, 0), dz(ntr, 0), Dz(ntr, 0), dDz(ntr, 0);
    for (int i = 0; i < ntr; ++i)
        for (int j = 0; j < nt; ++j) {
            size_t k = (size_t)i * nt + j;
            if (iter != 0) {
                dx[i] += (double)syn_x[k] * syn_x[k];
                dz[i] += (double)syn_z[k] * syn_z[k];
                dDx[i] += (double)syn_x[k] * obs_x[k];
                dDz[i] += (double)syn_z[k] * obs_z[k];
            }
            Dx[i] += (double)obs_x[k] * obs_x[k];
            Dz[i] += (double)obs_z[k] * obs_z[k];
        }
    if (iter != 0) {
        for (int i = 0; i < ntr; ++i) {
            double tx = (dx[i] > 0 && Dx[i] > 0) ? dDx[i] / (std::sqrt(dx[i]) * std::sqrt(Dx[i])) : 0.0;
            double tz = (dz[i] > 0 && Dz[i] > 0) ? dDz[i] / (std::sqrt(dz[i]) * std::sqrt(Dz[i])) : 0.0;
            *misfit -= (double)(1.0f / (2 * ntr)) * (tx + tz);
        }
    }
    for (int i = 0; i < ntr; ++i)
        for (int j = 0; j < nt; ++j) {
            size_t k = (size_t)i * nt + j;
            if (iter == 0) {
                double ax = Dx[i] > 0 ? 1.0 / (std::sqrt(Dx[i]) * std::sqrt(Dx[i])) : 0.0;
                double az = Dz[i] > 0 ? 1.0 / (std::sqrt(Dz[i]) * std::sqrt(Dz[i])) : 0.0;
                syn_x[k] = (R)(ax * (0.0 - (double)obs_x[k]));
                syn_z[k] = (R)(az * (0.0 - (double)obs_z[k]));
            } else {
                bool okx = dx[i] > 0 && Dx[i] > 0, okz = dz[i] > 0 && Dz[i] > 0;
                double ax = okx ? 1.0 / (std::sqrt(dx[i]) * std::sqrt(Dx[i])) : 0.0;
                double az = okz ? 1.0 / (std::sqrt(dz[i]) * std::sqrt(Dz[i])) : 0.0;
                double gx = okx ? dDx[i] / dx[i] : 0.0, gz = okz ? dDz[i] / dz[i] : 0.0;
                syn_x[k] = (R)(ax * (gx * (double)syn_x[k] - (double)obs_x[k]));
                syn_z[k] = (R)(az * (gz * (double)syn_z[k] - (double)obs_z[k]));
            }
        }
}

// ELSRTM_ROT_EXP_IMAGE_SINGLESHOT (FD2DGPU.cu:4184-4879), one shot, one
// iteration: forward bg/b/b2 with scatter records (and PML edge bands), NCC
// residual, then the reverse loop with source reconstruction, receiver-side
// adjoint, receiver-side Born adjoint and the three-term imaging condition.
// grads7 = [pp, ps, illum, b_pp, b_ps, b2_pp, b2_ps], each allnx*allnz, +=.
template <typename R>
void elsrtm_shot(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* wav, const R* img_pp,
                 const R* img_ps, const R* obs_z, const R* obs_x, int iter, int col0, R* grads7,
                 double* misfit) {
    Setup<R> S = make_setup<R>(s, vp, vs, rho);
    const Grid<R>& g = S.g;
    const size_t n = g.n();
    const int nt = s->nt, ntr = s->ntr, nz = S.nz;
    const bool rbc = s->boundary == 1;
    // buffer_image_extrapolation (FD2DGPU.cu:1324-1333): padded image window.
    std::vector<R> dmp(n, 0), dms(n, 0);
    pad_window(img_pp, s->allnz, col0, s->nx, nz, s->pml, dmp.data());
    pad_window(img_ps, s->allnz, col0, s->nx, nz, s->pml, dms.data());

    Elastic<R> bg, b, b2, bk, bbk;
    bg.init(n); b.init(n); b2.init(n); bk.init(n); bbk.init(n);
    Strain<R> Sg, Sb;
    Sg.init(n); Sb.init(n);
    std::vector<R> rs_z((size_t)ntr * nt, 0), rs_x(rs_z.size(), 0), rs2_z(rs_z.size(), 0), rs2_x(rs_z.size(), 0);
    const int N = NOP;
    EdgeBands<R> e_vx, e_vz, e_bvx, e_bvz;
    if (!rbc) {
        e_vx.init(nt, S.nx, nz, N); e_vz.init(nt, S.nx, nz, N);
        e_bvx.init(nt, S.nx, nz, N); e_bvz.init(nt, S.nx, nz, N);
    }
    const size_t src = g.at(s->isx, s->isz);
    for (int it = 0; it < nt; ++it) {
        bg.txx[src] += wav[it];
        bg.tzz[src] += wav[it];
        update_v(g, bg, S.m, 1);
        update_s(g, bg, S.m, 1);
        add_born_source(g, bg.vx, bg.vz, b, dmp, dms, S.m, 1);
        update_v(g, b, S.m, 1);
        update_s(g, b, S.m, 1);
        add_born_source(g, b.vx, b.vz, b2, dmp, dms, S.m, 1);
        update_v(g, b2, S.m, 1);
        update_s(g, b2, S.m, 1);
        record_pair(g, s, b.vz, b.vx, it, rs_z.data(), rs_x.data());
        record_pair(g, s, b2.vz, b2.vx, it, rs2_z.data(), rs2_x.data());
        if (!rbc) {
            edge_io(g, bg.vx, e_vx, it, true); edge_io(g, bg.vz, e_vz, it, true);
            edge_io(g, b.vx, e_bvx, it, true); edge_io(g, b.vz, e_bvz, it, true);
        }
    }
    // scatter_add (FD2DGPU.cu:3143-3149): s = 1*s + s2.
    for (size_t k = 0; k < rs_z.size(); ++k) {
        rs_z[k] = (R)1 * rs_z[k] + rs2_z[k];
        rs_x[k] = (R)1 * rs_x[k] + rs2_x[k];
    }
    ncc_residual<R>(ntr, nt, iter, rs_z.data(), rs_x.data(), obs_z, obs_x, misfit);

    const size_t A = (size_t)s->allnx * s->allnz;
    R *gpp = grads7, *gps = grads7 + A, *gil = grads7 + 2 * A, *bpp = grads7 + 3 * A, *bps = grads7 + 4 * A,
      *b2pp = grads7 + 5 * A, *b2ps = grads7 + 6 * A;
    for (int it = nt - 1; it >= 0; --it) {
        if (!rbc) {
            edge_io(g, bg.vx, e_vx, it, false); edge_io(g, bg.vz, e_vz, it, false);
            edge_io(g, b.vx, e_bvx, it, false); edge_io(g, b.vz, e_bvz, it, false);
        }
        recon_s(g, bg, S.m, -1, Sg);
        recon_v(g, bg, S.m, -1);
        recon_s(g, b, S.m, -1, Sb);
        add_born_source(g, bg.vx, bg.vz, b, dmp, dms, S.m, 1);
        recon_v(g, b, S.m, -1);
        for (int r = 0; r < ntr; ++r) {
            size_t c = g.at(s->rec_ix[r], s->igz);
            bk.vz[c] += rs_z[(size_t)r * nt + it];
            bk.vx[c] += rs_x[(size_t)r * nt + it];
        }
        adjoint_s(g, bk, S.m, -1);
        adjoint_v(g, bk, S.m, -1);
        adjoint_v(g, bbk, S.m, -1);
        add_born_source_adjoint(g, bk, bbk.vx, bbk.vz, dmp, dms, S.m, -1);
        adjoint_s(g, bbk, S.m, -1);
        image_gradient(g, Sg, bk, S.m, col0, s->allnz, gpp, gps);
        image_gradient(g, Sb, bk, S.m, col0, s->allnz, bpp, bps);
        image_gradient(g, Sg, bbk, S.m, col0, s->allnz, b2pp, b2ps);
        image_illum(g, bg, col0, s->allnz, gil);
    }
}

// Born modelling alone (ELSRTM_CALCULATE_ALPHA's forward loop, FD2DGPU.cu:
// 4915-4921: born_modeling mode 1 + record_scatter): first-order scatter
// records Lm for order 1; for order 2 the schedule of ELSRTM_ROT_EXP_IMAGE_
// SINGLESHOT's forward loop (modes 0/3/4/5) and scatter_add, Lm + Lm2.
template <typename R>
void elsrtm_born(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* wav, const R* img_pp,
                 const R* img_ps, int col0, int order, R* rz, R* rxx) {
    Setup<R> S = make_setup<R>(s, vp, vs, rho);
    const Grid<R>& g = S.g;
    const size_t n = g.n();
    const int nt = s->nt, ntr = s->ntr;
    std::vector<R> dmp(n, 0), dms(n, 0);
    pad_window(img_pp, s->allnz, col0, s->nx, S.nz, s->pml, dmp.data());
    pad_window(img_ps, s->allnz, col0, s->nx, S.nz, s->pml, dms.data());
    Elastic<R> bg, b, b2;
    bg.init(n); b.init(n); b2.init(n);
    std::vector<R> r2z((size_t)ntr * nt, 0), r2x(r2z.size(), 0);
    const size_t src = g.at(s->isx, s->isz);
    for (int it = 0; it < nt; ++it) {
        bg.txx[src] += wav[it];
        bg.tzz[src] += wav[it];
        update_v(g, bg, S.m, 1);
        update_s(g, bg, S.m, 1);
        add_born_source(g, bg.vx, bg.vz, b, dmp, dms, S.m, 1);
        update_v(g, b, S.m, 1);
        update_s(g, b, S.m, 1);
        if (order == 2) {
            add_born_source(g, b.vx, b.vz, b2, dmp, dms, S.m, 1);
            update_v(g, b2, S.m, 1);
            update_s(g, b2, S.m, 1);
            record_pair(g, s, b2.vz, b2.vx, it, r2z.data(), r2x.data());
        }
        record_pair(g, s, b.vz, b.vx, it, rz, rxx);
    }
    if (order == 2)
        for (size_t k = 0; k < r2z.size(); ++k) {
            rz[k] = (R)1 * rz[k] + r2z[k];
            rxx[k] = (R)1 * rxx[k] + r2x[k];
        }
}

// Single kernels on caller-provided padded arrays (15-array elastic state in
// the order vx, vz, txx, tzz, txz, vx_x, vx_z, vz_x, vz_z, txx_x, txx_z,
// tzz_x, tzz_z, txz_x, txz_z), for the pin tests.
template <typename R>
void load15(Elastic<R>& f, const R* a, size_t n) {
    f.init(n);
    auto all = f.all();
    for (size_t k = 0; k < all.size(); ++k) std::copy(a + k * n, a + (k + 1) * n, all[k]->begin());
}
template <typename R>
void store15(Elastic<R>& f, R* a, size_t n) {
    auto all = f.all();
    for (size_t k = 0; k < all.size(); ++k) std::copy(all[k]->begin(), all[k]->end(), a + k * n);
}
template <typename R>
Medium<R> medium_of(size_t n, const R* vp, const R* vs, const R* rho, const R* lam, const R* mu, const R* dampx,
                    const R* dampz) {
    Medium<R> m;
    auto cp = [&](std::vector<R>& v, const R* a) { if (a) v.assign(a, a + n); else v.assign(n, 0); };
    cp(m.vp, vp); cp(m.vs, vs); cp(m.rho, rho); cp(m.lam, lam); cp(m.mu, mu); cp(m.dampx, dampx); cp(m.dampz, dampz);
    return m;
}
// which: 0 update_v, 1 update_s, 2 recon_v, 3 recon_s (strain4 out), 4 adjoint_v,
// 5 adjoint_s, 6 add_born_source (velocities of src15 into the stress sums of
// f15), 7 add_born_source_adjoint (stresses of src15 into the velocity sums of f15).
template <typename R>
void kernel(int which, int nxp, int nzp, int pml, R dx, R dz, R dt, const R* vp, const R* vs, const R* rho,
            const R* lam, const R* mu, const R* dampx, const R* dampz, const R* dmp, const R* dms, int dir,
            const R* src15, R* f15, R* strain4) {
    Grid<R> g{nxp, nzp, pml, dx, dz, dt};
    const size_t n = g.n();
    Medium<R> m = medium_of<R>(n, vp, vs, rho, lam, mu, dampx, dampz);
    Elastic<R> f;
    load15(f, f15, n);
    std::vector<R> Dp(n, 0), Ds(n, 0);
    if (dmp) Dp.assign(dmp, dmp + n);
    if (dms) Ds.assign(dms, dms + n);
    Strain<R> st;
    st.init(n);
    Elastic<R> sf;
    if (src15) load15(sf, src15, n);
    switch (which) {
        case 0: update_v(g, f, m, dir); break;
        case 1: update_s(g, f, m, dir); break;
        case 2: recon_v(g, f, m, dir); break;
        case 3: recon_s(g, f, m, dir, st); break;
        case 4: adjoint_v(g, f, m, dir); break;
        case 5: adjoint_s(g, f, m, dir); break;
        case 6: add_born_source(g, sf.vx, sf.vz, f, Dp, Ds, m, dir); break;
        case 7: add_born_source_adjoint(g, sf, f.vx, f.vz, Dp, Ds, m, dir); break;
        default: break;
    }
    store15(f, f15, n);
    if (strain4) {
        std::copy(st.gx_vx.begin(), st.gx_vx.end(), strain4);
        std::copy(st.gz_vz.begin(), st.gz_vz.end(), strain4 + n);
        std::copy(st.gz_vx.begin(), st.gz_vx.end(), strain4 + 2 * n);
        std::copy(st.gx_vz.begin(), st.gx_vz.end(), strain4 + 3 * n);
    }
}
// edge_io on one padded field; bands [nt][up | down | left | right] per field.
template <typename R>
void edge_io_x(int nxp, int nzp, int pml, int nt, R* field, R* bands, int it, int save) {
    Grid<R> g{nxp, nzp, pml, 1, 1, 1};
    const int nx = nxp - 2 * pml, nz = nzp - 2 * pml;
    EdgeBands<R> e;
    e.init(nt, nx, nz, NOP);
    const size_t nu = e.up.size(), nl = e.left.size();
    std::copy(bands, bands + nu, e.up.begin());
    std::copy(bands + nu, bands + 2 * nu, e.down.begin());
    std::copy(bands + 2 * nu, bands + 2 * nu + nl, e.left.begin());
    std::copy(bands + 2 * nu + nl, bands + 2 * nu + 2 * nl, e.right.begin());
    std::vector<R> f(field, field + g.n());
    edge_io(g, f, e, it, save != 0);
    std::copy(f.begin(), f.end(), field);
    std::copy(e.up.begin(), e.up.end(), bands);
    std::copy(e.down.begin(), e.down.end(), bands + nu);
    std::copy(e.left.begin(), e.left.end(), bands + 2 * nu);
    std::copy(e.right.begin(), e.right.end(), bands + 2 * nu + nl);
}
// image_gradient (+ image_illum when il) of one window into global images.
template <typename R>
void image_x(int nxp, int nzp, int pml, const R* strain4, const R* bk15, const R* vp, const R* vs, const R* rho,
             int col0, int allnz, R* pp, R* ps, R* il) {
    Grid<R> g{nxp, nzp, pml, 1, 1, 1};
    const size_t n = g.n();
    Strain<R> st;
    st.gx_vx.assign(strain4, strain4 + n);
    st.gz_vz.assign(strain4 + n, strain4 + 2 * n);
    st.gz_vx.assign(strain4 + 2 * n, strain4 + 3 * n);
    st.gx_vz.assign(strain4 + 3 * n, strain4 + 4 * n);
    Elastic<R> bk;
    load15(bk, bk15, n);
    Medium<R> m = medium_of<R>(n, vp, vs, rho, nullptr, nullptr, nullptr, nullptr);
    image_gradient(g, st, bk, m, col0, allnz, pp, ps);
    if (il) image_illum(g, bk, col0, allnz, il);
}

// ---------------------------------------------------------------------------
// Acoustic, reference layout (FD2DGPU.cu:253-340, 496-544): split p, single
// damping array (dampX then dampZ into the same array, rr = 1e-6), rho = 1000,
// source on vz at (sxnum+pml, pml), pressure recorded at depth pml.  The
// velocity stagger is mirrored relative to the elastic kernels and vz uses
// dt/dx (FD2DGPU.cu:318) — kept.
// ---------------------------------------------------------------------------
// FD2DGPU.cu:505-519: dampX then dampZ write the same array, so cells in both
// strips (the corners) keep the z profile.
template <typename R>
void acoustic_ref_damp(const Grid<R>& g, const R* vt, R rr, R* damp) {
    const int pml = g.pml;
    const size_t n = g.n();
    std::vector<R> dxa(n, 0), dza(n, 0);
    damp_profile(g, vt, rr, dxa.data(), dza.data());
    for (int ix = 0; ix < g.nxp; ++ix)
        for (int iz = 0; iz < g.nzp; ++iz) {
            size_t c = g.at(ix, iz);
            damp[c] = 0;
            if (ix < pml || ix > g.nxp - pml) damp[c] = dxa[c];
            if (iz < pml || iz > g.nzp - pml) damp[c] = dza[c];
        }
}

// One step of FD2DGPU.cu:253-325: GPUcalculate_p on [nop-1, n-nop-1]^2, then
// GPUcalculate_vx (ix in [nop, nxp-nop], iz in [nop, nzp-nop-1]) and
// GPUcalculate_vz (ix in [nop, nxp-nop-1], iz in [nop, nzp-nop]); vz uses
// dt/dx (FD2DGPU.cu:318).
template <typename R>
void acoustic_ref_step(const Grid<R>& g, const R* k, const R* rr, const R* damp, R* px, R* pz, R* p, R* vx,
                       R* vz) {
    const R dt = g.dt, dx = g.dx, dz = g.dz;
    const int nxp = g.nxp, nzp = g.nzp;
#pragma omp parallel for
    for (int ix = NOP - 1; ix <= nxp - NOP - 1; ++ix)
        for (int iz = NOP - 1; iz <= nzp - NOP - 1; ++iz) {
            size_t c = g.at(ix, iz);
            R d1 = 1 - dt * damp[c] / 2, d2 = 1 + dt * damp[c] / 2;
            R tvx = 0, tvz = 0;
            for (int i = 1; i <= NOP; ++i) {
                tvx += (R)C8[i] * (vx[c + (size_t)i * nzp] - vx[c - (size_t)(i - 1) * nzp]);
                tvz += (R)C8[i] * (vz[c + i] - vz[c - i + 1]);
            }
            px[c] = (d1 * px[c] - k[c] * (dt / dx) * tvx) / d2;
            pz[c] = (d1 * pz[c] - k[c] * (dt / dz) * tvz) / d2;
            p[c] = px[c] + pz[c];
        }
#pragma omp parallel for
    for (int ix = NOP; ix <= nxp - NOP; ++ix)
        for (int iz = NOP; iz <= nzp - NOP; ++iz) {
            size_t c = g.at(ix, iz);
            R d1 = 1 - dt * damp[c] / 2, d2 = 1 + dt * damp[c] / 2;
            if (iz <= nzp - NOP - 1) {  // GPUcalculate_vx bounds
                R tp = 0;
                for (int i = 1; i <= NOP; ++i)
                    tp += (R)C8[i] * (p[c + (size_t)(i - 1) * nzp] - p[c - (size_t)i * nzp]);
                vx[c] = (d1 * vx[c] - ((R)1 / rr[c]) * (dt / dx) * tp) / d2;
            }
            if (ix <= nxp - NOP - 1) {  // GPUcalculate_vz bounds
                R tp = 0;
                for (int i = 1; i <= NOP; ++i) tp += (R)C8[i] * (p[c + i - 1] - p[c - i]);
                vz[c] = (d1 * vz[c] - ((R)1 / rr[c]) * (dt / dx) * tp) / d2;
            }
        }
}

template <typename R>
void acoustic_ref_forward(const orc_shot* s, const R* vel, const R* rho, const R* wav, int sxnum, R* rec_p) {
    const int pml = s->pml, nx = s->nx, nz = s->allnz;
    Grid<R> g{nx + 2 * pml, nz + 2 * pml, pml, (R)s->dx, (R)s->dz, (R)s->dt};
    const size_t n = g.n();
    std::vector<R> vt(n, 0), rr(n, 0), k(n, 0), damp(n, 0), px(n, 0), pz(n, 0), p(n, 0), vx(n, 0), vz(n, 0);
    pad_window(vel, s->allnz, s->ngx_min, nx, nz, pml, vt.data());
    std::vector<R> rhop(n, 0);
    if (rho) pad_window(rho, s->allnz, s->ngx_min, nx, nz, pml, rhop.data());
    acoustic_ref_damp(g, vt.data(), (R)s->rr, damp.data());
    for (size_t c = 0; c < n; ++c) {  // FD2DGPU.cu:496-502
        rr[c] = s->rho_const > 0 ? (R)s->rho_const : rhop[c];
        k[c] = vt[c] * vt[c] * rr[c];
    }
    for (int it = 0; it < s->nt; ++it) {
        acoustic_ref_step(g, k.data(), rr.data(), damp.data(), px.data(), pz.data(), p.data(), vx.data(), vz.data());
        vz[g.at(sxnum + pml, pml)] += wav[it];
        for (int ix = 0; ix < nx; ++ix) rec_p[(size_t)ix * s->nt + it] = p[g.at(ix + pml, pml)];
    }
}

// ---------------------------------------------------------------------------
// Plain-L2 acoustic Born modelling and its exact adjoint (the build's own
// formulation; no reference counterpart).  It is the mu = 0 specialisation of
// the elastic live stencils (born.cu:215-349: same stagger, same [nop,n-nop)
// bounds, txx = tzz = p, txz = 0), random boundary, no damping.
//   step n:  v  += b dt D+ (p + s_n delta_src)         (source transient)
//            dp  = k dt D- v ;  p += dp
//            dv += b dt D+ dpb ; dpb += k dt D- dv + 2 m dp   (Born, m = dv/v)
//            record p, dpb at (rec_ix, igz)
// The adjoint runs backward with bg reconstruction by exact reversal and is
// the transpose of the Born map above (dot-product test).
// ---------------------------------------------------------------------------
template <typename R>
struct Acou {
    Grid<R> g;
    std::vector<R> b, k, m;
    int nx, nz;
};

template <typename R>
Acou<R> acou_setup(const orc_shot* s, const R* vp, const R* rho, const R* refl) {
    Acou<R> A;
    A.nx = s->nx;
    A.nz = s->allnz;
    A.g = Grid<R>{s->nx + 2 * s->pml, s->allnz + 2 * s->pml, s->pml, (R)s->dx, (R)s->dz, (R)s->dt};
    const size_t n = A.g.n();
    std::vector<R> vpp(n, 0), vsp(n, 0), rp(n, 0);
    pad_window(vp, s->allnz, s->ngx_min, s->nx, A.nz, s->pml, vpp.data());
    if (rho) pad_window(rho, s->allnz, s->ngx_min, s->nx, A.nz, s->pml, rp.data());
    if (s->boundary == 1) random_boundary(A.g, (uint64_t)s->rb_seed, s->rb_shot, (R)s->rb_amp, vpp.data(), vsp.data());
    A.b.assign(n, 0); A.k.assign(n, 0); A.m.assign(n, 0);
    for (size_t c = 0; c < n; ++c) {
        R r = s->rho_const > 0 ? (R)s->rho_const : rp[c];
        A.b[c] = (R)1 / r;
        A.k[c] = r * vpp[c] * vpp[c];
    }
    if (refl)  // reflectivity window, zero outside the physical domain
        for (int ix = 0; ix < s->nx; ++ix)
            for (int iz = 0; iz < A.nz; ++iz)
                A.m[A.g.at(ix + s->pml, iz + s->pml)] = refl[(size_t)(s->ngx_min + ix) * s->allnz + iz];
    return A;
}

template <typename R>
inline R dplus_x(const R* f, size_t c, int nzp) {
    R a = 0;
    for (int i = 1; i <= NOP; ++i) a += (R)C8[i] * (f[c + (size_t)i * nzp] - f[c - (size_t)(i - 1) * nzp]);
    return a;
}
template <typename R>
inline R dplus_z(const R* f, size_t c) {
    R a = 0;
    for (int i = 1; i <= NOP; ++i) a += (R)C8[i] * (f[c + i] - f[c - i + 1]);
    return a;
}
template <typename R>
inline R dminus_x(const R* f, size_t c, int nzp) {
    R a = 0;
    for (int i = 1; i <= NOP; ++i) a += (R)C8[i] * (f[c + (size_t)(i - 1) * nzp] - f[c - (size_t)i * nzp]);
    return a;
}
template <typename R>
inline R dminus_z(const R* f, size_t c) {
    R a = 0;
    for (int i = 1; i <= NOP; ++i) a += (R)C8[i] * (f[c + i - 1] - f[c - i]);
    return a;
}

// v update with the transient source: v += (dt/d) b D+ (p + s delta_src).
template <typename R>
void acou_v(const Grid<R>& g, std::vector<R>& vx, std::vector<R>& vz, std::vector<R>& p, const std::vector<R>& b,
            size_t src, R sval, R sign) {
    R keep = p[src];
    p[src] = keep + sval;
    const int nzp = g.nzp;
    const R ax = g.dt / g.dx, az = g.dt / g.dz;
#pragma omp parallel for
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            size_t c = g.at(ix, iz);
            vx[c] += sign * (ax * b[c] * dplus_x(p.data(), c, nzp));
            vz[c] += sign * (az * b[c] * dplus_z(p.data(), c));
        }
    p[src] = keep;
}

template <typename R>
void acou_forward(const orc_shot* s, const R* vp, const R* rho, const R* refl, const R* wav, R* rec_bg,
                  R* rec_born, R* final3) {
    Acou<R> A = acou_setup<R>(s, vp, rho, refl);
    const Grid<R>& g = A.g;
    const size_t n = g.n();
    const int nzp = g.nzp;
    const R ax = g.dt / g.dx, az = g.dt / g.dz;
    std::vector<R> vx(n, 0), vz(n, 0), p(n, 0), bvx(n, 0), bvz(n, 0), bp(n, 0), dp(n, 0);
    const size_t src = g.at(s->isx, s->isz);
    for (int it = 0; it < s->nt; ++it) {
        acou_v(g, vx, vz, p, A.b, src, wav[it], (R)1);
#pragma omp parallel for
        for (int ix = NOP; ix < g.nxp - NOP; ++ix)
            for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
                size_t c = g.at(ix, iz);
                bvx[c] += ax * A.b[c] * dplus_x(bp.data(), c, nzp);
                bvz[c] += az * A.b[c] * dplus_z(bp.data(), c);
            }
#pragma omp parallel for
        for (int ix = NOP; ix < g.nxp - NOP; ++ix)
            for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
                size_t c = g.at(ix, iz);
                R d = A.k[c] * (ax * dminus_x(vx.data(), c, nzp) + az * dminus_z(vz.data(), c));
                dp[c] = d;
            }
#pragma omp parallel for
        for (int ix = NOP; ix < g.nxp - NOP; ++ix)
            for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
                size_t c = g.at(ix, iz);
                R bd = A.k[c] * (ax * dminus_x(bvx.data(), c, nzp) + az * dminus_z(bvz.data(), c));
                p[c] += dp[c];
                bp[c] += bd + 2 * A.m[c] * dp[c];
            }
        for (int r = 0; r < s->ntr; ++r) {
            size_t c = g.at(s->rec_ix[r], s->igz);
            if (rec_bg) rec_bg[(size_t)r * s->nt + it] = p[c];
            if (rec_born) rec_born[(size_t)r * s->nt + it] = bp[c];
        }
    }
    if (final3) {
        std::copy(vx.begin(), vx.end(), final3);
        std::copy(vz.begin(), vz.end(), final3 + n);
        std::copy(p.begin(), p.end(), final3 + 2 * n);
    }
}

// Adjoint of acou_forward's Born map: grad (global allnx*allnz, +=).
template <typename R>
void acou_adjoint(const orc_shot* s, const R* vp, const R* rho, const R* wav, const R* res, R* grad_glob,
                  R* grad_pad_out) {
    Acou<R> A = acou_setup<R>(s, vp, rho, nullptr);
    const Grid<R>& g = A.g;
    const size_t n = g.n();
    const int nzp = g.nzp;
    const R ax = g.dt / g.dx, az = g.dt / g.dz;
    std::vector<R> vx(n, 0), vz(n, 0), p(n, 0), lvx(n, 0), lvz(n, 0), lp(n, 0), w(n, 0), grad(n, 0), dp(n, 0);
    const size_t src = g.at(s->isx, s->isz);
    // background forward to t = nt
    for (int it = 0; it < s->nt; ++it) {
        acou_v(g, vx, vz, p, A.b, src, wav[it], (R)1);
#pragma omp parallel for
        for (int ix = NOP; ix < g.nxp - NOP; ++ix)
            for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
                size_t c = g.at(ix, iz);
                dp[c] = A.k[c] * (ax * dminus_x(vx.data(), c, nzp) + az * dminus_z(vz.data(), c));
            }
#pragma omp parallel for
        for (int ix = NOP; ix < g.nxp - NOP; ++ix)
            for (int iz = NOP; iz < g.nzp - NOP; ++iz) p[g.at(ix, iz)] += dp[g.at(ix, iz)];
    }
    auto inU = [&](int ix, int iz) { return ix >= NOP && ix < g.nxp - NOP && iz >= NOP && iz < g.nzp - NOP; };
    for (int it = s->nt - 1; it >= 0; --it) {
        for (int r = 0; r < s->ntr; ++r) lp[g.at(s->rec_ix[r], s->igz)] += res[(size_t)r * s->nt + it];
        // dp_n = k D- v^{n+1}; image; p^n = p^{n+1} - dp_n
#pragma omp parallel for
        for (int ix = NOP; ix < g.nxp - NOP; ++ix)
            for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
                size_t c = g.at(ix, iz);
                R d = A.k[c] * (ax * dminus_x(vx.data(), c, nzp) + az * dminus_z(vz.data(), c));
                grad[c] += 2 * d * lp[c];
                dp[c] = d;
            }
        // A_p^T: lv += -(dt/d) D+ (k lp) with k lp masked to U
        for (int ix = 0; ix < g.nxp; ++ix)
            for (int iz = 0; iz < g.nzp; ++iz) {
                size_t c = g.at(ix, iz);
                w[c] = inU(ix, iz) ? A.k[c] * lp[c] : 0;
            }
#pragma omp parallel for
        for (int ix = NOP; ix < g.nxp - NOP; ++ix)
            for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
                size_t c = g.at(ix, iz);
                p[c] -= dp[c];
                lvx[c] += -(ax * dplus_x(w.data(), c, nzp));
                lvz[c] += -(az * dplus_z(w.data(), c));
            }
        // bg velocity reversal with the transient source
        acou_v(g, vx, vz, p, A.b, src, wav[it], (R)-1);
        // A_v^T: lp += -(dt/d) D- (b lv) with b lv masked to U
        std::vector<R> ux(n, 0), uz(n, 0);
        for (int ix = 0; ix < g.nxp; ++ix)
            for (int iz = 0; iz < g.nzp; ++iz) {
                size_t c = g.at(ix, iz);
                if (inU(ix, iz)) { ux[c] = A.b[c] * lvx[c]; uz[c] = A.b[c] * lvz[c]; }
            }
#pragma omp parallel for
        for (int ix = NOP; ix < g.nxp - NOP; ++ix)
            for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
                size_t c = g.at(ix, iz);
                lp[c] += -(ax * dminus_x(ux.data(), c, nzp) + az * dminus_z(uz.data(), c));
            }
    }
    if (grad_glob)
        for (int ix = 0; ix < s->nx; ++ix)
            for (int iz = 0; iz < A.nz; ++iz)
                grad_glob[(size_t)(s->ngx_min + ix) * s->allnz + iz] += grad[g.at(ix + s->pml, iz + s->pml)];
    if (grad_pad_out) std::copy(grad.begin(), grad.end(), grad_pad_out);
}

// ---------------------------------------------------------------------------
// Plain-L2 elastic Born pair (random boundary, no damping, unsplit fields).
//
// Forward, per step: background velocity with the transient source on txx,
// tzz at the source cell (FD2DGPU.cu:2039-2049; update_v, born.cu:215-279),
// background stress (update_txx_tzz/update_txz, born.cu:282-349), scattered
// velocity, scattered stress plus the Born source of add_born_source
// (born.cu:351-405, direction 1) on [pml, n-pml)^2 from the new background
// velocity, then the scattered vz, vx at the receivers (FD2DGPU.cu:3093-3101).
// Without damping the split fields sum to the unsplit update.  The adjoint is
// the exact transpose of this linear map: its imaging products are those of
// image_gradient (FD2DGPU.cu:3336-3339, -2 vp^2 rho and -2 vs^2 rho forms).
template <typename R>
struct EState {
    std::vector<R> vx, vz, txx, tzz, txz;
    void init(size_t n) {
        vx.assign(n, 0); vz.assign(n, 0); txx.assign(n, 0); tzz.assign(n, 0); txz.assign(n, 0);
    }
};

template <typename R>
struct EBorn {
    Setup<R> S;
    std::vector<R> dmp, dms;  // padded image windows, zero outside the physical window
};

template <typename R>
EBorn<R> eborn_setup(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* ipp, const R* ips) {
    EBorn<R> E;
    E.S = make_setup<R>(s, vp, vs, rho);
    const Grid<R>& g = E.S.g;
    E.dmp.assign(g.n(), 0);
    E.dms.assign(g.n(), 0);
    for (int ix = 0; ix < s->nx; ++ix)
        for (int iz = 0; iz < s->allnz; ++iz) {
            const size_t c = g.at(ix + s->pml, iz + s->pml), q = (size_t)(s->ngx_min + ix) * s->allnz + iz;
            if (ipp) E.dmp[c] = ipp[q];
            if (ips) E.dms[c] = ips[q];
        }
    return E;
}

// v += (dt/d) (1/rho) D sigma, the source w added to txx and tzz at src for
// this update only (src = n: no source).
template <typename R>
void eb_vel(const Grid<R>& g, const Medium<R>& m, EState<R>& f, size_t src, R w) {
    const int nzp = g.nzp;
    const R ax = g.dt / g.dx, az = g.dt / g.dz;
    const bool has = src < g.n();
    R k1 = 0, k2 = 0;
    if (has) { k1 = f.txx[src]; k2 = f.tzz[src]; f.txx[src] = k1 + w; f.tzz[src] = k2 + w; }
#pragma omp parallel for schedule(static)
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            const size_t c = g.at(ix, iz);
            const R br = (R)1 / m.rho[c];
            f.vx[c] += br * ax * dplus_x(f.txx.data(), c, nzp) + br * az * dminus_z(f.txz.data(), c);
            f.vz[c] += br * ax * dminus_x(f.txz.data(), c, nzp) + br * az * dplus_z(f.tzz.data(), c);
        }
    if (has) { f.txx[src] = k1; f.tzz[src] = k2; }
}

template <typename R>
void eb_str(const Grid<R>& g, const Medium<R>& m, EState<R>& f) {
    const int nzp = g.nzp;
    const R ax = g.dt / g.dx, az = g.dt / g.dz;
#pragma omp parallel for schedule(static)
    for (int ix = NOP; ix < g.nxp - NOP; ++ix)
        for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
            const size_t c = g.at(ix, iz);
            const R dvx = dminus_x(f.vx.data(), c, nzp), dvz = dminus_z(f.vz.data(), c);
            const R dvxz = dplus_z(f.vx.data(), c), dvzx = dplus_x(f.vz.data(), c, nzp);
            const R l = m.lam[c], mu = m.mu[c], l2m = l + 2 * mu;
            f.txx[c] += ax * l2m * dvx + az * l * dvz;
            f.tzz[c] += ax * l * dvx + az * l2m * dvz;
            f.txz[c] += mu * (ax * dvzx + az * dvxz);
        }
}

// Background strains of the Born source: exx = D-x vx / dx, ezz = D-z vz / dz,
// exz = D+x vz / dx + D+z vx / dz (add_born_source's vxx/dx, vzz/dz, vzx/dx + vxz/dz).
template <typename R>
void eb_strain(const Grid<R>& g, const EState<R>& f, size_t c, R& exx, R& ezz, R& exz) {
    exx = dminus_x(f.vx.data(), c, g.nzp) / g.dx;
    ezz = dminus_z(f.vz.data(), c) / g.dz;
    exz = dplus_x(f.vz.data(), c, g.nzp) / g.dx + dplus_z(f.vx.data(), c) / g.dz;
}

template <typename R>
void eborn_forward(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* ipp, const R* ips,
                   const R* wav, R* rec_z, R* rec_x) {
    EBorn<R> E = eborn_setup<R>(s, vp, vs, rho, ipp, ips);
    const Grid<R>& g = E.S.g;
    const Medium<R>& m = E.S.m;
    const size_t n = g.n(), src = g.at(s->isx, s->isz);
    EState<R> bg, b;
    bg.init(n);
    b.init(n);
    const int p = g.pml;
    for (int it = 0; it < s->nt; ++it) {
        eb_vel(g, m, bg, src, wav[it]);
        eb_str(g, m, bg);
        eb_vel(g, m, b, n, (R)0);
        eb_str(g, m, b);
#pragma omp parallel for schedule(static)
        for (int ix = p; ix < g.nxp - p; ++ix)
            for (int iz = p; iz < g.nzp - p; ++iz) {
                const size_t c = g.at(ix, iz);
                R exx, ezz, exz;
                eb_strain(g, bg, c, exx, ezz, exz);
                const R l2m = m.lam[c] + 2 * m.mu[c], mu = m.mu[c];
                const R mp = -E.dmp[c], ms = -E.dms[c];
                b.txx[c] += 2 * l2m * mp * exx + 2 * (l2m * mp - 2 * mu * ms) * ezz;
                b.tzz[c] += 2 * (l2m * mp - 2 * mu * ms) * exx + 2 * l2m * mp * ezz;
                b.txz[c] += 2 * mu * ms * exz;
            }
        record_pair(g, s, b.vz, b.vx, it, rec_z, rec_x);
    }
}

// Exact transpose of eborn_forward: residual (res_z, res_x) -> (g_pp, g_ps),
// added into the global images.  The background strains of every step are
// kept from one forward sweep.
template <typename R>
void eborn_adjoint(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* wav, const R* res_z,
                   const R* res_x, R* gpp_glob, R* gps_glob) {
    EBorn<R> E = eborn_setup<R>(s, vp, vs, rho, nullptr, nullptr);
    const Grid<R>& g = E.S.g;
    const Medium<R>& m = E.S.m;
    const size_t n = g.n(), src = g.at(s->isx, s->isz);
    const int p = g.pml, nx = s->nx, nz = s->allnz, nzp = g.nzp, nt = s->nt;
    const size_t nI = (size_t)nx * nz;
    std::vector<R> hist((size_t)nt * nI * 3);
    EState<R> bg;
    bg.init(n);
    for (int it = 0; it < nt; ++it) {
        eb_vel(g, m, bg, src, wav[it]);
        eb_str(g, m, bg);
        for (int ix = 0; ix < nx; ++ix)
            for (int iz = 0; iz < nz; ++iz) {
                R* h = &hist[((size_t)it * nI + (size_t)ix * nz + iz) * 3];
                eb_strain(g, bg, g.at(ix + p, iz + p), h[0], h[1], h[2]);
            }
    }
    auto inU = [&](int ix, int iz) { return ix >= NOP && ix < g.nxp - NOP && iz >= NOP && iz < g.nzp - NOP; };
    const R ax = g.dt / g.dx, az = g.dt / g.dz;
    EState<R> lm;
    lm.init(n);
    std::vector<R> wa(n), wb(n), wc(n), ux(n), uz(n), gpp(n, 0), gps(n, 0);
    for (int it = nt - 2; it >= 0; --it) {
        // receiver injection of step it+1 (the transpose of its recorder)
        for (int r = 0; r < s->ntr; ++r) {
            const size_t c = g.at(s->rec_ix[r], s->igz);
            lm.vx[c] += res_x[(size_t)r * nt + it + 1];
            lm.vz[c] += res_z[(size_t)r * nt + it + 1];
        }
        // transpose of the stress update: v* += -(dt/d) D (C sigma*), C sigma* masked to U
        for (int ix = 0; ix < g.nxp; ++ix)
            for (int iz = 0; iz < g.nzp; ++iz) {
                const size_t c = g.at(ix, iz);
                const bool u = inU(ix, iz);
                const R l = m.lam[c], mu = m.mu[c], l2m = l + 2 * mu;
                wa[c] = u ? l2m * lm.txx[c] + l * lm.tzz[c] : 0;
                wb[c] = u ? l * lm.txx[c] + l2m * lm.tzz[c] : 0;
                wc[c] = u ? mu * lm.txz[c] : 0;
            }
#pragma omp parallel for schedule(static)
        for (int ix = NOP; ix < g.nxp - NOP; ++ix)
            for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
                const size_t c = g.at(ix, iz);
                lm.vx[c] += -(ax * dplus_x(wa.data(), c, nzp) + az * dminus_z(wc.data(), c));
                lm.vz[c] += -(az * dplus_z(wb.data(), c) + ax * dminus_x(wc.data(), c, nzp));
            }
        // transpose of the velocity update: sigma* += -(dt/d) D ((1/rho) v*), masked to U
        for (int ix = 0; ix < g.nxp; ++ix)
            for (int iz = 0; iz < g.nzp; ++iz) {
                const size_t c = g.at(ix, iz);
                const bool u = inU(ix, iz);
                const R br = (R)1 / m.rho[c];
                ux[c] = u ? br * lm.vx[c] : 0;
                uz[c] = u ? br * lm.vz[c] : 0;
            }
#pragma omp parallel for schedule(static)
        for (int ix = NOP; ix < g.nxp - NOP; ++ix)
            for (int iz = NOP; iz < g.nzp - NOP; ++iz) {
                const size_t c = g.at(ix, iz);
                lm.txx[c] += -(ax * dminus_x(ux.data(), c, nzp));
                lm.tzz[c] += -(az * dminus_z(uz.data(), c));
                lm.txz[c] += -(az * dplus_z(ux.data(), c) + ax * dplus_x(uz.data(), c, nzp));
            }
        // transpose of the Born source of step it
#pragma omp parallel for schedule(static)
        for (int ix = 0; ix < nx; ++ix)
            for (int iz = 0; iz < nz; ++iz) {
                const size_t c = g.at(ix + p, iz + p);
                const R* h = &hist[((size_t)it * nI + (size_t)ix * nz + iz) * 3];
                const R l2m = m.lam[c] + 2 * m.mu[c], mu = m.mu[c];
                gpp[c] += -2 * l2m * (h[0] + h[1]) * (lm.txx[c] + lm.tzz[c]);
                gps[c] += 4 * mu * (h[1] * lm.txx[c] + h[0] * lm.tzz[c]) - 2 * mu * h[2] * lm.txz[c];
            }
    }
    for (int ix = 0; ix < nx; ++ix)
        for (int iz = 0; iz < nz; ++iz) {
            const size_t c = g.at(ix + p, iz + p), q = (size_t)(s->ngx_min + ix) * s->allnz + iz;
            if (gpp_glob) gpp_glob[q] += gpp[c];
            if (gps_glob) gps_glob[q] += gps[c];
        }
}

// laplace.cpp:82-106: 5-point Laplacian on [2, n-2) with a zero frame.
template <typename R>
void laplace_order(int nz, int nx, R dx, R* dat) {
    std::vector<R> t1((size_t)nx * nz), t2((size_t)nx * nz, 0);
    std::copy(dat, dat + (size_t)nx * nz, t1.begin());
    for (int ix = 2; ix < nx - 2; ++ix)
        for (int iz = 2; iz < nz - 2; ++iz) {
            size_t c = (size_t)ix * nz + iz;
            // laplace.cpp:95: `1.0/dx/dx` is a double expression; the float sum is
            // widened, multiplied in double and rounded once to R.
            const R sum = t1[c + nz] + t1[c - nz] - 2 * t1[c] + t1[c + 1] + t1[c - 1] - 2 * t1[c];
            t2[c] = (R)((1.0 / (double)dx / (double)dx) * (double)sum);
        }
    std::copy(t2.begin(), t2.end(), dat);
}

// laplace.cpp:3-57 band filter, stated intent (padded stride nz+4, zero frame;
// the reference indexes its (nx+4)(nz+4) buffers with stride nz and reads
// uninitialised frame cells):  out = Sx( orig - (Sx o Sz)^4 orig ), S = (1/4,1/2,1/4).
// The n3 = 2 loop re-reads the same tmp2, so it equals one Sx pass.
template <typename R>
void laplace_band(int nz, int nx, R* dat) {
    const int s = nz + 4;
    const R a = (R)0.5, b = (R)0.25;
    std::vector<R> t0((size_t)(nx + 4) * s, 0), t1(t0.size(), 0), t2(t0.size(), 0);
    auto I = [&](int ix, int iz) { return (size_t)(ix + 2) * s + (iz + 2); };
    for (int ix = 0; ix < nx; ++ix)
        for (int iz = 0; iz < nz; ++iz) t0[I(ix, iz)] = t1[I(ix, iz)] = dat[(size_t)ix * nz + iz];
    for (int k = 0; k < 4; ++k) {
        for (int ix = 0; ix < nx; ++ix)
            for (int iz = 0; iz < nz; ++iz)
                t2[I(ix, iz)] = t1[I(ix, iz)] * a + t1[I(ix, iz - 1)] * b + t1[I(ix, iz + 1)] * b;
        for (int ix = 0; ix < nx; ++ix)
            for (int iz = 0; iz < nz; ++iz)
                t1[I(ix, iz)] = t2[I(ix, iz)] * a + t2[I(ix - 1, iz)] * b + t2[I(ix + 1, iz)] * b;
    }
    for (int ix = 0; ix < nx; ++ix)
        for (int iz = 0; iz < nz; ++iz) t2[I(ix, iz)] = t0[I(ix, iz)] - t1[I(ix, iz)];
    for (int ix = 0; ix < nx; ++ix)
        for (int iz = 0; iz < nz; ++iz)
            t1[I(ix, iz)] = t2[I(ix, iz)] * a + t2[I(ix - 1, iz)] * b + t2[I(ix + 1, iz)] * b;
    for (int ix = 0; ix < nx; ++ix)
        for (int iz = 0; iz < nz; ++iz) dat[(size_t)ix * nz + iz] = t1[I(ix, iz)];
}

// Generic elastic stepper for invariant tests: nsteps of (v then s) for
// dir = +1 or (s then v) for dir = -1 with the given damping, 15-array state.
template <typename R>
void elastic_run(int nxp, int nzp, R dx, R dz, R dt, const R* lam, const R* mu, const R* rho, const R* dampx,
                 const R* dampz, int nsteps, int dir, R* fields15) {
    Grid<R> g{nxp, nzp, 0, dx, dz, dt};
    const size_t n = g.n();
    Medium<R> m;
    m.lam.assign(lam, lam + n); m.mu.assign(mu, mu + n); m.rho.assign(rho, rho + n);
    m.dampx.assign(dampx, dampx + n); m.dampz.assign(dampz, dampz + n);
    Elastic<R> f;
    f.init(n);
    auto all = f.all();
    for (size_t a = 0; a < all.size(); ++a) std::copy(fields15 + a * n, fields15 + (a + 1) * n, all[a]->begin());
    for (int k = 0; k < nsteps; ++k) {
        if (dir > 0) { update_v(g, f, m, 1); update_s(g, f, m, 1); }
        else { update_s(g, f, m, -1); update_v(g, f, m, -1); }
    }
    for (size_t a = 0; a < all.size(); ++a) std::copy(all[a]->begin(), all[a]->end(), fields15 + a * n);
}

// Adam update of lsrtm_all_iteration (main.cpp:1092-1130) on g + b_g + b_g2.
template <typename R>
void adam_update(int n, int iter, const R* g, const R* bg, const R* bg2, R* s0, R* r0, R* image) {
    const float beta1 = 0.8f, beta2 = 0.999f, alpha = 3e-5f, eps = 5e8f, gama = 0.99f;
    float alpha_k = iter == 0 ? alpha : (float)((alpha / 2) * std::pow((double)gama, (double)iter));
    int k = iter + 1;
    for (int i = 0; i < n; ++i) {
        R sum = g[i] + bg[i] + bg2[i];
        R sk = beta1 * s0[i] + (1 - beta1) * sum;
        R rk = beta2 * r0[i] + (1 - beta2) * (sum * sum);
        R sm = sk / (R)(1 - std::pow((double)beta1, (double)k));
        R rm = rk / (R)(1 - std::pow((double)beta2, (double)k));
        R d = alpha_k * (sm / (std::sqrt(rm) + eps));
        image[i] = image[i] - d;
        s0[i] = sk;
        r0[i] = rk;
    }
}

// ---------------------------------------------------------------------------
// P/S-decoupled (rotation) propagation and decomposed imaging
// ---------------------------------------------------------------------------

// Rotation forward modelling, the forward loop of ROT_EXP_IMAGE_SINGLESHOT
// (FD2DGPU.cu:3544-3551): theta += s[it] (elastic_rotation_addS_for injects
// into txx, which is theta, FD2DGPU.cu:2046-2049 and 1407), rot_v, rot_p;
// then vz, vx at the receivers (the elastic recorder, generalised to a
// receiver list).  The medium needs no density: the scheme carries vp^2 and vs^2.
// One source-side step of the rotation scheme: forward (dir 1) theta += w,
// rot_v, rot_p; reversed (dir -1) rot_p(-1), theta += w, rot_v(-1), which
// undoes the forward step exactly when there is no damping (rot_p rewrites
// theta from its parts, so the injected sample must be back in theta when
// rot_v(-1) reads it).
template <typename R>
void rot_source_step(const Setup<R>& S, Rot<R>& f, size_t src, R w, int dir) {
    const Medium<R>& m = S.m;
    if (dir > 0) {
        f.th[src] += w;
        rot_v(S.g, f, m.vp.data(), m.vs.data(), m.dampx.data(), m.dampz.data(), 1);
        rot_p(S.g, f, m.dampx.data(), m.dampz.data(), 1);
    } else {
        rot_p(S.g, f, m.dampx.data(), m.dampz.data(), -1);
        f.th[src] += w;
        rot_v(S.g, f, m.vp.data(), m.vs.data(), m.dampx.data(), m.dampz.data(), -1);
    }
}

template <typename R>
void copy_rot4(const Rot<R>& f, size_t n, R* out4) {
    const std::vector<R>* out[4] = {&f.th, &f.om, &f.vx, &f.vz};
    for (int k = 0; k < 4; ++k) std::copy(out[k]->begin(), out[k]->end(), out4 + k * n);
}

template <typename R>
void rot_forward(const orc_shot* s, const R* vp, const R* vs, const R* wav, R* rz, R* rx, R* final4) {
    Setup<R> S = make_setup<R>(s, vp, vs, nullptr);
    Rot<R> f;
    f.init(S.g.n());
    const size_t src = S.g.at(s->isx, s->isz);
    for (int it = 0; it < s->nt; ++it) {
        rot_source_step(S, f, src, wav[it], 1);
        record_pair(S.g, s, f.vz, f.vx, it, rz, rx);
    }
    if (final4) copy_rot4(f, S.g.n(), final4);
}

// The RTM's source side after reversing the last nback of nt steps: the
// state of step nt-1-nback (theta, omega, vx, vz), for the reversal pin.
template <typename R>
void rot_reconstruct(const orc_shot* s, const R* vp, const R* vs, const R* wav, int nback, R* out4) {
    Setup<R> S = make_setup<R>(s, vp, vs, nullptr);
    Rot<R> f;
    f.init(S.g.n());
    const size_t src = S.g.at(s->isx, s->isz);
    for (int it = 0; it < s->nt; ++it) rot_source_step(S, f, src, wav[it], 1);
    for (int it = s->nt - 1; it >= s->nt - nback; --it) rot_source_step(S, f, src, wav[it], -1);
    copy_rot4(f, S.g.n(), out4);
}

// P/S-decomposed RTM with the rotation propagator (ROT_EXP_IMAGE_SINGLESHOT,
// FD2DGPU.cu:3505-3585, commented out in the reference; restated):
//   forward as rot_forward (random boundary: no damping);
//   for it = nt-1 .. 0:
//     source side reversed: rot_p(-1), theta += s[it], rot_v(-1).  The
//       reference re-injects the source after rot_v(-1) (FD2DGPU.cu:3562-3564),
//       where rot_p overwrites theta before anything reads it, so its reversal
//       injects the source a second time instead of removing it; injecting
//       before rot_v(-1) makes the reversal exact (documented deviation);
//     receiver side: rot_v(+1), vz += rec_z[tr][it] and vx += rec_x[tr][it]
//       at the receivers (GPUapplysource_backward_from_record with stack,
//       FD2DGPU.cu:3306-3327), rot_p(+1);
//     image on the physical window (doimage, FD2DGPU.cu:3218-3233):
//       PP += (vpx vpx' + vpz vpz') / (vpx^2 + vpz^2 + eps),
//       PS += (vpx vsx' + vpz vsz') / (vpx^2 + vpz^2 + eps)   (' receiver side),
//     stacked at global column ngx_min + ix (stack_image, FD2DGPU.cu:3082-3091).
template <typename R>
void rot_rtm(const orc_shot* s, const R* vp, const R* vs, const R* wav, const R* oz, const R* ox, R* img_pp,
             R* img_ps) {
    Setup<R> S = make_setup<R>(s, vp, vs, nullptr);
    const Grid<R>& g = S.g;
    const Medium<R>& m = S.m;
    const R *P = m.vp.data(), *Sv = m.vs.data(), *dxp = m.dampx.data(), *dzp = m.dampz.data();
    Rot<R> f, b;
    f.init(g.n());
    b.init(g.n());
    const size_t src = g.at(s->isx, s->isz);
    for (int it = 0; it < s->nt; ++it) rot_source_step(S, f, src, wav[it], 1);
    const int pml = g.pml, nx = S.nx, nz = S.nz;
    const R eps = image_eps<R>();
    for (int it = s->nt - 1; it >= 0; --it) {
        rot_source_step(S, f, src, wav[it], -1);
        rot_v(g, b, P, Sv, dxp, dzp, 1);
        for (int r = 0; r < s->ntr; ++r) {
            const size_t c = g.at(s->rec_ix[r], s->igz);
            b.vz[c] += oz[(size_t)r * s->nt + it];
            b.vx[c] += ox[(size_t)r * s->nt + it];
        }
        rot_p(g, b, dxp, dzp, 1);
#pragma omp parallel for schedule(static)
        for (int ix = 0; ix < nx; ++ix)
            for (int iz = 0; iz < nz; ++iz) {
                const size_t c = g.at(ix + pml, iz + pml);
                const R sp = f.vpx[c] * f.vpx[c] + f.vpz[c] * f.vpz[c];
                const R pp = f.vpx[c] * b.vpx[c] + f.vpz[c] * b.vpz[c];
                const R ps = f.vpx[c] * b.vsx[c] + f.vpz[c] * b.vsz[c];
                const size_t o = (size_t)(s->ngx_min + ix) * nz + iz;
                img_pp[o] += pp / (sp + eps);
                img_ps[o] += ps / (sp + eps);
            }
    }
}

// Floats per cell and step of the imaging quantity of each decomposed image
// kind: 0 PP with the Poynting up/down split of vector P (up xz, down xz);
// 1 PP, 2 PS, 3 SS vector images; 4 PP, 5 PS, 6 SS scalar images.
inline int decomp_floats(int kind) { return kind == 0 ? 4 : kind <= 3 ? 2 : 1; }

// The imaging quantity of one side at cell c (side 0 source, 1 receiver):
// PS images the source's P against the receiver's S.
template <typename R>
void decomp_quantity(int kind, int side, const Grid<R>& g, size_t c, const R* VP, const R* VS,
                     const Elastic<R>& f, R* q) {
    const bool use_s = kind == 3 || kind == 6 || ((kind == 2 || kind == 5) && side == 1);
    if (kind == 0) {
        R vec[2];
        vector_at(g, c, VP, VS, vec, (R*)nullptr);
        poynting_split_at(f.txz[c], f.tzz[c], f.vx[c], f.vz[c], vec, q, q + 2);
    } else if (kind <= 3) {
        if (use_s)
            vector_at(g, c, VP, VS, (R*)nullptr, q);
        else
            vector_at(g, c, VP, VS, q, (R*)nullptr);
    } else {
        q[0] = use_s ? VS[c] : VP[c];
    }
}

template <typename R>
R decomp_image(int kind, const R* h, const R* q) {
    if (kind == 0) return image_poynting(h, h + 2, q, q + 2);
    if (kind <= 3) return image_vector(h, q);
    return image_scalar(h[0], q[0]);
}

// Elastic RTM with decomposed imaging (FD_ELASTIC, FD2DGPU.cu:2241-2517,
// which no longer compiles against FD2DGPU.cuh; restated):
//   forward, per step: txx, tzz += s[it] at the source (the live order,
//   FD2DGPU.cu:2046), update_v, Helmholtz VP/VS (FD2DGPU.cu:1963-1986), the
//   source quantity of `kind` on the physical window, kept per step (the
//   reference writes 100-step blocks to disk, FD2DGPU.cu:2330-2336), update_s;
//   backward, for it = nt-1 .. 0: txx = tzz = rec_z[tr][it] at each receiver
//   (GPUapplysource_backward assigns, FD2DGPU.cu:2128-2134), update_v,
//   Helmholtz, the receiver quantity, the imaging condition against the
//   source quantity of step it, update_s.  Step nt-1 adds nothing: its
//   backward slot is never written (FD2DGPU.cu:2420) and stays zero.
// The image is stacked at global column ngx_min + ix.
template <typename R>
void decomp_rtm(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* wav, int kind, const R* oz,
                R* img) {
    Setup<R> S = make_setup<R>(s, vp, vs, rho);
    const Grid<R>& g = S.g;
    const int pml = g.pml, nx = S.nx, nz = S.nz, K = decomp_floats(kind);
    const size_t n = g.n(), win = (size_t)nx * nz;
    std::vector<R> hist((size_t)s->nt * win * K, R(0));
    std::vector<R> VP(n, R(0)), VS(n, R(0));
    Elastic<R> f;
    f.init(n);
    const size_t src = g.at(s->isx, s->isz);
    for (int it = 0; it < s->nt; ++it) {
        f.txx[src] += wav[it];
        f.tzz[src] += wav[it];
        update_v(g, f, S.m, 1);
        helmholtz(g, f.vx.data(), f.vz.data(), VP.data(), VS.data());
        R* h = hist.data() + (size_t)it * win * K;
#pragma omp parallel for schedule(static)
        for (int ix = 0; ix < nx; ++ix)
            for (int iz = 0; iz < nz; ++iz)
                decomp_quantity(kind, 0, g, g.at(ix + pml, iz + pml), VP.data(), VS.data(), f,
                                h + ((size_t)ix * nz + iz) * K);
        update_s(g, f, S.m, 1);
    }
    f.init(n);
    std::fill(VP.begin(), VP.end(), R(0));
    std::fill(VS.begin(), VS.end(), R(0));
    for (int it = s->nt - 1; it >= 0; --it) {
        for (int r = 0; r < s->ntr; ++r) {
            const size_t c = g.at(s->rec_ix[r], s->igz);
            f.txx[c] = oz[(size_t)r * s->nt + it];
            f.tzz[c] = oz[(size_t)r * s->nt + it];
        }
        update_v(g, f, S.m, 1);
        helmholtz(g, f.vx.data(), f.vz.data(), VP.data(), VS.data());
        if (it < s->nt - 1) {
            const R* h = hist.data() + (size_t)it * win * K;
#pragma omp parallel for schedule(static)
            for (int ix = 0; ix < nx; ++ix)
                for (int iz = 0; iz < nz; ++iz) {
                    R q[4];
                    decomp_quantity(kind, 1, g, g.at(ix + pml, iz + pml), VP.data(), VS.data(), f, q);
                    img[(size_t)(s->ngx_min + ix) * nz + iz] +=
                        decomp_image(kind, h + ((size_t)ix * nz + iz) * K, q);
                }
        }
        update_s(g, f, S.m, 1);
    }
}

// nsteps rotation steps on a bare grid (dir 1: rot_v then rot_p; dir -1 the
// reversal rot_p(-1) then rot_v(-1)); f12 holds the 12 fields of Rot in order.
template <typename R>
void rot_run(int nxp, int nzp, R dx, R dz, R dt, const R* vp, const R* vs, const R* dampx, const R* dampz,
             int nsteps, int dir, R* f12) {
    Grid<R> g{nxp, nzp, 0, dx, dz, dt};
    const size_t n = g.n();
    Rot<R> f;
    f.init(n);
    auto all = f.all();
    for (size_t a = 0; a < all.size(); ++a) std::copy(f12 + a * n, f12 + (a + 1) * n, all[a]->begin());
    for (int k = 0; k < nsteps; ++k) {
        if (dir > 0) {
            rot_v(g, f, vp, vs, dampx, dampz, 1);
            rot_p(g, f, dampx, dampz, 1);
        } else {
            rot_p(g, f, dampx, dampz, -1);
            rot_v(g, f, vp, vs, dampx, dampz, -1);
        }
    }
    for (size_t a = 0; a < all.size(); ++a) std::copy(all[a]->begin(), all[a]->end(), f12 + a * n);
}

}  // namespace

// ---------------------------------------------------------------------------
// extern "C" surface (d = FP64 truth, f = FP32 restatement)
// ---------------------------------------------------------------------------
extern "C" {

int orc_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

#define DEF_T(R, SUF)                                                                                        \
    int orc_ricker_##SUF(int nt, double f, double dt, R* s) { ricker<R>(nt, (R)f, (R)dt, s); return 0; }     \
    int orc_pad_window_##SUF(const R* glob, int allnz, int ngx_min, int nx, int nz, int pml, R* out) {        \
        pad_window<R>(glob, allnz, ngx_min, nx, nz, pml, out); return 0; }                                     \
    int orc_damp_profile_##SUF(int nxp, int nzp, int pml, const R* vp, double rr, R* dx_, R* dz_) {           \
        Grid<R> g{nxp, nzp, pml, 1, 1, 1}; damp_profile<R>(g, vp, (R)rr, dx_, dz_); return 0; }               \
    int orc_random_boundary_##SUF(int nxp, int nzp, int pml, unsigned long long seed, int shot, double amp,    \
                                  R* vp, R* vs) {                                                              \
        Grid<R> g{nxp, nzp, pml, 1, 1, 1}; random_boundary<R>(g, seed, shot, (R)amp, vp, vs); return 0; }     \
    int orc_elastic_forward_##SUF(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* wav,    \
                                  R* rz, R* rx, R* final5) {                                                   \
        elastic_forward<R>(s, vp, vs, rho, wav, rz, rx, final5); return 0; }                                   \
    int orc_elsrtm_shot_##SUF(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* wav,        \
                              const R* ipp, const R* ips, const R* oz, const R* ox, int iter, int col0,        \
                              R* grads7, double* misfit) {                                                     \
        elsrtm_shot<R>(s, vp, vs, rho, wav, ipp, ips, oz, ox, iter, col0, grads7, misfit); return 0; }        \
    int orc_acoustic_ref_forward_##SUF(const orc_shot* s, const R* vel, const R* rho, const R* wav, int sxnum, \
                                       R* rec) {                                                               \
        acoustic_ref_forward<R>(s, vel, rho, wav, sxnum, rec); return 0; }                                     \
    int orc_elsrtm_born_##SUF(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* wav,         \
                              const R* ipp, const R* ips, int col0, int order, R* rz, R* rx) {                 \
        elsrtm_born<R>(s, vp, vs, rho, wav, ipp, ips, col0, order, rz, rx); return 0; }                        \
    int orc_kernel_##SUF(int which, int nxp, int nzp, int pml, double dx, double dz, double dt, const R* vp,   \
                         const R* vs, const R* rho, const R* lam, const R* mu, const R* dampx, const R* dampz, \
                         const R* dmp, const R* dms, int dir, const R* src15, R* f15, R* strain4) {             \
        kernel<R>(which, nxp, nzp, pml, (R)dx, (R)dz, (R)dt, vp, vs, rho, lam, mu, dampx, dampz, dmp, dms, dir, \
                  src15, f15, strain4); return 0; }                                                             \
    int orc_edge_io_##SUF(int nxp, int nzp, int pml, int nt, R* field, R* bands, int it, int save) {           \
        edge_io_x<R>(nxp, nzp, pml, nt, field, bands, it, save); return 0; }                                   \
    int orc_image_##SUF(int nxp, int nzp, int pml, const R* st4, const R* bk15, const R* vp, const R* vs,       \
                        const R* rho, int col0, int allnz, R* pp, R* ps, R* il) {                              \
        image_x<R>(nxp, nzp, pml, st4, bk15, vp, vs, rho, col0, allnz, pp, ps, il); return 0; }                \
    int orc_acoustic_ref_damp_##SUF(int nxp, int nzp, int pml, const R* vt, double rr, R* damp) {              \
        Grid<R> g{nxp, nzp, pml, 1, 1, 1}; acoustic_ref_damp<R>(g, vt, (R)rr, damp); return 0; }              \
    int orc_acoustic_ref_step_##SUF(int nxp, int nzp, double dx, double dz, double dt, const R* k,             \
                                    const R* rr, const R* damp, R* px, R* pz, R* p, R* vx, R* vz) {           \
        Grid<R> g{nxp, nzp, 0, (R)dx, (R)dz, (R)dt};                                                          \
        acoustic_ref_step<R>(g, k, rr, damp, px, pz, p, vx, vz); return 0; }                                   \
    int orc_acou_forward_##SUF(const orc_shot* s, const R* vp, const R* rho, const R* refl, const R* wav,     \
                               R* rbg, R* rborn, R* final3) {                                                  \
        acou_forward<R>(s, vp, rho, refl, wav, rbg, rborn, final3); return 0; }                               \
    int orc_acou_adjoint_##SUF(const orc_shot* s, const R* vp, const R* rho, const R* wav, const R* res,      \
                               R* grad, R* gpad) {                                                             \
        acou_adjoint<R>(s, vp, rho, wav, res, grad, gpad); return 0; }                                         \
    int orc_eborn_forward_##SUF(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* ipp,      \
                                const R* ips, const R* wav, R* rz, R* rx) {                                     \
        eborn_forward<R>(s, vp, vs, rho, ipp, ips, wav, rz, rx); return 0; }                                   \
    int orc_eborn_adjoint_##SUF(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* wav,      \
                                const R* res_z, const R* res_x, R* gpp, R* gps) {                             \
        eborn_adjoint<R>(s, vp, vs, rho, wav, res_z, res_x, gpp, gps); return 0; }                            \
    int orc_laplace_order_##SUF(int nz, int nx, double dx, R* dat) { laplace_order<R>(nz, nx, (R)dx, dat); return 0; } \
    int orc_laplace_band_##SUF(int nz, int nx, R* dat) { laplace_band<R>(nz, nx, dat); return 0; }            \
    int orc_elastic_run_##SUF(int nxp, int nzp, double dx, double dz, double dt, const R* lam, const R* mu,   \
                              const R* rho, const R* dampx, const R* dampz, int nsteps, int dir, R* f15) {     \
        elastic_run<R>(nxp, nzp, (R)dx, (R)dz, (R)dt, lam, mu, rho, dampx, dampz, nsteps, dir, f15); return 0; } \
    int orc_ncc_residual_##SUF(int ntr, int nt, int iter, R* sz, R* sx, const R* oz, const R* ox, double* mf) { \
        ncc_residual<R>(ntr, nt, iter, sz, sx, oz, ox, mf); return 0; }                                        \
    int orc_adam_update_##SUF(int n, int iter, const R* g, const R* bg, const R* bg2, R* s0, R* r0, R* img) { \
        adam_update<R>(n, iter, g, bg, bg2, s0, r0, img); return 0; }                                           \
    int orc_rot_forward_##SUF(const orc_shot* s, const R* vp, const R* vs, const R* wav, R* rz, R* rx,        \
                              R* final4) {                                                                     \
        rot_forward<R>(s, vp, vs, wav, rz, rx, final4); return 0; }                                            \
    int orc_rot_reconstruct_##SUF(const orc_shot* s, const R* vp, const R* vs, const R* wav, int nback,       \
                                  R* out4) {                                                                   \
        rot_reconstruct<R>(s, vp, vs, wav, nback, out4); return 0; }                                           \
    int orc_rot_rtm_##SUF(const orc_shot* s, const R* vp, const R* vs, const R* wav, const R* oz, const R* ox, \
                          R* ipp, R* ips) {                                                                    \
        rot_rtm<R>(s, vp, vs, wav, oz, ox, ipp, ips); return 0; }                                              \
    int orc_decomp_rtm_##SUF(const orc_shot* s, const R* vp, const R* vs, const R* rho, const R* wav,         \
                             int kind, const R* oz, R* img) {                                                  \
        if (kind < 0 || kind > 6) return -1;                                                                   \
        decomp_rtm<R>(s, vp, vs, rho, wav, kind, oz, img); return 0; }                                         \
    int orc_rot_run_##SUF(int nxp, int nzp, double dx, double dz, double dt, const R* vp, const R* vs,        \
                          const R* dampx, const R* dampz, int nsteps, int dir, R* f12) {                       \
        rot_run<R>(nxp, nzp, (R)dx, (R)dz, (R)dt, vp, vs, dampx, dampz, nsteps, dir, f12); return 0; }         \
    int orc_helmholtz_##SUF(int nxp, int nzp, double dx, double dz, const R* vx, const R* vz, R* VP, R* VS) { \
        Grid<R> g{nxp, nzp, 0, (R)dx, (R)dz, 1}; helmholtz<R>(g, vx, vz, VP, VS); return 0; }                 \
    int orc_vector_##SUF(int nxp, int nzp, double dx, double dz, const R* VP, const R* VS, R* vecp, R* vecs) { \
        Grid<R> g{nxp, nzp, 0, (R)dx, (R)dz, 1};                                                              \
        for (int ix = NOP; ix <= nxp - NOP; ++ix)                                                              \
            for (int iz = NOP; iz <= nzp - NOP; ++iz) {                                                        \
                const size_t c = g.at(ix, iz);                                                                 \
                vector_at<R>(g, c, VP, VS, vecp + 2 * c, vecs + 2 * c);                                        \
            }                                                                                                  \
        return 0; }                                                                                            \
    int orc_poynting_split_##SUF(long long n, const R* txz, const R* tzz, const R* vx, const R* vz,           \
                                 const R* vec2, R* up2, R* down2) {                                            \
        for (long long c = 0; c < n; ++c)                                                                      \
            poynting_split_at<R>(txz[c], tzz[c], vx[c], vz[c], vec2 + 2 * c, up2 + 2 * c, down2 + 2 * c);     \
        return 0; }                                                                                            \
    int orc_image_cond_##SUF(int kind, long long n, const R* fw, const R* bw, R* out) {                       \
        if (kind < 0 || kind > 6) return -1;                                                                   \
        const int K = decomp_floats(kind);                                                                     \
        for (long long c = 0; c < n; ++c) out[c] = decomp_image<R>(kind, fw + c * K, bw + c * K);              \
        return 0; }

DEF_T(double, d)
DEF_T(float, f)

// laplace.cpp:59-78 cutM: zero the frame outside [xi,xe) x [zi,ze).
int orc_cutM_f(float* x, int nz, int nx, int zi, int ze, int xi, int xe) {
    for (int ix = 0; ix < xi; ++ix) for (int iz = 0; iz < nz; ++iz) x[(size_t)ix * nz + iz] = 0.f;
    for (int ix = xe; ix < nx; ++ix) for (int iz = 0; iz < nz; ++iz) x[(size_t)ix * nz + iz] = 0.f;
    for (int ix = 0; ix < nx; ++ix) for (int iz = 0; iz < zi; ++iz) x[(size_t)ix * nz + iz] = 0.f;
    for (int ix = 0; ix < nx; ++ix) for (int iz = ze; iz < nz; ++iz) x[(size_t)ix * nz + iz] = 0.f;
    return 0;
}

double orc_rb_uniform(unsigned long long seed, int shot, int ix, int iz) { return rb_uniform(seed, shot, ix, iz); }

}  // extern "C"
