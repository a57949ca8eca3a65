// Host-side launchers of the DS2D gfx950 kernels.
#pragma once
#include "common.hpp"

namespace ds2d {

// Geometry of one launch batch (device pointers to the slot table and the
// receiver-column arena; scalar sizes for the grids).
struct BatchShape {
    const SlotDesc* slots;  // device
    const int* rix;         // device, receiver columns (padded)
    int nslots, max_nxp, nzp, nz, max_ntr, nt;
    long long total_cells;  // sum of nxp*nzp
    float dt, dx, dz;
};

// --- reference-compatible elastic kernels (kernels_ref.hip) ---------------
// Overloads on EFieldD run the wavefield in FP64 (the source-side fields of
// the reference-compat gradient); media, strains and records stay FP32.
void launch_ref_v(hipStream_t st, const BatchShape& b, EField f, Med m, float dir, int recon);
void launch_ref_v(hipStream_t st, const BatchShape& b, EFieldD f, Med m, float dir, int recon);
void launch_ref_s(hipStream_t st, const BatchShape& b, EField f, Med m, Strain s, float dir, int recon);
void launch_ref_s(hipStream_t st, const BatchShape& b, EFieldD f, Med m, Strain s, float dir, int recon);
void launch_adj_v(hipStream_t st, const BatchShape& b, EField f, Med m, float dir);
void launch_adj_s(hipStream_t st, const BatchShape& b, EField f, Med m, float dir);
void launch_born_src(hipStream_t st, const BatchShape& b, const float* vx, const float* vz, EField tgt, Med m,
                     float dir);
void launch_born_src(hipStream_t st, const BatchShape& b, const double* vx, const double* vz, EFieldD tgt, Med m,
                     float dir);
void launch_born_src(hipStream_t st, const BatchShape& b, const double* vx, const double* vz, EField tgt, Med m,
                     float dir);
void launch_born_adj(hipStream_t st, const BatchShape& b, EField src, EField tgt, Med m, float dir);
void launch_inject_src(hipStream_t st, const BatchShape& b, float* txx, float* tzz, const float* wav, int it);
void launch_inject_src(hipStream_t st, const BatchShape& b, double* txx, double* tzz, const float* wav, int it);
void launch_record(hipStream_t st, const BatchShape& b, const float* vz, const float* vx, float* rz, float* rx,
                   const float* vz2, const float* vx2, float* rz2, float* rx2, int it);
void launch_record(hipStream_t st, const BatchShape& b, const double* vz, const double* vx, float* rz, float* rx,
                   const float* vz2, const float* vx2, float* rz2, float* rx2, int it);
void launch_inject_res(hipStream_t st, const BatchShape& b, float* vz, float* vx, const float* rz, const float* rx,
                       int it);
void launch_edge_io(hipStream_t st, const BatchShape& b, float* fa, float* fb, float* ea, float* eb, int it,
                    int save);
void launch_edge_io(hipStream_t st, const BatchShape& b, double* fa, double* fb, double* ea, double* eb, int it,
                    int save);
void launch_image(hipStream_t st, const BatchShape& b, Strain sg, Strain sb, EField bk, EField bb, EField bg, Med m,
                  float* w7, long long wstride);
void launch_image(hipStream_t st, const BatchShape& b, Strain sg, Strain sb, EField bk, EField bb, EFieldD bg, Med m,
                  float* w7, long long wstride);

// --- setup and reductions (kernels_setup.hip) ------------------------------
struct SetupArgs {
    const float *vp_all, *vs_all, *rho_all;  // global model (device); rho_all may be null
    const float *img_pp, *img_ps;            // global image (device); may be null
    int allnx, allnz;
    float pml_rr, rho_const, rb_amp;
    unsigned long long rb_seed;
    int acoustic;                            // 1: also emit buoyancy/modulus for the acoustic kernels
};
// Writes vp, vs, rho, lam, mu, dampx, dampz, dmp, dms of every slot.
void launch_setup(hipStream_t st, const BatchShape& b, const SetupArgs& a, Med m);
// Same, plus buoyancy 1/rho, modulus rho*vp^2 and the reflectivity window with
// a zero frame for the plain acoustic kernels.
void launch_setup_acoustic(hipStream_t st, const BatchShape& b, const SetupArgs& a, Med m, float* bu, float* kk,
                           float* refl);
// Reference acoustic damping: one array, z profile over the x profile (FD2DGPU.cu:505-519).
void launch_combine_damp(hipStream_t st, const BatchShape& b, const float* dampx, const float* dampz, float* damp);
// Deterministic gather of per-slot window images into global images:
// glob[k][gx][z] += sum over slots covering gx (slot order) of win[k].
void launch_gather(hipStream_t st, const BatchShape& b, const float* win, long long wstride, float* glob,
                   long long gstride, int nimg, int allnx, int allnz, const int* skip = nullptr);
// Per-slot non-finite flags of records (windows = 0) or image windows (windows = 1).
void launch_slot_nonfinite(hipStream_t st, const BatchShape& b, const float* a, long long astride, int narr,
                           int windows, int* flags);
// Per-trace FP64 sums for the NCC residual: out[slot_trace][6] =
// {|d_x|^2, |D_x|^2, d_x.D_x, |d_z|^2, |D_z|^2, d_z.D_z}.
void launch_ncc_sums(hipStream_t st, const BatchShape& b, const float* sz, const float* sx, const float* oz,
                     const float* ox, double* sums, const int* trace_base);
void launch_ncc_apply(hipStream_t st, const BatchShape& b, float* sz, float* sx, const float* oz, const float* ox,
                      const double* sums, const int* trace_base, int iter);
void launch_fill(hipStream_t st, float* p, long long n, float v);
void launch_axpby(hipStream_t st, float* y, const float* x, long long n, float a, float b);  // y = a*x + b*y
// Partial FP64 dot products, finished on the host in a fixed order.
void launch_dot(hipStream_t st, const float* x, const float* y, long long n, double* partial, int nblocks);

// --- acoustic kernels (kernels_acoustic.hip) -------------------------------
struct AcouField {
    float *vx, *vz, *p;
};
// The background of the plain acoustic Born pair is carried in FP64: its
// increment drives the Born source at every step, and FP32 rounding of a
// 2500-step background alone puts the Born records 1e-5 off FP64.  Only the
// increment, rounded once to FP32, reaches the Born field and the history.
struct AcouFieldD {
    double *vx, *vz, *p;
};
struct AcouMed {
    const float *b, *k, *m;       // buoyancy, modulus, reflectivity window (zero frame)
    const float *damp, *rho;      // reference acoustic layout only
};
void launch_acou_v(hipStream_t st, const BatchShape& b, AcouFieldD bg, AcouField born, AcouMed m, const float* wav,
                   int it, float sign, int do_bg, int do_born);
// hist_mode: 0 none, 1 store the background increment, 2 read it (do_bg = 0).
// Records (rec_bg / rec_born, may be null) are taken inside, through rmap.
void launch_acou_p(hipStream_t st, const BatchShape& b, AcouFieldD bg, AcouField born, AcouMed m, float* hist,
                   float* rec_bg, float* rec_born, int it, int do_bg, int do_born, int hist_mode, const int* rmap);
// Backward step it: (injection of res[it] already in lam.p) adj_p, then adj_v,
// which also injects res[it - 1] for the next step.  Step nt - 1 is injected
// by launch_acou_inj.
void launch_acou_inj(hipStream_t st, const BatchShape& b, float* lp, const float* res, int it);
// hist != null: gradient from the stored increment, background untouched
void launch_acou_adj_p(hipStream_t st, const BatchShape& b, AcouFieldD bg, AcouField lam, AcouMed m, float* grad,
                       const float* hist, int it);
void launch_acou_adj_v(hipStream_t st, const BatchShape& b, AcouFieldD bg, AcouField lam, AcouMed m,
                       const float* wav, int it, int do_bg, const int* rmap, const float* res);
struct ChunkDesc;
// Fused plain acoustic steps (kernels_acoustic_fused.hip): state A -> B, each
// 3 fields (vx, vz, p) of n floats from the base pointer.
struct AcouFusedArgs {
    const float* A;
    float* B;
    long long n;
    const float *b, *k, *m;  // buoyancy, modulus, reflectivity window
    const float* hist;       // resident background increments
    float* rec;              // Born step: records of p (or null)
    float* grad;             // adjoint step: padded gradient (read-modify-write)
    const float* res;        // adjoint step: residual records
    float ax, az;            // dt/dx, dt/dz
};
bool acou_fused_threads_ok(int threads);
void launch_acou_born_fused(hipStream_t st, const BatchShape& b, const AcouFusedArgs& a, const ChunkDesc* chunks,
                            int nchunks, const int* rmap, int it, int threads);
void launch_acou_adj_fused(hipStream_t st, const BatchShape& b, const AcouFusedArgs& a, const ChunkDesc* chunks,
                           int nchunks, const int* rmap, int it, int threads);
void launch_acou_ref_step(hipStream_t st, const BatchShape& b, float* px, float* pz, float* p, float* vx, float* vz,
                          const float* k, const float* rho, const float* damp, const float* wav, int sxnum,
                          float* rec, int it);

// --- fast fused elastic forward (kernels_fast.hip) -------------------------
struct FastArgs {
    float *vx, *vz, *txx, *tzz, *txz;          // state
    float *vx_x, *vx_z, *vz_x, *vz_z;          // split memory (used in damping strips only)
    float *txx_x, *txx_z, *tzz_x, *tzz_z, *txz_x, *txz_z;
    const float *bx, *lam, *mu;                // buoyancy*dt, lambda*dt, mu*dt
    const float *dampx, *dampz;
    const float* wav;
    float *rec_z, *rec_x;
};
void launch_fast_v(hipStream_t st, const BatchShape& b, const FastArgs& a, const int* rmap, int it);
void launch_fast_s(hipStream_t st, const BatchShape& b, const FastArgs& a);
void launch_fast_params(hipStream_t st, const float* rho, const float* lam, const float* mu, float* bx, float* ldt,
                        float* mdt, long long n, float dt);
void launch_fast_rmap(hipStream_t st, const BatchShape& b, int* rmap);

// --- plain-L2 elastic Born pair (kernels_eborn.hip) --------------------------
// Unsplit elastic wavefield (random boundary, no damping).
struct EbField {
    float *vx, *vz, *txx, *tzz, *txz;
};
struct EbMed {
    const float *rho, *lam, *mu, *dmp, *dms;  // dmp/dms: padded PP/PS image windows
};
// Adjoint products handed between the two adjoint kernels (zero outside the
// update region): C s* as wa = (l+2m) sxx* + l szz*, wb = l sxx* + (l+2m) szz*,
// wc = m sxz*; and u = (1/rho) v*.
struct EbAux {
    float *wa, *wb, *wc, *ux, *uz;
};
// One fused forward step of the pair (kernels_eborn_fused.hip): A and B hold
// 10 fields of n floats each (background vx vz txx tzz txz, then the
// scattered five); A is read, B written.  hist_mode 1 stores the background
// strains as k_eb_s does (0: none).
struct EbFusedArgs {
    const float* A;
    float* B;
    long long n;
    EbMed m;
    float* hist;
    long long hstride;
    float *rec_z, *rec_x;
};
// Forward step: k_eb_v (background with source sample w and/or scattered
// velocity), then k_eb_s (stresses, Born source, recorder).  sign = -1 undoes a
// background step.  hist_mode: 1 store the background strains (3 arrays of
// hstride floats), 2 read them (no background).
void launch_eb_v(hipStream_t st, const BatchShape& b, EbField bg, EbField born, const float* rho, float w,
                 float sign, int do_bg, int do_born);
void launch_eb_s(hipStream_t st, const BatchShape& b, EbField bg, EbField born, EbMed m, float* hist,
                 long long hstride, int it, float sign, int do_bg, int do_born, int hist_mode, const int* rmap,
                 float* rec_z, float* rec_x);
// Chunks as the fused elastic step's (rows of a block: threads - 2*FUSED_HALO);
// threads 256 or 512.  The result equals launch_eb_v + launch_eb_s with
// do_bg = do_born = 1 up to FP32 rounding (FMA contraction).
bool eb_fused_threads_ok(int threads);
// One fused step of the rotation scheme (kernels_rot_fused.hip): A and B hold
// 8 fields of n floats (theta omega vpx vpz vsx vsz, then theta_x omega_x of
// the damping cells); records of vz, vx through the column -> trace map.
struct RotFusedArgs {
    const float* A;
    float* B;
    long long n;
    const float *vp, *vs, *dampx, *dampz, *wav;
    float *rec_z, *rec_x;
};
void launch_rot_fused(hipStream_t st, const BatchShape& b, const RotFusedArgs& a, const struct ChunkDesc* chunks,
                      int nchunks, const int* rmap, int it, int threads, bool pml);
// One fused adjoint step (k_eb_adj_v with the residual of step it_inj, then
// k_eb_adj_s of step it): adjoint state and background read from A, written
// to B (5 fields each: vx vz txx tzz txz); gradients accumulated in place.
// undo: no strain history, the background steps back alongside; otherwise
// the strains come from hist (3 arrays of hstride floats).
struct EbAdjArgs {
    const float *Alm, *Abg;
    float *Blm, *Bbg;
    long long n;
    EbMed m;
    const float* hist;
    long long hstride;
    float *gpp, *gps;
    const float *res_z, *res_x;
};
void launch_eb_adj_fused(hipStream_t st, const BatchShape& b, const EbAdjArgs& a, const struct ChunkDesc* chunks,
                         int nchunks, const int* rmap, float w, int it, int it_inj, bool undo, int threads);
void launch_eb_fused(hipStream_t st, const BatchShape& b, const EbFusedArgs& a, const struct ChunkDesc* chunks,
                     int nchunks, const int* rmap, float w, float sign, int it, int hist_mode, int threads);
// Adjoint step: k_eb_adj_v (receiver injection of step it_inj, -1 none, and the
// transpose of the stress update), then k_eb_adj_s (transpose of the velocity
// update and the gradient of step it; strains from hist, or from bg when
// hist is null).  undo = 1 (no history): k_eb_adj_v also takes the background
// velocity one step back (source sample w) and k_eb_adj_s the background
// stress, so the background is rebuilt alongside the adjoint.
void launch_eb_adj_v(hipStream_t st, const BatchShape& b, EbField lm, EbAux x, EbField bg, EbMed m,
                     const int* rmap, const float* res_z, const float* res_x, int it_inj, int undo, float w);
void launch_eb_adj_s(hipStream_t st, const BatchShape& b, EbField lm, EbAux x, EbField bg, EbMed m,
                     const float* hist, long long hstride, float* gpp, float* gps, int it, int undo);

// --- P/S-decoupled rotation scheme and decomposed imaging (kernels_rot.hip) --
// The rotation wavefield (the reference aliases it onto the split elastic
// arrays, FD2DGPU.cu:1402-1412): dilatation theta and rotation omega with
// their x/z splits, the velocity and its P and S parts.
template <typename T>
struct RotFieldT {
    T *th, *om, *th_x, *th_z, *om_x, *om_z, *vx, *vz, *vpx, *vpz, *vsx, *vsz;
};
using RotField = RotFieldT<float>;
using RotFieldD = RotFieldT<double>;
// Velocity parts from theta/omega, then theta/omega from the velocity (uses
// Med vp, vs, dampx, dampz).  dir = -1 with rot_p first reverses a step.
void launch_rot_v(hipStream_t st, const BatchShape& b, RotField f, Med m, float dir);
void launch_rot_v(hipStream_t st, const BatchShape& b, RotFieldD f, Med m, float dir);
void launch_rot_p(hipStream_t st, const BatchShape& b, RotField f, Med m, float dir);
void launch_rot_p(hipStream_t st, const BatchShape& b, RotFieldD f, Med m, float dir);
void launch_rot_src(hipStream_t st, const BatchShape& b, float* th, const float* wav, int it);
void launch_rot_src(hipStream_t st, const BatchShape& b, double* th, const float* wav, int it);
// PP and PS of the rotation RTM into the window images w2 (2 x wstride).
void launch_rot_image(hipStream_t st, const BatchShape& b, RotFieldD src, RotFieldD rcv, float* w2, long long wstride);
// vz += rec_z, vx += rec_x at the receivers (FP64 receiver field of the rotation RTM).
void launch_rot_inject(hipStream_t st, const BatchShape& b, double* vz, double* vx, const float* rz, const float* rx,
                       int it);
void launch_helmholtz(hipStream_t st, const BatchShape& b, const float* vx, const float* vz, float* VP, float* VS);
// Floats per window cell and step of a decomposed image kind (0..6).
int decomp_floats(int kind);
// store = 1: the source quantity into hist (this step's block of
// decomp_floats(kind) x total_win floats); store = 0: the receiver quantity
// imaged against hist into the window image win.  Needs pml >= NOP.
void launch_decomp(hipStream_t st, const BatchShape& b, int kind, const float* VP, const float* VS, EField f,
                   float* hist, float* win, int store);
// txx = tzz = rec_z[tr][it] at each receiver (assignment).
void launch_assign_rec(hipStream_t st, const BatchShape& b, float* txx, float* tzz, const float* rz, int it);

// --- fused single-kernel elastic step (kernels_fused.hip) -------------------
struct ChunkDesc {
    int slot, x0, x1, z0;  // columns [x0, x1) of slot, rows from z0 (block of threads - 2*FUSED_HALO rows)
};
struct FusedArgs {
    // state read (A) and written (B): 15 fields of n floats each, in EField
    // order vx vz txx tzz txz | vx_x vx_z vz_x vz_z txx_x txx_z tzz_x tzz_z
    // txz_x txz_z (the split fields are used in the damping strips only)
    const float* A;
    float* B;
    long long n;
    const float* par;  // dt/rho | lambda*dt | mu*dt, n floats each
    const float *dampx, *dampz;
    const float* wav;
    float *rec_z, *rec_x;
};
// Fused addressing uses 32-bit byte offsets within each field.
constexpr long long FUSED_MAX_CELLS = (1LL << 30) - (1LL << 20);
// A fused workgroup is `threads` lanes (256, 512, 704 or 768), one z row per lane:
// FUSED_HALO rows above and below its block are loaded to feed the z taps,
// so it owns threads - 2*FUSED_HALO rows.
constexpr int FUSED_HALO = 8;
// Widest x chunk of the fused acoustic steps (the adjoint stages the
// receiver residuals of its chunk's columns in LDS).
constexpr int ACOU_FUSED_MAX_W = 1024;
bool fused_threads_ok(int threads);
// depth: columns of operands in flight per lane (2 or 4)
// pml = false compiles the damping-strip code out; valid only when no window
// of the batch has a PML (boundary != 0 for every slot).
void launch_fused_step(hipStream_t st, const BatchShape& b, const FusedArgs& a, const ChunkDesc* chunks,
                       int nchunks, const int* rmap, int it, int threads, int depth, bool pml);
// Two time steps per launch (steps it and it+1, it+1 < nt): workgroups own
// threads - 2*FUSED2_HALO rows.  scratch: 5 fields of n floats (the damped
// cells' intermediate splits; unused when pml is false).  depth: columns of
// operands in flight per lane (2).
constexpr int FUSED2_HALO = 16;
bool fused2_threads_ok(int threads);
void launch_fused2_step(hipStream_t st, const BatchShape& b, const FusedArgs& a, float* scratch,
                        const ChunkDesc* chunks, int nchunks, const int* rmap, int it, int threads, int depth,
                        bool pml);

}  // namespace ds2d
