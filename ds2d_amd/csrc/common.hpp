// Shared definitions of the MI355X (gfx950) DS2D kernels and their host side.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace ds2d {

constexpr int NOP = 4;
// FD2DGPU.cuh:16 coeff2: exact 8th-order staggered weights, rounded to float
// from the double quotient as the reference's __constant__ initialiser does.
#define DS2D_C1 ((float)(1225.0 / 1024.0))
#define DS2D_C2 ((float)(-245.0 / 3072.0))
#define DS2D_C3 ((float)(49.0 / 5120.0))
#define DS2D_C4 ((float)(-5.0 / 7168.0))

// One shot window ("slot") of a batch.  Every per-cell field of a batch is one
// allocation of sum(n_s) floats; slot s owns [off, off + nxp*nzp).
struct SlotDesc {
    long long off;       // first cell of this slot in every field arena
    long long rec_off;   // first sample of this slot in every record arena ([ntr][nt])
    long long win_off;   // first cell of this slot's image window (nx*nz)
    long long edge_off;  // first float of this slot's edge-band arena
    int nxp, nzp;        // padded window
    int nx, nz;          // physical window (nz = allnz)
    int pml, col0;       // frame width, global column of physical column 0
    int isx, isz;        // source cell (padded)
    int igz, ntr;        // receiver row (padded), trace count
    int rix_off;         // first receiver column of this slot in the rec_ix arena
    int shot_id;
    int boundary;        // 0 PML, 1 random boundary
    int rmap_off;        // first column of this slot in the receiver-map arena (nxp ints)
    long long hist_off;  // first float of this slot's background-increment history (nt*nx*nz)
};

// The reference's split-PML elastic wavefield (15 arrays, FD2DGPU.cuh:280-294).
// The reference-compat gradient keeps the two source-side wavefields (the
// background and the first-order Born field, whose velocities drive a Born
// source) in FP64; every other wavefield is FP32.
template <typename T>
struct EFieldT {
    T *vx, *vz, *txx, *tzz, *txz;
    T *vx_x, *vx_z, *vz_x, *vz_z;
    T *txx_x, *txx_z, *tzz_x, *tzz_z, *txz_x, *txz_z;
};
using EField = EFieldT<float>;
using EFieldD = EFieldT<double>;

// Strain outputs of the reconstruction stress kernels (vx_gx, vz_gz, vx_gz, vz_gx).
struct Strain {
    float *gx_vx, *gz_vz, *gz_vx, *gx_vz;
};

// Per-cell medium of the padded window.
struct Med {
    float *vp, *vs, *rho, *lam, *mu, *dampx, *dampz, *dmp, *dms;
};

// Block shape of the per-cell kernels: 64 lanes along contiguous z (one
// wavefront per row of the tile), 4 rows along x.
constexpr int BZ = 64;
constexpr int BX = 4;

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace ds2d
