// C-ABI shim over the reference's own laplace.cpp, compiled from the source
// where it lies (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY: tests use it
// to pin the oracle's laplace_order against the real reference routine.
#include "laplace.h"

extern "C" {
int ref_laplace_order(int nz, int nx, int order, float dx, float* dat) {
    laplace_filter_2d(nz, nx, order, dx, dat);  // laplace.cpp:82-106
    return 0;
}
int ref_cutM(float* x, int nz, int nx, int zi, int ze, int xi, int xe) {
    cutM(x, nz, nx, zi, ze, xi, xe);  // laplace.cpp:59-78
    return 0;
}
}
