// jur_mie.h -- Mie kernels (kernels_mie.hip) and their inputs.
#pragma once
#include <hip/hip_runtime.h>

// Radius nodes of the aerosol configuration's Mie grid (option "mie_grid").
#define MIE_GRID_NR 64

// One sphere of the Gauss-Hermite radius quadrature at one channel.
struct MieItem {
  double x, nr, ni;  // size parameter, refractive index
  int nstop, nmx;    // series length and start of the D_n recurrence
};
// One radius node of a mode: radius [um], K1 = N 1e-3 sqrt(pi), weight.
struct MieNode {
  double rad, K1, w;
};

// grid 0: the reference's 181 angles; 1: 180 angles (k + 0.5) deg resampled
// to the 181-node layout (option "mie_grid").
void launch_bhmie(const MieItem* items, int item0, int nitems, double2* scratch, size_t stride,
                  double* qext, double* qsca, double* phase, int grid, hipStream_t s);
void launch_mie_mode(const MieNode* nodes, int nnode, const int* item_of, int nd, const double* qext,
                     const double* qsca, const double* phase, double* mbe, double* mbs, double* mp,
                     hipStream_t s);
void launch_mie_layers(int nl, const int* mode0, const int* nmod, int nd, const double* mbe,
                       const double* mbs, const double* mp, double* be, double* bs, double* ba, double* p,
                       hipStream_t s);
