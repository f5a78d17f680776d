// jur_kernels_misc.h -- launchers of the small kernels in kernels_misc.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "jur_kernels.h"

enum { JUR_ST_FOV = 16 };  // "Cannot apply FOV convolution!"

void launch_hydrostatic(double* z, double* p, const double* t, const double* qh2o,
                        const double* lat, const int* prof0, const int* prof1, int nprof,
                        double hydz, hipStream_t s);
void launch_fov(int nr, int seg, int nd, const double* time, const double* tpz, const double* rad_in,
                const double* tau_in, const double* dz, const double* w, int n, double* rad_out,
                double* tau_out, int* status, hipStream_t s);
void launch_brightness(int nr, int nd, const double* nu, double* rad, hipStream_t s);
void launch_f2d(const float* a, double* b, size_t n, hipStream_t s);
void launch_los64(const DevAtm& M, int a, const double4* geo, const float* rec, int R, int np,
                  int ng, int nw, double* zll, double* pt, double* q, double* k, double* u,
                  double* ds, double* af, int* ai, hipStream_t s);
