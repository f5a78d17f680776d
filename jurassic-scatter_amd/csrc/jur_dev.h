// jur_dev.h -- device data layouts and __device__ helpers shared by the
// HIP kernels of the forward model (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jurassic_types.h"

// Fields of one FP32 line-of-sight point record (per ray, per point):
//   p [hPa], T [K], ds [km], aerofac, aeroi, q(H2O), k[nw], u[ng]
enum { REC_P = 0, REC_T = 1, REC_DS = 2, REC_AF = 3, REC_AI = 4, REC_QH2O = 5, REC_K = 6 };
__host__ __device__ inline int rec_len(int ng, int nw) { return REC_K + nw + ng; }
__host__ __device__ inline int rec_u(int nw) { return REC_K + nw; }

#define JUR_NPLANCK TBLNSMAX  // 1201 source-function temperatures, 100..400 K

// Emissivity tables on the device.  Table index tb = id*ng + ig (all
// emitters of a channel adjacent: channel-contiguous).  Per table:
//   np[tb]; p[tb][ip] (float); nt[tb][ip]; t[tb][ip][it] (float);
//   ncol[tb][ip][it] (entries of the column); col[tb][ip][it] (offset of
//   the column in ue, in float2 units).  ue holds (u, eps) pairs.
struct DevTables {
  const int* np;
  const float* p;      // [ntbl][TBLNPMAX]
  const int* nt;       // [ntbl][TBLNPMAX]
  const float* t;      // [ntbl][TBLNPMAX][TBLNTMAX]
  const int* ncol;     // [ntbl][TBLNPMAX][TBLNTMAX]
  const int* col;      // [ntbl][TBLNPMAX][TBLNTMAX]
  const float2* ue;    // (u, eps) pairs
  unsigned ue_bytes;   // size of ue in bytes
  // When every channel of an emitter has the same increasing (p, T) grids
  // and every column is strictly increasing (shared != 0): the grids once
  // per emitter, and (ncol, col) packed per column.
  int shared;
  const int* gnp;      // [ng]
  const float* gp;     // [ng][TBLNPMAX]
  const int* gnt;      // [ng][TBLNPMAX]
  const float* gt;     // [ng][TBLNPMAX][TBLNTMAX]
  const int2* meta;    // [ntbl][TBLNPMAX][TBLNTMAX] = (ncol, col)
  unsigned meta_bytes;
  // the pressure and temperature grids as read (FP64), for the FP64 path
  const double* p64;   // [ntbl][TBLNPMAX]
  const double* t64;   // [ntbl][TBLNPMAX][TBLNTMAX]
};

// Per-channel constants of the continua and the channel's window (filled on
// the host from the continuum coefficient tables, forwardmodel.c:120-152).
struct ChanConst {
  float nu;
  int window;
  // CO2 (continua.c:854-869): ctw(T) = (T-260)(T-296) k230 - (T-230)(T-296) k260
  //                                  + (T-230)(T-260) k296
  int co2_on, co2_ig;
  float co2_k230, co2_k260, co2_k296;
  // H2O self/foreign (continua.c:1898-1928)
  int h2o_on, h2o_ig;
  float h2o_s296;      // sfac*cw296
  float h2o_lnratio;   // log(cw260/cw296)
  float h2o_frn;       // cwfrn*fscal
  float h2o_tanhk;     // .7193876*nu
  // N2 / O2 (continua.c:1980-1992, 2037-2048)
  int n2_on;
  float n2_b, n2_beta;
  int o2_on;
  float o2_b, o2_beta;
};

// The same channel data in FP64 for the FP64 path, as the continuum
// functions form them (continua.c:854-869, 1898-1928, 1980-1992, 2037-2048).
struct ChanConst64 {
  double nu;
  int window;
  int co2_on, co2_ig;
  double co2_cw230, co2_cw260, co2_cw296;
  int h2o_on, h2o_ig;
  double h2o_sfac, h2o_cw296, h2o_cw260, h2o_ctwfrn;
  int n2_on;
  double n2_b, n2_beta;
  int o2_on;
  double o2_b, o2_beta;
};

// ---------------------------------------------------------------------------
// Small numerics helpers with the reference's semantics
// ---------------------------------------------------------------------------

// LIN (jurassic.h:73-74)
template <typename T>
__host__ __device__ inline T lin(T x0, T y0, T x1, T y1, T x) {
  return y0 + (y1 - y0) / (x1 - x0) * (x - x0);
}

// GSL_MAX(GSL_MIN(x, 1), 0): a NaN becomes 1 (first comparison fails).
__host__ __device__ inline float clamp01(float x) {
  float a = (x < 1.f) ? x : 1.f;
  return (a > 0.f) ? a : 0.f;
}

// Bisection over a monotone array, direction from xx[mid] vs xx[mid+1]
// (misc.c:237-265).
template <typename T>
__host__ __device__ inline int locate_mono(const T* xx, int n, T x) {
  int ilo = 0, ihi = n - 1, i = (ihi + ilo) >> 1;
  if (xx[i] < xx[i + 1]) {
    while (ihi > ilo + 1) {
      i = (ihi + ilo) >> 1;
      if (xx[i] > x) ihi = i; else ilo = i;
    }
  } else {
    while (ihi > ilo + 1) {
      i = (ihi + ilo) >> 1;
      if (xx[i] <= x) ihi = i; else ilo = i;
    }
  }
  return ilo;
}

// Bisection over an increasing float2 column, key .x or .y
// (locate_tbl, forwardmodel.c:546-565).
template <bool KEY_EPS, class C>
__device__ inline int locate_col(const C& c, int n, float x) {
  int ilo = 0, ihi = n - 1;
  while (ihi > ilo + 1) {
    int i = (ihi + ilo) >> 1;
    float v = KEY_EPS ? c[i].y : c[i].x;
    if (v > x) ihi = i; else ilo = i;
  }
  return ilo;
}

// ---------------------------------------------------------------------------
// FP64 geometry (misc.c:5-16, 61-72)
// ---------------------------------------------------------------------------
__host__ __device__ inline void geo2cart(double z, double lon, double lat, double* x) {
  double radius = z + RE;
  double clat = cos(lat / 180 * M_PI);
  x[0] = radius * clat * cos(lon / 180 * M_PI);
  x[1] = radius * clat * sin(lon / 180 * M_PI);
  x[2] = radius * sin(lat / 180 * M_PI);
}
__host__ __device__ inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__host__ __device__ inline double norm3(const double* a) { return sqrt(dot3(a, a)); }
__host__ __device__ inline void cart2geo(const double* x, double* z, double* lon, double* lat) {
  double radius = norm3(x);
  *lat = asin(x[2] / radius) * 180 / M_PI;
  *lon = atan2(x[1], x[0]) * 180 / M_PI;
  *z = radius - RE;
}
// The altitude part of cart2geo alone.
__host__ __device__ inline double cart2z(const double* x) { return norm3(x) - RE; }
__host__ __device__ inline double dist3(const double* a, const double* b) {
  double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
  return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}
// ANGLE (jurassic.h:32-34)
__host__ __device__ inline double angle3(const double* a, const double* b) {
  double c = dot3(a, b) / norm3(a) / norm3(b);
  c = (c > -1) ? c : -1;
  c = (c < 1) ? c : 1;
  return acos(c);
}

// Atmosphere set on the device: profiles of nlev levels, ray-independent
// layout [a][lev], q [a][ig][lev], k [a][iw][lev] (FP64).
struct DevAtm {
  const double* z;
  const double* p;
  const double* t;
  const double* q;
  const double* k;
  int nlev, ng, nw;
};

// intpol_atm_1d (atmosphere.c:298-320): p exponential (EXP, jurassic.h:55-58),
// the rest linear.  Only the fields the caller asks for.
struct AtmPoint {
  double p, t;
};
// locate on an increasing altitude grid starting from bracket h (the
// previous evaluation of the same line of sight): the bracket
// z[h] <= z0 < z[h + 1] is unique, so this returns locate_mono's index;
// a miss after one step up or down falls back to the bisection.
__device__ inline int atm_locate_h(const DevAtm& A, int a, double z0, int h) {
  const double* z = A.z + (size_t)a * A.nlev;
  const int n = A.nlev;
  if (!(z[0] < z[1]) || !(z0 == z0)) return locate_mono<double>(z, n, z0);
  h = h < 0 ? 0 : (h > n - 2 ? n - 2 : h);
  const bool okl = (h == 0) || (z[h] <= z0), okh = (h == n - 2) || (z0 < z[h + 1]);
  if (okl && okh) return h;
  if (okl) {
    if (h + 1 == n - 2 || z0 < z[h + 2]) return h + 1;
  } else {
    if (h - 1 == 0 || z[h - 1] <= z0) return h - 1;
  }
  return locate_mono<double>(z, n, z0);
}
__device__ inline int atm_locate(const DevAtm& A, int a, double z0) {
  return locate_mono<double>(A.z + (size_t)a * A.nlev, A.nlev, z0);
}
__device__ inline AtmPoint atm_pt(const DevAtm& A, int a, double z0, int ip) {
  const double* z = A.z + (size_t)a * A.nlev;
  const double* p = A.p + (size_t)a * A.nlev;
  const double* t = A.t + (size_t)a * A.nlev;
  AtmPoint r;
  double y0 = p[ip], y1 = p[ip + 1];
  if (y0 > 0 && y1 > 0)
    r.p = y0 * exp(log(y1 / y0) / (z[ip + 1] - z[ip]) * (z0 - z[ip]));
  else
    r.p = lin(z[ip], y0, z[ip + 1], y1, z0);
  r.t = lin(z[ip], t[ip], z[ip + 1], t[ip + 1], z0);
  return r;
}
__device__ inline double atm_q(const DevAtm& A, int a, int ig, double z0, int ip) {
  const double* z = A.z + (size_t)a * A.nlev;
  const double* q = A.q + ((size_t)a * A.ng + ig) * A.nlev;
  return lin(z[ip], q[ip], z[ip + 1], q[ip + 1], z0);
}
__device__ inline double atm_k(const DevAtm& A, int a, int iw, double z0, int ip) {
  const double* z = A.z + (size_t)a * A.nlev;
  const double* k = A.k + ((size_t)a * A.nw + iw) * A.nlev;
  return lin(z[ip], k[ip], z[ip + 1], k[ip + 1], z0);
}
