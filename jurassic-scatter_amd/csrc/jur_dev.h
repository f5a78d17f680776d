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
  int gnp_max, gnt_max;  // largest np and nt of the shared grids
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
// IP 2 and 3 (atmosphere.c:324-483): the profile is a set of points with
// geographic positions.  Surface positions are built on the host; the
// value arrays of DevAtm hold the points ([a][np] per copy a).  ip <= 1
// (the zero default): a single vertical profile.
struct DevGeo {
  int ip;
  int nx;              // IP 2: profiles (runs of equal lon, lat)
  const int* pidx;     // IP 2: first point of each profile
  const int* pnz;      // IP 2: points of each profile
  const double* px;    // IP 2: [nx][3]; IP 3: [np][3] surface positions
  const double* lat;   // IP 3: latitude of every point
  double cz, cx;       // IP 3: vertical / horizontal influence ranges [km]
};

struct DevAtm {
  const double* z;
  const double* p;
  const double* t;
  const double* q;
  const double* k;
  int nlev, ng, nw;
  DevGeo geo;
  // optional: log(p[j + 1] / p[j]) / (z[j + 1] - z[j]) of every interval
  // (k_log_slopes; the same operations as the interpolation's), [a][lev]
  const double* lps = nullptr;
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
// locate on an increasing grid from a guess by the grid's mean spacing:
// the guessed bracket is checked and one step tried either way before the
// bisection, so near-uniform grids need one or two loads instead of
// log2(n); the bracket z[h] <= z0 < z[h + 1] is unique, so this returns
// locate_mono's index.
__device__ inline int locate_guess(const double* z, int n, double z0) {
  if (n < 3 || !(z[0] < z[1]) || !(z[0] < z[n - 1]) || !(z0 == z0)) return locate_mono<double>(z, n, z0);
  const double g = (z0 - z[0]) / (z[n - 1] - z[0]) * (n - 1);
  int h = g < 0 ? 0 : (g >= n - 2 ? n - 2 : (int)g);
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
    r.p = y0 * exp((A.lps ? A.lps[(size_t)a * A.nlev + ip] : log(y1 / y0) / (z[ip + 1] - z[ip])) * (z0 - z[ip]));
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

// Where one evaluation at (z0, lon0, lat0) reads the point set
// (intpol_atm_geo, atmosphere.c:269-294).  IP 1 and 2: brackets j0 (and j1)
// of the one or two profiles, weight r of the second (intpol_atm_2d,
// :324-418: the two profiles nearest at the surface, the point projected
// onto the chord between them).  IP 3: the point itself (:422-483).
struct GeoLoc {
  int j0, j1;
  double r;
  double z0, lat0, x0[3];
};
__device__ inline double dist2_3(const double* a, const double* b) {
  double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
  return d0 * d0 + d1 * d1 + d2 * d2;
}
__device__ inline GeoLoc geo_locate(const DevAtm& A, int a, double z0, double lon0, double lat0) {
  GeoLoc L;
  L.j0 = L.j1 = 0;
  L.r = 0;
  L.z0 = z0;
  L.lat0 = lat0;
  if (A.geo.ip >= 2) geo2cart(0, lon0, lat0, L.x0);
  const double* z = A.z + (size_t)a * A.nlev;
  if (A.geo.ip <= 1) {
    L.j0 = L.j1 = locate_guess(z, A.nlev, z0);
  } else if (A.geo.ip == 2) {
    double d0 = 1e99, d1 = 1e99;
    int i0 = 0, i1 = 0;
    for (int ix = 0; ix < A.geo.nx; ix++) {
      const double dh = dist2_3(L.x0, A.geo.px + 3 * ix);
      if (dh <= d0) {
        d1 = d0;
        i1 = i0;
        d0 = dh;
        i0 = ix;
      } else if (dh <= d1) {
        d1 = dh;
        i1 = ix;
      }
    }
    const int b0 = A.geo.pidx[i0], b1 = A.geo.pidx[i1];
    L.j0 = b0 + locate_guess(z + b0, A.geo.pnz[i0], z0);
    L.j1 = b1 + locate_guess(z + b1, A.geo.pnz[i1], z0);
    const double x2 = dist2_3(A.geo.px + 3 * i0, A.geo.px + 3 * i1), x = sqrt(x2);
    const double r0 = (d0 - d1 + x2) / (2 * x), r1 = x - r0;
    L.r = r0 <= 0 ? 0.0 : (r1 <= 0 ? 1.0 : r0 / (r0 + r1));
  }
  return L;
}
// A field of copy a (f points at its first value) at the located point:
// exponential in altitude for the pressure (EXP, jurassic.h:55-58), linear
// otherwise.
// lp: the copy's precomputed log-slopes (pressure only; may be null).
template <bool EXPO>
__device__ inline double geo_1d(const double* z, const double* f, int j, double z0, const double* lp) {
  const double y0 = f[j], y1 = f[j + 1];
  if (EXPO && y0 > 0 && y1 > 0) return y0 * exp((lp ? lp[j] : log(y1 / y0) / (z[j + 1] - z[j])) * (z0 - z[j]));
  return lin(z[j], y0, z[j + 1], y1, z0);
}
template <bool EXPO>
__device__ inline double geo_field(const DevAtm& A, int a, const GeoLoc& L, const double* f) {
  const double* z = A.z + (size_t)a * A.nlev;
  const double* lp = EXPO && A.lps ? A.lps + (size_t)a * A.nlev : nullptr;
  if (A.geo.ip <= 1) return geo_1d<EXPO>(z, f, L.j0, L.z0, lp);
  if (A.geo.ip == 2)
    return (1 - L.r) * geo_1d<EXPO>(z, f, L.j0, L.z0, lp) + L.r * geo_1d<EXPO>(z, f, L.j1, L.z0, lp);
  // IP 3: weights (1 - dz/CZ)(1 - dh/CX) of the points within range, in
  // the reference's point order
  double v = 0, wsum = 0;
  const double cz = A.geo.cz, cx = A.geo.cx;
  for (int ip = 0; ip < A.nlev; ip++) {
    const double dz = fabs(z[ip] - L.z0);
    if (dz > cz) continue;
    if (fabs(A.geo.lat[ip] - L.lat0) * 111.13 > cx) continue;
    const double dx2 = dist2_3(L.x0, A.geo.px + 3 * ip);
    if (dx2 > cx * cx) continue;
    const double w = (1 - dz / cz) * (1 - sqrt(dx2) / cx);
    wsum += w;
    v += w * f[ip];
  }
  return v / wsum;
}
__device__ inline AtmPoint geo_pt(const DevAtm& A, int a, const GeoLoc& L) {
  AtmPoint r;
  r.p = geo_field<true>(A, a, L, A.p + (size_t)a * A.nlev);
  r.t = geo_field<false>(A, a, L, A.t + (size_t)a * A.nlev);
  return r;
}
__device__ inline double geo_q(const DevAtm& A, int a, int ig, const GeoLoc& L) {
  return geo_field<false>(A, a, L, A.q + ((size_t)a * A.ng + ig) * A.nlev);
}
__device__ inline double geo_k(const DevAtm& A, int a, int iw, const GeoLoc& L) {
  return geo_field<false>(A, a, L, A.k + ((size_t)a * A.nw + iw) * A.nlev);
}
