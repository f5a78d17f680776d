// jur_scatter_dev.h -- data blocks and launchers of kernels_scatter.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "jur_dev.h"

// Secondary pencils per scattering point: the reference's nalpha = 28
// (scatter.c:680); its Fibonacci grid construction (scatter.c:698-715)
// gives increasing angles in [0, 180] deg for every even count from 10 to 28;
// 30 and 32 take this project's grid (alpha_grid in scatter.cpp) (option
// "fan").
#define NALPHA_REF 28
#define NALPHA 32
#define NALPHA_MIN 10

// One scattering point: Cartesian position, LOS tangent, layer, atmosphere
// and time of its pencil.
struct ScaPt {
  double x[3];
  double dx[3];
  double time;
  int il;
  int atm;
  float af;  // aerosol factor of the point
  int pad;
};

// Ray geometry arrays of a ray batch (device).
struct RayArrays {
  double *time, *obsz, *obslon, *obslat, *vpz, *vplon, *vplat;
  int* atm;
};

struct SunArgs {
  DevAtm atm;
  double rayds, raydz;
  int refrac;
  const double* alti;  // layer-edge altitudes (descending)
  int nalti;
  double altimax, altimin;
  int* status;
};

struct AngArgs {
  int npt, nd;
  int nalpha;                // pencils per fan
  const ScaPt* pts;
  const double* alpha;       // [nalpha] fan zenith angles
  const double* phase;       // [NLMAX][nd][NTHETA] layer phase functions (FP64)
  const float* rad;          // [nd][nchild] radiances of the fans (child 28 k + j)
  int nchild;
  const int* sun_idx;        // [npt] solar pencil of the point or -1
  const double* sun_theta;   // [npt] scattering angle of the solar beam
  const float* tau_sun;      // [nd][nsun] solar pencil transmittances
  int nsun;
  const double* theta_grid;  // [NTHETA] pi i / 180
  const double* bsun;        // [nd] Planck radiance of the Sun (TSUN)
  float* J;                  // [npt][nd]
  const int* d_npt;          // optional device-side count of the points (npt: capacity)
  // surrogate: sub_idx[k] >= 0 is the point's fan block in rad (plain
  // angular sum); -1 takes the thermal source from the network output
  // ymlp[k][nd] (normalised ln(J / B(T)), J = exp(y ysd + ymu + lnb))
  // instead.  Null: every point is plain.
  const int* sub_idx;
  const float* ymlp;
  const float* lnb;          // [npt][nd] ln B(T) at the point
  float ymu, ysd;
  float* jth;                // [npt][nd] thermal part (no sun), or null
};

void launch_count_points(int nray, const int* np, const float* rec, int R, int cap, int* count, hipStream_t s);
// Points at pts[jbase[r] ...] in LOS order; an index at or beyond pts_cap
// raises JUR_ST_PLAN in *status instead of writing.
void launch_emit_points(int nray, const int* np, const float* rec, int R, int cap, const double4* geo,
                        const double4* geo2, const int* geo_final, const int* jbase, const double* time,
                        const int* atm_idx, ScaPt* pts, long pts_cap, int* status, hipStream_t s);
// npt: the points' capacity (grid); d_npt (optional): their device-side
// count, and d_nchild (optional) receives the fan rays' count.
void launch_children(int npt, int nalpha, const ScaPt* pts, const double* alpha, RayArrays out, hipStream_t s,
                     const int* d_npt = nullptr, int* d_nchild = nullptr);
void launch_sun(int npt, const ScaPt* pts, const SunArgs& S, int* valid, double* theta, RayArrays out,
                hipStream_t s, const int* d_npt = nullptr);
// Exclusive prefix sum with a device base (graph-capture safe): out[i] =
// *base + sum_{j<i} in[j], *total = *base + sum (base may be null); tmp holds
// scan_base_temp_ints(n) ints.
size_t scan_base_temp_ints(int n);
void launch_scan_base(const int* in, int n, int* out, int* tmp, const int* base, int* total, hipStream_t s);
void launch_angular(const AngArgs& G, hipStream_t s);
// srcfunc_sca_3d (scatter.c:811-895), the method for IP 2 / 3: a direct
// fan of DIR3D_NT x DIR3D_NPHI pencils per point (children k DIR3D_N +
// itheta DIR3D_NPHI + iphi) and its weighted sum (thermal source) plus the
// solar term.  G.nalpha is not used.
#define DIR3D_NT 180
#define DIR3D_NPHI 180
#define DIR3D_N (DIR3D_NT * DIR3D_NPHI)
void launch_children3d(int npt, const ScaPt* pts, RayArrays out, hipStream_t s);
void launch_angular3d(const AngArgs& G, hipStream_t s);
void launch_sun_compact(int npt, const int* valid, const int* pos, int* sun_idx, hipStream_t s);
void launch_gather_rays(int npt, const int* valid, const int* pos, RayArrays in, RayArrays out, hipStream_t s);
size_t scan_temp_bytes(int n);
void launch_exclusive_scan(const int* in, int* out, void* tmp, size_t bytes, int n, hipStream_t s);
