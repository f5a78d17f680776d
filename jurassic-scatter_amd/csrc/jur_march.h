// jur_march.h -- the FP64 refracted ray march of one line of sight as a
// device function, shared by the ray tracer (every point stored) and the
// solar-ray search of the scattering source (only the tail kept).
// Reference: src/lineofsight.c:5-167 (raytrace), :376-381 (refractivity),
// :385-407 (intersection_point).
#pragma once
#include "jur_dev.h"
#include "jur_kernels.h"

struct G4 {
  double z, lon, lat, ds;
};

// intersection_point (lineofsight.c:385-407): the point at altitude znew on
// the chord between a (earlier) and b.
__device__ inline G4 chord_point(const G4& a, const G4& b, double znew) {
  double x1[3], x2[3], x[3];
  double frac = (a.z - znew) / (a.z - b.z);
  geo2cart(a.z, a.lon, a.lat, x1);
  geo2cart(b.z, b.lon, b.lat, x2);
  for (int i = 0; i < 3; i++) x[i] = x1[i] + frac * (x2[i] - x1[i]);
  G4 r;
  cart2geo(x, &r.z, &r.lon, &r.lat);
  r.ds = 0;
  return r;
}

__device__ inline double refractivity(const AtmPoint& a) { return 7.753e-05 * a.p / a.t; }

// gsl_stats_min / gsl_stats_max of an atmosphere's altitude grid.
__device__ inline void atm_zrange(const DevAtm& M, int a, double& zmin, double& zmax) {
  const double* az = M.z + (size_t)a * M.nlev;
  zmin = zmax = az[0];
  for (int i = 0; i < M.nlev; i++) {
    double v = az[i];
    if (v < zmin || isnan(v)) zmin = v;
    if (v > zmax || isnan(v)) zmax = v;
  }
}

struct MarchOut {
  int np;
  int status;    // JUR_ST_* bits
  double tsurf;  // temperature of the last atmosphere evaluation at the ground, else -999
};

// March from the observer towards the view point (lineofsight.c:19-167).
// Calls sink.put(index, point) for every LOS point; stops at the ground or
// the top of the atmosphere, or when `cap` points are stored (JUR_ST_CAP).
template <class Sink>
__device__ MarchOut march_ray(const DevAtm& M, int a, double obsz, double obslon, double obslat,
                              double vpz, double vplon, double vplat, double rayds, double raydz,
                              int refrac, int cap, Sink& sink) {
  MarchOut R = {0, 0, -999};
  double zmin, zmax;
  atm_zrange(M, a, zmin, zmax);
  if (obsz < zmin) {
    R.status |= JUR_ST_BELOW_SURFACE;
    return R;
  }
  if (vpz > zmax - 0.001) return R;
  double xobs[3], xvp[3], ex0[3], ex1[3], x[3], xh[3], ng[3];
  geo2cart(obsz, obslon, obslat, xobs);
  geo2cart(vpz, vplon, vplat, xvp);
  for (int i = 0; i < 3; i++) ex0[i] = xvp[i] - xobs[i];
  double norm = norm3(ex0);
  for (int i = 0; i < 3; i++) ex0[i] /= norm;
  for (int i = 0; i < 3; i++) x[i] = xobs[i];
  double z = 1e99, lon, lat;
  // Entry point from above the atmosphere (lineofsight.c:53-67).
  if (obsz > zmax) {
    double dmin = 0, dmax = norm;
    while (fabs(dmin - dmax) > 0.001) {
      double d = (dmax + dmin) / 2;
      for (int i = 0; i < 3; i++) x[i] = xobs[i] + d * ex0[i];
      cart2geo(x, &z, &lon, &lat);
      if (z <= zmax && z > zmax - 0.001) break;
      if (z < zmax - 0.0005) dmax = d; else dmin = d;
    }
  }
  // (pl, tl): last atmosphere evaluation, as the reference's locals p, t.
  double pl = 0, tl = 0;
  int ih = 0;  // altitude bracket of the previous evaluation (search hint)
  G4 prev = {0, 0, 0, 0};
  int stop = 0, np = 0;
  while (1) {
    double ds = rayds;
    if (raydz > 0) {
      double nx = norm3(x);
      for (int i = 0; i < 3; i++) xh[i] = x[i] / nx;
      double cosa = fabs(dot3(ex0, xh));
      if (cosa != 0) ds = (rayds < raydz / cosa) ? rayds : raydz / cosa;
    }
    cart2geo(x, &z, &lon, &lat);
    G4 cur = {z, lon, lat, ds};
    if (z < zmin + 0.001 || z > zmax + 0.001) {
      stop = (z < zmin + 0.001 ? 2 : 1);
      // With no previous point the reference reads z[-1]; the boundary
      // point is kept as found here.
      if (np > 0) cur = chord_point(prev, cur, (stop == 2 ? zmin : zmax));
      cur.ds = 0.;
    } else if (M.geo.ip >= 2) {
      AtmPoint ap = geo_pt(M, a, geo_locate(M, a, z, lon, lat));
      pl = ap.p;
      tl = ap.t;
    } else {
      ih = atm_locate_h(M, a, z, ih);
      AtmPoint ap = atm_pt(M, a, z, ih);
      pl = ap.p;
      tl = ap.t;
    }
    if (np >= cap) {
      R.status |= JUR_ST_CAP;
      break;
    }
    sink.put(np, cur);
    prev = cur;
    if ((++np) > NLOS) {
      R.status |= JUR_ST_NLOS;
      break;
    }
    if (stop) {
      R.tsurf = (stop == 2 ? tl : -999);
      break;
    }
    double n = (refrac && z <= 60) ? 1 + refractivity({pl, tl}) : 1;
    for (int i = 0; i < 3; i++) ex1[i] = ex0[i] * n;
    if (refrac && z <= 60 && M.geo.ip >= 2) {
      // IP 2 / 3: the gradient points' own positions (lineofsight.c:123-146)
      auto at = [&](const double* xp) {
        double zz, lo, la;
        cart2geo(xp, &zz, &lo, &la);
        return geo_pt(M, a, geo_locate(M, a, zz, lo, la));
      };
      for (int i = 0; i < 3; i++) xh[i] = x[i] + 0.5 * ds * ex0[i];
      AtmPoint ap = at(xh);
      double n0 = refractivity(ap);
      for (int i = 0; i < 3; i++) {
        xh[i] += 0.02;
        ap = at(xh);
        ng[i] = (refractivity(ap) - n0) / 0.02;
        xh[i] -= 0.02;
      }
      pl = ap.p;
      tl = ap.t;
    } else if (refrac && z <= 60) {
      // the refractivity gradient needs only the altitudes of the points
      for (int i = 0; i < 3; i++) xh[i] = x[i] + 0.5 * ds * ex0[i];
      double zz = cart2z(xh);
      int jh = atm_locate_h(M, a, zz, ih);
      AtmPoint ap = atm_pt(M, a, zz, jh);
      double n0 = refractivity(ap);
      for (int i = 0; i < 3; i++) {
        xh[i] += 0.02;
        zz = cart2z(xh);
        ap = atm_pt(M, a, zz, atm_locate_h(M, a, zz, jh));
        ng[i] = (refractivity(ap) - n0) / 0.02;
        xh[i] -= 0.02;
      }
      pl = ap.p;
      tl = ap.t;
    } else {
      ng[0] = ng[1] = ng[2] = 0;
    }
    for (int i = 0; i < 3; i++) ex1[i] += ds * ng[i];
    double n1 = norm3(ex1);
    for (int i = 0; i < 3; i++) ex1[i] /= n1;
    for (int i = 0; i < 3; i++) x[i] += 0.5 * ds * (ex0[i] + ex1[i]);
    for (int i = 0; i < 3; i++) ex0[i] = ex1[i];
  }
  R.np = np;
  return R;
}

// Layer-edge crossings between two consecutive LOS points in the order
// add_aerosol_layers inserts them (lineofsight.c:263-277); calls
// emit(point) for each.  alti is sorted descending.
template <class Emit>
__device__ inline void edge_points(const G4& g0, const G4& g1, const double* alti, int jl,
                                   double altimax, double altimin, Emit& emit) {
  if ((g0.z < altimax || g1.z < altimax) && (g0.z > altimin || g1.z > altimin)) {
    for (int il = 0; il < jl; il++) {
      if (g0.z > alti[il] && g1.z < alti[il]) emit(chord_point(g0, g1, alti[il]));
      double ab = alti[jl - il - 1];
      if (g0.z < ab && g1.z > ab) emit(chord_point(g0, g1, ab));
    }
  }
}
