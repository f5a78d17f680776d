// kernels_scatter.hip -- device side of the scattering source
// (reference src/scatter.c:643-1041, forwardmodel.c:316-360).
//
// The reference recurses: each LOS point inside an aerosol layer traces 28
// secondary pencils (srcfunc_sca_1d) and a solar pencil (srcfunc_sca_sun),
// and its CPU work queue only defers the leaf pencils.  Here the recursion
// is unrolled into levels of ray batches: level L+1 holds the fans of all
// scattering points of level L (k_children), generated from the traced
// level-L rays (k_count_points, k_emit_points); solar rays are searched per
// point (k_sun); once the level below is integrated, k_angular turns its
// radiances into the source J of every point.
#include <hipcub/hipcub.hpp>

#include "jur_march.h"
#include "jur_scatter_dev.h"

namespace {

// bascoord (scatter.c:6-52)
__device__ void bascoord(const double* dz, const double* dy, double* ex, double* ey, double* ez) {
  double norm = norm3(dz);
  for (int i = 0; i < 3; i++) ez[i] = dz[i] / norm;
  norm = norm3(dy);
  for (int i = 0; i < 3; i++) ey[i] = dy[i] / norm;
  double dotp = dot3(ey, ez);
  if (dotp == 1) {
    ey[0] = 1; ey[1] = 0; ey[2] = 0;
    dotp = ez[0];
    if (dotp == 1) { ey[0] = 0; ey[1] = 1; dotp = ez[1]; }
  }
  for (int i = 0; i < 3; i++) ey[i] -= dotp * ez[i];
  norm = norm3(ey);
  for (int i = 0; i < 3; i++) ey[i] /= norm;
  ex[0] = ey[1] * ez[2] - ey[2] * ez[1];
  ex[1] = ey[2] * ez[0] - ey[0] * ez[2];
  ex[2] = ey[0] * ez[1] - ey[1] * ez[0];
}

// Local frame at x: lz up, ly towards north (scatter.c:692-696, 925-929).
__device__ void local_frame(const double* x, double* lx, double* ly, double* lz) {
  double dn[3] = {-x[0], -x[1], 2 * RE - x[2]};
  bascoord(x, dn, lx, ly, lz);
}

// suncoord (scatter.c:999-1041)
__device__ void suncoord(double sec, double lon, double lat, double* azi, double* sza) {
  double D = sec / 86400 - 0.5;
  double g = (357.529 + 0.98560028 * D) * M_PI / 180;
  double q = 280.459 + 0.98564736 * D;
  double L = (q + 1.915 * sin(g) + 0.020 * sin(2 * g)) * M_PI / 180;
  double e = (23.439 - 0.00000036 * D) * M_PI / 180;
  double dec = asin(sin(e) * sin(L));
  double ra = atan2(cos(e) * sin(L), cos(L));
  double GMST = 18.697374558 + 24.06570982441908 * D;
  double LST = GMST + lon / 15;
  double h = LST / 12 * M_PI - ra;
  lat *= M_PI / 180;
  *azi = atan2(-sin(h), cos(lat) * tan(dec) - sin(lat) * cos(h)) * 180 / M_PI;
  *sza = acos(sin(lat) * sin(dec) + cos(lat) * cos(dec) * cos(h)) * 180 / M_PI;
}

// The last three march points (enough for the last-segment rule and the
// final two points after layer-edge insertion).
struct TailSink {
  G4 t[3];
  __device__ void put(int i, const G4& g) {
    t[0] = t[1];
    t[1] = t[2];
    t[2] = g;
    (void)i;
  }
};
struct LastEmit {
  G4 last;
  int n = 0;
  __device__ void operator()(const G4& g) {
    last = g;
    n++;
  }
};

}  // namespace

// Points with aerosol (aerofac > 0) per traced ray.
__global__ void k_count_points(int nray, const int* np, const float* rec, int R, int cap, int* count) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nray) return;
  const float* pr = rec + (size_t)r * cap * R;
  int c = 0;
  for (int ip = 0; ip < np[r]; ip++) c += pr[(size_t)ip * R + REC_AF] > 0.f;
  count[r] = c;
}

// Scattering points of each ray in LOS order at jbase[r]: position x and
// LOS tangent dx = x(ip+1) - x(ip-1) (forwardmodel.c:320-327; the point
// after the last is the last point itself).
__global__ void k_emit_points(int nray, const int* np, const float* rec, int R, int cap, const double4* geo,
                              const double4* geo2, const int* geo_final, const int* jbase, const double* time,
                              const int* atm_idx, ScaPt* pts, long pts_cap, int* status) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nray) return;
  const float* pr = rec + (size_t)r * cap * R;
  const double4* G = (geo_final[r] ? geo2 : geo) + (size_t)r * cap;
  const int n = np[r];
  int k = jbase[r];
  for (int ip = 0; ip < n; ip++) {
    if (!(pr[(size_t)ip * R + REC_AF] > 0.f)) continue;
    int ip0 = ip > 0 ? ip - 1 : ip, ip1 = ip < n - 1 ? ip + 1 : ip;
    double x[3], x0[3], x1[3];
    geo2cart(G[ip].x, G[ip].y, G[ip].z, x);
    geo2cart(G[ip0].x, G[ip0].y, G[ip0].z, x0);
    geo2cart(G[ip1].x, G[ip1].y, G[ip1].z, x1);
    ScaPt P;
    for (int i = 0; i < 3; i++) {
      P.x[i] = x[i];
      P.dx[i] = x1[i] - x0[i];
    }
    P.il = (int)pr[(size_t)ip * R + REC_AI];
    P.atm = atm_idx ? atm_idx[r] : 0;
    P.time = time[r];
    P.af = pr[(size_t)ip * R + REC_AF];
    P.pad = 0;
    if (k >= pts_cap) {
      atomicOr(status, (int)JUR_ST_PLAN);
      return;
    }
    pts[k++] = P;
  }
}

// The fan of nalpha secondary pencils of each point (scatter.c:717-740):
// view points 10 km away at zenith angles alpha in the local vertical-north
// plane.
__global__ void k_children(int npt, int nalpha, const ScaPt* pts, const double* alpha, RayArrays out,
                           const int* d_npt, int* d_nchild) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  // npt: the points' capacity; d_npt (optional): their device-side count
  const int nv = d_npt ? min(*d_npt, npt) : npt;
  if (d_nchild && i == 0) *d_nchild = nv * nalpha;
  if (i >= (long)nv * nalpha) return;
  int k = (int)(i / nalpha), ir = (int)(i % nalpha);
  const ScaPt P = pts[k];
  double lx[3], ly[3], lz[3], xv[3];
  local_frame(P.x, lx, ly, lz);
  for (int j = 0; j < 3; j++) xv[j] = P.x[j] + 10. * cos(alpha[ir]) * (-1) * lz[j] + 10. * sin(alpha[ir]) * ly[j];
  double z, lon, lat;
  cart2geo(P.x, &z, &lon, &lat);
  out.obsz[i] = z;
  out.obslon[i] = lon;
  out.obslat[i] = lat;
  cart2geo(xv, &z, &lon, &lat);
  out.vpz[i] = z;
  out.vplon[i] = lon;
  out.vplat[i] = lat;
  out.time[i] = P.time;
  out.atm[i] = P.atm;
}

// Solar ray search of srcfunc_sca_sun (scatter.c:931-967): up to 10 refracted
// traces towards the Sun, correcting the geometric zenith angle by the
// golden ratio of the beam's end-point mismatch.  The sun contributes when
// the final trace does not hit the ground.
__global__ void k_sun(int npt, const ScaPt* pts, SunArgs S, int* valid, double* theta, RayArrays out,
                      const int* d_npt) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npt) return;
  if (d_npt && k >= *d_npt) {
    valid[k] = 0;
    return;
  }
  const ScaPt P = pts[k];
  double lx[3], ly[3], lz[3], ek[3] = {0, 0, 0}, x1[3], x0[3], dout[3];
  local_frame(P.x, lx, ly, lz);
  double oz, olon, olat, azi, sza;
  cart2geo(P.x, &oz, &olon, &olat);
  suncoord(P.time, olon, olat, &azi, &sza);
  double sza_cor = sza, vz = 0, vlon = 0, vlat = 0;
  double tsurf = -999;
  int status = 0;
  for (int i2 = 0; i2 < 10; i2++) {
    double s = sin(sza_cor * M_PI / 180), c = cos(sza_cor * M_PI / 180);
    double sa = sin(azi * M_PI / 180), ca = cos(azi * M_PI / 180);
    for (int i = 0; i < 3; i++) {
      ek[i] = s * sa * lx[i] + s * ca * ly[i] + c * lz[i];
      x1[i] = P.x[i] + 10 * ek[i];
    }
    cart2geo(x1, &vz, &vlon, &vlat);
    TailSink tail;
    MarchOut mo = march_ray(S.atm, P.atm, oz, olon, olat, vz, vlon, vlat, S.rayds, S.raydz, S.refrac,
                            NLOS + 1, tail);
    status |= mo.status;
    tsurf = mo.tsurf;
    int np = mo.np;
    if (status) break;
    // last-segment rule (lineofsight.c:170-171) on the raw points
    G4 a = tail.t[1], b = tail.t[2];
    if (np >= 2 && tail.t[1].ds < 1e-3 && np - 1 > 1) {
      np--;
      a = tail.t[0];
      b = tail.t[1];
    }
    if (np < 2) break;
    // with aerosol layers, an edge point may end up between the last two
    LastEmit le;
    if (S.nalti > 0) edge_points(a, b, S.alti, S.nalti, S.altimax, S.altimin, le);
    if (le.n > 0) a = le.last;
    geo2cart(a.z, a.lon, a.lat, x0);
    geo2cart(b.z, b.lon, b.lat, x1);
    for (int i = 0; i < 3; i++) dout[i] = x1[i] - x0[i];
    double sza_beam = angle3(P.x, dout) * 180 / M_PI;
    if (fabs(sza_beam - sza) < 0.01) break;
    sza_cor -= 0.61803 * (sza_beam - sza);
    sza_cor = sza_cor > 0 ? sza_cor : 0;
    sza_cor = sza_cor < 180 ? sza_cor : 180;
  }
  if (status) atomicOr(S.status, status);
  int ok = (tsurf < 0) && !status;
  valid[k] = ok;
  theta[k] = angle3(ek, P.dx);
  out.obsz[k] = oz;
  out.obslon[k] = olon;
  out.obslat[k] = olat;
  out.vpz[k] = vz;
  out.vplon[k] = vlon;
  out.vplat[k] = vlat;
  out.time[k] = P.time;
  out.atm[k] = P.atm;
}

// Scattering source J of each point (scatter.c:743-803, 969-990): the
// reference's 180 x 180 (theta, phi) sum of sin(theta) P(theta) rad(alpha)
// normalised by the sum of the weights.  The alpha-interpolation weights do
// not depend on the channel, so per point the kernel first accumulates
// A[theta][k] = sum over phi of the weight of fan ray k, then
// J = sum_theta w(theta) sum_k A[theta][k] rad_k / (180 sum_theta w(theta)).
template <int NA>
__global__ void __launch_bounds__(256) k_angular(AngArgs G) {
  __shared__ double Aw[180][NA];
  __shared__ double alpha[NA];
  __shared__ double sphi[180], cphi[180];  // sin and cos of the 180 azimuths
  const int k = blockIdx.x;
  if (G.d_npt && k >= *G.d_npt) return;  // beyond the device-side count of the points
  const ScaPt P = G.pts[k];
  const int sk = G.sub_idx ? G.sub_idx[k] : k;
  if (sk < 0) {
    // surrogate point: thermal source from the network, solar term as below
    const double* ph = G.phase + (size_t)P.il * G.nd * NTHETA;
    for (int id = threadIdx.x; id < G.nd; id += blockDim.x) {
      const size_t o = (size_t)k * G.nd + id;
      double src = exp((double)G.ymlp[o] * G.ysd + G.ymu + (double)G.lnb[o]);
      if (G.jth) G.jth[(size_t)k * G.nd + id] = (float)src;
      const int si = G.sun_idx[k];
      if (si >= 0) {
        const double* p = ph + (size_t)id * NTHETA;
        const double t2 = G.sun_theta[k];
        int idx = locate_mono<double>(G.theta_grid, NTHETA, t2);
        double p2 = lin(G.theta_grid[idx], p[idx], G.theta_grid[idx + 1], p[idx + 1], t2);
        src += 6.764e-5 * p2 * G.bsun[id] * (double)G.tau_sun[(size_t)id * G.nsun + si];
      }
      G.J[(size_t)k * G.nd + id] = (float)src;
    }
    return;
  }
  for (int i = threadIdx.x; i < 180 * NA; i += blockDim.x) (&Aw[0][0])[i] = 0.0;
  if (threadIdx.x < NA) alpha[threadIdx.x] = G.alpha[threadIdx.x];
  for (int i = threadIdx.x; i < 180; i += blockDim.x) {
    const double phi = 2. * (0.5 + i) * M_PI / 180.;
    sphi[i] = sin(phi);
    cphi[i] = cos(phi);
  }
  __syncthreads();
  double lx[3], ly[3], lz[3], sx[3], sy[3], sz[3];
  local_frame(P.x, lx, ly, lz);
  bascoord(P.dx, P.x, sx, sy, sz);
  const double mlz[3] = {-lz[0], -lz[1], -lz[2]};
  const double nmlz = norm3(mlz);
  for (int it = threadIdx.x; it < 180; it += blockDim.x) {
    const double th = (0.5 + it) * M_PI / 180.;
    const double st = sin(th), ct = cos(th);
    for (int iphi = 0; iphi < 180; iphi++) {
      double ek[3];
      const double sp = sphi[iphi], cp = cphi[iphi];
      for (int i = 0; i < 3; i++) ek[i] = st * sp * sx[i] + st * cp * sy[i] + ct * sz[i];
      // angle3(mlz, ek) with the fan's constant norm hoisted
      double c = dot3(mlz, ek) / nmlz / norm3(ek);
      c = (c > -1) ? c : -1;
      c = (c < 1) ? c : 1;
      const double a2 = acos(c);
      const int idx = locate_mono<double>(alpha, NA, a2);
      const double f = (a2 - alpha[idx]) / (alpha[idx + 1] - alpha[idx]);
      Aw[it][idx] += 1.0 - f;
      Aw[it][idx + 1] += f;
    }
  }
  __syncthreads();
  const size_t c0 = (size_t)sk * NA;
  const double* ph = G.phase + (size_t)P.il * G.nd * NTHETA;
  for (int id = threadIdx.x; id < G.nd; id += blockDim.x) {
    double r[NA];
    for (int j = 0; j < NA; j++) r[j] = G.rad[(size_t)id * G.nchild + c0 + j];
    const double* p = ph + (size_t)id * NTHETA;
    double num = 0, den = 0;
    for (int it = 0; it < 180; it++) {
      const double th = (0.5 + it) * M_PI / 180.;
      // phase function on the 1 deg grid theta_i = i pi / 180 (idp = it)
      const double t0 = (double)it * M_PI / 180., t1 = (double)(it + 1) * M_PI / 180.;
      const double w = sin(th) * lin(t0, p[it], t1, p[it + 1], th);
      double s = 0;
      for (int j = 0; j < NA; j++) s += Aw[it][j] * r[j];
      num += w * s;
      den += 180.0 * w;
    }
    double src = num / den;
    if (G.jth) G.jth[(size_t)k * G.nd + id] = (float)src;
    const int si = G.sun_idx[k];
    if (si >= 0) {
      const double t2 = G.sun_theta[k];
      // scatter.c:921-923: theta_i = pi i / 180
      int idx = locate_mono<double>(G.theta_grid, NTHETA, t2);
      double p2 = lin(G.theta_grid[idx], p[idx], G.theta_grid[idx + 1], p[idx + 1], t2);
      src += 6.764e-5 * p2 * G.bsun[id] * (double)G.tau_sun[(size_t)id * G.nsun + si];
    }
    G.J[(size_t)k * G.nd + id] = (float)src;
  }
}

__global__ void k_sun_compact(int npt, const int* valid, const int* pos, int* sun_idx) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npt) return;
  sun_idx[k] = valid[k] ? pos[k] : -1;
}

__global__ void k_gather_rays(int npt, const int* valid, const int* pos, RayArrays in, RayArrays out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npt || !valid[k]) return;
  int j = pos[k];
  out.time[j] = in.time[k];
  out.obsz[j] = in.obsz[k];
  out.obslon[j] = in.obslon[k];
  out.obslat[j] = in.obslat[k];
  out.vpz[j] = in.vpz[k];
  out.vplon[j] = in.vplon[k];
  out.vplat[j] = in.vplat[k];
  out.atm[j] = in.atm[k];
}

// ---- srcfunc_sca_3d (scatter.c:811-895) ---------------------------------------
// The direct fan of each point: view points 10 km away in the directions of
// polar angle theta2 = (0.5 + itheta) pi / 180 from the LOS tangent and
// azimuth phi = (0.5 + iphi) 2 pi / 180 (basis of bascoord(dx, x)).
__global__ void k_children3d(int npt, const ScaPt* pts, RayArrays out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)npt * DIR3D_N) return;
  const int k = (int)(i / DIR3D_N), c = (int)(i % DIR3D_N);
  const int itheta = c / DIR3D_NPHI, iphi = c % DIR3D_NPHI;
  const ScaPt P = pts[k];
  double sx[3], sy[3], sz[3], xv[3];
  bascoord(P.dx, P.x, sx, sy, sz);
  const double theta2 = (0.5 + itheta) / DIR3D_NT * M_PI, phi = (0.5 + iphi) / DIR3D_NPHI * 2 * M_PI;
  for (int j = 0; j < 3; j++)
    xv[j] = P.x[j] + 10 * sin(theta2) * sin(phi) * sx[j] + 10 * sin(theta2) * cos(phi) * sy[j] +
            10 * cos(theta2) * sz[j];
  double z, lon, lat;
  cart2geo(P.x, &z, &lon, &lat);
  out.obsz[i] = z;
  out.obslon[i] = lon;
  out.obslat[i] = lat;
  cart2geo(xv, &z, &lon, &lat);
  out.vpz[i] = z;
  out.vplon[i] = lon;
  out.vplat[i] = lat;
  out.time[i] = P.time;
  out.atm[i] = P.atm;
}

// Weight of polar angle bin it of one channel: solid angle times the phase
// function interpolated at theta2 (scatter.c:869-879).
__device__ inline double w3d(const double* theta_grid, const double* p, int it) {
  const double theta2 = (0.5 + it) / DIR3D_NT * M_PI;
  const int idx = locate_mono<double>(theta_grid, NTHETA, theta2);
  const double phase2 = lin(theta_grid[idx], p[idx], theta_grid[idx + 1], p[idx + 1], theta2);
  return M_PI / DIR3D_NT * 2 * M_PI * sin(theta2) / DIR3D_NPHI * phase2;
}
__device__ inline double block_sum256(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}
// One block per (point, channel): J = sum_k w_k rad_k / W, W the sum of the
// weights over the fan AND over all channels (scatter.c:883, 891, as the
// reference normalises), then the solar term as for srcfunc_sca_1d.
__global__ void __launch_bounds__(256) k_angular3d(AngArgs G) {
  __shared__ double wl[DIR3D_NT];
  __shared__ double red[4];
  const int k = blockIdx.x, id = blockIdx.y;
  const ScaPt P = G.pts[k];
  const double* ph = G.phase + (size_t)P.il * G.nd * NTHETA;
  const double* p = ph + (size_t)id * NTHETA;
  for (int it = threadIdx.x; it < DIR3D_NT; it += blockDim.x) wl[it] = w3d(G.theta_grid, p, it);
  double wpart = 0;
  for (int j = threadIdx.x; j < G.nd * DIR3D_NT; j += blockDim.x)
    wpart += w3d(G.theta_grid, ph + (size_t)(j / DIR3D_NT) * NTHETA, j % DIR3D_NT);
  __syncthreads();
  const float* r = G.rad + (size_t)id * G.nchild + (size_t)k * DIR3D_N;
  double num = 0;
  for (int c = threadIdx.x; c < DIR3D_N; c += blockDim.x) num += wl[c / DIR3D_NPHI] * (double)r[c];
  num = block_sum256(num, red);
  __syncthreads();
  const double wsum = DIR3D_NPHI * block_sum256(wpart, red);
  if (threadIdx.x != 0) return;
  double src = num / wsum;
  if (G.jth) G.jth[(size_t)k * G.nd + id] = (float)src;
  const int si = G.sun_idx[k];
  if (si >= 0) {
    const double t2 = G.sun_theta[k];
    int idx = locate_mono<double>(G.theta_grid, NTHETA, t2);
    double p2 = lin(G.theta_grid[idx], p[idx], G.theta_grid[idx + 1], p[idx + 1], t2);
    src += 6.764e-5 * p2 * G.bsun[id] * (double)G.tau_sun[(size_t)id * G.nsun + si];
  }
  G.J[(size_t)k * G.nd + id] = (float)src;
}

// ---- launchers --------------------------------------------------------------
static inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_count_points(int nray, const int* np, const float* rec, int R, int cap, int* count, hipStream_t s) {
  if (nray > 0) hipLaunchKernelGGL(k_count_points, dim3(nblk(nray, 128)), dim3(128), 0, s, nray, np, rec, R, cap, count);
}
void launch_emit_points(int nray, const int* np, const float* rec, int R, int cap, const double4* geo,
                        const double4* geo2, const int* geo_final, const int* jbase, const double* time,
                        const int* atm_idx, ScaPt* pts, long pts_cap, int* status, hipStream_t s) {
  if (nray > 0)
    hipLaunchKernelGGL(k_emit_points, dim3(nblk(nray, 128)), dim3(128), 0, s, nray, np, rec, R, cap, geo, geo2,
                       geo_final, jbase, time, atm_idx, pts, pts_cap, status);
}
void launch_children(int npt, int nalpha, const ScaPt* pts, const double* alpha, RayArrays out, hipStream_t s,
                     const int* d_npt, int* d_nchild) {
  long n = (long)npt * nalpha;
  if (n > 0)
    hipLaunchKernelGGL(k_children, dim3(nblk(n, 128)), dim3(128), 0, s, npt, nalpha, pts, alpha, out, d_npt,
                       d_nchild);
  else if (d_nchild)
    launch_zero_int(d_nchild, 1, s);
}
void launch_sun(int npt, const ScaPt* pts, const SunArgs& S, int* valid, double* theta, RayArrays out,
                hipStream_t s, const int* d_npt) {
  if (npt > 0)
    hipLaunchKernelGGL(k_sun, dim3(nblk(npt, 64)), dim3(64), 0, s, npt, pts, S, valid, theta, out, d_npt);
}

// ---- exclusive prefix sums with a carried base (graph-capture safe) --------
// out[i] = *base + sum_{j<i} in[j], *total = *base + sum in (base null: 0);
// three kernels over blocks of 1024 elements, no library temporaries and no
// memsets.
namespace {
constexpr int SCAN_B = 1024;
__global__ void __launch_bounds__(SCAN_B) k_scan_blocks(const int* in, int n, int* out, int* bsum) {
  __shared__ int v[SCAN_B];
  const int t = threadIdx.x;
  const long i = (long)blockIdx.x * SCAN_B + t;
  const int x = i < n ? in[i] : 0;
  v[t] = x;
  __syncthreads();
  for (int o = 1; o < SCAN_B; o <<= 1) {
    const int y = t >= o ? v[t - o] : 0;
    __syncthreads();
    v[t] += y;
    __syncthreads();
  }
  if (i < n) out[i] = v[t] - x;  // exclusive within the block
  if (t == SCAN_B - 1) bsum[blockIdx.x] = v[t];
}
// One block: block sums -> exclusive block offsets (plus the base), total.
__global__ void __launch_bounds__(SCAN_B) k_scan_sums(int* bsum, int nb, const int* base, int* total) {
  __shared__ int part[SCAN_B];
  const int t = threadIdx.x, per = (nb + SCAN_B - 1) / SCAN_B;
  int sum = 0;
  for (int j = 0; j < per; j++) {
    const int b = t * per + j;
    if (b < nb) sum += bsum[b];
  }
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < SCAN_B; o <<= 1) {
    const int y = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += y;
    __syncthreads();
  }
  const int b0 = base ? *base : 0;
  int run = b0 + part[t] - sum;
  for (int j = 0; j < per; j++) {
    const int b = t * per + j;
    if (b < nb) {
      const int c = bsum[b];
      bsum[b] = run;
      run += c;
    }
  }
  if (t == SCAN_B - 1) *total = b0 + part[t];
}
__global__ void k_scan_add(int* out, int n, const int* boff) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += boff[i / SCAN_B];
}
}  // namespace

size_t scan_base_temp_ints(int n) { return (size_t)(n + SCAN_B - 1) / SCAN_B + 1; }
void launch_scan_base(const int* in, int n, int* out, int* tmp, const int* base, int* total, hipStream_t s) {
  const int nb = (n + SCAN_B - 1) / SCAN_B;
  if (n > 0) hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(SCAN_B), 0, s, in, n, out, tmp);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_B), 0, s, tmp, nb, base, total);
  if (n > 0) hipLaunchKernelGGL(k_scan_add, dim3(nblk(n, 256)), dim3(256), 0, s, out, n, tmp);
}
template <int NA>
void launch_angular_t(const AngArgs& G, hipStream_t s) {
  hipLaunchKernelGGL(k_angular<NA>, dim3(G.npt), dim3(256), 0, s, G);
}
void launch_angular(const AngArgs& G, hipStream_t s) {
  if (G.npt <= 0) return;
  switch (G.nalpha) {
    case 10: launch_angular_t<10>(G, s); break;
    case 12: launch_angular_t<12>(G, s); break;
    case 14: launch_angular_t<14>(G, s); break;
    case 16: launch_angular_t<16>(G, s); break;
    case 18: launch_angular_t<18>(G, s); break;
    case 20: launch_angular_t<20>(G, s); break;
    case 22: launch_angular_t<22>(G, s); break;
    case 24: launch_angular_t<24>(G, s); break;
    case 26: launch_angular_t<26>(G, s); break;
    case 30: launch_angular_t<30>(G, s); break;
    case 32: launch_angular_t<32>(G, s); break;
    default: launch_angular_t<28>(G, s); break;
  }
}
void launch_children3d(int npt, const ScaPt* pts, RayArrays out, hipStream_t s) {
  const long n = (long)npt * DIR3D_N;
  if (n > 0) hipLaunchKernelGGL(k_children3d, dim3(nblk(n, 128)), dim3(128), 0, s, npt, pts, out);
}
void launch_angular3d(const AngArgs& G, hipStream_t s) {
  if (G.npt > 0 && G.nd > 0) hipLaunchKernelGGL(k_angular3d, dim3(G.npt, G.nd), dim3(256), 0, s, G);
}
void launch_sun_compact(int npt, const int* valid, const int* pos, int* sun_idx, hipStream_t s) {
  if (npt > 0) hipLaunchKernelGGL(k_sun_compact, dim3(nblk(npt, 256)), dim3(256), 0, s, npt, valid, pos, sun_idx);
}
void launch_gather_rays(int npt, const int* valid, const int* pos, RayArrays in, RayArrays out, hipStream_t s) {
  if (npt > 0) hipLaunchKernelGGL(k_gather_rays, dim3(nblk(npt, 256)), dim3(256), 0, s, npt, valid, pos, in, out);
}
size_t scan_temp_bytes(int n) {
  size_t b = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, (const int*)nullptr, (int*)nullptr, n);
  return b;
}
void launch_exclusive_scan(const int* in, int* out, void* tmp, size_t bytes, int n, hipStream_t s) {
  (void)hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, s);
}
