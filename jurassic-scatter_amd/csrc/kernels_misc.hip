// kernels_misc.hip -- small device kernels around the forward model:
// hydrostatic equilibrium, field-of-view convolution, brightness
// temperature, precision conversion and the FP64 LOS dump.
#include "jur_dev.h"
#include "jur_kernels_misc.h"

// hydrostatic_1d (reference atmosphere.c:188-243): one thread per profile
// [ip0, ip1) of a single atmosphere; gravity from atmosphere.c:153-159.
__device__ inline double gravity_d(double z, double lat) {
  double s1 = sin(lat / 180 * M_PI), s2 = sin(2 * lat / 180 * M_PI);
  return 9.780318 * (1 + 0.0053024 * s1 * s1 - 0.0000058 * s2 * s2) - 3.086e-3 * z;
}

__global__ void k_hydrostatic(double* z, double* p, const double* t, const double* qh2o,
                              const double* lat, const int* prof0, const int* prof1, int nprof,
                              double hydz) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nprof) return;
  const int ip0 = prof0[k], ip1 = prof1[k];
  const double mmair = 28.96456e-3, mmh2o = 18.0153e-3;
  const int ipts = 20;
  double dzmin = 1e99;
  int ipref = 0;
  for (int ip = ip0; ip < ip1; ip++)
    if (fabs(z[ip] - hydz) < dzmin) {
      dzmin = fabs(z[ip] - hydz);
      ipref = ip;
    }
  for (int ip = ipref + 1; ip < ip1; ip++) {
    double mean = 0, e = 0;
    for (int i = 0; i < ipts; i++) {
      double zz = lin(0.0, z[ip - 1], ipts - 1.0, z[ip], (double)i);
      if (qh2o) e = lin(0.0, qh2o[ip - 1], ipts - 1.0, qh2o[ip], (double)i);
      mean += (e * mmh2o + (1 - e) * mmair) * gravity_d(zz, lat[ipref]) / JUR_RGAS /
              lin(0.0, t[ip - 1], ipts - 1.0, t[ip], (double)i) / ipts;
    }
    p[ip] = exp(log(p[ip - 1]) - mean * 1000 * (z[ip] - z[ip - 1]));
  }
  for (int ip = ipref - 1; ip >= ip0; ip--) {
    double mean = 0, e = 0;
    for (int i = 0; i < ipts; i++) {
      double zz = lin(0.0, z[ip + 1], ipts - 1.0, z[ip], (double)i);
      if (qh2o) e = lin(0.0, qh2o[ip + 1], ipts - 1.0, qh2o[ip], (double)i);
      mean += (e * mmh2o + (1 - e) * mmair) * gravity_d(zz, lat[ipref]) / JUR_RGAS /
              lin(0.0, t[ip + 1], ipts - 1.0, t[ip], (double)i) / ipts;
    }
    p[ip] = exp(log(p[ip + 1]) - mean * 1000 * (z[ip] - z[ip + 1]));
  }
}

void launch_hydrostatic(double* z, double* p, const double* t, const double* qh2o,
                        const double* lat, const int* prof0, const int* prof1, int nprof,
                        double hydz, hipStream_t s) {
  hipLaunchKernelGGL(k_hydrostatic, dim3((nprof + 63) / 64), dim3(64), 0, s, z, p, t, qh2o, lat,
                     prof0, prof1, nprof, hydz);
}

// formod_fov (reference forwardmodel.c:156-220): convolution of each ray's
// radiance and transmittance with the field-of-view shape over neighbouring
// rays (+-NFOV) of the same time.  One thread per ray; rays are [nr] with
// independent runs of seg rays (one observation set each).
__global__ void k_fov(int nr, int seg, int nd, const double* time, const double* tpz, const double* rad_in,
                      const double* tau_in, const double* dz, const double* w, int n,
                      double* rad_out, double* tau_out, int* status) {
  int ir = blockIdx.x * blockDim.x + threadIdx.x;
  if (ir >= nr) return;
  int idx[2 * NFOV + 1];
  double z[2 * NFOV + 1];
  int nz = 0;
  const int s0 = ir / seg * seg, s1 = s0 + seg < nr ? s0 + seg : nr;
  int lo = ir - NFOV > s0 ? ir - NFOV : s0, hi = ir + 1 + NFOV < s1 ? ir + 1 + NFOV : s1;
  for (int ir2 = lo; ir2 < hi; ir2++)
    if (time[ir2] == time[ir]) {
      z[nz] = tpz[ir2];
      idx[nz++] = ir2;
    }
  if (nz < 2) {
    atomicOr(status, JUR_ST_FOV);
    return;
  }
  for (int id = 0; id < nd; id++) {
    double r = 0, tt = 0, wsum = 0;
    for (int i = 0; i < n; i++) {
      double zf = tpz[ir] + dz[i];
      int k = locate_mono<double>(z, nz, zf);
      double r0 = rad_in[(size_t)id * nr + idx[k]], r1 = rad_in[(size_t)id * nr + idx[k + 1]];
      double t0 = tau_in[(size_t)id * nr + idx[k]], t1 = tau_in[(size_t)id * nr + idx[k + 1]];
      r += w[i] * lin(z[k], r0, z[k + 1], r1, zf);
      tt += w[i] * lin(z[k], t0, z[k + 1], t1, zf);
      wsum += w[i];
    }
    rad_out[(size_t)id * nr + ir] = r / wsum;
    tau_out[(size_t)id * nr + ir] = tt / wsum;
  }
}

void launch_fov(int nr, int seg, int nd, const double* time, const double* tpz, const double* rad_in,
                const double* tau_in, const double* dz, const double* w, int n, double* rad_out,
                double* tau_out, int* status, hipStream_t s) {
  if (nr <= 0 || seg <= 0) return;
  hipLaunchKernelGGL(k_fov, dim3((nr + 63) / 64), dim3(64), 0, s, nr, seg, nd, time, tpz, rad_in, tau_in,
                     dz, w, n, rad_out, tau_out, status);
}

// brightness (reference forwardmodel.c:7-11) with gsl_log1p's formula.
__global__ void k_brightness(int nr, int nd, const double* nu, double* rad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr * nd) return;
  int id = i / nr;
  double v = nu[id];
  double x = C1 * (v * v * v) / rad[i];
  volatile double y = 1 + x;
  double zz = y - 1;
  rad[i] = C2 * v / (log(y) - (zz - x) / y);
}

void launch_brightness(int nr, int nd, const double* nu, double* rad, hipStream_t s) {
  int n = nr * nd;
  if (n > 0) hipLaunchKernelGGL(k_brightness, dim3((n + 255) / 256), dim3(256), 0, s, nr, nd, nu, rad);
}

__global__ void k_f2d(const float* a, double* b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}
void launch_f2d(const float* a, double* b, size_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_f2d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, n);
}

// FP64 LOS of one traced ray (for the raytrace() entry point): altitude,
// position, atmosphere, column densities and segment lengths at each final
// point, from the FP64 point geometry of the ray tracer.
__global__ void k_los64(DevAtm M, int a, const double4* geo, const float* rec, int R, int np,
                        int ng, int nw, double* out_zll, double* out_pt, double* out_q,
                        double* out_k, double* out_u, double* out_ds, double* out_af, int* out_ai) {
  int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  double4 g = geo[ip];
  const GeoLoc L = geo_locate(M, a, g.x, g.y, g.z);
  AtmPoint ap = geo_pt(M, a, L);
  out_zll[3 * ip] = g.x;
  out_zll[3 * ip + 1] = g.y;
  out_zll[3 * ip + 2] = g.z;
  out_pt[2 * ip] = ap.p;
  out_pt[2 * ip + 1] = ap.t;
  out_ds[ip] = g.w;
  for (int ig = 0; ig < ng; ig++) {
    double q = geo_q(M, a, ig, L);
    out_q[ip * NGMAX + ig] = q;
    out_u[ip * NGMAX + ig] = 10 * q * ap.p / (JUR_KBOLTZ * ap.t) * g.w;
  }
  for (int iw = 0; iw < nw; iw++) out_k[ip * NWMAX + iw] = geo_k(M, a, iw, L);
  out_af[ip] = rec[(size_t)ip * R + REC_AF];
  out_ai[ip] = (int)rec[(size_t)ip * R + REC_AI];
}

void launch_los64(const DevAtm& M, int a, const double4* geo, const float* rec, int R, int np,
                  int ng, int nw, double* zll, double* pt, double* q, double* k, double* u,
                  double* ds, double* af, int* ai, hipStream_t s) {
  if (np > 0)
    hipLaunchKernelGGL(k_los64, dim3((np + 63) / 64), dim3(64), 0, s, M, a, geo, rec, R, np, ng, nw,
                       zll, pt, q, k, u, ds, af, ai);
}
