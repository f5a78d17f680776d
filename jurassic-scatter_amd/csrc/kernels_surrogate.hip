// kernels_surrogate.hip -- inputs of the scattering-source surrogate: the
// 24 features of every (scattering point, channel) pair, their
// normalisation to BF16 network rows, the training-hull test, and the
// compaction of the points that fall back to the plain angular integral.
#include "jur_surrogate.h"

namespace {

// ln B(T, nu) of the Planck radiance in the reference's units
// (forwardmodel.c:569-573).
__device__ inline double ln_planck(double t, double nu) {
  return log(C1 * nu * nu * nu) - log(expm1(C2 * nu / t));
}

__device__ inline double temp_at(const DevAtm& A, int a, double z) {
  const double* zz = A.z + (size_t)a * A.nlev;
  z = z < zz[0] ? zz[0] : (z > zz[A.nlev - 1] ? zz[A.nlev - 1] : z);
  const int ip = locate_mono<double>(zz, A.nlev, z);
  return atm_pt(A, a, z, ip).t;
}

}  // namespace

__global__ void k_features(FeatArgs F) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)F.npt * F.nd) return;
  const int k = (int)(t / F.nd), id = (int)(t % F.nd);
  const ScaPt P = F.pts[k];
  const DevAtm& A = F.atm;
  const int a = P.atm, nl = A.nlev;
  const double* zz = A.z + (size_t)a * nl;
  const double* pp = A.p + (size_t)a * nl;
  const double* tt = A.t + (size_t)a * nl;
  double z, lon, lat;
  cart2geo(P.x, &z, &lon, &lat);
  const double zc = z < zz[0] ? zz[0] : (z > zz[nl - 1] ? zz[nl - 1] : z);
  const int ip = locate_mono<double>(zz, nl, zc);
  const AtmPoint ap = atm_pt(A, a, zc, ip);
  double tmin = ap.t, colup = 0, coldn = 0;
  for (int l = 1; l < nl; l++) {
    if (zz[l] > zc && tt[l] < tmin) tmin = tt[l];
    if (F.ig_h2o >= 0) {
      const double q = A.q[((size_t)a * A.ng + F.ig_h2o) * nl + l];
      const double c = q * pp[l] / tt[l] * (zz[l] - zz[l - 1]);
      if (0.5 * (zz[l] + zz[l - 1]) > zc) colup += c;
      else coldn += c;
    }
  }
  const double r = norm3(P.x), rd = norm3(P.dx);
  const double cz = rd > 0 ? dot3(P.dx, P.x) / (r * rd) : 0.0;
  const int il = P.il;
  const double be = F.beta_e[il * F.nd + id], bs = F.beta_s[il * F.nd + id];
  const double nu = F.nu[id];
  const double ext = (double)P.af * be;
  float f[MLP_IN];
  f[0] = (float)log(ap.p);
  f[1] = (float)ap.t;
  f[2] = (float)z;
  f[3] = (float)log(ext > 1e-20 ? ext : 1e-20);
  f[4] = (float)(be > 0 ? bs / be : 0.0);
  f[5] = F.asym[il * F.nd + id];
  f[6] = logf(fmaxf(F.pfwd[il * F.nd + id], 1e-20f));
  f[7] = logf(fmaxf(F.pback[il * F.nd + id], 1e-20f));
  f[8] = (float)cz;
  f[9] = (float)(F.top[il] - z);
  f[10] = (float)(z - F.bottom[il]);
  f[11] = (float)nu;
  f[12] = (float)ln_planck(ap.t, nu);
  f[13] = (float)ln_planck(tt[0], nu);
  f[14] = (float)ln_planck(temp_at(A, a, zc + 3.0), nu);
  f[15] = (float)ln_planck(temp_at(A, a, zc - 3.0), nu);
  f[16] = (float)ln_planck(tmin, nu);
  f[17] = (float)log(colup + 1e-30);
  f[18] = (float)log(coldn + 1e-30);
  const double tau_l = ext * (F.top[il] - F.bottom[il]);
  f[19] = (float)log(tau_l > 1e-20 ? tau_l : 1e-20);
  const size_t nud = 2 * (size_t)F.npt;
  const double lup = F.ud[(size_t)id * nud + 2 * k], ldn = F.ud[(size_t)id * nud + 2 * k + 1];
  f[20] = (float)(log(fmax(lup, 1e-30)) - f[12]);
  f[21] = (float)(log(fmax(ldn, 1e-30)) - f[12]);
  // vertical optical depth of the layer above and below the point
  const double ta = ext * fmax(F.top[il] - z, 0.0), tb = ext * fmax(z - F.bottom[il], 0.0);
  f[22] = (float)log(ta + 1e-6);
  f[23] = (float)log(tb + 1e-6);
  if (F.raw)
    for (int i = 0; i < MLP_IN; i++) F.raw[t * MLP_IN + i] = f[i];
  if (F.lnb) F.lnb[t] = f[12];
  if (F.x) {
    const MlpNorm& N = *F.norm;
    bool out = false;
    uint16_t* row = F.x + t * MLP_INP;
    for (int i = 0; i < MLP_IN; i++) {
      const float v = (f[i] - N.mu[i]) / N.sd[i];
      out |= !(v >= N.lo[i] - F.margin && v <= N.hi[i] + F.margin);
      row[i] = mlp_f2bf(v);
    }
    for (int i = MLP_IN; i < MLP_INP; i++) row[i] = 0;
    if (out) F.out_hull[k] = 1;
  }
}

// sub[k] = position of point k among the fallback points (flag set) or -1;
// the fallback points are copied to pts_sub in order.
__global__ void k_fallback_compact(int npt, const int* flag, const int* pos, const ScaPt* pts, int* sub,
                                   ScaPt* pts_sub) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npt) return;
  if (flag[k]) {
    sub[k] = pos[k];
    pts_sub[pos[k]] = pts[k];
  } else {
    sub[k] = -1;
  }
}

void launch_features(const FeatArgs& a, hipStream_t s) {
  const long n = (long)a.npt * a.nd;
  if (n > 0) hipLaunchKernelGGL(k_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
}
void launch_fallback_compact(int npt, const int* flag, const int* pos, const ScaPt* pts, int* sub,
                             ScaPt* pts_sub, hipStream_t s) {
  if (npt > 0)
    hipLaunchKernelGGL(k_fallback_compact, dim3((npt + 255) / 256), dim3(256), 0, s, npt, flag, pos, pts, sub,
                       pts_sub);
}
