// kernels_mie.hip -- Lorenz-Mie optics of log-normal particle modes in FP64.
//
// k_bhmie: one workgroup per (mode, radius node, channel) sphere, the
// Bohren & Huffman series of the reference (src/scatter.c:56-234); with
// grid 1 the phase function of the aerosol configuration's 180 angles
// (k + 0.5) deg resampled to the 181 nodes, Q_ext from the forward
// amplitude.  Lane 0
// runs the two inherently sequential recurrences (logarithmic derivative
// D_n downward from nmx, Riccati-Bessel psi/chi upward with the
// coefficients a_n, b_n and Q_sca), keeping D_n, a_n, b_n in a global
// scratch slice; then the lanes split the 91 scattering angles of the
// half-range and accumulate S1, S2 over n with their own pi_n / tau_n
// recurrences.  Complex division is GSL's (scaled by 1/|b|).
//
// k_mie_mode: the Gauss-Hermite sum over radius nodes for one mode, per
// (channel, angle), in node order (opt_prop_mie_log, scatter.c:476-507).
// k_mie_layers: the sum of the modes of each layer (get_opt_prop,
// scatter.c:367-417).
#include "jur_dev.h"
#include "jur_mie.h"

namespace {

struct cxd {
  double re, im;
};
__device__ inline cxd mk(double a, double b) { return {a, b}; }
__device__ inline cxd cadd(cxd a, cxd b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline cxd csub(cxd a, cxd b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline cxd cmul(cxd a, cxd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline double cabs_(cxd a) { return hypot(a.re, a.im); }
__device__ inline cxd cdiv(cxd a, cxd b) {
  double s = 1.0 / cabs_(b), sbr = s * b.re, sbi = s * b.im;
  return {(a.re * sbr + a.im * sbi) * s, (a.im * sbr - a.re * sbi) * s};
}

constexpr int NT = (NTHETA + 1) / 2;  // 91 angles of the half range
constexpr int MIE_CELLS = 180;        // the configuration grid's angles (k + 0.5) deg

}  // namespace

__global__ void __launch_bounds__(128) k_bhmie(const MieItem* items, int item0, double2* scratch,
                                               size_t stride, double* qext_o, double* qsca_o,
                                               double* phase_o, int grid) {
  const int it = item0 + blockIdx.x;
  const MieItem I = items[it];
  double2* D = scratch + (size_t)blockIdx.x * 3 * stride;
  double2* An = D + stride;
  double2* Bn = An + stride;
  __shared__ double qsca_s, qext_s;
  __shared__ double cell[MIE_CELLS];  // grid 1: the phase function at the 180 cell centres
  const double x = I.x;
  const cxd ref = mk(I.nr, I.ni);
  const int nstop = I.nstop, nmx = I.nmx;
  if (threadIdx.x == 0) {
    const cxd y = cmul(mk(x, 0.0), ref);
    // D_n downward from D_nmx = 0 (scatter.c:95-106); keep n <= nstop.
    cxd d = mk(0, 0);
    for (int n = 1; n <= nmx - 1; n++) {
      double rn = nmx - n + 1;
      cxd t = cadd(d, cdiv(mk(rn, 0.0), y));
      t = cdiv(mk(1.0, 0.0), t);
      d = csub(cdiv(mk(rn, 0.0), y), t);
      int idx = nmx - n;
      if (idx <= nstop) D[idx] = make_double2(d.re, d.im);
    }
    // Upward Riccati-Bessel recurrence, a_n, b_n, Q_sca (scatter.c:121-162).
    double psi0 = cos(x), psi1 = sin(x), chi0 = -sin(x), chi1 = cos(x), apsi1 = psi1;
    cxd xi1 = mk(apsi1, -chi1);
    double qs = 0;
    cxd sf = mk(0, 0);  // forward amplitude sum (2n + 1)/2 (a_n + b_n)
    for (int n = 1; n <= nstop; n++) {
      double rn = n;
      double psi = (2.0 * rn - 1.0) * psi1 / x - psi0;
      double chi = (2.0 * rn - 1.0) * chi1 / x - chi0;
      cxd xi = mk(psi, -chi);
      cxd dn = mk(D[n].x, D[n].y);
      cxd a = cdiv(dn, ref);
      a = cadd(a, mk(rn / x, 0.0));
      cxd an = csub(cmul(a, mk(psi, 0.0)), mk(apsi1, 0.0));
      cxd ad = csub(cmul(a, xi), xi1);
      an = cdiv(an, ad);
      cxd b = cmul(ref, dn);
      b = cadd(b, mk(rn / x, 0.0));
      cxd bn = csub(cmul(b, mk(psi, 0.0)), mk(apsi1, 0.0));
      cxd bd = csub(cmul(b, xi), xi1);
      bn = cdiv(bn, bd);
      double aa = cabs_(an), bb = cabs_(bn);
      qs = qs + (2. * rn + 1.) * (aa * aa + bb * bb);
      sf = cadd(sf, cmul(mk((2. * rn + 1.) / 2., 0.0), cadd(an, bn)));
      An[n] = make_double2(an.re, an.im);
      Bn[n] = make_double2(bn.re, bn.im);
      psi0 = psi1;
      psi1 = psi;
      apsi1 = psi1;
      chi0 = chi1;
      chi1 = chi;
      xi1 = mk(apsi1, -chi1);
    }
    qsca_s = (2.0 / (x * x)) * qs;
    qext_s = (4.0 / (x * x)) * sf.re;
  }
  __syncthreads();
  const double qsca = qsca_s;
  if (grid == 1) {
    // The configuration's angles (k + 0.5) deg: 90 of the half range and
    // their mirrors, no angle on 90 deg; then the 181-node layout, node i
    // the mean of cells i - 1 and i (the end nodes the end cells).
    for (int j = threadIdx.x; j < MIE_CELLS / 2; j += blockDim.x) {
      const double amu = cos((j + 0.5) * M_PI / 180.);
      double pi0 = 0.0, pi1 = 1.0;
      cxd s1 = mk(0, 0), s2 = mk(0, 0), s1m = mk(0, 0), s2m = mk(0, 0);
      for (int n = 1; n <= nstop; n++) {
        const double rn = n, fn = (2.0 * rn + 1.0) / (rn * (rn + 1.0));
        const double pi = pi1, tau = rn * amu * pi - (rn + 1.0) * pi0;
        const cxd an = mk(An[n].x, An[n].y), bn = mk(Bn[n].x, Bn[n].y);
        s1 = cadd(s1, cmul(mk(fn, 0.0), cadd(cmul(an, mk(pi, 0.0)), cmul(bn, mk(tau, 0.0)))));
        s2 = cadd(s2, cmul(mk(fn, 0.0), cadd(cmul(an, mk(tau, 0.0)), cmul(bn, mk(pi, 0.0)))));
        const double p = (n & 1) ? 1.0 : -1.0, t = -p;  // (-1)^(n-1), (-1)^n
        s1m = cadd(s1m, cmul(mk(fn, 0.0), cadd(cmul(an, mk(pi * p, 0.0)), cmul(bn, mk(tau * t, 0.0)))));
        s2m = cadd(s2m, cmul(mk(fn, 0.0), cadd(cmul(an, mk(tau * t, 0.0)), cmul(bn, mk(pi * p, 0.0)))));
        pi1 = ((2. * rn + 1.) * amu * pi - (rn + 1.) * pi0) / rn;
        pi0 = pi;
      }
      const double f = 2 / (x * x * qsca);
      cell[j] = f * (s1.re * s1.re + s1.im * s1.im + s2.re * s2.re + s2.im * s2.im);
      cell[MIE_CELLS - 1 - j] = f * (s1m.re * s1m.re + s1m.im * s1m.im + s2m.re * s2m.re + s2m.im * s2m.im);
    }
    __syncthreads();
    double* ph = phase_o + (size_t)it * NTHETA;
    for (int i = threadIdx.x; i < NTHETA; i += blockDim.x)
      ph[i] = i == 0 ? cell[0] : (i == NTHETA - 1 ? cell[MIE_CELLS - 1] : 0.5 * (cell[i - 1] + cell[i]));
    if (threadIdx.x == 0) {
      qext_o[it] = qext_s;
      qsca_o[it] = qsca;
    }
    return;
  }
  // Scattering amplitudes, one half-range angle per lane (scatter.c:163-214).
  for (int j = 1 + threadIdx.x; j <= NT; j += blockDim.x) {
    const int jj = 2 * NT - j;
    const double amu = cos((double)(j - 1) * (.5 * M_PI / (double)(NT - 1)));
    double pi0 = 0.0, pi1 = 1.0;
    cxd s1 = mk(0, 0), s2 = mk(0, 0), s1m = mk(0, 0), s2m = mk(0, 0);
    for (int n = 1; n <= nstop; n++) {
      double rn = n;
      double fn = (2.0 * rn + 1.0) / (rn * (rn + 1.0));
      double pi = pi1;
      double tau = rn * amu * pi - (rn + 1.0) * pi0;
      cxd an = mk(An[n].x, An[n].y), bn = mk(Bn[n].x, Bn[n].y);
      cxd t = cadd(cmul(an, mk(pi, 0.0)), cmul(bn, mk(tau, 0.0)));
      s1 = cadd(s1, cmul(mk(fn, 0.0), t));
      t = cadd(cmul(an, mk(tau, 0.0)), cmul(bn, mk(pi, 0.0)));
      s2 = cadd(s2, cmul(mk(fn, 0.0), t));
      if (j != jj) {
        double p = (n & 1) ? 1.0 : -1.0;   // (-1)^(n-1)
        double tt = -p;                     // (-1)^n
        t = cadd(cmul(an, mk(pi * p, 0.0)), cmul(bn, mk(tau * tt, 0.0)));
        s1m = cadd(s1m, cmul(mk(fn, 0.0), t));
        t = cadd(cmul(an, mk(tau * tt, 0.0)), cmul(bn, mk(pi * p, 0.0)));
        s2m = cadd(s2m, cmul(mk(fn, 0.0), t));
      }
      pi1 = ((2. * rn + 1.) * amu * pi - (rn + 1.) * pi0) / rn;
      pi0 = pi;
    }
    double* ph = phase_o + (size_t)it * NTHETA;
    const double f = 2 / (x * x * qsca);
    ph[j - 1] = f * (s1.re * s1.re + s1.im * s1.im + s2.re * s2.re + s2.im * s2.im);
    if (j != jj) ph[jj - 1] = f * (s1m.re * s1m.re + s1m.im * s1m.im + s2m.re * s2m.re + s2m.im * s2m.im);
    if (j == 1) {
      qext_o[it] = (4.0 / (x * x)) * s1.re;
      qsca_o[it] = qsca;
    }
  }
}

// One mode: beta_ext, beta_sca and the phase function per channel, summed
// over the radius nodes in node order; thread (id, jj), jj == NTHETA for
// the extinction / scattering coefficients.
__global__ void k_mie_mode(const MieNode* nodes, int nnode, const int* item_of, int nd,
                           const double* qext, const double* qsca, const double* phase,
                           double* mbe, double* mbs, double* mp) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nd * (NTHETA + 1)) return;
  int id = k / (NTHETA + 1), jj = k % (NTHETA + 1);
  double be = 0, bs = 0, ph = 0;
  for (int nn = 0; nn < nnode; nn++) {
    int it = item_of[nn * nd + id];
    if (it < 0) continue;
    const MieNode N = nodes[nn];
    double r2 = N.rad * N.rad;
    if (jj == NTHETA) {
      be += N.K1 * r2 * qext[it] * N.w;
      bs += N.K1 * r2 * qsca[it] * N.w;
    } else {
      ph += N.K1 * qsca[it] * r2 * phase[(size_t)it * NTHETA + jj] * N.w;
    }
  }
  if (jj == NTHETA) {
    mbe[id] = be;
    mbs[id] = bs;
  } else {
    mp[id * NTHETA + jj] = ph;  // normalised by beta_sca in k_mie_norm
  }
}

__global__ void k_mie_norm(int nd, const double* mbs, double* mp) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nd * NTHETA) return;
  mp[k] /= mbs[k / NTHETA];
}

// Layer sums over the layer's modes in order; thread (il, id, jj).
__global__ void k_mie_layers(int nl, const int* mode0, const int* nmod, int nd, const double* mbe,
                             const double* mbs, const double* mp, double* be, double* bs, double* ba,
                             double* p) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nl * nd * (NTHETA + 1)) return;
  int il = k / (nd * (NTHETA + 1));
  int r = k % (nd * (NTHETA + 1));
  int id = r / (NTHETA + 1), jj = r % (NTHETA + 1);
  if (jj == NTHETA) {
    double e = 0, s = 0, a = 0;
    for (int m = mode0[il]; m < mode0[il] + nmod[il]; m++) {
      e += mbe[m * nd + id];
      s += mbs[m * nd + id];
      a += (mbe[m * nd + id] - mbs[m * nd + id]);
    }
    be[il * nd + id] = e;
    bs[il * nd + id] = s;
    ba[il * nd + id] = a;
  } else {
    double s = 0;
    for (int m = mode0[il]; m < mode0[il] + nmod[il]; m++) s += mp[((size_t)m * nd + id) * NTHETA + jj];
    p[((size_t)il * nd + id) * NTHETA + jj] = s;
  }
}

void launch_bhmie(const MieItem* items, int item0, int nitems, double2* scratch, size_t stride,
                  double* qext, double* qsca, double* phase, int grid, hipStream_t s) {
  if (nitems > 0)
    hipLaunchKernelGGL(k_bhmie, dim3(nitems), dim3(128), 0, s, items, item0, scratch, stride, qext, qsca,
                       phase, grid);
}
void launch_mie_mode(const MieNode* nodes, int nnode, const int* item_of, int nd, const double* qext,
                     const double* qsca, const double* phase, double* mbe, double* mbs, double* mp,
                     hipStream_t s) {
  int n = nd * (NTHETA + 1);
  hipLaunchKernelGGL(k_mie_mode, dim3((n + 127) / 128), dim3(128), 0, s, nodes, nnode, item_of, nd, qext,
                     qsca, phase, mbe, mbs, mp);
  int m = nd * NTHETA;
  hipLaunchKernelGGL(k_mie_norm, dim3((m + 127) / 128), dim3(128), 0, s, nd, mbs, mp);
}
void launch_mie_layers(int nl, const int* mode0, const int* nmod, int nd, const double* mbe,
                       const double* mbs, const double* mp, double* be, double* bs, double* ba, double* p,
                       hipStream_t s) {
  int n = nl * nd * (NTHETA + 1);
  hipLaunchKernelGGL(k_mie_layers, dim3((n + 127) / 128), dim3(128), 0, s, nl, mode0, nmod, nd, mbe, mbs,
                     mp, be, bs, ba, p);
}
