// kernels_raytrace.hip -- refracted ray tracing of many lines of sight.
//
// One thread traces one ray in FP64 (reference src/lineofsight.c:5-190):
// the march of jur_march.h, then the reference's post-processing: drop of a
// vanishing last segment, tangent point (:411-452), recursive half-sum of
// the step lengths (:177-179), and, with aerosol layers, the layer-edge point
// insertion and chord-based segment lengths of add_aerosol_layers
// (:193-372).  The march keeps only FP64 (z, lon, lat, ds) per point;
// pressure, temperature, mixing ratios and extinction are re-interpolated at
// each final point's altitude (the same function of z the reference
// evaluates there), and stored as the FP32 point records of the
// path-integration kernel together with the column densities
// u = 10 q p / (k_B T) ds (:182-185, :330-332).
#include "jur_march.h"

namespace {

struct BufSink {
  double4* G;
  __device__ void put(int i, const G4& g) { G[i] = make_double4(g.z, g.lon, g.lat, g.ds); }
};

__device__ inline G4 load_g(const double4* G, int i) {
  double4 v = G[i];
  return {v.x, v.y, v.z, v.w};
}

struct InsertSink {
  double4* G2;
  int n, cap, status;
  __device__ void operator()(const G4& g) {
    if (n >= cap) {
      status |= JUR_ST_CAP;
      return;
    }
    G2[n++] = make_double4(g.z, g.lon, g.lat, 0.0);
  }
};

}  // namespace

__global__ void __launch_bounds__(128) k_raytrace(TraceArgs A) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= A.nray) return;
  const int r = A.order ? A.order[tid] : tid;
  if (A.nvalid && r + A.nvalid_off >= *A.nvalid) {
    // beyond the device-side count of a batch sized by its capacity: an
    // empty ray (no LOS points, no surface)
    A.np_out[r] = 0;
    A.tsurf_out[r] = -999.f;
    if (A.tsurf64) A.tsurf64[r] = -999.0;
    A.tpz[r] = A.tplon[r] = A.tplat[r] = NAN;
    if (A.geo_final) A.geo_final[r] = 0;
    return;
  }
  const int a = A.atm_idx ? A.atm_idx[r] : 0;
  const DevAtm& M = A.atm;
  double4* G = A.geo + (size_t)r * A.cap;
  double tpz = A.vpz[r], tplon = A.vplon[r], tplat = A.vplat[r];

  BufSink sink{G};
  MarchOut mo = march_ray(M, a, A.obsz[r], A.obslon[r], A.obslat[r], A.vpz[r], A.vplon[r], A.vplat[r],
                          A.rayds, A.raydz, A.refrac, A.cap, sink);
  int np = mo.np, status = mo.status;
  const int fatal = JUR_ST_CAP | JUR_ST_NLOS | JUR_ST_BELOW_SURFACE;
  if (!(status & fatal) && np > 0) {
    // Drop a vanishing last segment (lineofsight.c:170-171).
    if (np >= 2 && G[np - 2].w < 1e-3 && np - 1 > 1) np--;
    // Tangent point from the pre-trapezoid steps (lineofsight.c:411-452):
    // first minimum of z; limb rays get a quadratic fit.
    int im = 0;
    double zm = G[0].x;
    for (int i = 0; i < np; i++) {
      double zi = G[i].x;
      if (zi < zm) { zm = zi; im = i; }
      if (isnan(zi)) { im = i; break; }
    }
    if (im <= 0 || im >= np - 1) {
      G4 l = load_g(G, np - 1);
      tpz = l.z; tplon = l.lon; tplat = l.lat;
    } else {
      G4 g0 = load_g(G, im - 1), g1 = load_g(G, im), g2 = load_g(G, im + 1);
      double yy0 = g0.z, yy1 = g1.z, yy2 = g2.z;
      double x1 = sqrt(g1.ds * g1.ds - (yy1 - yy0) * (yy1 - yy0));
      double x2 = x1 + sqrt(g2.ds * g2.ds - (yy2 - yy1) * (yy2 - yy1));
      double ca = 1 / (x1 - x2) * (-(yy0 - yy1) / x1 + (yy0 - yy2) / x2);
      double cb = -(yy0 - yy1) / x1 - ca * x1;
      double xt = -cb / (2 * ca);
      tpz = ca * xt * xt + cb * xt + yy0;
      double v0[3], v2[3], v[3], dummy;
      geo2cart(g0.z, g0.lon, g0.lat, v0);
      geo2cart(g2.z, g2.lon, g2.lat, v2);
      for (int i = 0; i < 3; i++) v[i] = lin(0.0, v0[i], x2, v2[i], xt);
      cart2geo(v, &dummy, &tplon, &tplat);
    }
    // Recursive half-sum of steps, as the reference (lineofsight.c:177-179):
    // each ds is averaged with its already updated predecessor.
    double dprev = 0.5 * G[0].w;
    G[0].w = dprev;
    for (int i = 1; i < np; i++) {
      dprev = 0.5 * (dprev + G[i].w);
      G[i].w = dprev;
    }
  }

  // Aerosol-layer point insertion (lineofsight.c:193-372) into geo2; the
  // segment lengths, layer membership and point records of the final
  // points follow in k_records (one thread per point).
  const double4* F = G;
  int nf = np;
  if (A.sca_n > 0 && np > 0 && !(status & fatal)) {
    double4* G2 = A.geo2 + (size_t)r * A.cap;
    const int jl = A.nalti;
    const double altimax = jl > 0 ? A.altimax : NAN, altimin = jl > 0 ? A.altimin : NAN;
    InsertSink ins{G2, 0, A.cap, 0};
    G2[ins.n++] = G[0];
    for (int ip = 1; ip < np && !ins.status; ip++) {
      G4 g0 = load_g(G, ip - 1), g1 = load_g(G, ip);
      edge_points(g0, g1, A.alti, jl, altimax, altimin, ins);
      if (ins.status) break;
      if (ins.n >= A.cap) { ins.status |= JUR_ST_CAP; break; }
      if (ins.n > NLOS) { ins.status |= JUR_ST_NLOS; break; }
      G2[ins.n++] = G[ip];
    }
    status |= ins.status;
    if (!(status & fatal)) {
      F = G2;
      nf = ins.n;
    }
  }
  if (status & fatal) nf = 0;

  A.np_out[r] = nf;
  A.tsurf_out[r] = (float)mo.tsurf;
  if (A.tsurf64) A.tsurf64[r] = mo.tsurf;
  A.tpz[r] = tpz;
  A.tplon[r] = tplon;
  A.tplat[r] = tplat;
  if (A.geo_final) A.geo_final[r] = (F == G) ? 0 : 1;
  if (status) atomicOr(A.status, status);
}

// Cartesian positions of a point and its two neighbours on the ray, and the
// previous point's altitude (aerosol layers only).
struct Nbr {
  double xo[3], xp[3], xn[3];
  double zp;
};

// Point ip of ray r (nf final points): with aerosol layers the segment
// length from Cartesian half-chords (lineofsight.c:310-324; the positions
// come in nb, computed once per point by k_records) and the layer
// membership / transition factor (:334-363); then the atmosphere at the
// point's altitude and the FP32 record with the column densities
// u = 10 q p / (k_B T) ds (:182-185, :330-332).
__device__ inline void record_point(const TraceArgs& A, int r, int ip, int nf, const Nbr& nb) {
  const int a = A.atm_idx ? A.atm_idx[r] : 0;
  const DevAtm& M = A.atm;
  // k_raytrace inserted layer-edge points (into geo2) for every ray with
  // points when aerosol layers are on
  const bool ins = A.sca_n > 0 && A.geo2 != nullptr;
  double4* F = (ins ? A.geo2 : A.geo) + (size_t)r * A.cap;
  float* pr = A.rec + ((size_t)r * A.cap + ip) * A.R;
  const double4 g = F[ip];
  double ds = g.w;
  float af = 0.f, ai = -999.f;
  double af64 = 0.0;
  if (ins) {
    if (nf >= 2) {
      if (ip == 0 || ip == nf - 1)
        ds = 0.5 * dist3(nb.xo, ip == 0 ? nb.xn : nb.xp);
      else
        ds = 0.5 * (dist3(nb.xp, nb.xo) + dist3(nb.xo, nb.xn));
      F[ip].w = ds;
    }
    const double zi = g.x;
    const double altimax = A.nalti > 0 ? A.altimax : NAN, altimin = A.nalti > 0 ? A.altimin : NAN;
    const bool pre = (ip == 0) || ((nb.zp < altimax || zi < altimax) && (nb.zp > altimin || zi > altimin));
    double afd = 0;
    int aii = -999;
    if (pre) {
      for (int il = 0; il < A.nl; il++) {
        double top = A.top[il], bot = A.bottom[il], tr = A.trans[il];
        if (zi <= top && zi >= bot) { aii = il; afd = 1.; }
        if (tr > A.transs && zi <= top + tr && zi > top) { aii = il; afd = (top + tr - zi) / tr; }
        if (tr > A.transs && zi < bot && zi >= bot - tr) { aii = il; afd = fabs(bot - tr - zi) / tr; }
      }
    }
    af = (float)afd;
    af64 = afd;
    ai = (float)aii;
  }
  pr[REC_AF] = af;
  pr[REC_AI] = ai;
  const int uo = rec_u(A.nw);
  // the atmosphere at the point (IP 1: its altitude bracket; IP 2 / 3 also
  // its longitude and latitude)
  const GeoLoc L = geo_locate(M, a, g.x, g.y, g.z);
  const AtmPoint ap = geo_pt(M, a, L);
  pr[REC_P] = (float)ap.p;
  pr[REC_T] = (float)ap.t;
  pr[REC_DS] = (float)ds;
  double* pd = A.rec64 ? A.rec64 + ((size_t)r * A.cap + ip) * A.R : nullptr;
  if (pd) {
    pd[REC_AF] = af64;
    pd[REC_AI] = ai;
    pd[REC_P] = ap.p;
    pd[REC_T] = ap.t;
    pd[REC_DS] = ds;
  }
  double qh = 0;
  for (int ig = 0; ig < M.ng; ig++) {
    const double q = geo_q(M, a, ig, L);
    if (ig == A.ig_h2o) qh = q;
    const double u = 10 * q * ap.p / (JUR_KBOLTZ * ap.t) * ds;
    pr[uo + ig] = (float)u;
    if (pd) pd[uo + ig] = u;
  }
  pr[REC_QH2O] = (float)qh;
  if (pd) pd[REC_QH2O] = qh;
  for (int iw = 0; iw < M.nw; iw++) {
    const double k = geo_k(M, a, iw, L);
    pr[REC_K + iw] = (float)k;
    if (pd) pd[REC_K + iw] = k;
  }
}

// One wave per ray, its lanes over the ray's final points (a thread per
// point up to the LOS capacity left most threads without a point: the
// capacity is sized for the longest ray of the batch).  With aerosol layers
// the wave walks the ray in windows of 62 points with a halo point on each
// side: every lane converts its own point to Cartesian coordinates once and
// takes its neighbours' from the adjacent lanes, instead of converting all
// three points itself.
__global__ void __launch_bounds__(256) k_records(TraceArgs A) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= A.nray) return;
  const int nf = A.np_out[r];
  const int lane = threadIdx.x & 63;
  if (!(A.sca_n > 0 && A.geo2 != nullptr) || nf < 2) {
    const Nbr none{};
    for (int ip = lane; ip < nf; ip += 64) record_point(A, r, ip, nf, none);
    return;
  }
  const double4* F = A.geo2 + (size_t)r * A.cap;
  for (int base = 0; base < nf; base += 62) {  // uniform across the wave
    const int ip = base - 1 + lane;
    Nbr nb{};
    double z = 0;
    if (ip >= 0 && ip < nf) {
      const double4 g = F[ip];
      geo2cart(g.x, g.y, g.z, nb.xo);
      z = g.x;
    }
    for (int i = 0; i < 3; i++) {
      nb.xp[i] = __shfl_up(nb.xo[i], 1, 64);
      nb.xn[i] = __shfl_down(nb.xo[i], 1, 64);
    }
    nb.zp = __shfl_up(z, 1, 64);
    if (lane >= 1 && lane <= 62 && ip < nf) record_point(A, r, ip, nf, nb);
  }
}

// March-order key of each ray from its straight line observer -> view
// point: the march of a wave lasts as long as its longest ray, and rays in
// input order mix limb paths of several hundred steps with nadir paths of
// about a hundred.
__global__ void k_trace_key(TraceArgs A, int* key) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= A.nray) return;
  double xo[3], xv[3], e[3], c[3];
  geo2cart(A.obsz[r], A.obslon[r], A.obslat[r], xo);
  geo2cart(A.vpz[r], A.vplon[r], A.vplat[r], xv);
  for (int i = 0; i < 3; i++) e[i] = xv[i] - xo[i];
  const double n = norm3(e);
  for (int i = 0; i < 3; i++) e[i] /= n;
  const double sd = -dot3(xo, e);  // distance to the closest approach
  double ht = A.obsz[r];
  if (sd > 0) {
    for (int i = 0; i < 3; i++) c[i] = xo[i] + sd * e[i];
    ht = norm3(c) - RE;
  }
  int k = 0;
  if (sd > 0 && ht >= 0 && ht == ht) k = 1000 - (int)(ht < 999.0 ? ht : 999.0);
  key[r] = k;
}
void launch_trace_key(const TraceArgs& A, int* key, hipStream_t s) {
  if (A.nray > 0) hipLaunchKernelGGL(k_trace_key, dim3((A.nray + 255) / 256), dim3(256), 0, s, A, key);
}

// log(p[j + 1] / p[j]) / (z[j + 1] - z[j]) of every level interval of every
// copy (0 where the pressure interpolation is linear), the factor of the
// exponential interpolation (atmosphere.c:298-320 with EXP, jurassic.h:55-58)
// computed once per interval instead of once per evaluation.
__global__ void k_log_slopes(const double* z, const double* p, long n, double* lps) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n) return;
  const double y0 = p[i], y1 = p[i + 1];
  lps[i] = (y0 > 0 && y1 > 0) ? log(y1 / y0) / (z[i + 1] - z[i]) : 0.0;
}
void launch_log_slopes(const double* z, const double* p, long n, double* lps, hipStream_t s) {
  if (n > 1) hipLaunchKernelGGL(k_log_slopes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, z, p, n, lps);
}

void launch_raytrace(const TraceArgs& A, hipStream_t s) {
  if (A.nray <= 0) return;
  int bs = 128;
  hipLaunchKernelGGL(k_raytrace, dim3((A.nray + bs - 1) / bs), dim3(bs), 0, s, A);
  hipLaunchKernelGGL(k_records, dim3((unsigned)((A.nray + 3) / 4)), dim3(256), 0, s, A);
}
