// scatter.cpp -- host orchestration of the scattering source (reference
// src/scatter.c:643-995 with the recursion of forwardmodel.c:224-405).
//
// The reference evaluates the source J recursively, pencil by pencil; its
// optional CPU work queue (forwardmodel.c:38-100, workqueue.c) only defers
// the leaf pencils.  Here the whole recursion tree is built level by level
// on the device (a device-side queue of ray batches): level 0 holds the
// primaries; the scattering points of level L (LOS points with aerofac > 0)
// each spawn nalpha secondary rays in level L+1 and one solar ray.  The tree
// is then integrated bottom-up: leaves (depth 0, extinction only) and solar
// rays first, then per level the angular integral gives J at every point
// and the level's rays are integrated with that source.  Large levels are
// traced and integrated in chunks that fit a memory budget; the scattering
// points of a level are emitted chunk by chunk behind a running device-side
// offset, so every ray is traced once.
//
// Two ways to run a call.  Eager: the host reads each level's counts back
// and sizes the next level exactly (and grows the LOS capacity on
// overflow).  Planned: a later call of the same shape reuses the eager
// call's sizes with a margin as capacities; every count stays on the device
// (kernels launched over a capacity read the count and leave the rest
// empty), the host waits once at the end, and a call that outgrows a
// capacity (a flag in the status word) is run again eagerly.  A planned
// call is a fixed sequence of launches, so with option "graph" it is
// recorded once as a hipGraph and replayed.
#include <cmath>
#include <cstring>
#include <string>

#include "jur_scatter.h"
#include "jur_scatter_dev.h"
#include "jur_surrogate.h"

namespace {

struct RayBuf {
  int n = 0;
  DBuf time, obsz, obslon, obslat, vpz, vplon, vplat, atm;
  void alloc(int nr) {
    n = nr;
    size_t b = (size_t)(nr > 0 ? nr : 1) * sizeof(double);
    time.need(b);
    obsz.need(b);
    obslon.need(b);
    obslat.need(b);
    vpz.need(b);
    vplon.need(b);
    vplat.need(b);
    atm.need((size_t)(nr > 0 ? nr : 1) * sizeof(int));
  }
  RayArrays arrays() {
    return {time.as<double>(), obsz.as<double>(), obslon.as<double>(), obslat.as<double>(),
            vpz.as<double>(), vplon.as<double>(), vplat.as<double>(), atm.as<int>()};
  }
  RaySetDev view() {
    RaySetDev R;
    R.nray = n;
    R.time = time.as<double>();
    R.obsz = obsz.as<double>();
    R.obslon = obslon.as<double>();
    R.obslat = obslat.as<double>();
    R.vpz = vpz.as<double>();
    R.vplon = vplon.as<double>();
    R.vplat = vplat.as<double>();
    R.atm_idx = atm.as<int>();
    return R;
  }
};

struct Level {
  RaySetDev rays;            // the level's rays (level 0: the caller's); planned: capacity
  const int* d_nray = nullptr;  // planned levels >= 1: device-side count of the valid rays
  RayBuf own;                // storage of levels >= 1
  long npt = 0;              // scattering points (eager: exact; planned: capacity)
  DBuf count, jbase, ctot, J;
  DBuf pts;
  long nsun = 0;             // solar rays (eager: exact; planned: capacity)
  DBuf sun_valid, sun_pos, sun_theta, sun_idx, d_nsun;
  RayBuf sun_tmp, sun;
  DBuf rad, tau, srad, stau, d_nchild;
  const int* d_npt = nullptr;   // device-side count of the points (in ctot)
};

struct State {
  DBuf geo, geo2, rec, npts, tsurf, gf, scan, alpha, theta_grid, bsun, phase, status, tpd;
  int nalpha = 0;  // fan size of the uploaded alpha grid
  int nd_phase = 0;
  // layer phase-function moments for the surrogate features ([nl][nd])
  DBuf asym, pfwd, pback;
  // surrogate: features, network output, hull flags, fallback compaction
  DBuf fx, fy, fflag, fpos, fsub, fpts, flnb, fnfan;
  // the two vertical pencils of every surrogate point (radiance summary
  // features): zenith angles, rays, radiances
  DBuf ud_alpha, ud_rad, ud_tau;
  RayBuf ud;
  // collected training pairs of the last call (raw features, thermal J)
  DBuf craw, cjth;
  long ncollect = 0;
  DBuf counts_h;  // device staging of the planned call's counts
  std::vector<Level> lv;
  // plan of the device-resident call: capacities from the last eager call
  struct Plan {
    std::string key;
    int cap = 0;
    std::vector<long> nray, npt;  // [S + 1] ray capacities, [S] point capacities
  } plan;
  // the planned call recorded as a graph
  std::string g_key;
  hipGraph_t g_graph = nullptr;
  hipGraphExec_t g_exec = nullptr;
};
State& ST() {
  static State s;
  return s;
}

RaySetDev sub(const RaySetDev& R, int r0, int n) {
  RaySetDev S = R;
  S.nray = n;
  S.time += r0;
  S.obsz += r0;
  S.obslon += r0;
  S.obslat += r0;
  S.vpz += r0;
  S.vplon += r0;
  S.vplat += r0;
  if (S.atm_idx) S.atm_idx += r0;
  return S;
}

// Rays per chunk under the memory budget for the trace buffers (march
// geometry, point records, their packed copy with the table brackets) at
// LOS capacity cap.
int chunk_rays(const ctl_t* ctl, long n, int cap) {
  const int R = rec_len(ctl->ng, ctl->nw);
  const size_t per = (size_t)cap * (2 * sizeof(double4) + 2 * R * sizeof(float) + ctl->ng * sizeof(int4)) + 64;
  const size_t budget = engine_chunk_bytes();
  size_t c = budget / per;
  if (c < 64) c = 64;
  return (int)(c < (size_t)n ? c : (size_t)n);
}

TraceBufs trace_bufs(State& st, int cap) {
  return TraceBufs{&st.geo, &st.geo2, &st.rec, &st.npts, &st.tsurf, nullptr, &st.gf, cap};
}

// How the current call runs: eager (host-sized, cap grown on demand) or
// planned (fixed capacities, device-side counts, flags in status).
struct Mode {
  bool planned = false;
  int cap = 0;          // planned: LOS capacity of every trace
  int* status = nullptr;
};

// Trace one chunk of rays: eager with capacity growth, planned at the fixed
// capacity against the device-side count nvalid (rays r0.. of the set).
int trace_chunk(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& C, TraceBufs& tb, double* tpz,
                double* tplon, double* tplat, const Mode& M, const int* nvalid, int r0, hipStream_t s) {
  if (!M.planned) return engine_trace(ctl, atm, C, 1, tb, tpz, tplon, tplat, s);
  engine_trace_async(ctl, atm, C, 1, tb, M.cap, tpz, tplon, tplat, M.status, nvalid, r0, s);
  return M.cap;
}

// Trace + integrate all rays of a set in chunks; outputs [nd][ld] (ld 0: n).
// nvalid (planned): device-side count of the set's valid rays.
void integrate_set(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& R, int scattering,
                   const float* J, const int* jbase, float* rad, float* tau, double* tpz, double* tplon,
                   double* tplat, hipStream_t s, const Mode& M, const int* nvalid = nullptr,
                   double* rad64 = nullptr, double* tau64 = nullptr, int ld = 0, long nJ = 0) {
  State& st = ST();
  const int n = R.nray;
  if (n <= 0) return;
  if (ld <= 0) ld = n;
  double* tpd = (double*)st.tpd.need((size_t)3 * n * sizeof(double));
  for (int r0 = 0; r0 < n;) {
    const int m = chunk_rays(ctl, n - r0, M.planned ? M.cap : engine_cap());
    TraceBufs tb = trace_bufs(st, 0);
    RaySetDev C = sub(R, r0, m);
    const int cap = trace_chunk(ctl, atm, C, tb, tpz ? tpz + r0 : tpd, tplon ? tplon + r0 : tpd + n,
                                tplat ? tplat + r0 : tpd + 2 * n, M, nvalid, r0, s);
    engine_integrate(ctl, m, cap, st.rec.as<float>(), st.npts.as<int>(), st.tsurf.as<float>(), scattering, J,
                     jbase ? jbase + r0 : nullptr, rad, tau, ld, r0, s, rad64, tau64, nJ);
    r0 += m;
  }
}

// Grow a device buffer to n bytes keeping its first `used` bytes.
void* need_keep(DBuf& b, size_t used, size_t n, hipStream_t s) {
  if (n <= b.bytes) return b.p;
  DBuf nb;
  nb.need(n);
  if (used) HIP_OK(hipMemcpyAsync(nb.p, b.p, used, hipMemcpyDeviceToDevice, s));
  HIP_OK(hipStreamSynchronize(s));
  b.release();
  b.p = nb.p;
  b.bytes = nb.bytes;
  return b.p;
}

// The fan zenith angles of srcfunc_sca_1d (scatter.c:698-715): 83..90 deg
// in 1 deg steps, then steps growing along 3, 3, 5, 8, 13, 13, ... deg
// towards 0 and 180 deg (the reference's 28 and every even count down to
// 10).  30 and 32 pencils (the cirrus configuration's 32 would leave
// 0-180 deg in that construction): the horizon band and the first four
// steps on either side (to 111 and 62 deg), then nalpha/2 - 10 equal steps
// to 180 and 0 deg.
void alpha_grid(double* alpha, int nalpha) {
  int n1 = 1, n2 = 2;
  double midang = 83, up = 92, down = 81, step = 1 + 2;
  alpha[0] = 0;
  for (int ir = nalpha / 2 - 4; ir < nalpha / 2 + 4; ++ir) {
    alpha[ir] = midang * M_PI / 180.;
    midang++;
  }
  const int nref = nalpha <= 28 ? nalpha / 2 - 5 : 5;
  for (int ir = 0; ir < nref; ++ir) {
    alpha[nalpha / 2 + ir + 4] = up * M_PI / 180.;
    alpha[nalpha / 2 - ir - 5] = down * M_PI / 180.;
    up += step;
    down -= step;
    if (step < 13) {
      step = (double)n1 + (double)n2;
      n1 = n2;
      n2 = (int)step;
    }
  }
  if (nalpha > 28) {
    const int m = nalpha / 2 - 10;
    for (int k = 1; k <= m; k++) {
      alpha[nalpha / 2 + 8 + k] = (111. + k * (180. - 111.) / (m + 1)) * M_PI / 180.;
      alpha[nalpha / 2 - 9 - k] = (62. - k * 62. / (m + 1)) * M_PI / 180.;
    }
  }
  alpha[nalpha - 1] = M_PI;
}

template <typename T>
T* upload_vec(DBuf& b, const std::vector<T>& v, hipStream_t s) {
  T* d = (T*)b.need(v.size() * sizeof(T) + sizeof(T));
  if (!v.empty()) HIP_OK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return d;
}

int read_int(const int* dptr, hipStream_t s) {
  int v = 0;
  HIP_OK(hipMemcpyAsync(&v, dptr, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  return v;
}

}  // namespace

void scatter_set_phase(const ctl_t* ctl, const aero_t* aero, hipStream_t s) {
  State& st = ST();
  const int nd = ctl->nd;
  std::vector<double> ph((size_t)NLMAX * nd * NTHETA, 0.0);
  if (aero && ctl->sca_n > 0)
    for (int il = 0; il < aero->nl && il < NLMAX; il++)
      for (int id = 0; id < nd; id++)
        for (int it = 0; it < NTHETA; it++) ph[((size_t)il * nd + id) * NTHETA + it] = aero->p[il][id][it];
  upload_vec(st.phase, ph, s);
  // asymmetry parameter g = int P cos sin / int P sin (trapezoid on the
  // 1-degree grid), forward and backward phase values
  std::vector<float> g((size_t)NLMAX * nd, 0.f), pf((size_t)NLMAX * nd, 0.f), pb((size_t)NLMAX * nd, 0.f);
  for (int il = 0; il < NLMAX; il++)
    for (int id = 0; id < nd; id++) {
      const double* q = &ph[((size_t)il * nd + id) * NTHETA];
      double num = 0, den = 0;
      for (int it = 0; it < NTHETA - 1; it++) {
        const double t0 = M_PI * it / (NTHETA - 1), t1 = M_PI * (it + 1) / (NTHETA - 1);
        num += 0.5 * (q[it] * cos(t0) * sin(t0) + q[it + 1] * cos(t1) * sin(t1)) * (t1 - t0);
        den += 0.5 * (q[it] * sin(t0) + q[it + 1] * sin(t1)) * (t1 - t0);
      }
      g[(size_t)il * nd + id] = den > 0 ? (float)(num / den) : 0.f;
      pf[(size_t)il * nd + id] = (float)q[0];
      pb[(size_t)il * nd + id] = (float)q[NTHETA - 1];
    }
  upload_vec(st.asym, g, s);
  upload_vec(st.pfwd, pf, s);
  upload_vec(st.pback, pb, s);
  std::vector<double> tg(NTHETA), bs(nd);
  for (int i = 0; i < NTHETA; i++) tg[i] = M_PI * (double)i / (NTHETA - 1);
  for (int id = 0; id < nd; id++) bs[id] = planck(TSUN, ctl->nu[id]);
  upload_vec(st.theta_grid, tg, s);
  upload_vec(st.bsun, bs, s);
  st.nd_phase = nd;
}

namespace {

// Enqueue the scattering tree of one call (eager: also size each level from
// the counts read back, and record the exact sizes in act_npt / act_nsun).
void run_tree(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, int S, const OutDev& out,
              hipStream_t s, const Mode& M, int nalpha, bool use_sur, bool collect, bool dir3d,
              std::vector<long>* act_npt, std::vector<long>* act_nsun, long* pencils_eager) {
  State& st = ST();
  const int nd = ctl->nd, Rl = rec_len(ctl->ng, ctl->nw);
  auto& lv = st.lv;
  if ((int)lv.size() < S + 1) lv.resize(S + 1);
  lv[0].rays = rays;
  lv[0].d_nray = nullptr;
  const State::Plan& P = st.plan;
  int* status = M.status;
  AeroGeom ag = engine_aero_geom();
  long pencils = rays.nray;

  // ---- build the tree top-down ------------------------------------------
  for (int L = 0; L < S; L++) {
    Level& V = lv[L];
    const long n = V.rays.nray;
    int* count = (int*)V.count.need((size_t)(n + 1) * sizeof(int));
    int* jbase = (int*)V.jbase.need((size_t)(n + 1) * sizeof(int));
    // running point offsets: ctot[c] = points of the chunks before chunk c
    int* ctot = (int*)V.ctot.need(((size_t)n / 64 + 3) * sizeof(int));
    launch_zero_int(ctot, 1, s);
    int* scantmp = (int*)st.scan.need(scan_base_temp_ints((int)(n > 0 ? n : 1)) * sizeof(int) + 256);
    long pts_cap = M.planned ? P.npt[L] : 0, used = 0;
    ScaPt* pts = (ScaPt*)V.pts.need((size_t)(pts_cap > 0 ? pts_cap : 1) * sizeof(ScaPt));
    // trace each chunk once: count its points, place them behind the
    // running offset, emit them
    int nch = 0, last_cap = 0;
    for (long r0 = 0; r0 < n; nch++) {
      const int m = chunk_rays(ctl, n - r0, M.planned ? M.cap : engine_cap());
      TraceBufs tb = trace_bufs(st, 0);
      double* tpd = (double*)st.tpd.need((size_t)3 * m * sizeof(double));
      RaySetDev C = sub(V.rays, (int)r0, m);
      last_cap = trace_chunk(ctl, atm, C, tb, tpd, tpd + m, tpd + 2 * m, M, V.d_nray, (int)r0, s);
      launch_count_points(m, st.npts.as<int>(), st.rec.as<float>(), Rl, last_cap, count + r0, s);
      launch_scan_base(count + r0, m, jbase + r0, scantmp, ctot + nch, ctot + nch + 1, s);
      if (!M.planned) {
        const long tot = read_int(ctot + nch + 1, s);
        pts = (ScaPt*)need_keep(V.pts, (size_t)used * sizeof(ScaPt), (size_t)(tot + 1) * sizeof(ScaPt), s);
        pts_cap = used = tot;
      }
      launch_emit_points(m, st.npts.as<int>(), st.rec.as<float>(), Rl, last_cap, st.geo.as<double4>(),
                         st.geo2.as<double4>(), st.gf.as<int>(), jbase + r0, C.time, C.atm_idx, pts, pts_cap,
                         status, s);
      r0 += m;
    }
    V.d_npt = ctot + nch;
    V.npt = M.planned ? P.npt[L] : used;
    if (act_npt) (*act_npt)[L] = used;
    // surrogate / training-pair collection at the primaries' points (eager)
    const ScaPt* fan_pts = pts;
    long nfan = V.npt;
    if (L == 0 && (use_sur || collect) && V.npt > 0) {
      const long nrow = (long)V.npt * nd;
      // radiance summary of every point: the extinction-only radiances of
      // its zenith and nadir pencils (the fan's first and last directions)
      const std::vector<double> ud_ang = {0.0, M_PI};
      upload_vec(st.ud_alpha, ud_ang, s);
      st.ud.alloc(2 * (int)V.npt);
      launch_children((int)V.npt, 2, pts, st.ud_alpha.as<double>(), st.ud.arrays(), s);
      float* ud_rad = (float*)st.ud_rad.need((size_t)nd * 2 * V.npt * sizeof(float));
      float* ud_tau = (float*)st.ud_tau.need((size_t)nd * 2 * V.npt * sizeof(float));
      integrate_set(ctl, atm, st.ud.view(), 0, nullptr, nullptr, ud_rad, ud_tau, nullptr, nullptr, nullptr, s, M);
      pencils += 2L * V.npt;
      FeatArgs F;
      memset(&F, 0, sizeof F);
      F.ud = ud_rad;
      F.npt = (int)V.npt;
      F.nd = nd;
      F.pts = pts;
      F.atm = {atm.z, atm.p, atm.t, atm.q, atm.k, atm.nlev, ctl->ng, ctl->nw, atm.geo};
      F.ig_h2o = engine_find_emitter(ctl, "H2O");
      F.beta_e = engine_beta_e();
      F.beta_s = engine_beta_s();
      F.asym = st.asym.as<float>();
      F.pfwd = st.pfwd.as<float>();
      F.pback = st.pback.as<float>();
      engine_layer_edges(F.top, F.bottom);
      F.nu = engine_nu_dev();
      F.margin = (float)engine_option_double("hull_margin", 0.25);
      if (collect) F.raw = (float*)st.craw.need((size_t)nrow * MLP_IN * sizeof(float));
      int* flag = (int*)st.fflag.need((size_t)(V.npt + 1) * sizeof(int));
      if (use_sur) {
        F.norm = (const MlpNorm*)mlp_norm_dev();
        F.x = (uint16_t*)st.fx.need((size_t)nrow * MLP_INP * 2);
        F.lnb = (float*)st.flnb.need((size_t)nrow * sizeof(float));
        F.out_hull = flag;
        launch_zero_int(flag, (int)V.npt, s);
      }
      launch_features(F, s);
      st.ncollect = collect ? nrow : 0;
      if (use_sur) {
        hipEvent_t m0, m1;
        HIP_OK(hipEventCreate(&m0));
        HIP_OK(hipEventCreate(&m1));
        HIP_OK(hipEventRecord(m0, s));
        mlp_infer_dev(F.x, nrow, (float*)st.fy.need((size_t)nrow * sizeof(float)), s);
        HIP_OK(hipEventRecord(m1, s));
        HIP_OK(hipEventSynchronize(m1));
        float mms = 0.f;
        HIP_OK(hipEventElapsedTime(&mms, m0, m1));
        HIP_OK(hipEventDestroy(m0));
        HIP_OK(hipEventDestroy(m1));
        engine_set_stat("mlp_ms", mms);
        engine_set_stat("mlp_samples", (double)nrow);
        int* pos = (int*)st.fpos.need((size_t)(V.npt + 1) * sizeof(int));
        int* tot = (int*)st.fnfan.need(2 * sizeof(int));
        int* tmp = (int*)st.scan.need(scan_base_temp_ints((int)V.npt) * sizeof(int) + 256);
        launch_scan_base(flag, (int)V.npt, pos, tmp, nullptr, tot, s);
        nfan = read_int(tot, s);
        ScaPt* sub_pts = (ScaPt*)st.fpts.need((size_t)(nfan + 1) * sizeof(ScaPt));
        launch_fallback_compact((int)V.npt, flag, pos, pts, (int*)st.fsub.need((size_t)(V.npt + 1) * sizeof(int)),
                                sub_pts, s);
        fan_pts = sub_pts;
        engine_set_stat("surrogate_points", (double)(V.npt - nfan));
        engine_set_stat("fallback_points", (double)nfan);
      }
    }
    // secondary fans -> level L+1 (planned: over the points' capacity, the
    // device-side count decides which rays exist)
    Level& W = lv[L + 1];
    const long nchild = nfan * (dir3d ? DIR3D_N : nalpha);
    if (nchild > 0x7fffffffL) JUR_FAIL("Too many queue items!");
    W.own.alloc((int)nchild);
    int* d_nchild = (int*)W.d_nchild.need(2 * sizeof(int));
    if (dir3d)
      launch_children3d((int)nfan, fan_pts, W.own.arrays(), s);
    else
      launch_children((int)nfan, nalpha, fan_pts, st.alpha.as<double>(), W.own.arrays(), s,
                      M.planned ? V.d_npt : nullptr, d_nchild);
    W.rays = W.own.view();
    W.d_nray = M.planned ? d_nchild : nullptr;
    pencils += nchild;
    // solar rays of the points, compacted
    const long np_ = V.npt;
    int* valid = (int*)V.sun_valid.need((size_t)(np_ + 1) * sizeof(int));
    int* pos = (int*)V.sun_pos.need((size_t)(np_ + 1) * sizeof(int));
    double* theta = (double*)V.sun_theta.need((size_t)(np_ + 1) * sizeof(double));
    int* sidx = (int*)V.sun_idx.need((size_t)(np_ + 1) * sizeof(int));
    int* d_nsun = (int*)V.d_nsun.need(2 * sizeof(int));
    V.sun_tmp.alloc((int)np_);
    SunArgs SA;
    SA.atm = {atm.z, atm.p, atm.t, atm.q, atm.k, atm.nlev, ctl->ng, ctl->nw, atm.geo};
    SA.rayds = ctl->rayds;
    SA.raydz = ctl->raydz;
    SA.refrac = ctl->refrac;
    SA.alti = ag.alti;
    SA.nalti = ag.nalti;
    SA.altimax = ag.altimax;
    SA.altimin = ag.altimin;
    SA.status = status;
    launch_sun((int)np_, pts, SA, valid, theta, V.sun_tmp.arrays(), s, M.planned ? V.d_npt : nullptr);
    int* tmp = (int*)st.scan.need(scan_base_temp_ints((int)(np_ > 0 ? np_ : 1)) * sizeof(int) + 256);
    launch_scan_base(valid, (int)np_, pos, tmp, nullptr, d_nsun, s);
    V.nsun = M.planned ? np_ : read_int(d_nsun, s);
    if (act_nsun) (*act_nsun)[L] = V.nsun;
    launch_sun_compact((int)np_, valid, pos, sidx, s);
    V.sun.alloc((int)V.nsun);
    launch_gather_rays((int)np_, valid, pos, V.sun_tmp.arrays(), V.sun.arrays(), s);
    pencils += V.nsun;
  }
  if (!M.planned && S > 0) {
    const int st_h = read_int(status, s);
    if (st_h & JUR_ST_BELOW_SURFACE) JUR_FAIL("Observer below surface!");
    if (st_h & JUR_ST_NLOS) JUR_FAIL("Too many LOS points!");
  }
  if (pencils_eager) *pencils_eager = pencils;

  // ---- integrate bottom-up ---------------------------------------------------
  {
    Level& V = lv[S];
    float* rad = (float*)V.rad.need((size_t)nd * (V.rays.nray + 1) * sizeof(float));
    float* tau = (float*)V.tau.need((size_t)nd * (V.rays.nray + 1) * sizeof(float));
    integrate_set(ctl, atm, V.rays, 0, nullptr, nullptr, rad, tau, nullptr, nullptr, nullptr, s, M, V.d_nray);
  }
  for (int L = S - 1; L >= 0; L--) {
    Level& V = lv[L];
    float* srad = (float*)V.srad.need((size_t)nd * (V.nsun + 1) * sizeof(float));
    float* stau = (float*)V.stau.need((size_t)nd * (V.nsun + 1) * sizeof(float));
    integrate_set(ctl, atm, V.sun.view(), 0, nullptr, nullptr, srad, stau, nullptr, nullptr, nullptr, s, M,
                  M.planned ? V.d_nsun.as<int>() : nullptr);
    float* J = (float*)V.J.need((size_t)(V.npt + 1) * nd * sizeof(float));
    AngArgs G;
    memset(&G, 0, sizeof G);
    G.npt = (int)V.npt;
    G.d_npt = M.planned ? V.d_npt : nullptr;
    G.nd = nd;
    G.pts = V.pts.as<ScaPt>();
    G.alpha = st.alpha.as<double>();
    G.nalpha = nalpha;
    G.phase = st.phase.as<double>();
    G.rad = lv[L + 1].rad.as<float>();
    G.nchild = lv[L + 1].rays.nray;
    G.sun_idx = V.sun_idx.as<int>();
    G.sun_theta = V.sun_theta.as<double>();
    G.tau_sun = stau;
    G.nsun = (int)V.nsun;
    G.theta_grid = st.theta_grid.as<double>();
    G.bsun = st.bsun.as<double>();
    G.J = J;
    G.ysd = 1.f;
    if (L == 0 && use_sur && V.npt > 0) {
      G.sub_idx = st.fsub.as<int>();
      G.ymlp = st.fy.as<float>();
      G.lnb = st.flnb.as<float>();
      G.ymu = mlp_norm().ymu;
      G.ysd = mlp_norm().ysd;
    }
    if (L == 0 && collect && V.npt > 0) G.jth = (float*)st.cjth.need((size_t)V.npt * nd * sizeof(float));
    if (dir3d)
      launch_angular3d(G, s);
    else
      launch_angular(G, s);
    if (L == 0) {
      integrate_set(ctl, atm, V.rays, S, J, V.jbase.as<int>(), out.rad, out.tau, out.tpz, out.tplon, out.tplat, s,
                    M, nullptr, out.rad64, out.tau64, out.ld, V.npt);
      if (out.npts) {
        // LOS points of the primaries (of the last chunk's trace when chunked)
        if (V.rays.nray <= chunk_rays(ctl, V.rays.nray, M.planned ? M.cap : engine_cap()))
          HIP_OK(hipMemcpyAsync(out.npts, st.npts.p, V.rays.nray * sizeof(int), hipMemcpyDeviceToDevice, s));
      }
    } else {
      float* rad = (float*)V.rad.need((size_t)nd * (V.rays.nray + 1) * sizeof(float));
      float* tau = (float*)V.tau.need((size_t)nd * (V.rays.nray + 1) * sizeof(float));
      integrate_set(ctl, atm, V.rays, S - L, J, V.jbase.as<int>(), rad, tau, nullptr, nullptr, nullptr, s, M,
                    V.d_nray, nullptr, nullptr, 0, V.npt);
    }
  }
}

// The plan's identity: rays, depth, fan, channels, emitters.
std::string plan_key(const ctl_t* ctl, const RaySetDev& rays, int S, int nalpha) {
  char k[256];
  snprintf(k, sizeof k, "%d|%d|%d|%d|%d|%d|%d", rays.nray, S, nalpha, ctl->nd, ctl->ng, ctl->nw, ctl->sca_n);
  return k;
}

// Everything a recorded planned call depends on: the caller's arrays, the
// engine's tables, layer geometry and buffers, and every buffer of the tree.
std::string tree_key(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, const OutDev& out, int S,
                     hipStream_t s) {
  State& st = ST();
  std::string k = engine_call_key(ctl, atm, rays, out, S, s) + "|" + st.plan.key;
  char b[64];
  auto add = [&](const void* p) {
    snprintf(b, sizeof b, "%p,", p);
    k += b;
  };
  for (const DBuf* d : {&st.geo, &st.geo2, &st.rec, &st.npts, &st.tsurf, &st.gf, &st.scan, &st.alpha,
                        &st.theta_grid, &st.bsun, &st.phase, &st.status, &st.tpd})
    add(d->p);
  for (Level& V : st.lv) {
    for (const DBuf* d : {&V.count, &V.jbase, &V.ctot, &V.J, &V.pts, &V.sun_valid, &V.sun_pos, &V.sun_theta,
                          &V.sun_idx, &V.d_nsun, &V.rad, &V.tau, &V.srad, &V.stau, &V.d_nchild})
      add(d->p);
    for (RayBuf* r : {&V.own, &V.sun_tmp, &V.sun}) {
      add(r->time.p);
      add(r->atm.p);
    }
  }
  return k;
}

void drop_tree_graph() {
  State& st = ST();
  if (st.g_exec) HIP_OK(hipGraphExecDestroy(st.g_exec));
  if (st.g_graph) HIP_OK(hipGraphDestroy(st.g_graph));
  st.g_exec = nullptr;
  st.g_graph = nullptr;
  st.g_key.clear();
}

// After a planned call (or replay) has completed: false when it outgrew a
// capacity; the reference's error exits; the call's counts as statistics.
bool planned_done(int S, long nray, hipStream_t s) {
  State& st = ST();
  std::vector<int> h(1 + 3 * S, 0);
  HIP_OK(hipMemcpyAsync(&h[0], st.status.p, sizeof(int), hipMemcpyDeviceToHost, s));
  for (int L = 0; L < S; L++) {
    Level& V = st.lv[L];
    HIP_OK(hipMemcpyAsync(&h[1 + 3 * L], V.d_npt, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(&h[2 + 3 * L], V.d_nsun.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(&h[3 + 3 * L], st.lv[L + 1].d_nchild.p, sizeof(int), hipMemcpyDeviceToHost, s));
  }
  HIP_OK(hipStreamSynchronize(s));
  if (h[0] & JUR_ST_BELOW_SURFACE) JUR_FAIL("Observer below surface!");
  if (h[0] & JUR_ST_NLOS) JUR_FAIL("Too many LOS points!");
  if (h[0] & (JUR_ST_PLAN | JUR_ST_CAP)) return false;
  long pencils = nray, leaves = 0, npts = 0;
  for (int L = 0; L < S; L++) {
    pencils += (long)h[3 + 3 * L] + h[2 + 3 * L];
    leaves += h[2 + 3 * L];
    npts += h[1 + 3 * L];
  }
  leaves += S > 0 ? h[3 + 3 * (S - 1)] : nray;
  engine_set_stat("pencils", (double)pencils);
  engine_set_stat("leaf_pencils", (double)leaves);
  engine_set_stat("scattering_points", (double)npts);
  engine_set_stat("scattering_points0", S > 0 ? (double)h[1] : 0.0);
  return true;
}

}  // namespace

void scatter_forward(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, int S, const OutDev& out,
                     hipStream_t s) {
  State& st = ST();
  // IP 2 / 3: srcfunc_sca_3d's direct fans (scatter.c:654-657)
  const bool dir3d = ctl->ip != 1;
  // surrogate for the thermal source at the primaries' points (option
  // "surrogate", once a network is trained or loaded); "collect" keeps the
  // plain method's (features, thermal J) pairs of those points; the
  // surrogate learns srcfunc_sca_1d's source: vertical profiles only
  const bool use_sur = engine_option_int("surrogate", 0) && mlp_trained() && !dir3d;
  const bool collect = engine_option_int("collect", 0) != 0 && !dir3d;
  st.ncollect = 0;
  engine_set_stat("surrogate_points", 0);
  engine_set_stat("fallback_points", 0);
  // pencils per fan (option "fan", default the reference's 28)
  const int nalpha = engine_option_int("fan", NALPHA_REF);
  if (nalpha < NALPHA_MIN || nalpha > NALPHA || nalpha % 2) JUR_FAIL("Unsupported number of secondary pencils!");
  if (st.nalpha != nalpha) {
    std::vector<double> alpha(nalpha);
    alpha_grid(alpha.data(), nalpha);
    upload_vec(st.alpha, alpha, s);
    st.nalpha = nalpha;
    drop_tree_graph();
  }
  int* status = (int*)st.status.need(2 * sizeof(int));
  // device-resident calls: option "sca_plan" (default 1), the plain method
  // on vertical profiles in FP32
  const bool plannable = !dir3d && !use_sur && !collect && engine_option_int("sca_plan", 1) &&
                         !engine_option_int("fp64", 0);
  const bool graph = plannable && engine_option_int("graph", 0);
  const std::string pkey = plan_key(ctl, rays, S, nalpha);
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  auto finish = [&](const char* how) {
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    engine_set_stat("scatter_ms", ms);
    engine_set_stat("sca_mode", how[0] == 'e' ? 0 : (how[0] == 'p' ? 1 : 2));
    HIP_OK(hipEventDestroy(e0));
    HIP_OK(hipEventDestroy(e1));
  };
  if (plannable && st.plan.key == pkey) {
    const Mode M{true, st.plan.cap, status};
    if (graph && st.g_exec && st.g_key == tree_key(ctl, atm, rays, out, S, s)) {
      HIP_OK(hipEventRecord(e0, s));
      HIP_OK(hipGraphLaunch(st.g_exec, s));
      HIP_OK(hipEventRecord(e1, s));
      if (planned_done(S, rays.nray, s)) {
        finish("graph");
        engine_set_stat("graph_ms", engine_stat("scatter_ms"));
        return;
      }
      drop_tree_graph();
    } else {
      HIP_OK(hipEventRecord(e0, s));
      launch_zero_int(status, 1, s);
      engine_set_mode(true, false);
      run_tree(ctl, atm, rays, S, out, s, M, nalpha, false, false, false, nullptr, nullptr, nullptr);
      engine_set_mode(false, false);
      HIP_OK(hipEventRecord(e1, s));
      if (planned_done(S, rays.nray, s)) {
        finish("planned");
        if (graph) {
          // record the same launch sequence for the calls that follow
          drop_tree_graph();
          engine_set_mode(true, true);
          HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
          launch_zero_int(status, 1, s);
          run_tree(ctl, atm, rays, S, out, s, M, nalpha, false, false, false, nullptr, nullptr, nullptr);
          HIP_OK(hipStreamEndCapture(s, &st.g_graph));
          engine_set_mode(false, false);
          HIP_OK(hipGraphInstantiate(&st.g_exec, st.g_graph, nullptr, nullptr, 0));
          st.g_key = tree_key(ctl, atm, rays, out, S, s);
        }
        return;
      }
    }
    // a count or a LOS outgrew the plan: run the call eagerly and plan again
    const double n_over = engine_stat("plan_overflows");
    engine_set_stat("plan_overflows", (std::isfinite(n_over) ? n_over : 0.0) + 1);
  }
  HIP_OK(hipEventRecord(e0, s));
  launch_zero_int(status, 1, s);
  std::vector<long> act_npt(S, 0), act_nsun(S, 0);
  long pencils = 0;
  run_tree(ctl, atm, rays, S, out, s, Mode{false, 0, status}, nalpha, use_sur, collect, dir3d, &act_npt, &act_nsun,
           &pencils);
  HIP_OK(hipEventRecord(e1, s));
  HIP_OK(hipEventSynchronize(e1));
  finish("eager");
  long leaves = st.lv[S].rays.nray, npts0 = 0;
  for (int L = 0; L < S; L++) {
    leaves += act_nsun[L];
    npts0 += act_npt[L];
  }
  engine_set_stat("pencils", (double)pencils);
  engine_set_stat("leaf_pencils", (double)leaves);
  engine_set_stat("scattering_points", (double)npts0);
  engine_set_stat("scattering_points0", S > 0 ? (double)act_npt[0] : 0.0);
  if (plannable) {
    // capacities of the planned calls: these sizes with a margin
    State::Plan& P = st.plan;
    P.key = pkey;
    P.cap = engine_cap();
    P.nray.assign(S + 1, 0);
    P.npt.assign(S, 0);
    P.nray[0] = rays.nray;
    for (int L = 0; L < S; L++) {
      P.npt[L] = act_npt[L] + act_npt[L] / 4 + 64;
      P.nray[L + 1] = P.npt[L] * nalpha;
    }
    drop_tree_graph();
  }
}

long scatter_collected(float* feat, float* jth, long cap) {
  State& st = ST();
  const long n = st.ncollect < cap ? st.ncollect : cap;
  if (n <= 0) return st.ncollect;
  HIP_OK(hipMemcpy(feat, st.craw.p, (size_t)n * MLP_IN * sizeof(float), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(jth, st.cjth.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
  return st.ncollect;
}
