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
// traced and integrated in chunks that fit a memory budget.
#include <cmath>
#include <cstring>

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
  RaySetDev rays;        // the level's rays (level 0: the caller's)
  RayBuf own;            // storage of levels >= 1
  int npt = 0;           // scattering points
  DBuf count, jbase, pts, J;
  int nsun = 0;
  DBuf sun_valid, sun_pos, sun_theta, sun_idx;
  RayBuf sun_tmp, sun;   // solar rays (all points, then compacted)
  DBuf rad, tau, srad, stau;
};

struct State {
  DBuf geo, geo2, rec, npts, tsurf, gf, scan, alpha, theta_grid, bsun, phase, status, tpd;
  int nalpha = 0;  // fan size of the uploaded alpha grid
  int nd_phase = 0;
  // layer phase-function moments for the surrogate features ([nl][nd])
  DBuf asym, pfwd, pback;
  // surrogate: features, network output, hull flags, fallback compaction
  DBuf fx, fy, fflag, fpos, fsub, fpts, flnb;
  // the two vertical pencils of every surrogate point (radiance summary
  // features): zenith angles, rays, radiances
  DBuf ud_alpha, ud_rad, ud_tau;
  RayBuf ud;
  // collected training pairs of the last call (raw features, thermal J)
  DBuf craw, cjth;
  long ncollect = 0;
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

// Rays per chunk under the memory budget for the trace buffers.
int chunk_rays(const ctl_t* ctl, int n) {
  const int R = rec_len(ctl->ng, ctl->nw);
  const size_t per = (size_t)engine_cap() * (2 * sizeof(double4) + R * sizeof(float)) + 64;
  const size_t budget = (size_t)engine_option_int("chunk_mb", 6144) << 20;
  size_t c = budget / per;
  if (c < 64) c = 64;
  return (int)(c < (size_t)n ? c : (size_t)n);
}

TraceBufs trace_bufs(State& st, int cap) {
  return TraceBufs{&st.geo, &st.geo2, &st.rec, &st.npts, &st.tsurf, nullptr, &st.gf, cap};
}

// Trace + integrate all rays of a set in chunks; outputs [nd][ld] (ld 0: n).
void integrate_set(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& R, int scattering,
                   const float* J, const int* jbase, float* rad, float* tau, double* tpz, double* tplon,
                   double* tplat, hipStream_t s, double* rad64 = nullptr, double* tau64 = nullptr, int ld = 0) {
  State& st = ST();
  const int n = R.nray;
  if (n <= 0) return;
  if (ld <= 0) ld = n;
  double* tpd = (double*)st.tpd.need((size_t)3 * n * sizeof(double));
  for (int r0 = 0; r0 < n;) {
    int m = chunk_rays(ctl, n - r0);
    TraceBufs tb = trace_bufs(st, 0);
    RaySetDev C = sub(R, r0, m);
    int cap = engine_trace(ctl, atm, C, 1, tb, tpz ? tpz + r0 : tpd, tplon ? tplon + r0 : tpd + n,
                           tplat ? tplat + r0 : tpd + 2 * n, s);
    engine_integrate(ctl, m, cap, st.rec.as<float>(), st.npts.as<int>(), st.tsurf.as<float>(), scattering, J,
                     jbase ? jbase + r0 : nullptr, rad, tau, ld, r0, s, rad64, tau64);
    r0 += m;
  }
}

// The 28 fan zenith angles of srcfunc_sca_1d (scatter.c:698-715): 83..90
// deg in 1 deg steps, then steps growing along 3, 3, 5, 8, 13, 13, ... deg
// towards 0 and 180 deg.
void alpha_grid(double* alpha, int nalpha) {
  int n1 = 1, n2 = 2;
  double midang = 83, up = 92, down = 81, step = 1 + 2;
  alpha[0] = 0;
  for (int ir = nalpha / 2 - 4; ir < nalpha / 2 + 4; ++ir) {
    alpha[ir] = midang * M_PI / 180.;
    midang++;
  }
  for (int ir = 0; ir < nalpha / 2 - 5; ++ir) {
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

void scatter_forward(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, int S, const OutDev& out,
                     hipStream_t s) {
  State& st = ST();
  const int nd = ctl->nd, Rl = rec_len(ctl->ng, ctl->nw);
  // IP 2 / 3: srcfunc_sca_3d's direct fans (scatter.c:654-657)
  const bool dir3d = ctl->ip != 1;
  // surrogate for the thermal source at the primaries' points (option
  // "surrogate", once a network is trained or loaded); "collect" keeps the
  // plain method's (features, thermal J) pairs of those points
  // the surrogate learns srcfunc_sca_1d's source: vertical profiles only
  const bool use_sur = engine_option_int("surrogate", 0) && mlp_trained() && !dir3d;
  const bool collect = engine_option_int("collect", 0) != 0 && !dir3d;
  st.ncollect = 0;
  engine_set_stat("surrogate_points", 0);
  engine_set_stat("fallback_points", 0);
  // pencils per fan (option "fan", default the reference's 28)
  const int nalpha = engine_option_int("fan", NALPHA);
  if (nalpha < NALPHA_MIN || nalpha > NALPHA || nalpha % 2) JUR_FAIL("Unsupported number of secondary pencils!");
  if (st.nalpha != nalpha) {
    std::vector<double> alpha(nalpha);
    alpha_grid(alpha.data(), nalpha);
    upload_vec(st.alpha, alpha, s);
    st.nalpha = nalpha;
  }
  static std::vector<Level> lv;
  if ((int)lv.size() < S + 1) lv.resize(S + 1);
  lv[0].rays = rays;
  int* status = (int*)st.status.need(sizeof(int));
  launch_zero_int(status, 1, s);
  AeroGeom ag = engine_aero_geom();
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipEventRecord(e0, s));
  long pencils = rays.nray;

  // ---- build the tree top-down ------------------------------------------
  for (int L = 0; L < S; L++) {
    Level& V = lv[L];
    const int n = V.rays.nray;
    int* count = (int*)V.count.need((size_t)(n + 1) * sizeof(int));
    int* jbase = (int*)V.jbase.need((size_t)(n + 1) * sizeof(int));
    // pass 1: trace and count the scattering points of every ray
    int nchunks = 0, last_cap = 0;
    for (int r0 = 0; r0 < n; nchunks++) {
      int m = chunk_rays(ctl, n - r0);
      TraceBufs tb = trace_bufs(st, 0);
      double* tpd = (double*)st.tpd.need((size_t)3 * m * sizeof(double));
      last_cap = engine_trace(ctl, atm, sub(V.rays, r0, m), 1, tb, tpd, tpd + m, tpd + 2 * m, s);
      launch_count_points(m, st.npts.as<int>(), st.rec.as<float>(), Rl, last_cap, count + r0, s);
      r0 += m;
    }
    if (n == 0) {
      V.npt = 0;
    } else {
      size_t tb = scan_temp_bytes(n);
      launch_exclusive_scan(count, jbase, st.scan.need(tb), tb, n, s);
      V.npt = read_int(jbase + n - 1, s) + read_int(count + n - 1, s);
    }
    ScaPt* pts = (ScaPt*)V.pts.need((size_t)(V.npt > 0 ? V.npt : 1) * sizeof(ScaPt));
    // pass 2: emit the points (re-trace unless one chunk is still in the buffers)
    for (int r0 = 0; r0 < n;) {
      int m = chunk_rays(ctl, n - r0);
      int cap = last_cap;
      if (nchunks > 1) {
        TraceBufs tb = trace_bufs(st, 0);
        double* tpd = (double*)st.tpd.need((size_t)3 * m * sizeof(double));
        cap = engine_trace(ctl, atm, sub(V.rays, r0, m), 1, tb, tpd, tpd + m, tpd + 2 * m, s);
      }
      RaySetDev C = sub(V.rays, r0, m);
      launch_emit_points(m, st.npts.as<int>(), st.rec.as<float>(), Rl, cap, st.geo.as<double4>(),
                         st.geo2.as<double4>(), st.gf.as<int>(), jbase + r0, C.time, C.atm_idx, pts, s);
      r0 += m;
    }
    // surrogate / training-pair collection at the primaries' points
    const ScaPt* fan_pts = pts;
    int nfan = V.npt;
    if (L == 0 && (use_sur || collect) && V.npt > 0) {
      const long nrow = (long)V.npt * nd;
      // radiance summary of every point: the extinction-only radiances of
      // its zenith and nadir pencils (the fan's first and last directions)
      const std::vector<double> ud_ang = {0.0, M_PI};
      upload_vec(st.ud_alpha, ud_ang, s);
      st.ud.alloc(2 * V.npt);
      launch_children(V.npt, 2, pts, st.ud_alpha.as<double>(), st.ud.arrays(), s);
      float* ud_rad = (float*)st.ud_rad.need((size_t)nd * 2 * V.npt * sizeof(float));
      float* ud_tau = (float*)st.ud_tau.need((size_t)nd * 2 * V.npt * sizeof(float));
      integrate_set(ctl, atm, st.ud.view(), 0, nullptr, nullptr, ud_rad, ud_tau, nullptr, nullptr, nullptr, s);
      pencils += 2L * V.npt;
      FeatArgs F;
      memset(&F, 0, sizeof F);
      F.ud = ud_rad;
      F.npt = V.npt;
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
        launch_zero_int(flag, V.npt, s);
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
        size_t tb = scan_temp_bytes(V.npt);
        launch_exclusive_scan(flag, pos, st.scan.need(tb), tb, V.npt, s);
        nfan = read_int(pos + V.npt - 1, s) + read_int(flag + V.npt - 1, s);
        ScaPt* sub_pts = (ScaPt*)st.fpts.need((size_t)(nfan + 1) * sizeof(ScaPt));
        launch_fallback_compact(V.npt, flag, pos, pts, (int*)st.fsub.need((size_t)(V.npt + 1) * sizeof(int)),
                                sub_pts, s);
        fan_pts = sub_pts;
        engine_set_stat("surrogate_points", (double)(V.npt - nfan));
        engine_set_stat("fallback_points", (double)nfan);
      }
    }
    // secondary fans -> level L+1
    Level& W = lv[L + 1];
    long nchild = (long)nfan * (dir3d ? DIR3D_N : nalpha);
    if (nchild > 0x7fffffffL) JUR_FAIL("Too many queue items!");
    W.own.alloc((int)nchild);
    if (dir3d)
      launch_children3d(nfan, fan_pts, W.own.arrays(), s);
    else
      launch_children(nfan, nalpha, fan_pts, st.alpha.as<double>(), W.own.arrays(), s);
    W.rays = W.own.view();
    pencils += nchild;
    // solar rays of the points
    int* valid = (int*)V.sun_valid.need((size_t)(V.npt + 1) * sizeof(int));
    int* pos = (int*)V.sun_pos.need((size_t)(V.npt + 1) * sizeof(int));
    double* theta = (double*)V.sun_theta.need((size_t)(V.npt + 1) * sizeof(double));
    int* sidx = (int*)V.sun_idx.need((size_t)(V.npt + 1) * sizeof(int));
    V.sun_tmp.alloc(V.npt);
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
    launch_sun(V.npt, pts, SA, valid, theta, V.sun_tmp.arrays(), s);
    V.nsun = 0;
    if (V.npt > 0) {
      size_t tb = scan_temp_bytes(V.npt);
      launch_exclusive_scan(valid, pos, st.scan.need(tb), tb, V.npt, s);
      V.nsun = read_int(pos + V.npt - 1, s) + read_int(valid + V.npt - 1, s);
    }
    launch_sun_compact(V.npt, valid, pos, sidx, s);
    V.sun.alloc(V.nsun);
    launch_gather_rays(V.npt, valid, pos, V.sun_tmp.arrays(), V.sun.arrays(), s);
    pencils += V.nsun;
  }
  int st_h = read_int(status, s);
  if (st_h & JUR_ST_BELOW_SURFACE) JUR_FAIL("Observer below surface!");
  if (st_h & JUR_ST_NLOS) JUR_FAIL("Too many LOS points!");

  // ---- integrate bottom-up ---------------------------------------------------
  long leaves = lv[S].rays.nray;
  {
    Level& V = lv[S];
    float* rad = (float*)V.rad.need((size_t)nd * (V.rays.nray + 1) * sizeof(float));
    float* tau = (float*)V.tau.need((size_t)nd * (V.rays.nray + 1) * sizeof(float));
    integrate_set(ctl, atm, V.rays, 0, nullptr, nullptr, rad, tau, nullptr, nullptr, nullptr, s);
  }
  for (int L = S - 1; L >= 0; L--) {
    Level& V = lv[L];
    float* srad = (float*)V.srad.need((size_t)nd * (V.nsun + 1) * sizeof(float));
    float* stau = (float*)V.stau.need((size_t)nd * (V.nsun + 1) * sizeof(float));
    integrate_set(ctl, atm, V.sun.view(), 0, nullptr, nullptr, srad, stau, nullptr, nullptr, nullptr, s);
    leaves += V.nsun;
    float* J = (float*)V.J.need((size_t)(V.npt + 1) * nd * sizeof(float));
    AngArgs G;
    G.npt = V.npt;
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
    G.nsun = V.nsun;
    G.theta_grid = st.theta_grid.as<double>();
    G.bsun = st.bsun.as<double>();
    G.J = J;
    G.sub_idx = nullptr;
    G.ymlp = nullptr;
    G.lnb = nullptr;
    G.ymu = 0.f;
    G.ysd = 1.f;
    G.jth = nullptr;
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
                    out.rad64, out.tau64, out.ld);
      if (out.npts) {
        // LOS points of the primaries (of the last chunk's trace when chunked)
        if (V.rays.nray <= chunk_rays(ctl, V.rays.nray))
          HIP_OK(hipMemcpyAsync(out.npts, st.npts.p, V.rays.nray * sizeof(int), hipMemcpyDeviceToDevice, s));
      }
    } else {
      float* rad = (float*)V.rad.need((size_t)nd * (V.rays.nray + 1) * sizeof(float));
      float* tau = (float*)V.tau.need((size_t)nd * (V.rays.nray + 1) * sizeof(float));
      integrate_set(ctl, atm, V.rays, S - L, J, V.jbase.as<int>(), rad, tau, nullptr, nullptr, nullptr, s);
    }
  }
  HIP_OK(hipEventRecord(e1, s));
  HIP_OK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  engine_set_stat("scatter_ms", ms);
  engine_set_stat("pencils", (double)pencils);
  engine_set_stat("leaf_pencils", (double)leaves);
  long npts0 = 0;
  for (int L = 0; L < S; L++) npts0 += lv[L].npt;
  engine_set_stat("scattering_points", (double)npts0);
  engine_set_stat("scattering_points0", (double)lv[0].npt);
  HIP_OK(hipEventDestroy(e0));
  HIP_OK(hipEventDestroy(e1));
}

long scatter_collected(float* feat, float* jth, long cap) {
  State& st = ST();
  const long n = st.ncollect < cap ? st.ncollect : cap;
  if (n <= 0) return st.ncollect;
  HIP_OK(hipMemcpy(feat, st.craw.p, (size_t)n * MLP_IN * sizeof(float), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(jth, st.cjth.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
  return st.ncollect;
}
