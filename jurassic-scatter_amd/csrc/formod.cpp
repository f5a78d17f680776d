// formod.cpp -- the C ABI entry points: the reference's formod /
// formod_pencil / raytrace / hydrostatic on the GPU, and the batched
// extension.  Host code here only stages data and orders the device work.
#include <cmath>
#include <cstddef>
#include <cstring>
#include <strings.h>

#include "jur_host.h"
#include "jur_kernels_misc.h"

namespace {

// Device copies of one reference-style atmosphere and of a ray list.
struct Staging {
  DBuf az, ap, at, aq, ak;             // atmosphere [nlev], q [ng][nlev], k [nw][nlev]
  DBuf time, obsz, obslon, obslat, vpz, vplon, vplat;
  DBuf rad, tau, tpz, tplon, tplat, npts;
  DBuf radd, taud, radd2, taud2, nu, fdz, fw, st, rad64, tau64;
  DBuf prof0, prof1, lat;
  DBuf zll, pt, q, k, u, ds, af, ai;   // FP64 LOS dump
  DBuf bz, bp, bt, bq, bk, bidx, bo[7];  // batched host path
};
Staging& S() {
  static Staging s;
  return s;
}

template <typename T>
T* put(DBuf& b, const T* h, size_t n, hipStream_t s) {
  T* d = (T*)b.need(n * sizeof(T) + sizeof(T));
  if (n) HIP_OK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return d;
}

}  // namespace

// intpol_atm_2d / _3d set-up (atmosphere.c:339-367, 441-445): profiles are
// runs of equal (lon, lat); surface positions in Cartesian coordinates.
DevGeo stage_geo(const ctl_t* ctl, const atm_t* atm, hipStream_t s) {
  static DBuf b_idx, b_nz, b_x, b_lat;
  DevGeo G{};
  G.ip = ctl->ip;
  if (ctl->ip == 1) return G;
  if (ctl->ip != 2 && ctl->ip != 3) JUR_FAIL("Unknown interpolation method, check IP!");
  const int n = atm->np;
  std::vector<double> x;
  if (ctl->ip == 2) {
    std::vector<int> idx, nz;
    double lon1 = -999, lat1 = -999;
    for (int ip = 0; ip < n; ip++) {
      if (atm->lon[ip] != lon1 || atm->lat[ip] != lat1) {
        if ((int)idx.size() >= NPMAX) JUR_FAIL("Too many profiles!");
        lon1 = atm->lon[ip];
        lat1 = atm->lat[ip];
        idx.push_back(ip);
        nz.push_back(0);
        double c[3];
        geo2cart(0, lon1, lat1, c);
        x.insert(x.end(), c, c + 3);
      }
      nz.back()++;
    }
    for (int v : nz)
      if (v <= 1) JUR_FAIL("Cannot identify profiles. Check ordering of data points!");
    G.nx = (int)idx.size();
    G.pidx = put(b_idx, idx.data(), idx.size(), s);
    G.pnz = put(b_nz, nz.data(), nz.size(), s);
  } else {
    for (int ip = 0; ip < n; ip++) {
      double c[3];
      geo2cart(0, atm->lon[ip], atm->lat[ip], c);
      x.insert(x.end(), c, c + 3);
    }
    G.lat = put(b_lat, atm->lat, n, s);
    G.cz = ctl->cz;
    G.cx = ctl->cx;
  }
  G.px = put(b_x, x.data(), x.size(), s);
  HIP_OK(hipStreamSynchronize(s));  // the host vectors end here
  return G;
}

namespace {

AtmSetDev stage_atm(const ctl_t* ctl, const atm_t* atm, hipStream_t s) {
  Staging& g = S();
  const int n = atm->np;
  std::vector<double> q((size_t)(ctl->ng > 0 ? ctl->ng : 1) * n), k((size_t)(ctl->nw > 0 ? ctl->nw : 1) * n);
  for (int ig = 0; ig < ctl->ng; ig++) memcpy(&q[(size_t)ig * n], atm->q[ig], n * sizeof(double));
  for (int iw = 0; iw < ctl->nw; iw++) memcpy(&k[(size_t)iw * n], atm->k[iw], n * sizeof(double));
  AtmSetDev A;
  A.natm = 1;
  A.nlev = n;
  A.z = put(g.az, atm->z, n, s);
  A.p = put(g.ap, atm->p, n, s);
  A.t = put(g.at, atm->t, n, s);
  A.q = put(g.aq, q.data(), q.size(), s);
  A.k = put(g.ak, k.data(), k.size(), s);
  A.geo = stage_geo(ctl, atm, s);
  HIP_OK(hipStreamSynchronize(s));
  return A;
}

RaySetDev stage_rays(const obs_t* obs, int ir0, int nray, hipStream_t s) {
  Staging& g = S();
  RaySetDev R;
  R.nray = nray;
  R.time = put(g.time, obs->time + ir0, nray, s);
  R.obsz = put(g.obsz, obs->obsz + ir0, nray, s);
  R.obslon = put(g.obslon, obs->obslon + ir0, nray, s);
  R.obslat = put(g.obslat, obs->obslat + ir0, nray, s);
  R.vpz = put(g.vpz, obs->vpz + ir0, nray, s);
  R.vplon = put(g.vplon, obs->vplon + ir0, nray, s);
  R.vplat = put(g.vplat, obs->vplat + ir0, nray, s);
  return R;
}

OutDev stage_out(const ctl_t* ctl, int nray) {
  Staging& g = S();
  OutDev O;
  O.rad = (float*)g.rad.need((size_t)ctl->nd * nray * sizeof(float) + 4);
  O.tau = (float*)g.tau.need((size_t)ctl->nd * nray * sizeof(float) + 4);
  O.tpz = (double*)g.tpz.need(nray * sizeof(double));
  O.tplon = (double*)g.tplon.need(nray * sizeof(double));
  O.tplat = (double*)g.tplat.need(nray * sizeof(double));
  O.npts = (int*)g.npts.need(nray * sizeof(int));
  if (engine_option_int("fp64", 0)) {
    O.rad64 = (double*)g.rad64.need((size_t)ctl->nd * nray * sizeof(double) + 8);
    O.tau64 = (double*)g.tau64.need((size_t)ctl->nd * nray * sizeof(double) + 8);
  }
  return O;
}

void prepare(ctl_t* ctl, aero_t* aero) {
  engine_prepare(ctl);
  engine_set_aero(ctl, ctl->sca_n > 0 ? aero : nullptr);
}

}  // namespace

// hydrostatic (atmosphere.c:163-184): profiles are runs of equal (lon, lat).
extern "C" void hydrostatic(ctl_t* ctl, atm_t* atm) { engine_hydrostatic(ctl, atm); }

void engine_hydrostatic(ctl_t* ctl, atm_t* atm) {
  if (ctl->hydz < 0 || atm->np < 1) return;
  hipStream_t s = engine_stream();
  Staging& g = S();
  std::vector<int> p0, p1;
  double lat0 = -999, lon0 = -999;
  int ip0 = -999;
  for (int ip = 0; ip < atm->np; ip++)
    if (atm->lon[ip] != lon0 || atm->lat[ip] != lat0) {
      if (ip > 0) {
        p0.push_back(ip0);
        p1.push_back(ip);
      }
      lon0 = atm->lon[ip];
      lat0 = atm->lat[ip];
      ip0 = ip;
    }
  p0.push_back(ip0);
  p1.push_back(atm->np);
  int ig_h2o = -1;
  for (int ig = 0; ig < ctl->ng; ig++)
    if (strcasecmp(ctl->emitter[ig], "H2O") == 0) {
      ig_h2o = ig;
      break;
    }
  const int n = atm->np;
  double* z = put(g.az, atm->z, n, s);
  double* p = put(g.ap, atm->p, n, s);
  double* t = put(g.at, atm->t, n, s);
  double* q = ig_h2o >= 0 ? put(g.aq, atm->q[ig_h2o], n, s) : nullptr;
  double* lat = put(g.lat, atm->lat, n, s);
  int* d0 = put(g.prof0, p0.data(), p0.size(), s);
  int* d1 = put(g.prof1, p1.data(), p1.size(), s);
  launch_hydrostatic(z, p, t, q, lat, d0, d1, (int)p0.size(), ctl->hydz, s);
  HIP_OK(hipMemcpyAsync(atm->p, p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
}

// FP64 epilogue of formod (forwardmodel.c:101-109) on the device: FOV
// convolution over rays of the same time within each run of seg consecutive
// rays (one obs_t per run), then the brightness temperature.  Starts from
// the FP64 outputs O.rad64/tau64 when present (modified in place), else
// from the FP32 ones.
void formod_epilogue(const ctl_t* ctl, int nray, int seg, const double* time, const OutDev& O, double** rad,
                     double** tau, hipStream_t s) {
  Staging& g = S();
  const int nd = ctl->nd;
  size_t n = (size_t)nd * nray;
  double *radd = O.rad64, *taud = O.tau64;
  if (!radd || !taud) {
    radd = (double*)g.radd.need(n * sizeof(double) + 8);
    taud = (double*)g.taud.need(n * sizeof(double) + 8);
    launch_f2d(O.rad, radd, n, s);
    launch_f2d(O.tau, taud, n, s);
  }
  if (ctl->fov[0] != '-') {
    std::vector<double> dz, w;
    read_shape_file(ctl->fov, dz, w);
    double* d_dz = put(g.fdz, dz.data(), dz.size(), s);
    double* d_w = put(g.fw, w.data(), w.size(), s);
    double* r2 = (double*)g.radd2.need(n * sizeof(double) + 8);
    double* t2 = (double*)g.taud2.need(n * sizeof(double) + 8);
    int* st = (int*)g.st.need(sizeof(int));
    launch_zero_int(st, 1, s);
    launch_fov(nray, seg, nd, time, O.tpz, radd, taud, d_dz, d_w, (int)dz.size(), r2, t2, st, s);
    int sth = 0;
    HIP_OK(hipMemcpyAsync(&sth, st, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (sth & JUR_ST_FOV) JUR_FAIL("Cannot apply FOV convolution!");
    radd = r2;
    taud = t2;
  }
  if (ctl->write_bbt) {
    double* nu = put(g.nu, ctl->nu, nd, s);
    launch_brightness(nray, nd, nu, radd, s);
  }
  *rad = radd;
  *tau = taud;
}

// formod (forwardmodel.c:15-116)
extern "C" void formod(ctl_t* ctl, atm_t* atm, obs_t* obs, aero_t* aero) {
  static int mask[NDMAX][NRMAX];
  hipStream_t s = engine_stream();
  for (int id = 0; id < ctl->nd; id++)
    for (int ir = 0; ir < obs->nr; ir++) mask[id][ir] = !std::isfinite(obs->rad[id][ir]);
  hydrostatic(ctl, atm);
  if (ctl->retnn || ctl->retrr || ctl->retss) get_opt_prop(ctl, aero);
  prepare(ctl, aero);
  const int nr = obs->nr, nd = ctl->nd;
  AtmSetDev A = stage_atm(ctl, atm, s);
  RaySetDev R = stage_rays(obs, 0, nr, s);
  OutDev O = stage_out(ctl, nr);
  engine_forward(ctl, A, R, ctl->sca_mult, O, s);
  // The reference's work queue (MAX_QUEUE > 0) is released after use and its
  // capacity left at 0 (forwardmodel.c:38-40, 98-99, workqueue.c:9-29).
  if (ctl->queue.capacity > 0) {
    double items = engine_stat("leaf_pencils");
    if (std::isfinite(items) && items > ctl->queue.capacity) JUR_FAIL("Too many queue items!");
    ctl->queue.capacity = 0;
    ctl->queue.items = nullptr;
  }
  ctl->queue.state = Queue_Inactive;
  double *radd, *taud;
  formod_epilogue(ctl, nr, nr, R.time, O, &radd, &taud, s);
  // one copy per [nd][nr] device array, then a host scatter into the
  // [NDMAX][NRMAX] rows (a copy per channel costs a staging blit and a wait)
  std::vector<double> h((size_t)2 * nd * nr);
  HIP_OK(hipMemcpyAsync(h.data(), radd, (size_t)nd * nr * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(h.data() + (size_t)nd * nr, taud, (size_t)nd * nr * sizeof(double), hipMemcpyDeviceToHost,
                        s));
  HIP_OK(hipMemcpyAsync(obs->tpz, O.tpz, nr * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(obs->tplon, O.tplon, nr * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(obs->tplat, O.tplat, nr * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  for (int id = 0; id < nd; id++) {
    std::memcpy(obs->rad[id], &h[(size_t)id * nr], nr * sizeof(double));
    std::memcpy(obs->tau[id], &h[(size_t)(nd + id) * nr], nr * sizeof(double));
  }
  for (int id = 0; id < nd; id++)
    for (int ir = 0; ir < nr; ir++)
      if (mask[id][ir]) obs->rad[id][ir] = NAN;
}

// formod_pencil (forwardmodel.c:224-405), one ray, non-queue semantics.
extern "C" void formod_pencil(ctl_t* ctl, atm_t* atm, obs_t* obs, aero_t* aero, int scattering, int ir) {
  hipStream_t s = engine_stream();
  prepare(ctl, aero);
  AtmSetDev A = stage_atm(ctl, atm, s);
  RaySetDev R = stage_rays(obs, ir, 1, s);
  OutDev O = stage_out(ctl, 1);
  engine_forward(ctl, A, R, scattering, O, s);
  std::vector<float> r(ctl->nd), t(ctl->nd);
  HIP_OK(hipMemcpyAsync(r.data(), O.rad, ctl->nd * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(t.data(), O.tau, ctl->nd * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(&obs->tpz[ir], O.tpz, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(&obs->tplon[ir], O.tplon, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(&obs->tplat[ir], O.tplat, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  for (int id = 0; id < ctl->nd; id++) {
    obs->rad[id][ir] = r[id];
    obs->tau[id][ir] = t[id];
  }
}

// raytrace (lineofsight.c:5-190): trace ray ir and fill los in FP64.
extern "C" void raytrace(ctl_t* ctl, atm_t* atm, obs_t* obs, aero_t* aero, los_t* los, int ir) {
  hipStream_t s = engine_stream();
  Staging& g = S();
  engine_set_aero(ctl, ctl->sca_n > 0 ? aero : nullptr);
  AtmSetDev A = stage_atm(ctl, atm, s);
  RaySetDev R = stage_rays(obs, ir, 1, s);
  OutDev O = stage_out(ctl, 1);
  DBuf geo, geo2, rec, npts, tsurf, ts64, gf;
  TraceBufs tb{&geo, &geo2, &rec, &npts, &tsurf, &ts64, &gf, 0};
  int cap = engine_trace(ctl, A, R, ctl->sca_n > 0, tb, O.tpz, O.tplon, O.tplat, s);
  (void)cap;
  int np = 0, final2 = 0;
  double tsv = 0;
  HIP_OK(hipMemcpyAsync(&np, npts.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(&final2, gf.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(&tsv, ts64.p, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  const int R_ = rec_len(ctl->ng, ctl->nw);
  DevAtm M = {A.z, A.p, A.t, A.q, A.k, A.nlev, ctl->ng, ctl->nw, A.geo};
  const double4* G = (const double4*)(final2 ? geo2.p : geo.p);
  double* zll = (double*)g.zll.need(3 * (size_t)np * sizeof(double) + 8);
  double* pt = (double*)g.pt.need(2 * (size_t)np * sizeof(double) + 8);
  double* q = (double*)g.q.need((size_t)np * NGMAX * sizeof(double) + 8);
  double* k = (double*)g.k.need((size_t)np * NWMAX * sizeof(double) + 8);
  double* u = (double*)g.u.need((size_t)np * NGMAX * sizeof(double) + 8);
  double* ds = (double*)g.ds.need((size_t)np * sizeof(double) + 8);
  double* af = (double*)g.af.need((size_t)np * sizeof(double) + 8);
  int* ai = (int*)g.ai.need((size_t)np * sizeof(int) + 8);
  launch_los64(M, 0, G, rec.as<float>(), R_, np, ctl->ng, ctl->nw, zll, pt, q, k, u, ds, af, ai, s);
  std::vector<double> hzll(3 * (size_t)np), hpt(2 * (size_t)np), hds(np), haf(np);
  std::vector<int> hai(np);
  if (np > 0) {
    HIP_OK(hipMemcpyAsync(hzll.data(), zll, hzll.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(hpt.data(), pt, hpt.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(hds.data(), ds, hds.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(haf.data(), af, haf.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(hai.data(), ai, hai.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(los->q, q, (size_t)np * NGMAX * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(los->k, k, (size_t)np * NWMAX * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(los->u, u, (size_t)np * NGMAX * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  HIP_OK(hipMemcpyAsync(&obs->tpz[ir], O.tpz, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(&obs->tplon[ir], O.tplon, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(&obs->tplat[ir], O.tplat, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  los->np = np;
  los->tsurf = tsv;
  for (int ip = 0; ip < np; ip++) {
    los->z[ip] = hzll[3 * ip];
    los->lon[ip] = hzll[3 * ip + 1];
    los->lat[ip] = hzll[3 * ip + 2];
    los->p[ip] = hpt[2 * ip];
    los->t[ip] = hpt[2 * ip + 1];
    los->ds[ip] = hds[ip];
    los->aerofac[ip] = haf[ip];
    los->aeroi[ip] = hai[ip];
  }
  geo.release();
  geo2.release();
  rec.release();
  npts.release();
  tsurf.release();
  ts64.release();
  gf.release();
}

extern "C" double brightness(double rad, double nu) {
  // forwardmodel.c:7-11 (gsl_log1p's formula)
  double x = C1 * (nu * nu * nu) / rad;
  volatile double y = 1 + x;
  double z = y - 1;
  return C2 * nu / (log(y) - (z - x) / y);
}

extern "C" double planck(double t, double nu) {
  // forwardmodel.c:569-573 (gsl_expm1 is exp(x)-1 above ln 2, a series below)
  double x = C2 * nu / t, em1;
  if (fabs(x) < M_LN2) {
    double i = 1.0, sum = x, term = x;
    do {
      i++;
      term *= x / i;
      sum += term;
    } while (fabs(term) > fabs(sum) * 2.2204460492503131e-16);
    em1 = sum;
  } else {
    em1 = exp(x) - 1;
  }
  return C1 * (nu * nu * nu) / em1;
}

// ---- batched extension ----------------------------------------------------

// The layout the Python binding (pyjurassic/lib.py) mirrors.
static_assert(offsetof(jur_batch_t, out_ld) == 168 && sizeof(jur_batch_t) == 176, "jur_batch_t layout");

static void check_batch(const ctl_t* ctl, const jur_batch_t* b) {
  if (!b || b->nray < 0 || b->nlev < 2 || b->natm < 1) JUR_FAIL("Invalid batch description!");
  if (!b->z || !b->p || !b->t || (!b->q && ctl->ng > 0) || !b->obsz || !b->vpz || !b->rad || !b->tau)
    JUR_FAIL("Missing batch array!");
  if (b->nlev > NPMAX) JUR_FAIL("Too many data points!");
  if (b->out_ld != 0 && b->out_ld < b->nray) JUR_FAIL("Invalid batch description!");
  if (ctl->ip != 1) JUR_FAIL("The batched entry point takes vertical profiles (IP=1)!");
}

extern "C" int jur_formod_batch_device(ctl_t* ctl, const jur_batch_t* b, const aero_t* aero, int scattering,
                                       void* stream) {
  check_batch(ctl, b);
  hipStream_t s = stream ? (hipStream_t)stream : engine_stream();
  engine_prepare(ctl);
  engine_set_aero(ctl, ctl->sca_n > 0 ? aero : nullptr);
  AtmSetDev A;
  A.natm = b->natm;
  A.nlev = b->nlev;
  A.z = b->z;
  A.p = b->p;
  A.t = b->t;
  A.q = b->q;
  A.k = b->k;
  static DBuf kzero, idx;
  if (!A.k && ctl->nw > 0) {
    size_t n = (size_t)b->natm * ctl->nw * b->nlev;
    double* kz = (double*)kzero.need(n * sizeof(double) + 8);
    launch_fill_f64(kz, n, 0.0, s);  // a kernel, not a memset node, should the caller capture this call
    A.k = kz;
  }
  RaySetDev R;
  R.nray = b->nray;
  R.time = b->time;
  R.obsz = b->obsz;
  R.obslon = b->obslon;
  R.obslat = b->obslat;
  R.vpz = b->vpz;
  R.vplon = b->vplon;
  R.vplat = b->vplat;
  R.atm_idx = b->atm_of_ray;
  if (!R.atm_idx && b->natm == b->nray && b->natm > 1) {
    int* d = (int*)idx.need((size_t)b->nray * sizeof(int) + 8);
    launch_iota(d, b->nray, s);
    R.atm_idx = d;
  }
  OutDev O;
  O.rad = b->rad;
  O.tau = b->tau;
  O.ld = b->out_ld;
  static DBuf tz, tlo, tla;
  O.tpz = b->tpz ? b->tpz : (double*)tz.need((size_t)b->nray * sizeof(double) + 8);
  O.tplon = b->tplon ? b->tplon : (double*)tlo.need((size_t)b->nray * sizeof(double) + 8);
  O.tplat = b->tplat ? b->tplat : (double*)tla.need((size_t)b->nray * sizeof(double) + 8);
  O.npts = b->npts;
  engine_forward(ctl, A, R, scattering, O, s);
  return 0;
}

extern "C" int jur_formod_batch(ctl_t* ctl, const jur_batch_t* b, const aero_t* aero, int scattering) {
  check_batch(ctl, b);
  hipStream_t s = engine_stream();
  Staging& g = S();
  const size_t na = b->natm, nl = b->nlev, nr = b->nray, nd = ctl->nd;
  jur_batch_t d = *b;
  d.z = put(g.bz, b->z, na * nl, s);
  d.p = put(g.bp, b->p, na * nl, s);
  d.t = put(g.bt, b->t, na * nl, s);
  d.q = ctl->ng > 0 ? put(g.bq, b->q, na * ctl->ng * nl, s) : nullptr;
  d.k = (b->k && ctl->nw > 0) ? put(g.bk, b->k, na * ctl->nw * nl, s) : nullptr;
  d.atm_of_ray = b->atm_of_ray ? put(g.bidx, b->atm_of_ray, nr, s) : nullptr;
  const double* src[7] = {b->time, b->obsz, b->obslon, b->obslat, b->vpz, b->vplon, b->vplat};
  const double* dst[7];
  std::vector<double> zeros(nr, 0.0);
  for (int i = 0; i < 7; i++) dst[i] = put(g.bo[i], src[i] ? src[i] : zeros.data(), nr, s);
  d.time = dst[0];
  d.obsz = dst[1];
  d.obslon = dst[2];
  d.obslat = dst[3];
  d.vpz = dst[4];
  d.vplon = dst[5];
  d.vplat = dst[6];
  OutDev O = stage_out(ctl, (int)nr);
  d.rad = O.rad;
  d.tau = O.tau;
  d.out_ld = 0;  // dense on the device; the copies back honour b->out_ld
  d.tpz = O.tpz;
  d.tplon = O.tplon;
  d.tplat = O.tplat;
  d.npts = O.npts;
  HIP_OK(hipStreamSynchronize(s));
  jur_formod_batch_device(ctl, &d, aero, scattering, s);
  if (b->out_ld > 0 && (size_t)b->out_ld != nr && nr > 0) {
    const size_t ld = b->out_ld;
    HIP_OK(hipMemcpy2DAsync(b->rad, ld * sizeof(float), d.rad, nr * sizeof(float), nr * sizeof(float), nd,
                            hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpy2DAsync(b->tau, ld * sizeof(float), d.tau, nr * sizeof(float), nr * sizeof(float), nd,
                            hipMemcpyDeviceToHost, s));
  } else {
    HIP_OK(hipMemcpyAsync(b->rad, d.rad, nd * nr * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(b->tau, d.tau, nd * nr * sizeof(float), hipMemcpyDeviceToHost, s));
  }
  if (b->tpz) HIP_OK(hipMemcpyAsync(b->tpz, d.tpz, nr * sizeof(double), hipMemcpyDeviceToHost, s));
  if (b->tplon) HIP_OK(hipMemcpyAsync(b->tplon, d.tplon, nr * sizeof(double), hipMemcpyDeviceToHost, s));
  if (b->tplat) HIP_OK(hipMemcpyAsync(b->tplat, d.tplat, nr * sizeof(double), hipMemcpyDeviceToHost, s));
  if (b->npts) HIP_OK(hipMemcpyAsync(b->npts, d.npts, nr * sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  return 0;
}
