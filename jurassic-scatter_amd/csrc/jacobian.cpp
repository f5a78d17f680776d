// jacobian.cpp -- kernel() (reference src/retrievalmodel.c:201-298): the
// forward-difference Jacobian of formod.  The reference runs one formod per
// state-vector element on a perturbed copy of the atmosphere; here the
// perturbed copies are the profiles of one ray batch (each copy's rays select
// their profile), traced and integrated together, with hydrostatic balance
// applied to every copy in one launch.  Aerosol elements (N, r, sigma of the
// first mode) change the layer optics and run as one batch each.
#include <chrono>
#include <cmath>
#include <cstring>
#include <strings.h>

#include "jur_host.h"
#include "jur_kernels_misc.h"
#include "jur_retrieval.h"

namespace {

// Rays per batched forward call (a group of atmosphere copies); option
// "kernel_rays".
constexpr int kRaysPerCall = 1 << 17;

struct JacBufs {
  DBuf z, p, t, q, k, h2o, lat, p0, p1, geo[7], idx, rad, tau, rad64, tau64, tpz, tplon, tplat;
};
JacBufs& B() {
  static JacBufs b;
  return b;
}

template <typename T>
T* upload(DBuf& b, const std::vector<T>& h, hipStream_t s) {
  T* d = (T*)b.need(h.size() * sizeof(T) + sizeof(T));
  if (!h.empty()) HIP_OK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return d;
}

// One batched formod over copies of atm: copy c has state element elem[c]
// increased by step[c] (elem < 0 or an aerosol element: the atmosphere
// unchanged).  Returns the FP64 radiances [nd][C * nr] after formod's
// epilogue, read back to y.
struct Copy0 {
  std::vector<double> tau;  // [nd][nr]
  std::vector<double> tp;   // [3][nr] tangent point z, lon, lat
};
void run_copies(ctl_t* ctl, const atm_t* atm, const obs_t* obs, const aero_t* aero, const StateVec& sv,
                const std::vector<long>& elem, const std::vector<double>& step, std::vector<double>& y,
                hipStream_t s, Copy0* c0 = nullptr) {
  JacBufs& b = B();
  const int C = (int)elem.size(), nl = atm->np, nr = obs->nr, ng = ctl->ng, nw = ctl->nw, nd = ctl->nd;
  const size_t N = (size_t)C * nr;
  std::vector<double> z((size_t)C * nl), p((size_t)C * nl), t((size_t)C * nl), lat((size_t)C * nl);
  std::vector<double> q((size_t)C * (ng > 0 ? ng : 1) * nl), k((size_t)C * (nw > 0 ? nw : 1) * nl);
  int ih2o = -1;
  for (int ig = 0; ig < ng; ig++)
    if (strcasecmp(ctl->emitter[ig], "H2O") == 0) {
      ih2o = ig;
      break;
    }
  std::vector<double> h2o(ih2o >= 0 ? (size_t)C * nl : 0);
  for (int c = 0; c < C; c++) {
    double* zc = &z[(size_t)c * nl];
    double* pc = &p[(size_t)c * nl];
    double* tc = &t[(size_t)c * nl];
    double* qc = &q[(size_t)c * ng * nl];
    double* kc = &k[(size_t)c * nw * nl];
    memcpy(zc, atm->z, nl * sizeof(double));
    memcpy(pc, atm->p, nl * sizeof(double));
    memcpy(tc, atm->t, nl * sizeof(double));
    memcpy(&lat[(size_t)c * nl], atm->lat, nl * sizeof(double));
    for (int ig = 0; ig < ng; ig++) memcpy(qc + (size_t)ig * nl, atm->q[ig], nl * sizeof(double));
    for (int iw = 0; iw < nw; iw++) memcpy(kc + (size_t)iw * nl, atm->k[iw], nl * sizeof(double));
    const long j = elem[c];
    if (j >= 0 && sv.ipa[j] >= 0) {
      const int iq = sv.iqa[j], ip = sv.ipa[j];
      double* v = iq == IDXP ? pc : iq == IDXT ? tc : iq < IDXK(0) ? qc + (size_t)(iq - IDXQ(0)) * nl
                                                                   : kc + (size_t)(iq - IDXK(0)) * nl;
      v[ip] = sv.x[j] + step[c];
    }
    if (ih2o >= 0) memcpy(&h2o[(size_t)c * nl], qc + (size_t)ih2o * nl, nl * sizeof(double));
  }
  AtmSetDev A;
  A.natm = C;
  A.nlev = nl;
  double* dz = upload(b.z, z, s);
  double* dp = upload(b.p, p, s);
  double* dt = upload(b.t, t, s);
  A.z = dz;
  A.p = dp;
  A.t = dt;
  A.q = upload(b.q, q, s);
  A.k = upload(b.k, k, s);
  A.geo = stage_geo(ctl, atm, s);
  if (ctl->hydz >= 0 && nl > 0) {
    // hydrostatic (atmosphere.c:163-184) on every copy: profiles are runs of
    // equal (lon, lat)
    std::vector<int> p0, p1;
    for (int c = 0; c < C; c++) {
      int ip0 = 0;
      for (int ip = 1; ip <= nl; ip++)
        if (ip == nl || atm->lon[ip] != atm->lon[ip0] || atm->lat[ip] != atm->lat[ip0]) {
          p0.push_back(c * nl + ip0);
          p1.push_back(c * nl + ip);
          ip0 = ip;
        }
    }
    launch_hydrostatic(dz, dp, dt, ih2o >= 0 ? upload(b.h2o, h2o, s) : nullptr, upload(b.lat, lat, s),
                       upload(b.p0, p0, s), upload(b.p1, p1, s), (int)p0.size(), ctl->hydz, s);
  }
  const double* src[7] = {obs->time, obs->obsz, obs->obslon, obs->obslat, obs->vpz, obs->vplon, obs->vplat};
  double* geo[7];
  std::vector<double> rep(N);
  for (int i = 0; i < 7; i++) {
    for (int c = 0; c < C; c++) memcpy(&rep[(size_t)c * nr], src[i], nr * sizeof(double));
    geo[i] = upload(b.geo[i], rep, s);
  }
  std::vector<int> idx(N);
  for (size_t i = 0; i < N; i++) idx[i] = (int)(i / nr);
  RaySetDev R;
  R.nray = (int)N;
  R.time = geo[0];
  R.obsz = geo[1];
  R.obslon = geo[2];
  R.obslat = geo[3];
  R.vpz = geo[4];
  R.vplon = geo[5];
  R.vplat = geo[6];
  R.atm_idx = C > 1 ? upload(b.idx, idx, s) : nullptr;
  OutDev O;
  O.rad = (float*)b.rad.need(nd * N * sizeof(float) + 4);
  O.tau = (float*)b.tau.need(nd * N * sizeof(float) + 4);
  if (engine_option_int("fp64", 0)) {
    O.rad64 = (double*)b.rad64.need(nd * N * sizeof(double) + 8);
    O.tau64 = (double*)b.tau64.need(nd * N * sizeof(double) + 8);
  }
  O.tpz = (double*)b.tpz.need(N * sizeof(double) + 8);
  O.tplon = (double*)b.tplon.need(N * sizeof(double) + 8);
  O.tplat = (double*)b.tplat.need(N * sizeof(double) + 8);
  HIP_OK(hipStreamSynchronize(s));
  engine_prepare(ctl);
  engine_set_aero(ctl, ctl->sca_n > 0 ? aero : nullptr);
  // the copies' rays repeat copy 0's up to the first point a perturbation
  // reaches, and each emitter's chain and the clear-air optics up to the
  // first point whose inputs to them differ: integrated from copy 0's path
  // state and factors (option "kernel_prefix", default 1)
  engine_set_prefix(C > 1 && engine_option_int("kernel_prefix", 1) ? nr : 0);
  engine_forward(ctl, A, R, ctl->sca_mult, O, s);
  engine_set_prefix(0);
  double *rad, *tau;
  formod_epilogue(ctl, (int)N, nr, R.time, O, &rad, &tau, s);
  y.resize(nd * N);
  HIP_OK(hipMemcpyAsync(y.data(), rad, nd * N * sizeof(double), hipMemcpyDeviceToHost, s));
  if (c0) {
    c0->tau.resize((size_t)nd * nr);
    c0->tp.resize((size_t)3 * nr);
    // copy 0's transmittances: the first nr of each channel's N, one strided copy
    HIP_OK(hipMemcpy2DAsync(c0->tau.data(), nr * sizeof(double), tau, N * sizeof(double), nr * sizeof(double), nd,
                            hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(&c0->tp[0], O.tpz, nr * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(&c0->tp[nr], O.tplon, nr * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(&c0->tp[2 * (size_t)nr], O.tplat, nr * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  HIP_OK(hipStreamSynchronize(s));
}

}  // namespace

extern "C" size_t jur_state_size(ctl_t* ctl, atm_t* atm, aero_t* aero) {
  static aero_t none;
  StateVec sv;
  state_vector(ctl, atm, aero ? aero : &none, sv);
  return sv.n;
}

extern "C" size_t jur_meas_size(ctl_t* ctl, obs_t* obs) { return meas_count(ctl, obs); }

// kernel (retrievalmodel.c:201-298)
extern "C" int jur_kernel(ctl_t* ctl, atm_t* atm, obs_t* obs, aero_t* aero, double* k, size_t m, size_t n,
                          size_t ldk) {
  static aero_t none, aero1;
  static int mask[NDMAX][NRMAX];
  if (!aero) aero = &none;
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = engine_stream();
  const int nr = obs->nr, nd = ctl->nd;
  // Path integration in FP64 for the perturbed calls (option "kernel_fp64",
  // default 1): a 1 % perturbation of one level can change a radiance by
  // less than FP32 resolves.
  const double fp64_saved = engine_option_double("fp64", 0);
  const bool fp64 = engine_option_int("kernel_fp64", 1) != 0;
  // The reference starts with formod on the unperturbed state (filling obs,
  // applying hydrostatic balance to atm, refreshing retrieved aerosol
  // optics).  Here the unperturbed state is copy 0 of the first batched
  // call, with formod's side effects applied around it; formod itself runs
  // when there is no state vector or a work queue is set (its item count is
  // checked per call).
  const bool direct = n == 0 || nr <= 0 || ctl->queue.capacity > 0;
  if (direct) {
    formod(ctl, atm, obs, aero);
  } else {
    for (int id = 0; id < nd; id++)
      for (int ir = 0; ir < nr; ir++) mask[id][ir] = !std::isfinite(obs->rad[id][ir]);
    hydrostatic(ctl, atm);
    if (ctl->retnn || ctl->retrr || ctl->retss) get_opt_prop(ctl, aero);
    ctl->queue.state = Queue_Inactive;
  }
  StateVec sv;
  state_vector(ctl, atm, aero, sv);
  if (sv.n != n) JUR_FAIL("Matrix size does not match the state and measurement vectors!");
  if (n > 0 && ldk < n) JUR_FAIL("Leading dimension of the kernel matrix is too small!");
  std::vector<double> h(n);
  for (size_t j = 0; j < n; j++) {
    h[j] = perturbation(ctl, sv.iqa[j], sv.x[j]);
    if (h[j] <= 0) JUR_FAIL("Cannot set perturbation size!");
  }
  if (direct) {
    if (meas_count(ctl, obs) != m) JUR_FAIL("Matrix size does not match the state and measurement vectors!");
    for (size_t i = 0; i < m; i++)
      for (size_t j = 0; j < n; j++) k[i * ldk + j] = 0;
    if (n == 0 || m == 0) return 0;
  }
  jur_set_option("fp64", fp64 ? 1 : fp64_saved);
  size_t natm = 0;
  while (natm < n && sv.ipa[natm] >= 0) natm++;
  const int rays_per_call = engine_option_int("kernel_rays", kRaysPerCall);
  const size_t per = (size_t)std::max(2, rays_per_call / std::max(nr, 1)) - 1;
  // measurement slots of obs2y (ray-major, channels whose unperturbed
  // radiance is finite; a perturbed formod masks the others)
  std::vector<int> slot_id, slot_ir;
  std::vector<double> yy1(m, NAN), y0(m), y;
  auto set_slots = [&]() {
    slot_id.clear();
    slot_ir.clear();
    for (int ir = 0; ir < nr; ir++)
      for (int id = 0; id < nd; id++)
        if (std::isfinite(obs->rad[id][ir])) {
          slot_id.push_back(id);
          slot_ir.push_back(ir);
        }
    if (slot_id.size() != m) JUR_FAIL("Matrix size does not match the state and measurement vectors!");
    for (size_t i = 0; i < m; i++)
      for (size_t j = 0; j < n; j++) k[i * ldk + j] = 0;
  };
  if (direct) set_slots();
  // copy 0 of the first call: formod's output, masked where the caller's
  // radiances were not finite (forwardmodel.c:26-28, 112-115)
  auto take_copy0 = [&](const Copy0& c0, size_t N) {
    for (int id = 0; id < nd; id++)
      for (int ir = 0; ir < nr; ir++) {
        obs->rad[id][ir] = mask[id][ir] ? NAN : y[(size_t)id * N + ir];
        obs->tau[id][ir] = c0.tau[(size_t)id * nr + ir];
      }
    for (int ir = 0; ir < nr; ir++) {
      obs->tpz[ir] = c0.tp[ir];
      obs->tplon[ir] = c0.tp[nr + ir];
      obs->tplat[ir] = c0.tp[2 * nr + ir];
    }
    set_slots();
  };
  // obs2y of copy c of a batched call with N rays: the finite radiances among
  // the slots (a perturbed radiance that turns non-finite shifts the rest up
  // and leaves the last slots as the previous column set them, as in the
  // reference)
  auto compact = [&](size_t N, size_t c, std::vector<double>& out) {
    size_t i = 0;
    for (size_t sl = 0; sl < m; sl++) {
      double v = y[(size_t)slot_id[sl] * N + c * nr + slot_ir[sl]];
      if (std::isfinite(v)) out[i++] = v;
    }
  };
  auto column = [&](size_t j, const std::vector<double>& base) {
    for (size_t r = 0; r < m; r++) k[r * ldk + j] = (yy1[r] - base[r]) / h[j];
  };
  bool first = !direct;
  // Atmospheric elements: groups of perturbed copies, each group with its own
  // unperturbed copy 0, so every difference is taken between rays of the
  // same batched call.
  for (size_t j0 = 0; j0 < natm; j0 += per) {
    std::vector<long> elem{-1};
    std::vector<double> step{0};
    for (size_t j = j0; j < std::min(natm, j0 + per); j++) {
      elem.push_back((long)j);
      step.push_back(h[j]);
    }
    Copy0 c0;
    run_copies(ctl, atm, obs, aero, sv, elem, step, y, s, first ? &c0 : nullptr);
    const size_t N = elem.size() * (size_t)nr;
    if (first) take_copy0(c0, N);
    first = false;
    for (size_t i = 0; i < m; i++) y0[i] = y[(size_t)slot_id[i] * N + slot_ir[i]];
    for (size_t c = 1; c < elem.size(); c++) {
      compact(N, c, yy1);
      column((size_t)elem[c], y0);
    }
  }
  // Aerosol elements (first mode's N, r, sigma): fresh layer optics for the
  // perturbed parameter, one call each.  Their base is an unperturbed call of
  // the same shape (the FP64 path gives every ray the same result in any
  // batch, so the first call's copy 0 serves when there was one).
  if (natm < n && (first || !fp64 || natm == 0)) {
    Copy0 c0;
    run_copies(ctl, atm, obs, aero, sv, {-1}, {0}, y, s, first ? &c0 : nullptr);
    if (first) take_copy0(c0, (size_t)nr);
    first = false;
    for (size_t i = 0; i < m; i++) y0[i] = y[(size_t)slot_id[i] * nr + slot_ir[i]];
  }
  for (size_t j = natm; j < n; j++) {
    aero1 = *aero;
    const double v = sv.x[j] + h[j];
    if (sv.iqa[j] == IDXNN) aero1.nn[0] = v;
    if (sv.iqa[j] == IDXRR) aero1.rr[0] = v;
    if (sv.iqa[j] == IDXSS) aero1.ss[0] = v;
    get_opt_prop(ctl, &aero1);
    run_copies(ctl, atm, obs, &aero1, sv, {-1}, {0}, y, s);
    compact((size_t)nr, 0, yy1);
    column(j, y0);
  }
  jur_set_option("fp64", fp64_saved);
  engine_set_stat("kernel_ms",
                  std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  engine_set_stat("kernel_columns", (double)n);
  return 0;
}

extern "C" void kernel(ctl_t* ctl, atm_t* atm, obs_t* obs, aero_t* aero, jur_gsl_matrix* k) {
  if (!k) JUR_FAIL("Missing kernel matrix!");
  jur_kernel(ctl, atm, obs, aero, k->data, k->size1, k->size2, k->tda);
}
