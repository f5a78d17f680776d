// jur_retrieval.h -- the retrieval's state and measurement vectors as the
// reference composes them (state vector atmosphere.c:5-95, its inverse
// retrievalmodel.c:1103-1162, measurement vector :379-404, perturbation
// sizes :242-268), shared by kernel() and optimal_estimation().
#pragma once
#include <cmath>
#include <vector>

#include "jurassic_types.h"

struct StateVec {
  std::vector<double> x;
  std::vector<int> iqa, ipa;
  size_t n = 0;
  void add(double v, int iq, int ip) {
    x.push_back(v);
    iqa.push_back(iq);
    ipa.push_back(ip);
    n++;
  }
};

// atm2x / atm2x_help (atmosphere.c:5-95): p, T, each q, each k on the levels
// zmin <= z <= zmax of their quantity, then the first mode's N, r, sigma.
inline void state_vector(const ctl_t* ctl, const atm_t* atm, const aero_t* aero, StateVec& sv) {
  auto levels = [&](double zmin, double zmax, const double* v, int iq) {
    for (int ip = 0; ip < atm->np; ip++)
      if (atm->z[ip] >= zmin && atm->z[ip] <= zmax) sv.add(v[ip], iq, ip);
  };
  levels(ctl->retp_zmin, ctl->retp_zmax, atm->p, IDXP);
  levels(ctl->rett_zmin, ctl->rett_zmax, atm->t, IDXT);
  for (int ig = 0; ig < ctl->ng; ig++) levels(ctl->retq_zmin[ig], ctl->retq_zmax[ig], atm->q[ig], IDXQ(ig));
  for (int iw = 0; iw < ctl->nw; iw++) levels(ctl->retk_zmin[iw], ctl->retk_zmax[iw], atm->k[iw], IDXK(iw));
  if (ctl->retnn) sv.add(aero->nn[0], IDXNN, -1);
  if (ctl->retrr) sv.add(aero->rr[0], IDXRR, -1);
  if (ctl->retss) sv.add(aero->ss[0], IDXSS, -1);
}

// Perturbation size (retrievalmodel.c:242-268); <= 0 when none is defined.
inline double perturbation(const ctl_t* ctl, int iq, double x) {
  double h = -999;
  if (iq == IDXP) h = std::fmax(std::fabs(0.01 * x), 1e-7);
  if (iq == IDXT) h = 1;
  if (iq >= IDXQ(0) && iq < IDXQ(ctl->ng)) h = std::fmax(std::fabs(0.01 * x), 1e-15);
  if (iq >= IDXK(0) && iq < IDXK(ctl->nw)) h = 1e-4;
  if (iq == IDXNN) h = std::fmax(std::fabs(0.5 * x), 0.1);
  if (iq == IDXRR) h = std::fmax(std::fabs(0.25 * x), 0.1);
  if (iq == IDXSS) h = std::fmax(std::fabs(0.1 * x), 0.01);
  return h;
}

// obs2y (retrievalmodel.c:379-404): finite radiances, ray-major.
inline size_t meas_count(const ctl_t* ctl, const obs_t* obs) {
  size_t m = 0;
  for (int ir = 0; ir < obs->nr; ir++)
    for (int id = 0; id < ctl->nd; id++) m += std::isfinite(obs->rad[id][ir]) ? 1 : 0;
  return m;
}

// x2atm (retrievalmodel.c:1103-1162): write state vector x back into atm and
// the first aerosol mode, in state_vector's order.
inline void state_to_atm(const ctl_t* ctl, const double* x, atm_t* atm, aero_t* aero) {
  size_t n = 0;
  auto levels = [&](double zmin, double zmax, double* v) {
    for (int ip = 0; ip < atm->np; ip++)
      if (atm->z[ip] >= zmin && atm->z[ip] <= zmax) v[ip] = x[n++];
  };
  levels(ctl->retp_zmin, ctl->retp_zmax, atm->p);
  levels(ctl->rett_zmin, ctl->rett_zmax, atm->t);
  for (int ig = 0; ig < ctl->ng; ig++) levels(ctl->retq_zmin[ig], ctl->retq_zmax[ig], atm->q[ig]);
  for (int iw = 0; iw < ctl->nw; iw++) levels(ctl->retk_zmin[iw], ctl->retk_zmax[iw], atm->k[iw]);
  if (ctl->retnn) aero->nn[0] = x[n++];
  if (ctl->retrr) aero->rr[0] = x[n++];
  if (ctl->retss) aero->ss[0] = x[n++];
}

// obs2y (retrievalmodel.c:379-404): the finite radiances ray-major, with
// their channel and ray indices.
inline size_t meas_vector(const ctl_t* ctl, const obs_t* obs, double* y, int* ida, int* ira) {
  size_t m = 0;
  for (int ir = 0; ir < obs->nr; ir++)
    for (int id = 0; id < ctl->nd; id++)
      if (std::isfinite(obs->rad[id][ir])) {
        if (y) y[m] = obs->rad[id][ir];
        if (ida) ida[m] = id;
        if (ira) ira[m] = ir;
        m++;
      }
  return m;
}
