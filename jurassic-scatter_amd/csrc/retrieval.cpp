// retrieval.cpp -- optimal_estimation (reference src/retrievalmodel.c:407-706):
// the Levenberg-Marquardt optimal-estimation retrieval with its a priori and
// measurement covariances, cost function, error analysis and outputs, plus
// read_ret (:710-757).  Forward model and Jacobian run on the GPU (formod,
// jur_kernel); the dense linear algebra (K^T S K products, gain and
// averaging-kernel matrices, Cholesky solves and inverses) runs on the GPU
// with rocBLAS / rocSOLVER on host-resident row-major matrices.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "jur_host.h"
#include "jur_retrieval.h"

namespace {

#define BLAS_OK(cmd)                                              \
  do {                                                            \
    rocblas_status st_ = (cmd);                                   \
    if (st_ != rocblas_status_success) {                          \
      char m_[256];                                               \
      snprintf(m_, sizeof m_, "rocBLAS error %d: %s", (int)st_, #cmd); \
      JUR_FAIL(m_);                                               \
    }                                                             \
  } while (0)

struct LinAlg {
  rocblas_handle h = nullptr;
  DBuf a, b, c, info;
};
LinAlg& la() {
  static LinAlg L;
  if (!L.h) {
    BLAS_OK(rocblas_create_handle(&L.h));
    BLAS_OK(rocblas_set_stream(L.h, engine_stream()));
  }
  return L;
}

double* put(DBuf& b, const double* h, size_t n) {
  double* d = (double*)b.need(n * sizeof(double) + 8);
  HIP_OK(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, engine_stream()));
  return d;
}
void get(double* h, const double* d, size_t n) {
  HIP_OK(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, engine_stream()));
  HIP_OK(hipStreamSynchronize(engine_stream()));
}

// C[M][N] = op(A) op(B) for row-major host matrices (A is M x K, or K x M
// when ta; B is K x N, or N x K when tb).  Column-major rocBLAS sees each
// row-major matrix transposed, so it computes C^T = op(B)^T op(A)^T.
void gemm(bool ta, bool tb, int M, int N, int K, const double* A, const double* B, double* C) {
  LinAlg& L = la();
  const int lda = ta ? M : K, ldb = tb ? K : N;
  double* dA = put(L.a, A, (size_t)M * K);
  double* dB = put(L.b, B, (size_t)K * N);
  double* dC = (double*)L.c.need((size_t)M * N * sizeof(double) + 8);
  const double one = 1.0, zero = 0.0;
  BLAS_OK(rocblas_dgemm(L.h, tb ? rocblas_operation_transpose : rocblas_operation_none,
                        ta ? rocblas_operation_transpose : rocblas_operation_none, N, M, K, &one, dB, ldb, dA, lda,
                        &zero, dC, N));
  get(C, dC, (size_t)M * N);
}

// Cholesky factorisation of the symmetric positive definite n x n matrix on
// the device (lower triangle in column-major = upper in row-major).
void potrf(double* dA, int n) {
  LinAlg& L = la();
  rocblas_int* info = (rocblas_int*)L.info.need(sizeof(rocblas_int));
  BLAS_OK(rocsolver_dpotrf(L.h, rocblas_fill_lower, n, dA, n, info));
  rocblas_int hinfo = 0;
  HIP_OK(hipMemcpyAsync(&hinfo, info, sizeof hinfo, hipMemcpyDeviceToHost, engine_stream()));
  HIP_OK(hipStreamSynchronize(engine_stream()));
  if (hinfo != 0) JUR_FAIL("Matrix is not positive definite!");
}

// Solve A x = b (A symmetric positive definite, n x n row-major) by Cholesky
// decomposition (gsl_linalg_cholesky_decomp / _solve).
void chol_solve(const double* A, const double* b, double* x, int n) {
  LinAlg& L = la();
  double* dA = put(L.a, A, (size_t)n * n);
  double* dB = put(L.b, b, (size_t)n);
  potrf(dA, n);
  BLAS_OK(rocsolver_dpotrs(L.h, rocblas_fill_lower, n, 1, dA, n, dB, n));
  get(x, dB, (size_t)n);
}

// matrix_invert (retrievalmodel.c:302-327): a diagonal matrix by its
// reciprocals, else by Cholesky decomposition and inversion.
void matrix_invert(std::vector<double>& a, int n) {
  bool diag = true;
  for (int i = 0; i < n && diag; i++)
    for (int j = i + 1; j < n; j++)
      if (a[(size_t)i * n + j] != 0) {
        diag = false;
        break;
      }
  if (diag) {
    for (int i = 0; i < n; i++) a[(size_t)i * n + i] = 1 / a[(size_t)i * n + i];
    return;
  }
  LinAlg& L = la();
  double* dA = put(L.a, a.data(), (size_t)n * n);
  potrf(dA, n);
  rocblas_int* info = (rocblas_int*)L.info.need(sizeof(rocblas_int));
  BLAS_OK(rocsolver_dpotri(L.h, rocblas_fill_lower, n, dA, n, info));
  get(a.data(), dA, (size_t)n * n);
  // column-major lower = row-major upper triangle holds the inverse
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++) a[(size_t)i * n + j] = a[(size_t)j * n + i];
}

// matrix_product (retrievalmodel.c:331-375): A^T B A (transpose 1) or A B A^T
// (transpose 2) for the m x n matrix A and the diagonal B = diag(b^2)
// given as b.
void matrix_product(const std::vector<double>& a, int m, int n, const std::vector<double>& b, int transpose,
                    std::vector<double>& c) {
  std::vector<double> aux((size_t)m * n);
  if (transpose == 1) {
    for (int i = 0; i < m; i++)
      for (int j = 0; j < n; j++) aux[(size_t)i * n + j] = b[i] * a[(size_t)i * n + j];
    c.assign((size_t)n * n, 0.0);
    gemm(true, false, n, n, m, aux.data(), aux.data(), c.data());
  } else {
    for (int i = 0; i < m; i++)
      for (int j = 0; j < n; j++) aux[(size_t)i * n + j] = a[(size_t)i * n + j] * b[j];
    c.assign((size_t)m * m, 0.0);
    gemm(false, true, m, m, n, aux.data(), aux.data(), c.data());
  }
}

// copy_atm / copy_obs (atmosphere.c:98-135, misc.c:20-57): the used points,
// optionally with the values cleared.
void copy_atm(const ctl_t* ctl, atm_t* d, const atm_t* s, int init) {
  const size_t n = (size_t)s->np * sizeof(double);
  d->np = s->np;
  memcpy(d->time, s->time, n);
  memcpy(d->z, s->z, n);
  memcpy(d->lon, s->lon, n);
  memcpy(d->lat, s->lat, n);
  memcpy(d->p, s->p, n);
  memcpy(d->t, s->t, n);
  for (int ig = 0; ig < ctl->ng; ig++) memcpy(d->q[ig], s->q[ig], n);
  for (int iw = 0; iw < ctl->nw; iw++) memcpy(d->k[iw], s->k[iw], n);
  d->init = s->init;
  if (init)
    for (int ip = 0; ip < d->np; ip++) {
      d->p[ip] = d->t[ip] = 0;
      for (int ig = 0; ig < ctl->ng; ig++) d->q[ig][ip] = 0;
      for (int iw = 0; iw < ctl->nw; iw++) d->k[iw][ip] = 0;
    }
}
void copy_obs(const ctl_t* ctl, obs_t* d, const obs_t* s, int init) {
  const size_t n = (size_t)s->nr * sizeof(double);
  d->nr = s->nr;
  memcpy(d->time, s->time, n);
  memcpy(d->obsz, s->obsz, n);
  memcpy(d->obslon, s->obslon, n);
  memcpy(d->obslat, s->obslat, n);
  memcpy(d->vpz, s->vpz, n);
  memcpy(d->vplon, s->vplon, n);
  memcpy(d->vplat, s->vplat, n);
  memcpy(d->tpz, s->tpz, n);
  memcpy(d->tplon, s->tplon, n);
  memcpy(d->tplat, s->tplat, n);
  for (int id = 0; id < ctl->nd; id++) {
    memcpy(d->rad[id], s->rad[id], n);
    memcpy(d->tau[id], s->tau[id], n);
  }
  if (init)
    for (int ir = 0; ir < d->nr; ir++)
      for (int id = 0; id < ctl->nd; id++)
        if (std::isfinite(d->rad[id][ir])) d->rad[id][ir] = d->tau[id][ir] = 0;
}

// idx2name (retrievalmodel.c:171-197)
void idx2name(const ctl_t* ctl, int idx, char* q) {
  if (idx == IDXP) sprintf(q, "PRESSURE");
  if (idx == IDXT) sprintf(q, "TEMPERATURE");
  for (int ig = 0; ig < ctl->ng; ig++)
    if (idx == IDXQ(ig)) sprintf(q, "%s", ctl->emitter[ig]);
  for (int iw = 0; iw < ctl->nw; iw++)
    if (idx == IDXK(iw)) sprintf(q, "EXTINCT_WINDOW%d", iw);
  if (idx == IDXNN) sprintf(q, "CONCENTRATION");
  if (idx == IDXRR) sprintf(q, "RADIUS");
  if (idx == IDXSS) sprintf(q, "PSD_WIDTH");
}

// write_matrix (retrievalmodel.c:946-1099), written only with WRITE_MATRIX.
void write_matrix(const char* dirname, const char* filename, const ctl_t* ctl, const std::vector<double>& mat,
                  const atm_t* atm, const aero_t* aero, const obs_t* obs, const char* rowspace,
                  const char* colspace, const char* sort) {
  if (!ctl->write_matrix) return;
  char file[2 * LEN + 2], quantity[LEN];
  if (dirname)
    snprintf(file, sizeof file, "%s/%s", dirname, filename);
  else
    snprintf(file, sizeof file, "%s", filename);
  printf("Write matrix: %s\n", file);
  FILE* out = fopen(file, "w");
  if (!out) JUR_FAIL("Cannot create file!");
  std::vector<int> rid(MMAX), rir(MMAX), cid(MMAX), cir(MMAX);
  StateVec rs, cs;
  size_t nr, nc;
  if (rowspace[0] == 'y') {
    fprintf(out,
            "# $1 = Row: index (measurement space)\n"
            "# $2 = Row: channel wavenumber [cm^-1]\n"
            "# $3 = Row: time (seconds since 2000-01-01T00:00Z)\n"
            "# $4 = Row: view point altitude [km]\n"
            "# $5 = Row: view point longitude [deg]\n"
            "# $6 = Row: view point latitude [deg]\n");
    nr = meas_vector(ctl, obs, nullptr, rid.data(), rir.data());
  } else {
    fprintf(out,
            "# $1 = Row: index (state space)\n"
            "# $2 = Row: name of quantity\n"
            "# $3 = Row: time (seconds since 2000-01-01T00:00Z)\n"
            "# $4 = Row: altitude [km]\n"
            "# $5 = Row: longitude [deg]\n"
            "# $6 = Row: latitude [deg]\n");
    state_vector(ctl, atm, aero, rs);
    nr = rs.n;
  }
  if (colspace[0] == 'y') {
    fprintf(out,
            "# $7 = Col: index (measurement space)\n"
            "# $8 = Col: channel wavenumber [cm^-1]\n"
            "# $9 = Col: time (seconds since 2000-01-01T00:00Z)\n"
            "# $10 = Col: view point altitude [km]\n"
            "# $11 = Col: view point longitude [deg]\n"
            "# $12 = Col: view point latitude [deg]\n");
    nc = meas_vector(ctl, obs, nullptr, cid.data(), cir.data());
  } else {
    fprintf(out,
            "# $7 = Col: index (state space)\n"
            "# $8 = Col: name of quantity\n"
            "# $9 = Col: time (seconds since 2000-01-01T00:00Z)\n"
            "# $10 = Col: altitude [km]\n"
            "# $11 = Col: longitude [deg]\n"
            "# $12 = Col: latitude [deg]\n");
    state_vector(ctl, atm, aero, cs);
    nc = cs.n;
  }
  fprintf(out, "# $13 = Matrix element\n\n");
  // aerosol elements have no level; their rows name level 0
  auto lev = [](int ip) { return ip < 0 ? 0 : ip; };
  size_t i = 0, j = 0;
  while (i < nr && j < nc) {
    const double v = mat[i * nc + j];
    if (v != 0) {
      if (rowspace[0] == 'y') {
        fprintf(out, "%d %g %.2f %g %g %g", (int)i, ctl->nu[rid[i]], obs->time[rir[i]], obs->vpz[rir[i]],
                obs->vplon[rir[i]], obs->vplat[rir[i]]);
      } else {
        idx2name(ctl, rs.iqa[i], quantity);
        const int ip = lev(rs.ipa[i]);
        fprintf(out, "%d %s %.2f %g %g %g", (int)i, quantity, atm->time[ip], atm->z[ip], atm->lon[ip],
                atm->lat[ip]);
      }
      if (colspace[0] == 'y') {
        fprintf(out, " %d %g %.2f %g %g %g", (int)j, ctl->nu[cid[j]], obs->time[cir[j]], obs->vpz[cir[j]],
                obs->vplon[cir[j]], obs->vplat[cir[j]]);
      } else {
        idx2name(ctl, cs.iqa[j], quantity);
        const int ip = lev(cs.ipa[j]);
        fprintf(out, " %d %s %.2f %g %g %g", (int)j, quantity, atm->time[ip], atm->z[ip], atm->lon[ip],
                atm->lat[ip]);
      }
      fprintf(out, " %g\n", v);
    }
    if (sort[0] == 'r') {
      if (++j >= nc) {
        j = 0;
        i++;
        fprintf(out, "\n");
      }
    } else {
      if (++i >= nr) {
        i = 0;
        j++;
        fprintf(out, "\n");
      }
    }
  }
  fclose(out);
}

// set_cov_apr (retrievalmodel.c:761-857)
void set_cov_apr(const ret_t* ret, const ctl_t* ctl, const atm_t* atm, const aero_t* aero, const StateVec& sv,
                 std::vector<double>& s_a) {
  const size_t n = sv.n;
  StateVec xa;
  state_vector(ctl, atm, aero, xa);
  std::vector<double> x = xa.x;
  for (size_t i = 0; i < n; i++) {
    const int iq = sv.iqa[i];
    if (iq == IDXP) x[i] = ret->err_press / 100 * x[i];
    if (iq == IDXT) x[i] = ret->err_temp;
    for (int ig = 0; ig < ctl->ng; ig++)
      if (iq == IDXQ(ig)) x[i] = ret->err_q[ig] / 100 * x[i];
    for (int iw = 0; iw < ctl->nw; iw++)
      if (iq == IDXK(iw)) x[i] = ret->err_k[iw];
    if (iq == IDXNN) x[i] = ret->err_nn;
    if (iq == IDXRR) x[i] = ret->err_rr;
    if (iq == IDXSS) x[i] = ret->err_ss;
  }
  s_a.assign(n * n, 0.0);
  for (size_t i = 0; i < n; i++) s_a[i * n + i] = x[i] * x[i];
  std::vector<double> xs(3 * (size_t)(atm->np > 0 ? atm->np : 1));
  for (int ip = 0; ip < atm->np; ip++) geo2cart(0, atm->lon[ip], atm->lat[ip], &xs[3 * (size_t)ip]);
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < n; j++)
      if (i != j && sv.iqa[i] == sv.iqa[j]) {
        const int iq = sv.iqa[i];
        double cz = 0, ch = 0;
        if (iq == IDXP) {
          cz = ret->err_press_cz;
          ch = ret->err_press_ch;
        }
        if (iq == IDXT) {
          cz = ret->err_temp_cz;
          ch = ret->err_temp_ch;
        }
        for (int ig = 0; ig < ctl->ng; ig++)
          if (iq == IDXQ(ig)) {
            cz = ret->err_q_cz[ig];
            ch = ret->err_q_ch[ig];
          }
        for (int iw = 0; iw < ctl->nw; iw++)
          if (iq == IDXK(iw)) {
            cz = ret->err_k_cz[iw];
            ch = ret->err_k_ch[iw];
          }
        if (cz > 0 && ch > 0) {
          // corr_function (retrievalmodel.c:91-108): exp(-d_h/ch - |dz|/cz),
          // d_h the chord between the surface points (computed once per level)
          const int a = sv.ipa[i], b = sv.ipa[j];
          const double* xa = &xs[3 * (size_t)a];
          const double* xb = &xs[3 * (size_t)b];
          const double d = std::sqrt((xa[0] - xb[0]) * (xa[0] - xb[0]) + (xa[1] - xb[1]) * (xa[1] - xb[1]) +
                                     (xa[2] - xb[2]) * (xa[2] - xb[2]));
          s_a[i * n + j] = x[i] * x[j] * std::exp(-d / ch - std::fabs(atm->z[a] - atm->z[b]) / cz);
        }
      }
}

// set_cov_meas (retrievalmodel.c:861-898)
void set_cov_meas(const ret_t* ret, const ctl_t* ctl, const obs_t* obs, std::vector<double>& sig_noise,
                  std::vector<double>& sig_formod, std::vector<double>& sig_eps_inv) {
  static obs_t err;
  const size_t m = sig_eps_inv.size();
  copy_obs(ctl, &err, obs, 1);
  for (int ir = 0; ir < err.nr; ir++)
    for (int id = 0; id < ctl->nd; id++) err.rad[id][ir] = std::isfinite(obs->rad[id][ir]) ? ret->err_noise[id] : NAN;
  sig_noise.assign(m, 0.0);
  meas_vector(ctl, &err, sig_noise.data(), nullptr, nullptr);
  copy_obs(ctl, &err, obs, 1);
  for (int ir = 0; ir < err.nr; ir++)
    for (int id = 0; id < ctl->nd; id++) err.rad[id][ir] = std::fabs(ret->err_formod[id] / 100 * obs->rad[id][ir]);
  sig_formod.assign(m, 0.0);
  meas_vector(ctl, &err, sig_formod.data(), nullptr, nullptr);
  for (size_t i = 0; i < m; i++)
    sig_eps_inv[i] = 1 / std::sqrt(sig_noise[i] * sig_noise[i] + sig_formod[i] * sig_formod[i]);
}

// cost_function (retrievalmodel.c:112-167): chi^2/m of the measurement and
// a priori terms, printed and appended to the cost file.
double cost_function(FILE* out, int it, const std::vector<double>& dx, const std::vector<double>& dy,
                     const std::vector<double>& s_a_inv, const std::vector<double>& sig_eps_inv) {
  const size_t m = dy.size(), n = dx.size();
  double chisq_m = 0, chisq_a = 0;
  for (size_t i = 0; i < m; i++) chisq_m += (dy[i] * sig_eps_inv[i]) * (dy[i] * sig_eps_inv[i]);
  chisq_m /= (double)m;
  for (size_t i = 0; i < n; i++) {
    double s = 0;
    for (size_t j = 0; j < n; j++) s += s_a_inv[i * n + j] * dx[j];
    chisq_a += dx[i] * s;
  }
  chisq_a /= (double)m;
  const double chisq = chisq_m + chisq_a;
  printf("it= %d / chi^2/m= %g (meas: %g / apr: %g)\n", it, chisq, chisq_m, chisq_a);
  if (it == 0)
    fprintf(out,
            "# $1 = iteration number\n"
            "# $2 = normalized cost function: total\n"
            "# $3 = normalized cost function: measurements\n"
            "# $4 = normalized cost function: a priori\n"
            "# $5 = number of measurements\n"
            "# $6 = number of state vector elements\n\n");
  fprintf(out, "%d %g %g %g %d %d\n", it, chisq, chisq_m, chisq_a, (int)m, (int)n);
  return chisq;
}

// write_stddev (retrievalmodel.c:902-942)
void write_stddev(const char* quantity, const ret_t* ret, const ctl_t* ctl, const atm_t* atm, const aero_t* aero,
                  const std::vector<double>& s, size_t n) {
  static atm_t atm_aux;
  static aero_t aero_aux;
  std::vector<double> x(n);
  for (size_t i = 0; i < n; i++) x[i] = std::sqrt(s[i * n + i]);
  copy_atm(ctl, &atm_aux, atm, 1);
  copy_aero((ctl_t*)ctl, &aero_aux, (aero_t*)aero, 1);
  state_to_atm(ctl, x.data(), &atm_aux, &aero_aux);
  char filename[LEN];
  snprintf(filename, sizeof filename, "atm_err_%s.tab", quantity);
  write_atm(ret->dir, filename, (ctl_t*)ctl, &atm_aux);
  if (ctl->retnn || ctl->retrr || ctl->retss) {
    snprintf(filename, sizeof filename, "aero_err_%s.tab", quantity);
    write_aero(ret->dir, filename, &aero_aux);
  }
}

// analyze_avk / analyze_avk_quantity (retrievalmodel.c:5-87): area of the
// averaging kernel (atm_cont.tab) and information density 1 / A_ii
// (atm_res.tab) per retrieved quantity.
void analyze_avk(const ret_t* ret, const ctl_t* ctl, const atm_t* atm, const aero_t* aero, const StateVec& sv,
                 const std::vector<double>& avk) {
  static atm_t cont, res;
  static aero_t acont, ares;
  const size_t n = sv.n;
  copy_atm(ctl, &cont, atm, 1);
  copy_atm(ctl, &res, atm, 1);
  copy_aero((ctl_t*)ctl, &acont, (aero_t*)aero, 1);
  copy_aero((ctl_t*)ctl, &ares, (aero_t*)aero, 1);
  auto quantity = [&](int iq, double* c, double* r) {
    size_t n0 = NMAX, n1 = 0;
    for (size_t i = 0; i < n; i++) {
      if (sv.iqa[i] == iq && n0 == NMAX) n0 = i;
      if (sv.iqa[i] == iq) n1 = i - n0 + 1;
    }
    if (n0 >= NMAX) return;
    for (size_t i = 0; i < n1; i++) {
      const int ip = sv.ipa[n0 + i] < 0 ? 0 : sv.ipa[n0 + i];
      for (size_t j = 0; j < n1; j++) c[ip] += avk[(n0 + i) * n + n0 + j];
      r[ip] = 1 / avk[(n0 + i) * n + n0 + i];
    }
  };
  quantity(IDXP, cont.p, res.p);
  quantity(IDXT, cont.t, res.t);
  quantity(IDXNN, acont.nn, ares.nn);
  quantity(IDXRR, acont.rr, ares.rr);
  quantity(IDXSS, acont.ss, ares.ss);
  for (int ig = 0; ig < ctl->ng; ig++) quantity(IDXQ(ig), cont.q[ig], res.q[ig]);
  for (int iw = 0; iw < ctl->nw; iw++) quantity(IDXK(iw), cont.k[iw], res.k[iw]);
  write_atm(ret->dir, "atm_cont.tab", (ctl_t*)ctl, &cont);
  write_atm(ret->dir, "atm_res.tab", (ctl_t*)ctl, &res);
  if (ctl->retnn || ctl->retrr || ctl->retss) write_aero(ret->dir, "aero_cont.tab", &acont);
}

// The retrieval's forward calls integrate in FP64 when the Jacobian does
// (option "kernel_fp64", default 1), so that F(x) and K(x) come from the
// same arithmetic; the caller's "fp64" setting returns afterwards.
struct Fp64Scope {
  double saved;
  Fp64Scope() : saved(engine_option_double("fp64", 0)) {
    if (engine_option_int("kernel_fp64", 1)) jur_set_option("fp64", 1);
  }
  ~Fp64Scope() { jur_set_option("fp64", saved); }
};

// wall-clock time of one retrieval's phases, reported through jur_stat as
// oe_formod_ms, oe_kernel_ms and oe_total_ms (the rest is host-side algebra,
// rocBLAS / rocSOLVER calls and output files)
double oe_formod_ms, oe_kernel_ms;

struct PhaseClock {
  double& acc;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit PhaseClock(double& a) : acc(a) {}
  ~PhaseClock() { acc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
};

void kernel_matrix(ctl_t* ctl, atm_t* atm, obs_t* obs, aero_t* aero, std::vector<double>& k, size_t m, size_t n) {
  PhaseClock clock(oe_kernel_ms);
  k.assign(m * n, 0.0);
  jur_kernel(ctl, atm, obs, aero, k.data(), m, n, n);
}

}  // namespace

// write_matrix for callers with a GSL-layout matrix (the kernel command)
extern "C" void write_matrix(const char* dirname, const char* filename, ctl_t* ctl, jur_gsl_matrix* matrix,
                             atm_t* atm, aero_t* aero, obs_t* obs, const char* rowspace, const char* colspace,
                             const char* sort) {
  static aero_t none;
  if (!aero) aero = &none;
  const size_t r = matrix->size1, c = matrix->size2;
  std::vector<double> m(r * c);
  for (size_t i = 0; i < r; i++)
    for (size_t j = 0; j < c; j++) m[i * c + j] = matrix->data[i * matrix->tda + j];
  write_matrix(dirname, filename, ctl, m, atm, aero, obs, rowspace, colspace, sort);
}

// read_ret (retrievalmodel.c:710-757)
extern "C" void read_ret(int argc, char* argv[], ctl_t* ctl, ret_t* ret) {
  auto num = [&](const char* key, int idx, const char* dflt) { return scan_ctl(argc, argv, key, idx, dflt, nullptr); };
  ret->kernel_recomp = (int)num("KERNEL_RECOMP", -1, "1");
  ret->conv_itmax = (int)num("CONV_ITMAX", -1, "20");
  ret->conv_dmin = num("CONV_DMIN", -1, "0.1");
  ret->resmax = num("RESMAX", -1, "-999");
  ret->err_ana = (int)num("ERR_ANA", -1, "1");
  for (int id = 0; id < ctl->nd; id++) ret->err_formod[id] = num("ERR_FORMOD", id, "0");
  for (int id = 0; id < ctl->nd; id++) ret->err_noise[id] = num("ERR_NOISE", id, "0");
  ret->err_press = num("ERR_PRESS", -1, "0");
  ret->err_press_cz = num("ERR_PRESS_CZ", -1, "-999");
  ret->err_press_ch = num("ERR_PRESS_CH", -1, "-999");
  ret->err_temp = num("ERR_TEMP", -1, "0");
  ret->err_temp_cz = num("ERR_TEMP_CZ", -1, "-999");
  ret->err_temp_ch = num("ERR_TEMP_CH", -1, "-999");
  ret->err_nn = num("ERR_NN", -1, "0");
  ret->err_rr = num("ERR_RR", -1, "0");
  ret->err_ss = num("ERR_SS", -1, "0");
  for (int ig = 0; ig < ctl->ng; ig++) {
    ret->err_q[ig] = num("ERR_Q", ig, "0");
    ret->err_q_cz[ig] = num("ERR_Q_CZ", ig, "-999");
    ret->err_q_ch[ig] = num("ERR_Q_CH", ig, "-999");
  }
  for (int iw = 0; iw < ctl->nw; iw++) {
    ret->err_k[iw] = num("ERR_K", iw, "0");
    ret->err_k_cz[iw] = num("ERR_K_CZ", iw, "-999");
    ret->err_k_ch[iw] = num("ERR_K_CH", iw, "-999");
  }
}

// optimal_estimation (retrievalmodel.c:407-706)
extern "C" void optimal_estimation(ret_t* ret, ctl_t* ctl, obs_t* obs_meas, obs_t* obs_i, atm_t* atm_apr,
                                   atm_t* atm_i, aero_t* aero_apr, aero_t* aero_i) {
  Fp64Scope fp64;
  const auto t_start = std::chrono::steady_clock::now();
  oe_formod_ms = oe_kernel_ms = 0;
  // clear air: the aerosol arguments may be NULL
  static aero_t none_apr, none_i;
  if (!aero_apr) aero_apr = &none_apr;
  if (!aero_i) aero_i = &none_i;
  StateVec sv;
  state_vector(ctl, atm_apr, aero_apr, sv);
  const size_t m = meas_count(ctl, obs_meas), n = sv.n;
  if (m <= 0 || n <= 0) JUR_FAIL("Check problem definition!");
  // the reference takes it modulo KERNEL_RECOMP (a zero divides by zero)
  if (ret->kernel_recomp <= 0) JUR_FAIL("KERNEL_RECOMP must be positive!");
  printf("Problem size: m= %d / n= %d (alloc= %.4g MB / stat= %.4g MB)\n", (int)m, (int)n,
         (double)(3 * m * n + 3 * n * n + 8 * m + 8 * n) * sizeof(double) / 1024. / 1024.,
         (double)(5 * sizeof(atm_t) + 3 * sizeof(obs_t) + 2 * NMAX * sizeof(int) + 3 * sizeof(aero_t)) / 1024. /
             1024.);
  std::vector<double> a, cov, k_i, s_a_inv, b(n), dx(n), dy(m), sig_eps_inv(m), sig_formod, sig_noise, x_a(n),
      x_i(n), x_step(n), y_aux(m), y_i(m), y_m(m);
  // initial state: the a priori
  copy_atm(ctl, atm_i, atm_apr, 0);
  copy_obs(ctl, obs_i, obs_meas, 0);
  copy_aero(ctl, aero_i, aero_apr, 0);
  {
    PhaseClock clock(oe_formod_ms);
    formod(ctl, atm_i, obs_i, aero_i);
  }
  {
    StateVec s;
    state_vector(ctl, atm_apr, aero_apr, s);
    x_a = s.x;
    StateVec t;
    state_vector(ctl, atm_i, aero_i, t);
    x_i = t.x;
  }
  meas_vector(ctl, obs_meas, y_m.data(), nullptr, nullptr);
  meas_vector(ctl, obs_i, y_i.data(), nullptr, nullptr);
  // a priori covariance and its inverse, measurement errors
  set_cov_apr(ret, ctl, atm_apr, aero_apr, sv, s_a_inv);
  write_matrix(ret->dir, "matrix_cov_apr.tab", ctl, s_a_inv, atm_i, aero_i, obs_i, "x", "x", "r");
  matrix_invert(s_a_inv, (int)n);
  set_cov_meas(ret, ctl, obs_meas, sig_noise, sig_formod, sig_eps_inv);
  char filename[2 * LEN + 16];
  snprintf(filename, sizeof filename, "%s/costs.tab", ret->dir);
  FILE* out = fopen(filename, "w");
  if (!out) JUR_FAIL("Cannot create cost function file!");
  for (size_t i = 0; i < n; i++) dx[i] = x_i[i] - x_a[i];
  for (size_t i = 0; i < m; i++) dy[i] = y_m[i] - y_i[i];
  double chisq = cost_function(out, 0, dx, dy, s_a_inv, sig_eps_inv);
  if (!std::isfinite(chisq)) {
    printf("Retrieval failed!\n");
    fclose(out);
    return;
  }
  kernel_matrix(ctl, atm_i, obs_i, aero_i, k_i, m, n);
  // Levenberg-Marquardt minimisation
  double lmpar = 0.001, disq = 0;
  for (int it = 1; it <= ret->conv_itmax; it++) {
    const double chisq_old = chisq;
    if (it % ret->kernel_recomp == 0 && !(ret->kernel_recomp == 1 && it == 1))
      kernel_matrix(ctl, atm_i, obs_i, aero_i, k_i, m, n);
    if (it % ret->kernel_recomp == 0 || it == 1) matrix_product(k_i, (int)m, (int)n, sig_eps_inv, 1, cov);
    for (size_t i = 0; i < n; i++) dx[i] = x_i[i] - x_a[i];
    for (size_t i = 0; i < m; i++) dy[i] = y_m[i] - y_i[i];
    // b = K^T S_eps^-1 dy - S_a^-1 dx
    for (size_t i = 0; i < m; i++) y_aux[i] = dy[i] * sig_eps_inv[i] * sig_eps_inv[i];
    for (size_t j = 0; j < n; j++) {
      double s = 0;
      for (size_t i = 0; i < m; i++) s += k_i[i * n + j] * y_aux[i];
      double t = 0;
      for (size_t i = 0; i < n; i++) t += s_a_inv[j * n + i] * dx[i];
      b[j] = s - t;
    }
    for (int it2 = 0; it2 < 20; it2++) {
      // A = (1 + lmpar) S_a^-1 + K^T S_eps^-1 K, solved by Cholesky
      a.resize(n * n);
      for (size_t i = 0; i < n * n; i++) a[i] = s_a_inv[i] * (1 + lmpar) + cov[i];
      chol_solve(a.data(), b.data(), x_step.data(), (int)n);
      for (size_t i = 0; i < n; i++) x_i[i] += x_step[i];
      copy_atm(ctl, atm_i, atm_apr, 0);
      copy_obs(ctl, obs_i, obs_meas, 0);
      copy_aero(ctl, aero_i, aero_apr, 0);
      state_to_atm(ctl, x_i.data(), atm_i, aero_i);
      // keep the state physical
      for (int ip = 0; ip < atm_i->np; ip++) {
        atm_i->p[ip] = std::fmin(std::fmax(atm_i->p[ip], 5e-7), 5e4);
        atm_i->t[ip] = std::fmin(std::fmax(atm_i->t[ip], 100.0), 400.0);
        for (int ig = 0; ig < ctl->ng; ig++) atm_i->q[ig][ip] = std::fmin(std::fmax(atm_i->q[ig][ip], 0.0), 1.0);
        for (int iw = 0; iw < ctl->nw; iw++) atm_i->k[iw][ip] = std::fmax(atm_i->k[iw][ip], 0.0);
      }
      for (int im = 0; im < aero_i->nm; im++) {
        aero_i->nn[im] = std::fmin(std::fmax(aero_i->nn[im], 1e-6), 1e4);
        aero_i->rr[im] = std::fmin(std::fmax(aero_i->rr[im], 1e-2), 100.0);
        aero_i->ss[im] = std::fmin(std::fmax(aero_i->ss[im], 1.1), 3.0);
      }
      {
        PhaseClock clock(oe_formod_ms);
        formod(ctl, atm_i, obs_i, aero_i);
      }
      meas_vector(ctl, obs_i, y_i.data(), nullptr, nullptr);
      for (size_t i = 0; i < n; i++) dx[i] = x_i[i] - x_a[i];
      for (size_t i = 0; i < m; i++) dy[i] = y_m[i] - y_i[i];
      chisq = cost_function(out, it, dx, dy, s_a_inv, sig_eps_inv);
      if (chisq > chisq_old) {
        lmpar *= 10;
        for (size_t i = 0; i < n; i++) x_i[i] -= x_step[i];
      } else {
        lmpar /= 10;
        break;
      }
    }
    disq = 0;
    for (size_t i = 0; i < n; i++) disq += x_step[i] * b[i];
    disq /= (double)n;
    if (disq < ret->conv_dmin) break;
  }
  fclose(out);
  write_obs(ret->dir, "obs_final.tab", ctl, obs_i);
  write_atm(ret->dir, "atm_final.tab", ctl, atm_i);
  write_aero(ret->dir, "aero_final.tab", aero_i);
  write_matrix(ret->dir, "matrix_kernel.tab", ctl, k_i, atm_i, aero_i, obs_i, "y", "x", "r");
  if (ret->err_ana) {
    // retrieval covariance (S_a^-1 + K^T S_eps^-1 K)^-1
    matrix_product(k_i, (int)m, (int)n, sig_eps_inv, 1, cov);
    for (size_t i = 0; i < n * n; i++) cov[i] += s_a_inv[i];
    matrix_invert(cov, (int)n);
    write_matrix(ret->dir, "matrix_cov_ret.tab", ctl, cov, atm_i, aero_i, obs_i, "x", "x", "r");
    write_stddev("total", ret, ctl, atm_i, aero_i, cov, n);
    // gain matrix G = cov K^T S_eps^-1
    std::vector<double> auxnm(n * m), gain(n * m);
    for (size_t i = 0; i < n; i++)
      for (size_t j = 0; j < m; j++) auxnm[i * m + j] = k_i[j * n + i] * sig_eps_inv[j] * sig_eps_inv[j];
    gemm(false, false, (int)n, (int)m, (int)n, cov.data(), auxnm.data(), gain.data());
    write_matrix(ret->dir, "matrix_gain.tab", ctl, gain, atm_i, aero_i, obs_i, "x", "y", "c");
    // retrieval errors due to noise and forward model errors: G S G^T
    matrix_product(gain, (int)n, (int)m, sig_noise, 2, a);
    write_stddev("noise", ret, ctl, atm_i, aero_i, a, n);
    matrix_product(gain, (int)n, (int)m, sig_formod, 2, a);
    write_stddev("formod", ret, ctl, atm_i, aero_i, a, n);
    // averaging kernel A = G K
    a.assign(n * n, 0.0);
    gemm(false, false, (int)n, (int)n, (int)m, gain.data(), k_i.data(), a.data());
    write_matrix(ret->dir, "matrix_avk.tab", ctl, a, atm_i, aero_i, obs_i, "x", "x", "r");
    analyze_avk(ret, ctl, atm_i, aero_i, sv, a);
  }
  engine_set_stat("oe_formod_ms", oe_formod_ms);
  engine_set_stat("oe_kernel_ms", oe_kernel_ms);
  engine_set_stat("oe_total_ms",
                  std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count());
}
