// mie.cpp -- layer optical properties from log-normal aerosol / cloud modes
// (reference src/scatter.c:330-508): host bookkeeping (layer grouping,
// refractive-index files, quadrature nodes) and the device Mie kernels.
#include <cmath>
#include <cstring>
#include <strings.h>

#include "jur_host.h"
#include "jur_mie.h"

namespace {

// Gauss-Hermite nodes and weights, n = NRAD (scatter.c:275-327; Newton
// iteration from the Numerical Recipes 3rd ed. initial guesses).
void gauher_nodes(double* x, double* w) {
  const double EPS = 1.0e-14, PIM4 = 0.7511255444649425;
  const int MAXIT = 10000, n = NRAD, m = (NRAD + 1) / 2;
  double p1, p2, p3, pp = 0, z = 0, z1;
  for (int i = 0; i < m; i++) {
    if (i == 0)
      z = sqrt((double)(2 * n + 1)) - 1.85575 * pow((double)(2 * n + 1), -0.16667);
    else if (i == 1)
      z -= 1.14 * pow((double)n, 0.426) / z;
    else if (i == 2)
      z = 1.86 * z - 0.86 * x[0];
    else if (i == 3)
      z = 1.91 * z - 0.91 * x[1];
    else
      z = 2.0 * z - x[i - 2];
    int its;
    for (its = 0; its < MAXIT; its++) {
      p1 = PIM4;
      p2 = 0.0;
      for (int j = 0; j < n; j++) {
        p3 = p2;
        p2 = p1;
        p1 = z * sqrt(2.0 / ((double)j + 1)) * p2 - sqrt((double)j / (j + 1.)) * p3;
      }
      pp = sqrt(2. * (double)n) * p2;
      z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) <= EPS) break;
    }
    if (its >= MAXIT) JUR_FAIL("Too many iterations in gauher.");
    x[i] = z;
    x[n - 1 - i] = -z;
    w[i] = 2.0 / (pp * pp);
    w[n - 1 - i] = w[i];
  }
}

// Series length and D_n start of one sphere (scatter.c:83-88).
MieItem mie_item(double x, double nr, double ni) {
  MieItem I;
  I.x = x;
  I.nr = nr;
  I.ni = ni;
  double xstop = x + 4.E0 * pow(x, 0.3333) + 2.0;
  I.nstop = (int)xstop;
  double ymod = hypot(x * nr, x * ni);
  I.nmx = (int)((xstop > ymod ? xstop : ymod) + 15);
  if (I.nmx > 10000) JUR_FAIL("Too many Mie terms!");
  return I;
}

struct MieBufs {
  DBuf items, scratch, qext, qsca, phase, nodes, item_of, mbe, mbs, mp, mode0, nmod, be, bs, ba, p;
};
MieBufs& MB() {
  static MieBufs b;
  return b;
}

// Run k_bhmie over all items, in chunks that keep the D_n / a_n / b_n scratch
// under ~1 GB.
void run_bhmie(const std::vector<MieItem>& items, hipStream_t s, int grid = 0) {
  MieBufs& b = MB();
  const int n = (int)items.size();
  if (n == 0) return;
  int maxn = 1;
  for (auto& I : items) maxn = I.nstop > maxn ? I.nstop : maxn;
  size_t stride = (size_t)maxn + 2;
  MieItem* di = (MieItem*)b.items.need(n * sizeof(MieItem));
  HIP_OK(hipMemcpyAsync(di, items.data(), n * sizeof(MieItem), hipMemcpyHostToDevice, s));
  b.qext.need(n * sizeof(double));
  b.qsca.need(n * sizeof(double));
  b.phase.need((size_t)n * NTHETA * sizeof(double));
  size_t per = 3 * stride * sizeof(double2);
  int chunk = (int)((size_t)1 << 30) / (int)per;
  if (chunk < 1) chunk = 1;
  if (chunk > n) chunk = n;
  double2* scr = (double2*)b.scratch.need(per * chunk);
  for (int i0 = 0; i0 < n; i0 += chunk) {
    int m = n - i0 < chunk ? n - i0 : chunk;
    launch_bhmie(di, i0, m, scr, stride, b.qext.as<double>(), b.qsca.as<double>(), b.phase.as<double>(), grid,
                 s);
  }
}

}  // namespace

// bhmie (scatter.c:56-234): one sphere on the device.
extern "C" void bhmie(double x, double n_real, double n_imag, double* phase, double* qext, double* qsca) {
  hipStream_t s = engine_stream();
  std::vector<MieItem> items = {mie_item(x, n_real, n_imag)};
  run_bhmie(items, s);
  MieBufs& b = MB();
  HIP_OK(hipMemcpyAsync(qext, b.qext.p, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(qsca, b.qsca.p, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(phase, b.phase.p, NTHETA * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
}

// opt_prop_external (scatter.c:512-579): the mode's optical properties from
// a table of rows "nu beta_e beta_s P(0 deg) ... P(180 deg)", for each
// channel the row nearest its wavenumber (bracket, then the closer of its two
// rows; a tie keeps the lower).  NN = 0 selects this, NN = 1 stops with the
// reference's "Implement me!", other values leave the mode at zero.  Lines
// follow the reference's TOK rules (an unparsable wavenumber or coefficient
// skips the line, a missing token is an error); an unparsable phase value is
// an error here (the reference keeps whatever its static table held there).
// A mode naming the same file as the previous mode reuses the last table; a
// one-row table serves every channel.
static void ext_mode(const ctl_t* ctl, const aero_t* aero, int m, double* be, double* bs, double* ph) {
  static std::vector<double> nu, ne, ns, np_;
  const int nd = ctl->nd;
  for (int id = 0; id < nd; id++) {
    be[id] = bs[id] = 0;
    for (int ia = 0; ia < NTHETA; ia++) ph[(size_t)id * NTHETA + ia] = 0;
  }
  if (aero->nn[m] == 0) {
    if (m == 0 || strcmp(aero->filepath[m], aero->filepath[m - 1]) != 0) {
      printf("Read non-spherical optical properties: %s\n", aero->filepath[m]);
      FILE* in = fopen(aero->filepath[m], "r");
      if (!in) JUR_FAIL("Cannot open file!");
      nu.clear();
      ne.clear();
      ns.clear();
      np_.clear();
      static char line[LEN];
      double row[3 + NTHETA];
      while (fgets(line, LEN, in)) {
        bool skip = false;
        char* tok = strtok(line, " \t");
        for (int k = 0; k < 3 + NTHETA && !skip; k++) {
          if (k > 0) tok = strtok(nullptr, " \t");
          if (!tok) JUR_FAIL("Error while reading!");
          if (sscanf(tok, "%lg", &row[k]) != 1) {
            if (k >= 3) JUR_FAIL("Error while reading!");
            skip = true;
          }
        }
        if (skip) continue;
        nu.push_back(row[0]);
        ne.push_back(row[1]);
        ns.push_back(row[2]);
        np_.insert(np_.end(), row + 3, row + 3 + NTHETA);
        if (nu.size() > (size_t)REFMAX) JUR_FAIL("Too many data points!");
      }
      fclose(in);
      if (nu.empty()) JUR_FAIL("Could not read any data!");
    }
    const int n = (int)nu.size();
    if (n < 1) JUR_FAIL("Could not read any data!");
    for (int id = 0; id < nd; id++) {
      const double x = ctl->nu[id];
      int im = n >= 2 ? locate_mono<double>(nu.data(), n, x) : 0;
      if (n >= 2 && im != n && fabs(nu[im] - x) > fabs(nu[im + 1] - x)) im = im + 1;
      be[id] = ne[im];
      bs[id] = ns[im];
      for (int ia = 0; ia < NTHETA; ia++) ph[(size_t)id * NTHETA + ia] = np_[(size_t)im * NTHETA + ia];
    }
  }
  if (aero->nn[m] == 1) JUR_FAIL("Implement me!");
}

extern "C" void get_opt_prop(ctl_t* ctl, aero_t* aero) {
  engine_get_opt_prop(ctl, aero);
}

void engine_get_opt_prop(ctl_t* ctl, aero_t* aero) {
  hipStream_t s = engine_stream();
  MieBufs& b = MB();
  const int nd = ctl->nd;
  // Layers: modes sharing a top altitude form one layer (scatter.c:339-365).
  if (aero->top_mod[0] <= aero->bottom_mod[0])
    JUR_FAIL("Aerosol top altitude is smaller than bottom altitude. Please check aero.tab.");
  int nl = 1, nm = 1;
  aero->top[0] = aero->top_mod[0];
  aero->bottom[0] = aero->bottom_mod[0];
  aero->trans[0] = aero->trans_mod[0];
  for (int ii = 1; ii < aero->nm; ii++) {
    if (aero->top_mod[ii] > aero->top_mod[ii - 1] || aero->top_mod[ii] <= aero->bottom_mod[ii] ||
        aero->bottom_mod[ii] > aero->bottom_mod[ii - 1])
      JUR_FAIL("Aerosol/Cloud altitudes and/or transition layers are wrong. Please check aero.tab.");
    if (aero->top_mod[ii] != aero->top_mod[ii - 1]) {
      if (nl >= NLMAX) JUR_FAIL("Too many aerosol/cloud layers!");
      aero->top[nl] = aero->top_mod[ii];
      aero->bottom[nl] = aero->bottom_mod[ii];
      aero->trans[nl] = aero->trans_mod[ii];
      aero->nmod[nl - 1] = nm;
      nl++;
      nm = 0;
    }
    nm++;
  }
  aero->nmod[nl - 1] = nm;
  aero->nl = nl;

  static double zs[NRAD], wts[NRAD];
  static bool gh = false;
  if (!gh) {
    gauher_nodes(zs, wts);
    gh = true;
  }
  const int nmode = aero->nm;
  // radius quadrature: the reference's NRAD Gauss-Hermite nodes, or with
  // option "mie_grid" 1 the aerosol configuration's 64 radii log-spaced
  // over 0.01-20 um (trapezoid rule in ln r over the log-normal density)
  const int grid = engine_option_int("mie_grid", 0) ? 1 : 0;
  const int NR = grid ? MIE_GRID_NR : NRAD;
  std::vector<MieItem> items;
  std::vector<int> item_of((size_t)nmode * NR * nd, -1);
  std::vector<MieNode> nodes((size_t)nmode * NR);
  std::vector<double> nr(nd), ni(nd);
  // tabulated modes ("Ext"): their properties, host side, into the mode rows below
  std::vector<char> is_ext(nmode, 0);
  std::vector<double> ext_be((size_t)nmode * nd), ext_bs((size_t)nmode * nd),
      ext_p((size_t)nmode * nd * NTHETA);
  for (int m = 0; m < nmode; m++) {
    if (strcasecmp(aero->type[m], "MIE") != 0) {
      if (strcasecmp(aero->type[m], "Ext") == 0) {
        is_ext[m] = 1;
        ext_mode(ctl, aero, m, &ext_be[(size_t)m * nd], &ext_bs[(size_t)m * nd], &ext_p[(size_t)m * nd * NTHETA]);
        continue;
      }
      if (strcasecmp(aero->type[m], "Const") == 0) {
        printf("Using constant extinction [1/km]: %g\n", aero->nn[m]);
        JUR_FAIL("Implement me!");
      }
      JUR_FAIL("Please give valid scattering model (MIE, Ext, Const)!");
    }
    // Refractive indices, linear in wavenumber (scatter.c:444-462).
    printf("Read refractive indices: %s\n", aero->filepath[m]);
    FILE* in = fopen(aero->filepath[m], "r");
    if (!in) JUR_FAIL("Cannot open file!");
    std::vector<double> fnu, fnr, fni;
    static char line[LEN];
    double a0, a1, a2;
    while (fgets(line, LEN, in))
      if (sscanf(line, "%lg %lg %lg", &a0, &a1, &a2) == 3) {
        fnu.push_back(a0);
        fnr.push_back(a1);
        fni.push_back(a2);
        if ((int)fnu.size() > REFMAX) JUR_FAIL("Too many data points!");
      }
    fclose(in);
    if (fnu.size() < 2) JUR_FAIL("Could not read any data!");
    for (int id = 0; id < nd; id++) {
      int idx = locate_mono<double>(fnu.data(), (int)fnu.size(), ctl->nu[id]);
      nr[id] = lin(fnu[idx], fnr[idx], fnu[idx + 1], fnr[idx + 1], ctl->nu[id]);
      ni[id] = lin(fnu[idx], fni[idx], fnu[idx + 1], fni[idx + 1], ctl->nu[id]);
    }
    if (aero->nn[m] <= 0 || aero->rr[m] <= 0 || aero->ss[m] <= 1)
      JUR_FAIL("The log-normal parameters are nonsense. ((?_?)) ");
    const double K1 = aero->nn[m] * 1e-3 * sqrt(M_PI);
    for (int nn = 0; nn < NR; nn++) {
      double rad, wn;
      if (grid) {
        // the Gauss-Hermite form's weight becomes exp(-z^2) dz, z = (ln r -
        // ln r_m) / (sqrt 2 ln s), with the trapezoid step in ln r
        const double l0 = log(0.01), dl = (log(20.) - l0) / (MIE_GRID_NR - 1), sg = sqrt(2) * log(aero->ss[m]);
        rad = exp(l0 + nn * dl);
        const double z = (log(rad) - log(aero->rr[m])) / sg;
        wn = exp(-z * z) * dl / sg * ((nn == 0 || nn == MIE_GRID_NR - 1) ? 0.5 : 1.0);
      } else {
        rad = exp(sqrt(2) * log(aero->ss[m]) * zs[nn] + log(aero->rr[m]));
        wn = wts[nn];
      }
      nodes[(size_t)m * NR + nn] = {rad, K1, wn};
      if (rad >= 0.001 && rad <= 1000. && K1 > 0.) {
        for (int id = 0; id < nd; id++) {
          double lambda = 1. / (ctl->nu[id]) * pow(10, 4.);
          double x = 2 * M_PI * rad / lambda;
          item_of[((size_t)m * NR + nn) * nd + id] = (int)items.size();
          items.push_back(mie_item(x, nr[id], ni[id]));
        }
      }
    }
  }
  run_bhmie(items, s, grid);
  // Mode sums over radius nodes, then layer sums over modes.
  MieNode* dn = (MieNode*)b.nodes.need(nodes.size() * sizeof(MieNode));
  int* dio = (int*)b.item_of.need(item_of.size() * sizeof(int));
  HIP_OK(hipMemcpyAsync(dn, nodes.data(), nodes.size() * sizeof(MieNode), hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(dio, item_of.data(), item_of.size() * sizeof(int), hipMemcpyHostToDevice, s));
  double* mbe = (double*)b.mbe.need((size_t)nmode * nd * sizeof(double));
  double* mbs = (double*)b.mbs.need((size_t)nmode * nd * sizeof(double));
  double* mp = (double*)b.mp.need((size_t)nmode * nd * NTHETA * sizeof(double));
  for (int m = 0; m < nmode; m++) {
    if (is_ext[m]) {
      HIP_OK(hipMemcpyAsync(mbe + (size_t)m * nd, &ext_be[(size_t)m * nd], nd * sizeof(double),
                            hipMemcpyHostToDevice, s));
      HIP_OK(hipMemcpyAsync(mbs + (size_t)m * nd, &ext_bs[(size_t)m * nd], nd * sizeof(double),
                            hipMemcpyHostToDevice, s));
      HIP_OK(hipMemcpyAsync(mp + (size_t)m * nd * NTHETA, &ext_p[(size_t)m * nd * NTHETA],
                            (size_t)nd * NTHETA * sizeof(double), hipMemcpyHostToDevice, s));
      continue;
    }
    launch_mie_mode(dn + (size_t)m * NR, NR, dio + (size_t)m * NR * nd, nd, b.qext.as<double>(),
                    b.qsca.as<double>(), b.phase.as<double>(), mbe + (size_t)m * nd, mbs + (size_t)m * nd,
                    mp + (size_t)m * nd * NTHETA, s);
  }
  std::vector<int> mode0(nl), nmod(nl);
  for (int il = 0, c = 0; il < nl; il++) {
    mode0[il] = c;
    nmod[il] = aero->nmod[il];
    c += aero->nmod[il];
  }
  int* dm0 = (int*)b.mode0.need(nl * sizeof(int));
  int* dnm = (int*)b.nmod.need(nl * sizeof(int));
  HIP_OK(hipMemcpyAsync(dm0, mode0.data(), nl * sizeof(int), hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(dnm, nmod.data(), nl * sizeof(int), hipMemcpyHostToDevice, s));
  double* be = (double*)b.be.need((size_t)nl * nd * sizeof(double));
  double* bs = (double*)b.bs.need((size_t)nl * nd * sizeof(double));
  double* ba = (double*)b.ba.need((size_t)nl * nd * sizeof(double));
  double* pp = (double*)b.p.need((size_t)nl * nd * NTHETA * sizeof(double));
  launch_mie_layers(nl, dm0, dnm, nd, mbe, mbs, mp, be, bs, ba, pp, s);
  std::vector<double> hbe((size_t)nl * nd), hbs((size_t)nl * nd), hba((size_t)nl * nd),
      hp((size_t)nl * nd * NTHETA);
  HIP_OK(hipMemcpyAsync(hbe.data(), be, hbe.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(hbs.data(), bs, hbs.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(hba.data(), ba, hba.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(hp.data(), pp, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  for (int il = 0; il < nl; il++)
    for (int id = 0; id < nd; id++) {
      aero->beta_e[il][id] = hbe[(size_t)il * nd + id];
      aero->beta_s[il][id] = hbs[(size_t)il * nd + id];
      aero->beta_a[il][id] = hba[(size_t)il * nd + id];
      for (int it = 0; it < NTHETA; it++) aero->p[il][id][it] = hp[((size_t)il * nd + id) * NTHETA + it];
    }
  engine_set_stat("mie_spheres", (double)items.size());
}
