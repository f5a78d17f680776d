// control_io.cpp -- control keys and the text / binary file formats of the
// forward model (host side; no compute).
//
// Formats follow the reference: control files and argv overrides
// (src/control.c:7-171), atm.tab (src/atmosphere.c:487-598), obs.tab
// (src/misc.c:269-390), aero.tab (src/scatter.c:583-639, 1045-1092),
// emissivity tables .bin/.tab (src/forwardmodel.c:610-739), two-column
// shape/filter files (src/forwardmodel.c:577-606).
#include <cstring>
#include <strings.h>

#include "jur_host.h"

namespace {

void make_path(char* file, const char* dirname, const char* filename) {
  if (dirname != nullptr)
    snprintf(file, LEN, "%s/%s", dirname, filename);
  else
    snprintf(file, LEN, "%s", filename);
}

// A line parsed with the reference's TOK convention (jurassic.h:93-98):
// tokens split on blanks/tabs; a token that does not parse skips the line;
// a line with too few tokens is an error.
struct Tok {
  char* line;
  bool first = true;
  bool bad = false;
  explicit Tok(char* l) : line(l) {}
  char* next() {
    char* t = strtok(first ? line : nullptr, " \t");
    first = false;
    if (!t) JUR_FAIL("Error while reading!");
    return t;
  }
  void dbl(double& v) {
    if (bad) return;
    if (sscanf(next(), "%lg", &v) != 1) bad = true;
  }
  void str(char* v) {
    if (bad) return;
    if (sscanf(next(), "%s", v) != 1) bad = true;
  }
};

}  // namespace

extern "C" double scan_ctl(int argc, char* argv[], const char* varname, int arridx,
                           const char* defvalue, char* value) {
  FILE* in = nullptr;
  static char line[LEN], rname[LEN], dummy[LEN], rval[LEN], name1[LEN], name2[LEN];
  bool found = false;
  if (argv[1][0] != '-')
    if (!(in = fopen(argv[1], "r"))) JUR_FAIL("Cannot open file!");
  if (arridx >= 0) {
    snprintf(name1, LEN, "%s[%d]", varname, arridx);
    snprintf(name2, LEN, "%s[*]", varname);
  } else {
    snprintf(name1, LEN, "%s", varname);
    snprintf(name2, LEN, "%s", varname);
  }
  if (in) {
    while (fgets(line, LEN, in))
      if (sscanf(line, "%s %s %s", rname, dummy, rval) == 3 &&
          (strcasecmp(rname, name1) == 0 || strcasecmp(rname, name2) == 0)) {
        found = true;
        break;
      }
    fclose(in);
  }
  for (int i = 1; i < argc - 1; i++)
    if (strcasecmp(argv[i], name1) == 0 || strcasecmp(argv[i], name2) == 0) {
      snprintf(rval, LEN, "%s", argv[i + 1]);
      found = true;
      break;
    }
  if (!found) {
    if (strlen(defvalue) > 0) {
      snprintf(rval, LEN, "%s", defvalue);
    } else {
      static char msg[2 * LEN];
      snprintf(msg, sizeof msg, "Missing variable %s!\n", name1);
      JUR_FAIL(msg);
    }
  }
  printf("%s = %s\n", name1, rval);
  if (value) snprintf(value, LEN, "%s", rval);
  return atof(rval);
}

extern "C" void read_ctl(int argc, char* argv[], ctl_t* ctl) {
  printf("\nJURASSIC - Version with Scattering (MI355X)\nExecutable: %s\nControl parameters: %s\n",
         argv[0], argv[1]);
  auto num = [&](const char* k, const char* d) { return scan_ctl(argc, argv, k, -1, d, nullptr); };
  ctl->ng = (int)num("NG", "0");
  if (ctl->ng < 0 || ctl->ng > NGMAX) JUR_FAIL("Set 0 <= NG <= MAX!");
  for (int ig = 0; ig < ctl->ng; ig++) scan_ctl(argc, argv, "EMITTER", ig, "", ctl->emitter[ig]);
  ctl->nd = (int)num("ND", "0");
  if (ctl->nd < 0 || ctl->nd > NDMAX) JUR_FAIL("Set 0 <= ND <= MAX!");
  for (int id = 0; id < ctl->nd; id++) ctl->nu[id] = scan_ctl(argc, argv, "NU", id, "", nullptr);
  ctl->nw = (int)num("NW", "1");
  if (ctl->nw < 0 || ctl->nw > NWMAX) JUR_FAIL("Set 0 <= NW <= NWMAX!");
  for (int id = 0; id < ctl->nd; id++)
    ctl->window[id] = (int)scan_ctl(argc, argv, "WINDOW", id, "0", nullptr);
  scan_ctl(argc, argv, "TBLBASE", -1, "-", ctl->tblbase);
  ctl->hydz = num("HYDZ", "-999");
  ctl->ctm_co2 = (int)num("CTM_CO2", "1");
  ctl->ctm_h2o = (int)num("CTM_H2O", "1");
  ctl->ctm_n2 = (int)num("CTM_N2", "1");
  ctl->ctm_o2 = (int)num("CTM_O2", "1");
  ctl->sca_n = (int)num("SCA_N", "0");
  if (ctl->sca_n < 0 || ctl->sca_n > SCAMOD) JUR_FAIL("Set 0 <= SCA_NMOD <= MAX!");
  ctl->sca_mult = (int)num("SCA_MULT", "1");
  scan_ctl(argc, argv, "SCA_EXT", -1, "beta_a", ctl->sca_ext);
  ctl->ip = (int)num("IP", "1");
  ctl->cz = num("CZ", "0");
  ctl->cx = num("CX", "0");
  ctl->refrac = (int)num("REFRAC", "1");
  ctl->rayds = num("RAYDS", "10");
  ctl->raydz = num("RAYDZ", "1");
  ctl->transs = num("TRANSS", "0.02");
  scan_ctl(argc, argv, "FOV", -1, "-", ctl->fov);
  ctl->retp_zmin = num("RETP_ZMIN", "-999");
  ctl->retp_zmax = num("RETP_ZMAX", "-999");
  ctl->rett_zmin = num("RETT_ZMIN", "-999");
  ctl->rett_zmax = num("RETT_ZMAX", "-999");
  for (int ig = 0; ig < ctl->ng; ig++) {
    ctl->retq_zmin[ig] = scan_ctl(argc, argv, "RETQ_ZMIN", ig, "-999", nullptr);
    ctl->retq_zmax[ig] = scan_ctl(argc, argv, "RETQ_ZMAX", ig, "-999", nullptr);
  }
  for (int iw = 0; iw < ctl->nw; iw++) {
    ctl->retk_zmin[iw] = scan_ctl(argc, argv, "RETK_ZMIN", iw, "-999", nullptr);
    ctl->retk_zmax[iw] = scan_ctl(argc, argv, "RETK_ZMAX", iw, "-999", nullptr);
  }
  ctl->retnn = (int)num("RETNN", "0");
  ctl->retrr = (int)num("RETRR", "0");
  ctl->retss = (int)num("RETSS", "0");
  ctl->retnn_zmin = (int)num("RETNN_ZMIN", "-999");
  ctl->retnn_zmax = (int)num("RETNN_ZMAX", "-999");
  ctl->write_bbt = (int)num("WRITE_BBT", "0");
  ctl->write_matrix = (int)num("WRITE_MATRIX", "0");
  ctl->queue.capacity = (int)num("MAX_QUEUE", "-1");
  ctl->queue.items = nullptr;
  ctl->queue.begin = ctl->queue.end = 0;
  ctl->queue.state = Queue_Inactive;
}

extern "C" void read_atm(const char* dirname, const char* filename, ctl_t* ctl, atm_t* atm) {
  static char file[LEN], line[LEN];
  make_path(file, dirname, filename);
  printf("Read atmospheric data: %s\n", file);
  FILE* in = fopen(file, "r");
  if (!in) JUR_FAIL("Cannot open file!");
  atm->init = 0;
  atm->np = 0;
  while (fgets(line, LEN, in)) {
    Tok tk(line);
    int n = atm->np;
    tk.dbl(atm->time[n]);
    tk.dbl(atm->z[n]);
    tk.dbl(atm->lon[n]);
    tk.dbl(atm->lat[n]);
    tk.dbl(atm->p[n]);
    tk.dbl(atm->t[n]);
    for (int ig = 0; ig < ctl->ng; ig++) tk.dbl(atm->q[ig][n]);
    for (int iw = 0; iw < ctl->nw; iw++) tk.dbl(atm->k[iw][n]);
    if (tk.bad) continue;
    if ((++atm->np) > NPMAX) JUR_FAIL("Too many data points!");
  }
  fclose(in);
  if (atm->np < 1) JUR_FAIL("Could not read any data!");
}

extern "C" void write_atm(const char* dirname, const char* filename, ctl_t* ctl, atm_t* atm) {
  static char file[LEN];
  make_path(file, dirname, filename);
  printf("Write atmospheric data: %s\n", file);
  FILE* out = fopen(file, "w");
  if (!out) JUR_FAIL("Cannot create file!");
  fprintf(out,
          "# $1 = time (seconds since 2000-01-01T00:00Z)\n"
          "# $2 = altitude [km]\n"
          "# $3 = longitude [deg]\n"
          "# $4 = latitude [deg]\n"
          "# $5 = pressure [hPa]\n"
          "# $6 = temperature [K]\n");
  int n = 6;
  for (int ig = 0; ig < ctl->ng; ig++) fprintf(out, "# $%d = %s volume mixing ratio\n", ++n, ctl->emitter[ig]);
  for (int iw = 0; iw < ctl->nw; iw++) fprintf(out, "# $%d = window %d: extinction [1/km]\n", ++n, iw);
  for (int ip = 0; ip < atm->np; ip++) {
    if (ip == 0 || atm->lat[ip] != atm->lat[ip - 1] || atm->lon[ip] != atm->lon[ip - 1]) fprintf(out, "\n");
    fprintf(out, "%.2f %g %g %g %g %g", atm->time[ip], atm->z[ip], atm->lon[ip], atm->lat[ip], atm->p[ip],
            atm->t[ip]);
    for (int ig = 0; ig < ctl->ng; ig++) fprintf(out, " %g", atm->q[ig][ip]);
    for (int iw = 0; iw < ctl->nw; iw++) fprintf(out, " %g", atm->k[iw][ip]);
    fprintf(out, "\n");
  }
  fclose(out);
}

extern "C" void read_obs(const char* dirname, const char* filename, ctl_t* ctl, obs_t* obs) {
  static char file[LEN], line[LEN];
  make_path(file, dirname, filename);
  printf("Read observation data: %s\n", file);
  FILE* in = fopen(file, "r");
  if (!in) JUR_FAIL("Cannot open file!");
  obs->nr = 0;
  while (fgets(line, LEN, in)) {
    Tok tk(line);
    int n = obs->nr;
    double* cols[10] = {&obs->time[n], &obs->obsz[n], &obs->obslon[n], &obs->obslat[n], &obs->vpz[n],
                        &obs->vplon[n], &obs->vplat[n], &obs->tpz[n], &obs->tplon[n], &obs->tplat[n]};
    for (double* c : cols) tk.dbl(*c);
    for (int id = 0; id < ctl->nd; id++) tk.dbl(obs->rad[id][n]);
    for (int id = 0; id < ctl->nd; id++) tk.dbl(obs->tau[id][n]);
    if (tk.bad) continue;
    if ((++obs->nr) > NRMAX) JUR_FAIL("Too many rays!");
  }
  fclose(in);
  if (obs->nr < 1) JUR_FAIL("Could not read any data!");
}

extern "C" void write_obs(const char* dirname, const char* filename, ctl_t* ctl, obs_t* obs) {
  static char file[LEN];
  make_path(file, dirname, filename);
  printf("Write observation data: %s\n", file);
  FILE* out = fopen(file, "w");
  if (!out) JUR_FAIL("Cannot create file!");
  fprintf(out,
          "# $1 = time (seconds since 2000-01-01T00:00Z)\n"
          "# $2 = observer altitude [km]\n"
          "# $3 = observer longitude [deg]\n"
          "# $4 = observer latitude [deg]\n"
          "# $5 = view point altitude [km]\n"
          "# $6 = view point longitude [deg]\n"
          "# $7 = view point latitude [deg]\n"
          "# $8 = tangent point altitude [km]\n"
          "# $9 = tangent point longitude [deg]\n"
          "# $10 = tangent point latitude [deg]\n");
  int n = 10;
  for (int id = 0; id < ctl->nd; id++)
    fprintf(out, "# $%d = channel %g: radiance [W/(m^2 sr cm^-1)]\n", ++n, ctl->nu[id]);
  for (int id = 0; id < ctl->nd; id++) fprintf(out, "# $%d = channel %g: transmittance\n", ++n, ctl->nu[id]);
  for (int ir = 0; ir < obs->nr; ir++) {
    if (ir == 0 || obs->time[ir] != obs->time[ir - 1]) fprintf(out, "\n");
    fprintf(out, "%.2f %g %g %g %g %g %g %g %g %g", obs->time[ir], obs->obsz[ir], obs->obslon[ir],
            obs->obslat[ir], obs->vpz[ir], obs->vplon[ir], obs->vplat[ir], obs->tpz[ir], obs->tplon[ir],
            obs->tplat[ir]);
    for (int id = 0; id < ctl->nd; id++) fprintf(out, " %g", obs->rad[id][ir]);
    for (int id = 0; id < ctl->nd; id++) fprintf(out, " %g", obs->tau[id][ir]);
    fprintf(out, "\n");
  }
  fclose(out);
}

extern "C" void read_aero(const char* dirname, const char* filename, ctl_t* ctl, aero_t* aero) {
  static char file[LEN], line[LEN];
  make_path(file, dirname, filename);
  printf("Read aerosol data: %s\n", file);
  FILE* in = fopen(file, "r");
  if (!in) JUR_FAIL("Cannot open file!");
  aero->nm = 0;
  while (fgets(line, LEN, in)) {
    Tok tk(line);
    int n = aero->nm;
    tk.dbl(aero->top_mod[n]);
    tk.dbl(aero->bottom_mod[n]);
    tk.dbl(aero->trans_mod[n]);
    tk.str(aero->type[n]);
    tk.str(aero->filepath[n]);
    tk.dbl(aero->nn[n]);
    tk.dbl(aero->rr[n]);
    tk.dbl(aero->ss[n]);
    if (tk.bad) continue;
    if ((++aero->nm) > SCAMOD) JUR_FAIL("Too many aerosol models!");
  }
  fclose(in);
  if (aero->nm < 1) JUR_FAIL("Could not read any data!");
  if (ctl->sca_n != aero->nm)
    JUR_FAIL("Number of scattering models in control file and aerosol file does not match.");
}

extern "C" void write_aero(const char* dirname, const char* filename, aero_t* aero) {
  static char file[LEN];
  make_path(file, dirname, filename);
  printf("Write particle data: %s\n", file);
  FILE* out = fopen(file, "w");
  if (!out) JUR_FAIL("Cannot create file!");
  fprintf(out,
          "# $1 = aerosol layer top altitude [km]\n"
          "# $2 = aerosol layer bottom altitude [km]\n"
          "# $3 = transition layer thickness [km]\n"
          "# $4 = source for optical properties\n"
          "# $5 = refractive index file\n"
          "# $6 = particle concentration of log-normal mode [cm-3]\n"
          "# $7 = median radius of log-normal mode [mum]\n"
          "# $8 = width of log-normal mode\n");
  for (int im = 0; im < aero->nm; im++) {
    if (im == 0) fprintf(out, "\n");
    fprintf(out, "%g %g %g %s %s %g %g %g\n", aero->top_mod[im], aero->bottom_mod[im], aero->trans_mod[im],
            aero->type[im], aero->filepath[im], aero->nn[im], aero->rr[im], aero->ss[im]);
  }
  fclose(out);
}

void read_shape_file(const char* filename, std::vector<double>& x, std::vector<double>& y) {
  FILE* in = fopen(filename, "r");
  if (!in) JUR_FAIL("Cannot open file!");
  static char line[LEN];
  x.clear();
  y.clear();
  double a, b;
  while (fgets(line, LEN, in))
    if (sscanf(line, "%lg %lg", &a, &b) == 2) {
      x.push_back(a);
      y.push_back(b);
      if ((int)x.size() > NSHAPE) JUR_FAIL("Too many data points!");
    }
  fclose(in);
  if (x.empty()) JUR_FAIL("Could not read any data!");
}

// read_shape (forwardmodel.c:577-606): two columns into x, y (NSHAPE max).
extern "C" void read_shape(const char* filename, double* x, double* y, int* n) {
  std::vector<double> a, b;
  read_shape_file(filename, a, b);
  for (size_t i = 0; i < a.size(); i++) {
    x[i] = a[i];
    y[i] = b[i];
  }
  *n = (int)a.size();
}

// copy_aero (scatter.c:237-271): copy; with init, clear the mode and layer
// parameters and the optics of the ctl->nd channels (ss = 1).
extern "C" void copy_aero(ctl_t* ctl, aero_t* dst, aero_t* src, int init) {
  memmove(dst, src, sizeof(aero_t));
  if (!init) return;
  for (int im = 0; im < dst->nm; im++) {
    dst->top_mod[im] = dst->bottom_mod[im] = dst->trans_mod[im] = 0;
    dst->nn[im] = dst->rr[im] = 0;
    dst->ss[im] = 1;
  }
  for (int il = 0; il < dst->nl; il++) {
    dst->nmod[il] = 0;
    dst->top[il] = dst->bottom[il] = dst->trans[il] = 0;
    for (int id = 0; id < ctl->nd; id++) {
      dst->beta_e[il][id] = dst->beta_s[il][id] = dst->beta_a[il][id] = 0;
      for (int ia = 0; ia < NTHETA; ia++) dst->p[il][id][ia] = 0;
    }
  }
}

// Binary table: int np; double p[np]; int nt[np]; per p: double t[nt];
// int nu[nt]; per T: float u[nu]; float eps[nu] (forwardmodel.c:636-651).
bool read_table_bin(const char* filename, HostTable& tb) {
  FILE* in = fopen(filename, "r");
  if (!in) return false;
  auto rd = [&](void* ptr, size_t sz, size_t n) {
    if (fread(ptr, sz, n, in) != n) JUR_FAIL("Error while reading!");
  };
  rd(&tb.np, sizeof(int), 1);
  if (tb.np > TBLNPMAX) JUR_FAIL("Too many pressure levels!");
  if (tb.np < 0) tb.np = 0;
  tb.p.resize(tb.np);
  tb.nt.resize(tb.np);
  rd(tb.p.data(), sizeof(double), tb.np);
  rd(tb.nt.data(), sizeof(int), tb.np);
  tb.t.resize(tb.np);
  tb.nu.resize(tb.np);
  tb.u.resize(tb.np);
  tb.eps.resize(tb.np);
  for (int ip = 0; ip < tb.np; ip++) {
    if (tb.nt[ip] > TBLNTMAX) JUR_FAIL("Too many temperatures!");
    int nt = tb.nt[ip] > 0 ? tb.nt[ip] : 0;
    tb.t[ip].resize(nt);
    tb.nu[ip].resize(nt);
    rd(tb.t[ip].data(), sizeof(double), nt);
    rd(tb.nu[ip].data(), sizeof(int), nt);
    tb.u[ip].resize(nt);
    tb.eps[ip].resize(nt);
    for (int it = 0; it < nt; it++) {
      int n = tb.nu[ip][it] < TBLNUMAX ? tb.nu[ip][it] : TBLNUMAX;
      if (n < 0) n = 0;
      tb.nu[ip][it] = n;
      tb.u[ip][it].resize(n);
      tb.eps[ip][it].resize(n);
      rd(tb.u[ip][it].data(), sizeof(float), n);
      rd(tb.eps[ip][it].data(), sizeof(float), n);
    }
  }
  fclose(in);
  return true;
}

// ASCII table "p T u eps" rows; a row is kept only if u and eps both
// increase within its (p, T) column (forwardmodel.c:658-737): a
// non-increasing row overwrites the column's last node, a row beyond
// TBLNUMAX nodes is dropped.  A new pressure level always opens a new
// temperature column (the reference indexes column -1 when its first
// temperature equals the previous row's).
bool read_table_ascii(const char* filename, HostTable& tb) {
  FILE* in = fopen(filename, "r");
  if (!in) return false;
  printf("Read emissivity table: %s\n", filename);
  static char line[LEN];
  double press, temp, u, eps, eps_old = -999, press_old = -999, temp_old = -999, u_old = -999;
  tb = HostTable();
  tb.np = 0;
  while (fgets(line, LEN, in)) {
    if (sscanf(line, "%lg %lg %lg %lg", &press, &temp, &u, &eps) != 4) continue;
    if (press != press_old) {
      press_old = press;
      temp_old = NAN;
      tb.p.push_back(press);
      tb.t.emplace_back();
      tb.u.emplace_back();
      tb.eps.emplace_back();
      tb.nu.emplace_back();
      if ((int)tb.p.size() > TBLNPMAX) JUR_FAIL("Too many pressure levels!");
    }
    int ip = (int)tb.p.size() - 1;
    tb.p[ip] = press;
    if (temp != temp_old) {
      temp_old = temp;
      tb.t[ip].push_back(temp);
      tb.u[ip].emplace_back();
      tb.eps[ip].emplace_back();
      if ((int)tb.t[ip].size() > TBLNTMAX) JUR_FAIL("Too many temperatures!");
    }
    int it = (int)tb.t[ip].size() - 1;
    tb.t[ip][it] = temp;
    auto& cu = tb.u[ip][it];
    auto& ce = tb.eps[ip][it];
    if ((eps > eps_old && u > u_old) || cu.empty()) {
      eps_old = eps;
      u_old = u;
      if ((int)cu.size() + 1 >= TBLNUMAX + 1) continue;  // column full: drop the row
      cu.push_back((float)u);
      ce.push_back((float)eps);
    } else {
      // a non-increasing row overwrites the current last node
      cu.back() = (float)u;
      ce.back() = (float)eps;
    }
  }
  fclose(in);
  tb.np = (int)tb.p.size();
  tb.nt.resize(tb.np);
  for (int ip = 0; ip < tb.np; ip++) {
    tb.nt[ip] = (int)tb.t[ip].size();
    tb.nu[ip].resize(tb.nt[ip]);
    for (int it = 0; it < tb.nt[ip]; it++) tb.nu[ip][it] = (int)tb.u[ip][it].size();
  }
  return true;
}

// tab2bin (reference src/tab2bin.c): every table of ctl -- binary if present,
// else ASCII -- written in the binary layout (forwardmodel.c:636-651) as
// TBLBASE_<nu>_<emitter>.bin.  A missing table is reported and skipped (the
// reference would write one with np = -1).
extern "C" int jur_tab2bin(ctl_t* ctl) {
  static char fn[2 * LEN + 64];
  for (int ig = 0; ig < ctl->ng; ig++)
    for (int id = 0; id < ctl->nd; id++) {
      HostTable tb;
      snprintf(fn, sizeof fn, "%s_%.4f_%s.bin", ctl->tblbase, ctl->nu[id], ctl->emitter[ig]);
      bool ok = read_table_bin(fn, tb);
      if (!ok) {
        snprintf(fn, sizeof fn, "%s_%.4f_%s.tab", ctl->tblbase, ctl->nu[id], ctl->emitter[ig]);
        ok = read_table_ascii(fn, tb);
      }
      if (!ok) {
        printf("Missing emissivity table: %s\n", fn);
        continue;
      }
      snprintf(fn, sizeof fn, "%s_%.4f_%s.bin", ctl->tblbase, ctl->nu[id], ctl->emitter[ig]);
      printf("Write binary table: %s\n", fn);
      FILE* out = fopen(fn, "wb");
      if (!out) JUR_FAIL("Cannot create file!");
      auto put = [&](const void* d, size_t size, size_t n) {
        if (n && fwrite(d, size, n, out) != n) JUR_FAIL("Error while writing!");
      };
      put(&tb.np, sizeof(int), 1);
      put(tb.p.data(), sizeof(double), tb.np);
      put(tb.nt.data(), sizeof(int), tb.np);
      for (int ip = 0; ip < tb.np; ip++) {
        put(tb.t[ip].data(), sizeof(double), tb.nt[ip]);
        put(tb.nu[ip].data(), sizeof(int), tb.nt[ip]);
        for (int it = 0; it < tb.nt[ip]; it++) {
          put(tb.u[ip][it].data(), sizeof(float), tb.nu[ip][it]);
          put(tb.eps[ip][it].data(), sizeof(float), tb.nu[ip][it]);
        }
      }
      fclose(out);
    }
  return 0;
}
