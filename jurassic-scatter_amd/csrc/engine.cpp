// engine.cpp -- device context of the forward model: emissivity tables,
// source-function tables and continuum constants resident in HBM, aerosol
// layer optics, work buffers, and the clear-air / extinction pipeline
// (ray tracing -> sort by path length -> path integration).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <strings.h>

#include "continua_coeffs.inc"
#include "jur_host.h"
#include "jur_kernels_misc.h"
#include "jur_scatter.h"

namespace {

struct Engine {
  int prefix_nr = 0;          // FP64 prefix reuse: rays are copies of the first prefix_nr
  DBuf pstart, prem, pids, snap, fdiff, fcont, pevals, pkeys, ev_acc;
  int ev_ng = -1;  // >= 0: the call reused base rays, ev_acc holds its emitter evaluations
  int call_ng = 0;  // emitters of the call's path integration
  bool ready = false;
  int device = 0;
  hipStream_t stream = nullptr;
  // table set
  std::string tbl_key, chan_key;
  int ng = 0, nd = 0;
  DBuf t_np, t_p, t_nt, t_t, t_ncol, t_col, t_ue, planck, chan, nu_d;
  DBuf t_p64, t_t64, planck64, chan64;  // FP64 grids, Planck table, channel data
  DBuf g_np, g_p, g_nt, g_t, t_meta;
  int shared = 0, gnp_max = 0, gnt_max = 0;
  unsigned ue_bytes = 0, meta_bytes = 0;
  // aerosol layers
  DBuf beta_e, beta_s, beta_a, alti;
  DBuf beta_e64, beta_s64, beta_a64;  // FP64 copies for the FP64 path
  int nalti = 0;
  double altimax = 0, altimin = 0;
  int nl = 0;
  double top[NLMAX] = {0}, bottom[NLMAX] = {0}, trans[NLMAX] = {0};
  // work buffers of the primary pass
  DBuf geo, geo2, rec, npts, tsurf, status, ids, order, keys, sorttmp;
  DBuf pack;  // packed point records and (p, T) brackets of the path integration
  DBuf lps;   // log-slopes of the traced atmosphere set's pressure
  DBuf tkey, torder, thist;  // march order of the traced rays
  int cap = 512;
  // FP64 point records and surface temperatures of the last trace (option
  // "fp64"); the next path integration reads them
  DBuf rec64, tsurf64;
  const double* last_rec64 = nullptr;
  const double* last_tsurf64 = nullptr;
  int last_nray = 0, last_cap = 0;
  std::map<std::string, double> opt, stats;
  // per-call accounting of the path-integration launches: event pairs
  // around each launch and the LOS points integrated (device counter)
  std::vector<std::pair<hipEvent_t, hipEvent_t>> path_ev;
  size_t path_ev_used = 0;
  DBuf pts_acc;
  // captured clear-air forward call (option "graph"): the key holds every
  // address and size the graph was recorded with
  bool capturing = false;
  // device-resident (planned) calls: kernels only between the host's
  // launches -- the counting sort, no autotuning runs
  bool planned = false;
  DBuf hist;
  // inputs of the last engine_set_aero
  bool aero_ready = false;
  std::vector<char> aero_snap;
  std::string g_key;
  hipGraph_t g_graph = nullptr;
  hipGraphExec_t g_exec = nullptr;
  size_t hbm_bytes = 0;  // the device's memory
};

Engine& E() {
  static Engine e;
  if (!e.ready) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      JUR_FAIL("No GPU found: the forward model needs an MI355X (gfx950) device.");
    HIP_OK(hipSetDevice(e.device));
    HIP_OK(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
    size_t fr = 0;
    HIP_OK(hipMemGetInfo(&fr, &e.hbm_bytes));
    e.ready = true;
  }
  return e;
}

int find_emitter(const ctl_t* ctl, const char* name) {
  for (int ig = 0; ig < ctl->ng; ig++)
    if (strcasecmp(ctl->emitter[ig], name) == 0) return ig;
  return -1;
}

template <typename T>
T* upload(DBuf& b, const std::vector<T>& v, hipStream_t s) {
  T* d = (T*)b.need(v.size() * sizeof(T) + sizeof(T));
  if (!v.empty()) HIP_OK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return d;
}

}  // namespace

// Continuum coefficients of channel id (continua.c): the nu-dependent table
// interpolation is done once here in FP64; the kernel evaluates the T, p
// and column-density dependence per point.
ChanConst make_chan_const(const ctl_t* ctl, int id) {
  ChanConst c;
  memset(&c, 0, sizeof c);
  const double nu = ctl->nu[id];
  c.nu = (float)nu;
  c.window = ctl->window[id];
  int ig_co2 = find_emitter(ctl, "CO2"), ig_h2o = find_emitter(ctl, "H2O");
  double xw = nu / 2 + 1;  // ctmco2 (continua.c:854-869)
  if (ctl->ctm_co2 && ig_co2 >= 0 && xw >= 1 && xw < 2001) {
    int iw = (int)xw, jw = iw + 1;
    double dw = xw - iw, ew = 1 - dw;
    double cw296 = ew * JUR_CTM_CO2_296[iw - 1] + dw * JUR_CTM_CO2_296[jw - 1];
    double cw260 = ew * JUR_CTM_CO2_260[iw - 1] + dw * JUR_CTM_CO2_260[jw - 1];
    double cw230 = ew * JUR_CTM_CO2_230[iw - 1] + dw * JUR_CTM_CO2_230[jw - 1];
    c.co2_on = 1;
    c.co2_ig = ig_co2;
    c.co2_k230 = (float)(5.050505e-4 * cw230);
    c.co2_k260 = (float)(9.259259e-4 * cw260);
    c.co2_k296 = (float)(4.208754e-4 * cw296);
  }
  xw = nu / 10 + 1;  // ctmh2o (continua.c:1898-1928)
  if (ctl->ctm_h2o && ig_h2o >= 0 && xw >= 1 && xw < 2001) {
    int iw = (int)xw, jw = iw + 1;
    double dw = xw - iw, ew = 1 - dw, sfac;
    double cw296 = ew * JUR_CTM_H2O_SELF296[iw - 1] + dw * JUR_CTM_H2O_SELF296[jw - 1];
    double cw260 = ew * JUR_CTM_H2O_SELF260[iw - 1] + dw * JUR_CTM_H2O_SELF260[jw - 1];
    double cwfrn = ew * JUR_CTM_H2O_FOREIGN[iw - 1] + dw * JUR_CTM_H2O_FOREIGN[jw - 1];
    if (nu <= 820 || nu >= 960) {
      sfac = 1;
    } else {
      double xx = (nu - 820) / 10;
      int ix = (int)xx;
      double dx = xx - ix;
      sfac = (1 - dx) * JUR_CTM_H2O_XFCREV[ix] + dx * JUR_CTM_H2O_XFCREV[ix + 1];
    }
    double vf2 = (nu - 370) * (nu - 370), vf6 = vf2 * vf2 * vf2;
    double fscal = 36100 / (vf2 + vf6 * 1e-8 + 36100) * -.25 + 1;
    c.h2o_on = 1;
    c.h2o_ig = ig_h2o;
    c.h2o_s296 = (float)(sfac * cw296);
    c.h2o_lnratio = (float)log(cw260 / cw296);
    c.h2o_frn = (float)(cwfrn * fscal);
    c.h2o_tanhk = (float)(.7193876 * nu);
  }
  auto band = [&](const double* nua, const double* ba, const double* bb, int n, float& b, float& beta) {
    if (nu < nua[0] || nu > nua[n - 1]) return 0;
    int idx = locate_mono<double>(nua, n, nu);
    b = (float)lin(nua[idx], ba[idx], nua[idx + 1], ba[idx + 1], nu);
    beta = (float)lin(nua[idx], bb[idx], nua[idx + 1], bb[idx + 1], nu);
    return 1;
  };
  if (ctl->ctm_n2) c.n2_on = band(JUR_CTM_N2_NU, JUR_CTM_N2_B, JUR_CTM_N2_BETA, 98, c.n2_b, c.n2_beta);
  if (ctl->ctm_o2) c.o2_on = band(JUR_CTM_O2_NU, JUR_CTM_O2_B, JUR_CTM_O2_BETA, 90, c.o2_b, c.o2_beta);
  return c;
}

// The channel's continuum data in FP64 (continua.c:854-869, 1898-1928,
// 1980-1992, 2037-2048), for the FP64 path.
ChanConst64 make_chan_const64(const ctl_t* ctl, int id) {
  ChanConst64 c;
  memset(&c, 0, sizeof c);
  const double nu = ctl->nu[id];
  c.nu = nu;
  c.window = ctl->window[id];
  int ig_co2 = find_emitter(ctl, "CO2"), ig_h2o = find_emitter(ctl, "H2O");
  double xw = nu / 2 + 1;
  if (ctl->ctm_co2 && ig_co2 >= 0 && xw >= 1 && xw < 2001) {
    int iw = (int)xw, jw = iw + 1;
    double dw = xw - iw, ew = 1 - dw;
    c.co2_on = 1;
    c.co2_ig = ig_co2;
    c.co2_cw296 = ew * JUR_CTM_CO2_296[iw - 1] + dw * JUR_CTM_CO2_296[jw - 1];
    c.co2_cw260 = ew * JUR_CTM_CO2_260[iw - 1] + dw * JUR_CTM_CO2_260[jw - 1];
    c.co2_cw230 = ew * JUR_CTM_CO2_230[iw - 1] + dw * JUR_CTM_CO2_230[jw - 1];
  }
  xw = nu / 10 + 1;
  if (ctl->ctm_h2o && ig_h2o >= 0 && xw >= 1 && xw < 2001) {
    int iw = (int)xw, jw = iw + 1;
    double dw = xw - iw, ew = 1 - dw, sfac;
    c.h2o_cw296 = ew * JUR_CTM_H2O_SELF296[iw - 1] + dw * JUR_CTM_H2O_SELF296[jw - 1];
    c.h2o_cw260 = ew * JUR_CTM_H2O_SELF260[iw - 1] + dw * JUR_CTM_H2O_SELF260[jw - 1];
    double cwfrn = ew * JUR_CTM_H2O_FOREIGN[iw - 1] + dw * JUR_CTM_H2O_FOREIGN[jw - 1];
    if (nu <= 820 || nu >= 960) {
      sfac = 1;
    } else {
      double xx = (nu - 820) / 10;
      int ix = (int)xx;
      double dx = xx - ix;
      sfac = (1 - dx) * JUR_CTM_H2O_XFCREV[ix] + dx * JUR_CTM_H2O_XFCREV[ix + 1];
    }
    double vf2 = (nu - 370) * (nu - 370), vf6 = vf2 * vf2 * vf2;
    double fscal = 36100 / (vf2 + vf6 * 1e-8 + 36100) * -.25 + 1;
    c.h2o_on = 1;
    c.h2o_ig = ig_h2o;
    c.h2o_sfac = sfac;
    c.h2o_ctwfrn = cwfrn * fscal;
  }
  auto band = [&](const double* nua, const double* ba, const double* bb, int n, double& b, double& beta) {
    if (nu < nua[0] || nu > nua[n - 1]) return 0;
    int idx = locate_mono<double>(nua, n, nu);
    b = lin(nua[idx], ba[idx], nua[idx + 1], ba[idx + 1], nu);
    beta = lin(nua[idx], bb[idx], nua[idx + 1], bb[idx + 1], nu);
    return 1;
  };
  if (ctl->ctm_n2) c.n2_on = band(JUR_CTM_N2_NU, JUR_CTM_N2_B, JUR_CTM_N2_BETA, 98, c.n2_b, c.n2_beta);
  if (ctl->ctm_o2) c.o2_on = band(JUR_CTM_O2_NU, JUR_CTM_O2_B, JUR_CTM_O2_BETA, 90, c.o2_b, c.o2_beta);
  return c;
}

hipStream_t engine_stream() { return E().stream; }

// Trace-buffer budget of a scattering level's chunks: option "chunk_mb", by
// default a sixth of the device's memory, at most 48 GiB (48 GiB of 288 on
// an MI355X: ~400,000 pencils of 512 LOS points per chunk, several waves
// per SIMD for the thread-per-pencil march).
size_t engine_chunk_bytes() {
  const int mb = engine_option_int("chunk_mb", 0);
  if (mb > 0) return (size_t)mb << 20;
  const size_t cap = (size_t)48 << 30, sixth = E().hbm_bytes / 6;
  return sixth > 0 && sixth < cap ? sixth : cap;
}
int engine_option_int(const char* key, int dflt) {
  auto& m = E().opt;
  auto it = m.find(key);
  return it == m.end() ? dflt : (int)it->second;
}
double engine_stat(const char* key) {
  auto& m = E().stats;
  auto it = m.find(key);
  return it == m.end() ? NAN : it->second;
}
void engine_set_stat(const char* key, double v) { E().stats[key] = v; }

// Keys of the table set (table base, emitters, wavenumbers) and of the
// channel constants (plus the continua switches and windows) of ctl.
static std::string table_key(const ctl_t* ctl) {
  std::string key = std::string(ctl->tblbase) + "|";
  for (int ig = 0; ig < ctl->ng; ig++) key += std::string(ctl->emitter[ig]) + ",";
  char buf[64];
  for (int id = 0; id < ctl->nd; id++) {
    snprintf(buf, sizeof buf, "%.17g,", ctl->nu[id]);
    key += buf;
  }
  return key;
}
static std::string channel_key(const ctl_t* ctl, const std::string& key) {
  std::string ck = key;
  char buf[64];
  snprintf(buf, sizeof buf, "|%d%d%d%d", ctl->ctm_co2, ctl->ctm_h2o, ctl->ctm_n2, ctl->ctm_o2);
  ck += buf;
  for (int id = 0; id < ctl->nd; id++) ck += std::to_string(ctl->window[id]) + ",";
  return ck;
}

static std::vector<char> aero_snapshot(const ctl_t* ctl, const aero_t* aero);

bool engine_resident(const ctl_t* ctl, const aero_t* aero) {
  Engine& e = E();
  const std::string key = table_key(ctl);
  return key == e.tbl_key && channel_key(ctl, key) == e.chan_key && e.aero_ready &&
         aero_snapshot(ctl, ctl->sca_n > 0 ? aero : nullptr) == e.aero_snap;
}

void engine_prepare(const ctl_t* ctl) {
  Engine& e = E();
  hipStream_t s = e.stream;
  // Tables and source-function tables: loaded once per table set (the
  // reference loads them at its first call, forwardmodel.c:286-292).
  const std::string key = table_key(ctl);
  if (key != e.tbl_key) {
    const int ng = ctl->ng, nd = ctl->nd, ntbl = ng * nd;
    std::vector<int> np(ntbl, -1), nt((size_t)ntbl * TBLNPMAX, 0), ncol((size_t)ntbl * TBLNPMAX * TBLNTMAX, 0),
        col((size_t)ntbl * TBLNPMAX * TBLNTMAX, 0);
    std::vector<float> p((size_t)ntbl * TBLNPMAX, 0.f), t((size_t)ntbl * TBLNPMAX * TBLNTMAX, 0.f);
    std::vector<double> p64((size_t)ntbl * TBLNPMAX, 0.0), t64((size_t)ntbl * TBLNPMAX * TBLNTMAX, 0.0);
    std::vector<float2> ue;
    static char fn[2 * LEN + 64];
    for (int id = 0; id < nd; id++)
      for (int ig = 0; ig < ng; ig++) {
        HostTable tb;
        snprintf(fn, sizeof fn, "%s_%.4f_%s.bin", ctl->tblbase, ctl->nu[id], ctl->emitter[ig]);
        bool ok = read_table_bin(fn, tb);
        if (!ok) {
          snprintf(fn, sizeof fn, "%s_%.4f_%s.tab", ctl->tblbase, ctl->nu[id], ctl->emitter[ig]);
          ok = read_table_ascii(fn, tb);
          if (!ok) printf("Missing emissivity table: %s\n", fn);
        }
        const int k = id * ng + ig;  // channel-contiguous table order
        if (!ok) continue;
        np[k] = tb.np;
        for (int ip = 0; ip < tb.np; ip++) {
          p[(size_t)k * TBLNPMAX + ip] = (float)tb.p[ip];
          p64[(size_t)k * TBLNPMAX + ip] = tb.p[ip];
          nt[(size_t)k * TBLNPMAX + ip] = tb.nt[ip];
          for (int it = 0; it < tb.nt[ip]; it++) {
            size_t c = ((size_t)k * TBLNPMAX + ip) * TBLNTMAX + it;
            t[c] = (float)tb.t[ip][it];
            t64[c] = tb.t[ip][it];
            ncol[c] = tb.nu[ip][it];
            col[c] = (int)ue.size();
            for (int iu = 0; iu < tb.nu[ip][it]; iu++) ue.push_back(make_float2(tb.u[ip][it][iu], tb.eps[ip][it][iu]));
            ue.push_back(make_float2(0.f, 0.f));  // pad so col[i+1] stays in bounds
          }
        }
      }
    // Shared-grid fast path: every channel of an emitter has the same
    // increasing p and T grids, and all columns increase strictly in u and eps.
    bool shared = true;
    std::vector<int> gnp(ng, -1), gnt((size_t)ng * TBLNPMAX, 0);
    std::vector<float> gp((size_t)ng * TBLNPMAX, 0.f), gt((size_t)ng * TBLNPMAX * TBLNTMAX, 0.f);
    for (int ig = 0; ig < ng && shared; ig++) {
      const int k0 = ig;  // channel 0
      for (int id = 1; id < nd && shared; id++) {
        const int k = id * ng + ig;
        if (np[k] != np[k0]) shared = false;
        for (int ip = 0; shared && ip < np[k0]; ip++) {
          if (p[(size_t)k * TBLNPMAX + ip] != p[(size_t)k0 * TBLNPMAX + ip] ||
              nt[(size_t)k * TBLNPMAX + ip] != nt[(size_t)k0 * TBLNPMAX + ip])
            shared = false;
          for (int it = 0; shared && it < nt[(size_t)k0 * TBLNPMAX + ip]; it++)
            if (t[((size_t)k * TBLNPMAX + ip) * TBLNTMAX + it] != t[((size_t)k0 * TBLNPMAX + ip) * TBLNTMAX + it])
              shared = false;
        }
      }
      gnp[ig] = np[k0];
      for (int ip = 0; ip < np[k0]; ip++) {
        gp[(size_t)ig * TBLNPMAX + ip] = p[(size_t)k0 * TBLNPMAX + ip];
        if (ip > 0 && !(gp[(size_t)ig * TBLNPMAX + ip] > gp[(size_t)ig * TBLNPMAX + ip - 1])) shared = false;
        gnt[(size_t)ig * TBLNPMAX + ip] = nt[(size_t)k0 * TBLNPMAX + ip];
        for (int it = 0; it < nt[(size_t)k0 * TBLNPMAX + ip]; it++) {
          float v = t[((size_t)k0 * TBLNPMAX + ip) * TBLNTMAX + it];
          gt[((size_t)ig * TBLNPMAX + ip) * TBLNTMAX + it] = v;
          if (it > 0 && !(v > gt[((size_t)ig * TBLNPMAX + ip) * TBLNTMAX + it - 1])) shared = false;
        }
      }
    }
    for (size_t c = 0; c < col.size() && shared; c++)
      for (int i = 1; i < ncol[c]; i++) {
        const float2 a = ue[col[c] + i - 1], b = ue[col[c] + i];
        if (!(b.x > a.x && b.y > a.y)) {
          shared = false;
          break;
        }
      }
    std::vector<int2> meta(ncol.size());
    for (size_t c = 0; c < ncol.size(); c++) meta[c] = make_int2(ncol[c], col[c]);
    e.shared = shared ? 1 : 0;
    e.gnp_max = e.gnt_max = 0;
    for (int ig = 0; ig < ng; ig++) {
      e.gnp_max = std::max(e.gnp_max, gnp[ig]);
      for (int ip = 0; ip < gnp[ig]; ip++) e.gnt_max = std::max(e.gnt_max, gnt[(size_t)ig * TBLNPMAX + ip]);
    }
    upload(e.g_np, gnp, s);
    upload(e.g_p, gp, s);
    upload(e.g_nt, gnt, s);
    upload(e.g_t, gt, s);
    upload(e.t_meta, meta, s);
    e.meta_bytes = (unsigned)(meta.size() * sizeof(int2));
    upload(e.t_np, np, s);
    upload(e.t_p, p, s);
    upload(e.t_nt, nt, s);
    upload(e.t_t, t, s);
    upload(e.t_p64, p64, s);
    upload(e.t_t64, t64, s);
    upload(e.t_ncol, ncol, s);
    upload(e.t_col, col, s);
    upload(e.t_ue, ue, s);
    e.ue_bytes = (unsigned)(ue.size() * sizeof(float2));
    // Filter functions -> Planck table on the device.
    std::vector<double> fnu, fw;
    std::vector<int> foff, fnn;
    std::vector<double> x, y;
    for (int id = 0; id < nd; id++) {
      snprintf(fn, sizeof fn, "%s_%.4f.filt", ctl->tblbase, ctl->nu[id]);
      read_shape_file(fn, x, y);
      foff.push_back((int)fnu.size());
      fnn.push_back((int)x.size());
      fnu.insert(fnu.end(), x.begin(), x.end());
      fw.insert(fw.end(), y.begin(), y.end());
    }
    DBuf dfnu, dfw, dfoff, dfn;
    upload(dfnu, fnu, s);
    upload(dfw, fw, s);
    upload(dfoff, foff, s);
    upload(dfn, fnn, s);
    e.planck.need((size_t)nd * JUR_NPLANCK * sizeof(float));
    e.planck64.need((size_t)nd * JUR_NPLANCK * sizeof(double));
    launch_planck_table(dfnu.as<double>(), dfw.as<double>(), dfoff.as<int>(), dfn.as<int>(), nd,
                        e.planck.as<float>(), e.planck64.as<double>(), s);
    HIP_OK(hipStreamSynchronize(s));
    dfnu.release();
    dfw.release();
    dfoff.release();
    dfn.release();
    e.tbl_key = key;
    e.stats["shared_grids"] = e.shared;
    e.ng = ng;
    e.nd = nd;
    printf("Initialize source function table...\n");
  }
  // Channel constants (depend on the continua switches and windows too).
  const std::string ck = channel_key(ctl, key);
  if (ck != e.chan_key) {
    std::vector<ChanConst> cc(ctl->nd);
    std::vector<ChanConst64> cc64(ctl->nd);
    std::vector<double> nu(ctl->nd);
    for (int id = 0; id < ctl->nd; id++) {
      cc[id] = make_chan_const(ctl, id);
      cc64[id] = make_chan_const64(ctl, id);
      nu[id] = ctl->nu[id];
    }
    upload(e.chan, cc, s);
    upload(e.chan64, cc64, s);
    upload(e.nu_d, nu, s);
    HIP_OK(hipStreamSynchronize(s));
    e.chan_key = ck;
  }
}

namespace {
void drop_forward_graph();
}

// Aerosol layers for the ray tracer (layer-edge altitudes of
// add_aerosol_layers, lineofsight.c:211-244) and layer optics for the path
// integration.
// Every input engine_set_aero reads, as bytes.
static std::vector<char> aero_snapshot(const ctl_t* ctl, const aero_t* aero) {
  const int nd = ctl->nd;
  std::vector<char> snap;
  auto add = [&snap](const void* p, size_t n) {
    snap.insert(snap.end(), (const char*)p, (const char*)p + n);
  };
  const int on = (aero && ctl->sca_n > 0) ? 1 : 0;
  add(&on, sizeof on);
  add(&nd, sizeof nd);
  add(ctl->nu, sizeof(double) * nd);
  if (on) {
    const int nl = aero->nl < NLMAX ? aero->nl : NLMAX;
    add(&nl, sizeof nl);
    add(&ctl->transs, sizeof ctl->transs);
    add(aero->top, sizeof(double) * nl);
    add(aero->bottom, sizeof(double) * nl);
    add(aero->trans, sizeof(double) * nl);
    for (int il = 0; il < nl; il++) {
      add(aero->beta_e[il], sizeof(double) * nd);
      add(aero->beta_s[il], sizeof(double) * nd);
      add(aero->beta_a[il], sizeof(double) * nd);
      for (int id = 0; id < nd; id++) add(aero->p[il][id], sizeof(double) * NTHETA);
    }
  }
  return snap;
}

void engine_set_aero(const ctl_t* ctl, const aero_t* aero) {
  Engine& e = E();
  hipStream_t s = e.stream;
  const int nd = ctl->nd;
  // Layer optics and phase functions are uploaded only when they change:
  // the snapshot holds every input this function reads.
  std::vector<char> snap = aero_snapshot(ctl, aero);
  if (e.aero_ready && snap == e.aero_snap) return;
  // a recorded forward call holds the old layer geometry and buffers
  drop_forward_graph();
  e.aero_snap.swap(snap);
  e.aero_ready = true;
  std::vector<float> be((size_t)NLMAX * nd, 0.f), bs((size_t)NLMAX * nd, 0.f), ba((size_t)NLMAX * nd, 0.f);
  std::vector<double> be64((size_t)NLMAX * nd, 0.0), bs64((size_t)NLMAX * nd, 0.0), ba64((size_t)NLMAX * nd, 0.0);
  std::vector<double> alti;
  e.nl = 0;
  if (aero && ctl->sca_n > 0) {
    e.nl = aero->nl;
    const double eps = 0.005;
    for (int il = 0; il < aero->nl && il < NLMAX; il++) {
      e.top[il] = aero->top[il];
      e.bottom[il] = aero->bottom[il];
      e.trans[il] = aero->trans[il];
      for (int id = 0; id < nd; id++) {
        be[(size_t)il * nd + id] = (float)aero->beta_e[il][id];
        bs[(size_t)il * nd + id] = (float)aero->beta_s[il][id];
        ba[(size_t)il * nd + id] = (float)aero->beta_a[il][id];
        be64[(size_t)il * nd + id] = aero->beta_e[il][id];
        bs64[(size_t)il * nd + id] = aero->beta_s[il][id];
        ba64[(size_t)il * nd + id] = aero->beta_a[il][id];
      }
      alti.push_back(aero->top[il] + eps);
      alti.push_back(aero->top[il] - eps);
      alti.push_back(aero->bottom[il] + eps);
      alti.push_back(aero->bottom[il] - eps);
      if (aero->trans[il] > ctl->transs) {
        double tt = aero->trans[il] / ctl->transs;
        alti.push_back(aero->top[il] + aero->trans[il] + eps);
        alti.push_back(aero->top[il] + aero->trans[il] - eps);
        alti.push_back(aero->bottom[il] - aero->trans[il] + eps);
        alti.push_back(aero->bottom[il] - aero->trans[il] - eps);
        for (int it = 1; it < (int)tt; it++) {
          alti.push_back(aero->top[il] + aero->trans[il] - eps - it * ctl->transs);
          alti.push_back(aero->bottom[il] - aero->trans[il] + eps + it * ctl->transs);
        }
      }
    }
    // descending order (the reference sorts the negated values ascending)
    for (size_t i = 1; i < alti.size(); i++) {
      double v = alti[i];
      size_t j = i;
      while (j > 0 && alti[j - 1] < v) {
        alti[j] = alti[j - 1];
        j--;
      }
      alti[j] = v;
    }
  }
  e.nalti = (int)alti.size();
  e.altimax = e.nalti ? alti.front() : 0;
  e.altimin = e.nalti ? alti.back() : 0;
  upload(e.beta_e, be, s);
  upload(e.beta_s, bs, s);
  upload(e.beta_a, ba, s);
  upload(e.beta_e64, be64, s);
  upload(e.beta_s64, bs64, s);
  upload(e.beta_a64, ba64, s);
  upload(e.alti, alti, s);
  scatter_set_phase(ctl, aero, s);
  HIP_OK(hipStreamSynchronize(s));
}

namespace {
// Trace arguments of a ray set (everything but the buffers).
TraceArgs trace_args(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, int sca_on, int cap) {
  Engine& e = E();
  TraceArgs A;
  memset(&A, 0, sizeof A);
  A.nray = rays.nray;
  A.obsz = rays.obsz;
  A.obslon = rays.obslon;
  A.obslat = rays.obslat;
  A.vpz = rays.vpz;
  A.vplon = rays.vplon;
  A.vplat = rays.vplat;
  A.atm_idx = rays.atm_idx;
  A.atm = {atm.z, atm.p, atm.t, atm.q, atm.k, atm.nlev, ctl->ng, ctl->nw, atm.geo};
  A.rayds = ctl->rayds;
  A.raydz = ctl->raydz;
  A.transs = ctl->transs;
  A.refrac = ctl->refrac;
  A.sca_n = sca_on ? ctl->sca_n : 0;
  A.alti = e.alti.as<double>();
  A.nalti = e.nalti;
  A.altimax = e.altimax;
  A.altimin = e.altimin;
  A.nl = e.nl;
  for (int il = 0; il < NLMAX; il++) {
    A.top[il] = e.top[il];
    A.bottom[il] = e.bottom[il];
    A.trans[il] = e.trans[il];
  }
  A.nw = ctl->nw;
  A.ig_h2o = find_emitter(ctl, "H2O");
  A.R = rec_len(ctl->ng, ctl->nw);
  A.cap = cap;
  return A;
}

// Enqueue one ray trace of a ray set with per-ray capacity cap (status
// flags cleared first); no host synchronisation.
// status: the word the kernels raise flags in (null: the engine's own,
// cleared first); nvalid / nvalid_off: device-side count of the valid rays.
void enqueue_trace(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, int sca_on, TraceBufs& tb,
                   int cap, double* tpz, double* tplon, double* tplat, hipStream_t s, int* status = nullptr,
                   const int* nvalid = nullptr, int nvalid_off = 0) {
  Engine& e = E();
  TraceArgs A = trace_args(ctl, atm, rays, sca_on, cap);
  A.nvalid = nvalid;
  A.nvalid_off = nvalid_off;
  const int R = A.R;
  const size_t npt = (size_t)rays.nray * cap;
  A.geo = (double4*)tb.geo->need(npt * sizeof(double4));
  A.geo2 = A.sca_n > 0 ? (double4*)tb.geo2->need(npt * sizeof(double4)) : nullptr;
  A.rec = (float*)tb.rec->need(npt * R * sizeof(float));
  A.np_out = (int*)tb.npts->need(rays.nray * sizeof(int));
  A.tsurf_out = (float*)tb.tsurf->need(rays.nray * sizeof(float));
  A.tsurf64 = tb.tsurf64 ? (double*)tb.tsurf64->need(rays.nray * sizeof(double)) : nullptr;
  A.geo_final = tb.geo_final ? (int*)tb.geo_final->need(rays.nray * sizeof(int)) : nullptr;
  A.tpz = tpz;
  A.tplon = tplon;
  A.tplat = tplat;
  A.rec64 = nullptr;
  if (engine_option_int("fp64", 0)) {
    A.rec64 = (double*)e.rec64.need(npt * R * sizeof(double));
    if (!A.tsurf64) A.tsurf64 = (double*)e.tsurf64.need(rays.nray * sizeof(double));
  }
  e.last_rec64 = A.rec64;
  e.last_tsurf64 = A.tsurf64;
  e.last_nray = rays.nray;
  e.last_cap = cap;
  if (status) {
    A.status = status;
  } else {
    A.status = (int*)e.status.need(sizeof(int) * 4);
    launch_zero_int(A.status, 1, s);
  }
  // the pressure interpolation's log-slopes of this atmosphere set
  const long nlv = (long)(atm.natm > 0 ? atm.natm : 1) * atm.nlev;
  double* lps = (double*)e.lps.need((size_t)nlv * sizeof(double));
  launch_log_slopes(atm.z, atm.p, nlv, lps, s);
  A.atm.lps = lps;
  // march order: rays of similar straight-line length in one wave (option
  // "trace_sort", default 1), for calls of at least four waves per SIMD
  // (262,144 rays): with fewer every wave marches at once and the call
  // lasts as long as its longest ray whatever the order
  if (engine_option_int("trace_sort", 1) && rays.nray >= 262144) {
    int* key = (int*)e.tkey.need((size_t)rays.nray * sizeof(int));
    int* ord = (int*)e.torder.need((size_t)rays.nray * sizeof(int));
    launch_trace_key(A, key, s);
    launch_count_sort(key, rays.nray, TRACE_KEY_BINS, (int*)e.thist.need(TRACE_KEY_BINS * sizeof(int)), ord, s);
    A.order = ord;
  }
  launch_raytrace(A, s);
}
}  // namespace

void engine_trace_async(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, int sca_on, TraceBufs& tb,
                        int cap, double* tpz, double* tplon, double* tplat, int* status, const int* nvalid,
                        int nvalid_off, hipStream_t s) {
  enqueue_trace(ctl, atm, rays, sca_on, tb, cap, tpz, tplon, tplat, s, status, nvalid, nvalid_off);
  tb.cap = cap;
}

// Trace a set of rays into the given buffers; grows the per-ray capacity
// until every ray fits.  Returns the capacity used.
int engine_trace(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, int sca_on,
                 TraceBufs& tb, double* tpz, double* tplon, double* tplat, hipStream_t s) {
  Engine& e = E();
  int cap = tb.cap > 0 ? tb.cap : e.cap;
  while (true) {
    enqueue_trace(ctl, atm, rays, sca_on, tb, cap, tpz, tplon, tplat, s);
    int st = 0;
    HIP_OK(hipMemcpyAsync(&st, e.status.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (st & JUR_ST_BELOW_SURFACE) JUR_FAIL("Observer below surface!");
    if (st & JUR_ST_NLOS) JUR_FAIL("Too many LOS points!");
    if (st & JUR_ST_CAP) {
      if (cap > NLOS) JUR_FAIL("Too many LOS points!");
      cap = cap * 2 > NLOS + 2 ? NLOS + 2 : cap * 2;
      continue;
    }
    tb.cap = cap;
    if (cap > e.cap) e.cap = cap;
    return cap;
  }
}

namespace {
void begin_call(hipStream_t s) {
  Engine& e = E();
  e.path_ev_used = 0;
  e.ev_ng = -1;
  launch_zero_u64((unsigned long long*)e.pts_acc.need(sizeof(unsigned long long)), s);
}
// After the stream has reached the end of the call: path-kernel time and
// the LOS points integrated, summed over the call's launches.
void end_call(hipStream_t s) {
  Engine& e = E();
  double ms = 0;
  for (size_t i = 0; i < e.path_ev_used; i++) {
    float t = 0;
    HIP_OK(hipEventElapsedTime(&t, e.path_ev[i].first, e.path_ev[i].second));
    ms += t;
  }
  unsigned long long pts = 0, ev[24] = {0};
  HIP_OK(hipMemcpyAsync(&pts, e.pts_acc.p, sizeof pts, hipMemcpyDeviceToHost, s));
  if (e.ev_ng >= 0) HIP_OK(hipMemcpyAsync(ev, e.ev_acc.p, sizeof ev, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  // a replayed graph has no per-launch events
  e.stats["path_ms"] = e.path_ev_used ? ms : NAN;
  e.stats["path_launches"] = (double)e.path_ev_used;
  e.stats["integrated_points"] = (double)pts;
  // (point, emitter) EGA evaluations made: with base-ray reuse the copies'
  // own and every point of the bases; else NaN (every integrated point
  // evaluates every emitter)
  e.stats["emitter_points"] = e.ev_ng >= 0 ? (double)ev[0] + (double)e.ev_ng * (double)ev[1] : NAN;
  // the reuse pass as its waves run it (per channel): point, emitter and
  // clear-air optics iterations
  e.stats["reuse_wave_points"] = e.ev_ng >= 0 ? (double)ev[2] : NAN;
  e.stats["reuse_wave_emitters"] = e.ev_ng >= 0 ? (double)ev[3] : NAN;
  e.stats["reuse_wave_optics"] = e.ev_ng >= 0 ? (double)ev[4] : NAN;
  e.stats["reuse_all_points"] = e.ev_ng >= 0 ? (double)ev[23] : NAN;  // LOS points of every ray
  for (int k = 0; k <= 8; k++) {
    e.stats["reuse_waves_e" + std::to_string(k)] = e.ev_ng >= 0 ? (double)ev[5 + 2 * k] : NAN;
    e.stats["reuse_wave_points_e" + std::to_string(k)] = e.ev_ng >= 0 ? (double)ev[6 + 2 * k] : NAN;
  }
  // running totals over calls (a retrieval makes many)
  e.stats["calls_total"] += 1;
  e.stats["path_ms_total"] += e.path_ev_used ? ms : engine_stat("scatter_ms");
  e.stats["integrated_points_total"] += (double)pts;
  e.stats["emitter_points_total"] += e.ev_ng >= 0 ? e.stats["emitter_points"] : (double)e.call_ng * (double)pts;
  e.stats["pencils_total"] += e.stats["pencils"];
}
}  // namespace

namespace {
// Every address, size and setting a captured clear-air call depends on:
// inputs, outputs, work buffers, the tables and layer optics it reads, the
// ray-tracing keys, the layer geometry copied into the trace arguments and
// the kernel options.  A change of the aerosol layers also drops the graph
// (engine_set_aero).
std::string graph_key(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, const OutDev& out,
                      int scattering, hipStream_t s) {
  Engine& e = E();
  char k[2048];
  int n = snprintf(k, sizeof k, "%s|%s|%d %d %d %d|%p %p %p %p %p|%p %p %p %p %p %p %p %p|%p %p %p %p %p %p|"
                   "%p %p %p %p %p %p %p %p %p %p %p %p|%p|%d %.17g %.17g %.17g %d %d|%s|%d %d %d|",
                   e.tbl_key.c_str(), e.chan_key.c_str(), rays.nray, atm.natm, atm.nlev, scattering,
                   (const void*)atm.z, (const void*)atm.p, (const void*)atm.t, (const void*)atm.q,
                   (const void*)atm.k, (const void*)rays.time, (const void*)rays.obsz, (const void*)rays.obslon,
                   (const void*)rays.obslat, (const void*)rays.vpz, (const void*)rays.vplon,
                   (const void*)rays.vplat, (const void*)rays.atm_idx, (void*)out.rad, (void*)out.tau,
                   (void*)out.tpz, (void*)out.tplon, (void*)out.tplat, (void*)out.npts, e.geo.p, e.geo2.p, e.rec.p,
                   e.npts.p, e.tsurf.p, e.order.p, e.hist.p, e.status.p, e.pts_acc.p, e.pack.p, e.lps.p, e.torder.p, (void*)s, ctl->sca_n,
                   ctl->rayds, ctl->raydz, ctl->transs, ctl->refrac, ctl->nw + 1000 * out.ld, ctl->sca_ext,
                   engine_option_int("fast", 1), engine_option_int("variant", -1), engine_option_int("wpb", -1));
  std::string key(k, n < (int)sizeof k ? n : (int)sizeof k - 1);
  // layer geometry and optics (addresses; their contents are in the
  // snapshot, whose change drops the graph)
  snprintf(k, sizeof k, "%p %d %.17g %.17g %d|%p %p %p|%p %p %p %p", e.alti.p, e.nalti, e.altimax, e.altimin, e.nl,
           e.beta_e.p, e.beta_s.p, e.beta_a.p, e.chan.p, e.planck.p, e.t_ue.p, e.t_meta.p);
  key += k;
  for (int il = 0; il < e.nl && il < NLMAX; il++) {
    snprintf(k, sizeof k, "|%.17g %.17g %.17g", e.top[il], e.bottom[il], e.trans[il]);
    key += k;
  }
  return key;
}

}  // namespace

// graph_key plus the option / autotuner entries: what a recorded call of
// the scattering tree depends on besides its own buffers.
std::string engine_call_key(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, const OutDev& out,
                            int scattering, hipStream_t s) {
  std::string k = graph_key(ctl, atm, rays, out, scattering, s);
  char b[160];
  for (const auto& kv : E().opt) {
    snprintf(b, sizeof b, "|%s=%.17g", kv.first.c_str(), kv.second);
    k += b;
  }
  return k;
}

namespace {
void drop_forward_graph() {
  Engine& e = E();
  if (e.g_exec) HIP_OK(hipGraphExecDestroy(e.g_exec));
  if (e.g_graph) HIP_OK(hipGraphDestroy(e.g_graph));
  e.g_exec = nullptr;
  e.g_graph = nullptr;
  e.g_key.clear();
}

// Record trace -> sort -> integrate of the call just made eagerly (so the
// LOS capacity and the kernel variant are settled) as one hipGraph.
void capture_forward(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, const OutDev& out,
                     int scattering, int cap, hipStream_t s) {
  Engine& e = E();
  drop_forward_graph();
  e.hist.need((size_t)(cap + 1) * sizeof(int));  // sized before the key is taken
  TraceBufs tb{&e.geo, &e.geo2, &e.rec, &e.npts, &e.tsurf, nullptr, nullptr, cap};
  e.capturing = true;
  HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  launch_zero_u64((unsigned long long*)e.pts_acc.need(sizeof(unsigned long long)), s);
  enqueue_trace(ctl, atm, rays, ctl->sca_n > 0, tb, cap, out.tpz, out.tplon, out.tplat, s);
  engine_integrate(ctl, rays.nray, cap, e.rec.as<float>(), e.npts.as<int>(), e.tsurf.as<float>(), 0, nullptr,
                   nullptr, out.rad, out.tau, out.ld > 0 ? out.ld : rays.nray, 0, s);
  if (out.npts) HIP_OK(hipMemcpyAsync(out.npts, e.npts.p, rays.nray * sizeof(int), hipMemcpyDeviceToDevice, s));
  HIP_OK(hipStreamEndCapture(s, &e.g_graph));
  e.capturing = false;
  HIP_OK(hipGraphInstantiate(&e.g_exec, e.g_graph, nullptr, nullptr, 0));
  e.g_key = graph_key(ctl, atm, rays, out, scattering, s);
}

// Replay the captured call; false (graph dropped) when a ray overflowed the
// captured LOS capacity, and the caller then runs the call eagerly.
bool replay_forward(const RaySetDev& rays, hipStream_t s) {
  Engine& e = E();
  hipEvent_t a, b;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  HIP_OK(hipEventRecord(a, s));
  HIP_OK(hipGraphLaunch(e.g_exec, s));
  HIP_OK(hipEventRecord(b, s));
  int st = 0;
  unsigned long long pts = 0;
  HIP_OK(hipMemcpyAsync(&st, e.status.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(&pts, e.pts_acc.p, sizeof pts, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, a, b));
  HIP_OK(hipEventDestroy(a));
  HIP_OK(hipEventDestroy(b));
  if (st & JUR_ST_BELOW_SURFACE) JUR_FAIL("Observer below surface!");
  if (st & JUR_ST_NLOS) JUR_FAIL("Too many LOS points!");
  if (st & JUR_ST_CAP) {
    drop_forward_graph();
    e.stats["graph_overflows"] += 1;
    return false;
  }
  e.stats["graph_ms"] = ms;
  // events recorded inside a graph cannot be timed here: no per-kernel time
  e.stats["path_ms"] = NAN;
  e.stats["trace_ms"] = NAN;
  e.stats["path_launches"] = 1;
  e.stats["integrated_points"] = (double)pts;
  e.stats["pencils"] = rays.nray;
  e.stats["leaf_pencils"] = rays.nray;  // recorded calls are clear air / extinction (depth 0)
  e.stats["scattering_points"] = 0;
  // running totals (a replay adds its whole time)
  e.stats["calls_total"] += 1;
  e.stats["path_ms_total"] += ms;
  e.stats["integrated_points_total"] += (double)pts;
  e.stats["emitter_points_total"] += (double)e.call_ng * (double)pts;
  e.stats["pencils_total"] += rays.nray;
  return true;
}
}  // namespace

void engine_forward(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, int scattering,
                    const OutDev& out, hipStream_t s) {
  Engine& e = E();
  if (rays.nray <= 0) return;
  if (ctl->ip < 1 || ctl->ip > 3) JUR_FAIL("Unknown interpolation method, check IP!");
  if (ctl->ip != atm.geo.ip && !(ctl->ip == 1 && atm.geo.ip == 0))
    JUR_FAIL("The atmosphere was staged for another interpolation method (IP)!");
  const int sca_on = ctl->sca_n > 0;
  begin_call(s);
  if (sca_on && scattering > 0) {
    scatter_forward(ctl, atm, rays, scattering, out, s);
    end_call(s);
    return;
  }
  TraceBufs tb{&e.geo, &e.geo2, &e.rec, &e.npts, &e.tsurf, nullptr, nullptr, 0};
  // FP64 calls are not recorded (the graph key does not hold the precision)
  const bool graph = engine_option_int("graph", 0) != 0 && engine_option_int("fp64", 0) == 0 && ctl->ip == 1;
  if (graph && e.g_exec && e.g_key == graph_key(ctl, atm, rays, out, scattering, s)) {
    if (replay_forward(rays, s)) return;
  }
  hipEvent_t ev[3];
  for (auto& x : ev) HIP_OK(hipEventCreate(&x));
  HIP_OK(hipEventRecord(ev[0], s));
  int cap = engine_trace(ctl, atm, rays, sca_on, tb, out.tpz, out.tplon, out.tplat, s);
  HIP_OK(hipEventRecord(ev[1], s));
  engine_integrate(ctl, rays.nray, cap, e.rec.as<float>(), e.npts.as<int>(), e.tsurf.as<float>(), 0, nullptr,
                   nullptr, out.rad, out.tau, out.ld > 0 ? out.ld : rays.nray, 0, s, out.rad64, out.tau64);
  HIP_OK(hipEventRecord(ev[2], s));
  if (out.npts) HIP_OK(hipMemcpyAsync(out.npts, e.npts.p, rays.nray * sizeof(int), hipMemcpyDeviceToDevice, s));
  HIP_OK(hipEventSynchronize(ev[2]));
  float t1 = 0;
  HIP_OK(hipEventElapsedTime(&t1, ev[0], ev[1]));
  e.stats["trace_ms"] = t1;
  e.stats["pencils"] = rays.nray;
  // pencils the reference's queue would hold: the primaries are leaves only
  // at depth 0 (forwardmodel.c:245-262)
  e.stats["leaf_pencils"] = scattering == 0 ? rays.nray : 0;
  e.stats["scattering_points"] = 0;
  end_call(s);
  for (auto& x : ev) HIP_OK(hipEventDestroy(x));
  if (graph) capture_forward(ctl, atm, rays, out, scattering, cap, s);
}

namespace {
// Autotuner key of a call shape: emitters, channels, size class (cls),
// table layout.
std::string tune_key(const ctl_t* ctl, int nray, int shared, int* cls) {
  *cls = nray >= (1 << 16) ? 3 : nray >= (1 << 14) ? 2 : nray >= 4096 ? 1 : 0;
  char key[96];
  snprintf(key, sizeof key, "tune:%d:%d:%d:%d", ctl->ng, ctl->nd, *cls, shared);
  return key;
}

// Arguments of the path integration over traced rays (kernel variant left
// to the caller).
PathArgs path_args(const ctl_t* ctl, int nray, int cap, const float* rec, const int* npts, const float* tsurf,
                   const int* order, int scattering, const float* J, const int* jbase, float* rad, float* tau,
                   int out_ld, int out_off, double* rad64, double* tau64) {
  Engine& e = E();
  PathArgs P;
  memset(&P, 0, sizeof P);
  P.nray = nray;
  P.nd = ctl->nd;
  P.ng = ctl->ng;
  P.nw = ctl->nw;
  P.R = rec_len(ctl->ng, ctl->nw);
  P.cap = cap;
  P.rec = rec;
  P.np = npts;
  P.tsurf = tsurf;
  P.order = order;
  P.T = engine_tables();
  if (engine_option_int("fast", 1) == 0) P.T.shared = 0;
  P.cc = e.chan.as<ChanConst>();
  P.planck = e.planck.as<float>();
  P.cc64 = e.chan64.as<ChanConst64>();
  P.planck64 = e.planck64.as<double>();
  P.beta_e = e.beta_e.as<float>();
  P.beta_s = e.beta_s.as<float>();
  P.beta_a = e.beta_a.as<float>();
  P.beta_e64 = e.beta_e64.as<double>();
  P.beta_s64 = e.beta_s64.as<double>();
  P.beta_a64 = e.beta_a64.as<double>();
  P.sca_n = ctl->sca_n;
  P.ext_beta_e = strcmp(ctl->sca_ext, "beta_a") != 0;
  P.scattering = scattering;
  P.J = J;
  P.jbase = jbase;
  P.rad = rad;
  P.tau = tau;
  P.out_ld = out_ld;
  P.out_off = out_off;
  P.fp64 = engine_option_int("fp64", 0) != 0;
  P.fp64_split = engine_option_int("fp64_split", -1);
  P.rad64 = P.fp64 ? rad64 : nullptr;
  P.tau64 = P.fp64 ? tau64 : nullptr;
  // the FP64 records belong to the trace just made of these rays
  P.rec64 = (P.fp64 && e.last_nray == nray && e.last_cap == cap) ? e.last_rec64 : nullptr;
  P.tsurf64 = P.rec64 ? e.last_tsurf64 : nullptr;
  return P;
}
}  // namespace

// Sort rays by point count and run the path integration.
void engine_integrate(const ctl_t* ctl, int nray, int cap, const float* rec, const int* npts,
                      const float* tsurf, int scattering, const float* J, const int* jbase, float* rad,
                      float* tau, int out_ld, int out_off, hipStream_t s, double* rad64, double* tau64,
                      long nJ) {
  Engine& e = E();
  e.call_ng = ctl->ng;
  int* order = (int*)e.order.need(nray * sizeof(int));
  // FP64 reuse of base rays (engine_set_prefix): clear air, FP64 records of
  // this trace, rays = copies of the first pnr, snapshot buffer of modest
  // size
  const int R_ = rec_len(ctl->ng, ctl->nw);
  const bool fp64 = engine_option_int("fp64", 0) != 0;
  const int pnr = e.prefix_nr;
  const size_t snap_bytes =
      (size_t)(pnr > 0 ? pnr : 0) * ctl->nd * (size_t)(cap + 1) * (2 * ctl->ng + 4) * sizeof(double);
  const bool prefix = fp64 && scattering == 0 && pnr > 0 && nray > pnr && nray % pnr == 0 && !e.capturing &&
                      e.last_rec64 && e.last_nray == nray && e.last_cap == cap && snap_bytes < ((size_t)1 << 30);
  int *pstart = nullptr, *prem = nullptr, *fdiff = nullptr, *fcont = nullptr, *pevals = nullptr, *pkeys = nullptr;
  if (prefix) {
    pstart = (int*)e.pstart.need(nray * sizeof(int));
    prem = (int*)e.prem.need(nray * sizeof(int));
    fdiff = (int*)e.fdiff.need((size_t)nray * (ctl->ng > 0 ? ctl->ng : 1) * sizeof(int));
    fcont = (int*)e.fcont.need(nray * sizeof(int));
    pevals = (int*)e.pevals.need(nray * sizeof(int));
    pkeys = (int*)e.pkeys.need(nray * sizeof(int));
    // group a wave's lanes by base ray and the emitters they compute (option
    // "reuse_group", default 1; the key packs (class, base) into 19 bits)
    const int keyed = engine_option_int("reuse_group", 1) != 0 && (long)(NGMAX + 2) * pnr < (1L << 19);
    // emitters whose column amounts enter the continua
    unsigned long long cmask = 0;
    for (const char* name : {"CO2", "H2O"}) {
      const int ig = find_emitter(ctl, name);
      if (ig >= 0) cmask |= 1ull << ig;
    }
    launch_reuse_start(nray, pnr, R_, ctl->ng, ctl->nw, cap, cmask, e.last_rec64, npts, pstart, prem, fdiff, fcont,
                       pevals, pkeys, keyed, s);
  }
  if (e.capturing || e.planned) {
    // counting sort: kernels only (no library temporaries or memsets)
    launch_count_sort(npts, nray, cap + 1, (int*)e.hist.need((size_t)(cap + 1) * sizeof(int)), order, s);
  } else {
    int* ids = (int*)e.ids.need(nray * sizeof(int));
    int* keys = (int*)e.keys.need(nray * sizeof(int));
    size_t tb = sort_temp_bytes(nray);
    void* tmp = e.sorttmp.need(tb);
    launch_iota(ids, nray, s);
    // by the points each ray integrates, or for base-ray reuse by the key above
    launch_sort_rays(prefix ? pkeys : npts, nullptr, keys, ids, order, tmp, tb, nray, s);
  }
  PathArgs P = path_args(ctl, nray, cap, rec, npts, tsurf, order, scattering, J, jbase, rad, tau, out_ld, out_off,
                         rad64, tau64);
  P.nJ = J ? nJ : 0;
  // packed point records and (p, T) brackets for the FP32 kernel variant
  // that reads them (R + 4 ng floats per LOS point slot of every ray group)
  if (!P.fp64 && P.T.shared && ctl->ng > 0)
    P.pack = (float*)e.pack.need((size_t)(nray + 63) / 64 * 64 * cap * (P.R + 4 * ctl->ng) * sizeof(float));
  int variant = engine_option_int("variant", -1), wpb = engine_option_int("wpb", -1);
  if (P.fp64) {
    // the FP64 kernels have one configuration (variant and waves per block
    // apply to the FP32 kernels only)
    if (variant < 0) variant = 2;
    if (wpb < 0) wpb = 4;
  }
  if (variant < 0 || wpb < 0) {
    // Autotuner: the first large batch of a (emitters, channels, size class)
    // shape times each kernel variant (channels per wave x occupancy) and
    // waves per block once, with HIP events on this stream, and keeps the
    // fastest.  Small batches use the default.
    int cls = 0;
    const std::string key = tune_key(ctl, nray, P.T.shared, &cls);
    auto it = e.opt.find(key);
    int best = 2 * 8 + 4;  // variant 2, 4 waves per block
    if (it != e.opt.end()) {
      best = (int)it->second;
    } else if (cls >= 2 && engine_option_int("autotune", 1) && !e.capturing && !e.planned) {
      hipEvent_t a, b;
      HIP_OK(hipEventCreate(&a));
      HIP_OK(hipEventCreate(&b));
      float tbest = 1e30f;
      const int wpbs[3] = {1, 2, 4};
      for (int v = 0; v < path_variants(); v++)
        for (int w : wpbs) {
          P.cpw = v;
          HIP_OK(hipEventRecord(a, s));
          launch_path(P, w, s);
          HIP_OK(hipEventRecord(b, s));
          HIP_OK(hipEventSynchronize(b));
          float ms = 0;
          HIP_OK(hipEventElapsedTime(&ms, a, b));
          if (ms < tbest) {
            tbest = ms;
            best = v * 8 + w;
          }
        }
      HIP_OK(hipEventDestroy(a));
      HIP_OK(hipEventDestroy(b));
      e.opt[key] = best;
      e.stats["tuned_variant"] = best / 8;
      e.stats["tuned_wpb"] = best % 8;
      e.stats["tuned_ms"] = tbest;
    }
    if (variant < 0) variant = best / 8;
    if (wpb < 0) wpb = best % 8;
  }
  P.cpw = variant;
  // LOS points integrated: with prefix reuse the base rays' and every ray's
  // remainder
  unsigned long long* acc = (unsigned long long*)e.pts_acc.need(sizeof(unsigned long long));
  if (prefix) {
    launch_sum_points(prem, nray, acc, s);
    launch_sum_points(npts, pnr, acc, s);
    // emitter evaluations: the copies' (those not taken from the base) and
    // the bases' (all emitters at every point, added on the host)
    unsigned long long* ev = (unsigned long long*)e.ev_acc.need(24 * sizeof(unsigned long long));
    launch_zero_u64(ev, s, 24);
    launch_sum_points(pevals, nray, ev, s);
    launch_sum_points(npts, pnr, ev + 1, s);
    launch_reuse_waves(order, pstart, npts, fdiff, fcont, nray, ctl->ng, ev + 2, s);
    launch_sum_points(npts, nray, ev + 23, s);
    e.ev_ng = ctl->ng;  // read at the end of the call
  } else {
    launch_sum_points(npts, nray, acc, s);
  }
  if (e.capturing) {
    launch_path(P, wpb, s);
    return;
  }
  if (e.path_ev.size() <= e.path_ev_used) {
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    e.path_ev.push_back({a, b});
  }
  auto& ev = e.path_ev[e.path_ev_used++];
  HIP_OK(hipEventRecord(ev.first, s));
  if (prefix) {
    // the base rays first, writing their path state at every point; then
    // every ray from its first differing point (base rays: their end)
    PathArgs B = P;
    int* pids = (int*)e.pids.need(pnr * sizeof(int));
    launch_iota(pids, pnr, s);
    B.nray = pnr;
    B.order = pids;
    B.snap_mode = 1;
    B.pnr = pnr;
    B.snap = (double*)e.snap.need(snap_bytes);
    launch_path(B, wpb, s);
    P.snap_mode = 2;
    P.pnr = pnr;
    P.pstart = pstart;
    P.fdiff = fdiff;
    P.fcont = fcont;
    P.snap = B.snap;
  }
  launch_path(P, wpb, s);
  HIP_OK(hipEventRecord(ev.second, s));
}

namespace {
// Bump allocation of 256-byte aligned pieces from a fixed region.
struct Carve {
  char* p;
  size_t left;
  void* take(size_t n) {
    const size_t a = (n + 255) & ~(size_t)255;
    if (a > left) return nullptr;
    void* r = p;
    p += a;
    left -= a;
    return r;
  }
};
}  // namespace

int engine_forward_noalloc(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, const OutDev& out,
                           void* base, size_t bytes, hipStream_t s) {
  const int n = rays.nray;
  if (n <= 0) return 0;
  const int sca_on = ctl->sca_n > 0;
  const int R = rec_len(ctl->ng, ctl->nw);
  Carve c{(char*)base, bytes};
  int* npts = (int*)c.take((size_t)n * sizeof(int));
  float* tsurf = (float*)c.take((size_t)n * sizeof(float));
  int* order = (int*)c.take((size_t)n * sizeof(int));
  int* status = (int*)c.take(4 * sizeof(int));
  if (!npts || !tsurf || !order || !status) return -1;
  // per LOS point: march geometry, geometry after layer-edge insertion,
  // FP32 record; per ray and capacity step a histogram bin of the sort
  const size_t per_pt = sizeof(double4) * (sca_on ? 2 : 1) + (size_t)R * sizeof(float);
  const size_t room = c.left > 4 * 256 ? c.left - 4 * 256 : 0;  // alignment slack of the four pieces
  const long fit_l = (long)((room / ((size_t)n * per_pt + sizeof(int))) - 1);
  const int fit = (int)(fit_l > NLOS + 2 ? NLOS + 2 : fit_l);
  if (fit < 16) return -1;
  int cap = engine_cap();
  if (cap > fit) cap = fit;
  TraceArgs A = trace_args(ctl, atm, rays, sca_on, cap);
  while (true) {
    Carve cc = c;
    const size_t npt = (size_t)n * cap;
    A = trace_args(ctl, atm, rays, sca_on, cap);
    A.geo = (double4*)cc.take(npt * sizeof(double4));
    A.geo2 = sca_on ? (double4*)cc.take(npt * sizeof(double4)) : nullptr;
    A.rec = (float*)cc.take(npt * R * sizeof(float));
    int* hist = (int*)cc.take((size_t)(cap + 1) * sizeof(int));
    if (!A.geo || (sca_on && !A.geo2) || !A.rec || !hist) return -1;
    A.np_out = npts;
    A.tsurf_out = tsurf;
    A.tpz = out.tpz;
    A.tplon = out.tplon;
    A.tplat = out.tplat;
    A.status = status;
    launch_zero_int(status, 1, s);
    launch_raytrace(A, s);
    int st = 0;
    HIP_OK(hipMemcpyAsync(&st, status, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (st & JUR_ST_BELOW_SURFACE) JUR_FAIL("Observer below surface!");
    if (st & JUR_ST_NLOS) JUR_FAIL("Too many LOS points!");
    if (st & JUR_ST_CAP) {
      if (cap > NLOS) JUR_FAIL("Too many LOS points!");
      if (cap >= fit) return -1;  // the rays need more LOS points than the region holds
      cap = cap * 2 > fit ? fit : cap * 2;
      continue;
    }
    launch_count_sort(npts, n, cap + 1, hist, order, s);
    break;
  }
  PathArgs P = path_args(ctl, n, cap, A.rec, npts, tsurf, order, 0, nullptr, nullptr, out.rad, out.tau,
                         out.ld > 0 ? out.ld : n, 0, nullptr, nullptr);
  // the variant the autotuner chose for this shape, if any, else the default
  // (no timing runs here)
  int variant = engine_option_int("variant", -1), wpb = engine_option_int("wpb", -1);
  int cls = 0, best = 2 * 8 + 4;
  auto it = E().opt.find(tune_key(ctl, n, P.T.shared, &cls));
  if (it != E().opt.end()) best = (int)it->second;
  P.cpw = variant >= 0 ? variant : best / 8;
  launch_path(P, wpb >= 0 ? wpb : best % 8, s);
  if (out.npts) HIP_OK(hipMemcpyAsync(out.npts, npts, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s));
  HIP_OK(hipStreamSynchronize(s));
  return cap;
}

void engine_set_prefix(int nr) { E().prefix_nr = nr; }
void engine_set_mode(bool planned, bool capturing) {
  E().planned = planned;
  E().capturing = capturing;
}
bool engine_capturing() { return E().capturing; }
void engine_begin_call(hipStream_t s) { begin_call(s); }
void engine_end_call(hipStream_t s) { end_call(s); }

extern "C" int jur_set_option(const char* key, double value) {
  Engine& e = E();
  e.opt[key] = value;
  if (strcmp(key, "cap") == 0) e.cap = value < 16 ? 16 : (value > NLOS + 2 ? NLOS + 2 : (int)value);
  return 0;
}
extern "C" double jur_stat(const char* key) { return engine_stat(key); }
extern "C" int jur_init(int device) {
  Engine& e = E();
  if (device != e.device) {
    e.device = device;
    HIP_OK(hipSetDevice(device));
    HIP_OK(hipStreamDestroy(e.stream));
    HIP_OK(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
    size_t fr = 0;
    HIP_OK(hipMemGetInfo(&fr, &e.hbm_bytes));
    e.tbl_key.clear();
    e.chan_key.clear();
    e.aero_ready = false;
  }
  noalloc_scratch(false);  // this device's no-allocation work region (looked up here once)
  return 0;
}
extern "C" int jur_sync(void) {
  Engine& e = E();
  HIP_OK(hipStreamSynchronize(e.stream));
  if (e.status.p) {
    int st = 0;
    HIP_OK(hipMemcpy(&st, e.status.p, sizeof(int), hipMemcpyDeviceToHost));
    if (st & JUR_ST_BELOW_SURFACE) JUR_FAIL("Observer below surface!");
    if (st & (JUR_ST_NLOS | JUR_ST_CAP)) JUR_FAIL("Too many LOS points!");
    if (st & JUR_ST_QUEUE) JUR_FAIL("Too many queue items!");
  }
  return 0;
}

DevTables engine_tables() {
  Engine& e = E();
  DevTables T;
  T.np = e.t_np.as<int>();
  T.p = e.t_p.as<float>();
  T.nt = e.t_nt.as<int>();
  T.t = e.t_t.as<float>();
  T.ncol = e.t_ncol.as<int>();
  T.col = e.t_col.as<int>();
  T.ue = e.t_ue.as<float2>();
  T.ue_bytes = e.ue_bytes;
  T.shared = e.shared;
  T.gnp = e.g_np.as<int>();
  T.gp = e.g_p.as<float>();
  T.gnt = e.g_nt.as<int>();
  T.gt = e.g_t.as<float>();
  T.meta = e.t_meta.as<int2>();
  T.meta_bytes = e.meta_bytes;
  T.gnp_max = e.gnp_max;
  T.gnt_max = e.gnt_max;
  T.p64 = e.t_p64.as<double>();
  T.t64 = e.t_t64.as<double>();
  return T;
}
const double* engine_nu_dev() { return E().nu_d.as<double>(); }
int engine_find_emitter(const ctl_t* ctl, const char* name) { return find_emitter(ctl, name); }
const float* engine_beta_e() { return E().beta_e.as<float>(); }
const float* engine_beta_s() { return E().beta_s.as<float>(); }
void engine_layer_edges(double* top, double* bottom) {
  Engine& e = E();
  for (int il = 0; il < NLMAX; il++) {
    top[il] = e.top[il];
    bottom[il] = e.bottom[il];
  }
}
double engine_option_double(const char* key, double dflt) {
  auto& m = E().opt;
  auto it = m.find(key);
  return it == m.end() ? dflt : it->second;
}
AeroGeom engine_aero_geom() {
  Engine& e = E();
  return {e.alti.as<double>(), e.nalti, e.altimax, e.altimin};
}
int engine_cap() { return E().cap; }
