// kernels_path.hip -- path integration: emissivity growth approximation on
// the look-up tables, continua, Planck source and the radiance /
// transmittance recursion along each line of sight (reference
// src/forwardmodel.c:303-399 with intpol_tbl :409-565, formod_continua
// :120-152 and srcfunc_planck :743-794), in FP32 with Kahan-compensated
// radiance.
//
// Work mapping: one wavefront = 64 rays (lanes) of the ray list sorted by
// descending point count, times CPW channels (wave-uniform).  All lanes of a
// wave read the SAME emissivity tables (those of its channels), and
// neighbouring rays in the sorted list traverse similar altitudes, so the
// table columns a wave touches overlap strongly in L1/L2.  Blocks are
// mapped XCD-aware: the blocks covering the channel chunks of one ray group
// are dispatched to the same XCD so they share its L2 copy of the group's
// point records.
//
// EGA in FP32: per (channel, emitter) the lane carries the cumulative path
// emissivity ep (= 1 - tau_path of the reference) and the path
// transmittance tp.  The inverse lookup uses ep directly, and the segment
// emissivity is formed as (eps_ext - ep) / tp, which equals the reference's
// 1 - (1 - eps_ext) / tau_path without subtracting numbers close to one.
#include <hipcub/hipcub.hpp>

#include "jur_dev.h"
#include "jur_kernels.h"

namespace {

// Emissivity and its complement 1 - eps, carried side by side through the
// interpolations: eps is exact for transparent paths, the complement for
// nearly opaque ones (1 - eps of FP32 table nodes >= 0.5 is exact).
struct EC {
  float e, c;
};

// GSL_MAX(GSL_MIN(eps, 1), 0) on the pair; NaN becomes 1 like the reference:
// eps -> NaN ? 1 : eps <= 0 ? 0 : 1 - eps <= 0 ? 1 : min(eps, 1), and the
// complement to match.  Written with the NaN test last (a NaN eps fails
// eps <= 0, and fminf(NaN, 1) is 1), which needs no NaN-dependent constants.
__device__ inline EC clamp_ec(float e, float c) {
  const bool lo = e <= 0.f, hi = c <= 0.f, nan = !(e == e);
  const float re = lo ? 0.f : (hi ? 1.f : fminf(e, 1.f));
  const float rc = lo ? 1.f : ((hi || nan) ? 0.f : c);
  return {re, rc};
}

// LIN between two (eps, 1 - eps) nodes.
__device__ inline EC lin_ec(float x0, EC y0, float x1, EC y1, float x) {
  float w = (x - x0) / (x1 - x0);
  return clamp_ec(y0.e + (y1.e - y0.e) * w, y0.c + (y1.c - y0.c) * w);
}

// One table corner: u at the path emissivity (inverse lookup), then the
// emissivity of the extended path (forward lookup), clamped to [0, 1]
// (forwardmodel.c:460-474, 506-542).
__device__ inline EC corner(const float2* col, int n, float ep, float du) {
  int i = locate_col<true>(col, n, ep);
  float2 a = col[i], b = col[i + 1];
  float u = lin(a.y, a.x, b.y, b.x, ep);
  float ux = u + du;
  int j = locate_col<false>(col, n, ux);
  float2 c = col[j], d = col[j + 1];
  float s = (d.y - c.y) / (d.x - c.x), dx = ux - c.x;
  return clamp_ec(c.y + s * dx, (1.f - c.y) - s * dx);
}

// One EGA step of one emitter (forwardmodel.c:433-500).  ep, tp: path
// emissivity and transmittance.  Returns the segment emissivity `seg` and
// transmittance `fac` (= 1 - seg), each formed in the precise domain: the
// reference's eps = 1 - (1 - eps_ext) / tau_path.
__device__ inline void ega_step(const DevTables& T, int tb, float p, float t, float du, float& ep,
                                float& tp, float& seg, float& fac) {
  const int ntp = T.np[tb];
  seg = 0.f;
  fac = 1.f;
  if (ntp < 2) return;
  if (tp < 1e-9f) {
    seg = 1.f;
    fac = 0.f;
    ep = 1.f;
    tp = 0.f;
    return;
  }
  const float* pg = T.p + (size_t)tb * TBLNPMAX;
  const int ipr = locate_mono<float>(pg, ntp, p);
  const size_t b0 = (size_t)tb * TBLNPMAX + ipr;
  const int nt0 = T.nt[b0], nt1 = T.nt[b0 + 1];
  if (nt0 < 2 || nt1 < 2) return;
  const float* t0g = T.t + b0 * TBLNTMAX;
  const float* t1g = t0g + TBLNTMAX;
  const int it0 = locate_mono<float>(t0g, nt0, t);
  const int it1 = locate_mono<float>(t1g, nt1, t);
  const size_t k00 = b0 * TBLNTMAX + it0, k10 = (b0 + 1) * TBLNTMAX + it1;
  const int n00 = T.ncol[k00], n01 = T.ncol[k00 + 1], n10 = T.ncol[k10], n11 = T.ncol[k10 + 1];
  if (n00 < 2 || n01 < 2 || n10 < 2 || n11 < 2) return;
  EC e00 = corner(T.ue + T.col[k00], n00, ep, du);
  EC e01 = corner(T.ue + T.col[k00 + 1], n01, ep, du);
  EC e10 = corner(T.ue + T.col[k10], n10, ep, du);
  EC e11 = corner(T.ue + T.col[k10 + 1], n11, ep, du);
  e00 = lin_ec(t0g[it0], e00, t0g[it0 + 1], e01, t);
  e11 = lin_ec(t1g[it1], e10, t1g[it1 + 1], e11, t);
  e00 = lin_ec(pg[ipr], e00, pg[ipr + 1], e11, p);
  if (e00.e < 0.5f) {
    seg = (e00.e - ep) / tp;
    fac = 1.f - seg;
  } else {
    fac = e00.c / tp;
    seg = 1.f - fac;
  }
  ep = e00.e;
  tp = e00.c;
}

__device__ inline void kahan_add(float& sum, float& comp, float term) {
  float y = term - comp;
  float s = sum + y;
  comp = (s - sum) - y;
  sum = s;
}

// Continuum extinction of one point for one channel [1/km]
// (forwardmodel.c:120-152; continua.c ctmco2 :854-869, ctmh2o :1898-1928,
// ctmn2 :1980-1992, ctmo2 :2037-2048).
__device__ inline float continua(const ChanConst& c, float p, float t, float ds, float qh2o,
                                 float uco2, float uh2o, float kw, float rt, float rds) {
  float beta = kw;
  // A zero-length segment (an aerosol edge point inserted onto the
  // observer; u = 0) adds no CO2/H2O continuum: the reference's u/ds is
  // 0/0 there.
  const bool seg = ds > 0.f;
  if (c.co2_on && seg) {
    float d230 = t - 230.f, d260 = t - 260.f, d296 = t - 296.f;
    float ctw = d260 * d296 * c.co2_k230 - d230 * d296 * c.co2_k260 + d230 * d260 * c.co2_k296;
    beta += uco2 * (float)(1.0 / (JUR_AVOGADRO * 1000.0 * P0)) * p * ctw * rds;
  }
  if (c.h2o_on && seg) {
    float ctwslf = c.h2o_s296 * expf(c.h2o_lnratio * ((296.f - t) * (1.f / 36.f)));
    float a1 = c.nu * uh2o * tanhf(c.h2o_tanhk * rt);
    float a2 = 296.f * rt;
    float a3 = p * (float)(1.0 / P0) * (qh2o * ctwslf + (1.f - qh2o) * c.h2o_frn) * 1e-20f;
    beta += a1 * a2 * a3 * rds;
  }
  if (c.n2_on || c.o2_on) {
    float pp = p * (float)(1.0 / P0), tt = 273.f * rt;
    float base = 0.1f * pp * pp * tt * tt;
    float ex = (1.f / 296.f - rt);
    if (c.n2_on)
      beta += base * 0.79f * c.n2_b * expf(c.n2_beta * ex) *
              (0.79f + 0.21f * (1.294f - 0.4545f * t / 296.f));
    if (c.o2_on) beta += base * 0.21f * c.o2_b * expf(c.o2_beta * ex);
  }
  return beta;
}

// Planck table row lookup; the temperature grid is 100 + 0.25 i, i < 1201
// (forwardmodel.c:771-793: locate on that grid, then linear).
__device__ inline void planck_index(float t, int& it, float& dt) {
  float x = (t - 100.f) * 4.f;
  if (!(x == x)) {
    it = JUR_NPLANCK - 2;
  } else {
    x = floorf(x);
    it = x < 0.f ? 0 : (x > (float)(JUR_NPLANCK - 2) ? JUR_NPLANCK - 2 : (int)x);
  }
  dt = t - (100.f + 0.25f * (float)it);
}
__device__ inline float planck_at(const float* row, int it, float dt) {
  float y0 = row[it], y1 = row[it + 1];
  return y0 + (y1 - y0) / 0.25f * dt;
}

// Work of a wave: ray group (64 consecutive rays of the sorted list) and
// channel chunk (CPW channels).  A block's waves take consecutive ray
// groups of ONE channel chunk, so they read the same tables (shared L1);
// logical blocks are chunk-major and the hardware's round-robin over the 8
// XCDs is undone (L = xcd * per + slot), so each XCD works on a contiguous
// range of channel chunks and its L2 holds only those channels' tables.
__device__ inline bool map_work(const PathArgs& A, int cpw, int& chunk, int& group) {
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wpb = blockDim.x >> 6;
  const int per = gridDim.x >> 3;
  const int L = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const int ngroups = (A.nray + 63) >> 6;
  const int bpc = (ngroups + wpb - 1) / wpb;  // blocks per channel chunk
  chunk = L / bpc;
  group = (L - chunk * bpc) * wpb + wv;
  const int nchunk = (A.nd + cpw - 1) / cpw;
  return chunk < nchunk && group < ngroups;
}

// Tiled variant: each XCD's contiguous range of logical blocks cycles
// through TC = nchunk / 8 channel chunks for one ray group before the next
// group, so the point records a ray group reads are shared in the XCD's L2
// by its TC channel chunks (read once per XCD instead of once per chunk),
// while the XCD's table working set grows to TC chunks' tables.
__device__ inline bool map_work_tiled(const PathArgs& A, int cpw, int& chunk, int& group) {
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wpb = blockDim.x >> 6;
  const int per = gridDim.x >> 3;
  const int L = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const int ngroups = (A.nray + 63) >> 6;
  const int bpc = (ngroups + wpb - 1) / wpb;  // blocks per channel chunk
  const int nchunk = (A.nd + cpw - 1) / cpw;
  const int TC = nchunk >= 16 ? nchunk >> 3 : 1;
  const int tile = L / (TC * bpc), r = L - tile * TC * bpc;
  const int tc = min(TC, nchunk - tile * TC);  // chunks in this tile
  if (tc <= 0) return false;
  chunk = tile * TC + r % tc;
  group = (r / tc) * wpb + wv;
  return chunk < nchunk && group < ngroups && r < tc * bpc;
}

// Radiance and transmittance update of one ray x channel at one LOS point
// (forwardmodel.c:332-389) given the gas segment transmittance tg = tau_gas
// and its complement et = 1 - tau_gas.  pr: the point's record, fields rs
// floats apart.
__device__ inline void point_radiance(const PathArgs& A, const ChanConst& cc, int id, const float* pr, int uo,
                                      float p, float t, float ds, float qh, float af, int ai, bool scpt, int kj,
                                      int itp, float dtp, float tg, float et, float rt, float rds, float& radv,
                                      float& comp, float& tau, int rs = 1) {
  const float uco2 = cc.co2_on ? pr[(uo + cc.co2_ig) * rs] : 0.f;
  const float uh2o = cc.h2o_on ? pr[(uo + cc.h2o_ig) * rs] : 0.f;
  const float kw = A.nw > 0 ? pr[(REC_K + cc.window) * rs] : 0.f;
  const float beta = continua(cc, p, t, ds, qh, uco2, uh2o, kw, rt, rds);
  const float B = planck_at(A.planck + (size_t)id * JUR_NPLANCK, itp, dtp);
  if (!(tg > 0.f)) return;
  if (scpt) {
    // a zero-length segment adds nothing (the reference forms -log(1)/0)
    if (!(ds > 0.f)) return;
    // scattering source (forwardmodel.c:332-358)
    const float be = A.beta_e[ai * A.nd + id], bs = A.beta_s[ai * A.nd + id];
    const float J = kj < A.nJ ? A.J[(size_t)kj * A.nd + id] : 0.f;  // a batch over its planned capacity is re-run
    const float lg = et < 0.5f ? -log1pf(-et) : -logf(tg);
    const float bt = lg / ds + beta + af * be;
    const float x = bt * ds;
    const float eps = -expm1f(-x);
    const float src = ((bt - af * bs) * B + af * bs * J) / bt;
    kahan_add(radv, comp, tau * eps * src);
    tau *= expf(-x);
  } else {
    // gas, continua and aerosol extinction (forwardmodel.c:365-388)
    float bx = beta;
    if (A.sca_n > 0 && af != 0.f)
      bx += af * (A.ext_beta_e ? A.beta_e[ai * A.nd + id] : A.beta_a[ai * A.nd + id]);
    const float x = bx * ds;
    const float eps = et - tg * expm1f(-x);
    kahan_add(radv, comp, B * eps * tau);
    tau *= tg * expf(-x);
  }
}

// Surface term and store (forwardmodel.c:393-399).
template <int CPW>
__device__ inline void finish_ray(const PathArgs& A, int ray, int c0, const float* radv, float* comp,
                                  const float* tau) {
  const float ts = A.tsurf[ray];
  int its = 0;
  float dts = 0.f;
  if (ts > 0.f) planck_index(ts, its, dts);
#pragma unroll
  for (int c = 0; c < CPW; c++) {
    const int id = c0 + c;
    if (id < A.nd) {
      float r = radv[c];
      if (ts > 0.f) kahan_add(r, comp[c], planck_at(A.planck + (size_t)id * JUR_NPLANCK, its, dts) * tau[c]);
      A.rad[(size_t)id * A.out_ld + A.out_off + ray] = r;
      A.tau[(size_t)id * A.out_ld + A.out_off + ray] = tau[c];
    }
  }
}

template <int NGT, int CPW>
__global__ void __launch_bounds__(256) k_path(PathArgs A) {
  const int lane = threadIdx.x & 63;
  int chunk, group;
  if (!map_work(A, CPW, chunk, group)) return;
  const int slot = group * 64 + lane;
  const bool valid = slot < A.nray;
  const int ray = valid ? A.order[slot] : 0;
  const int npts = valid ? A.np[ray] : 0;
  const int c0 = chunk * CPW;
  const int uo = rec_u(A.nw);

  float ep[CPW][NGT], tp[CPW][NGT], radv[CPW], comp[CPW], tau[CPW];
#pragma unroll
  for (int c = 0; c < CPW; c++) {
#pragma unroll
    for (int g = 0; g < NGT; g++) {
      ep[c][g] = 0.f;
      tp[c][g] = 1.f;
    }
    radv[c] = 0.f;
    comp[c] = 0.f;
    tau[c] = 1.f;
  }
  int kj = (valid && A.scattering > 0) ? A.jbase[ray] : 0;
  const float* rp = A.rec + (size_t)ray * A.cap * A.R;

  for (int ip = 0;; ip++) {
    const bool act = ip < npts;
    if (!__any(act)) break;
    if (act) {
      const float* pr = rp + (size_t)ip * A.R;
      const float p = pr[REC_P], t = pr[REC_T], ds = pr[REC_DS], af = pr[REC_AF];
      const float qh = pr[REC_QH2O];
      float uu[NGT];
#pragma unroll
      for (int g = 0; g < NGT; g++) uu[g] = (g < A.ng) ? pr[uo + g] : 0.f;
      int itp;
      float dtp;
      planck_index(t, itp, dtp);
      const bool scpt = A.scattering > 0 && af > 0.f;
      const int ai = scpt || (A.sca_n > 0 && af != 0.f) ? (int)pr[REC_AI] : 0;
#pragma unroll
      for (int c = 0; c < CPW; c++) {
        const int id = c0 + c;
        if (id < A.nd) {
          const ChanConst& cc = A.cc[id];
          float tg = 1.f, et = 0.f;
#pragma unroll
          for (int g = 0; g < NGT; g++) {
            if (g < A.ng) {
              float s, f;
              ega_step(A.T, id * A.ng + g, p, t, uu[g], ep[c][g], tp[c][g], s, f);
              et = et + s - et * s;
              tg *= f;
            }
          }
          point_radiance(A, cc, id, pr, uo, p, t, ds, qh, af, ai, scpt, kj, itp, dtp, tg, et, 1.f / t, 1.f / ds,
                         radv[c], comp[c],
                         tau[c]);
        }
      }
      if (scpt) kj++;
    }
  }
  if (!valid) return;
  finish_ray<CPW>(A, ray, c0, radv, comp, tau);
}


// ---------------------------------------------------------------------------
// Fast path: shared (p, T) grids and strictly increasing columns.  Brackets
// are searched from the previous point's bracket (the LOS state moves
// slowly), which returns the same index as the reference's bisection
// (misc.c:237-265, forwardmodel.c:546-565) for increasing data; a miss after
// a few steps falls back to the bisection.
// ---------------------------------------------------------------------------
__device__ inline int near_grid(const float* xx, int n, float x, int h) {
  h = h < 0 ? 0 : (h > n - 2 ? n - 2 : h);
  const float a = xx[h], b = xx[h + 1];
  const bool okl = (h == 0) || (a <= x), okh = (h == n - 2) || (x < b);
  // a NaN passes only when n == 2, where its bracket (n - 2) is h anyway
  if (okl && okh) return h;
  if (!(x == x)) return n - 2;
  if (okl) {
    const int g = h + 1;
    if (g == n - 2 || x < xx[g + 1]) return g;
  } else {
    const int g = h - 1;
    if (g == 0 || xx[g] <= x) return g;
  }
  return locate_mono<float>(xx, n, x);
}

template <bool KEY_EPS>
__device__ inline float keyof(float2 v) { return KEY_EPS ? v.y : v.x; }

// A table column read through a buffer resource: one wave-uniform
// descriptor for the whole (u, eps) array and 32-bit per-lane offsets
// (buffer_load_dwordx2 offen), so no 64-bit address arithmetic per gather;
// the hardware range check returns zeros beyond the array.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
struct ColB {
  __amdgpu_buffer_rsrc_t r;
  int base;  // element offset of the column in the array
  __device__ float2 operator[](int i) const {
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (base + i) * 8, 0, 0);
    return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
  }
};

__device__ inline int2 ldmeta(__amdgpu_buffer_rsrc_t r, int i) {
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, i * 8, 0, 0);
  return make_int2((int)v.x, (int)v.y);
}

template <bool KEY_EPS, class C>
__device__ inline int near_col(const C& c, int n, float x, int h, float2& lo, float2& hi) {
  h = h < 0 ? 0 : (h > n - 2 ? n - 2 : h);
  lo = c[h];
  hi = c[h + 1];
  const bool okl = (h == 0) || (keyof<KEY_EPS>(lo) <= x);
  const bool okh = (h == n - 2) || (x < keyof<KEY_EPS>(hi));
  // a NaN passes only when n == 2, where its bracket (n - 2) is h anyway
  if (okl && okh) return h;
  if (!(x == x)) {
    lo = c[n - 2];
    hi = c[n - 1];
    return n - 2;
  }
  if (okl) {
    for (int s = 0; s < 3; s++) {
      h++;
      lo = hi;
      hi = c[h + 1];
      if (h == n - 2 || x < keyof<KEY_EPS>(hi)) return h;
    }
  } else {
    h--;
    hi = lo;
    lo = c[h];
    if (h == 0 || keyof<KEY_EPS>(lo) <= x) return h;
  }
  h = locate_col<KEY_EPS>(c, n, x);
  lo = c[h];
  hi = c[h + 1];
  return h;
}

// Quotients of the fast kernel as a multiply by the hardware reciprocal
// (v_rcp_f32, 1 ulp): a full IEEE division costs about ten instructions and
// the kernel issues nine per (point, channel, emitter).  The resulting
// radiance differences stay orders of magnitude below the 1e-4 budget.
__device__ inline float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ inline float lin_r(float x0, float y0, float x1, float y1, float x) {
  return y0 + fdiv(y1 - y0, x1 - x0) * (x - x0);
}

// Corner lookup (inverse then forward) starting from bracket hint h; returns
// the inverse bracket in h.  The forward search starts in the inverse
// bracket with its two nodes already loaded: u + du usually stays there
// (du >= 0), and then no further load is on the dependent chain.
__device__ inline EC corner_h(const ColB& col, int n, float ep, float du, int& h) {
  float2 a, b, c, d;
  const int i = near_col<true>(col, n, ep, h, a, b);
  h = i;
  const float u = lin_r(a.y, a.x, b.y, b.x, ep);
  const float ux = u + du;
  // same bracket as near_col<false> would return from hint i: x in [lo, hi)
  // (or beyond an end bracket), NaN excluded
  if (ux == ux && (i == 0 || a.x <= ux) && (i == n - 2 || ux < b.x)) {
    c = a;
    d = b;
  } else {
    near_col<false>(col, n, ux, i, c, d);
  }
  const float s = fdiv(d.y - c.y, d.x - c.x), dx = ux - c.x;
  return clamp_ec(c.y + s * dx, (1.f - c.y) - s * dx);
}

__device__ inline EC mix_ec(EC y0, EC y1, float w) {
  return clamp_ec(y0.e + (y1.e - y0.e) * w, y0.c + (y1.c - y0.c) * w);
}

template <int NGT, int CPW, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC))) k_path_fast(PathArgs A) {
  const int lane = threadIdx.x & 63;
  int chunk, group;
  if (!map_work(A, CPW, chunk, group)) return;
  const int slot = group * 64 + lane;
  const bool valid = slot < A.nray;
  const int ray = valid ? A.order[slot] : 0;
  const int npts = valid ? A.np[ray] : 0;
  const int c0 = chunk * CPW;
  const int uo = rec_u(A.nw);
  const DevTables& T = A.T;
  const __amdgpu_buffer_rsrc_t ur = __builtin_amdgcn_make_buffer_rsrc((void*)T.ue, (short)0, (int)T.ue_bytes,
                                                                      0x00020000);
  const __amdgpu_buffer_rsrc_t mr = __builtin_amdgcn_make_buffer_rsrc((void*)T.meta, (short)0,
                                                                      (int)T.meta_bytes, 0x00020000);

  float ep[CPW][NGT], tp[CPW][NGT], radv[CPW], comp[CPW], tau[CPW];
  // column-search hints of the previous point live in LDS (one slot per lane
  // and (channel, emitter); consecutive lanes, consecutive banks), which
  // keeps them out of the register budget
  __shared__ int hint_lds[4][CPW * NGT][64];
  int (*hint)[64] = hint_lds[threadIdx.x >> 6];
  int hpt[NGT];  // packed (ipr, it0, it1) brackets of the previous point
#pragma unroll
  for (int c = 0; c < CPW; c++) {
#pragma unroll
    for (int g = 0; g < NGT; g++) {
      ep[c][g] = 0.f;
      tp[c][g] = 1.f;
      hint[c * NGT + g][lane] = 0;
    }
    radv[c] = 0.f;
    comp[c] = 0.f;
    tau[c] = 1.f;
  }
#pragma unroll
  for (int g = 0; g < NGT; g++) hpt[g] = 0;
  int kj = (valid && A.scattering > 0) ? A.jbase[ray] : 0;
  const float* rp = A.rec + (size_t)ray * A.cap * A.R;

  for (int ip = 0;; ip++) {
    const bool act = ip < npts;
    if (!__any(act)) break;
    if (act) {
      const float* pr = rp + (size_t)ip * A.R;
      const float p = pr[REC_P], t = pr[REC_T], ds = pr[REC_DS], af = pr[REC_AF];
      const float qh = pr[REC_QH2O];
      int itp;
      float dtp;
      planck_index(t, itp, dtp);
      const bool scpt = A.scattering > 0 && af > 0.f;
      const int ai = scpt || (A.sca_n > 0 && af != 0.f) ? (int)pr[REC_AI] : 0;
      const float rtp = __builtin_amdgcn_rcpf(t), rdsp = __builtin_amdgcn_rcpf(ds);
      // Emitter-outer loop: the (p, T) brackets and weights of an emitter are
      // shared by all channels of the wave; only the per-channel segment
      // transmittance products (tg, et) stay live across emitters.
      float tg[CPW], et[CPW];
#pragma unroll
      for (int c = 0; c < CPW; c++) {
        tg[c] = 1.f;
        et[c] = 0.f;
      }
#pragma unroll
      for (int g = 0; g < NGT; g++) {
        if (g >= A.ng) continue;
        const float du = pr[uo + g];
        const int np_ = T.gnp[g];
        if (np_ < 2) continue;
        const float* pg = T.gp + g * TBLNPMAX;
        const int ipr = near_grid(pg, np_, p, hpt[g] & 0xff);
        const int nt0 = T.gnt[g * TBLNPMAX + ipr], nt1 = T.gnt[g * TBLNPMAX + ipr + 1];
        if (nt0 < 2 || nt1 < 2) continue;
        const float* t0g = T.gt + (size_t)(g * TBLNPMAX + ipr) * TBLNTMAX;
        const float* t1g = t0g + TBLNTMAX;
        const int it0 = near_grid(t0g, nt0, t, (hpt[g] >> 8) & 0xff);
        const int it1 = near_grid(t1g, nt1, t, (hpt[g] >> 16) & 0xff);
        hpt[g] = ipr | (it0 << 8) | (it1 << 16);
        const float wt0 = fdiv(t - t0g[it0], t0g[it0 + 1] - t0g[it0]);
        const float wt1 = fdiv(t - t1g[it1], t1g[it1 + 1] - t1g[it1]);
        const float wp = fdiv(p - pg[ipr], pg[ipr + 1] - pg[ipr]);
        const int gb0 = ipr * TBLNTMAX + it0, gb1 = (ipr + 1) * TBLNTMAX + it1;
#pragma unroll
        for (int c = 0; c < CPW; c++) {
          const int id = c0 + c;
          if (id >= A.nd) continue;
          float s = 0.f, f = 1.f;
          if (tp[c][g] < 1e-9f) {
            s = 1.f;
            f = 0.f;
            ep[c][g] = 1.f;
            tp[c][g] = 0.f;
          } else {
            const int tb = (id * A.ng + g) * TBLNPMAX * TBLNTMAX;
            const int2 m00 = ldmeta(mr, tb + gb0), m01 = ldmeta(mr, tb + gb0 + 1);
            const int2 m10 = ldmeta(mr, tb + gb1), m11 = ldmeta(mr, tb + gb1 + 1);
            if (m00.x >= 2 && m01.x >= 2 && m10.x >= 2 && m11.x >= 2) {
              const float e = ep[c][g];
              const int h0 = hint[c * NGT + g][lane];
              int h = h0;
              EC e00 = corner_h(ColB{ur, m00.y}, m00.x, e, du, h);
              hint[c * NGT + g][lane] = h;
              h = h0;
              EC e01 = corner_h(ColB{ur, m01.y}, m01.x, e, du, h);
              h = h0;
              EC e10 = corner_h(ColB{ur, m10.y}, m10.x, e, du, h);
              h = h0;
              EC e11 = corner_h(ColB{ur, m11.y}, m11.x, e, du, h);
              e00 = mix_ec(e00, e01, wt0);
              e11 = mix_ec(e10, e11, wt1);
              e00 = mix_ec(e00, e11, wp);
              // transparent: segment emissivity first; opaque: transmittance
              const float rt = __builtin_amdgcn_rcpf(tp[c][g]);
              const bool tr = e00.e < 0.5f;
              const float s1 = (e00.e - e) * rt, f1 = e00.c * rt;
              s = tr ? s1 : 1.f - f1;
              f = tr ? 1.f - s1 : f1;
              ep[c][g] = e00.e;
              tp[c][g] = e00.c;
            }
          }
          et[c] = et[c] + s - et[c] * s;
          tg[c] *= f;
        }
      }
#pragma unroll
      for (int c = 0; c < CPW; c++) {
        const int id = c0 + c;
        if (id < A.nd)
          point_radiance(A, A.cc[id], id, pr, uo, p, t, ds, qh, af, ai, scpt, kj, itp, dtp, tg[c], et[c], rtp, rdsp,
                         radv[c],
                         comp[c], tau[c]);
      }
      if (scpt) kj++;
    }
  }
  if (!valid) return;
  finish_ray<CPW>(A, ray, c0, radv, comp, tau);
}

// ---------------------------------------------------------------------------
// Filter-weighted Planck source-function table, FP64 (forwardmodel.c:
// 756-786, planck :569-573 with gsl_expm1's series below ln 2).
// ---------------------------------------------------------------------------
__device__ inline double expm1_gsl(double x) {
  if (fabs(x) < M_LN2) {
    double i = 1.0, sum = x, term = x;
    do {
      i++;
      term *= x / i;
      sum += term;
    } while (fabs(term) > fabs(sum) * 2.2204460492503131e-16);
    return sum;
  }
  return exp(x) - 1;
}

__global__ void k_planck_table(const double* fnu, const double* fw, const int* foff, const int* fn,
                               int nd, float* table, double* table64) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nd * JUR_NPLANCK) return;
  int id = idx / JUR_NPLANCK, it = idx % JUR_NPLANCK;
  double temp = lin(0.0, 100.0, JUR_NPLANCK - 1.0, 400.0, (double)it);
  const double* nu = fnu + foff[id];
  const double* w = fw + foff[id];
  double fsum = 0, acc = 0;
  for (int i = 0; i < fn[id]; i++) {
    fsum += w[i];
    double v = nu[i];
    acc += w[i] * (C1 * (v * v * v) / expm1_gsl(C2 * v / temp));
  }
  table[(size_t)id * JUR_NPLANCK + it] = (float)(acc / fsum);
  table64[(size_t)id * JUR_NPLANCK + it] = acc / fsum;
}

__global__ void k_iota(int* ids, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ids[i] = i;
}
__global__ void k_zero_int(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Staged path kernel: the fast kernel's arithmetic (same helpers, same
// operation order: bit-identical radiances) with its memory accesses
// reorganised for latency.  The shared (p, T) grids of all emitters are
// staged in LDS once per block; a column's two bracket nodes are one 16-byte
// load; every corner of an emitter keeps its own bracket hint (16 bits in
// LDS, the forward bracket of the previous point, where the next inverse
// lookup lands 94 % of the time and otherwise one node off); the metadata
// and node loads of all corners and channels of an emitter are issued
// together, and the four corners search in lockstep: one load round per
// step for all of them, where the fast kernel walks each corner's search
// chain in turn (a wave pays for the slowest lane of each).  Searches from
// a hint return the same bracket as the reference's bisection on strictly
// increasing columns.  The next point's record is loaded one iteration
// ahead.
// ---------------------------------------------------------------------------
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// nodes h and h + 1 of the column at element b: (u0, eps0, u1, eps1)
__device__ inline float4 ld_pair(__amdgpu_buffer_rsrc_t r, int b, int h) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (b + h) * 8, 0, 0);
  return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
}
// two neighbouring (ncol, col) entries
__device__ inline int4 ld_meta2(__amdgpu_buffer_rsrc_t r, int i) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, i * 8, 0, 0);
  return make_int4((int)v.x, (int)v.y, (int)v.z, (int)v.w);
}

// A contiguous column read through the table's buffer resource: pairs are
// one 16-byte load.
struct ColP : ColB {
  __device__ float4 pair(int h) const { return ld_pair(r, base, h); }
};
// Step of a column search: -1 / +1 when x lies below / above the bracket
// (h, q), 0 inside it (or NaN).
// X: float (FP32 path) or double (FP64 path: FP32 nodes compared in double,
// as the reference's locate on the table's float columns does).
template <bool KEY_EPS, class X>
__device__ inline int step_dir(float4 q, int h, int n, X x) {
  const X lo = KEY_EPS ? q.y : q.x, hi = KEY_EPS ? q.w : q.z;
  return (h > 0 && x < lo) ? -1 : ((h < n - 2 && hi <= x) ? 1 : 0);
}

// Bisection of a column for key x (locate, misc.c:237-265, for increasing
// data; NaN gives the last bracket).
template <bool KEY_EPS, class C, class X>
__device__ inline int locate_colx(const C& c, int n, X x) {
  int ilo = 0, ihi = n - 1;
  while (ihi > ilo + 1) {
    const int i = (ihi + ilo) >> 1;
    const X v = KEY_EPS ? c[i].y : c[i].x;
    if (v > x) ihi = i; else ilo = i;
  }
  return ilo;
}

// Brackets of x[k] in the four corner columns (element base b[k], length
// n[k]) from the hints h[k] with their pairs q[k] loaded: all corners step
// together, one load round per step for the four of them (a wave takes a
// step when any of its lanes needs it), three steps at most; what is still
// off then, and NaN, is found by bisection.  Returns the reference's
// bracket (the unique one in a strictly increasing column).  Lanes that are
// not live step too (within [0, n - 2], their results unused) but neither
// ask for a round nor go to the bisection: no per-lane branch around the
// step tests.
template <bool KEY_EPS, class Col, class X>
__device__ inline void lockstep(const Col (&col)[4], const int (&n)[4], const X (&x)[4], bool live, int (&h)[4],
                                float4 (&q)[4]) {
  int d[4];
  bool any = false;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    d[k] = step_dir<KEY_EPS>(q[k], h[k], n[k], x[k]);
    any = any || (live && d[k] != 0);
  }
#pragma unroll
  for (int s = 0; s < 3; s++) {
    if (!__any(any)) break;
    // every corner reloads (a corner that stays reloads its own pair): no
    // per-corner branches, the four loads of a round issue together
#pragma unroll
    for (int k = 0; k < 4; k++) {
      h[k] += d[k];
      q[k] = col[k].pair(h[k]);
    }
    any = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      d[k] = step_dir<KEY_EPS>(q[k], h[k], n[k], x[k]);
      any = any || (live && d[k] != 0);
    }
  }
  bool slow = any;
#pragma unroll
  for (int k = 0; k < 4; k++) slow = slow || (live && !(x[k] == x[k]));
  if (__any(slow)) {
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (live && (d[k] != 0 || !(x[k] == x[k]))) {
        h[k] = locate_colx<KEY_EPS>(col[k], n[k], x[k]);
        q[k] = col[k].pair(h[k]);
      }
  }
}

// Bracket in a strictly increasing grid held in LDS, from hint h.
__device__ inline int grid_bracket(const float* g, int n, float x, int h) {
  h = h < 0 ? 0 : (h > n - 2 ? n - 2 : h);
  const float a = g[h], b = g[h + 1];
  if ((h == 0 || a <= x) && (h == n - 2 || x < b)) return h;
  return locate_mono<float>(g, n, x);
}

// PACK: the point records come from the ray group's packed stream written
// by k_pack (A.pack: lane-contiguous, so every per-point load of the wave is
// one contiguous 256-byte or 1-KiB access instead of 64 scattered ones), and
// the (p, T) brackets and weights of every (point, emitter) come precomputed
// in that stream instead of from the grids in LDS.
template <int NGT, int CPW, int OCC, bool TILE = false, bool PACK = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC))) k_path_stage(PathArgs A) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const DevTables& T = A.T;
  const int NPS = PACK ? 0 : T.gnp_max, NTS = PACK ? 0 : T.gnt_max;
  float* sgp = (float*)lds_raw;                 // [NGT][NPS] pressure grids
  int* sgnt = (int*)(sgp + NGT * NPS);          // [NGT][NPS] temperatures per pressure
  float* sgt = (float*)(sgnt + NGT * NPS);      // [NGT][NPS][NTS] temperature grids
  const int goff = (2 * NGT * NPS + NGT * NPS * NTS + 1) & ~1;
  if (!PACK) {
    for (int i = threadIdx.x; i < NGT * NPS; i += blockDim.x) {
      const int g = i / NPS, ip = i - g * NPS;
      const bool in = g < A.ng;
      sgp[i] = in ? T.gp[g * TBLNPMAX + ip] : 0.f;
      sgnt[i] = in ? T.gnt[g * TBLNPMAX + ip] : 0;
    }
    for (int i = threadIdx.x; i < NGT * NPS * NTS; i += blockDim.x) {
      const int g = i / (NPS * NTS), r = i - g * NPS * NTS, ip = r / NTS, it = r - ip * NTS;
      sgt[i] = g < A.ng ? T.gt[(size_t)(g * TBLNPMAX + ip) * TBLNTMAX + it] : 0.f;
    }
    __syncthreads();
  }

  const int lane = threadIdx.x & 63;
  int chunk, group;
  if (!(TILE ? map_work_tiled(A, CPW, chunk, group) : map_work(A, CPW, chunk, group))) return;
  // per-corner bracket hints: [wave][channel x emitter][lane], 4 x 16 bits
  uint2* hints = (uint2*)((float*)lds_raw + goff) + (threadIdx.x >> 6) * (CPW * NGT * 64) + lane;
  const int slot = group * 64 + lane;
  const bool valid = slot < A.nray;
  const int ray = valid ? A.order[slot] : 0;
  const int npts = valid ? A.np[ray] : 0;
  const int c0 = chunk * CPW;
  const int uo = rec_u(A.nw);
  const __amdgpu_buffer_rsrc_t ur = __builtin_amdgcn_make_buffer_rsrc((void*)T.ue, (short)0, (int)T.ue_bytes,
                                                                      0x00020000);
  const __amdgpu_buffer_rsrc_t mr = __builtin_amdgcn_make_buffer_rsrc((void*)T.meta, (short)0,
                                                                      (int)T.meta_bytes, 0x00020000);
  int gnp[NGT];
#pragma unroll
  for (int g = 0; g < NGT; g++) gnp[g] = g < A.ng ? T.gnp[g] : 0;

  float ep[CPW][NGT], tp[CPW][NGT], radv[CPW], comp[CPW], tau[CPW];
#pragma unroll
  for (int c = 0; c < CPW; c++) {
#pragma unroll
    for (int g = 0; g < NGT; g++) {
      ep[c][g] = 0.f;
      tp[c][g] = 1.f;
      hints[(c * NGT + g) * 64] = make_uint2(0u, 0u);
    }
    radv[c] = 0.f;
    comp[c] = 0.f;
    tau[c] = 1.f;
  }
  int hpt[NGT];  // packed (ipr, it0, it1) brackets of the previous point
#pragma unroll
  for (int g = 0; g < NGT; g++) hpt[g] = 0;
  int kj = (valid && A.scattering > 0) ? A.jbase[ray] : 0;
  // point records: the ray's own [cap][R] (stride 1 between fields), or
  // the group's packed stream [cap][R + 4 ng][64] (stride 64)
  const int RS = PACK ? 64 : 1;
  const size_t PF = PACK ? (size_t)(A.R + 4 * A.ng) * 64 : (size_t)A.R;  // floats per point slot
  const float* rp = PACK ? A.pack + (size_t)group * A.cap * PF + lane : A.rec + (size_t)ray * A.cap * A.R;

  // without the packed stream: the next point's record, loaded one
  // iteration ahead (the grid searches need p and T first)
  float np_p = 0.f, np_t = 0.f, np_ds = 0.f, np_af = 0.f, np_qh = 0.f, np_ai = 0.f, np_du[NGT];
  if constexpr (!PACK) {
    const float* pr = rp;
    np_p = pr[REC_P * RS];
    np_t = pr[REC_T * RS];
    np_ds = pr[REC_DS * RS];
    np_af = pr[REC_AF * RS];
    np_qh = pr[REC_QH2O * RS];
    np_ai = pr[REC_AI * RS];
#pragma unroll
    for (int g = 0; g < NGT; g++) np_du[g] = g < A.ng ? pr[(uo + g) * RS] : 0.f;
  }

  for (int ip = 0;; ip++) {
    const bool act = ip < npts;
    if (!__any(act)) break;
    float p, t, ds, af, qh, aif;
    float dus[NGT];
    int4 bks[NGT];
    if constexpr (PACK) {
      // this point's brackets and record (ip < cap: some lane of the wave
      // still has points), loaded together: the first use of either, the
      // metadata loads, needs the brackets, and the record's first use
      // comes after the node loads, so no prefetch
      const float* pc = rp + (size_t)ip * PF;
      const int4* bq = (const int4*)(pc - lane + (size_t)A.R * 64) + lane;
#pragma unroll
      for (int g = 0; g < NGT; g++) bks[g] = g < A.ng ? bq[g * 64] : make_int4(0, 0, 0, 0);
      p = pc[REC_P * RS];
      t = pc[REC_T * RS];
      ds = pc[REC_DS * RS];
      af = pc[REC_AF * RS];
      qh = pc[REC_QH2O * RS];
      aif = pc[REC_AI * RS];
#pragma unroll
      for (int g = 0; g < NGT; g++) dus[g] = g < A.ng ? pc[(uo + g) * RS] : 0.f;
    } else {
      p = np_p;
      t = np_t;
      ds = np_ds;
      af = np_af;
      qh = np_qh;
      aif = np_ai;
#pragma unroll
      for (int g = 0; g < NGT; g++) {
        dus[g] = np_du[g];
        bks[g] = make_int4(0, 0, 0, 0);
      }
      // prefetch (the last point of the ray reloads itself)
      const int in = ip + 1 < npts ? ip + 1 : (npts > 0 ? npts - 1 : 0);
      const float* pn = rp + (size_t)in * PF;
      np_p = pn[REC_P * RS];
      np_t = pn[REC_T * RS];
      np_ds = pn[REC_DS * RS];
      np_af = pn[REC_AF * RS];
      np_qh = pn[REC_QH2O * RS];
      np_ai = pn[REC_AI * RS];
#pragma unroll
      for (int g = 0; g < NGT; g++) np_du[g] = g < A.ng ? pn[(uo + g) * RS] : 0.f;
    }
    if (act) {
      const float* pr = rp + (size_t)ip * PF;
      int itp;
      float dtp;
      planck_index(t, itp, dtp);
      const bool scpt = A.scattering > 0 && af > 0.f;
      const int ai = scpt || (A.sca_n > 0 && af != 0.f) ? (int)aif : 0;
      const float rtp = __builtin_amdgcn_rcpf(t), rdsp = __builtin_amdgcn_rcpf(ds);
      float tg[CPW], et[CPW];
#pragma unroll
      for (int c = 0; c < CPW; c++) {
        tg[c] = 1.f;
        et[c] = 0.f;
      }
#pragma unroll
      for (int g = 0; g < NGT; g++) {
        if (g >= A.ng) continue;
        const int4 bc = bks[g];
        const int np_ = gnp[g];
        if (np_ < 2) continue;
        const float du = dus[g];
        int gb0, gb1;  // metadata rows of the (p, T) corners (ipr, it0) and (ipr + 1, it1)
        bool gok;
        float wt0, wt1, wp;
        if (PACK) {
          gb0 = bc.x & 0x7ff;
          gb1 = (bc.x >> 11) & 0x7ff;
          gok = (bc.x >> 22) != 0;
          wt0 = __int_as_float(bc.y);
          wt1 = __int_as_float(bc.z);
          wp = __int_as_float(bc.w);
        } else {
          // (p, T) brackets and weights from the staged grids
          const float* pg = sgp + g * NPS;
          const int ipr = grid_bracket(pg, np_, p, hpt[g] & 0xff);
          const int nt0 = sgnt[g * NPS + ipr], nt1 = sgnt[g * NPS + ipr + 1];
          gok = nt0 >= 2 && nt1 >= 2;
          const float* t0g = sgt + (g * NPS + ipr) * NTS;
          const float* t1g = t0g + NTS;
          const int it0 = grid_bracket(t0g, gok ? nt0 : 2, t, (hpt[g] >> 8) & 0xff);
          const int it1 = grid_bracket(t1g, gok ? nt1 : 2, t, (hpt[g] >> 16) & 0xff);
          if (gok) hpt[g] = ipr | (it0 << 8) | (it1 << 16);
          wt0 = fdiv(t - t0g[it0], t0g[it0 + 1] - t0g[it0]);
          wt1 = fdiv(t - t1g[it1], t1g[it1 + 1] - t1g[it1]);
          wp = fdiv(p - pg[ipr], pg[ipr + 1] - pg[ipr]);
          gb0 = ipr * TBLNTMAX + it0;
          gb1 = (ipr + 1) * TBLNTMAX + it1;
        }
        // all channels: metadata, hints and the hinted bracket nodes
        int nn[CPW][4], bb[CPW][4], hh[CPW][4];
        float4 qa[CPW][4];
#pragma unroll
        for (int c = 0; c < CPW; c++) {
          const int id = c0 + c < A.nd ? c0 + c : A.nd - 1;
          const int tb = (id * A.ng + g) * TBLNPMAX * TBLNTMAX;
          const int4 m0 = ld_meta2(mr, tb + gb0), m1 = ld_meta2(mr, tb + gb1);
          const uint2 hv = hints[(c * NGT + g) * 64];
          nn[c][0] = m0.x; bb[c][0] = m0.y; nn[c][1] = m0.z; bb[c][1] = m0.w;
          nn[c][2] = m1.x; bb[c][2] = m1.y; nn[c][3] = m1.z; bb[c][3] = m1.w;
          hh[c][0] = (int)(hv.x & 0xffffu); hh[c][1] = (int)(hv.x >> 16);
          hh[c][2] = (int)(hv.y & 0xffffu); hh[c][3] = (int)(hv.y >> 16);
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const int hm = nn[c][k] - 2;
            hh[c][k] = hh[c][k] > hm ? hm : hh[c][k];
            hh[c][k] = hh[c][k] < 0 ? 0 : hh[c][k];
            qa[c][k] = ld_pair(ur, bb[c][k], hh[c][k]);
          }
        }
#pragma unroll
        for (int c = 0; c < CPW; c++) {
          const int id = c0 + c;
          if (id >= A.nd) continue;
          const bool sat = tp[c][g] < 1e-9f;
          const bool live = gok && !sat && nn[c][0] >= 2 && nn[c][1] >= 2 && nn[c][2] >= 2 && nn[c][3] >= 2;
          const float e = ep[c][g];
          const float ex[4] = {e, e, e, e};
          const ColP cols[4] = {{{ur, bb[c][0]}}, {{ur, bb[c][1]}}, {{ur, bb[c][2]}}, {{ur, bb[c][3]}}};
          lockstep<true>(cols, nn[c], ex, live, hh[c], qa[c]);
          float ux[4];
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const float4 q = qa[c][k];
            ux[k] = lin_r(q.y, q.x, q.w, q.z, e) + du;
          }
          // the forward search starts in the inverse bracket, in place
          lockstep<false>(cols, nn[c], ux, live, hh[c], qa[c]);
          // stored every time (cheaper than testing whether a bracket moved)
          hints[(c * NGT + g) * 64] = make_uint2((unsigned)hh[c][0] | ((unsigned)hh[c][1] << 16),
                                                 (unsigned)hh[c][2] | ((unsigned)hh[c][3] << 16));
          EC ec[4];
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const float4 q = qa[c][k];
            const float s = fdiv(q.w - q.y, q.z - q.x), dx = ux[k] - q.x;
            ec[k] = clamp_ec(q.y + s * dx, (1.f - q.y) - s * dx);
          }
          EC e00 = mix_ec(ec[0], ec[1], wt0);
          const EC e11 = mix_ec(ec[2], ec[3], wt1);
          e00 = mix_ec(e00, e11, wp);
          const float rt = __builtin_amdgcn_rcpf(tp[c][g]);
          const bool tr = e00.e < 0.5f;
          const float s1 = (e00.e - e) * rt, f1 = e00.c * rt;
          float s = tr ? s1 : 1.f - f1;
          float f = tr ? 1.f - s1 : f1;
          if (sat) {
            s = 1.f;
            f = 0.f;
            ep[c][g] = 1.f;
            tp[c][g] = 0.f;
          } else if (live) {
            ep[c][g] = e00.e;
            tp[c][g] = e00.c;
          } else {
            s = 0.f;
            f = 1.f;
          }
          et[c] = et[c] + s - et[c] * s;
          tg[c] *= f;
        }
      }
#pragma unroll
      for (int c = 0; c < CPW; c++) {
        const int id = c0 + c;
        if (id < A.nd)
          point_radiance(A, A.cc[id], id, pr, uo, p, t, ds, qh, af, ai, scpt, kj, itp, dtp, tg[c], et[c], rtp, rdsp,
                         radv[c], comp[c], tau[c], RS);
      }
      if (scpt) kj++;
    }
  }
  if (!valid) return;
  finish_ray<CPW>(A, ray, c0, radv, comp, tau);
}

// LDS bytes of k_path_stage: staged grids + hints; 0 if the grids are too
// large to stage (the fast kernel runs instead).  With the packed stream
// (precomputed brackets) the hints only.
size_t stage_lds_bytes(const PathArgs& A, int ngt, int cpw, int wpb, bool pack = false) {
  if (A.T.gnp_max < 2 || A.T.gnt_max < 2) return 0;
  const size_t hints = (size_t)wpb * cpw * ngt * 64 * 8;
  if (pack) return hints;
  const size_t gw = (size_t)2 * ngt * A.T.gnp_max + (size_t)ngt * A.T.gnp_max * A.T.gnt_max;
  if (gw * 4 > 24576) return 0;
  return ((gw + 1) & ~(size_t)1) * 4 + hints;
}

// Packed stream of a ray group for k_path_stage<..., PACK>: per group of 64
// consecutive rays of the sorted list and LOS point slot, the rays' point
// records lane-contiguous ([R][64] floats) followed by the (p, T) brackets
// and interpolation weights of every emitter ([ng][64] int4: gb0 | gb1 << 11
// | gok << 22, wt0, wt1, wp; gb0 = ipr TBLNTMAX + it0 and gb1 = (ipr + 1)
// TBLNTMAX + it1 are the metadata rows of the (p, T) corners, gok = both
// pressure rows hold >= 2 temperatures) on the shared table grids (forwardmodel.c:442-446 and the
// weights of the LIN calls at :477-491), computed once per (point, emitter)
// for all channels.  The searches are the staged kernel's own (grid_bracket:
// on a strictly increasing grid the unique bracket, whatever the hint) and
// the quotients the same, so radiances are bit-identical.  Block (group,
// y): its 4 waves write point slots 32 y .. 32 y + 31 of the group, a wave
// one slot's 64 lanes per step; blocks past the group's longest ray leave
// at once.
static_assert(TBLNPMAX * TBLNTMAX <= 2048, "metadata rows are packed in 11 bits");
constexpr int PACK_SLOTS = 32;  // point slots per block of k_pack
__device__ inline void pack_point(const PathArgs& A, int group, int lane, int ray, int ip);
__global__ void __launch_bounds__(256) k_pack(PathArgs A) {
  const int group = blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int slot = group * 64 + lane;
  const int ray = slot < A.nray ? A.order[slot] : -1;
  const int np = ray >= 0 ? A.np[ray] : 0;
  int mx = np;
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
  const int ip1 = min(mx, (int)(blockIdx.y + 1) * PACK_SLOTS);
  for (int ip = blockIdx.y * PACK_SLOTS + wv; ip < ip1; ip += 4)
    if (ip < np) pack_point(A, group, lane, ray, ip);
}
__device__ inline void pack_point(const PathArgs& A, int group, int lane, int ray, int ip) {
  const DevTables& T = A.T;
  const float* pr = A.rec + ((size_t)ray * A.cap + ip) * A.R;
  float* out = A.pack + ((size_t)group * A.cap + ip) * (size_t)(A.R + 4 * A.ng) * 64 + lane;
  if ((A.R & 3) == 0 && ((uintptr_t)A.rec & 15) == 0) {
    // every record 16-byte aligned: a quarter of the loads, each touching
    // the same 64 lines a scalar one does
    for (int f = 0; f < A.R; f += 4) {
      const float4 v = *(const float4*)(pr + f);
      out[f * 64] = v.x;
      out[(f + 1) * 64] = v.y;
      out[(f + 2) * 64] = v.z;
      out[(f + 3) * 64] = v.w;
    }
  } else {
    for (int f = 0; f < A.R; f++) out[f * 64] = pr[f];
  }
  int4* ob = (int4*)(out - lane + (size_t)A.R * 64) + lane;
  const float p = pr[REC_P], tt = pr[REC_T];
  for (int g = 0; g < A.ng; g++) {
    const int np_ = T.gnp[g];
    if (np_ < 2) {
      ob[g * 64] = make_int4(0, 0, 0, 0);
      continue;
    }
    const float* pg = T.gp + g * TBLNPMAX;
    const int ipr = grid_bracket(pg, np_, p, 0);
    const int nt0 = T.gnt[g * TBLNPMAX + ipr], nt1 = T.gnt[g * TBLNPMAX + ipr + 1];
    const bool gok = nt0 >= 2 && nt1 >= 2;
    const float* t0g = T.gt + (size_t)(g * TBLNPMAX + ipr) * TBLNTMAX;
    const float* t1g = t0g + TBLNTMAX;
    const int it0 = grid_bracket(t0g, gok ? nt0 : 2, tt, 0);
    const int it1 = grid_bracket(t1g, gok ? nt1 : 2, tt, 0);
    const float wt0 = fdiv(tt - t0g[it0], t0g[it0 + 1] - t0g[it0]);
    const float wt1 = fdiv(tt - t1g[it1], t1g[it1 + 1] - t1g[it1]);
    const float wp = fdiv(p - pg[ipr], pg[ipr + 1] - pg[ipr]);
    const int gb0 = ipr * TBLNTMAX + it0, gb1 = (ipr + 1) * TBLNTMAX + it1;
    ob[g * 64] = make_int4(gb0 | (gb1 << 11) | ((gok ? 1 : 0) << 22), __float_as_int(wt0),
                           __float_as_int(wt1), __float_as_int(wp));
  }
}

// ---------------------------------------------------------------------------
// FP64 path integration: the reference's arithmetic and operation order
// (forwardmodel.c:303-399 with intpol_tbl :409-502 and formod_continua
// :120-152) in double precision, on the FP32 (u, eps) columns the reference
// also stores in single precision and on FP64 copies of everything it keeps
// in double (table grids, Planck table, continuum coefficients, point
// records).  Used where radiance DIFFERENCES matter: the finite-difference
// Jacobian perturbs one level by 1 % and its radiance change can be far
// below FP32 resolution.  One channel per wave.
// ---------------------------------------------------------------------------

// LIN on float table nodes as the reference's macro evaluates it: under C's
// usual arithmetic conversions the differences and their quotient are
// single precision, the offset and the sum double (jurassic.h:73-74).
__device__ inline double lin_f(float x0, float y0, float x1, float y1, double x) {
  const float sl = (y1 - y0) / (x1 - x0);
  return (double)y0 + (double)sl * (x - (double)x0);
}
// GSL_MAX(GSL_MIN(x, 1), 0): a NaN becomes 1.
__device__ inline double clamp01d(double x) {
  const double a = x < 1.0 ? x : 1.0;
  return a > 0.0 ? a : 0.0;
}
// locate_tbl (forwardmodel.c:546-565) on an increasing column, double key.
template <bool KEY_EPS>
__device__ inline int locate_col64(const float2* c, int n, double x) {
  int ilo = 0, ihi = n - 1;
  while (ihi > ilo + 1) {
    const int i = (ihi + ilo) >> 1;
    const double v = KEY_EPS ? c[i].y : c[i].x;
    if (v > x) ihi = i; else ilo = i;
  }
  return ilo;
}

// Bracket search from a hint (the previous point's bracket) on a strictly
// increasing column: the same index as the bisection (locate_col64), found
// in a step or two while the LOS state moves slowly.
template <bool KEY_EPS>
__device__ inline int near_col64(const float2* c, int n, double x, int h) {
  h = h < 0 ? 0 : (h > n - 2 ? n - 2 : h);
  const bool okl = (h == 0) || ((double)(KEY_EPS ? c[h].y : c[h].x) <= x);
  const bool okh = (h == n - 2) || (x < (double)(KEY_EPS ? c[h + 1].y : c[h + 1].x));
  if (okl && okh) return h;
  if (!(x == x)) return n - 2;
  if (okl) {
    for (int s = 0; s < 3; s++) {
      h++;
      if (h == n - 2 || x < (double)(KEY_EPS ? c[h + 1].y : c[h + 1].x)) return h;
    }
  } else {
    h--;
    if (h == 0 || (double)(KEY_EPS ? c[h].y : c[h].x) <= x) return h;
  }
  return locate_col64<KEY_EPS>(c, n, x);
}
// The same on an increasing grid (locate, misc.c:237-265).
__device__ inline int near_grid64(const double* xx, int n, double x, int h) {
  h = h < 0 ? 0 : (h > n - 2 ? n - 2 : h);
  const bool okl = (h == 0) || (xx[h] <= x), okh = (h == n - 2) || (x < xx[h + 1]);
  if (okl && okh) return h;
  if (!(x == x)) return n - 2;
  if (okl) {
    const int g = h + 1;
    if (g == n - 2 || x < xx[g + 1]) return g;
  } else {
    const int g = h - 1;
    if (g == 0 || xx[g] <= x) return g;
  }
  return locate_mono<double>(xx, n, x);
}
// intpol_tbl_u then intpol_tbl_eps of one (p, T) corner (forwardmodel.c:460-474).
// HINT: strictly increasing columns; h holds the inverse lookup's bracket.
// Per-lane hint slots in LDS: slot k of a lane at base[64 k].
// (16-bit: brackets are below TBLNUMAX = 320)
struct LaneHints {
  short* base;
  struct Ref {
    short* p;
    __device__ operator int() const { return *p; }
    __device__ Ref& operator=(int v) {
      *p = (short)v;
      return *this;
    }
  };
  __device__ Ref operator[](int k) const { return Ref{base + k * 64}; }
};
// Hint slots in registers (one emitter per lane: 7 slots).
struct RegHints {
  int* h;
  __device__ int& operator[](int k) const { return h[k]; }
};


template <bool HINT, class H>
__device__ inline double corner64(const float2* col, int n, double e_path, double du, H&& h) {
  const int i = HINT ? near_col64<true>(col, n, e_path, h) : locate_col64<true>(col, n, e_path);
  h = i;
  const double uu = lin_f(col[i].y, col[i].x, col[i + 1].y, col[i + 1].x, e_path) + du;
  const int j = HINT ? near_col64<false>(col, n, uu, i) : locate_col64<false>(col, n, uu);
  return clamp01d(lin_f(col[j].x, col[j].y, col[j + 1].x, col[j + 1].y, uu));
}
// Segment emissivity of one emitter; updates the path transmittance
// (forwardmodel.c:433-500).  hint: brackets of the previous point (p, T at
// both pressures, the four corners' inverse lookups), used when HINT
// (increasing grids and columns).
template <bool HINT, class Hints>
__device__ inline double ega64(const DevTables& T, int tb, double p, double t, double du, double& tau_path,
                               Hints hint) {
  double eps = 0.0;
  const int ntp = T.np[tb];
  if (ntp < 2) {
    eps = 0.0;
  } else if (tau_path < 1e-9) {
    eps = 1.0;
  } else {
    const double* pg = T.p64 + (size_t)tb * TBLNPMAX;
    const int ipr = HINT ? near_grid64(pg, ntp, p, hint[0]) : locate_mono<double>(pg, ntp, p);
    hint[0] = ipr;
    const size_t b0 = (size_t)tb * TBLNPMAX + ipr;
    const int nt0 = T.nt[b0], nt1 = T.nt[b0 + 1];
    if (nt0 >= 2 && nt1 >= 2) {
      const double* t0g = T.t64 + b0 * TBLNTMAX;
      const double* t1g = t0g + TBLNTMAX;
      const int it0 = HINT ? near_grid64(t0g, nt0, t, hint[1]) : locate_mono<double>(t0g, nt0, t);
      const int it1 = HINT ? near_grid64(t1g, nt1, t, hint[2]) : locate_mono<double>(t1g, nt1, t);
      hint[1] = it0;
      hint[2] = it1;
      const size_t k00 = b0 * TBLNTMAX + it0, k10 = (b0 + 1) * TBLNTMAX + it1;
      const int n00 = T.ncol[k00], n01 = T.ncol[k00 + 1], n10 = T.ncol[k10], n11 = T.ncol[k10 + 1];
      if (n00 >= 2 && n01 >= 2 && n10 >= 2 && n11 >= 2) {
        const double ep = 1.0 - tau_path;
        double e00, e01, e10, e11;
        if (HINT) {
          // the four corners searched in lockstep from their hints (the
          // previous point's forward brackets), as in the FP32 staged kernel
          const __amdgpu_buffer_rsrc_t ur =
              __builtin_amdgcn_make_buffer_rsrc((void*)T.ue, (short)0, (int)T.ue_bytes, 0x00020000);
          const ColP col[4] = {{{ur, T.col[k00]}}, {{ur, T.col[k00 + 1]}}, {{ur, T.col[k10]}},
                               {{ur, T.col[k10 + 1]}}};
          const int n[4] = {n00, n01, n10, n11};
          int h[4];
          float4 q[4];
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const int hk = hint[3 + k];
            h[k] = hk < 0 ? 0 : (hk > n[k] - 2 ? n[k] - 2 : hk);
            q[k] = col[k].pair(h[k]);
          }
          const double ex[4] = {ep, ep, ep, ep};
          lockstep<true>(col, n, ex, true, h, q);
          double uu[4];
#pragma unroll
          for (int k = 0; k < 4; k++) uu[k] = lin_f(q[k].y, q[k].x, q[k].w, q[k].z, ep) + du;
          lockstep<false>(col, n, uu, true, h, q);
          double ek[4];
#pragma unroll
          for (int k = 0; k < 4; k++) {
            hint[3 + k] = h[k];
            ek[k] = clamp01d(lin_f(q[k].x, q[k].y, q[k].z, q[k].w, uu[k]));
          }
          e00 = ek[0];
          e01 = ek[1];
          e10 = ek[2];
          e11 = ek[3];
        } else {
          e00 = corner64<HINT>(T.ue + T.col[k00], n00, ep, du, hint[3]);
          e01 = corner64<HINT>(T.ue + T.col[k00 + 1], n01, ep, du, hint[4]);
          e10 = corner64<HINT>(T.ue + T.col[k10], n10, ep, du, hint[5]);
          e11 = corner64<HINT>(T.ue + T.col[k10 + 1], n11, ep, du, hint[6]);
        }
        e00 = clamp01d(lin<double>(t0g[it0], e00, t0g[it0 + 1], e01, t));
        e11 = clamp01d(lin<double>(t1g[it1], e10, t1g[it1 + 1], e11, t));
        e00 = clamp01d(lin<double>(pg[ipr], e00, pg[ipr + 1], e11, p));
        eps = 1.0 - (1.0 - e00) / tau_path;
      }
    }
  }
  tau_path *= 1.0 - eps;
  return eps;
}
// Continuum extinction [1/km] (forwardmodel.c:120-152) in FP64, each term
// in the operation order of ctmco2, ctmh2o, ctmn2, ctmo2 (continua.c:854-869,
// 1898-1928, 1980-1992, 2037-2048).
__device__ inline double continua64(const ChanConst64& c, double p, double t, double ds, double qh2o,
                                    double uco2, double uh2o, double kw) {
  double beta = kw;
  if (c.co2_on && ds > 0.0) {
    const double dt230 = t - 230, dt260 = t - 260, dt296 = t - 296;
    const double ctw = dt260 * 5.050505e-4 * dt296 * c.co2_cw230 - dt230 * 9.259259e-4 * dt296 * c.co2_cw260 +
                       dt230 * 4.208754e-4 * dt260 * c.co2_cw296;
    beta += uco2 / JUR_AVOGADRO / 1000 * p / P0 * ctw / ds;
  }
  if (c.h2o_on && ds > 0.0) {
    const double ctwslf = c.h2o_sfac * c.h2o_cw296 * pow(c.h2o_cw260 / c.h2o_cw296, (296 - t) / (296 - 260));
    const double a1 = c.nu * uh2o * tanh(.7193876 / t * c.nu);
    const double a2 = 296 / t;
    const double a3 = p / P0 * (qh2o * ctwslf + (1 - qh2o) * c.h2o_ctwfrn) * 1e-20;
    beta += a1 * a2 * a3 / ds;
  }
  if (c.n2_on) {
    const double x = p / P0, y = 273 / t;
    beta += 0.1 * (x * x) * (y * y) * 0.79 * c.n2_b * exp(c.n2_beta * (1 / 296.0 - 1 / t)) *
            (0.79 + (1 - 0.79) * (1.294 - 0.4545 * t / 296));
  }
  if (c.o2_on) {
    const double x = p / P0, y = 273 / t;
    beta += 0.1 * (x * x) * (y * y) * 0.21 * c.o2_b * exp(c.o2_beta * (1 / 296.0 - 1 / t));
  }
  return beta;
}
// Filter-weighted Planck function from the FP64 table: locate on the
// 1201-point grid 100-400 K, then LIN (forwardmodel.c:786-793).
// The grid nodes 100 + 0.25 i are exact in FP64, so the bisection's bracket
// is floor(4 (t - 100)) clamped to [0, 1199] (NaN: the last bracket).
__device__ inline double planck64(const double* row, double t) {
  const double x = floor((t - 100.0) * 4.0);
  const int ilo = !(x == x) ? JUR_NPLANCK - 2 : (x < 0.0 ? 0 : (x > JUR_NPLANCK - 2.0 ? JUR_NPLANCK - 2 : (int)x));
  const double t0 = lin(0.0, 100.0, JUR_NPLANCK - 1.0, 400.0, (double)ilo);
  const double t1 = lin(0.0, 100.0, JUR_NPLANCK - 1.0, 400.0, (double)(ilo + 1));
  return lin(t0, row[ilo], t1, row[ilo + 1], t);
}

// Radiance and transmittance update at one LOS point given the gas
// segment transmittance tau_gas (forwardmodel.c:317-390).  Clear-air
// optics of the point, E = exp(-beta_x ds) and the source B: PM 0 computed;
// PM 1 computed and written to eb[0], eb[1] (also where tau_gas = 0); PM 2
// read from eb when use_eb (a base ray's, same inputs), else computed.
template <int PM = 0, class Rec>
__device__ __forceinline__ void point64(const PathArgs& A, const ChanConst64& cc, const double* prow, int id,
                                        int uo, const Rec& rec, int ip, double p, double t, double ds, double af,
                                        double qh, double tau_gas, int& kj, double& rad, double& tau,
                                        double* eb = nullptr, bool use_eb = false) {
  const bool scpt = A.scattering > 0 && af > 0.0;
  const int ai = scpt || (A.sca_n > 0 && af != 0.0) ? (int)rec(ip, REC_AI) : 0;
  double beta = 0.0, B, E = 0.0;
  if (PM == 2 && use_eb) {
    E = eb[0];
    B = eb[1];
  } else {
    const double uco2 = cc.co2_on ? rec(ip, uo + cc.co2_ig) : 0.0;
    const double uh2o = cc.h2o_on ? rec(ip, uo + cc.h2o_ig) : 0.0;
    const double kw = A.nw > 0 ? rec(ip, REC_K + cc.window) : 0.0;
    beta = continua64(cc, p, t, ds, qh, uco2, uh2o, kw);
    B = planck64(prow, t);
    if (PM != 0 && !scpt) {
      double bx = beta;
      if (A.sca_n > 0 && af != 0.0)
        bx += af * (A.ext_beta_e ? A.beta_e64[ai * A.nd + id] : A.beta_a64[ai * A.nd + id]);
      E = exp(-bx * ds);
    }
    if (PM == 1) {
      eb[0] = E;
      eb[1] = B;
    }
  }
  if (scpt) {
    // scattering source (forwardmodel.c:317-360); a zero-length segment
    // adds nothing (the reference forms -log(1)/0 there)
    if (tau_gas > 0.0 && ds > 0.0) {
      const double be = A.beta_e64[ai * A.nd + id], bs = A.beta_s64[ai * A.nd + id];
      const double J = kj < A.nJ ? (double)A.J[(size_t)kj * A.nd + id] : 0.0;
      const double bt = -log(tau_gas) / ds + beta + af * be;
      const double eps = 1.0 - exp(-bt * ds);
      const double src = ((bt - af * bs) * B + af * bs * J) / bt;
      rad += tau * eps * src;
      tau *= exp(-bt * ds);
    }
    kj++;
  } else if (tau_gas > 0.0) {
    // gas, continua and aerosol extinction (forwardmodel.c:363-390)
    if (PM == 0) {
      double bx = beta;
      if (A.sca_n > 0 && af != 0.0)
        bx += af * (A.ext_beta_e ? A.beta_e64[ai * A.nd + id] : A.beta_a64[ai * A.nd + id]);
      E = exp(-bx * ds);
    }
    const double eps = 1.0 - tau_gas * E;
    rad += B * eps * tau;
    tau *= 1.0 - eps;
  }
}
// Surface term and outputs of one ray x channel (forwardmodel.c:393-399).
__device__ __forceinline__ void finish64(const PathArgs& A, const double* prow, int id, int ray, double rad,
                                         double tau) {
  const double ts = A.tsurf64 ? A.tsurf64[ray] : (double)A.tsurf[ray];
  if (ts > 0.0) rad += planck64(prow, ts) * tau;
  const size_t o = (size_t)id * A.out_ld + A.out_off + ray;
  if (A.rad64) {
    A.rad64[o] = rad;
    A.tau64[o] = tau;
  }
  A.rad[o] = (float)rad;
  A.tau[o] = (float)tau;
}

// MODE 0: plain; 1: write the snapshots of base rays (snap_mode 1); 2: reuse
// base rays (snap_mode 2).  Modes 1 and 2 read FP64 records.
template <int NGT, bool HINT, bool R64, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) k_path64(PathArgs A) {
  constexpr bool REUSE = MODE == 2, SNAP = MODE == 1;
  const int lane = threadIdx.x & 63;
  int chunk, group;
  if (!map_work(A, 1, chunk, group)) return;
  const int slot = group * 64 + lane;
  const bool valid = slot < A.nray;
  const int ray = valid ? A.order[slot] : 0;
  const int npts = valid ? A.np[ray] : 0;
  const int id = chunk;
  const int uo = rec_u(A.nw);
  const ChanConst64& cc = A.cc64[id];
  const double* prow = A.planck64 + (size_t)id * JUR_NPLANCK;
  double tau_path[NGT];
  // bracket hints of the previous point, 7 per emitter (p, T at both
  // pressures, four corners), in LDS
  __shared__ short hint_lds[HINT ? 4 : 1][HINT ? NGT * 7 : 1][64];
  short* hbase = &hint_lds[HINT ? (threadIdx.x >> 6) : 0][0][lane];
#pragma unroll
  for (int g = 0; g < NGT; g++) {
    tau_path[g] = 1.0;
    if (HINT)
      for (int i = 0; i < 7; i++) hbase[(g * 7 + i) * 64] = 0;
  }
  double rad = 0.0, tau = 1.0;
  int kj = (valid && A.scattering > 0) ? A.jbase[ray] : 0;
  // reuse of a base ray's path (snap_mode 2): its state before point ip0,
  // and per emitter the first point fd[g] whose inputs to that emitter's
  // chain differ from the base's (before it, the base's factor 1 - eps)
  const int snw = 2 * A.ng + 4;
  double* sp = nullptr;
  int ip0 = 0, fc = 0;
  int fd[NGT];
#pragma unroll
  for (int g = 0; g < NGT; g++) fd[g] = 0;
  if (MODE != 0 && valid) {
    sp = A.snap + ((size_t)(ray % A.pnr) * A.nd + id) * (size_t)(A.cap + 1) * snw;
    if (REUSE) {
      ip0 = A.pstart[ray];
      const double* st = sp + (size_t)ip0 * snw;
#pragma unroll
      for (int g = 0; g < NGT; g++)
        if (g < A.ng) {
          tau_path[g] = st[g];
          fd[g] = A.fdiff[(size_t)ray * A.ng + g];
        }
      rad = st[A.ng];
      tau = st[A.ng + 1];
      fc = A.fcont[ray];
    }
  }
  auto save = [&](int ip) {
    double* st = sp + (size_t)ip * snw;
#pragma unroll
    for (int g = 0; g < NGT; g++)
      if (g < A.ng) st[g] = tau_path[g];
    st[A.ng] = rad;
    st[A.ng + 1] = tau;
  };
  // copies of one base ray share a wave (sorted so): its lanes walk the same
  // point index, each from its own first point
  int lo = ip0, hi = npts;
  if (REUSE) {
    lo = valid ? ip0 : 0x7fffffff;
    for (int o = 32; o > 0; o >>= 1) {
      lo = min(lo, __shfl_xor(lo, o, 64));
      hi = max(hi, __shfl_xor(hi, o, 64));
    }
  }
  // point records: the FP64 copy when the trace wrote one (R64)
  const size_t rbase = (size_t)ray * A.cap * A.R;
  auto rec = [&](int ip, int f) -> double {
    const size_t i = rbase + (size_t)ip * A.R + f;
    return R64 ? A.rec64[i] : (double)A.rec[i];
  };
  for (int ip = lo; ip < hi; ip++) {
    if (REUSE && (ip < ip0 || ip >= npts)) continue;
    if (SNAP) save(ip);
    const double p = rec(ip, REC_P), t = rec(ip, REC_T), ds = rec(ip, REC_DS), af = rec(ip, REC_AF);
    const double qh = rec(ip, REC_QH2O);
    const double* st = REUSE ? sp + (size_t)ip * snw + A.ng + 2 : nullptr;  // the base's factors here
    double tau_gas = 1.0;
#pragma unroll
    for (int g = 0; g < NGT; g++)
      if (g < A.ng) {
        double f;
        if (REUSE && ip < fd[g]) {
          // the base's chain: the same factor, the same path transmittance
          f = st[g];
          tau_path[g] *= f;
        } else {
          f = 1.0 - ega64<HINT>(A.T, id * A.ng + g, p, t, rec(ip, uo + g), tau_path[g],
                                LaneHints{hbase + g * 7 * 64});
        }
        if (SNAP) sp[(size_t)ip * snw + A.ng + 2 + g] = f;
        tau_gas *= f;
      }
    if (MODE == 0)
      point64(A, cc, prow, id, uo, rec, ip, p, t, ds, af, qh, tau_gas, kj, rad, tau);
    else
      point64<MODE>(A, cc, prow, id, uo, rec, ip, p, t, ds, af, qh, tau_gas, kj, rad, tau,
                    sp + (size_t)ip * snw + 2 * A.ng + 2, ip < fc);
  }
  if (!valid) return;
  if (SNAP) save(npts);
  finish64(A, prow, id, ray, rad, tau);
}

// The same integration with a ray's emitters on separate lanes: 8 lanes per
// ray (lane g < ng runs emitter g's EGA chain, its hints in registers),
// 8 rays per wave.  The segment transmittance is the product of the lanes'
// 1 - eps in emitter order (shuffles), so the arithmetic and its order are
// k_path64's; the dependent chain per LOS point is one emitter long instead
// of ng.  For calls too small to fill the GPU with one lane per ray (the
// few-ray forward models of a retrieval).  Lanes of a ray stay converged:
// they share its point count.
template <bool HINT>
__global__ void __launch_bounds__(256) k_path64e(PathArgs A) {
  const int lane = threadIdx.x & 63, sub = lane & 7;
  const int wv = threadIdx.x >> 6, wpb = blockDim.x >> 6;
  const int ngroups = (A.nray + 7) >> 3;
  const int bpc = (ngroups + wpb - 1) / wpb;
  const int id = blockIdx.x / bpc;
  const int group = (blockIdx.x - id * bpc) * wpb + wv;
  if (id >= A.nd || group >= ngroups) return;
  const int slot = group * 8 + (lane >> 3);
  const bool valid = slot < A.nray;
  const int ray = valid ? A.order[slot] : 0;
  const int npts = valid ? A.np[ray] : 0;
  const bool gl = sub < A.ng;
  const int uo = rec_u(A.nw);
  const ChanConst64& cc = A.cc64[id];
  const double* prow = A.planck64 + (size_t)id * JUR_NPLANCK;
  int hv[7] = {0, 0, 0, 0, 0, 0, 0};
  double tau_path = 1.0, rad = 0.0, tau = 1.0;
  int kj = (valid && A.scattering > 0) ? A.jbase[ray] : 0;
  const size_t rbase = (size_t)ray * A.cap * A.R;
  auto rec = [&](int ip, int f) -> double {
    const size_t i = rbase + (size_t)ip * A.R + f;
    return A.rec64 ? A.rec64[i] : (double)A.rec[i];
  };
  const int gbase = lane & ~7;
  // snap_mode 1: the path state before every point and after the last, and
  // each emitter's factor 1 - eps (the layout of k_path64)
  const int snw = 2 * A.ng + 4;
  double* sp = (A.snap_mode == 1 && valid) ? A.snap + ((size_t)(ray % A.pnr) * A.nd + id) * (size_t)(A.cap + 1) * snw
                                           : nullptr;
  auto save = [&](int ip) {
    double* st = sp + (size_t)ip * snw;
    if (gl) st[sub] = tau_path;
    if (sub == 0) {
      st[A.ng] = rad;
      st[A.ng + 1] = tau;
    }
  };
  for (int ip = 0; ip < npts; ip++) {
    if (sp) save(ip);
    const double p = rec(ip, REC_P), t = rec(ip, REC_T), ds = rec(ip, REC_DS), af = rec(ip, REC_AF);
    const double qh = rec(ip, REC_QH2O);
    double f = 1.0;
    if (gl) f = 1.0 - ega64<HINT>(A.T, id * A.ng + sub, p, t, rec(ip, uo + sub), tau_path, RegHints{hv});
    if (sp && gl) sp[(size_t)ip * snw + A.ng + 2 + sub] = f;
    double tau_gas = 1.0;
    for (int g = 0; g < A.ng; g++) tau_gas *= __shfl(f, gbase + g, 64);
    if (sp) {
      // every lane of the ray computes E and B; lane 0 writes them
      double eb[2];
      point64<1>(A, cc, prow, id, uo, rec, ip, p, t, ds, af, qh, tau_gas, kj, rad, tau, eb);
      if (sub == 0) {
        sp[(size_t)ip * snw + 2 * A.ng + 2] = eb[0];
        sp[(size_t)ip * snw + 2 * A.ng + 3] = eb[1];
      }
    } else {
      point64(A, cc, prow, id, uo, rec, ip, p, t, ds, af, qh, tau_gas, kj, rad, tau);
    }
  }
  if (sp) save(npts);
  if (!valid || sub != 0) return;
  finish64(A, prow, id, ray, rad, tau);
}

// Reuse of base rays (ray i >= pnr repeats ray i % pnr up to where its FP64
// records differ): one wave per ray compares 64 points at a time.  start[i]:
// the first point with any field different (or the shorter point count);
// fdiff[i][g]: the first point where emitter g's inputs (p, T, u_g) differ;
// rem[i] = np[i] - start[i]; evals[i] = sum_g (np[i] - fdiff[i][g]).  Bases:
// start = fdiff = np, rem = evals = 0.  key[i] (sort key, descending):
// keyed != 0: ((class * pnr + base) << 12) | min(start, 4095), class 0 when
// no emitter differs, 1 + g when only emitter g does, ng + 1 otherwise, so
// that a wave holds copies of one base ray that compute the same emitters;
// keyed == 0: rem.
__global__ void __launch_bounds__(256) k_reuse_start(int nray, int pnr, int R, int ng, int uo, int cap,
                                                     unsigned long long cmask, const double* rec64, const int* np,
                                                     int* start, int* rem, int* fdiff, int* fcont, int* evals,
                                                     int* key, int keyed) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= nray) return;
  const int n = np[i];
  if (i < pnr) {
    if (lane < ng) fdiff[(size_t)i * ng + lane] = n;
    if (lane == 0) {
      start[i] = n;
      fcont[i] = n;
      rem[i] = 0;
      evals[i] = 0;
      key[i] = keyed ? ((i << 12) | min(n, 4095)) : 0;
    }
    return;
  }
  const int b = i % pnr, nb = np[b], nmin = n < nb ? n : nb;
  const double* ri = rec64 + (size_t)i * cap * R;
  const double* rb = rec64 + (size_t)b * cap * R;
  int f_any = nmin, f_c = nmin;
  int f_g = nmin;  // lane g < ng: emitter g's first differing point
  unsigned long long open = (ng < 64 ? (1ull << ng) - 1 : ~0ull);  // emitters not yet found
  bool any_open = true, c_open = true;
  for (int f0 = 0; f0 < nmin && (any_open || c_open || open); f0 += 64) {
    const int f = f0 + lane;
    bool d_any = false, d_pt = false, d_c = false;
    unsigned long long d_u = 0;  // bit g: u_g differs
    if (f < nmin) {
      const double* a = ri + (size_t)f * R;
      const double* c = rb + (size_t)f * R;
      for (int k = 0; k < R; k++) {
        const bool d = __double_as_longlong(a[k]) != __double_as_longlong(c[k]);
        d_any |= d;
        if (k == REC_P || k == REC_T) d_pt |= d;
        if (k >= uo && d) d_u |= 1ull << (k - uo);
        // the point's clear-air optics: every field but the column
        // amounts of emitters without a continuum
        if (k < uo || ((cmask >> (k - uo)) & 1ull)) d_c |= d;
      }
    }
    if (any_open) {
      const unsigned long long m = __ballot(d_any);
      if (m) {
        f_any = f0 + __builtin_ctzll(m);
        any_open = false;
      }
    }
    if (c_open) {
      const unsigned long long m = __ballot(d_c);
      if (m) {
        f_c = f0 + __builtin_ctzll(m);
        c_open = false;
      }
    }
    for (int g = 0; g < ng; g++) {
      if (!((open >> g) & 1ull)) continue;
      const unsigned long long m = __ballot(d_pt || ((d_u >> g) & 1ull));
      if (m) {
        if (lane == g) f_g = f0 + __builtin_ctzll(m);
        open &= ~(1ull << g);
      }
    }
  }
  // a differing emitter input is a differing field: f_g >= f_any
  int ev = lane < ng ? n - f_g : 0, cnt = lane < ng && f_g < n ? 1 : 0;
  int who = lane < ng && f_g < n ? lane : 0;
  for (int o = 32; o > 0; o >>= 1) {
    ev += __shfl_xor(ev, o, 64);
    cnt += __shfl_xor(cnt, o, 64);
    who = max(who, __shfl_xor(who, o, 64));
  }
  if (lane < ng) fdiff[(size_t)i * ng + lane] = f_g;
  if (lane == 0) {
    start[i] = f_any;
    fcont[i] = f_c;
    rem[i] = n - f_any;
    evals[i] = ev;
    const int cls = cnt == 0 ? 0 : (cnt == 1 ? 1 + who : ng + 1);
    key[i] = keyed ? (((cls * pnr + b) << 12) | min(f_any, 4095)) : n - f_any;
  }
}

void launch_64e(const PathArgs& A, int wpb, hipStream_t s) {
  const int ngroups = (A.nray + 7) / 8;
  const long nb = (long)A.nd * ((ngroups + wpb - 1) / wpb);
  if (A.T.shared)
    hipLaunchKernelGGL(k_path64e<true>, dim3((unsigned)nb), dim3(64 * wpb), 0, s, A);
  else
    hipLaunchKernelGGL(k_path64e<false>, dim3((unsigned)nb), dim3(64 * wpb), 0, s, A);
}

template <int NGT>
void launch_64(const PathArgs& A, int wpb, hipStream_t s) {
  const int ngroups = (A.nray + 63) / 64;
  long nb = (long)A.nd * ((ngroups + wpb - 1) / wpb);
  nb = (nb + 7) / 8 * 8;
  const bool r64 = A.rec64 != nullptr;
  const bool hint = A.T.shared && NGT <= 8;
  if (A.snap_mode == 2) {
    // reuse of base rays (FP64 records only)
    if (hint)
      hipLaunchKernelGGL((k_path64<NGT, NGT <= 8, true, 2>), dim3((unsigned)nb), dim3(64 * wpb), 0, s, A);
    else
      hipLaunchKernelGGL((k_path64<NGT, false, true, 2>), dim3((unsigned)nb), dim3(64 * wpb), 0, s, A);
  } else if (A.snap_mode == 1) {
    if (hint)
      hipLaunchKernelGGL((k_path64<NGT, NGT <= 8, true, 1>), dim3((unsigned)nb), dim3(64 * wpb), 0, s, A);
    else
      hipLaunchKernelGGL((k_path64<NGT, false, true, 1>), dim3((unsigned)nb), dim3(64 * wpb), 0, s, A);
  } else if (hint) {
    if (r64)
      hipLaunchKernelGGL((k_path64<NGT, NGT <= 8, true, 0>), dim3((unsigned)nb), dim3(64 * wpb), 0, s, A);
    else
      hipLaunchKernelGGL((k_path64<NGT, NGT <= 8, false, 0>), dim3((unsigned)nb), dim3(64 * wpb), 0, s, A);
  } else {
    if (r64)
      hipLaunchKernelGGL((k_path64<NGT, false, true, 0>), dim3((unsigned)nb), dim3(64 * wpb), 0, s, A);
    else
      hipLaunchKernelGGL((k_path64<NGT, false, false, 0>), dim3((unsigned)nb), dim3(64 * wpb), 0, s, A);
  }
}

template <int NGT, int CPW, int OCC, bool STAGE = false, bool TILE = false, bool PACK = false>
void launch_t(const PathArgs& A, int wpb, hipStream_t s) {
  int nchunk = (A.nd + CPW - 1) / CPW;
  int ngroups = (A.nray + 63) / 64;
  long nb = (long)nchunk * ((ngroups + wpb - 1) / wpb);
  nb = (nb + 7) / 8 * 8;  // a multiple of the XCD count for the logical remap
  // compile-time branches: only the staged variants instantiate the staged
  // kernel, and only the packed one its packed form
  if constexpr (PACK) {
    const size_t lds = A.T.shared && A.pack ? stage_lds_bytes(A, NGT, CPW, wpb, true) : 0;
    if (lds > 0) {
      hipLaunchKernelGGL(k_pack, dim3((unsigned)ngroups, (unsigned)((A.cap + PACK_SLOTS - 1) / PACK_SLOTS)),
                         dim3(256), 0, s, A);
      hipLaunchKernelGGL((k_path_stage<NGT, CPW, OCC, TILE, true>), dim3((unsigned)nb), dim3(64 * wpb),
                         (unsigned)lds, s, A);
      return;
    }
  }
  if constexpr (STAGE) {
    const size_t lds = A.T.shared ? stage_lds_bytes(A, NGT, CPW, wpb) : 0;
    if (lds > 0) {
      hipLaunchKernelGGL((k_path_stage<NGT, CPW, OCC, TILE>), dim3((unsigned)nb), dim3(64 * wpb), (unsigned)lds,
                         s, A);
      return;
    }
  }
  if (A.T.shared)
    hipLaunchKernelGGL((k_path_fast<NGT, CPW, OCC>), dim3((unsigned)nb), dim3(64 * wpb), 0, s, A);
  else
    hipLaunchKernelGGL((k_path<NGT, CPW>), dim3((unsigned)nb), dim3(64 * wpb), 0, s, A);
}

// Variants (kernel, channels per wave, occupancy target) picked by the
// tuner: 0-6 the fast kernel, 7-9 the staged kernel (the fast one where the
// grids are too large to stage), 10 the staged kernel on the packed stream
// of k_pack (variant 9 where no stream buffer is given).
template <int NGT>
void launch_ng(const PathArgs& A, int wpb, int variant, hipStream_t s) {
  switch (variant) {
    case 0: launch_t<NGT, 1, 4>(A, wpb, s); break;
    case 1: launch_t<NGT, 1, 8>(A, wpb, s); break;
    case 2: launch_t<NGT, 2, 4>(A, wpb, s); break;
    case 3: launch_t<NGT, 4, 2>(A, wpb, s); break;
    case 4: launch_t<NGT, 4, 4>(A, wpb, s); break;
    case 5: launch_t<NGT, 2, 3>(A, wpb, s); break;
    case 7: launch_t<NGT, 1, 4, true>(A, wpb, s); break;
    case 8: launch_t<NGT, 2, 3, true>(A, wpb, s); break;
    case 9: launch_t<NGT, 1, 4, true, true>(A, wpb, s); break;
    case 10: launch_t<NGT, 1, 4, true, true, true>(A, wpb, s); break;
    default: launch_t<NGT, 2, 5>(A, wpb, s); break;
  }
}

}  // namespace

int path_variants() { return 11; }

void launch_path(const PathArgs& A, int wpb, hipStream_t s) {
  if (A.nray <= 0 || A.nd <= 0) return;
  if (A.fp64) {
    // emitters on separate lanes when one lane per ray leaves the GPU
    // mostly idle (fewer than 2048 waves: 8 per CU)
    const long waves = (long)A.nd * ((A.nray + 63) / 64);
    if (A.ng <= 8 && A.snap_mode != 2 && (A.fp64_split > 0 || (A.fp64_split < 0 && waves < 2048))) {
      launch_64e(A, 4, s);
      return;
    }
    if (A.ng <= 2) launch_64<2>(A, 4, s);
    else if (A.ng <= 5) launch_64<5>(A, 4, s);
    else launch_64<NGMAX>(A, 4, s);
    return;
  }
  const int v = A.cpw;  // variant index
  if (A.ng <= 2) launch_ng<2>(A, wpb, v, s);
  else if (A.ng <= 3) launch_ng<3>(A, wpb, v, s);
  else if (A.ng <= 5) launch_ng<5>(A, wpb, v, s);
  else if (A.ng <= 8) launch_ng<8>(A, wpb, v < 2 ? v : 0, s);
  else launch_t<NGMAX, 1, 2>(A, wpb, s);
}

void launch_planck_table(const double* fnu, const double* fw, const int* foff, const int* fn, int nd,
                         float* table, double* table64, hipStream_t s) {
  int n = nd * JUR_NPLANCK;
  hipLaunchKernelGGL(k_planck_table, dim3((n + 255) / 256), dim3(256), 0, s, fnu, fw, foff, fn, nd,
                     table, table64);
}

size_t sort_temp_bytes(int n) {
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, (const int*)nullptr, (int*)nullptr,
                                               (const int*)nullptr, (int*)nullptr, n);
  return bytes;
}

void launch_sort_rays(const int* np, int* keys_tmp, int* keys_out, int* ids_in, int* ids_out,
                      void* temp, size_t temp_bytes, int n, hipStream_t s) {
  (void)keys_tmp;
  (void)hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, np, keys_out, ids_in, ids_out, n,
                                               0, 32, s);
}

void launch_iota(int* ids, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, ids, n);
}
// Sum of the LOS point counts of a ray set into *acc (one atomic per block).
__global__ void k_sum_points(const int* np, int n, unsigned long long* acc) {
  __shared__ unsigned long long part[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = i < n ? (unsigned long long)np[i] : 0ull;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(acc, part[0] + part[1] + part[2] + part[3]);
}
__global__ void k_zero_u64(unsigned long long* p, int n) {
  const int i = threadIdx.x;
  if (i < n) p[i] = 0ull;
}

void launch_sum_points(const int* np, int n, unsigned long long* acc, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_sum_points, dim3((n + 255) / 256), dim3(256), 0, s, np, n, acc);
}
void launch_zero_u64(unsigned long long* p, hipStream_t s, int n) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, p, n);
}
// Counting sort of rays by descending LOS point count (keys 0..nbins-1), for
// captured graphs: three kernels, no library temporaries or memsets.  The
// order within one count is arbitrary (rays are independent).
__global__ void k_count_hist(const int* np, int n, int nbins, int* hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&hist[min(max(np[i], 0), nbins - 1)], 1);
}
// In place: hist[b] <- number of rays with a larger count (one block).
__global__ void __launch_bounds__(1024) k_hist_offsets(int* hist, int nbins) {
  __shared__ int part[1024];
  const int t = threadIdx.x, per = (nbins + 1023) / 1024;
  // thread t owns bins hi..lo in descending order: b = nbins - 1 - (t per + j)
  int sum = 0;
  for (int j = 0; j < per; j++) {
    const int b = nbins - 1 - (t * per + j);
    if (b >= 0) sum += hist[b];
  }
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - sum;  // exclusive prefix of thread t
  for (int j = 0; j < per; j++) {
    const int b = nbins - 1 - (t * per + j);
    if (b >= 0) {
      const int c = hist[b];
      hist[b] = run;
      run += c;
    }
  }
}
__global__ void k_scatter_order(const int* np, int n, int nbins, int* cursor, int* order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) order[atomicAdd(&cursor[min(max(np[i], 0), nbins - 1)], 1)] = i;
}
// The same two passes with a block-local histogram in LDS (nbins <=
// SORT_LDS_BINS): rays of equal count are frequent, and one global atomic
// per (block, bin) replaces one per ray.  A block takes SORT_PER rays per
// thread.
constexpr int SORT_LDS_BINS = 4096, SORT_PER = 8;
__global__ void __launch_bounds__(256) k_count_hist_lds(const int* np, int n, int nbins, int* hist) {
  __shared__ int lh[SORT_LDS_BINS];
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) lh[b] = 0;
  __syncthreads();
  const int i0 = blockIdx.x * blockDim.x * SORT_PER + threadIdx.x;
  for (int j = 0; j < SORT_PER; j++) {
    const int i = i0 + j * blockDim.x;
    if (i < n) atomicAdd(&lh[min(max(np[i], 0), nbins - 1)], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += blockDim.x)
    if (lh[b]) atomicAdd(&hist[b], lh[b]);
}
__global__ void __launch_bounds__(256) k_scatter_order_lds(const int* np, int n, int nbins, int* cursor,
                                                           int* order) {
  __shared__ int lh[SORT_LDS_BINS];
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) lh[b] = 0;
  __syncthreads();
  const int i0 = blockIdx.x * blockDim.x * SORT_PER + threadIdx.x;
  int bin[SORT_PER], rank[SORT_PER];
  for (int j = 0; j < SORT_PER; j++) {
    const int i = i0 + j * blockDim.x;
    bin[j] = i < n ? min(max(np[i], 0), nbins - 1) : -1;
    rank[j] = bin[j] >= 0 ? atomicAdd(&lh[bin[j]], 1) : 0;
  }
  __syncthreads();
  // the block's range of each bin, reserved with one global atomic
  for (int b = threadIdx.x; b < nbins; b += blockDim.x)
    if (lh[b]) lh[b] = atomicAdd(&cursor[b], lh[b]);
  __syncthreads();
  for (int j = 0; j < SORT_PER; j++)
    if (bin[j] >= 0) order[lh[bin[j]] + rank[j]] = i0 + j * blockDim.x;
}

void launch_count_sort(const int* np, int n, int nbins, int* hist, int* order, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_int, dim3((nbins + 255) / 256), dim3(256), 0, s, hist, nbins);
  if (nbins <= SORT_LDS_BINS) {
    const int nb = (n + 256 * SORT_PER - 1) / (256 * SORT_PER);
    hipLaunchKernelGGL(k_count_hist_lds, dim3(nb), dim3(256), 0, s, np, n, nbins, hist);
    hipLaunchKernelGGL(k_hist_offsets, dim3(1), dim3(1024), 0, s, hist, nbins);
    hipLaunchKernelGGL(k_scatter_order_lds, dim3(nb), dim3(256), 0, s, np, n, nbins, hist, order);
    return;
  }
  hipLaunchKernelGGL(k_count_hist, dim3((n + 255) / 256), dim3(256), 0, s, np, n, nbins, hist);
  hipLaunchKernelGGL(k_hist_offsets, dim3(1), dim3(1024), 0, s, hist, nbins);
  hipLaunchKernelGGL(k_scatter_order, dim3((n + 255) / 256), dim3(256), 0, s, np, n, nbins, hist, order);
}

void launch_zero_int(int* p, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_zero_int, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}

__global__ void k_fill_f64(double* p, size_t n, double v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void launch_fill_f64(double* p, size_t n, double v, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

// Work of the reuse pass as the waves see it (lanes = 64 consecutive rays of
// the sorted order): per wave the point iterations max(np) - min(start), the
// emitter iterations sum_g max(np) - min(fdiff_g) over lanes still computing
// g, and the clear-air optics iterations, summed into acc[0..2]; waves and
// their point iterations by the number k of emitters computed (k = 0 .. 8,
// 8 for 8 or more) into acc[3 + 2 k], acc[4 + 2 k].
__global__ void __launch_bounds__(256) k_reuse_waves(const int* order, const int* start, const int* np,
                                                     const int* fdiff, const int* fcont, int nray, int ng,
                                                     unsigned long long* acc) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w * 64 >= nray) return;
  const int slot = w * 64 + lane;
  const bool v = slot < nray;
  const int r = v ? order[slot] : 0;
  const int n = v ? np[r] : 0;
  int lo = v ? start[r] : 0x7fffffff, hi = n, clo = v && fcont[r] < n ? fcont[r] : 0x7fffffff;
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o, 64));
    hi = max(hi, __shfl_xor(hi, o, 64));
    clo = min(clo, __shfl_xor(clo, o, 64));
  }
  long ge = 0;
  int ne = 0;
  for (int g = 0; g < ng; g++) {
    int f = v && fdiff[(size_t)r * ng + g] < n ? fdiff[(size_t)r * ng + g] : 0x7fffffff;
    for (int o = 32; o > 0; o >>= 1) f = min(f, __shfl_xor(f, o, 64));
    if (f < hi) {
      ge += hi - f;
      ne++;
    }
  }
  if (lane == 0) {
    const int span = hi > lo ? hi - lo : 0;
    atomicAdd(&acc[0], (unsigned long long)span);
    atomicAdd(&acc[1], (unsigned long long)ge);
    atomicAdd(&acc[2], (unsigned long long)(clo < hi ? hi - clo : 0));
    // by the number of emitters the wave computes (0 .. 8+)
    const int k = ne < 8 ? ne : 8;
    atomicAdd(&acc[3 + 2 * k], 1ull);
    atomicAdd(&acc[4 + 2 * k], (unsigned long long)span);
  }
}
void launch_reuse_waves(const int* order, const int* start, const int* np, const int* fdiff, const int* fcont,
                        int nray, int ng, unsigned long long* acc, hipStream_t s) {
  const int waves = (nray + 63) / 64;
  if (waves > 0)
    hipLaunchKernelGGL(k_reuse_waves, dim3((waves + 3) / 4), dim3(256), 0, s, order, start, np, fdiff, fcont, nray,
                       ng, acc);
}
void launch_reuse_start(int nray, int pnr, int R, int ng, int nw, int cap, unsigned long long cmask,
                        const double* rec64, const int* np, int* start, int* rem, int* fdiff, int* fcont, int* evals,
                        int* key, int keyed, hipStream_t s) {
  if (nray > 0)
    hipLaunchKernelGGL(k_reuse_start, dim3((nray + 3) / 4), dim3(256), 0, s, nray, pnr, R, ng, rec_u(nw), cap,
                       cmask, rec64, np, start, rem, fdiff, fcont, evals, key, keyed);
}
