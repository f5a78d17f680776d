// jur_kernels.h -- kernel argument blocks and launchers (host <-> device).
#pragma once
#include <hip/hip_runtime.h>

#include "jur_dev.h"

// Status bits a kernel raises in a device word; the host turns them into the
// reference's error exits.
enum {
  JUR_ST_CAP = 1,            // LOS buffer too small (host retries bigger)
  JUR_ST_NLOS = 2,           // "Too many LOS points!"
  JUR_ST_BELOW_SURFACE = 4,  // "Observer below surface!"
  JUR_ST_QUEUE = 8,          // device queue overflow
  JUR_ST_PLAN = 16,          // a batch outgrew the capacities it was launched with (host re-runs it)
};

struct TraceArgs {
  int nray;
  const double *obsz, *obslon, *obslat, *vpz, *vplon, *vplat;
  const int* atm_idx;  // atmosphere of each ray (null: all use atmosphere 0)
  DevAtm atm;
  double rayds, raydz, transs;
  int refrac;
  int sca_n;           // > 0: insert aerosol-layer edge points
  const double* alti;  // layer-edge altitudes sorted descending (device)
  int nalti;
  double altimax, altimin;
  int nl;
  double top[NLMAX], bottom[NLMAX], trans[NLMAX];
  int nw, ig_h2o, R, cap;
  double4* geo;        // [nray][cap] march points (z, lon, lat, ds)
  double4* geo2;       // [nray][cap] points after layer insertion
  float* rec;          // [nray][cap][R] FP32 point records
  double* rec64;       // optional FP64 copy of the records
  int* np_out;
  float* tsurf_out;
  double* tsurf64;     // optional
  double *tpz, *tplon, *tplat;
  int* geo_final;      // optional: 1 if the final points are in geo2
  int* status;
  // optional device-side count of the valid rays of a batch sized by its
  // capacity: ray r is empty when r + nvalid_off >= *nvalid
  const int* nvalid;
  int nvalid_off;
  // optional march order: thread i of k_raytrace traces ray order[i]
  // (rays of similar straight-line length share a wave)
  const int* order;
};
void launch_raytrace(const TraceArgs& A, hipStream_t s);
// Straight-line length class of every ray for the march order
// (k_trace_key): 0 for lines that reach the ground or leave upwards, else
// 1000 - the tangent altitude in km (longer marches larger); keys < 1001.
constexpr int TRACE_KEY_BINS = 1001;
void launch_trace_key(const TraceArgs& A, int* key, hipStream_t s);
// Log-slopes of the pressure of n consecutive levels (k_log_slopes).
void launch_log_slopes(const double* z, const double* p, long n, double* lps, hipStream_t s);

struct PathArgs {
  int nray, nd, ng, nw, R, cap;
  const float* rec;
  const int* np;
  const float* tsurf;
  const int* order;        // rays sorted by descending point count
  DevTables T;
  const ChanConst* cc;     // [nd]
  const float* planck;     // [nd][JUR_NPLANCK]
  const float *beta_e, *beta_s, *beta_a;  // [NLMAX][nd] layer optics
  const double *beta_e64, *beta_s64, *beta_a64;  // the same in FP64 (FP64 kernels)
  int sca_n;
  int ext_beta_e;          // leaf extinction: 1 = beta_e, 0 = beta_a
  int scattering;          // > 0: points with aerofac > 0 use the source J
  const float* J;          // [nJ][nd]
  long nJ;                 // rows of J (a point at or beyond them reads no source)
  const int* jbase;        // [nray] index of the ray's first scattering point
  float* rad;              // [nd][out_ld], ray r at column out_off + r
  float* tau;
  int out_ld, out_off;
  int cpw;                 // kernel variant (channels per wave x occupancy), see launch_path
  int fp64;                // 1: FP64 arithmetic (k_path64); also writes rad64/tau64 if set
  int fp64_split;          // FP64: emitters on separate lanes (k_path64e): 1 on, 0 off, -1 small calls
  // FP64 reuse of base rays (clear air): rays [0, pnr) are bases, ray i >=
  // pnr repeats base i % pnr up to its first differing LOS point.
  // snap_mode 1: the (base) rays write their path state before every point
  // and after the last, each emitter's factor 1 - eps and the point's
  // clear-air optics at every point, into snap [pnr][nd][cap + 1][2 ng + 4]
  // (tau_path per emitter, rad, tau, factor per emitter, exp(-beta ds), B);
  // 2: every ray starts at pstart[ray] from its base's state there, takes
  // emitter g's factor from the base before point fdiff[ray][g] and the
  // optics before point fcont[ray].
  int snap_mode, pnr;
  const int* pstart;
  const int* fdiff;
  const int* fcont;
  double* snap;
  double *rad64, *tau64;   // [nd][out_ld] (optional)
  const double* rec64;     // FP64 records / surface temperatures (optional, fp64)
  const double* tsurf64;
  const ChanConst64* cc64; // [nd] (fp64)
  const double* planck64;  // [nd][JUR_NPLANCK] (fp64)
  // packed records and (p, T) brackets per ray group (64 sorted rays) and
  // point slot, [ngroups][cap][R + 4 ng][64] floats, written by the path
  // launch itself for variant 10 (optional: without it variant 10 runs as
  // variant 9)
  float* pack;
};
int path_variants();
// Launch the path integration with the given waves per block.
void launch_path(const PathArgs& A, int waves_per_block, hipStream_t s);

// Filter-weighted Planck table [nd][JUR_NPLANCK] (forwardmodel.c:743-786).
void launch_planck_table(const double* filt_nu, const double* filt_w, const int* filt_off,
                         const int* filt_n, int nd, float* table, double* table64, hipStream_t s);

// Sort ray ids by descending point count (stable).  temp sized by
// sort_temp_bytes.
size_t sort_temp_bytes(int n);
void launch_sort_rays(const int* np, int* keys_tmp, int* keys_out, int* ids_in, int* ids_out,
                      void* temp, size_t temp_bytes, int n, hipStream_t s);
void launch_iota(int* ids, int n, hipStream_t s);
// Where ray i >= pnr starts to differ from ray i % pnr (FP64 records, point
// counts): start[i] (any field; np[i] for the bases), the points left
// rem[i] = np[i] - start[i], per emitter fdiff[i][g] (p, T or u_g), the
// emitter evaluations left evals[i], the first point fcont[i] whose
// clear-air optics inputs differ (every field but the column amounts of
// emitters outside cmask, the continuum emitters), and the sort key key[i]
// (keyed: grouped by the emitters computed and the base ray; else rem).
void launch_reuse_start(int nray, int pnr, int R, int ng, int nw, int cap, unsigned long long cmask,
                        const double* rec64, const int* np, int* start, int* rem, int* fdiff, int* fcont, int* evals,
                        int* key, int keyed, hipStream_t s);
// Iterations of the reuse pass per wave of the sorted order (points,
// emitter evaluations, clear-air optics), summed into acc[0..2]; waves and
// point iterations by emitters computed into acc[3..20].
void launch_reuse_waves(const int* order, const int* start, const int* np, const int* fdiff, const int* fcont,
                        int nray, int ng, unsigned long long* acc, hipStream_t s);
void launch_zero_int(int* p, int n, hipStream_t s);
void launch_fill_f64(double* p, size_t n, double v, hipStream_t s);
// Rays by descending point count into order (hist: nbins ints of scratch).
void launch_count_sort(const int* np, int n, int nbins, int* hist, int* order, hipStream_t s);
void launch_sum_points(const int* np, int n, unsigned long long* acc, hipStream_t s);
// p[0 .. n) = 0 (n <= 64) with a kernel (capturable, no memset node)
void launch_zero_u64(unsigned long long* p, hipStream_t s, int n = 1);
