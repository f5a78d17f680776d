/*
 * jurassic_amd.h -- C ABI of the MI355X forward-model library
 * (libjurassic_amd.so).
 *
 * The first group of functions has the reference's names and signatures, so
 * a caller of the reference forward model links against this library
 * instead of the reference objects: the structs are layout-identical
 * (jurassic_types.h), errors stop the process with the reference's message
 * format, and the work runs on the GPU (device 0 unless jur_init() picked
 * another).  Not reentrant: one caller thread, like the reference.
 *
 * The second group (jur_*) is new: a batched entry point with one atmosphere
 * per ray and any number of rays (the reference's obs_t stops at NRMAX=1000
 * rays and one atmosphere), its device-pointer variant, and control / stats.
 */
#ifndef JURASSIC_AMD_H
#define JURASSIC_AMD_H

#include "jurassic_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Drop-in entry points (reference interfaces) ---------------------- */

/* Forward model for all rays of obs (reference src/forwardmodel.h:16-19,
   forwardmodel.c:15-116): NaN mask, hydrostatic equilibrium, optional Mie
   refresh, ray tracing, path integration with single/multiple scattering,
   FOV convolution, brightness temperature.  Writes obs->rad, obs->tau and
   the tangent points. */
void formod(ctl_t *ctl, atm_t *atm, obs_t *obs, aero_t *aero);

/* One pencil beam, ray ir, with the given scattering depth (reference
   src/forwardmodel.h:32-37, forwardmodel.c:224-405, non-queue semantics). */
void formod_pencil(ctl_t *ctl, atm_t *atm, obs_t *obs, aero_t *aero, int scattering, int ir);

/* Ray tracing of ray ir into los (reference src/lineofsight.h, lineofsight.c:5-190,
   including add_aerosol_layers :193-372 when SCA_N > 0). */
void raytrace(ctl_t *ctl, atm_t *atm, obs_t *obs, aero_t *aero, los_t *los, int ir);

/* Layer optical properties from the aerosol models with Mie theory
   (reference src/scatter.h, scatter.c:330-508). */
void get_opt_prop(ctl_t *ctl, aero_t *aero);

/* Mie efficiencies and phase function of one sphere (reference scatter.c:56-234). */
void bhmie(double x, double n_real, double n_imag, double *phase, double *qext, double *qsca);

/* Hydrostatic pressure profile (reference src/atmosphere.c:163-184). */
void hydrostatic(ctl_t *ctl, atm_t *atm);

/* Control file and command-line keys (reference src/control.c:7-171). */
void read_ctl(int argc, char *argv[], ctl_t *ctl);
double scan_ctl(int argc, char *argv[], const char *varname, int arridx, const char *defvalue,
                char *value);

/* Text file formats (reference src/atmosphere.c:487-598, src/misc.c:269-390,
   src/scatter.c:583-639 and 1045-1092). */
void read_atm(const char *dirname, const char *filename, ctl_t *ctl, atm_t *atm);
void write_atm(const char *dirname, const char *filename, ctl_t *ctl, atm_t *atm);
void read_obs(const char *dirname, const char *filename, ctl_t *ctl, obs_t *obs);
void write_obs(const char *dirname, const char *filename, ctl_t *ctl, obs_t *obs);
void read_aero(const char *dirname, const char *filename, ctl_t *ctl, aero_t *aero);
void write_aero(const char *dirname, const char *filename, aero_t *aero);

/* Two-column shape file, e.g. filter or FOV function, at most NSHAPE points
   (reference src/forwardmodel.h:73-76, forwardmodel.c:577-606). */
void read_shape(const char *filename, double *x, double *y, int *n);

/* Copy an aerosol model; init clears its mode/layer parameters and optics
   (reference src/scatter.h:34-37, scatter.c:237-271). */
void copy_aero(ctl_t *ctl, aero_t *aero_dest, aero_t *aero_src, int init);

/* Point formulas (reference src/forwardmodel.c:7-11 and :569-573). */
double brightness(double rad, double nu);
double planck(double t, double nu);

/* ---- Retrieval: finite-difference Jacobian ------------------------------ */

/* Layout of GSL's gsl_matrix (gsl 1.15, matrix/gsl_matrix_double.h): a
   caller's gsl_matrix * may be passed where this type is expected. */
typedef struct {
  size_t size1, size2, tda;
  double *data;
  void *block;
  int owner;
} jur_gsl_matrix;

/* kernel (reference src/retrievalmodel.h, src/retrievalmodel.c:201-298):
   k[i][j] = dy_i/dx_j by forward differences.  Runs formod on the
   unperturbed state first (filling obs, applying hydrostatic balance to atm
   and refreshing retrieved aerosol optics), then all perturbed states as
   batched forward calls.  k->size1 / size2 must be the measurement / state
   vector sizes (jur_meas_size after formod, jur_state_size). */
void kernel(ctl_t *ctl, atm_t *atm, obs_t *obs, aero_t *aero, jur_gsl_matrix *k);
/* The same with a plain row-major matrix k [m][ldk]. Returns 0. */
int jur_kernel(ctl_t *ctl, atm_t *atm, obs_t *obs, aero_t *aero, double *k, size_t m, size_t n,
               size_t ldk);
/* Sizes of atm2x's state vector (reference src/atmosphere.c:5-68; aero may
   be NULL) and of obs2y's measurement vector (finite radiances, reference
   src/retrievalmodel.c:379-404). */
size_t jur_state_size(ctl_t *ctl, atm_t *atm, aero_t *aero);
size_t jur_meas_size(ctl_t *ctl, obs_t *obs);

/* ---- Retrieval: optimal estimation -------------------------------------- */

/* read_ret (reference src/retrievalmodel.h, src/retrievalmodel.c:710-757):
   iteration control and error keys (KERNEL_RECOMP, CONV_ITMAX, CONV_DMIN,
   RESMAX, ERR_ANA, ERR_FORMOD[], ERR_NOISE[], ERR_PRESS*, ERR_TEMP*, ERR_Q*,
   ERR_K*, ERR_NN/RR/SS). */
void read_ret(int argc, char *argv[], ctl_t *ctl, ret_t *ret);
/* write_matrix (reference src/retrievalmodel.h, src/retrievalmodel.c:946-1099):
   the non-zero elements of a matrix in measurement ("y") or state ("x")
   space along rows and columns, sorted by row ("r") or column ("c"); only
   with WRITE_MATRIX = 1, like the reference. */
void write_matrix(const char *dirname, const char *filename, ctl_t *ctl, jur_gsl_matrix *matrix, atm_t *atm,
                  aero_t *aero, obs_t *obs, const char *rowspace, const char *colspace, const char *sort);
/* optimal_estimation (reference src/retrievalmodel.h,
   src/retrievalmodel.c:407-706): Levenberg-Marquardt retrieval of atm_i /
   aero_i from obs_meas with a priori atm_apr / aero_apr; writes costs.tab,
   obs_final.tab, atm_final.tab, aero_final.tab and (ERR_ANA) the error
   analysis files into ret->dir, the matrices with WRITE_MATRIX.  aero_apr
   and aero_i may be NULL for clear air. */
void optimal_estimation(ret_t *ret, ctl_t *ctl, obs_t *obs_meas, obs_t *obs_i, atm_t *atm_apr, atm_t *atm_i,
                        aero_t *aero_apr, aero_t *aero_i);

/* ---- Batched extension (no reference counterpart) --------------------- */

/* Rays with their own atmospheres.  Profiles: z, p, t [natm][nlev];
   q [natm][ng][nlev]; k [natm][nw][nlev] or NULL for zero extinction.
   atm_of_ray [nray] selects a profile per ray; NULL means ray i uses profile
   i when natm == nray, else profile 0.  Outputs: rad, tau [nd][nray]
   (channel-major like obs_t, FP32) with leading dimension out_ld (0: nray;
   otherwise >= nray, so that consecutive calls can fill column blocks of
   one larger array through rad + offset); tpz/tplon/tplat [nray] and npts
   [nray] (LOS points) are optional.  Vertical profiles only: ctl->ip must
   be 1. */
typedef struct {
  int nray, nlev, natm;
  const int *atm_of_ray;
  const double *z, *p, *t, *q, *k;
  const double *time, *obsz, *obslon, *obslat, *vpz, *vplon, *vplat;
  float *rad, *tau;
  double *tpz, *tplon, *tplat;
  int *npts;
  int out_ld;
} jur_batch_t;

/* All pointers of b are host pointers.  aero may be NULL when ctl->sca_n is 0;
   scattering is the recursion depth (ctl->sca_mult for formod semantics).
   Returns 0. */
int jur_formod_batch(ctl_t *ctl, const jur_batch_t *b, const aero_t *aero, int scattering);

/* All pointers of b are device pointers already resident in GPU memory;
   stream is a hipStream_t (NULL: the library's stream).  The work runs on
   that stream; the call returns once the outputs are complete (the LOS
   buffers are sized from a device-side overflow flag, and the scattering
   tree from device-side point counts, both read back during the call). */
int jur_formod_batch_device(ctl_t *ctl, const jur_batch_t *b, const aero_t *aero, int scattering,
                            void *stream);

/* No counterpart in the reference.  For callers that cannot allocate
   device memory (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc): jur_formod_batch_device for clear air and
   extinction (scattering depth 0 when ctl->sca_n > 0), FP32 path, whose
   work buffers come from a fixed device array of JUR_NOALLOC_SCRATCH_BYTES
   per GPU (held until the process ends) instead of being allocated; it calls
   no hipMalloc / hipFree.  The caller passes inputs and outputs as for
   jur_formod_batch_device; optional ones it leaves NULL (tangent points,
   atm_of_ray, k) take room in the array too.  The array holds nray x
   (LOS capacity) x (32 B of geometry, 64 B with aerosol layers, + 4 B per
   record element: 12 at five emitters and one window) bytes, e.g. 6,000
   clear-air rays at 512 LOS points; larger batches, scattering depths > 0
   and option fp64 stop with an error.  The tables and layer optics of
   ctl / aero must already be resident (one allocating call with the same
   ctl and aero first), else it stops with an error.  All calls on one GPU
   share its array: calls that overlap on one GPU (other streams or host
   threads) race.  Returns 0 once the outputs are complete. */
#define JUR_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int jur_formod_batch_device_noalloc(ctl_t *ctl, const jur_batch_t *b, const aero_t *aero, int scattering,
                                    void *stream);

/* Wait for the library stream and check device-side error flags. */
int jur_sync(void);

/* Every emissivity table of ctl (binary if present, else ASCII) written in
   the binary layout as TBLBASE_<nu>_<emitter>.bin: what the reference's
   tab2bin tool does (src/tab2bin.c, through its padded tbl_t); the tab2bin
   command calls it.  Host only.  Returns 0. */
int jur_tab2bin(ctl_t *ctl);

/* Select the GPU (call before anything else; default device 0). */
int jur_init(int device);

/* Tuning.  Options: "variant" FP32 path-kernel variant (0-6 the fast
   kernel at several channels per wave x occupancy targets, 7-9 the staged
   kernel, 10 the staged kernel on a packed lane-contiguous copy of the
   point records with the (p, T) brackets precomputed once per LOS point and
   emitter; all give bit-identical radiances; default: autotuned),
   "wpb" waves per block (1, 2, 4; default: autotuned), "autotune" (0/1,
   default 1), "fast" (0: always use the generic table search), "cap"
   initial LOS points per ray, "chunk_mb" trace-buffer budget of the
   scattering levels in MiB (default: a sixth of the device's memory,
   at most 48 GiB),
   "graph" (1: a clear-air call is recorded once as a hipGraph -- trace,
   sort, path integration -- and replayed while the inputs, outputs and
   work buffers keep their addresses; a replay whose rays overflow the
   recorded LOS capacity is detected and the call re-run), "surrogate",
   "hull_margin", "collect" (learned scattering source, below),
   "trace_sort" (1, the default: the rays of a trace of at least 262,144
   rays are marched in the order of their straight-line tangent altitude,
   so that a wave's rays march about as long; results are identical),
   "mlp_infer" (the network's fused inference kernel: 1, the default, a
   wave per 128 samples x 64 neurons of a layer; 0 a wave per 64 samples x
   128 neurons),
   "fan" (secondary pencils per scattering point, even, 10-32; default the
   reference's 28; 10-28 use the reference's zenith-angle construction, 30
   and 32 keep its horizon band and first steps and continue in equal steps
   to 0 and 180 deg), "mie_grid" (1: get_opt_prop integrates each mode over
   64 radii log-spaced 0.01-20 um with the trapezoid rule in ln r and
   evaluates the phase function at 180 angles (k + 0.5) deg, resampled to the
   181 nodes of aero_t; 0, the default: the reference's 170 Gauss-Hermite
   nodes and 181 angles), "kernel_rays" (rays per batched call of kernel/jur_kernel, default
   131072: the perturbed atmospheres are grouped into calls of about this
   many rays), "fp64" (1: path integration in FP64 arithmetic, the
   reference's operation order, on FP64 tables grids, Planck table,
   continuum data and layer optics, with FP64 radiances for formod; the
   scattering source J of scattering calls is still formed in FP32 by the
   angular integral; default 0: FP32 with a compensated radiance sum), "kernel_fp64" (default 1: kernel
   and jur_kernel integrate the perturbed paths in FP64, since a 1 %
   perturbation of one level can change a radiance by less than FP32
   resolves; optimal_estimation's forward calls follow the same setting),
   "fp64_split" (FP64 path with a ray's emitters on separate lanes: -1, the
   default, for calls too small to fill the GPU with one lane per ray; 1
   always; 0 never; results are identical), "kernel_prefix" (default 1:
   kernel integrates each perturbed copy's rays from their first LOS point
   whose FP64 records differ from the unperturbed ray's, starting from its
   path state there, and takes each emitter's factor 1 - eps and the
   point's clear-air optics from the unperturbed ray up to the first point
   whose inputs to them differ -- bit-identical; 0: every copy integrated
   from the observer), "reuse_group" (default 1: the copies are ordered so
   that a wave holds copies of one ray that compute the same emitters).
   Returns 0. */
int jur_set_option(const char *key, double value);
/* Statistics of the last call: "path_ms" (path-integration kernel time,
   summed over its launches), "path_launches", "integrated_points" (LOS
   points integrated, all pencils), "emitter_points" (point x emitter
   evaluations made when kernel_prefix reused an unperturbed ray, else NaN),
   "reuse_all_points" (LOS points of all the call's rays, reused or not),
   "reuse_wave_points", "reuse_wave_emitters", "reuse_wave_optics" (that
   pass's iterations as its waves run them, per channel),
   "reuse_waves_e<k>", "reuse_wave_points_e<k>" (its waves computing k
   emitters, k = 0..8, and their point iterations), "trace_ms" (clear-air ray tracing),
   "pencils" (primaries + secondary + solar pencils), "leaf_pencils",
   "scattering_points", "scatter_ms" (whole scattering call),
   "tuned_variant", "tuned_wpb", "tuned_ms" (autotuner choice),
   "graph_ms" (replayed call), "graph_overflows", "surrogate_points",
   "fallback_points", "mlp_ms", "mlp_samples", "train_ms",
   "shared_grids" (1: the fast table layout is in use), "mie_spheres"
   (spheres of the last get_opt_prop), "kernel_ms" (wall time of the last
   kernel call), "kernel_columns", "oe_formod_ms", "oe_kernel_ms",
   "oe_total_ms" (wall time of the last optimal_estimation call: its
   forward models, its Jacobians, in total); running totals over all calls since
   the library loaded: "calls_total", "path_ms_total",
   "integrated_points_total", "emitter_points_total" (point x emitter
   evaluations), "pencils_total".  A replayed graph reports
   "graph_ms" (the whole call; it also adds that to "path_ms_total") and no
   per-kernel times ("path_ms", "trace_ms" NaN).
   Unknown keys give NaN. */
double jur_stat(const char *key);

/* ---- Scattering-source surrogate (no reference counterpart) ---------- */

/* A perceptron 24 -> 4 x 256 (SiLU) -> 1 predicting ln(J / B(T)), J the
   thermal scattering source at a scattering point and B(T) the Planck
   radiance at its temperature (feature 12), from 24 local features (listed
   in the library's jur_surrogate.h); BF16 weights and
   activations, FP32 accumulation, trained with Adam on the GPU.  With
   jur_set_option("surrogate", 1) and a trained or loaded network, the
   scattering forward model takes J at the primaries' scattering points
   from the network; points whose features leave the training hull
   (option "hull_margin", normalised units, default 0.25) fall back to the
   plain secondary-pencil method, which also keeps the solar term.  The
   parameter layout is jur_mlp_nparam() floats (W1 [32][256], b1, W2..W4
   [256][256], b2..b4, W5 [256], b5). */
int jur_mlp_init(unsigned seed);
/* Training pairs: x [n][24] raw features, y [n] targets (ln J - feature 12
   for the scattering source).  normalize:
   1 sets the feature / target normalisation and training hull from these
   data and shuffles the rows (seed); 0 keeps the current normalisation. */
int jur_mlp_set_data(const float *x, const float *y, long n, int normalize, unsigned seed);
/* steps Adam steps (beta 0.9 / 0.999, eps 1e-8) on batches of `batch` rows
   (multiple of 32), replaying one captured hipGraph per step; loss_hist
   [steps] (optional) receives the batch mean squared error (normalised
   units). */
int jur_mlp_train(int steps, int batch, float lr, float *loss_hist);
/* Gradients of the mean squared error on one batch (x raw features, y
   targets) without updating: grads [jur_mlp_nparam()], loss (optional). */
int jur_mlp_grads(const float *x, const float *y, int batch, float *grads, float *loss);
/* Network output (de-normalised target) for n feature rows, fused
   inference kernel. */
int jur_mlp_forward(const float *x, long n, float *y);
int jur_mlp_get_params(float *p);
int jur_mlp_set_params(const float *p);
int jur_mlp_nparam(void);
int jur_mlp_save(const char *path);
int jur_mlp_load(const char *path);
/* Training pairs of the last scattering call made with
   jur_set_option("collect", 1): raw features feat [n][24] and the plain
   method's thermal source jth [n] (no solar term) for every (primary
   scattering point, channel); copies at most cap rows, returns n. */
long jur_sca_collect(float *feat, float *jth, long cap);

#ifdef __cplusplus
}
#endif
#endif
