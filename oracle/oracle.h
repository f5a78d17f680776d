/*
 * oracle.h -- FP64 CPU restatement of the reference forward model.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU library is
 * compared against (and the CPU baseline that the benchmark times beside it).
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load it; the product library never links or calls it.
 *
 * Every function restates the arithmetic of the reference file:line cited at
 * its definition in oracle.c, in FP64, without GSL.  The functions that touch
 * the reference's GSL call sites restate GSL 1.15's published algorithms
 * (gsl_expm1, gsl_log1p, gsl_complex_div/abs, gsl_stats_min_index, gsl_sort).
 *
 * Parity pinning (what checks each part; see tests/test_oracle_*.py):
 *   formod clear air, raytrace, intpol_tbl, planck table, continua (CO2/H2O):
 *       reference fixtures examples/org/rad_clear.tab, los.0, obs*.tab.
 *   add_aerosol_layers, get_opt_prop, bhmie, gauher (beta_a path):
 *       examples/org/rad_ext1.tab; bhmie also against Rayleigh limit,
 *       k=0 energy conservation, phase normalisation; gauher against exact
 *       Gaussian moments.
 *   srcfunc_sca_1d, srcfunc_sca_sun at night, scattering radiance (depth 1):
 *       examples/org/rad_aero0.tab, rad_aero1.tab; isotropic-field closure.
 *   suncoord: solstice/equinox declination and local-noon checks.
 *   brightness: closed-form inverse of planck (no reference fixture uses it).
 *   hydrostatic: closed form for an isothermal dry column.
 *   ctmn2, ctmo2: parity unpinned by fixtures (no fixture channel lies in
 *       their bands); checked against direct evaluation at table nodes.
 *   formod_fov: no reference fixture uses a FOV; pinned by closed forms
 *       (tests/test_oracle_fov_sun.py): a delta-shaped FOV reproduces the
 *       unconvolved radiance, radiances linear in tangent height give the
 *       value at the weight-averaged offset (inside and beyond the profile,
 *       un-normalised weights), other shapes the weighted mean of numpy's
 *       linear interpolation of the rays with the same time only.
 *   srcfunc_sca_sun by day: no reference fixture is by day; pinned in a
 *       transparent atmosphere with a phase function linear in the
 *       scattering angle: J = 6.764e-5 P(theta) B(5780 K) for LOS tangents at
 *       known angles from the sun direction (built from suncoord's zenith
 *       and azimuth in a textbook east-north-up frame, the same Earth-fixed
 *       vector at every point), 0 when the sun ray meets the ground, the
 *       full term when it grazes past the horizon below 90 deg + dip.
 *   Scattering depth >= 2 and fan sizes other than 28: parity unpinned by
 *       fixtures (no reference test uses SCA_MULT >= 2); checked by the
 *       closure of an isothermal opaque atmosphere (the source J equals the
 *       Planck radiance, radiance B(T) (1 - tau) or B(T) over the ground).
 *   kernel (finite-difference Jacobian): parity unpinned by fixtures (the
 *       reference ships no kernel output); its columns are checked against
 *       formod on independently perturbed atmospheres, first-order
 *       linearity, and the zero pattern of levels a ray never reaches.
 *   optimal_estimation, cov_apr, cov_meas, read_ret: parity unpinned by
 *       fixtures (no retrieval output ships); checked by closed forms with
 *       numpy: covariance elements (vertical and surface-chord
 *       correlations), the first damped step, the cost function, the error
 *       analysis against a numpy inverse, a vanishing gradient at
 *       convergence.
 *   intpol_atm_2d / 3d (IP 2, 3): parity unpinned by fixtures (every
 *       example uses IP 1); checked by projection weights computed with a dot
 *       product, clamping, IP 2 = IP 1 for identical profiles, a symmetric
 *       IP 3 window reproducing a linear profile, chord-distance weights.
 *   srcfunc_sca_3d: parity unpinned by fixtures; checked by the isothermal
 *       opaque closure (J = B(T) with one channel) and the reference's
 *       normalisation over all channels (J = B(T) / 2 for two identical
 *       channels).
 *   opt_prop_external (Ext aerosol type): parity unpinned by fixtures (no
 *       table ships); checked by the nearest-row rule with its tie, the
 *       mode sums and the NN / Const exits.
 */
#ifndef JURASSIC_ORACLE_H
#define JURASSIC_ORACLE_H

#include "jurassic_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Control file parsing (reference src/control.c:7-171). */
void oracle_read_ctl(int argc, char *argv[], ctl_t *ctl);
double oracle_scan_ctl(int argc, char *argv[], const char *varname, int arridx,
                       const char *defvalue, char *value);

/* File I/O (reference src/atmosphere.c:487-541, src/misc.c:269-390,
   src/scatter.c:583-639). */
void oracle_read_atm(const char *dirname, const char *filename, ctl_t *ctl, atm_t *atm);
void oracle_read_obs(const char *dirname, const char *filename, ctl_t *ctl, obs_t *obs);
void oracle_write_obs(const char *dirname, const char *filename, ctl_t *ctl, obs_t *obs);
void oracle_read_aero(const char *dirname, const char *filename, ctl_t *ctl, aero_t *aero);

/* Optical properties of the aerosol layers (reference src/scatter.c:330-508). */
void oracle_get_opt_prop(ctl_t *ctl, aero_t *aero);
void oracle_bhmie(double x, double n_real, double n_imag, double *phase,
                  double *qext, double *qsca);
void oracle_gauher(double *x, double *w);

/* The forward model (reference src/forwardmodel.c:15-116). */
void oracle_formod(ctl_t *ctl, atm_t *atm, obs_t *obs, aero_t *aero);
/* One pencil beam (reference src/forwardmodel.c:224-405, non-queue path). */
void oracle_formod_pencil(ctl_t *ctl, atm_t *atm, obs_t *obs, aero_t *aero,
                          int scattering, int ir);
/* Ray tracing into a caller-owned los_t (reference src/lineofsight.c:5-190). */
void oracle_raytrace(ctl_t *ctl, atm_t *atm, obs_t *obs, aero_t *aero, los_t *los, int ir);

/* Point functions. */
double oracle_planck(double t, double nu);          /* forwardmodel.c:569-573 */
double oracle_brightness(double rad, double nu);    /* forwardmodel.c:7-11 */
void oracle_suncoord(double sec, double lon, double lat, double *azi, double *sza); /* scatter.c:999-1041 */
void oracle_hydrostatic(ctl_t *ctl, atm_t *atm);    /* atmosphere.c:163-243 */
/* intpol_atm_geo (atmosphere.c:269-483): IP 1 vertical profile, 2 nearest two
   profiles, 3 distance-weighted average within CZ / CX. */
void oracle_intpol_atm_geo(const ctl_t *ctl, const atm_t *atm, double z0, double lon0, double lat0, double *p,
                           double *t, double *q, double *k);
/* Continuum absorption of one LOS point for all channels (forwardmodel.c:120-152). */
void oracle_continua(ctl_t *ctl, double p, double t, const double *q, const double *u,
                     double ds, const double *k, double *beta);
double oracle_ctmn2(double nu, double p, double t);  /* continua.c:1935-1993 */
double oracle_ctmo2(double nu, double p, double t);  /* continua.c:1997-2049 */
/* Filter-weighted Planck source function for all channels (forwardmodel.c:743-794). */
void oracle_srcfunc_planck(ctl_t *ctl, double t, double *src);

/*
 * Batched clear-air / extinction forward model with one atmosphere per ray
 * (the shape of the batched GPU entry point).  Profiles are ray-major:
 * z,p,t [nray][nlev], q [nray][ng][nlev], k [nray][nw][nlev].  Outputs rad,
 * tau [nray][nd] and the tangent point [nray].  Rays ray0..ray1-1 only.
 * nthreads <= 0 uses the OpenMP default.  aero (NULL: clear air) must hold
 * the layer optics (oracle_get_opt_prop); scattering is the recursion depth
 * of formod_pencil (0 without aero).
 */
void oracle_formod_batch(ctl_t *ctl, int nray, int nlev, const double *z,
                         const double *p, const double *t, const double *q,
                         const double *k, const double *time, const double *obsz,
                         const double *obslon, const double *obslat, const double *vpz,
                         const double *vplon, const double *vplat, int ray0, int ray1,
                         int nthreads, double *rad, double *tau, double *tpz,
                         double *tplon, double *tplat, int *npts, const aero_t *aero,
                         int scattering);

/* Retrieval: state vector (reference src/atmosphere.c:5-95), its inverse
   (src/retrievalmodel.c:1103-1162), measurement vector (:379-404), the
   perturbation sizes (:242-268) and the finite-difference Jacobian
   k [m][ldk] (:201-298).  x/iqa/ipa/y/ida/ira may be NULL (size query). */
size_t oracle_atm2x(const ctl_t *ctl, const atm_t *atm, const aero_t *aero, double *x, int *iqa,
                    int *ipa);
void oracle_x2atm(const ctl_t *ctl, const double *x, atm_t *atm, aero_t *aero);
size_t oracle_obs2y(const ctl_t *ctl, const obs_t *obs, double *y, int *ida, int *ira);
double oracle_perturbation(const ctl_t *ctl, int iq, double x);
void oracle_kernel(ctl_t *ctl, atm_t *atm, obs_t *obs, aero_t *aero, double *k, size_t m, size_t n,
                   size_t ldk);
void oracle_kernel_columns(ctl_t *ctl, atm_t *atm, obs_t *obs, aero_t *aero, double *k, size_t m,
                           size_t n, size_t ldk, const int *cols, int ncols);

/* The scattering source at Cartesian point x with LOS tangent dx in aerosol
   layer il (srcfunc_sca, scatter.c:643-662: srcfunc_sca_1d for IP 1,
   srcfunc_sca_3d otherwise, plus the solar term). */
void oracle_srcfunc_sca(ctl_t *ctl, const atm_t *atm, const aero_t *aero, double time, const double *x,
                        const double *dx, int il, double *src_sca, int scattering);

/* Optimal estimation (reference src/retrievalmodel.c:407-706, read_ret
   :710-757) without its files: chisq[] receives every cost-function value
   (at most 1 + 20 conv_itmax), *nchisq their number, err_total[n] the
   retrieval errors (square roots of the covariance diagonal, ERR_ANA). */
void oracle_read_ret(int argc, char *argv[], ctl_t *ctl, ret_t *ret);
/* The a priori covariance S_a (n x n) and the measurement errors (m each)
   of the retrieval (set_cov_apr, set_cov_meas: retrievalmodel.c:761-898). */
void oracle_cov_apr(const ret_t *ret, ctl_t *ctl, atm_t *atm_apr, aero_t *aero_apr, double *s_a);
void oracle_cov_meas(const ret_t *ret, ctl_t *ctl, obs_t *obs_meas, double *sig_noise, double *sig_formod,
                     double *sig_eps_inv);
void oracle_optimal_estimation(const ret_t *ret, ctl_t *ctl, obs_t *obs_meas, obs_t *obs_i, atm_t *atm_apr,
                               atm_t *atm_i, aero_t *aero_apr, aero_t *aero_i, double *chisq, int *nchisq,
                               double *err_total);

/* The solar term of srcfunc_sca alone (srcfunc_sca_sun, scatter.c:899-995),
   added to src_sca[nd]. */
void oracle_srcfunc_sca_sun(ctl_t *ctl, const atm_t *atm, const aero_t *aero, double sec, const double *x,
                            const double *dx, int il, double *src_sca);
/* The FOV convolution of obs's radiances and transmittances alone
   (formod_fov, forwardmodel.c:156-220; FOV shape file ctl->fov). */
void oracle_formod_fov(ctl_t *ctl, obs_t *obs);

/* Secondary pencils per scattering point (reference: 28); even, 10-32
   (30 and 32: this project's grid for the cirrus configuration, see
   oracle.c), and the fan's zenith angles [rad]. */
void oracle_set_fan(int nalpha);
void oracle_alpha_grid(int nalpha, double *alpha);
/* 1: the aerosol configuration's Mie grid (64 log-spaced radii 0.01-20 um,
   phase function at 180 angles resampled to the 181-node layout); 0: the
   reference's (170 Gauss-Hermite nodes, 181 angles). */
void oracle_set_mie_grid(int grid);
/* One sphere at the 180 angles (k + 0.5) deg (phase180, optional) and
   resampled to the 181 nodes (phase), with Q_ext from the forward amplitude. */
void oracle_bhmie180(double x, double n_real, double n_imag, double *phase180, double *phase, double *qext,
                     double *qsca);

/* Drop the cached tables, filter functions and Planck tables. */
void oracle_reset(void);

#ifdef __cplusplus
}
#endif
#endif
