/*
 * jurassic_types.h -- data types shared across the drop-in boundary.
 *
 * The structs below are layout-identical to the ones a caller of the
 * reference forward model allocates (reference src/jurassic.h:228-570 and
 * src/workqueue.h:13-25): same members, same order, same fixed dimensions.
 * A caller compiled against the reference header can therefore hand its
 * ctl_t / atm_t / obs_t / aero_t objects straight to this library.  The
 * static assertions at the end pin the sizes and a few offsets.
 *
 * Plain C (C99 or C++), no dependency on HIP, torch or GSL.
 */
#ifndef JURASSIC_TYPES_H
#define JURASSIC_TYPES_H

/* ------------------------------------------------------------------------
   Dimensions (reference src/jurassic.h:135-196).  They size the caller's
   structs, so they must not change.
   ------------------------------------------------------------------------ */
#define LEN 5000
#define NWMAX 5
#define NDMAX 130
#define NGMAX 25
#define NLOS 10000
#define NPMAX 1000
#define NRMAX 1000
#define NSHAPE 10000
#define NFOV 50
#define TBLNPMAX 50
#define TBLNSMAX 1201
#define TBLNTMAX 30
#define TBLNUMAX 320
#define SCAMOD 30
#define NLMAX 10
#define NTHETA 181
#define NRAD 170
#define REFMAX 5000

/* Retrieval vector sizes and state-vector quantity indices (reference
   src/jurassic.h:139-221).  IDXK/IDXNN/IDXRR/IDXSS need `ctl` in scope. */
#define NQMAX (2 + NGMAX + NWMAX)
#define MMAX (NRMAX * NDMAX)
#define NMAX (NQMAX * NPMAX)
#define IDXP 0
#define IDXT 1
#define IDXQ(ig) (2 + (ig))
#define IDXK(iw) (2 + ctl->ng + (iw))
#define IDXNN (2 + ctl->ng + ctl->nw + 1)
#define IDXRR (2 + ctl->ng + ctl->nw + 2)
#define IDXSS (2 + ctl->ng + ctl->nw + 3)

/* ------------------------------------------------------------------------
   Physical constants (reference src/jurassic.h:113-129 and the GSL 1.15
   constants the reference takes from gsl_const_mksa.h / gsl_const_num.h).
   ------------------------------------------------------------------------ */
#define C1 1.19104259e-8          /* 2 h c^2 [W/(m^2 sr cm^-4)] */
#define C2 1.43877506             /* h c / k [K/cm^-1] */
#define P0 1013.25                /* standard pressure [hPa] */
#define RE 6367.421               /* mean Earth radius [km] */
#define ME 5.976e24               /* Earth mass [kg] */
#define TSUN 5780.                /* solar brightness temperature [K] */
#define JUR_KBOLTZ 1.3806504e-23  /* Boltzmann constant, GSL 1.15 value */
#define JUR_AVOGADRO 6.02214199e23
#define JUR_RGAS 8.314472         /* molar gas constant, GSL 1.15 value */

/* Work-queue states (reference src/workqueue.h:8-11). */
#define Queue_Inactive -1
#define Queue_Prepare 2
#define Queue_Execute 8
#define Queue_Collect 32

#ifdef __cplusplus
extern "C" {
#endif

/* Work-queue item and queue (reference src/workqueue.h:13-25). */
typedef struct {
  void *input;
  void *result;
  int ir;
} queue_item_t;

typedef struct {
  queue_item_t *items;
  int capacity;
  int begin;
  int end;
  int state;
} queue_t;

/* Forward-model control parameters (reference src/jurassic.h:228-348). */
typedef struct {
  int ng;                       /* number of emitters */
  char emitter[NGMAX][LEN];     /* emitter names */
  int nd;                       /* number of channels */
  int nw;                       /* number of spectral windows */
  double nu[NDMAX];             /* channel centroid wavenumbers [cm^-1] */
  int window[NDMAX];            /* window index per channel */
  char tblbase[LEN];            /* basename of table and filter files */
  double hydz;                  /* hydrostatic reference height [km] (<0: off) */
  int ctm_co2;                  /* continua switches */
  int ctm_h2o;
  int ctm_n2;
  int ctm_o2;
  int sca_n;                    /* number of scattering models */
  int sca_mult;                 /* scattering recursion depth */
  char sca_ext[LEN];            /* leaf extinction: "beta_a" or "beta_e" */
  int ip;                       /* atmosphere interpolation method */
  double cz;
  double cx;
  int refrac;                   /* refraction on/off */
  double rayds;                 /* maximum raytrace step [km] */
  double raydz;                 /* vertical raytrace step [km] */
  double transs;                /* transition-layer sampling step [km] */
  char fov[LEN];                /* field-of-view file ("-": none) */
  double retp_zmin;
  double retp_zmax;
  double rett_zmin;
  double rett_zmax;
  double retq_zmin[NGMAX];
  double retq_zmax[NGMAX];
  double retk_zmin[NWMAX];
  double retk_zmax[NWMAX];
  double retnn_zmin;
  double retnn_zmax;
  int retnn;
  int retrr;
  int retss;
  int write_bbt;                /* output brightness temperature */
  int write_matrix;
  queue_t queue;                /* secondary-ray work queue */
} ctl_t;

/* Line-of-sight data (reference src/jurassic.h:352-393). */
typedef struct {
  int np;
  double z[NLOS];
  double lon[NLOS];
  double lat[NLOS];
  double p[NLOS];
  double t[NLOS];
  double q[NLOS][NGMAX];
  double k[NLOS][NWMAX];
  int aeroi[NLOS];
  double aerofac[NLOS];
  double tsurf;
  double ds[NLOS];
  double u[NLOS][NGMAX];
} los_t;

/* Observation geometry and radiances (reference src/jurassic.h:397-438). */
typedef struct {
  int nr;
  double time[NRMAX];
  double obsz[NRMAX];
  double obslon[NRMAX];
  double obslat[NRMAX];
  double vpz[NRMAX];
  double vplon[NRMAX];
  double vplat[NRMAX];
  double tpz[NRMAX];
  double tplon[NRMAX];
  double tplat[NRMAX];
  double tau[NDMAX][NRMAX];
  double rad[NDMAX][NRMAX];
} obs_t;

/* Aerosol / cloud models and layer optics (reference src/jurassic.h:475-535). */
typedef struct {
  int nm;
  double top_mod[SCAMOD];
  double bottom_mod[SCAMOD];
  double trans_mod[SCAMOD];
  char type[SCAMOD][LEN];
  char filepath[SCAMOD][LEN];
  double nn[SCAMOD];
  double rr[SCAMOD];
  double ss[SCAMOD];
  int nl;
  int nmod[NLMAX];
  double top[NLMAX];
  double bottom[NLMAX];
  double trans[NLMAX];
  double beta_e[NLMAX][NDMAX];
  double beta_s[NLMAX][NDMAX];
  double beta_a[NLMAX][NDMAX];
  double p[NLMAX][NDMAX][NTHETA];
} aero_t;

/* Atmospheric profiles (reference src/jurassic.h:538-570). */
typedef struct {
  int np;
  double time[NPMAX];
  double z[NPMAX];
  double lon[NPMAX];
  double lat[NPMAX];
  double p[NPMAX];
  double t[NPMAX];
  double q[NGMAX][NPMAX];
  double k[NWMAX][NPMAX];
  int init;
} atm_t;

/* Retrieval control parameters (reference src/jurassic.h:573-645). */
typedef struct {
  char dir[LEN];             /* working directory */
  int kernel_recomp;         /* recompute the kernel every n iterations */
  int conv_itmax;            /* maximum number of iterations */
  double conv_dmin;          /* minimum normalised step size in state space */
  double resmax;             /* radiance residual threshold [%] (-999: off) */
  int err_ana;               /* error analysis (0 = no, 1 = yes) */
  double err_formod[NDMAX];  /* forward model error [%] */
  double err_noise[NDMAX];   /* noise error [W/(m^2 sr cm^-1)] */
  double err_press;          /* pressure error [%] */
  double err_press_cz;       /* vertical / horizontal correlation lengths [km] */
  double err_press_ch;
  double err_temp;           /* temperature error [K] */
  double err_temp_cz;
  double err_temp_ch;
  double err_nn;             /* particle concentration error [cm^-3] */
  double err_rr;             /* particle radius error */
  double err_ss;             /* size distribution width error */
  double err_q[NGMAX];       /* volume mixing ratio error [%] */
  double err_q_cz[NGMAX];
  double err_q_ch[NGMAX];
  double err_k[NWMAX];       /* extinction error [1/km] */
  double err_k_cz[NWMAX];
  double err_k_ch[NWMAX];
} ret_t;

#ifdef __cplusplus
}
#endif

/* Layout pins: sizes and offsets of the reference structs on x86-64 / LP64. */
#include <stddef.h>
#ifdef __cplusplus
#define JUR_STATIC_ASSERT static_assert
#else
#define JUR_STATIC_ASSERT _Static_assert
#endif
JUR_STATIC_ASSERT(sizeof(queue_t) == 24, "queue_t layout");
JUR_STATIC_ASSERT(sizeof(ctl_t) == 142240, "ctl_t layout");
JUR_STATIC_ASSERT(offsetof(ctl_t, queue) == 142216, "ctl_t.queue offset");
JUR_STATIC_ASSERT(sizeof(atm_t) == 288016, "atm_t layout");
JUR_STATIC_ASSERT(sizeof(obs_t) == 2160008, "obs_t layout");
JUR_STATIC_ASSERT(offsetof(obs_t, rad) == 1120008, "obs_t.rad offset");
JUR_STATIC_ASSERT(sizeof(los_t) == 5000016, "los_t layout");
JUR_STATIC_ASSERT(sizeof(aero_t) == 2215336, "aero_t layout");
JUR_STATIC_ASSERT(offsetof(aero_t, p) == 332936, "aero_t.p offset");
JUR_STATIC_ASSERT(sizeof(ret_t) == 7904, "ret_t layout");
JUR_STATIC_ASSERT(offsetof(ret_t, err_formod) == 5032, "ret_t.err_formod offset");

#endif
