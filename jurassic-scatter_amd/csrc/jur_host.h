// jur_host.h -- host-side internals of the GPU forward model library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "jur_kernels.h"
#include "jurassic_amd.h"

// Error exit with the reference's message format (jurassic.h:48-52).
#define JUR_FAIL(msg)                                                            \
  do {                                                                          \
    printf("\nError (%s, %s, l%d): %s\n\n", __FILE__, __func__, __LINE__, (msg)); \
    fflush(stdout);                                                             \
    exit(EXIT_FAILURE);                                                         \
  } while (0)

#define HIP_OK(cmd)                                                  \
  do {                                                               \
    hipError_t e_ = (cmd);                                           \
    if (e_ != hipSuccess) {                                          \
      char m_[512];                                                  \
      snprintf(m_, sizeof m_, "HIP error %s: %s", #cmd, hipGetErrorString(e_)); \
      JUR_FAIL(m_);                                                  \
    }                                                                \
  } while (0)

// Growable device buffer (memory is kept across calls; growth reallocates,
// after the device has finished the work that may still read the old
// memory).
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void* need(size_t n) {
    if (n > bytes) {
      if (p) {
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipFree(p));
      }
      size_t nb = n + n / 4 + 256;
      HIP_OK(hipMalloc(&p, nb));
      bytes = nb;
    }
    return p;
  }
  template <typename T>
  T* as() const { return (T*)p; }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// One emissivity table as read from disk (forwardmodel.c:610-739).
struct HostTable {
  int np = -1;
  std::vector<double> p;                         // [np]
  std::vector<int> nt;                           // [np]
  std::vector<std::vector<double>> t;            // [np][nt]
  std::vector<std::vector<int>> nu;              // [np][nt]
  std::vector<std::vector<std::vector<float>>> u, eps;  // [np][nt][nu]
};
bool read_table_bin(const char* filename, HostTable& tb);
bool read_table_ascii(const char* filename, HostTable& tb);
void read_shape_file(const char* filename, std::vector<double>& x, std::vector<double>& y);

// Continuum coefficients of one channel (continua.c), host FP64 -> FP32.
ChanConst make_chan_const(const ctl_t* ctl, int id);

// Engine: device context and the forward-model pipeline.
struct RaySetDev {
  int nray = 0;
  const double *time = nullptr, *obsz = nullptr, *obslon = nullptr, *obslat = nullptr;
  const double *vpz = nullptr, *vplon = nullptr, *vplat = nullptr;
  const int* atm_idx = nullptr;  // null: all rays use atmosphere 0
};
struct AtmSetDev {
  int natm = 0, nlev = 0;
  const double *z = nullptr, *p = nullptr, *t = nullptr, *q = nullptr, *k = nullptr;
  DevGeo geo{};  // IP 2 / 3: geographic positions of the points (shared by the copies)
};
// The geographic part of an atmosphere for ctl->ip (2: profiles, 3: points;
// 1: none), on the device; exits like the reference on unknown IP or
// profiles that cannot be identified.
DevGeo stage_geo(const ctl_t* ctl, const atm_t* atm, hipStream_t s);
struct OutDev {
  float* rad = nullptr;   // [nd][ld]: ray r at column r
  float* tau = nullptr;   // [nd][ld]
  int ld = 0;             // leading dimension of rad / tau / rad64 / tau64 (0: the ray count)
  double *tpz = nullptr, *tplon = nullptr, *tplat = nullptr;  // [nray]
  int* npts = nullptr;    // [nray] (optional)
  // FP64 radiance / transmittance [nd][ld] (optional): with option "fp64"
  // the path integration runs in FP64 and also writes these
  double *rad64 = nullptr, *tau64 = nullptr;
};

// Make sure tables / Planck table / channel constants for ctl are on the
// device (loaded on first use and cached per table set).
void engine_prepare(const ctl_t* ctl);
// Upload layer optics of aero (beta_e/s/a, phase functions) for ctl.
void engine_set_aero(const ctl_t* ctl, const aero_t* aero);
// Run the forward model for a ray set on the engine's stream.  scattering is
// the recursion depth (SCA_MULT) used for the primaries.
void engine_forward(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays,
                    int scattering, const OutDev& out, hipStream_t s);
hipStream_t engine_stream();
int engine_option_int(const char* key, int dflt);
size_t engine_chunk_bytes();
// FP64 clear-air calls whose rays are copies of the first nr rays (the
// Jacobian's perturbed states): each ray is integrated from its first LOS
// point that differs from its base ray (0: off).
void engine_set_prefix(int nr);
double engine_stat(const char* key);
void engine_set_stat(const char* key, double v);

// Buffers one ray set is traced into (see engine_trace).
struct TraceBufs {
  DBuf *geo, *geo2, *rec, *npts, *tsurf, *tsurf64, *geo_final;
  int cap;
};
// Trace rays (FP64) into tb; grows the per-ray point capacity until every ray
// fits and returns it.  sca_on inserts aerosol-layer points.
int engine_trace(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, int sca_on,
                 TraceBufs& tb, double* tpz, double* tplon, double* tplat, hipStream_t s);
// Enqueue one trace at a fixed per-ray capacity without waiting for it:
// flags are raised in *status (JUR_ST_CAP when a ray needs more points), rays
// r with r + nvalid_off >= *nvalid (nvalid optional) are empty.
void engine_trace_async(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, int sca_on, TraceBufs& tb,
                        int cap, double* tpz, double* tplon, double* tplat, int* status, const int* nvalid,
                        int nvalid_off, hipStream_t s);
// Every address, size and setting a recorded forward call depends on
// (inputs, outputs, tables, layer geometry and optics, work buffers,
// options and autotuner choices).
std::string engine_call_key(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, const OutDev& out,
                            int scattering, hipStream_t s);
// planned: device-resident calls (counting sort, no autotuning runs);
// capturing: the calls are being recorded into a graph (no event timing).
void engine_set_mode(bool planned, bool capturing);
bool engine_capturing();
// Per-call accounting: path-kernel events and the LOS-point counter
// (end_call waits for the stream).
void engine_begin_call(hipStream_t s);
void engine_end_call(hipStream_t s);
// Sort by point count and integrate along the traced rays.
void engine_integrate(const ctl_t* ctl, int nray, int cap, const float* rec, const int* npts,
                      const float* tsurf, int scattering, const float* J, const int* jbase,
                      float* rad, float* tau, int out_ld, int out_off, hipStream_t s,
                      double* rad64 = nullptr, double* tau64 = nullptr, long nJ = 0);
// True when the tables, channel constants and layer optics of ctl / aero
// are already on the device (nothing to load or upload).
bool engine_resident(const ctl_t* ctl, const aero_t* aero);
// Clear-air / extinction forward call (scattering depth 0, FP32 path) whose
// work buffers are carved from the device region [base, base + bytes):
// allocates nothing.  Returns the LOS capacity used, or -1 when the rays do
// not fit the region.
int engine_forward_noalloc(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, const OutDev& out,
                           void* base, size_t bytes, hipStream_t s);
// Device address of the no-allocation work region on the current device
// (looked up once per device); required: stop with an error if it cannot be
// found, else return NULL then.
void* noalloc_scratch(bool required);
// Layer-edge altitudes of the aerosol layers (for the solar-ray search).
struct AeroGeom {
  const double* alti;
  int nalti;
  double altimax, altimin;
};
AeroGeom engine_aero_geom();
int engine_cap();
DevTables engine_tables();
const double* engine_nu_dev();
int engine_find_emitter(const ctl_t* ctl, const char* name);
const float* engine_beta_e();  // [nl][nd] layer extinction (FP32)
const float* engine_beta_s();
void engine_layer_edges(double* top, double* bottom);  // [NLMAX]
double engine_option_double(const char* key, double dflt);

// Mie optics on the device (scatter.c:330-508).
void engine_get_opt_prop(ctl_t* ctl, aero_t* aero);
// Hydrostatic equilibrium for a host atmosphere (atmosphere.c:163-243), on the device.
void engine_hydrostatic(ctl_t* ctl, atm_t* atm);
// formod's FP64 epilogue (FOV within runs of seg rays, brightness temperature)
// on outputs O of nray rays; returns device [nd][nray] radiance/transmittance.
void formod_epilogue(const ctl_t* ctl, int nray, int seg, const double* time, const OutDev& O, double** rad,
                     double** tau, hipStream_t s);
