// noalloc.hip -- the batched clear-air / extinction forward call for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc): its work buffers (LOS
// geometry and point records, point counts, the sort, the status word, and
// any optional input or output the caller leaves out) are carved from one
// fixed device array per GPU instead of the library's growable buffers.
#include <cstring>

#include "jur_host.h"

namespace {

// The work region.  Each GPU the library runs on holds its own copy until
// the process ends.  Host code never passes the array's name to a kernel or
// a copy: that name is a host-side stand-in, not a device address; the
// per-device address comes from hipGetSymbolAddress below.
static __device__ __align__(256) unsigned char scratch[JUR_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
void* scratch_dev[kMaxDevices] = {nullptr};

}  // namespace

// Device address of the work region on the current device, looked up once
// per device (with that device current).  required: stop with an error when
// it cannot be found (else return NULL).
void* noalloc_scratch(bool required) {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    if (required) JUR_FAIL("Device index out of range for the no-allocation work region!");
    return nullptr;
  }
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !p) {
      if (!required) return nullptr;
      char m[256];
      snprintf(m, sizeof m, "Cannot locate the no-allocation work region on device %d: %s", dev,
               hipGetErrorString(err));
      JUR_FAIL(m);
    }
    scratch_dev[dev] = p;
  }
  return scratch_dev[dev];
}

extern "C" int jur_formod_batch_device_noalloc(ctl_t* ctl, const jur_batch_t* b, const aero_t* aero,
                                                int scattering, void* stream) {
  if (!b || b->nray < 0 || b->nlev < 2 || b->natm < 1) JUR_FAIL("Invalid batch description!");
  if (!b->z || !b->p || !b->t || (!b->q && ctl->ng > 0) || !b->obsz || !b->vpz || !b->rad || !b->tau)
    JUR_FAIL("Missing batch array!");
  if (b->nlev > NPMAX) JUR_FAIL("Too many data points!");
  if (b->out_ld != 0 && b->out_ld < b->nray) JUR_FAIL("Invalid batch description!");
  if (ctl->ip != 1) JUR_FAIL("The batched entry point takes vertical profiles (IP=1)!");
  if (ctl->sca_n > 0 && scattering > 0)
    JUR_FAIL("The no-allocation entry point computes clear air and extinction only (scattering depth 0)!");
  if (engine_option_int("fp64", 0))
    JUR_FAIL("The no-allocation entry point runs the FP32 path only (option fp64 is set)!");
  if (!engine_resident(ctl, ctl->sca_n > 0 ? aero : nullptr))
    JUR_FAIL("Tables or aerosol layers not resident: make one jur_formod_batch(_device) call with this "
             "ctl and aero first!");
  void* base = noalloc_scratch(true);  // before any stream work: the first call looks the address up
  hipStream_t s = stream ? (hipStream_t)stream : engine_stream();
  char* p = (char*)base;
  size_t left = JUR_NOALLOC_SCRATCH_BYTES;
  auto take = [&p, &left](size_t n) -> void* {
    const size_t a = (n + 255) & ~(size_t)255;
    if (a > left) JUR_FAIL("Batch too large for the no-allocation work region (JUR_NOALLOC_SCRATCH_BYTES)!");
    void* r = p;
    p += a;
    left -= a;
    return r;
  };
  AtmSetDev A;
  A.natm = b->natm;
  A.nlev = b->nlev;
  A.z = b->z;
  A.p = b->p;
  A.t = b->t;
  A.q = b->q;
  A.k = b->k;
  if (!A.k && ctl->nw > 0) {
    const size_t n = (size_t)b->natm * ctl->nw * b->nlev;
    double* kz = (double*)take(n * sizeof(double));
    launch_fill_f64(kz, n, 0.0, s);
    A.k = kz;
  }
  RaySetDev R;
  R.nray = b->nray;
  R.time = b->time;
  R.obsz = b->obsz;
  R.obslon = b->obslon;
  R.obslat = b->obslat;
  R.vpz = b->vpz;
  R.vplon = b->vplon;
  R.vplat = b->vplat;
  R.atm_idx = b->atm_of_ray;
  if (!R.atm_idx && b->natm == b->nray && b->natm > 1) {
    int* d = (int*)take((size_t)b->nray * sizeof(int));
    launch_iota(d, b->nray, s);
    R.atm_idx = d;
  }
  OutDev O;
  O.rad = b->rad;
  O.tau = b->tau;
  O.ld = b->out_ld;
  O.tpz = b->tpz ? b->tpz : (double*)take((size_t)b->nray * sizeof(double));
  O.tplon = b->tplon ? b->tplon : (double*)take((size_t)b->nray * sizeof(double));
  O.tplat = b->tplat ? b->tplat : (double*)take((size_t)b->nray * sizeof(double));
  O.npts = b->npts;
  if (engine_forward_noalloc(ctl, A, R, O, p, left, s) < 0)
    JUR_FAIL("Batch too large for the no-allocation work region (JUR_NOALLOC_SCRATCH_BYTES)!");
  return 0;
}
