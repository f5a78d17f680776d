// jur_scatter.h -- scattering pipeline and Mie optics (host side).
#pragma once
#include "jur_host.h"

// Upload the layer phase functions of aero (used by the scattering source).
void scatter_set_phase(const ctl_t* ctl, const aero_t* aero, hipStream_t s);
// Forward model with the scattering source at depth `scattering` >= 1.
void scatter_forward(const ctl_t* ctl, const AtmSetDev& atm, const RaySetDev& rays, int scattering,
                     const OutDev& out, hipStream_t s);
// Training pairs (raw features [n][24], thermal J [n]) of the primaries'
// scattering points from the last call with option "collect"; copies at
// most cap rows and returns the number available.
long scatter_collected(float* feat, float* jth, long cap);
