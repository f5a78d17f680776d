// jur_surrogate.h -- the scattering-source surrogate as the scattering
// pipeline uses it: feature normalisation / training hull, inference on
// device features, and the feature kernel for scattering points.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jur_mlp.h"
#include "jur_scatter_dev.h"

// Normalisation (x - mu) / sd of the 24 features and of the target log J,
// and the training hull [lo, hi] of every normalised feature.
struct MlpNorm {
  float mu[MLP_IN], sd[MLP_IN], lo[MLP_IN], hi[MLP_IN];
  float ymu, ysd;
  int trained, pad;
};

const MlpNorm& mlp_norm();
const float* mlp_norm_dev();  // device copy (MlpNorm layout)
void mlp_infer_dev(const uint16_t* x, long n, float* y, hipStream_t s);
bool mlp_trained();

// The network predicts ln(J / B(T)) of the thermal source J, B(T) the
// Planck radiance at the point's temperature (feature 12).
//
// Features of every (scattering point, channel) pair, row k * nd + id:
//   0 ln p, 1 T, 2 z, 3 ln(af beta_e), 4 single-scattering albedo,
//   5 asymmetry g, 6 ln P(0), 7 ln P(180), 8 cos of the LOS zenith angle,
//   9 layer top - z, 10 z - layer bottom, 11 wavenumber,
//   12..16 ln B(T) at the point, the ground, z + 3 km, z - 3 km and the
//   coldest level above, 17 / 18 ln H2O column above / below,
//   19 ln(vertical aerosol optical depth of the layer),
//   20 / 21 ln(L_zenith / B(T)), ln(L_nadir / B(T)): the radiance summary,
//   extinction-only radiances of the point's vertical pencils,
//   22 / 23 ln of the layer's vertical optical depth above / below the
//   point (+1e-6).
struct FeatArgs {
  int npt, nd;
  const ScaPt* pts;
  DevAtm atm;
  int ig_h2o;
  const float* beta_e;  // [nl][nd]
  const float* beta_s;
  const float* asym;    // [nl][nd] asymmetry parameter
  const float* pfwd;    // [nl][nd] P(0)
  const float* pback;   // [nl][nd] P(180)
  double top[NLMAX], bottom[NLMAX];
  const double* nu;     // [nd]
  const MlpNorm* norm;  // device; null: raw features only
  float* raw;           // [npt nd][24] or null
  uint16_t* x;          // [npt nd][32] normalised BF16 or null
  int* out_hull;        // [npt]: set to 1 when a channel leaves the hull
  float margin;         // hull margin in normalised units
  float* lnb;           // [npt nd] ln B(T) at the point (feature 12), or null
  const float* ud;      // [nd][2 npt] radiances of the zenith (2k) and nadir (2k+1) pencils
};
void launch_features(const FeatArgs& a, hipStream_t s);
void launch_fallback_compact(int npt, const int* flag, const int* pos, const ScaPt* pts, int* sub,
                             ScaPt* pts_sub, hipStream_t s);
