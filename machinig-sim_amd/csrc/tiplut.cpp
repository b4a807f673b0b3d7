// tiplut.cpp — the tip-shape table of a cutter (tipLut, src/assembly.js:38-54), host side.
// The table gives the height of the cutter surface above the tip at radii 0, LUT_DR,
// 2 LUT_DR, ... R; Stock.stamp reads it for corner-radius and imported tool shapes.
#include <cmath>

#include "../../include/mcsim.h"
#include "mc_host.h"

extern "C" int mc_tip_lut(int32_t shape, double D, double cornerR, const double* edge_r,
                          int32_t edge_n, double edge_dz, float* out, int32_t cap, int32_t* n_out) {
  using mch::fail;
  if (!(D > 0) || !std::isfinite(D) || shape < MC_TIP_FLAT || shape > MC_TIP_EDGE)
    return fail(MC_ERR_ARG, "bad tip shape or diameter");
  if (shape == MC_TIP_EDGE && (!edge_r || edge_n < 1 || !std::isfinite(edge_dz)))
    return fail(MC_ERR_ARG, "edge profile needs edge_r[edge_n >= 1] and edge_dz");
  const double R = D / 2;
  const double nd = std::ceil(R / MC_LUT_DR) + 1;
  if (nd > 1e8) return fail(MC_ERR_LIMIT, "tip table of %.0f entries too large", nd);
  const int32_t n = (int32_t)nd;
  if (n_out) *n_out = n;
  if (!out) return MC_OK;
  if (cap < n) return fail(MC_ERR_ARG, "tip table needs %d entries, room for %d", n, cap);
  for (int32_t i = 0; i < n; i++) {
    const double r = std::fmin(R, i * MC_LUT_DR);
    double z = 0;
    if (shape == MC_TIP_EDGE) {
      // lowest height at which the envelope reaches radius r (linear between samples)
      z = edge_dz * (edge_n - 1);
      for (int32_t k = 0; k < edge_n; k++)
        if (edge_r[k] >= r - 1e-6) {
          if (k > 0) {
            const double span = edge_r[k] - edge_r[k - 1];
            z = (k - 1 + (r - edge_r[k - 1]) / (span > 1e-9 ? span : 1e-9)) * edge_dz;
          } else {
            z = 0;
          }
          break;
        }
    } else if (shape == MC_TIP_BALL) {
      const double u = R * R - r * r;
      z = R - std::sqrt(u > 0 ? u : 0.0);
    } else if (shape == MC_TIP_BULL && cornerR > 0) {
      const double rc = std::fmin(R, cornerR), r0 = R - rc;
      if (r > r0) {
        const double u = rc * rc - (r - r0) * (r - r0);
        z = rc - std::sqrt(u > 0 ? u : 0.0);
      }
    }
    out[i] = (float)z;
  }
  return MC_OK;
}
