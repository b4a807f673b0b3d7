// jsmath.cpp — V8's Math.* (the fdlibm ports of mc_math.h) on the host, for host code that
// must produce the reference's doubles: the path discretizer's arcs (atan2, acos, cos, sin;
// src/gcode.js:443-460) and callers checking the library's math.  No device work.
#include <cmath>

#include "../../include/mcsim.h"
#include "mc_host.h"
#include "mc_math.h"

extern "C" int mc_js_math(int32_t fn, const double* x, const double* y, double* out, int64_t n) {
  if (n < 0 || (n > 0 && (!x || !out))) return mch::fail(MC_ERR_ARG, "null arrays or negative count");
  const bool two = fn == MC_JS_ATAN2 || fn == MC_JS_POW || fn == MC_JS_HYPOT;
  if (two && n > 0 && !y) return mch::fail(MC_ERR_ARG, "this function takes two arguments");
  for (int64_t i = 0; i < n; i++) {
    const double a = x[i];
    switch (fn) {
      case MC_JS_SIN: out[i] = mcm::sin_js(a); break;
      case MC_JS_COS: out[i] = mcm::cos_js(a); break;
      case MC_JS_TAN: out[i] = mcm::tan_js(a); break;
      case MC_JS_ATAN: out[i] = mcm::atan_js(a); break;
      case MC_JS_ATAN2: out[i] = mcm::atan2_js(a, y[i]); break;
      case MC_JS_ACOS: out[i] = mcm::acos_js(a); break;
      case MC_JS_POW: out[i] = mcm::pow_js(a, y[i]); break;
      case MC_JS_HYPOT: out[i] = mcm::hypot_js(a, y[i]); break;
      default: return mch::fail(MC_ERR_ARG, "unknown function %d", fn);
    }
  }
  return MC_OK;
}
