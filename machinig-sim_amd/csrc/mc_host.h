// mc_host.h — library-internal host services shared by the source files of libmcsim.so:
// error reporting, the per-device context (stream, events, call lock) and timing.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/mcsim.h"

namespace mch {
int fail(int code, const char* fmt, ...);           // sets mc_last_error(), returns code
int init();                                         // the implicit mc_init(0)
hipStream_t stream();                               // the library's stream (default priority)
std::mutex& lock();                                 // serialises calls on the device
int cus();                                          // compute units of the device
void set_last_ms(double ms);                        // mc_last_kernel_time()

// a stock's device arrays, for kernels of other source files that write the stock
struct StockView {
  int nx, ny;
  double hx, hy, zTop, zBottom, band;
  float *H, *err, *wallD, *markS, *markF;
  int32_t *lineCut, *lineWall, *rect;   // rect: the dirty rectangle {i0, j0, i1, j1}
};
// fills v; the caller is about to change the heights, so the block maxima are marked stale
int stock_view(mc_stock* s, StockView* v);
// the stock's strip maxima for the sub-step kernel: one float per strip of MCH_STRIP_W x
// MCH_STRIP_H nodes, snx strips per row of strips (allocated on first use; their values
// are the kernel's to set)
constexpr int MCH_STRIP_W = 32, MCH_STRIP_H = 2;
int stock_strips(mc_stock* s, float** smax, int* snx, int* sny);
}  // namespace mch

#define MCH_HIPCHK(call)                                                                    \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return mch::fail(MC_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)
