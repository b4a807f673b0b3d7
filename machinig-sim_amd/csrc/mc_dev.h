// mc_dev.h — device-side data layouts of the chatter search (shared by the host glue in
// api.cpp and the kernels in chatter.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mcsim.h"

namespace mcd {

constexpr int HIST = 2048;   // chatter-index history kept per group (the index reads <= 1416)
constexpr int TS = 32;       // once-per-tooth samples kept (chatterIndex reads the last 24, src/cutting.js:299-306)

// Per-cell constants: CuttingEngine.configure (src/cutting.js:39-82) evaluated on the
// host for the cell's speed and modes.  Everything that depends on the trial depth
// (fluteLen = max(tool.fluteLen, ap + 1), dz, slices) is derived per trial on the device.
struct CellDev {
  double R, tanB, toolFluteLen, omega, dt, dphi, pitch, binScale, fz, hMax;
  double Ktc, Krc, Kte, Kre, mc, Ci;  // mc and Ci already 0 when nonlinear is off
  double kx[MC_MAX_MODES], mx[MC_MAX_MODES], cx[MC_MAX_MODES];
  double ky[MC_MAX_MODES], my[MC_MAX_MODES], cy[MC_MAX_MODES];
  double dR[MC_MAX_FLUTES];
  int32_t N, ball, spt, spr, decim, nmx, nmy, revs;
  float Hb[MC_BINS];
};

enum TaskMode : int32_t {
  TASK_SINGLE = 0,   // one chatterAt trial at depth ap
  TASK_BISECT = 1,   // timeDomainLimit (src/stabtd.js:24-32)
  TASK_BRACKET = 2,  // bracketed search from [lo, hi], falls back to TASK_BISECT
};

struct TaskDev {
  int32_t cell, mode, iters, pad;
  double ap;          // TASK_SINGLE depth
  double aMin, aMax;  // bisection range
  double lo, hi, tol; // TASK_BRACKET hint and stopping width
  double mu, sig;     // prior on ln(alim) ~ N(mu, sig^2) guiding the speculation; sig <= 0: none
  // sim != 0 (TASK_SINGLE only): the engine as the simulator drives it -- feed direction
  // (ux, uy), tip height tipZ, the tool's own flute length, velocity (vx0, vy0) set from
  // outside -- instead of chatterAt's u = (1, 0), tipZ = -ap, fluteLen >= ap + 1, vx = 0.01
  double ux, uy, tipZ, vx0, vy0;
  int32_t sim, pad2;
};

struct ResultDev {
  double alim, lo, hi;
  double index, amp, once;  // last trial's chatter index
  int32_t open, trials, chatter, nz, fell_back, pad;
  int64_t steps, visits, engaged;
  int64_t t_start, t_end;   // device wall clock (100 MHz) around the task
};

struct LaunchDev {
  const CellDev* cells;
  const TaskDev* tasks;
  ResultDev* out;
  int32_t* queue;        // next task (atomic counter)
  float* ring;           // delay lines, ring_cap floats per group
  float2* hist;          // chatter-index history, HIST per group
  double* trace;         // TASK_SINGLE trace: 4 doubles per step (Fx, Fy, x, y) or null
  int64_t trace_steps;
  int64_t ring_cap;
  int32_t ntasks;
  int32_t nz_cap;        // slice capacity of the per-group LDS tables
  int32_t nv_cap;        // (tooth, slice) visit capacity of the per-group LDS tables
  int32_t grp_bytes;     // LDS bytes per group
  const TaskDev* wtasks; // search kernel: tasks run by a whole workgroup (first), their
  ResultDev* wout;       //   results, queue and count
  int32_t* wqueue;
  int32_t nwide;
  int32_t* cstate;       // per cell: 0 aMax check pending, 1 it chattered, 2 it was stable
  double2* geo;          // edge geometry table (sin, cos), geo_cap entries per wave
  unsigned char* gbin;   // engagement-map bin of each table entry
  int64_t geo_cap;
};

}  // namespace mcd
