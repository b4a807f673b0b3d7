// napi_addon.cc — the thin N-API binding of libmcsim.so for the JavaScript host modules
// in machinig-sim_amd/js (which keep the reference's exported signatures).
//
// Data crosses as ArrayBuffers laid out exactly like the C structs of include/mcsim.h
// (the JS side packs mc_cell / mc_stamp / mc_stamp_opt records with a DataView), so the addon only
// forwards pointers and maps a non-zero mc_status to a thrown Error carrying
// mc_last_error().  bisectBatchAsync runs the search on a libuv worker thread and
// returns a Promise, for the stability-sweep worker protocol.
#define NAPI_VERSION 6
#include <node_api.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mcsim.h"

namespace {

#define NAPI_OK(call)                                              \
  do {                                                             \
    if ((call) != napi_ok) {                                       \
      napi_throw_error(env, nullptr, "N-API call failed: " #call); \
      return nullptr;                                              \
    }                                                              \
  } while (0)

napi_value throw_mc(napi_env env, int st) {
  std::string m = "libmcsim error " + std::to_string(st) + ": " + mc_last_error();
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < want) {
    napi_throw_type_error(env, nullptr, "missing arguments");
    return false;
  }
  return true;
}

// bytes of an ArrayBuffer or a TypedArray
bool buffer_of(napi_env env, napi_value v, void** data, size_t* len) {
  bool is_ab = false, is_ta = false;
  napi_is_arraybuffer(env, v, &is_ab);
  if (is_ab) return napi_get_arraybuffer_info(env, v, data, len) == napi_ok;
  napi_is_typedarray(env, v, &is_ta);
  if (is_ta) {
    napi_typedarray_type t;
    size_t n, off;
    napi_value ab;
    if (napi_get_typedarray_info(env, v, &t, &n, data, &ab, &off) != napi_ok) return false;
    size_t es = (t == napi_float64_array || t == napi_bigint64_array || t == napi_biguint64_array) ? 8
                : (t == napi_float32_array || t == napi_int32_array || t == napi_uint32_array) ? 4
                : (t == napi_int16_array || t == napi_uint16_array) ? 2 : 1;
    *len = n * es;
    return true;
  }
  napi_throw_type_error(env, nullptr, "expected an ArrayBuffer or a typed array");
  return false;
}

bool napi_is_exception_pending_bool(napi_env env) {
  bool p = false;
  napi_is_exception_pending(env, &p);
  return p;
}

double num(napi_env env, napi_value v) {
  double d = 0;
  napi_get_value_double(env, v, &d);
  return d;
}

napi_value new_ab(napi_env env, size_t bytes, void** data) {
  napi_value ab;
  if (napi_create_arraybuffer(env, bytes, data, &ab) != napi_ok) return nullptr;
  return ab;
}

napi_value Init(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  int dev = 0;
  size_t argc = 1;
  napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
  if (argc >= 1) dev = (int)num(env, argv[0]);
  int st = mc_init(dev);
  if (st) return throw_mc(env, st);
  return nullptr;
}

napi_value DeviceInfo(napi_env env, napi_callback_info) {
  char buf[256];
  int st = mc_device_info(buf, sizeof buf);
  if (st) return throw_mc(env, st);
  napi_value s;
  NAPI_OK(napi_create_string_utf8(env, buf, NAPI_AUTO_LENGTH, &s));
  return s;
}

napi_value LastKernelMs(napi_env env, napi_callback_info) {
  double ms = 0;
  int st = mc_last_kernel_time(&ms);
  if (st) return throw_mc(env, st);
  napi_value v;
  NAPI_OK(napi_create_double(env, ms, &v));
  return v;
}

// chatterBatch(cells: ArrayBuffer of n mc_cell, ap: Float64Array(n)) -> ArrayBuffer of mc_trial
napi_value ChatterBatch(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  void *cells, *ap;
  size_t cb, ab;
  if (!buffer_of(env, argv[0], &cells, &cb) || !buffer_of(env, argv[1], &ap, &ab)) return nullptr;
  const int n = (int)(cb / sizeof(mc_cell));
  if (ab / 8 < (size_t)n) { napi_throw_range_error(env, nullptr, "ap shorter than cells"); return nullptr; }
  void* out;
  napi_value res = new_ab(env, (size_t)n * sizeof(mc_trial) + 1, &out);
  int st = mc_chatter_batch((const mc_cell*)cells, (const double*)ap, n, (mc_trial*)out);
  if (st) return throw_mc(env, st);
  return res;
}

// bisectBatch(cells, aMin, aMax, iters) -> ArrayBuffer of mc_limit
napi_value BisectBatch(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return nullptr;
  void* cells;
  size_t cb;
  if (!buffer_of(env, argv[0], &cells, &cb)) return nullptr;
  const int n = (int)(cb / sizeof(mc_cell));
  mc_bisect_opts o{num(env, argv[1]), num(env, argv[2]), (int32_t)num(env, argv[3]), 0};
  void* out;
  napi_value res = new_ab(env, (size_t)n * sizeof(mc_limit) + 1, &out);
  int st = mc_bisect_batch((const mc_cell*)cells, n, &o, (mc_limit*)out);
  if (st) return throw_mc(env, st);
  return res;
}

struct AsyncBisect {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  std::vector<mc_cell> cells;
  std::vector<mc_limit> out;
  mc_bisect_opts o{};
  int st = 0;
  std::string err;
};

void bisect_exec(napi_env, void* data) {
  AsyncBisect* a = (AsyncBisect*)data;
  a->st = mc_bisect_batch(a->cells.data(), (int)a->cells.size(), &a->o, a->out.data());
  if (a->st) a->err = mc_last_error();
}

void bisect_done(napi_env env, napi_status, void* data) {
  AsyncBisect* a = (AsyncBisect*)data;
  if (a->st) {
    napi_value msg, e;
    std::string m = "libmcsim error " + std::to_string(a->st) + ": " + a->err;
    napi_create_string_utf8(env, m.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &e);
    napi_reject_deferred(env, a->deferred, e);
  } else {
    void* p;
    napi_value ab = new_ab(env, a->out.size() * sizeof(mc_limit) + 1, &p);
    std::memcpy(p, a->out.data(), a->out.size() * sizeof(mc_limit));
    napi_resolve_deferred(env, a->deferred, ab);
  }
  napi_delete_async_work(env, a->work);
  delete a;
}

// bisectBatchAsync(cells, aMin, aMax, iters) -> Promise<ArrayBuffer of mc_limit>
napi_value BisectBatchAsync(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return nullptr;
  void* cells;
  size_t cb;
  if (!buffer_of(env, argv[0], &cells, &cb)) return nullptr;
  AsyncBisect* a = new AsyncBisect();
  const size_t n = cb / sizeof(mc_cell);
  a->cells.assign((const mc_cell*)cells, (const mc_cell*)cells + n);
  a->out.resize(n);
  a->o = mc_bisect_opts{num(env, argv[1]), num(env, argv[2]), (int32_t)num(env, argv[3]), 0};
  napi_value promise, name;
  NAPI_OK(napi_create_promise(env, &a->deferred, &promise));
  NAPI_OK(napi_create_string_utf8(env, "mcsim.bisect", NAPI_AUTO_LENGTH, &name));
  NAPI_OK(napi_create_async_work(env, nullptr, name, bisect_exec, bisect_done, a, &a->work));
  NAPI_OK(napi_queue_async_work(env, a->work));
  return promise;
}

// bracketBatch(cells, aMin, aMax, iters, lo: Float64Array, hi: Float64Array, tol)
//   -> ArrayBuffer: n mc_limit followed by n int32 fell_back flags
napi_value BracketBatch(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return nullptr;
  void *cells, *lo, *hi;
  size_t cb, lb, hb;
  if (!buffer_of(env, argv[0], &cells, &cb) || !buffer_of(env, argv[4], &lo, &lb) ||
      !buffer_of(env, argv[5], &hi, &hb))
    return nullptr;
  const int n = (int)(cb / sizeof(mc_cell));
  if (lb / 8 < (size_t)n || hb / 8 < (size_t)n) { napi_throw_range_error(env, nullptr, "hint arrays too short"); return nullptr; }
  mc_bisect_opts o{num(env, argv[1]), num(env, argv[2]), (int32_t)num(env, argv[3]), 0};
  void* out;
  napi_value res = new_ab(env, (size_t)n * (sizeof(mc_limit) + 4) + 1, &out);
  int st = mc_bisect_bracket_batch((const mc_cell*)cells, n, &o, (const double*)lo, (const double*)hi,
                                   num(env, argv[6]), (mc_limit*)out,
                                   (int32_t*)((char*)out + (size_t)n * sizeof(mc_limit)));
  if (st) return throw_mc(env, st);
  return res;
}

// engineTrace(cell: ArrayBuffer(1 mc_cell), ap, nsteps) -> Float64Array(4*steps): Fx, Fy, x, y blocks
napi_value EngineTrace(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  void* cell;
  size_t cb;
  if (!buffer_of(env, argv[0], &cell, &cb)) return nullptr;
  if (cb < sizeof(mc_cell)) { napi_throw_range_error(env, nullptr, "cell buffer too small"); return nullptr; }
  const int64_t ns = (int64_t)num(env, argv[2]);
  if (ns < 0) { napi_throw_range_error(env, nullptr, "negative step count"); return nullptr; }
  std::vector<double> buf((size_t)(ns > 0 ? ns : 1) * 4);
  int64_t got = 0;
  int st = mc_engine_trace((const mc_cell*)cell, num(env, argv[1]), ns, buf.data(), buf.data() + ns,
                           buf.data() + 2 * ns, buf.data() + 3 * ns, &got);
  if (st) return throw_mc(env, st);
  void* p;
  napi_value ab = new_ab(env, (size_t)got * 4 * 8 + 1, &p);
  for (int k = 0; k < 4; k++) std::memcpy((double*)p + k * got, buf.data() + k * ns, (size_t)got * 8);
  napi_value ta;
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, (size_t)got * 4, ab, 0, &ta));
  return ta;
}

// ---- stock: an External wrapping mc_stock*, freed by stockDestroy or the GC
void stock_finalize(napi_env, void* data, void*) {
  mc_stock** h = (mc_stock**)data;
  if (*h) mc_stock_destroy(*h);
  delete h;
}

mc_stock* stock_of(napi_env env, napi_value v) {
  mc_stock** h = nullptr;
  if (napi_get_value_external(env, v, (void**)&h) != napi_ok || !h || !*h) {
    napi_throw_type_error(env, nullptr, "not a live stock handle");
    return nullptr;
  }
  return *h;
}

napi_value StockCreate(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return nullptr;
  mc_stock* s = nullptr;
  int st = mc_stock_create(num(env, argv[0]), num(env, argv[1]), num(env, argv[2]), num(env, argv[3]), &s);
  if (st) return throw_mc(env, st);
  mc_stock** h = new mc_stock*(s);
  napi_value ext;
  NAPI_OK(napi_create_external(env, h, stock_finalize, nullptr, &ext));
  return ext;
}

napi_value StockDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  mc_stock** h = nullptr;
  if (napi_get_value_external(env, argv[0], (void**)&h) == napi_ok && h && *h) {
    mc_stock_destroy(*h);
    *h = nullptr;
  }
  return nullptr;
}

napi_value StockDims(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  mc_stock* s = stock_of(env, argv[0]);
  if (!s) return nullptr;
  int32_t nx, ny;
  double hx, hy;
  int st = mc_stock_dims(s, &nx, &ny, &hx, &hy);
  if (st) return throw_mc(env, st);
  napi_value o, v;
  NAPI_OK(napi_create_object(env, &o));
  napi_create_int32(env, nx, &v); napi_set_named_property(env, o, "nx", v);
  napi_create_int32(env, ny, &v); napi_set_named_property(env, o, "ny", v);
  napi_create_double(env, hx, &v); napi_set_named_property(env, o, "hx", v);
  napi_create_double(env, hy, &v); napi_set_named_property(env, o, "hy", v);
  return o;
}

napi_value StockSetWallBand(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  mc_stock* s = stock_of(env, argv[0]);
  if (!s) return nullptr;
  int st = mc_stock_set_wall_band(s, num(env, argv[1]));
  if (st) return throw_mc(env, st);
  return nullptr;
}

// stockRead(handle, H, err, wallD, lineCut, lineWall): fills the given typed arrays (any may be null)
napi_value StockRead(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return nullptr;
  mc_stock* s = stock_of(env, argv[0]);
  if (!s) return nullptr;
  void* p[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int32_t nx, ny;
  mc_stock_dims(s, &nx, &ny, nullptr, nullptr);
  for (int k = 0; k < 5; k++) {
    napi_valuetype t;
    napi_typeof(env, argv[1 + k], &t);
    if (t == napi_null || t == napi_undefined) continue;
    size_t len;
    if (!buffer_of(env, argv[1 + k], &p[k], &len)) return nullptr;
    if (len < (size_t)nx * ny * 4) { napi_throw_range_error(env, nullptr, "array smaller than the grid"); return nullptr; }
  }
  int st = MC_OK;
  if (p[0]) st = mc_stock_read_h(s, (float*)p[0]);
  if (!st && (p[1] || p[2] || p[3] || p[4]))
    st = mc_stock_read_side(s, (float*)p[1], (float*)p[2], (int32_t*)p[3], (int32_t*)p[4]);
  if (st) return throw_mc(env, st);
  return nullptr;
}

napi_value StockWriteH(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  mc_stock* s = stock_of(env, argv[0]);
  if (!s) return nullptr;
  void* p;
  size_t len;
  if (!buffer_of(env, argv[1], &p, &len)) return nullptr;
  int32_t nx, ny;
  mc_stock_dims(s, &nx, &ny, nullptr, nullptr);
  if (len < (size_t)nx * ny * 4) { napi_throw_range_error(env, nullptr, "array smaller than the grid"); return nullptr; }
  int st = mc_stock_write_h(s, (const float*)p);
  if (st) return throw_mc(env, st);
  return nullptr;
}

bool is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_null || t == napi_undefined;
}

// stampBatch(handle, stamps: ArrayBuffer of n mc_stamp[, opts: ArrayBuffer of n mc_stamp_opt,
//            luts: Float32Array, lutOff: Int32Array, lutLen: Int32Array]) -> Float64Array(n)
napi_value StampBatch(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  size_t argc = 6;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < 2) {
    napi_throw_type_error(env, nullptr, "missing arguments");
    return nullptr;
  }
  mc_stock* s = stock_of(env, argv[0]);
  if (!s) return nullptr;
  void* p;
  size_t len;
  if (!buffer_of(env, argv[1], &p, &len)) return nullptr;
  const int n = (int)(len / sizeof(mc_stamp));
  void* q[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t ql[4] = {0, 0, 0, 0};
  for (int k = 0; k < 4; k++)
    if (argc > 2 + (size_t)k && !is_nullish(env, argv[2 + k]) && !buffer_of(env, argv[2 + k], &q[k], &ql[k]))
      return nullptr;
  if (q[0] && ql[0] < (size_t)n * sizeof(mc_stamp_opt)) {
    napi_throw_range_error(env, nullptr, "one mc_stamp_opt per stamp expected");
    return nullptr;
  }
  const int nluts = (int)(ql[2] / 4);
  if (ql[3] / 4 != (size_t)nluts) {
    napi_throw_range_error(env, nullptr, "lutOff and lutLen differ in length");
    return nullptr;
  }
  // every profile must lie inside the luts array handed over
  for (int t = 0; t < nluts; t++) {
    const int64_t end = (int64_t)((int32_t*)q[2])[t] + ((int32_t*)q[3])[t];
    if (((int32_t*)q[2])[t] < 0 || end > (int64_t)(ql[1] / 4)) {
      napi_throw_range_error(env, nullptr, "profile outside the luts array");
      return nullptr;
    }
  }
  void* vol;
  napi_value ab = new_ab(env, (size_t)n * 8 + 1, &vol);
  int st = mc_stamp_batch_ex(s, (const mc_stamp*)p, (const mc_stamp_opt*)q[0], n, (const float*)q[1],
                             (const int32_t*)q[2], (const int32_t*)q[3], nluts, (double*)vol);
  if (st) return throw_mc(env, st);
  napi_value ta;
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, n, ab, 0, &ta));
  return ta;
}

// stockReadMarks(handle, markS, markF): fills the given Float32Arrays (either may be null)
napi_value StockReadMarks(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  mc_stock* s = stock_of(env, argv[0]);
  if (!s) return nullptr;
  int32_t nx, ny;
  mc_stock_dims(s, &nx, &ny, nullptr, nullptr);
  void* p[2] = {nullptr, nullptr};
  for (int k = 0; k < 2; k++) {
    if (is_nullish(env, argv[1 + k])) continue;
    size_t len;
    if (!buffer_of(env, argv[1 + k], &p[k], &len)) return nullptr;
    if (len < (size_t)nx * ny * 4) { napi_throw_range_error(env, nullptr, "array smaller than the grid"); return nullptr; }
  }
  int st = mc_stock_read_marks(s, (float*)p[0], (float*)p[1]);
  if (st) return throw_mc(env, st);
  return nullptr;
}

// stockDirty(handle) -> Int32Array [i0, j0, i1, j1]
napi_value StockDirty(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  mc_stock* s = stock_of(env, argv[0]);
  if (!s) return nullptr;
  void* out;
  napi_value ab = new_ab(env, 16, &out);
  int st = mc_stock_dirty(s, (int32_t*)out);
  if (st) return throw_mc(env, st);
  napi_value ta;
  NAPI_OK(napi_create_typedarray(env, napi_int32_array, 4, ab, 0, &ta));
  return ta;
}

napi_value StockClearDirty(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  mc_stock* s = stock_of(env, argv[0]);
  if (!s) return nullptr;
  int st = mc_stock_clear_dirty(s);
  if (st) return throw_mc(env, st);
  return nullptr;
}

// stockBlockMax(handle) -> Float32Array(bnx * bny)
napi_value StockBlockMax(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  mc_stock* s = stock_of(env, argv[0]);
  if (!s) return nullptr;
  int32_t nx, ny;
  mc_stock_dims(s, &nx, &ny, nullptr, nullptr);
  const size_t nb = (size_t)((nx + MC_STOCK_B - 1) / MC_STOCK_B) * ((ny + MC_STOCK_B - 1) / MC_STOCK_B);
  void* out;
  napi_value ab = new_ab(env, nb * 4, &out);
  int st = mc_stock_block_max(s, (float*)out);
  if (st) return throw_mc(env, st);
  napi_value ta;
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, nb, ab, 0, &ta));
  return ta;
}

// stockMaxInDisk(handle, q: Float64Array(4n) {cx, cy, r, above}) -> Float64Array(3n) {h, x, y}
napi_value StockMaxInDisk(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  mc_stock* s = stock_of(env, argv[0]);
  if (!s) return nullptr;
  void* p;
  size_t len;
  if (!buffer_of(env, argv[1], &p, &len)) return nullptr;
  const int n = (int)(len / 32);
  void* out;
  napi_value ab = new_ab(env, (size_t)n * 24 + 1, &out);
  int st = mc_stock_max_in_disk(s, (const double*)p, n, (double*)out);
  if (st) return throw_mc(env, st);
  napi_value ta;
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, (size_t)n * 3, ab, 0, &ta));
  return ta;
}

// stockEngagement(handle, q: Float64Array(7n) {cx, cy, ux, uy, R, tipZ, fluteLen},
//                 Hb: Float32Array(180n) | null) -> ArrayBuffer of n mc_engage_result
napi_value StockEngagement(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  mc_stock* s = stock_of(env, argv[0]);
  if (!s) return nullptr;
  void* p;
  size_t len;
  if (!buffer_of(env, argv[1], &p, &len)) return nullptr;
  const int n = (int)(len / sizeof(mc_engage_query));
  void* hb = nullptr;
  size_t hbl = 0;
  if (!is_nullish(env, argv[2])) {
    if (!buffer_of(env, argv[2], &hb, &hbl)) return nullptr;
    if (hbl < (size_t)n * MC_BINS * 4) { napi_throw_range_error(env, nullptr, "Hb needs 180 floats per query"); return nullptr; }
  }
  void* out;
  napi_value ab = new_ab(env, (size_t)n * sizeof(mc_engage_result) + 1, &out);
  int st = mc_stock_engagement(s, (const mc_engage_query*)p, n, (float*)hb, (mc_engage_result*)out);
  if (st) return throw_mc(env, st);
  return ab;
}

// tipLut(shape, D, cornerR, edgeR: Float64Array | null, edgeDz) -> Float32Array
napi_value TipLut(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return nullptr;
  int32_t shape = 0;
  napi_get_value_int32(env, argv[0], &shape);
  void* er = nullptr;
  size_t erl = 0;
  if (!is_nullish(env, argv[3]) && !buffer_of(env, argv[3], &er, &erl)) return nullptr;
  int32_t n = 0;
  int st = mc_tip_lut(shape, num(env, argv[1]), num(env, argv[2]), (const double*)er, (int32_t)(erl / 8),
                      num(env, argv[4]), nullptr, 0, &n);
  if (st) return throw_mc(env, st);
  void* out;
  napi_value ab = new_ab(env, (size_t)n * 4, &out);
  st = mc_tip_lut(shape, num(env, argv[1]), num(env, argv[2]), (const double*)er, (int32_t)(erl / 8),
                  num(env, argv[4]), (float*)out, n, &n);
  if (st) return throw_mc(env, st);
  napi_value ta;
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, (size_t)n, ab, 0, &ta));
  return ta;
}

// compositeLobes(nodes: mc_lobe_args records, step, hi) -> Float64Array(2n): n speeds, then
// their n depths
napi_value CompositeLobes(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  void* nd = nullptr;
  size_t nl = 0;
  if (!buffer_of(env, argv[0], &nd, &nl)) return nullptr;
  const int n = (int)(nl / sizeof(mc_lobe_args));
  const double step = num(env, argv[1]), hi = num(env, argv[2]);
  int cnt = 0;
  int st = mc_composite_lobes((const mc_lobe_args*)nd, n, step, hi, nullptr, nullptr, 0, &cnt);
  if (st) return throw_mc(env, st);
  void* out;
  napi_value ab = new_ab(env, (size_t)cnt * 16, &out);
  if (!ab) return nullptr;
  st = mc_composite_lobes((const mc_lobe_args*)nd, n, step, hi, (double*)out, (double*)out + cnt, cnt, &cnt);
  if (st) return throw_mc(env, st);
  napi_value ta;
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, (size_t)cnt * 2, ab, 0, &ta));
  return ta;
}

// stockSample(handle, xy: Float64Array(2n)) -> Float64Array(n)
napi_value StockSample(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  mc_stock* s = stock_of(env, argv[0]);
  if (!s) return nullptr;
  void* p;
  size_t len;
  if (!buffer_of(env, argv[1], &p, &len)) return nullptr;
  const int n = (int)(len / 16);
  void* out;
  napi_value ab = new_ab(env, (size_t)n * 8 + 1, &out);
  int st = mc_stock_sample(s, (const double*)p, n, (double*)out);
  if (st) return throw_mc(env, st);
  napi_value ta;
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, n, ab, 0, &ta));
  return ta;
}

// ---- the device-resident CuttingEngine (mc_engine_*) for js/cutting.js
void engine_finalize(napi_env, void* data, void*) {
  mc_engine** h = (mc_engine**)data;
  if (h && *h) mc_engine_destroy(*h);
  delete h;
}

mc_engine* engine_of(napi_env env, napi_value v) {
  mc_engine** h = nullptr;
  if (napi_get_value_external(env, v, (void**)&h) != napi_ok || !h || !*h) {
    napi_throw_type_error(env, nullptr, "not a live engine handle");
    return nullptr;
  }
  return *h;
}

napi_value f64_array(napi_env env, const double* v, size_t n) {
  void* d;
  napi_value ab = new_ab(env, n * 8, &d);
  if (!ab) return nullptr;
  std::memcpy(d, v, n * 8);
  napi_value ta;
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, n, ab, 0, &ta));
  return ta;
}

napi_value EngineCreate(napi_env env, napi_callback_info) {
  mc_engine* e = nullptr;
  int st = mc_engine_create(&e);
  if (st) return throw_mc(env, st);
  mc_engine** h = new mc_engine*(e);
  napi_value ext;
  NAPI_OK(napi_create_external(env, h, engine_finalize, nullptr, &ext));
  return ext;
}

// engineConfigure(h, cell: one packed mc_cell, fnMax) -> Float64Array [dt, omega, Ttooth, dz, S,
// stepsPerTooth, stepsPerRev, decim, piStep, maxSlices]
napi_value EngineConfigure(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  mc_engine* e = engine_of(env, argv[0]);
  if (!e) return nullptr;
  void* cell;
  size_t cb;
  if (!buffer_of(env, argv[1], &cell, &cb)) return nullptr;
  if (cb < sizeof(mc_cell)) { napi_throw_range_error(env, nullptr, "cell record too short"); return nullptr; }
  mc_engine_info in;
  int st = mc_engine_configure(e, (const mc_cell*)cell, num(env, argv[2]), &in);
  if (st) return throw_mc(env, st);
  const double v[10] = {in.dt, in.omega, in.Ttooth, in.dz, in.S, (double)in.stepsPerTooth, (double)in.stepsPerRev,
                        (double)in.decim, (double)in.piStep, (double)in.maxSlices};
  return f64_array(env, v, 10);
}

napi_value EngineSetWear(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  mc_engine* e = engine_of(env, argv[0]);
  if (!e) return nullptr;
  int st = mc_engine_set_wear(e, num(env, argv[1]), num(env, argv[2]));
  if (st) return throw_mc(env, st);
  return nullptr;
}

// engineSetWork(h, null | Float64Array [k, m, zeta, nx, ny])
napi_value EngineSetWork(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  mc_engine* e = engine_of(env, argv[0]);
  if (!e) return nullptr;
  int st;
  if (is_nullish(env, argv[1])) {
    st = mc_engine_set_work(e, nullptr);
  } else {
    void* w;
    size_t wb;
    if (!buffer_of(env, argv[1], &w, &wb)) return nullptr;
    if (wb < sizeof(mc_wall)) { napi_throw_range_error(env, nullptr, "wall needs 5 values"); return nullptr; }
    st = mc_engine_set_work(e, (const mc_wall*)w);
  }
  if (st) return throw_mc(env, st);
  return nullptr;
}

// engineSetEngagement(h, ux, uy, fz, tipZ, ap, sliceR | null, sliceKap | null) -> nz
napi_value EngineSetEngagement(napi_env env, napi_callback_info info) {
  napi_value argv[8];
  if (!get_args(env, info, 8, argv)) return nullptr;
  mc_engine* e = engine_of(env, argv[0]);
  if (!e) return nullptr;
  void *r = nullptr, *k = nullptr;
  size_t rb = 0, kb = 0;
  if (!is_nullish(env, argv[6]) && !buffer_of(env, argv[6], &r, &rb)) return nullptr;
  if (!is_nullish(env, argv[7]) && !buffer_of(env, argv[7], &k, &kb)) return nullptr;
  if (r && rb < 8 * (size_t)MC_MAX_SLICES) { napi_throw_range_error(env, nullptr, "sliceR needs MC_MAX_SLICES entries"); return nullptr; }
  if (k && kb < 8 * (size_t)MC_MAX_SLICES) { napi_throw_range_error(env, nullptr, "sliceKap needs MC_MAX_SLICES entries"); return nullptr; }
  int32_t nz = 0;
  int st = mc_engine_set_engagement(e, num(env, argv[1]), num(env, argv[2]), num(env, argv[3]), num(env, argv[4]),
                                    num(env, argv[5]), (const double*)r, (const double*)k, &nz);
  if (st) return throw_mc(env, st);
  napi_value v;
  NAPI_OK(napi_create_int32(env, nz, &v));
  return v;
}

// engineSetEndEngagement(h, fza, rOut, rings: Float64Array(24) | null)
napi_value EngineSetEndEngagement(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return nullptr;
  mc_engine* e = engine_of(env, argv[0]);
  if (!e) return nullptr;
  void* g = nullptr;
  size_t gb = 0;
  if (!is_nullish(env, argv[3]) && !buffer_of(env, argv[3], &g, &gb)) return nullptr;
  if (g && gb < 8 * 4 * (size_t)MC_END_RINGS) { napi_throw_range_error(env, nullptr, "rings need 24 values"); return nullptr; }
  int st = mc_engine_set_end_engagement(e, num(env, argv[1]), num(env, argv[2]), (const double*)g);
  if (st) return throw_mc(env, st);
  return nullptr;
}

// engineRun(h, nSteps, Hb: Float32Array(180), Eb: Float32Array(216), set: Float64Array [x, y,
// vx, vy, w, wv, t, decim, hist_i], Kte, Kre) -> Float64Array: the state after the run [x, y,
// vx, vy, w, wv, t, runMeanW, wk, wnx, wny, stepCount, pr, decim, nz, hist_i, hist_count, ts_i,
// ts_count, ws_i, ws_count, wsp_i, wsp_count], the 23 stats, the number of chip events
// the engine's state, statistics and chip-event count: [x, y, vx, vy, w, wv, t, runMeanW, wk,
// wnx, wny, stepCount, pr, decim, nz, hist_i, hist_count, ts_i, ts_count, ws_i, ws_count,
// wsp_i, wsp_count, the 23 mc_engine_stats fields, events]
napi_value engine_snapshot(napi_env env, mc_engine* e) {
  mc_engine_state s;
  mc_engine_stats sts;
  int st = mc_engine_get_state(e, &s);
  if (!st) st = mc_engine_read_stats(e, &sts, 0);
  int64_t nev = 0;
  if (!st) st = mc_engine_chip_events(e, nullptr, 0, &nev);
  if (st) return throw_mc(env, st);
  double v[23 + 23 + 1];
  const double sv2[23] = {s.x, s.y, s.vx, s.vy, s.w, s.wv, s.t, s.runMeanW, s.wk, s.wnx, s.wny, (double)s.stepCount,
                          (double)s.pr, (double)s.decim, (double)s.nz, (double)s.hist_i, (double)s.hist_count,
                          (double)s.ts_i, (double)s.ts_count, (double)s.ws_i, (double)s.ws_count,
                          (double)s.wsp_i, (double)s.wsp_count};
  std::memcpy(v, sv2, sizeof sv2);
  std::memcpy(v + 23, &sts, sizeof sts);
  v[46] = (double)nev;
  return f64_array(env, v, 47);
}

napi_value EngineRun(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return nullptr;
  mc_engine* e = engine_of(env, argv[0]);
  if (!e) return nullptr;
  void *hb, *eb, *sv;
  size_t hbb, ebb, svb;
  if (!buffer_of(env, argv[2], &hb, &hbb) || !buffer_of(env, argv[3], &eb, &ebb) || !buffer_of(env, argv[4], &sv, &svb))
    return nullptr;
  if (hbb < 4 * (size_t)MC_BINS || ebb < 4 * (size_t)(MC_END_BINS * MC_END_RINGS) || svb < 8 * 9) {
    napi_throw_range_error(env, nullptr, "engine maps or state too short");
    return nullptr;
  }
  const double* set = (const double*)sv;
  mc_engine_state s;
  int st = mc_engine_get_state(e, &s);
  if (!st) {
    s.x = set[0]; s.y = set[1]; s.vx = set[2]; s.vy = set[3]; s.w = set[4]; s.wv = set[5]; s.t = set[6];
    s.decim = (int32_t)set[7]; s.hist_i = (int64_t)set[8];
    st = mc_engine_set_state(e, &s);
  }
  if (!st) st = mc_engine_set_maps(e, (const float*)hb, (const float*)eb);
  if (!st) st = mc_engine_set_wear(e, num(env, argv[5]), num(env, argv[6]));
  if (!st) st = mc_engine_run(e, (int64_t)num(env, argv[1]));
  if (st) return throw_mc(env, st);
  return engine_snapshot(env, e);
}

// engineState(h) -> the Float64Array(47) engineRun returns, without running: the state a
// device pipeline (simRun) left
napi_value EngineState(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  mc_engine* e = engine_of(env, argv[0]);
  if (!e) return nullptr;
  return engine_snapshot(env, e);
}

// engineTakeStats(h): restart the device statistics (takeStats)
napi_value EngineTakeStats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  mc_engine* e = engine_of(env, argv[0]);
  if (!e) return nullptr;
  int st = mc_engine_read_stats(e, nullptr, 1);
  if (st) return throw_mc(env, st);
  return nullptr;
}

// engineChipEvents(h) -> Float64Array [vol, phi, z] per event of the last run
napi_value EngineChipEvents(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  mc_engine* e = engine_of(env, argv[0]);
  if (!e) return nullptr;
  int64_t n = 0;
  int st = mc_engine_chip_events(e, nullptr, 0, &n);
  if (st) return throw_mc(env, st);
  std::vector<mc_chip_event> ev((size_t)std::max<int64_t>(n, 1));
  st = mc_engine_chip_events(e, ev.data(), n, &n);
  if (st) return throw_mc(env, st);
  std::vector<double> v((size_t)(3 * n));
  for (int64_t i = 0; i < n; i++) { v[3 * i] = ev[i].vol; v[3 * i + 1] = ev[i].phi; v[3 * i + 2] = ev[i].z; }
  return f64_array(env, v.data(), v.size());
}

napi_value EngineChatterIndex(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  mc_engine* e = engine_of(env, argv[0]);
  if (!e) return nullptr;
  double out[3];
  int st = mc_engine_chatter_index(e, out);
  if (st) return throw_mc(env, st);
  return f64_array(env, out, 3);
}

// engineRings(h) -> Float32Array: hist (5 x 4096: Fx Fy Fz x y), then toothSamples, wallSamples,
// wallSamplesPi (2 x 64 each: x, y)
napi_value EngineRings(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  mc_engine* e = engine_of(env, argv[0]);
  if (!e) return nullptr;
  const size_t n = 5 * MC_ENGINE_HIST + 6 * MC_ENGINE_RING;
  void* d;
  napi_value ab = new_ab(env, n * 4, &d);
  if (!ab) return nullptr;
  float* f = (float*)d;
  int st = mc_engine_read_rings(e, f, f + 5 * MC_ENGINE_HIST, f + 5 * MC_ENGINE_HIST + 2 * MC_ENGINE_RING,
                                f + 5 * MC_ENGINE_HIST + 4 * MC_ENGINE_RING);
  if (st) return throw_mc(env, st);
  napi_value ta;
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, n, ab, 0, &ta));
  return ta;
}


// bisectPriorBatch(cells, aMin, aMax, iters, mu: Float64Array(n), sig: Float64Array(n))
//   -> ArrayBuffer of mc_limit (the plain search's results; the prior only steers speculation)
napi_value BisectPriorBatch(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return nullptr;
  void *cells, *mu, *sig;
  size_t cb, mb, sb;
  if (!buffer_of(env, argv[0], &cells, &cb) || !buffer_of(env, argv[4], &mu, &mb) || !buffer_of(env, argv[5], &sig, &sb))
    return nullptr;
  const int n = (int)(cb / sizeof(mc_cell));
  if (mb / 8 < (size_t)n || sb / 8 < (size_t)n) { napi_throw_range_error(env, nullptr, "prior arrays too short"); return nullptr; }
  mc_bisect_opts o{num(env, argv[1]), num(env, argv[2]), (int32_t)num(env, argv[3]), 0};
  void* out;
  napi_value res = new_ab(env, (size_t)n * sizeof(mc_limit) + 1, &out);
  int st = mc_bisect_prior_batch((const mc_cell*)cells, n, &o, (const double*)mu, (const double*)sig, (mc_limit*)out);
  if (st) return throw_mc(env, st);
  return res;
}

// surrogateFeatures(cells) -> Float32Array(n * MC_SURR_IN)
napi_value SurrogateFeatures(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  void* cells;
  size_t cb;
  if (!buffer_of(env, argv[0], &cells, &cb)) return nullptr;
  const int n = (int)(cb / sizeof(mc_cell));
  void* out;
  napi_value ab = new_ab(env, (size_t)n * MC_SURR_IN * 4 + 4, &out);
  int st = mc_surrogate_features((const mc_cell*)cells, n, (float*)out);
  if (st) return throw_mc(env, st);
  napi_value ta;
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, (size_t)n * MC_SURR_IN, ab, 0, &ta));
  return ta;
}

// ---- the surrogate network: an External wrapping mc_mlp*
void mlp_finalize(napi_env, void* data, void*) {
  mc_mlp** h = (mc_mlp**)data;
  if (h && *h) mc_mlp_destroy(*h);
  delete h;
}

mc_mlp* mlp_of(napi_env env, napi_value v) {
  mc_mlp** h = nullptr;
  if (napi_get_value_external(env, v, (void**)&h) != napi_ok || !h || !*h) {
    napi_throw_type_error(env, nullptr, "not a live surrogate handle");
    return nullptr;
  }
  return *h;
}

// mlpCreate(params: Float32Array(mc_mlp_nparams()), xMean, xStd: Float32Array(MC_SURR_IN), yMean, yStd)
//   -> handle of a network with those weights and that input / output scaling
napi_value MlpCreate(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return nullptr;
  void *pp, *xm, *xs;
  size_t pb, mb, sb;
  if (!buffer_of(env, argv[0], &pp, &pb) || !buffer_of(env, argv[1], &xm, &mb) || !buffer_of(env, argv[2], &xs, &sb))
    return nullptr;
  if (pb < (size_t)mc_mlp_nparams() * 4 || mb < MC_SURR_IN * 4 || sb < MC_SURR_IN * 4) {
    napi_throw_range_error(env, nullptr, "surrogate parameters or scaling too short");
    return nullptr;
  }
  mc_mlp* m = nullptr;
  int st = mc_mlp_create(1, &m);
  if (!st) st = mc_mlp_params(m, (float*)pp, 1);
  if (!st) st = mc_mlp_set_scaling(m, (const float*)xm, (const float*)xs, (float)num(env, argv[3]), (float)num(env, argv[4]));
  if (st) {
    if (m) mc_mlp_destroy(m);
    return throw_mc(env, st);
  }
  mc_mlp** h = new mc_mlp*(m);
  napi_value ext;
  NAPI_OK(napi_create_external(env, h, mlp_finalize, nullptr, &ext));
  return ext;
}

// mlpPredict(h, X: Float32Array(n * MC_SURR_IN)) -> Float32Array(n): ln alim
napi_value MlpPredict(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  mc_mlp* m = mlp_of(env, argv[0]);
  if (!m) return nullptr;
  void* X;
  size_t xb;
  if (!buffer_of(env, argv[1], &X, &xb)) return nullptr;
  const int n = (int)(xb / (4 * MC_SURR_IN));
  void* out;
  napi_value ab = new_ab(env, (size_t)n * 4 + 4, &out);
  int st = mc_mlp_predict(m, (const float*)X, n, (float*)out);
  if (st) return throw_mc(env, st);
  napi_value ta;
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, (size_t)n, ab, 0, &ta));
  return ta;
}

napi_value MlpNparams(napi_env env, napi_callback_info) {
  napi_value v;
  NAPI_OK(napi_create_int32(env, mc_mlp_nparams(), &v));
  return v;
}

// pathDiscretize(moves: ArrayBuffer of mc_move, maxSeg) -> ArrayBuffer of mc_path_point
napi_value PathDiscretize(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  void* mv;
  size_t mb;
  if (!buffer_of(env, argv[0], &mv, &mb)) return nullptr;
  const int n = (int)(mb / sizeof(mc_move));
  const double maxSeg = num(env, argv[1]);
  int np = 0;
  int st = mc_path_discretize((const mc_move*)mv, n, maxSeg, nullptr, 0, &np);
  if (st) return throw_mc(env, st);
  void* out;
  napi_value ab = new_ab(env, (size_t)np * sizeof(mc_path_point), &out);
  if (np > 0) st = mc_path_discretize((const mc_move*)mv, n, maxSeg, (mc_path_point*)out, np, &np);
  if (st) return throw_mc(env, st);
  return ab;
}

// simSchedule(pts: ArrayBuffer of mc_path_point, hx, dtReal, speed, feedOverride, flutes, maxFrames)
//   -> [ArrayBuffer of mc_substep, Int32Array of (first, count) per frame]
napi_value SimSchedule(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return nullptr;
  void* pts;
  size_t pb;
  if (!buffer_of(env, argv[0], &pts, &pb)) return nullptr;
  const int np = (int)(pb / sizeof(mc_path_point));
  const double mf = num(env, argv[6]);
  mc_schedule_opts o{num(env, argv[1]), num(env, argv[2]), num(env, argv[3]), num(env, argv[4]),
                     (int32_t)num(env, argv[5]), (int32_t)std::min(mf, 2147483647.0)};
  int ns = 0, nf = 0;
  int st = mc_sim_schedule((const mc_path_point*)pts, np, &o, nullptr, 0, &ns, nullptr, 0, &nf);
  if (st) return throw_mc(env, st);
  void *sub, *fr;
  napi_value sab = new_ab(env, (size_t)ns * sizeof(mc_substep), &sub);
  napi_value fab = new_ab(env, (size_t)nf * 8, &fr);
  if (ns > 0 || nf > 0)
    st = mc_sim_schedule((const mc_path_point*)pts, np, &o, (mc_substep*)sub, ns, &ns, o.dt_real > 0 ? (int32_t*)fr : nullptr,
                         o.dt_real > 0 ? nf : 0, &nf);
  if (st) return throw_mc(env, st);
  napi_value fta, arr;
  NAPI_OK(napi_create_typedarray(env, napi_int32_array, (size_t)nf * 2, fab, 0, &fta));
  NAPI_OK(napi_create_array_with_length(env, 2, &arr));
  NAPI_OK(napi_set_element(env, arr, 0, sab));
  NAPI_OK(napi_set_element(env, arr, 1, fta));
  return arr;
}

double prop_num(napi_env env, napi_value o, const char* k, double def) {
  bool has = false;
  napi_value v;
  napi_valuetype t;
  if (napi_has_named_property(env, o, k, &has) != napi_ok || !has) return def;
  if (napi_get_named_property(env, o, k, &v) != napi_ok || napi_typeof(env, v, &t) != napi_ok) return def;
  if (t != napi_number) return def;
  return num(env, v);
}

bool prop_val(napi_env env, napi_value o, const char* k, napi_value* v) {
  bool has = false;
  napi_valuetype t;
  if (napi_has_named_property(env, o, k, &has) != napi_ok || !has) return false;
  if (napi_get_named_property(env, o, k, v) != napi_ok || napi_typeof(env, *v, &t) != napi_ok) return false;
  return t != napi_undefined && t != napi_null;
}

// simRun(programs: [{actual, ideal (stock handles), engine (engine handle), sub (ArrayBuffer of
//   mc_substep), R, fluteLen, rw, share, type, lut (Float32Array, bull-nose or profile), cornerR,
//   edgeR (Float64Array), edgeDz (an imported edge), dynBudget, fz, frames (true: the frames'
//   stats and chatter index), budgetLeft (the frame budget left when the call starts inside a
//   frame)}])
//   -> [[ArrayBuffer of mc_substep_out, ArrayBuffer of mc_frame_out, budget left after the call
//        (a number, or undefined without budgetLeft)], ...]
// the programs' sub-step loops in one device call (mc_sim_run)
napi_value SimRun(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  bool isarr = false;
  napi_is_array(env, argv[0], &isarr);
  if (!isarr) { napi_throw_type_error(env, nullptr, "expected an array of programs"); return nullptr; }
  uint32_t n = 0;
  NAPI_OK(napi_get_array_length(env, argv[0], &n));
  std::vector<mc_sim_program> progs(n);
  std::vector<napi_value> outs(n), frs(n);
  std::vector<double> left(n, 0.0);
  for (uint32_t i = 0; i < n; i++) {
    napi_value o, v;
    NAPI_OK(napi_get_element(env, argv[0], i, &o));
    mc_sim_program& g = progs[i];
    std::memset(&g, 0, sizeof g);
    if (prop_val(env, o, "budgetLeft", &v)) {
      left[i] = prop_num(env, o, "budgetLeft", 0);
      g.budgetLeft = &left[i];
    }
    if (!prop_val(env, o, "actual", &v) || !(g.actual = stock_of(env, v))) return nullptr;
    if (!prop_val(env, o, "ideal", &v) || !(g.ideal = stock_of(env, v))) return nullptr;
    if (!prop_val(env, o, "engine", &v) || !(g.engine = engine_of(env, v))) return nullptr;
    void* sub;
    size_t sb;
    if (!prop_val(env, o, "sub", &v) || !buffer_of(env, v, &sub, &sb)) {
      if (!napi_is_exception_pending_bool(env)) napi_throw_type_error(env, nullptr, "program without sub-steps");
      return nullptr;
    }
    g.sub = (const mc_substep*)sub;
    g.nsub = (int32_t)prop_num(env, o, "nsub", (double)(sb / sizeof(mc_substep)));
    if (g.nsub < 0 || (size_t)g.nsub > sb / sizeof(mc_substep)) {
      napi_throw_range_error(env, nullptr, "nsub outside the sub-step array");
      return nullptr;
    }
    g.R = prop_num(env, o, "R", 0);
    g.fluteLen = prop_num(env, o, "fluteLen", 0);
    g.rw = prop_num(env, o, "rw", 0);
    g.share = prop_num(env, o, "share", 1);
    g.type = (int32_t)prop_num(env, o, "type", MC_TOOL_FLAT);
    g.cornerR = prop_num(env, o, "cornerR", 0);
    g.dynBudget = prop_num(env, o, "dynBudget", 0);
    g.fz = prop_num(env, o, "fz", 0);
    if (prop_val(env, o, "lut", &v)) {
      void* lut;
      size_t lb;
      if (!buffer_of(env, v, &lut, &lb)) return nullptr;
      g.lut = (const float*)lut;
      g.lutn = (int32_t)(lb / 4);
    }
    if (prop_val(env, o, "edgeR", &v)) {   // an imported edge's radii (Float64Array)
      void* er;
      size_t eb;
      if (!buffer_of(env, v, &er, &eb)) return nullptr;
      g.edge_r = (const double*)er;
      g.edge_n = (int32_t)(eb / 8);
      g.edge_dz = prop_num(env, o, "edgeDz", 0);
    }
    bool want = false;
    if (prop_val(env, o, "frames", &v)) napi_get_value_bool(env, v, &want);
    int nf = 0;
    if (want)
      for (int k = 0; k < g.nsub; k++) nf += (g.sub[k].frame & MC_FRAME_LAST) ? 1 : 0;
    void *op, *fp;
    outs[i] = new_ab(env, (size_t)g.nsub * sizeof(mc_substep_out), &op);
    frs[i] = new_ab(env, (size_t)nf * sizeof(mc_frame_out), &fp);
    if (!outs[i] || !frs[i]) { napi_throw_error(env, nullptr, "out of memory"); return nullptr; }
    g.out = (mc_substep_out*)op;
    g.frames = nf ? (mc_frame_out*)fp : nullptr;
    g.nframes = nf;
  }
  int st = mc_sim_run(progs.data(), (int)n);
  if (st) return throw_mc(env, st);
  napi_value res;
  NAPI_OK(napi_create_array_with_length(env, n, &res));
  for (uint32_t i = 0; i < n; i++) {
    napi_value pair, bl;
    NAPI_OK(napi_create_array_with_length(env, 3, &pair));
    NAPI_OK(napi_set_element(env, pair, 0, outs[i]));
    NAPI_OK(napi_set_element(env, pair, 1, frs[i]));
    if (progs[i].budgetLeft) NAPI_OK(napi_create_double(env, left[i], &bl));
    else NAPI_OK(napi_get_undefined(env, &bl));
    NAPI_OK(napi_set_element(env, pair, 2, bl));
    NAPI_OK(napi_set_element(env, res, i, pair));
  }
  return res;
}

napi_value Sizes(napi_env env, napi_callback_info) {
  napi_value o, v;
  NAPI_OK(napi_create_object(env, &o));
  const struct { const char* k; size_t s; } t[] = {
      {"mc_cell", sizeof(mc_cell)}, {"mc_trial", sizeof(mc_trial)}, {"mc_limit", sizeof(mc_limit)},
      {"mc_stamp", sizeof(mc_stamp)}, {"mc_stamp_opt", sizeof(mc_stamp_opt)},
      {"mc_engage_query", sizeof(mc_engage_query)}, {"mc_engage_result", sizeof(mc_engage_result)}, {"mc_lobe_args", sizeof(mc_lobe_args)},
      {"mc_move", sizeof(mc_move)}, {"mc_path_point", sizeof(mc_path_point)}, {"mc_substep", sizeof(mc_substep)},
      {"mc_substep_out", sizeof(mc_substep_out)}, {"mc_frame_out", sizeof(mc_frame_out)},
      {"mc_sim_program", sizeof(mc_sim_program)},
      {"MC_SURR_IN", MC_SURR_IN}, {"MC_MAX_MODES", MC_MAX_MODES}, {"MC_MAX_FLUTES", MC_MAX_FLUTES}};
  for (auto& e : t) {
    napi_create_uint32(env, (uint32_t)e.s, &v);
    napi_set_named_property(env, o, e.k, v);
  }
  return o;
}

napi_value ModuleInit(napi_env env, napi_value exports) {
  const struct { const char* name; napi_callback cb; } fns[] = {
      {"init", Init}, {"deviceInfo", DeviceInfo}, {"lastKernelMs", LastKernelMs},
      {"chatterBatch", ChatterBatch}, {"bisectBatch", BisectBatch},
      {"bisectBatchAsync", BisectBatchAsync}, {"bracketBatch", BracketBatch},
      {"engineTrace", EngineTrace}, {"stockCreate", StockCreate}, {"stockDestroy", StockDestroy},
      {"stockDims", StockDims}, {"stockSetWallBand", StockSetWallBand}, {"stockRead", StockRead},
      {"stockWriteH", StockWriteH}, {"stampBatch", StampBatch}, {"stockSample", StockSample},
      {"stockReadMarks", StockReadMarks}, {"stockDirty", StockDirty}, {"stockClearDirty", StockClearDirty},
      {"stockBlockMax", StockBlockMax}, {"stockMaxInDisk", StockMaxInDisk}, {"tipLut", TipLut},
      {"stockEngagement", StockEngagement}, {"engineCreate", EngineCreate},
      {"engineConfigure", EngineConfigure}, {"engineSetWear", EngineSetWear}, {"engineSetWork", EngineSetWork},
      {"engineSetEngagement", EngineSetEngagement}, {"engineSetEndEngagement", EngineSetEndEngagement},
      {"engineRun", EngineRun}, {"engineTakeStats", EngineTakeStats}, {"engineChipEvents", EngineChipEvents},
      {"engineChatterIndex", EngineChatterIndex}, {"engineRings", EngineRings},
      {"compositeLobes", CompositeLobes}, {"engineState", EngineState}, {"bisectPriorBatch", BisectPriorBatch},
      {"surrogateFeatures", SurrogateFeatures}, {"mlpCreate", MlpCreate}, {"mlpPredict", MlpPredict},
      {"mlpNparams", MlpNparams}, {"pathDiscretize", PathDiscretize}, {"simSchedule", SimSchedule},
      {"simRun", SimRun}, {"sizes", Sizes}};
  for (auto& f : fns) {
    napi_value fn;
    if (napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn) != napi_ok ||
        napi_set_named_property(env, exports, f.name, fn) != napi_ok)
      return nullptr;
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, ModuleInit)
