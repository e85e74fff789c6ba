// C-ABI entry points of libksig_hip.so (declared in include/ksig_hip.h):
// argument validation, LDS sizing and dispatch to the gfx950 kernels.

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "ksig_common.h"
#include "ksig_hip.h"
#include "dp_forward.h"
#include "pde_common.h"
#include "dp_stream.h"
#include "gak.h"
#include "rws.h"
#include "lowrank.h"

namespace ksig {
hipError_t lowrank_scan_launch(int dtype, const void* z, int n_levels, int order,
                               int64_t n_seq, int steps, int n_comp, double* out,
                               hipStream_t s);
hipError_t lowrank_diag_scan_launch(int dtype, const void* du, int q, int n_levels, int order,
                                    int64_t n_seq, int steps, int half, double* out,
                                    hipStream_t s);
hipError_t lowrank_sketch_launch(int dtype, const void* du, int n_levels, int order,
                                 int64_t n_seq, int steps, int width, const int* hidx,
                                 const void* hbit, int m, int ng, double* out, hipStream_t s);
size_t lowrank_sketch_lds(int n_levels, int order, int m, int width, int ng);
int lowrank_sketch_groups(int n_levels, int order, int m, int width, size_t lds_limit);
int lowrank_sketch_max_comp();
hipError_t rff_diff_launch(int dtype, const void* p, int64_t n_seq, int L, int half, int ld,
                           int difference, void* out, hipStream_t s);
}  // namespace ksig

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  return fail((int)e, "%s: %s", where, hipGetErrorString(e));
}

size_t max_lds_bytes() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 64 * 1024;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return 64 * 1024;
  return (size_t)v;
}

int dmax_for(int d) { return d <= 8 ? 8 : (d <= 16 ? 16 : 32); }

int device_cus() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 1;
  return v > 0 ? v : 1;
}

// Resident row-stream waves per CU for a (kernel, block LDS size, waves per
// block), cached.
int stream_occupancy(const ksig::StreamEntry& e, size_t lds, int wpb) {
  struct Slot { const void* fn; size_t lds; int wpb; int occ; };
  static Slot cache[256];
  static int used = 0;
  static std::mutex mu;
  std::lock_guard<std::mutex> g(mu);
  for (int i = 0; i < used; ++i)
    if (cache[i].fn == e.fn && cache[i].lds == lds && cache[i].wpb == wpb) return cache[i].occ;
  const int occ = ksig::dp_stream_occupancy(e, lds, wpb);
  if (used < 256) cache[used++] = Slot{e.fn, lds, wpb, occ};
  return occ;
}

// Scale factors folded into the inputs: RBF uses exp2(-|x'-y'|^2) with
// x' = sqrt(log2(e) / max(2 sigma^2, 1e-12)) x (ksig/static/kernels.py:258-260,
// ksig/utils.py:16); linear uses <scale x, y> (ksig/static/kernels.py:155).
int static_scales(int kind, double param, double* xs, double* ys) {
  if (!(param > 0.0))
    return fail(KSIG_EINVAL, "static kernel parameter must be positive (got %g)", param);
  if (kind == ksig::kRBF) {
    double den = 2.0 * param * param;
    if (den < 1e-12) den = 1e-12;
    const double s = sqrt(1.4426950408889634 / den);
    *xs = s;
    *ys = s;
    return 0;
  }
  if (kind == ksig::kLinear) {
    *xs = param;
    *ys = 1.0;
    return 0;
  }
  return fail(KSIG_EINVAL, "unknown static kernel kind %d", kind);
}

int check_common(int dtype, const void* x, const void* y, int d, const int32_t* pairs,
                 int64_t n_pairs, double* out) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (n_pairs < 0) return fail(KSIG_EINVAL, "n_pairs < 0");
  if (n_pairs > 0x7fffffffLL) return fail(KSIG_EINVAL, "n_pairs exceeds 2^31-1");
  if (n_pairs > 0 && (!x || !y || !pairs || !out))
    return fail(KSIG_EINVAL, "null device pointer");
  if (d < 1) return fail(KSIG_EINVAL, "d must be >= 1 (got %d)", d);
  if (d > 32) return fail(KSIG_EUNSUPPORTED, "d = %d > 32 is not built into the fused kernels", d);
  return 0;
}

// Scratch of the allocation-free entry point
// (ksig_signature_levels_batched_noalloc): the row stream's global band FIFO.
// Each GPU the library runs on holds one copy, at its own address, until the
// process ends.  Host code never passes the array's name to a kernel (on the
// host it names a stand-in, not device memory): it uses the device address
// that hipGetSymbolAddress returns, looked up once per device.
static __device__ __align__(256) unsigned char scratch[KSIG_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of this GPU's scratch (the current device's), looked up on
// the first call on that device.
int scratch_address(void** ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(KSIG_EINVAL, "device %d beyond the %d the scratch table holds", dev, kMaxDevices);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(KSIG_EINVAL, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" {

const char* ksig_version(void) { return "ksig_hip 0.1.0 (gfx950)"; }

const char* ksig_last_error(void) { return g_err; }

int ksig_dp_supported(int n_levels, int order) {
  return ksig::dp_supported(n_levels, order) ? 1 : 0;
}

static int dp_levels_common(int dtype, ksig::DpFwdArgs& a, int cols, int dmax,
                            void* stream) {
  const size_t esz = dtype == KSIG_F32 ? 4 : 8;
  const int nf = ksig::dp_flow_count(a.n_levels, a.order);
  const int pc = (a.difference ? cols : cols + 1) > ksig::kWave
                     ? (a.difference ? cols : cols + 1) : ksig::kWave;
  const bool fused = a.kind != ksig::kGram;
  const size_t ycols = fused ? (size_t)cols : 0;
  const size_t xrows = fused ? 2 * (size_t)ksig::kWave : 0;
  const size_t lds = (xrows * (dmax + 1) + ycols * (dmax + 1) + (size_t)pc * nf) * esz;
  if (lds > max_lds_bytes())
    return fail(KSIG_ELDS, "per-pair LDS state of %zu bytes exceeds the device limit "
                "(shorter side length %d)", lds, cols);
  hipError_t e = ksig::dp_forward_launch(a, dtype, lds, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "dp_forward launch");
  return 0;
}

// Caller-matrix recursion: the register-resident kernel for the compiled
// (n_levels, order) table, the runtime-levels kernel for every other pair.
static int gram_levels_dispatch(int dtype, ksig::DpFwdArgs& a, int cols, void* stream) {
  if (ksig::dp_supported(a.n_levels, a.order)) return dp_levels_common(dtype, a, cols, 8, stream);
  const size_t lds = ksig::dp_generic_lds(dtype, a.n_levels, a.order, cols);
  if (lds > max_lds_bytes())
    return fail(KSIG_ELDS, "(n_levels=%d, order=%d) with %d columns needs %zu bytes of LDS "
                "state, above the device limit", a.n_levels, a.order, cols, lds);
  hipError_t e = ksig::dp_generic_launch(a, dtype, lds, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "dp_generic launch");
  return 0;
}

// Row-stream launch (float32, difference=1): one persistent pass, every wave a
// contiguous slice of the pair list.  stream_levels returns -1 when no
// compiled row-stream kernel or no LDS layout fits, so the caller can fall
// back.
struct StreamPlan {
  ksig::StreamEntry e;
  size_t lds;     // block LDS
  size_t region;  // LDS of one wave
  int wpb;        // waves per block (2 with the global FIFO: two-wave slices)
  bool gf;        // FIFO in global memory
  int pc;         // FIFO slots
  int cap;        // pair-table entries
  int64_t slots;  // resident waves on the device
};

// Kernel, LDS size and resident waves of the row stream for a problem with a
// pair table of `cap` entries; false if no kernel or LDS layout fits.
static bool stream_plan(int d, int kind, int n_levels, int order, int max_len_x,
                        int max_len_y, bool vec, bool gf, int cap, StreamPlan* p,
                        bool save = false) {
  const int dmax = dmax_for(d);
  if (save && (!vec || gf || order != 1)) return false;
  p->e = save ? ksig::dp_stream_save_entry(kind, dmax, n_levels)
              : ksig::dp_stream_entry(kind, dmax, vec, gf, n_levels, order);
  if (!p->e.fn) return false;
  const int nf = ksig::dp_flow_count(n_levels, order);
  int pc = max_len_x < max_len_y ? max_len_x : max_len_y;
  const int min_cols = gf ? ksig::kStreamMinColsGlobal : ksig::kStreamMinCols;
  if (pc < min_cols) pc = min_cols;
  p->gf = gf;
  p->cap = (cap + 3) & ~3;
  p->pc = pc;
  p->region = ksig::stream_region_bytes(ksig::stream_lds_bytes(dmax, p->e.r, pc, nf, p->cap, gf));
  p->wpb = gf ? 2 : 1;
  p->lds = ksig::stream_block_lds_bytes(p->region, p->wpb);
  if (p->lds > max_lds_bytes()) return false;
  const int occ = stream_occupancy(p->e, p->lds, p->wpb);
  p->slots = (int64_t)device_cus() * (occ > 0 ? occ : 1);
  return true;
}

static bool stream_vec(int d, const void* x, const void* y) {
  return d == dmax_for(d) && ((uintptr_t)x & 15) == 0 && ((uintptr_t)y & 15) == 0;
}

static int stream_launch(const StreamPlan& pl, const ksig::DpFwdArgs& f, double static_param,
                         const int32_t* sched_order, const int32_t* wave_beg, int n_waves,
                         int64_t per_wave, void* workspace, void* stream) {
  ksig::DpStreamArgs a{};
  a.save = f.save;
  a.save_off = f.save_off;
  a.gfifo = pl.gf ? workspace : nullptr;
  a.x = f.x;
  a.y = f.y;
  a.d = f.d;
  a.tab_cap = pl.cap;
  a.fifo_cols = pl.pc;
  a.pairs = f.pairs;
  a.n_pairs = f.n_pairs;
  a.order = sched_order;
  a.wave_beg = wave_beg;
  a.n_waves = n_waves;
  a.per_wave = (int)per_wave;
  a.lds_region = (int)pl.region;
  if (f.kind == ksig::kRBF) {
    double den = 2.0 * static_param * static_param;
    if (den < 1e-12) den = 1e-12;
    const double gn = 1.4426950408889634 / den;
    a.gn = (float)gn;
    a.gdot = (float)(2.0 * gn);
  } else {
    a.gn = 0.0f;
    a.gdot = (float)static_param;
  }
  a.out = f.out;
  hipError_t err = ksig::dp_stream_launch(pl.e, a, pl.lds, pl.wpb, (hipStream_t)stream);
  if (err != hipSuccess) return hip_fail(err, "dp_stream launch");
  return 0;
}

// Equal slices of the pair list: the smallest pair table (in steps of 4) whose
// resident waves take every pair in one round; past 128 pairs per wave, equal
// rounds of full tables.  Returns -1 when no row-stream kernel fits.
static int stream_levels(const ksig::DpFwdArgs& f, int max_len_x, int max_len_y,
                         double static_param, void* stream) {
  const bool vec = stream_vec(f.d, f.x.data, f.y.data);
  const bool save = f.save != nullptr;
  const int64_t n = f.n_pairs;
  StreamPlan pl;
  int64_t per = 0;
  for (int cap = 4; cap <= ksig::kStreamMaxPairs; cap += 4) {
    if (!stream_plan(f.d, f.kind, f.n_levels, f.order, max_len_x, max_len_y, vec, false, cap,
                     &pl, save))
      return -1;
    per = (n + pl.slots - 1) / pl.slots;
    if (per <= cap) break;
  }
  if (per > pl.cap) {
    const int64_t rounds = (n + pl.slots * pl.cap - 1) / (pl.slots * pl.cap);
    per = (n + pl.slots * rounds - 1) / (pl.slots * rounds);
  }
  if (per < 1) per = 1;
  const int64_t waves = (n + per - 1) / per;
  return stream_launch(pl, f, static_param, nullptr, nullptr, (int)waves, per, nullptr, stream);
}

int ksig_signature_levels(int dtype, const void* x, const int64_t* x_off,
                          const int32_t* x_len, int max_len_x, const void* y,
                          const int64_t* y_off, const int32_t* y_len,
                          int max_len_y, int d, const int32_t* pairs,
                          int64_t n_pairs, int static_kind, double static_param,
                          int n_levels, int order, int difference, int gram_mode,
                          double* out, void* stream) {
  int rc = check_common(dtype, x, y, d, pairs, n_pairs, out);
  if (rc) return rc;
  if (n_levels < 1) return fail(KSIG_EINVAL, "n_levels must be positive");
  if (order <= 0 || order >= n_levels) order = n_levels;  // ksig/algorithms.py:37
  if (!ksig::dp_supported(n_levels, order))
    return fail(KSIG_EUNSUPPORTED, "(n_levels=%d, order=%d) is not built", n_levels, order);
  ksig::DpFwdArgs a{};
  rc = static_scales(static_kind, static_param, &a.xscale, &a.yscale);
  if (rc) return rc;
  if (n_pairs == 0) return 0;
  if (max_len_x < 0 || max_len_y < 0) return fail(KSIG_EINVAL, "negative max length");
  a.x = {x, x_off, x_len};
  a.y = {y, y_off, y_len};
  a.d = d;
  a.pairs = pairs;
  a.n_pairs = n_pairs;
  a.kind = static_kind;
  a.n_levels = n_levels;
  a.order = order;
  a.difference = difference ? 1 : 0;
  a.out = out;
  int cols = max_len_x < max_len_y ? max_len_x : max_len_y;
  if (cols < 1) cols = 1;
  if (gram_mode < KSIG_GRAM_AUTO || gram_mode > KSIG_GRAM_MFMA_BF16X3)
    return fail(KSIG_EINVAL, "bad gram_mode %d", gram_mode);
  int mode = gram_mode;
  const bool mfma_ok = dtype == KSIG_F32 && a.difference && ksig::dp_mfma_supported(n_levels, order);
  if (mode == KSIG_GRAM_AUTO) mode = KSIG_GRAM_VALU;
  if (mode != KSIG_GRAM_VALU) {
    if (!mfma_ok)
      return fail(KSIG_EUNSUPPORTED, "MFMA Gram needs float32, difference=1 and a built "
                  "(n_levels=%d, order=%d)", n_levels, order);
    const int nf = ksig::dp_flow_count(n_levels, order);
    int pc = cols > 80 ? cols : 80;
    pc = (pc + 15) & ~15;
    const int dmax = dmax_for(d);
    const size_t lds = ((size_t)80 * 65 + (size_t)ksig::kWave * (dmax + 1) + (size_t)pc * nf) * 4;
    if (lds > max_lds_bytes())
      return fail(KSIG_ELDS, "per-pair LDS state of %zu bytes exceeds the device limit", lds);
    hipError_t e = ksig::dp_mfma_launch(a, mode, lds, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "dp_mfma launch");
    return 0;
  }
  if (a.difference && dtype == KSIG_F32) {
    rc = stream_levels(a, max_len_x, max_len_y, static_param, stream);
    if (rc != -1) return rc;  // -1: no row-stream kernel fits, use the per-pair sweep
  }
  if (a.difference) {
    const size_t lds = ksig::dp_sweep_lds(dtype, d, cols, n_levels, order);
    if (lds > max_lds_bytes())
      return fail(KSIG_ELDS, "per-pair LDS state of %zu bytes exceeds the device limit", lds);
    hipError_t e = ksig::dp_sweep_launch(a, dtype, lds, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "dp_sweep launch");
    return 0;
  }
  return dp_levels_common(dtype, a, cols, dmax_for(d), stream);
}

int ksig_stream_slots(int d, int static_kind, int n_levels, int order, int max_len_x,
                      int max_len_y, int max_slice, int global_fifo, int* slots,
                      int* max_pairs_per_wave, size_t* workspace_per_wave) {
  if (!slots || !max_pairs_per_wave || !workspace_per_wave)
    return fail(KSIG_EINVAL, "null output pointer");
  *slots = 0;
  *max_pairs_per_wave = ksig::kStreamMaxPairs;
  *workspace_per_wave = 0;
  if (d < 1 || d > 32) return 0;
  if (static_kind != ksig::kRBF && static_kind != ksig::kLinear) return 0;
  if (n_levels < 1) return fail(KSIG_EINVAL, "n_levels must be positive");
  if (order <= 0 || order >= n_levels) order = n_levels;
  if (max_slice < 1 || max_slice > ksig::kStreamMaxPairs) max_slice = ksig::kStreamMaxPairs;
  StreamPlan pl;
  if (!stream_plan(d, static_kind, n_levels, order, max_len_x, max_len_y, d == dmax_for(d),
                   global_fifo != 0, max_slice, &pl))
    return 0;
  *slots = (int)pl.slots;
  if (pl.gf)
    *workspace_per_wave = ksig::stream_gfifo_bytes(pl.pc, ksig::dp_flow_count(n_levels, order), 1);
  return 0;
}

// The batched row stream; noalloc: the band FIFO goes to the library's static
// scratch array (workspace is its device address, workspace_bytes its size).
static int batched_common(const void* x, const int64_t* x_off, const int32_t* x_len,
                          int max_len_x, const void* y, const int64_t* y_off,
                          const int32_t* y_len, int max_len_y, int d, const int32_t* pairs,
                          int64_t n_pairs, int static_kind, double static_param, int n_levels,
                          int order, const int32_t* sched_order, const int32_t* wave_beg,
                          int n_waves, int max_slice, void* workspace, size_t workspace_bytes,
                          double* out, void* stream, bool noalloc) {
  int rc = check_common(KSIG_F32, x, y, d, pairs, n_pairs, out);
  if (rc) return rc;
  if (n_levels < 1) return fail(KSIG_EINVAL, "n_levels must be positive");
  if (order <= 0 || order >= n_levels) order = n_levels;
  if (static_kind != ksig::kRBF && static_kind != ksig::kLinear)
    return fail(KSIG_EINVAL, "unknown static kernel kind %d", static_kind);
  double xs, ys;
  rc = static_scales(static_kind, static_param, &xs, &ys);
  if (rc) return rc;
  if (n_pairs == 0) return 0;
  if (max_len_x < 0 || max_len_y < 0) return fail(KSIG_EINVAL, "negative max length");
  if (wave_beg && n_waves < 1) return fail(KSIG_EINVAL, "n_waves must be positive");
  if (max_slice < 1 || max_slice > ksig::kStreamMaxPairs)
    return fail(KSIG_EINVAL, "max_slice must be in [1, %d] (got %d)", ksig::kStreamMaxPairs,
                max_slice);
  const bool gf = workspace != nullptr;
  if (gf && !stream_vec(d, x, y))
    return fail(KSIG_EUNSUPPORTED, "the global-memory FIFO needs d in {8, 16, 32} and 16-byte "
                "aligned inputs");
  StreamPlan pl;
  if (!stream_plan(d, static_kind, n_levels, order, max_len_x, max_len_y,
                   stream_vec(d, x, y), gf, max_slice, &pl))
    return fail(KSIG_EUNSUPPORTED, "no row-stream kernel for d=%d, (n_levels=%d, order=%d) "
                "and column length %d", d, n_levels, order,
                max_len_x < max_len_y ? max_len_x : max_len_y);
  ksig::DpFwdArgs f{};
  f.x = {x, x_off, x_len};
  f.y = {y, y_off, y_len};
  f.d = d;
  f.pairs = pairs;
  f.n_pairs = n_pairs;
  f.kind = static_kind;
  f.n_levels = n_levels;
  f.order = order;
  f.difference = 1;
  f.out = out;
  const int64_t waves = wave_beg ? n_waves : (n_pairs + max_slice - 1) / max_slice;
  if (gf) {
    const size_t need =
        ksig::stream_gfifo_bytes(pl.pc, ksig::dp_flow_count(n_levels, order), waves);
    if (noalloc && need > workspace_bytes)
      return fail(KSIG_EINVAL, "the band FIFO of %lld waves x %d columns needs %zu bytes, more "
                  "than the %zu-byte scratch of the allocation-free entry (cut fewer slices, or "
                  "pass a workspace to ksig_signature_levels_batched)", (long long)waves, pl.pc,
                  need, workspace_bytes);
    if (workspace_bytes < need || ((uintptr_t)workspace & 15) != 0)
      return fail(KSIG_EINVAL, "workspace of %zu bytes (16-byte aligned) needed, got %zu", need,
                  workspace_bytes);
  }
  return stream_launch(pl, f, static_param, sched_order, wave_beg, (int)waves,
                       wave_beg ? 0 : max_slice, workspace, stream);
}

int ksig_signature_levels_batched(const void* x, const int64_t* x_off, const int32_t* x_len,
                                  int max_len_x, const void* y, const int64_t* y_off,
                                  const int32_t* y_len, int max_len_y, int d,
                                  const int32_t* pairs, int64_t n_pairs, int static_kind,
                                  double static_param, int n_levels, int order,
                                  const int32_t* sched_order, const int32_t* wave_beg,
                                  int n_waves, int max_slice, void* workspace,
                                  size_t workspace_bytes, double* out, void* stream) {
  return batched_common(x, x_off, x_len, max_len_x, y, y_off, y_len, max_len_y, d, pairs,
                        n_pairs, static_kind, static_param, n_levels, order, sched_order,
                        wave_beg, n_waves, max_slice, workspace, workspace_bytes, out, stream,
                        false);
}

int ksig_signature_levels_batched_noalloc(const void* x, const int64_t* x_off,
                                          const int32_t* x_len, int max_len_x, const void* y,
                                          const int64_t* y_off, const int32_t* y_len,
                                          int max_len_y, int d, const int32_t* pairs,
                                          int64_t n_pairs, int static_kind, double static_param,
                                          int n_levels, int order, const int32_t* sched_order,
                                          const int32_t* wave_beg, int n_waves, int max_slice,
                                          double* out, void* stream) {
  void* fifo = nullptr;
  const int rc = scratch_address(&fifo);
  if (rc) return rc;
  return batched_common(x, x_off, x_len, max_len_x, y, y_off, y_len, max_len_y, d, pairs,
                        n_pairs, static_kind, static_param, n_levels, order, sched_order,
                        wave_beg, n_waves, max_slice, fifo, KSIG_NOALLOC_SCRATCH_BYTES, out,
                        stream, true);
}

size_t ksig_noalloc_scratch_bytes(void) { return KSIG_NOALLOC_SCRATCH_BYTES; }

size_t ksig_signature_levels_fwd_save_workspace(int dtype, int d, int n_levels,
                                                int max_len_x, int max_len_y,
                                                int64_t n_pairs) {
  if (dtype != KSIG_F32 || n_pairs <= 0 || !ksig::dp_fwd_mfma_supported(d, n_levels)) return 0;
  int cols = max_len_x < max_len_y ? max_len_x : max_len_y;
  if (cols < 1) cols = 1;
  return (size_t)n_pairs * ksig::dp_fwd_mfma_fifo_bytes(cols, n_levels);
}

int ksig_signature_levels_fwd_save(int dtype, const void* x, const int64_t* x_off,
                                   const int32_t* x_len, int max_len_x, const void* y,
                                   const int64_t* y_off, const int32_t* y_len,
                                   int max_len_y, int d, const int32_t* pairs,
                                   int64_t n_pairs, int static_kind, double static_param,
                                   int n_levels, double* out, void* save,
                                   const int64_t* save_off, void* workspace,
                                   size_t workspace_bytes, void* stream) {
  int rc = check_common(dtype, x, y, d, pairs, n_pairs, out);
  if (rc) return rc;
  if (n_levels < 1) return fail(KSIG_EINVAL, "n_levels must be positive");
  if (!ksig::dp_supported(n_levels, 1) || !ksig::dp_backward_supported(n_levels))
    return fail(KSIG_EUNSUPPORTED, "n_levels=%d has no adjoint built", n_levels);
  ksig::DpFwdArgs a{};
  rc = static_scales(static_kind, static_param, &a.xscale, &a.yscale);
  if (rc) return rc;
  if (n_pairs == 0) return 0;
  if (n_levels > 1 && save && !save_off) return fail(KSIG_EINVAL, "null save offsets");
  a.x = {x, x_off, x_len};
  a.y = {y, y_off, y_len};
  a.d = d;
  a.pairs = pairs;
  a.n_pairs = n_pairs;
  a.kind = static_kind;
  a.n_levels = n_levels;
  a.order = 1;
  a.difference = 1;
  a.out = out;
  a.save = n_levels > 1 ? save : nullptr;
  a.save_off = save_off;
  const int cols = max_len_x < max_len_y ? max_len_x : max_len_y;
  // float32, 8 < d <= 32, with the workspace: the MFMA-Gram forward (the row
  // stream's per-lane y rows bound it by the vector-memory path there)
  const size_t need = ksig_signature_levels_fwd_save_workspace(dtype, d, n_levels, max_len_x,
                                                               max_len_y, n_pairs);
  if (need > 0 && workspace && workspace_bytes >= need && ((uintptr_t)workspace & 15) == 0 &&
      (int64_t)max_len_x * max_len_y < (int64_t(1) << 31)) {
    a.fifo = workspace;
    a.fifo_cols = ksig::dp_fwd_mfma_fifo_cols(cols < 1 ? 1 : cols);
    const size_t lds = ksig::dp_fwd_mfma_lds();
    hipError_t e = ksig::dp_fwd_mfma_launch(a, a.save != nullptr, lds, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "dp_fwd_mfma launch");
    return 0;
  }
  if (dtype == KSIG_F32) {
    rc = stream_levels(a, max_len_x, max_len_y, static_param, stream);
    if (rc != -1) return rc;  // -1: no row-stream kernel fits, use the per-pair one
  }
  return dp_levels_common(dtype, a, cols < 1 ? 1 : cols, dmax_for(d), stream);
}

// float32: the adjoint with the static Gram on MFMA; float64: the VALU adjoint
// (no float64 MFMA Gram: the float64 path is the verification dtype).
static bool bwd_uses_mfma(int dtype, int d) { return dtype == KSIG_F32 && d > 0; }

size_t ksig_signature_levels_bwd_workspace(int dtype, int d, int n_levels, int max_len_x,
                                           int max_len_y, int64_t n_pairs) {
  if (!bwd_uses_mfma(dtype, d) || n_pairs <= 0 || n_levels < 1) return 0;
  int cols = max_len_x < max_len_y ? max_len_x : max_len_y;
  if (cols < 1) cols = 1;
  return (size_t)n_pairs * ksig::dp_backward_mfma_fifo_bytes(cols, n_levels);
}

int ksig_signature_levels_bwd(int dtype, const void* x, const int64_t* x_off,
                              const int32_t* x_len, int max_len_x, const void* y,
                              const int64_t* y_off, const int32_t* y_len,
                              int max_len_y, int d, const int32_t* pairs,
                              int64_t n_pairs, int static_kind, double static_param,
                              int n_levels, const double* grad_levels, const void* save,
                              const int64_t* save_off, void* omega,
                              const int64_t* omega_off, double* grad_param,
                              void* workspace, size_t workspace_bytes, void* stream) {
  // grad_param may be NULL (no parameter gradient); omega is the required output
  int rc = check_common(dtype, x, y, d, pairs, n_pairs, reinterpret_cast<double*>(omega));
  if (rc) return rc;
  if (n_levels < 1) return fail(KSIG_EINVAL, "n_levels must be positive");
  if (!ksig::dp_backward_supported(n_levels))
    return fail(KSIG_EUNSUPPORTED, "n_levels=%d has no adjoint built", n_levels);
  ksig::DpBwdArgs a{};
  rc = static_scales(static_kind, static_param, &a.xscale, &a.yscale);
  if (rc) return rc;
  if (n_pairs == 0) return 0;
  if (!grad_levels || !omega || !omega_off || !save_off || (n_levels > 1 && !save))
    return fail(KSIG_EINVAL, "null device pointer");
  a.x = {x, x_off, x_len};
  a.y = {y, y_off, y_len};
  a.d = d;
  a.pairs = pairs;
  a.n_pairs = n_pairs;
  a.kind = static_kind;
  a.n_levels = n_levels;
  a.gout = grad_levels;
  a.save = save;
  a.save_off = save_off;
  a.omega = omega;
  a.omega_off = omega_off;
  a.gparam = grad_param;
  int cols = max_len_x < max_len_y ? max_len_x : max_len_y;
  if (cols < 1) cols = 1;
  if ((int64_t)max_len_x * max_len_y >= (int64_t(1) << 31))
    return fail(KSIG_EINVAL, "a pair's omega block must stay below 2^31 entries");
  const bool mfma = bwd_uses_mfma(dtype, d);
  if (mfma && (int64_t)max_len_x * max_len_y >= (int64_t(1) << 29))
    return fail(KSIG_EINVAL, "a pair's float32 omega block must stay below 2^31 bytes");
  if (mfma) {
    const size_t need = ksig_signature_levels_bwd_workspace(dtype, d, n_levels, max_len_x,
                                                            max_len_y, n_pairs);
    if (!workspace || workspace_bytes < need || ((uintptr_t)workspace & 15))
      return fail(KSIG_EINVAL, "workspace of %zu bytes (16-byte aligned) required, got %zu",
                  need, workspace ? workspace_bytes : (size_t)0);
    a.fifo = workspace;
    a.fifo_cols = ksig::dp_backward_mfma_fifo_cols(cols);
  }
  const size_t lds = mfma ? ksig::dp_backward_mfma_lds(d, cols, n_levels)
                          : ksig::dp_backward_lds(dtype, d, cols, n_levels);
  if (lds > max_lds_bytes())
    return fail(KSIG_ELDS, "adjoint LDS state of %zu bytes exceeds the device limit", lds);
  hipError_t e = mfma ? ksig::dp_backward_mfma_launch(a, lds, (hipStream_t)stream)
                      : ksig::dp_backward_launch(a, dtype, lds, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "dp_backward launch");
  return 0;
}

int ksig_signature_pde_grad_contract(int dtype, const void* x, const int64_t* x_off,
                                     const int32_t* x_len, int max_len_x, const void* y,
                                     const int64_t* y_off, const int32_t* y_len,
                                     int max_len_y, int d, const int32_t* pairs,
                                     int64_t n_pairs, int static_kind, double static_param,
                                     int difference, const void* grad_m,
                                     const int64_t* m_off, void* grad_x, void* grad_y,
                                     double* grad_param, void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (n_pairs < 0 || n_pairs > 0x7fffffffLL) return fail(KSIG_EINVAL, "bad n_pairs");
  if (d < 1) return fail(KSIG_EINVAL, "d must be >= 1 (got %d)", d);
  if (!(static_param > 0.0))
    return fail(KSIG_EINVAL, "static kernel parameter must be positive (got %g)", static_param);
  if (static_kind != ksig::kRBF && static_kind != ksig::kLinear)
    return fail(KSIG_EINVAL, "unknown static kernel kind %d", static_kind);
  if (n_pairs == 0 || (!grad_x && !grad_y && !grad_param)) return 0;
  if (!x || !x_off || !x_len || !y || !y_off || !y_len || !pairs || !grad_m || !m_off)
    return fail(KSIG_EINVAL, "null device pointer");
  if ((int64_t)max_len_x * max_len_y >= (int64_t(1) << 31))
    return fail(KSIG_EINVAL, "a pair's dL/dM block must stay below 2^31 entries");
  ksig::ContractArgs a{};
  a.x = {x, x_off, x_len};
  a.y = {y, y_off, y_len};
  a.d = d;
  a.pairs = pairs;
  a.n_pairs = n_pairs;
  a.omega = grad_m;
  a.omega_off = m_off;
  a.gx = grad_x;
  a.gy = grad_y;
  a.mode = difference ? 1 : 2;
  a.kind = static_kind;
  a.gparam = grad_param;
  if (static_kind == ksig::kRBF) {
    const double den = 2.0 * static_param * static_param;
    const double s = 1.0 / (den > 1e-12 ? den : 1e-12);
    a.c_cross = 2.0 * s;
    a.c_self = -2.0 * s;
    a.gsc = 1.4426950408889634 * s;
    // d/dsigma of exp(-|x-y|^2 / (2 sigma^2)) = G |x-y|^2 / sigma^3 (0 once clamped)
    a.gp_scale = den > 1e-12 ? 1.0 / (static_param * static_param * static_param) : 0.0;
  } else {
    a.c_cross = static_param;
    a.c_self = 0.0;
    a.gsc = 0.0;
    a.gp_scale = 1.0;
  }
  hipError_t e = ksig::grad_contract_launch(a, dtype == KSIG_F64, max_len_x, max_len_y,
                                            device_cus(), (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "pde grad_contract launch");
  return 0;
}

int ksig_static_grad_contract(int dtype, const void* x, const int64_t* x_off,
                              const int32_t* x_len, int max_len_x, const void* y,
                              const int64_t* y_off, const int32_t* y_len, int max_len_y,
                              int d, const int32_t* pairs, int64_t n_pairs,
                              const void* omega, const int64_t* omega_off, double c_cross,
                              double c_self, void* grad_x, void* grad_y, void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (n_pairs < 0 || n_pairs > 0x7fffffffLL) return fail(KSIG_EINVAL, "bad n_pairs");
  if (d < 1) return fail(KSIG_EINVAL, "d must be >= 1 (got %d)", d);
  if (n_pairs == 0 || (!grad_x && !grad_y)) return 0;
  if (!x || !x_off || !x_len || !y || !y_off || !y_len || !pairs || !omega || !omega_off)
    return fail(KSIG_EINVAL, "null device pointer");
  ksig::ContractArgs a{};
  a.x = {x, x_off, x_len};
  a.y = {y, y_off, y_len};
  a.d = d;
  a.pairs = pairs;
  a.n_pairs = n_pairs;
  a.omega = omega;
  a.omega_off = omega_off;
  a.c_cross = c_cross;
  a.c_self = c_self;
  a.gx = grad_x;
  a.gy = grad_y;
  hipError_t e = ksig::grad_contract_launch(a, dtype == KSIG_F64, max_len_x, max_len_y,
                                            device_cus(), (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "grad_contract launch");
  return 0;
}

int ksig_signature_levels_gram(int dtype, const void* M, int64_t n_mats, int lx,
                               int ly, int n_levels, int order, int difference,
                               double* out, void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (n_mats < 0 || n_mats > 0x7fffffffLL) return fail(KSIG_EINVAL, "bad n_mats");
  if (lx < 0 || ly < 0) return fail(KSIG_EINVAL, "negative matrix shape");
  if (n_levels < 1) return fail(KSIG_EINVAL, "n_levels must be positive");
  if (order <= 0 || order >= n_levels) order = n_levels;
  if (n_mats == 0) return 0;
  if (!M || !out) return fail(KSIG_EINVAL, "null device pointer");
  ksig::DpFwdArgs a{};
  a.kind = ksig::kGram;
  a.d = 1;
  a.n_pairs = n_mats;
  a.n_levels = n_levels;
  a.order = order;
  a.difference = difference ? 1 : 0;
  a.gram = M;
  a.gram_lx = lx;
  a.gram_ly = ly;
  a.out = out;
  return gram_levels_dispatch(dtype, a, ly < 1 ? 1 : ly, stream);
}

int ksig_signature_levels_dots(int dtype, const void* dots, const int64_t* mat_off,
                               int64_t ld, const int32_t* lens, int lx, int ly, int max_ly,
                               int static_kind, const void* hx, const int64_t* hx_off,
                               const void* hy, const int64_t* hy_off, int64_t n_mats,
                               int n_levels, int order, int difference, double* out,
                               void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (n_mats < 0 || n_mats > 0x7fffffffLL) return fail(KSIG_EINVAL, "bad n_mats");
  if (n_levels < 1) return fail(KSIG_EINVAL, "n_levels must be positive");
  if (static_kind != ksig::kRBF && static_kind != ksig::kLinear)
    return fail(KSIG_EINVAL, "unknown static kernel kind %d", static_kind);
  if (order <= 0 || order >= n_levels) order = n_levels;
  if (n_mats == 0) return 0;
  if (!dots || !out || !mat_off) return fail(KSIG_EINVAL, "null device pointer");
  if (static_kind == ksig::kRBF && (!hx || !hx_off || !hy || !hy_off))
    return fail(KSIG_EINVAL, "the RBF transform needs both norm-term arrays and offsets");
  if (!lens && (lx < 0 || ly < 0)) return fail(KSIG_EINVAL, "negative matrix shape");
  const int cols = lens ? max_ly : ly;
  if (cols < 0) return fail(KSIG_EINVAL, "negative maximum column count");
  if (ld < (lens ? 0 : ly)) return fail(KSIG_EINVAL, "row stride below the column count");
  ksig::DpFwdArgs a{};
  a.kind = ksig::kGram;
  a.d = 1;
  a.n_pairs = n_mats;
  a.n_levels = n_levels;
  a.order = order;
  a.difference = difference ? 1 : 0;
  a.gram = dots;
  a.gram_lx = lx;
  a.gram_ly = ly;
  a.gram_off = mat_off;
  a.gram_ld = ld;
  a.gram_len = lens;
  a.gram_tf = static_kind == ksig::kRBF ? ksig::kGramRbfDots : ksig::kGramPlain;
  a.gram_hx = hx;
  a.gram_hx_off = hx_off;
  a.gram_hy = hy;
  a.gram_hy_off = hy_off;
  a.out = out;
  return gram_levels_dispatch(dtype, a, cols < 1 ? 1 : cols, stream);
}

static int pde_common(int dtype, ksig::PdeFwdArgs& a, int cols, int dmax, void* stream) {
  const size_t esz = dtype == KSIG_F32 ? 4 : 8;
  if (a.dyadic < 0 || a.dyadic > 6) return fail(KSIG_EINVAL, "dyadic_order must be in [0, 6]");
  const int nc = a.difference ? cols - 1 : cols;
  long pc = (long)(nc + 1) << a.dyadic;
  if (pc <= ksig::kWave) pc = ksig::kWave + 1;  // as the kernel's band pass
  const size_t ycols = a.kind == ksig::kGram ? 0 : (size_t)cols;
  size_t lds = (ycols * dmax + ycols + (size_t)pc * 4) * esz;
  if (a.dyadic <= 1 && nc >= 0) {  // phase-aligned kernel
    const int uesz = a.edges ? 8 : (int)esz;  // U and V in float64 when saving the edges
    lds = (size_t)ksig::pde_aligned_lds((int)esz, uesz, dmax, a.kind,
                                        ksig::pde_aligned_pc(nc, a.dyadic), a.dyadic).bytes;
  }
  if (lds > max_lds_bytes())
    return fail(KSIG_ELDS, "per-pair LDS state of %zu bytes exceeds the device limit", lds);
  hipError_t e = ksig::pde_forward_launch(a, dtype, lds, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "pde_forward launch");
  return 0;
}

int ksig_signature_pde(int dtype, const void* x, const int64_t* x_off,
                       const int32_t* x_len, int max_len_x, const void* y,
                       const int64_t* y_off, const int32_t* y_len,
                       int max_len_y, int d, const int32_t* pairs,
                       int64_t n_pairs, int static_kind, double static_param,
                       int dyadic_order, int difference, double* out,
                       void* stream) {
  int rc = check_common(dtype, x, y, d, pairs, n_pairs, out);
  if (rc) return rc;
  ksig::PdeFwdArgs a{};
  rc = static_scales(static_kind, static_param, &a.xscale, &a.yscale);
  if (rc) return rc;
  if (n_pairs == 0) return 0;
  a.x = {x, x_off, x_len};
  a.y = {y, y_off, y_len};
  a.d = d;
  a.pairs = pairs;
  a.n_pairs = n_pairs;
  a.kind = static_kind;
  a.dyadic = dyadic_order;
  a.difference = difference ? 1 : 0;
  a.out = out;
  const int cols = max_len_x < max_len_y ? max_len_x : max_len_y;
  return pde_common(dtype, a, cols < 1 ? 1 : cols, d <= 4 ? 4 : dmax_for(d), stream);
}

int ksig_signature_pde_fwd_edges(int dtype, const void* x, const int64_t* x_off,
                                 const int32_t* x_len, int max_len_x, const void* y,
                                 const int64_t* y_off, const int32_t* y_len,
                                 int max_len_y, int d, const int32_t* pairs,
                                 int64_t n_pairs, int static_kind, double static_param,
                                 int dyadic_order, int difference, double* out,
                                 void* edges, const int64_t* e_off, void* stream) {
  int rc = check_common(dtype, x, y, d, pairs, n_pairs, out);
  if (rc) return rc;
  ksig::PdeFwdArgs a{};
  rc = static_scales(static_kind, static_param, &a.xscale, &a.yscale);
  if (rc) return rc;
  if (n_pairs == 0) return 0;
  if (!edges || !e_off) return fail(KSIG_EINVAL, "null edge buffer");
  if (static_kind != ksig::kRBF && static_kind != ksig::kLinear)
    return fail(KSIG_EINVAL, "unknown static kernel kind %d", static_kind);
  a.x = {x, x_off, x_len};
  a.y = {y, y_off, y_len};
  a.d = d;
  a.pairs = pairs;
  a.n_pairs = n_pairs;
  a.kind = static_kind;
  a.dyadic = dyadic_order;
  a.difference = difference ? 1 : 0;
  a.out = out;
  a.edges = edges;
  a.e_off = e_off;
  const int cols = max_len_x < max_len_y ? max_len_x : max_len_y;
  return pde_common(dtype, a, cols < 1 ? 1 : cols, d <= 4 ? 4 : dmax_for(d), stream);
}

// Refined columns (shorter side) and rows (longer side) of the PDE grid.
static void pde_refined(int max_len_x, int max_len_y, int dyadic_order, int difference,
                        long* ncr, long* nrr) {
  const int lo = max_len_x < max_len_y ? max_len_x : max_len_y;
  const int hi = max_len_x < max_len_y ? max_len_y : max_len_x;
  long nc = difference ? lo - 1 : lo, nr = difference ? hi - 1 : hi;
  if (nc < 1) nc = 1;
  if (nr < 1) nr = 1;
  *ncr = nc << dyadic_order;
  *nrr = nr << dyadic_order;
}

size_t ksig_signature_pde_bwd_workspace(int dtype, int max_len_x, int max_len_y,
                                        int dyadic_order, int difference, int64_t n_pairs) {
  if (n_pairs <= 0 || dyadic_order < 0 || dyadic_order > 6) return 0;
  long ncr, nrr;
  pde_refined(max_len_x, max_len_y, dyadic_order, difference, &ncr, &nrr);
  return (size_t)n_pairs * ksig::pde_bwd_fifo_bytes(dtype, (int)ncr, dyadic_order);
}

int ksig_signature_pde_bwd_edges(int dtype, const void* x, const int64_t* x_off,
                                 const int32_t* x_len, int max_len_x, const void* y,
                                 const int64_t* y_off, const int32_t* y_len,
                                 int max_len_y, int d, const int32_t* pairs,
                                 int64_t n_pairs, int static_kind, double static_param,
                                 int dyadic_order, int difference, const double* grad_out,
                                 const void* edges, const int64_t* e_off, void* grad_m,
                                 const int64_t* m_off, void* workspace,
                                 size_t workspace_bytes, void* stream) {
  int rc = check_common(dtype, x, y, d, pairs, n_pairs, (double*)grad_m);
  if (rc) return rc;
  ksig::PdeBwdEdgeArgs a{};
  rc = static_scales(static_kind, static_param, &a.xscale, &a.yscale);
  if (rc) return rc;
  if (dyadic_order < 0 || dyadic_order > 6)
    return fail(KSIG_EINVAL, "dyadic_order must be in [0, 6]");
  if (static_kind != ksig::kRBF && static_kind != ksig::kLinear)
    return fail(KSIG_EINVAL, "unknown static kernel kind %d", static_kind);
  if (n_pairs == 0) return 0;
  if (!grad_out || !edges || !e_off || !grad_m || !m_off)
    return fail(KSIG_EINVAL, "null device pointer");
  a.x = {x, x_off, x_len};
  a.y = {y, y_off, y_len};
  a.d = d;
  a.pairs = pairs;
  a.n_pairs = n_pairs;
  a.kind = static_kind;
  a.dyadic = dyadic_order;
  a.difference = difference ? 1 : 0;
  a.gout = grad_out;
  a.edges = edges;
  a.e_off = e_off;
  a.gm = grad_m;
  a.m_off = m_off;
  const int ly = max_len_x < max_len_y ? max_len_x : max_len_y;  // the shorter side
  long ncr, nrr;
  pde_refined(max_len_x, max_len_y, dyadic_order, difference, &ncr, &nrr);
  const size_t lds =
      ksig::pde_bwd_edges_lds(dtype, d, ly < 1 ? 1 : ly, (int)ncr, (int)nrr, dyadic_order);
  if (lds > max_lds_bytes())
    return fail(KSIG_ELDS, "adjoint LDS state of %zu bytes exceeds the device limit", lds);
  const size_t need = ksig_signature_pde_bwd_workspace(dtype, max_len_x, max_len_y,
                                                       dyadic_order, difference, n_pairs);
  if (need > 0 && (!workspace || workspace_bytes < need || ((uintptr_t)workspace & 15)))
    return fail(KSIG_EINVAL, "workspace of %zu bytes (16-byte aligned) required, got %zu", need,
                workspace ? workspace_bytes : (size_t)0);
  a.fifo = workspace;
  a.fifo_cols = ksig::pde_bwd_fifo_cols((int)ncr, dyadic_order);
  hipError_t e = ksig::pde_bwd_edges_launch(a, dtype, lds, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "pde_bwd_edges launch");
  return 0;
}

static bool pde_fused_takes(int dtype, int d, int dyadic_order, int difference,
                            int static_kind) {
  return dtype == KSIG_F32 && difference && dyadic_order >= 0 && dyadic_order <= 1 &&
         ksig::pde_fused_dmax(d) > 0 &&
         (static_kind == ksig::kRBF || static_kind == ksig::kLinear);
}

size_t ksig_signature_pde_bwd_fused_workspace(int dtype, int d, int max_len_x, int max_len_y,
                                              int dyadic_order, int difference,
                                              int64_t n_pairs) {
  if (n_pairs <= 0 || d < 1 || !pde_fused_takes(dtype, d, dyadic_order, difference, ksig::kRBF))
    return 0;
  const int lo = max_len_x < max_len_y ? max_len_x : max_len_y;
  const int nc = lo - 1 < 1 ? 1 : lo - 1;
  return (size_t)n_pairs * ksig::pde_fused_pc(nc, dyadic_order) * ksig::pde_fused_slot_bytes(d);
}

int ksig_signature_pde_bwd_fused(int dtype, const void* x, const int64_t* x_off,
                                 const int32_t* x_len, int max_len_x, const void* y,
                                 const int64_t* y_off, const int32_t* y_len,
                                 int max_len_y, int d, const int32_t* pairs,
                                 int64_t n_pairs, int static_kind, double static_param,
                                 int dyadic_order, int difference, const double* grad_out,
                                 const void* edges, const int64_t* e_off, void* grad_x,
                                 void* grad_y, double* grad_param_pairs, void* workspace,
                                 size_t workspace_bytes, void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (n_pairs < 0 || n_pairs > 0x7fffffffLL) return fail(KSIG_EINVAL, "bad n_pairs");
  if (d < 1) return fail(KSIG_EINVAL, "d must be >= 1 (got %d)", d);
  ksig::PdeBwdFusedArgs a{};
  int rc = static_scales(static_kind, static_param, &a.xscale, &a.yscale);
  if (rc) return rc;
  if (dyadic_order < 0 || dyadic_order > 6)
    return fail(KSIG_EINVAL, "dyadic_order must be in [0, 6]");
  if (!pde_fused_takes(dtype, d, dyadic_order, difference, static_kind))
    return fail(KSIG_EUNSUPPORTED,
                "fused PDE adjoint: float32, difference=1, dyadic order 0/1, d <= 8 only");
  if (n_pairs == 0) return 0;
  if (!x || !x_off || !x_len || !y || !y_off || !y_len || !pairs || !grad_out || !edges ||
      !e_off)
    return fail(KSIG_EINVAL, "null device pointer");
  const size_t need = ksig_signature_pde_bwd_fused_workspace(dtype, d, max_len_x, max_len_y,
                                                             dyadic_order, difference, n_pairs);
  if (!workspace || workspace_bytes < need || ((uintptr_t)workspace & 15))
    return fail(KSIG_EINVAL, "workspace of %zu bytes (16-byte aligned) required, got %zu", need,
                workspace ? workspace_bytes : (size_t)0);
  a.x = {x, x_off, x_len};
  a.y = {y, y_off, y_len};
  a.d = d;
  a.pairs = pairs;
  a.n_pairs = n_pairs;
  a.kind = static_kind;
  a.dyadic = dyadic_order;
  a.gout = grad_out;
  a.edges = edges;
  a.e_off = e_off;
  a.gx = grad_x;
  a.gy = grad_y;
  a.gp = grad_param_pairs;
  if (static_kind == ksig::kRBF) {
    const double den = 2.0 * static_param * static_param;
    const double sc = 1.0 / (den > 1e-12 ? den : 1e-12);
    a.c_cross = 2.0 * sc;
    a.c_self = -2.0 * sc;
    // d/dsigma of exp(-|x-y|^2 / (2 sigma^2)) = G |x-y|^2 / sigma^3 (0 once clamped)
    a.gp_scale = den > 1e-12 ? 1.0 / (static_param * static_param * static_param) : 0.0;
  } else {
    a.c_cross = static_param;
    a.c_self = 0.0;
    a.gp_scale = 1.0;
  }
  const int lo = max_len_x < max_len_y ? max_len_x : max_len_y;
  a.fifo = workspace;
  a.fifo_cols = ksig::pde_fused_pc(lo - 1 < 1 ? 1 : lo - 1, dyadic_order);
  const size_t lds = ksig::pde_fused_lds(dtype, d, dyadic_order);
  hipError_t e = ksig::pde_bwd_fused_launch(a, dtype, lds, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "pde_bwd_fused launch");
  return 0;
}

int ksig_signature_pde_gram(int dtype, const void* M, int64_t n_mats, int lx,
                            int ly, int dyadic_order, int difference,
                            double* out, void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (n_mats < 0 || n_mats > 0x7fffffffLL) return fail(KSIG_EINVAL, "bad n_mats");
  if (lx < 0 || ly < 0) return fail(KSIG_EINVAL, "negative matrix shape");
  if (n_mats == 0) return 0;
  if (!M || !out) return fail(KSIG_EINVAL, "null device pointer");
  ksig::PdeFwdArgs a{};
  a.kind = ksig::kGram;
  a.d = 1;
  a.n_pairs = n_mats;
  a.dyadic = dyadic_order;
  a.difference = difference ? 1 : 0;
  a.gram = M;
  a.gram_lx = lx;
  a.gram_ly = ly;
  a.out = out;
  return pde_common(dtype, a, ly < 1 ? 1 : ly, 8, stream);
}

static int gak_common(int dtype, ksig::GakArgs& a, int cols, void* stream) {
  const size_t lds = ksig::gak_lds_bytes(dtype, a.kind, a.d, cols);
  if (lds > max_lds_bytes())
    return fail(KSIG_ELDS, "per-pair LDS state of %zu bytes exceeds the device limit", lds);
  hipError_t e = ksig::gak_forward_launch(a, dtype, lds, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gak_forward launch");
  return 0;
}

int ksig_global_align_log(int dtype, const void* x, const int64_t* x_off,
                          const int32_t* x_len, int max_len_x, const void* y,
                          const int64_t* y_off, const int32_t* y_len, int max_len_y,
                          int d, const int32_t* pairs, int64_t n_pairs, int static_kind,
                          double static_param, double* out, void* stream) {
  int rc = check_common(dtype, x, y, d, pairs, n_pairs, out);
  if (rc) return rc;
  ksig::GakArgs a{};
  rc = static_scales(static_kind, static_param, &a.xscale, &a.yscale);
  if (rc) return rc;
  if (max_len_x < 0 || max_len_y < 0) return fail(KSIG_EINVAL, "negative max length");
  if (n_pairs == 0) return 0;
  a.x = {x, x_off, x_len};
  a.y = {y, y_off, y_len};
  a.d = d;
  a.pairs = pairs;
  a.n_pairs = n_pairs;
  a.kind = static_kind;
  a.out = out;
  // the shorter side of a pair (the LDS side) is at most the smaller maximum,
  // the longer (the rows) at most the larger
  const int cols = max_len_x < max_len_y ? max_len_x : max_len_y;
  a.rows_max = max_len_x > max_len_y ? max_len_x : max_len_y;
  return gak_common(dtype, a, cols < 1 ? 1 : cols, stream);
}

int ksig_global_align_log_gram(int dtype, const void* M, int64_t n_mats, int lx, int ly,
                               double* out, void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (n_mats < 0 || n_mats > 0x7fffffffLL) return fail(KSIG_EINVAL, "bad n_mats");
  if (lx < 0 || ly < 0) return fail(KSIG_EINVAL, "negative matrix shape");
  if (n_mats == 0) return 0;
  if (!M || !out) return fail(KSIG_EINVAL, "null device pointer");
  ksig::GakArgs a{};
  a.kind = ksig::kGram;
  a.d = 1;
  a.n_pairs = n_mats;
  a.gram = M;
  a.gram_lx = lx;
  a.gram_ly = ly;
  a.rows_max = lx;
  a.out = out;
  return gak_common(dtype, a, ly < 1 ? 1 : ly, stream);
}

static int rws_common(int dtype, ksig::RwsArgs& a, int lx, void* stream) {
  if (a.n_seq < 0 || a.n_seq * (int64_t)(a.n_groups > 0 ? a.n_groups : 1) > 0x7fffffffLL)
    return fail(KSIG_EINVAL, "bad n_seq");
  if (a.n_warps < 0 || a.warp_rows < 0) return fail(KSIG_EINVAL, "negative warp counts");
  if (a.n_seq == 0 || a.n_warps == 0) return 0;
  if (a.n_groups < 1) return fail(KSIG_EINVAL, "n_groups must be positive");
  if (!a.warp_off || !a.groups || !a.out) return fail(KSIG_EINVAL, "null device pointer");
  const size_t lds = ksig::rws_lds_bytes(dtype, a.gram != nullptr, a.d, lx < 1 ? 1 : lx);
  if (lds > max_lds_bytes())
    return fail(KSIG_ELDS, "per-wave LDS state of %zu bytes exceeds the device limit", lds);
  hipError_t e = ksig::rws_forward_launch(a, dtype, lds, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "rws_forward launch");
  return 0;
}

int ksig_random_warping_series(int dtype, const void* x, const int64_t* x_off,
                               const int32_t* x_len, int max_len_x, int64_t n_seq, int d,
                               const void* warp, const int32_t* warp_off, int n_warps,
                               int warp_rows, const int32_t* groups, int n_groups,
                               double* out, void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (d < 1) return fail(KSIG_EINVAL, "d must be >= 1 (got %d)", d);
  if (d > 32) return fail(KSIG_EUNSUPPORTED, "d = %d > 32 is not built into the fused kernels", d);
  if (max_len_x < 0) return fail(KSIG_EINVAL, "negative max length");
  if (n_seq > 0 && n_warps > 0 && (!x || !x_off || !x_len || !warp))
    return fail(KSIG_EINVAL, "null device pointer");
  ksig::RwsArgs a{};
  a.x = {x, x_off, x_len};
  a.n_seq = n_seq;
  a.d = d;
  a.warp = warp;
  a.warp_off = warp_off;
  a.n_warps = n_warps;
  a.warp_rows = warp_rows;
  a.groups = groups;
  a.n_groups = n_groups;
  a.out = out;
  return rws_common(dtype, a, max_len_x, stream);
}

int ksig_random_warping_series_gram(int dtype, const void* D, int64_t n_seq, int lx,
                                    int warp_rows, const int32_t* warp_off, int n_warps,
                                    const int32_t* groups, int n_groups, double* out,
                                    void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (lx < 0) return fail(KSIG_EINVAL, "negative matrix shape");
  if (n_seq > 0 && n_warps > 0 && !D) return fail(KSIG_EINVAL, "null device pointer");
  ksig::RwsArgs a{};
  a.n_seq = n_seq;
  a.d = 1;
  a.warp_off = warp_off;
  a.n_warps = n_warps;
  a.warp_rows = warp_rows;
  a.groups = groups;
  a.n_groups = n_groups;
  a.gram = D;
  a.gram_lx = lx;
  a.out = out;
  return rws_common(dtype, a, lx, stream);
}

int ksig_lowrank_levels(int dtype, const void* z, int n_levels, int order, int64_t n_seq,
                        int steps, int n_comp, double* out, void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (n_levels < 1) return fail(KSIG_EINVAL, "n_levels must be positive");
  if (n_levels > 8) return fail(KSIG_EUNSUPPORTED, "n_levels = %d > 8 is not built", n_levels);
  // ksig/algorithms.py:165: a non-positive order, or one of n_levels or
  // more, is the full recursion
  if (order <= 0 || order > n_levels) order = n_levels;
  if (n_seq < 0 || steps < 0 || n_comp < 0) return fail(KSIG_EINVAL, "negative shape");
  if (n_seq * (int64_t)n_comp > 0x7fffffffLL * 256LL) return fail(KSIG_EINVAL, "too many sequences");
  if (n_seq == 0 || n_comp == 0) return 0;
  if (!out || (steps > 0 && !z)) return fail(KSIG_EINVAL, "null device pointer");
  hipError_t e = ksig::lowrank_scan_launch(dtype, z, n_levels, order, n_seq, steps, n_comp,
                                          out, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "lowrank_scan launch");
  return 0;
}

static int lowrank_diagonal(int dtype, const void* du, int n_levels, int order, int q,
                            int64_t n_seq, int steps, int half, double* out, void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (n_levels < 1) return fail(KSIG_EINVAL, "n_levels must be positive");
  if (q < 1) return fail(KSIG_EINVAL, "internal_size must be positive");
  if (order <= 0 || order > n_levels) order = n_levels;  // ksig/algorithms.py:165
  if (!ksig::diag_fits(q, n_levels, order))
    return fail(KSIG_EUNSUPPORTED,
                "diagonal scan with internal_size %d, %d levels, order %d is not built", q,
                n_levels, order);
  if (n_seq < 0 || steps < 0 || half < 0) return fail(KSIG_EINVAL, "negative shape");
  if (n_seq * (int64_t)half > 0x7fffffffLL * 256LL) return fail(KSIG_EINVAL, "too many sequences");
  if (n_seq == 0 || half == 0) return 0;
  if (!out || (steps > 0 && !du)) return fail(KSIG_EINVAL, "null device pointer");
  hipError_t e = ksig::lowrank_diag_scan_launch(dtype, du, q, n_levels, order, n_seq, steps, half,
                                               out, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "lowrank_diag_scan launch");
  return 0;
}

int ksig_lowrank_levels_diag(int dtype, const void* du, int n_levels, int64_t n_seq,
                             int steps, int half, double* out, void* stream) {
  if (n_levels > 5) return fail(KSIG_EUNSUPPORTED, "n_levels = %d > 5 is not built", n_levels);
  return lowrank_diagonal(dtype, du, n_levels, 1, 2, n_seq, steps, half, out, stream);
}

int ksig_lowrank_levels_diagonal(int dtype, const void* du, int n_levels, int order,
                                 int internal_size, int64_t n_seq, int steps, int half,
                                 double* out, void* stream) {
  return lowrank_diagonal(dtype, du, n_levels, order, internal_size, n_seq, steps, half, out,
                          stream);
}

int ksig_lowrank_levels_sketch(int dtype, const void* du, int n_levels, int order, int64_t n_seq,
                               int steps, int width, const int32_t* hash_idx,
                               const void* hash_bit, int n_comp, double* out, void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (n_levels < 1) return fail(KSIG_EINVAL, "n_levels must be positive");
  if (n_levels > 8) return fail(KSIG_EUNSUPPORTED, "n_levels = %d > 8 is not built", n_levels);
  if (order <= 0 || order > n_levels) order = n_levels;  // ksig/algorithms.py:165
  if (n_seq < 0 || steps < 0 || width < 0) return fail(KSIG_EINVAL, "negative shape");
  if (n_comp < 1) return fail(KSIG_EINVAL, "n_comp must be positive");
  if (n_seq > 0x7fffffffLL) return fail(KSIG_EINVAL, "too many sequences");
  if (n_comp > ksig::lowrank_sketch_max_comp())
    return fail(KSIG_EUNSUPPORTED, "n_comp = %d > %d is not built", n_comp,
                ksig::lowrank_sketch_max_comp());
  const size_t limit = max_lds_bytes();
  const int ng = ksig::lowrank_sketch_groups(n_levels, order, n_comp, width, limit);
  const size_t lds = ksig::lowrank_sketch_lds(n_levels, order, n_comp, width, ng);
  if (lds > limit)
    return fail(KSIG_EUNSUPPORTED, "sketch scan needs %zu bytes of LDS (> %zu)", lds, limit);
  if (n_seq == 0) return 0;
  if (!out || (steps > 0 && width > 0 && (!du || !hash_idx || !hash_bit)))
    return fail(KSIG_EINVAL, "null device pointer");
  hipError_t e = ksig::lowrank_sketch_launch(dtype, du, n_levels, order, n_seq, steps, width,
                                            hash_idx, hash_bit, n_comp, ng, out,
                                            (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "lowrank_sketch launch");
  return 0;
}

int ksig_rff_increments(int dtype, const void* proj, int64_t n_seq, int len, int half, int ld,
                        int difference, void* out, void* stream) {
  if (dtype != KSIG_F32 && dtype != KSIG_F64) return fail(KSIG_EINVAL, "bad dtype %d", dtype);
  if (n_seq < 0 || len < 0 || half < 0) return fail(KSIG_EINVAL, "negative shape");
  if (ld < half) return fail(KSIG_EINVAL, "row stride %d below the row length %d", ld, half);
  const int steps = difference ? len - 1 : len;
  if (n_seq == 0 || half == 0 || steps <= 0) return 0;
  if (!proj || !out) return fail(KSIG_EINVAL, "null device pointer");
  hipError_t e = ksig::rff_diff_launch(dtype, proj, n_seq, len, half, ld, difference ? 1 : 0,
                                      out, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "rff_diff launch");
  return 0;
}

}  // extern "C"
