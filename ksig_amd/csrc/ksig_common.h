// Shared device helpers for the signature-kernel HIP library (gfx950, wave64).
//
// Every hot kernel in this library is a skewed systolic sweep over one
// sequence pair: lane i of a 64-lane wavefront owns row (64*band + i) of the
// doubly-differenced Gram grid and walks the columns left to right, one column
// per step, one step behind lane i-1.  Values that flow down a column (column
// prefix sums, Gram values of the row above, PDE solution values) move from
// lane i-1 to lane i with a DPP wave_shr:1 at every step; lane 63's outputs are
// parked in an LDS FIFO (one slot per column) and read back by lane 0 when it
// starts the next 64-row band.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ksig {

constexpr int kWave = 64;

enum StaticKind : int { kRBF = 0, kLinear = 1, kGram = 2 };
// How a caller-matrix kernel forms static-kernel values from the stored ones
// (DpFwdArgs::gram_tf); zero, the value of a zero-initialised argument block,
// is the plain matrix.
enum GramTransform : int { kGramPlain = 0, kGramRbfDots = 1 };

// ---------------------------------------------------------------------------
// s_waitcnt vmcnt(N), expcnt and lgkmcnt not waited for (gfx9-family
// encoding: vmcnt in bits [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8]).
// These parts retire a wave's vector memory loads and stores on one counter in
// the order they were issued, so once vmcnt <= N every vector memory
// instruction of the wave older than its N most recent has completed.
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N <= 63, "vmcnt holds 6 bits");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Band hand-over FIFOs in global memory: lane 63 stores a slot and lane 0
// loads it back `steps` or more steps later, and every step issues at least
// `per_step` vector memory instructions.  vmcnt(fifo_wait_count(...)) right
// before lane 0's load therefore completes the store first (two steps of slack
// for the compiler's placement of accesses inside an unrolled pair of steps);
// a wave rarely has that many instructions in flight, so the wait is
// normally free.
__host__ __device__ constexpr int fifo_wait_count(int steps, int per_step) {
  const int n = (steps - 2) * per_step - 1;
  return n > 62 ? 62 : (n < 0 ? 0 : n);
}

// ---------------------------------------------------------------------------
// Lane shift: lane i receives lane i-1's value; lane 0 receives `fill`.
// DPP wave_shr:1 (dpp_ctrl 0x138) with bound_ctrl off keeps the `old` operand
// in lane 0, so the FIFO value for lane 0 rides in as `old`.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float shift_down(float v, float fill) {
  return __int_as_float(__builtin_amdgcn_update_dpp(
      __float_as_int(fill), __float_as_int(v), 0x138, 0xf, 0xf, false));
}

// The same shift when `fill` is zero on every lane but lane 0 (a load that
// only lane 0 keeps in range): lane 0 gets 0 + fill, lane i > 0 gets v of
// lane i-1 + 0.  The compiler folds the DPP move into the add (one
// v_add_f32_dpp per value); the adds are exact.
__device__ __forceinline__ float shift_add(float v, float fill0) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true)) + fill0;
}

__device__ __forceinline__ double shift_down(double v, double fill) {
  const int2 a = *reinterpret_cast<const int2*>(&v);
  const int2 f = *reinterpret_cast<const int2*>(&fill);
  int2 r;
  r.x = __builtin_amdgcn_update_dpp(f.x, a.x, 0x138, 0xf, 0xf, false);
  r.y = __builtin_amdgcn_update_dpp(f.y, a.y, 0x138, 0xf, 0xf, false);
  return *reinterpret_cast<const double*>(&r);
}

// Raw buffer descriptor (word 3 = 0x00020000: 32-bit data, no swizzle or
// stride) over `bytes` bytes at p.  A lane whose offset is at or past `bytes`
// reads zero and its store is dropped, so an offset of kBufOob masks a lane's
// store without a branch.  `p` must be wave-uniform.
constexpr unsigned kBufOob = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Full-wave sum in double (used once per pair for the level sums).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// v_exp_f32 (about 1 ulp); results below 2^-126 flush to zero, which only
// affects static-kernel values that are already negligible.
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ double fast_exp2(double x) { return exp2(x); }

// ---------------------------------------------------------------------------
// Sequence-set descriptor: packed rows [total_rows, d] plus per-sequence row
// offsets and lengths (dense [n, L, d] arrays are the case off = a*L, len = L).
// ---------------------------------------------------------------------------
struct SeqSet {
  const void* data;
  const int64_t* off;
  const int32_t* len;
};

}  // namespace ksig
