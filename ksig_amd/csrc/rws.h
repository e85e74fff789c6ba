// Random Warping Series (DTW against warping series): launch arguments and
// entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "ksig_common.h"

namespace ksig {

struct RwsArgs {
  SeqSet x;               // sequences (the columns of each DTW grid)
  int64_t n_seq;
  int d;
  const void* warp;       // warping-series points [warp_rows][d], dtype of x
  const int32_t* warp_off;  // [n_warps + 1] first row of each series (+ total)
  int n_warps;
  int warp_rows;          // warp_off[n_warps]
  const int32_t* groups;  // [n_groups + 1] series split points: one wave per (sequence, group)
  int n_groups;
  const void* gram;       // optional distances [n_seq][gram_lx][warp_rows] (replaces x, warp)
  int gram_lx;
  double* out;            // [n_seq][n_warps]
};

__device__ __forceinline__ float fminf_t(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double fminf_t(double a, double b) { return fmin(a, b); }

// Dynamic LDS of one wave for sequences of at most `lx` points.
size_t rws_lds_bytes(int dtype, bool gram, int d, int lx);
hipError_t rws_forward_launch(const RwsArgs& a, int dtype, size_t lds, hipStream_t s);

}  // namespace ksig
