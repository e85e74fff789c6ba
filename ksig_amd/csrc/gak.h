// Global alignment kernel (log space): launch arguments and entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ksig_common.h"

namespace ksig {

struct GakArgs {
  SeqSet x, y;
  int d;
  const int32_t* pairs;  // [n_pairs][2]
  int64_t n_pairs;
  int kind;              // kRBF, kLinear or kGram
  double xscale, yscale; // static-kernel scales folded into the inputs
  const void* gram;      // kGram: [n_pairs][gram_lx][gram_ly] static-kernel values
  int gram_lx, gram_ly;
  int rows_max;          // the longest row side of any pair (the longer sequence)
  double* out;           // [n_pairs] log GA values
};

// Dynamic LDS of one pair's wave for a column side of `cols` points.
size_t gak_lds_bytes(int dtype, int kind, int d, int cols);
hipError_t gak_forward_launch(const GakArgs& a, int dtype, size_t lds, hipStream_t s);

}  // namespace ksig
