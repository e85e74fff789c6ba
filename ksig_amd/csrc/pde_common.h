// Shared pieces of the phase-aligned signature-PDE forward and adjoint
// (dyadic order 0 and 1): the static-kernel value of one original cell edge,
// the scheme's coefficients, straight-line loads of a y row through a buffer
// descriptor, and the forward's LDS layout.  The forward that saves the
// edges and the adjoint that rebuilds U from them must form bit-identical
// static-kernel values and coefficients (the rebuild inverts the very
// coefficients the forward used), so both evaluate them here.
#pragma once

#include <type_traits>

#include "ksig_common.h"

namespace ksig {

__device__ __forceinline__ float pde_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double pde_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Row-side constants.  Every scale factor of the static kernel rides on the x
// row, so a y row is used as stored:
//   RBF    G = exp2(<xm, y> + nx - yn),  xm = 2 s_x s_y x,
//          nx = -(s_x^2 |x|^2 + 2 lam),  yn = s_y^2 |y|^2,
//   linear G = <xm, y>,                  xm = s_x s_y x / 4^lam,
// with s_x = s_y = sqrt(log2(e) / (2 sigma^2)) for RBF (exp2 of the scaled
// squared distance) and s_x = c, s_y = 1 for the linear kernel.  The factor
// 1/4^lam of the dyadic refinement (M' = M / 4^lam) is folded into G: into the
// exponent for RBF, into xm for the linear kernel.
template <typename T, int DMAX, bool RBF>
__device__ __forceinline__ void pde_xrow_prep(const T* raw, T* xm, T& nx, T xs, T ys, int lam) {
  T n2 = T(0);
  const T mul = RBF ? T(2) * xs * ys : xs * ys / T(1 << (2 * lam));
#pragma unroll
  for (int e = 0; e < DMAX; ++e) {
    const T v = raw[e] * xs;
    n2 = pde_fma(v, v, n2);
    xm[e] = raw[e] * mul;
  }
  nx = RBF ? -(n2 + T(2 * lam)) : T(0);
}

// Static-kernel value of one (x row, y row) from the prepared row and the y
// norm term; the same operation order wherever it is evaluated.
template <typename T, int DMAX, bool RBF>
__device__ __forceinline__ T pde_gval(const T* xm, T nx, const T* yv, T yn) {
  if constexpr (RBF) {
    T acc = nx - yn;
#pragma unroll
    for (int e = 0; e < DMAX; ++e) acc = pde_fma(xm[e], yv[e], acc);
    return fast_exp2(acc);
  } else {
    T acc = xm[0] * yv[0];
#pragma unroll
    for (int e = 1; e < DMAX; ++e) acc = pde_fma(xm[e], yv[e], acc);
    return acc;
  }
}

// Coefficients of the scheme for M' = m: a = m/2 + m^2/12, b = m^2/12.
template <typename T>
__device__ __forceinline__ void pde_coef(T m, T& ca, T& cb) {
  cb = (m * m) * T(1.0 / 12.0);
  ca = pde_fma(T(0.5), m, cb);
}

// y row at byte offset voff of a buffer descriptor (rows past either end of
// the sequence read zero).  VEC: d == DMAX and 16-byte aligned rows, one
// 16-byte load per quad; else one load per feature, features past d reading
// the row's last feature (their x weights are zero).  Straight-line code.
template <typename T, int DMAX, bool VEC>
__device__ __forceinline__ void pde_load_y(T* yv, __amdgpu_buffer_rsrc_t rs, int voff, int d) {
  if constexpr (VEC) {
    constexpr int NQ = DMAX * (int)sizeof(T) / 16;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const u32x4 t = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, 16 * q, 0);
      if constexpr (sizeof(T) == 4) {
        yv[4 * q + 0] = __uint_as_float(t.x);
        yv[4 * q + 1] = __uint_as_float(t.y);
        yv[4 * q + 2] = __uint_as_float(t.z);
        yv[4 * q + 3] = __uint_as_float(t.w);
      } else {
        yv[2 * q + 0] = __hiloint2double((int)t.y, (int)t.x);
        yv[2 * q + 1] = __hiloint2double((int)t.w, (int)t.z);
      }
    }
  } else {
#pragma unroll
    for (int e = 0; e < DMAX; ++e) {
      const int f = e < d ? e : d - 1;  // wave-uniform
      if constexpr (sizeof(T) == 4) {
        yv[e] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, voff, 4 * f, 0));
      } else {
        typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
        const u32x2 t = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, 8 * f, 0);
        yv[e] = __hiloint2double((int)t.y, (int)t.x);
      }
    }
  }
}

// LDS layout of the phase-aligned forward for a band pass of pc slots (host
// and device agree; regions 16-byte aligned): FIFO of (U, V) per slot, FIFO
// of G per slot, each lane's prepared x rows of its band and their norm
// terms, and the y norm terms of the band pass's original columns with 64
// leading zeros (the lanes' fetches before they start).
struct PdeAlignedLds {
  int uv, g, x, xn, yn, bytes;
};
__host__ __device__ inline int pde_al16(long b) { return (int)((b + 15) & ~15L); }
__host__ __device__ inline PdeAlignedLds pde_aligned_lds(int esz, int uesz, int dmax, int kind,
                                                          int pc, int lam) {
  const int no = 4 >> lam;
  const int ncols = pc >> lam;
  PdeAlignedLds L;
  L.uv = 0;
  L.g = pde_al16(2L * pc * uesz);
  L.x = L.g + pde_al16((long)pc * esz);
  L.xn = L.x + (kind == kGram ? 0 : pde_al16((long)kWave * no * dmax * esz));
  L.yn = L.xn + (kind == kRBF ? pde_al16((long)kWave * no * esz) : 0);
  L.bytes = L.yn + (kind == kRBF ? pde_al16((long)(kWave + ncols + 2) * esz) : 0);
  return L;
}

}  // namespace ksig
