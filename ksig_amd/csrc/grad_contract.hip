// Chain rule through the static kernel: contraction of a per-pair Gram
// adjoint with the inputs.
//
// For every pair p = (i, j) the adjoint kernels leave omega_p[a, b]
// (len_x[i] x len_y[j], row-major at omega_off[p]).  This pass adds
//   gx[x_i, a] += c_cross * sum_b omega[a, b] y_b + c_self * (sum_b omega[a, b]) x_a
//   gy[y_j, b] += c_cross * sum_a omega[a, b] x_a + c_self * (sum_a omega[a, b]) y_b
// which covers both static kernels of the signature kernels (RBF:
// c_cross = 1/sigma^2, c_self = -c_cross, omega = dL/dG * G; linear:
// c_cross = scale, c_self = 0, omega = dL/dG).
//
// float32: one workgroup (4 waves) per (run of consecutive pairs, 64-row
// block of x).  A 64 x 64 omega tile and the matching x / y row blocks are
// staged in LDS with coalesced loads; both products run on
// v_mfma_f32_16x16x4f32 (K = 64 per tile), the
// x-side accumulator stays in registers while consecutive pairs share their x
// sequence (pair lists in x-major order), and results are added to gx / gy
// with hardware float atomics.  float64 (the verification dtype) and pairs
// of 2^29 or more omega entries: a plain VALU reduction.
//
// For the signature-PDE adjoint the same kernel builds omega itself, tile by
// tile, from the adjoint's dL/dM (original cells, the adjoint's orientation):
// omega = ΔΔ^T(dL/dM) * G (RBF) or ΔΔ^T(dL/dM) (linear), from a 65 x 65
// haloed dL/dM tile in LDS and the 64 x 64 tile of <x_a, y_b> on MFMA, and it
// accumulates dL/dparam (RBF: sum omega |x-y|^2 / sigma^3, linear:
// sum omega <x, y>) as well.
#include "dp_forward.h"

namespace ksig {


namespace {

typedef float cf4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;
constexpr int kOS = kTile + 1;  // omega tile row stride in LDS (words)
constexpr int kRawBuf = 0x00020000;  // buffer descriptor word 3: raw 32-bit data

// MODE 0: omega given per pair.  MODE 1 / 2: omega built tile by tile from the
// signature-PDE adjoint's dL/dM (with / without the ΔΔ of difference=True),
// times the static kernel recomputed on MFMA (KIND).
template <int NB, int MODE, int KIND>  // NB: feature blocks of 16 (d <= 16 * NB)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODE == 0 ? 4 : 3))) void
contract_f32_kernel(const ContractArgs a) {
  constexpr int FS = 16 * NB;  // LDS row stride of the x / y row blocks
  constexpr int kGS = kTile + 2;  // dL/dM halo tile: 65 x 65 at row stride 66
  constexpr int OMW = MODE == 0 ? kTile * kOS : (kTile + 1) * kGS;
  constexpr int NRM = MODE == 0 ? 1 : kTile;
  __shared__ float om[OMW];          // omega tile [a][b] (MODE 1/2: first the dL/dM tile)
  __shared__ float xs[kTile * FS];   // x rows a0 .. a0+63 of the current x sequence
  __shared__ float ys[kTile * FS];   // y rows b0 .. b0+63 of the current tile
  __shared__ float xnorm[NRM];       // |x|^2 and |y|^2 of the staged rows (MODE 1/2)
  __shared__ float ynorm[NRM];
  __shared__ double gpart[4];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int c16 = lane & 15, grp = lane >> 4;
  const int a0 = blockIdx.y * kTile;
  const int64_t pb = (int64_t)blockIdx.x * a.group;
  const int64_t pe = pb + a.group < a.n_pairs ? pb + a.group : a.n_pairs;
  const float* X = static_cast<const float*>(a.x.data);
  const float* Y = static_cast<const float*>(a.y.data);
  const float* Om = static_cast<const float*>(a.omega);
  float* GX = static_cast<float*>(a.gx);
  float* GY = static_cast<float*>(a.gy);
  const int d = a.d;
  const float cc = float(a.c_cross), cself = float(a.c_self);

  // feature padding columns stay zero (staging writes f < d only)
  for (int e = tid; e < kTile * FS; e += 256) {
    xs[e] = 0.f;
    ys[e] = 0.f;
  }
  // Row-block staging: rows [row0, row0 + nrows) of a packed [*, d] array
  // (a contiguous block) into dst.  Raw buffer loads over exactly that block
  // return zero past its end, so rows beyond nrows come out zero; thread tid
  // moves elements e = tid + 256k, all loads in flight before the stores.
  constexpr int KS = 4 * NB;  // ceil(64 * 16NB / 256)
  int st_idx[KS];
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int e = tid + 256 * k;
    const int r = e / d;
    st_idx[k] = e < kTile * d ? r * FS + (e - r * d) : -1;
  }
  auto stage_rows = [&](float* dst, const float* src, int64_t row0, int nrows) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void*)(src + row0 * d), (short)0, nrows * d * 4, kRawBuf);
    float v[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k)
      v[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, tid * 4 + 1024 * k, 0, 0));
#pragma unroll
    for (int k = 0; k < KS; ++k)
      if (st_idx[k] >= 0) dst[st_idx[k]] = v[k];
  };

  cf4 gxa[NB];
  float rsum = 0.f;  // omega row sum, row 16w + c16, over columns = grp mod 4
  double gp_acc = 0.0;  // MODE 1/2: this thread's share of dL/dparam (unscaled)
  bool xnorm_due = false;
  const float gsc = float(a.gsc);  // RBF exponent factor log2(e) / max(2 sigma^2, 1e-12)
  int cur = -1;
  int curA = 0;
  int64_t cur_off = 0;

  auto flush_gx = [&]() {
    float r = rsum;
    r += __shfl_xor(r, 16, kWave);
    r += __shfl_xor(r, 32, kWave);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float rq = __shfl(r, 4 * grp + q, kWave);
      const int rl = 16 * w + 4 * grp + q;
      if (a0 + rl < curA) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          const int f = 16 * nb + c16;
          if (f < d)
            unsafeAtomicAdd(GX + (cur_off + a0 + rl) * d + f,
                            cc * gxa[nb][q] + cself * rq * xs[rl * FS + f]);
        }
      }
    }
  };

  for (int64_t p = pb; p < pe; ++p) {
    const int i = a.pairs[2 * p], j = a.pairs[2 * p + 1];
    const int A = a.x.len[i], B = a.y.len[j];
    if (a0 >= A || B <= 0) continue;  // uniform over the workgroup
    const int64_t xo = a.x.off[i], yo = a.y.off[j];
    if (i != cur) {
      if (GX && cur >= 0) flush_gx();
      __syncthreads();  // xs readers done
      stage_rows(xs, X, xo + a0, A - a0 < kTile ? A - a0 : kTile);
      xnorm_due = true;
      cur = i;
      curA = A;
      cur_off = xo;
      rsum = 0.f;
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) gxa[nb] = cf4{0.f, 0.f, 0.f, 0.f};
    }
    // the descriptor must sit in scalar registers: make the offset provably
    // wave-uniform (it is: p is)
    const int64_t wo = a.omega_off[p];
    const int64_t wou = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)wo) |
                        ((int64_t)__builtin_amdgcn_readfirstlane((int)(wo >> 32)) << 32);
    // MODE 1/2: dL/dM of this pair is [A - sh][B - sh] in the caller's
    // orientation, stored in the adjoint's orientation (rows = the longer
    // sequence): element (a', b') at a' * rsg + b' * csg.
    constexpr int sh = MODE == 1 ? 1 : 0;
    const int An = A - sh, Bn = B - sh;
    const int nck = B > A ? An : Bn;
    const int rsg = B > A ? 1 : nck, csg = B > A ? nck : 1;
    const int wbytes = MODE == 0 ? A * B * 4 : (An > 0 && Bn > 0 ? An * Bn * 4 : 0);
    const __amdgpu_buffer_rsrc_t wr =
        __builtin_amdgcn_make_buffer_rsrc((void*)(Om + wou), (short)0, wbytes, kRawBuf);
    for (int b0 = 0; b0 < B; b0 += kTile) {
      __syncthreads();  // previous tile's readers done
      if constexpr (MODE == 0) {
        // omega rows a0 + w + 4k, columns b0 + lane: one 256-byte row segment
        // per wave and load; out-of-range rows and columns read zero
        const int bc = b0 + lane;
        const unsigned base = bc < B ? unsigned(((a0 + w) * B + bc) * 4) : 0x80000000u;
        float v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k)
          v[k] = __uint_as_float(
              __builtin_amdgcn_raw_buffer_load_b32(wr, int(base + unsigned(16 * B * k)), 0, 0));
        stage_rows(ys, Y, yo + b0, B - b0 < kTile ? B - b0 : kTile);
#pragma unroll
        for (int k = 0; k < 16; ++k) om[(w + 4 * k) * kOS + lane] = v[k];
      } else {
        // dL/dM rows a0-1 .. a0+63, columns b0-1 .. b0+63 (zero outside): wave
        // w takes rows w, w+4, ..., lane c column b0-1+c, and column b0+63 as
        // a broadcast; batches of 4 rows bound the registers in flight
        stage_rows(ys, Y, yo + b0, B - b0 < kTile ? B - b0 : kTile);
        const int bpl = b0 - 1 + lane, bpe = b0 + kTile - 1;
        const bool cok = bpl >= 0 && bpl < Bn, eok = bpe < Bn;
#pragma unroll
        for (int k0 = 0; k0 < 17; k0 += 4) {
          float vl[4], ve[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int r = w + 4 * (k0 + k);
            const int ap = a0 - 1 + r;
            const bool rok = k0 + k < 17 && r <= kTile && ap >= 0 && ap < An;
            const unsigned ol = rok && cok ? unsigned((ap * rsg + bpl * csg) * 4) : 0x80000000u;
            const unsigned oe = rok && eok ? unsigned((ap * rsg + bpe * csg) * 4) : 0x80000000u;
            vl[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(wr, int(ol), 0, 0));
            ve[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(wr, int(oe), 0, 0));
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int r = w + 4 * (k0 + k);
            if (k0 + k < 17 && r <= kTile) {
              om[r * kGS + lane] = vl[k];
              if (lane == 0) om[r * kGS + kTile] = ve[k];
            }
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      __syncthreads();
      if constexpr (MODE != 0) {
        // squared norms of the staged rows
        if (tid < kTile) {
          float sy = 0.f;
#pragma unroll
          for (int f = 0; f < FS; ++f) sy += ys[tid * FS + f] * ys[tid * FS + f];
          ynorm[tid] = sy;
          if (xnorm_due) {
            float sx = 0.f;
#pragma unroll
            for (int f = 0; f < FS; ++f) sx += xs[tid * FS + f] * xs[tid * FS + f];
            xnorm[tid] = sx;
          }
        }
        xnorm_due = false;
        // <x_a, y_b> of the tile: wave w takes rows 16w .. 16w+15
        cf4 dt[4];
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) dt[nb] = cf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < FS / 4; ++kb) {
          const float av = xs[(16 * w + c16) * FS + 4 * kb + grp];
#pragma unroll
          for (int nb = 0; nb < 4; ++nb)
            dt[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                av, ys[(16 * nb + c16) * FS + 4 * kb + grp], dt[nb], 0, 0, 0);
        }
        __syncthreads();  // norms visible
        float wv[4][4];
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int al = 16 * w + 4 * grp + q, bl = 16 * nb + c16;
            const float* g0 = om + al * kGS + bl;
            const float gG = MODE == 1 ? (g0[0] - g0[1]) - (g0[kGS] - g0[kGS + 1]) : g0[kGS + 1];
            const float dot = dt[nb][q];
            float wval;
            if constexpr (KIND == kRBF) {
              const float d2 = xnorm[al] + ynorm[bl] - 2.f * dot;
              wval = gG * fast_exp2(-gsc * d2);
              gp_acc += (double)(wval * d2);
            } else {
              wval = gG;
              gp_acc += (double)(wval * dot);
            }
            wv[nb][q] = wval;
            if (q == 3) __builtin_amdgcn_sched_barrier(0);  // bound the LDS reads in flight
          }
        __syncthreads();  // dL/dM tile consumed
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
          for (int q = 0; q < 4; ++q) om[(16 * w + 4 * grp + q) * kOS + 16 * nb + c16] = wv[nb][q];
        __syncthreads();
      }
      if (GX) {
#pragma unroll
        for (int kb = 0; kb < 16; ++kb) {
          const float av = om[(16 * w + c16) * kOS + 4 * kb + grp];
          rsum += av;
#pragma unroll
          for (int nb = 0; nb < NB; ++nb)
            gxa[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                av, ys[(4 * kb + grp) * FS + 16 * nb + c16], gxa[nb], 0, 0, 0);
        }
      }
      if (GY) {
        cf4 gya[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) gya[nb] = cf4{0.f, 0.f, 0.f, 0.f};
        float csum = 0.f;  // omega column sum, column 16w + c16, over rows = grp mod 4
#pragma unroll
        for (int kb = 0; kb < 16; ++kb) {
          const float av = om[(4 * kb + grp) * kOS + 16 * w + c16];
          csum += av;
#pragma unroll
          for (int nb = 0; nb < NB; ++nb)
            gya[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                av, xs[(4 * kb + grp) * FS + 16 * nb + c16], gya[nb], 0, 0, 0);
        }
        csum += __shfl_xor(csum, 16, kWave);
        csum += __shfl_xor(csum, 32, kWave);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float cq = __shfl(csum, 4 * grp + q, kWave);
          const int rl = 16 * w + 4 * grp + q;
          if (b0 + rl < B) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
              const int f = 16 * nb + c16;
              if (f < d)
                unsafeAtomicAdd(GY + (yo + b0 + rl) * d + f,
                                cc * gya[nb][q] + cself * cq * ys[rl * FS + f]);
            }
          }
        }
      }
    }
  }
  if (GX && cur >= 0) flush_gx();
  if constexpr (MODE != 0) {
    double g = gp_acc;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) g += __shfl_xor(g, o, kWave);
    if (lane == 0) gpart[w] = g;
    __syncthreads();
    if (tid == 0 && a.gparam)
      unsafeAtomicAdd(a.gparam, (gpart[0] + gpart[1] + gpart[2] + gpart[3]) * a.gp_scale);
  }
}

// float64 and oversized float32 pairs, MODE 1 / 2: omega(a, b) recomputed per
// element from dL/dM and the static kernel; one thread per row of one side
// (side 0: rows of x with dL/dparam, side 1: rows of y).
template <typename T, int MODE>
__global__ __launch_bounds__(64) void contract_pde_simple_kernel(const ContractArgs a) {
  const int64_t p = blockIdx.x;
  const int side = blockIdx.z;
  const int r = blockIdx.y * kTile + threadIdx.x;
  const int i = a.pairs[2 * p], j = a.pairs[2 * p + 1];
  const int A = a.x.len[i], B = a.y.len[j];
  const int64_t xo = a.x.off[i], yo = a.y.off[j];
  const T* X = static_cast<const T*>(a.x.data);
  const T* Y = static_cast<const T*>(a.y.data);
  const T* GM = static_cast<const T*>(a.omega) + a.omega_off[p];
  const int d = a.d;
  const bool rbf = a.kind == kRBF;
  const T cc = T(a.c_cross), cself = T(a.c_self), gsc = T(a.gsc);
  constexpr int sh = MODE == 1 ? 1 : 0;
  const int An = A - sh, Bn = B - sh;
  const int nck = B > A ? An : Bn;
  const int64_t rsg = B > A ? 1 : nck, csg = B > A ? nck : 1;
  auto gm = [&](int ap, int bp) -> T {
    return (ap >= 0 && ap < An && bp >= 0 && bp < Bn) ? GM[ap * rsg + bp * csg] : T(0);
  };
  // omega(aa, bb) and the factor of dL/dparam (RBF: |x-y|^2, linear: <x, y>)
  auto omega_at = [&](int aa, int bb, T* fac) -> T {
    const T gG = MODE == 1 ? (gm(aa - 1, bb - 1) - gm(aa - 1, bb)) - (gm(aa, bb - 1) - gm(aa, bb))
                           : gm(aa, bb);
    T dot = T(0), xn = T(0), yn = T(0);
    for (int f = 0; f < d; ++f) {
      const T xv = X[(xo + aa) * d + f], yv = Y[(yo + bb) * d + f];
      dot += xv * yv;
      xn += xv * xv;
      yn += yv * yv;
    }
    if (rbf) {
      const T d2 = xn + yn - T(2) * dot;
      *fac = d2;
      return gG * exp2(-gsc * d2);
    }
    *fac = dot;
    return gG;
  };
  // features in chunks of 32 accumulators (omega recomputed per chunk: this
  // is the float64 / very wide fallback)
  for (int f0 = 0; f0 < d; f0 += 32) {
    T acc[32];
#pragma unroll
    for (int f = 0; f < 32; ++f) acc[f] = T(0);
    T sum = T(0), gp = T(0);
    if (side == 0) {
      if (r >= A) return;
      for (int b = 0; b < B; ++b) {
        T fac;
        const T wv = omega_at(r, b, &fac);
#pragma unroll
        for (int f = 0; f < 32; ++f)
          if (f0 + f < d) acc[f] += wv * Y[(yo + b) * d + f0 + f];
        sum += wv;
        gp += wv * fac;
      }
      T* G = static_cast<T*>(a.gx);
      if (G) {
#pragma unroll
        for (int f = 0; f < 32; ++f) {
          if (f0 + f >= d) break;
          const int64_t o = (xo + r) * d + f0 + f;
          unsafeAtomicAdd(G + o, cc * acc[f] + cself * sum * X[o]);
        }
      }
      if (a.gparam && f0 == 0) unsafeAtomicAdd(a.gparam, (double)gp * a.gp_scale);
    } else {
      T* G = static_cast<T*>(a.gy);
      if (!G || r >= B) return;
      for (int q = 0; q < A; ++q) {
        T fac;
        const T wv = omega_at(q, r, &fac);
#pragma unroll
        for (int f = 0; f < 32; ++f)
          if (f0 + f < d) acc[f] += wv * X[(xo + q) * d + f0 + f];
        sum += wv;
      }
#pragma unroll
      for (int f = 0; f < 32; ++f) {
        if (f0 + f >= d) break;
        const int64_t o = (yo + r) * d + f0 + f;
        unsafeAtomicAdd(G + o, cc * acc[f] + cself * sum * Y[o]);
      }
    }
  }
}

// float64, and float32 pairs too large for 32-bit tile offsets: thread (row,
// feature phase) of one pair; side 0 writes gx rows, side 1 gy rows.
template <typename T>
__global__ __launch_bounds__(256) void contract_simple_kernel(const ContractArgs a) {
  const int64_t p = blockIdx.x;
  const int side = blockIdx.z;
  const int r = blockIdx.y * kTile + (threadIdx.x & 63);
  const int ph = threadIdx.x >> 6;
  const int i = a.pairs[2 * p], j = a.pairs[2 * p + 1];
  const int A = a.x.len[i], B = a.y.len[j];
  const int64_t xo = a.x.off[i], yo = a.y.off[j];
  const T* X = static_cast<const T*>(a.x.data);
  const T* Y = static_cast<const T*>(a.y.data);
  const T* W = static_cast<const T*>(a.omega) + a.omega_off[p];
  const T cc = T(a.c_cross), cself = T(a.c_self);
  const int d = a.d;
  if (side == 0) {
    T* G = static_cast<T*>(a.gx);
    if (!G || r >= A) return;
    for (int f = ph; f < d; f += 4) {
      T s = T(0), rs = T(0);
      for (int b = 0; b < B; ++b) {
        const T wv = W[(int64_t)r * B + b];
        s += wv * Y[(yo + b) * d + f];
        rs += wv;
      }
      const int64_t o = (xo + r) * d + f;
      unsafeAtomicAdd(G + o, cc * s + cself * rs * X[o]);
    }
  } else {
    T* G = static_cast<T*>(a.gy);
    if (!G || r >= B) return;
    for (int f = ph; f < d; f += 4) {
      T s = T(0), cs = T(0);
      for (int q = 0; q < A; ++q) {
        const T wv = W[(int64_t)q * B + r];
        s += wv * X[(xo + q) * d + f];
        cs += wv;
      }
      const int64_t o = (yo + r) * d + f;
      unsafeAtomicAdd(G + o, cc * s + cself * cs * Y[o]);
    }
  }
}

}  // namespace

template <int MODE, int KIND>
static void launch_f32(const ContractArgs& b, dim3 grid, hipStream_t s) {
  if (b.d <= 16)
    hipLaunchKernelGGL((contract_f32_kernel<1, MODE, KIND>), grid, dim3(256), 0, s, b);
  else if (b.d <= 32)
    hipLaunchKernelGGL((contract_f32_kernel<2, MODE, KIND>), grid, dim3(256), 0, s, b);
  else if (b.d <= 48)
    hipLaunchKernelGGL((contract_f32_kernel<3, MODE, KIND>), grid, dim3(256), 0, s, b);
  else
    hipLaunchKernelGGL((contract_f32_kernel<4, MODE, KIND>), grid, dim3(256), 0, s, b);
}

// Features the MFMA tiles take (16 per block, up to 4 blocks); wider inputs
// take the VALU reduction, which loops over any d.
constexpr int kContractMfmaMaxD = 64;

hipError_t grad_contract_launch(const ContractArgs& a, bool f64, int max_len_x, int max_len_y,
                                int n_cus, hipStream_t s) {
  if (a.n_pairs <= 0 || (!a.gx && !a.gy && !(a.mode != 0 && a.gparam))) return hipSuccess;
  if (max_len_x < 1 || max_len_y < 1) return hipSuccess;
  const int mx = max_len_x > max_len_y ? max_len_x : max_len_y;
  if (f64 || a.d > kContractMfmaMaxD || (int64_t)max_len_x * max_len_y >= (int64_t(1) << 29)) {
    if (a.mode == 0) {
      dim3 grid((unsigned)a.n_pairs, (unsigned)((mx + kTile - 1) / kTile), 2);
      if (f64)
        hipLaunchKernelGGL(contract_simple_kernel<double>, grid, dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL(contract_simple_kernel<float>, grid, dim3(256), 0, s, a);
    } else {
      dim3 grid((unsigned)a.n_pairs, (unsigned)((mx + kTile - 1) / kTile), 2);
      if (f64) {
        if (a.mode == 1)
          hipLaunchKernelGGL((contract_pde_simple_kernel<double, 1>), grid, dim3(64), 0, s, a);
        else
          hipLaunchKernelGGL((contract_pde_simple_kernel<double, 2>), grid, dim3(64), 0, s, a);
      } else {
        if (a.mode == 1)
          hipLaunchKernelGGL((contract_pde_simple_kernel<float, 1>), grid, dim3(64), 0, s, a);
        else
          hipLaunchKernelGGL((contract_pde_simple_kernel<float, 2>), grid, dim3(64), 0, s, a);
      }
    }
    return hipGetLastError();
  }
  ContractArgs b = a;
  const int nab = (max_len_x + kTile - 1) / kTile;
  // about 32 workgroups per CU over the launch, at most 32 pairs each
  int64_t g = (a.n_pairs * nab) / ((int64_t)n_cus * 32);
  b.group = (int)(g < 1 ? 1 : g > 32 ? 32 : g);
  dim3 grid((unsigned)((a.n_pairs + b.group - 1) / b.group), (unsigned)nab);
  if (a.mode == 0) {
    launch_f32<0, kRBF>(b, grid, s);
  } else if (a.mode == 1) {
    if (a.kind == kRBF) launch_f32<1, kRBF>(b, grid, s);
    else launch_f32<1, kLinear>(b, grid, s);
  } else {
    if (a.kind == kRBF) launch_f32<2, kRBF>(b, grid, s);
    else launch_f32<2, kLinear>(b, grid, s);
  }
  return hipGetLastError();
}

}  // namespace ksig
