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
#include "dp_forward.h"

namespace ksig {


namespace {

typedef float cf4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;
constexpr int kOS = kTile + 1;  // omega tile row stride in LDS (words)
constexpr int kRawBuf = 0x00020000;  // buffer descriptor word 3: raw 32-bit data

template <int NB>  // feature blocks of 16 (d <= 16 * NB)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void
contract_f32_kernel(const ContractArgs a) {
  constexpr int FS = 16 * NB;  // LDS row stride of the x / y row blocks
  __shared__ float om[kTile * kOS];  // omega tile [a][b]
  __shared__ float xs[kTile * FS];   // x rows a0 .. a0+63 of the current x sequence
  __shared__ float ys[kTile * FS];   // y rows b0 .. b0+63 of the current tile
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
    const __amdgpu_buffer_rsrc_t wr =
        __builtin_amdgcn_make_buffer_rsrc((void*)(Om + wou), (short)0, A * B * 4, kRawBuf);
    for (int b0 = 0; b0 < B; b0 += kTile) {
      __syncthreads();  // previous tile's readers done
      {
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
      }
      __syncthreads();
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

hipError_t grad_contract_launch(const ContractArgs& a, bool f64, int max_len_x, int max_len_y,
                                int n_cus, hipStream_t s) {
  if (a.n_pairs <= 0 || (!a.gx && !a.gy)) return hipSuccess;
  if (max_len_x < 1 || max_len_y < 1) return hipSuccess;
  if (f64 || (int64_t)max_len_x * max_len_y >= (int64_t(1) << 29)) {
    const int mx = max_len_x > max_len_y ? max_len_x : max_len_y;
    dim3 grid((unsigned)a.n_pairs, (unsigned)((mx + kTile - 1) / kTile), 2);
    if (f64)
      hipLaunchKernelGGL(contract_simple_kernel<double>, grid, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL(contract_simple_kernel<float>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
  }
  ContractArgs b = a;
  const int nab = (max_len_x + kTile - 1) / kTile;
  // about 32 workgroups per CU over the launch, at most 32 pairs each
  int64_t g = (a.n_pairs * nab) / ((int64_t)n_cus * 32);
  b.group = (int)(g < 1 ? 1 : g > 32 ? 32 : g);
  dim3 grid((unsigned)((a.n_pairs + b.group - 1) / b.group), (unsigned)nab);
  if (a.d <= 16)
    hipLaunchKernelGGL(contract_f32_kernel<1>, grid, dim3(256), 0, s, b);
  else
    hipLaunchKernelGGL(contract_f32_kernel<2>, grid, dim3(256), 0, s, b);
  return hipGetLastError();
}

}  // namespace ksig
