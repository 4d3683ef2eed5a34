// pnet.hip -- the learned part of the dielectric solve: a 3-level periodic
// convolutional encoder-decoder that maps (phi_sin, beta phi_sin) to three
// potentials psi_d (the response to a unit field along d), trained and run
// on the GPU with BF16 MFMA (v_mfma_f32_16x16x16_bf16, FP32 accumulation).
//
// Architecture (weights W[tap][cin][cout] then b[cout] per layer, the
// order of the CPU oracle):
//   e1a 2(16)->16, e1b 16->16 (s1)   N^3     | pool
//   e2a 16->32,    e2b 32->32 (s2)   (N/2)^3 | pool
//   e3a 32->32,    e3b 32->32        (N/4)^3 | up, concat [up | s2]
//   d2  64->32                       (N/2)^3 | up, concat [up | s1]
//   d1  48->16, out 16->3 (linear)   N^3
// The input's 2 channels are stored padded to 16 (zero channels and zero
// weights) so every convolution is a 16-multiple implicit GEMM.
//
// Activations are channels-last BF16 a[v*C + c], v = (i*n + j)*n + k.
// Convolution = 27 taps x C/16 K-steps of 16x16x16 MFMA tiles: each
// wavefront owns 16 consecutive voxels x COUT outputs; lane l loads the 4
// channels 4*(l/16).. of voxel l%16 (8 contiguous bytes), so a K-step of a
// wavefront reads one contiguous 512-byte block.  The input gradient is the
// same kernel with the taps mirrored and W transposed; the weight gradient
// reduces over voxels into per-chunk FP32 partials, summed in a fixed order.
#include <cmath>
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace kapsel {

namespace {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ bf16 f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16)(u >> 16);
}

__device__ __forceinline__ int wrapi(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// voxel v shifted by (ox, oy, oz) with periodic wrap on an n^3 cube
__device__ __forceinline__ int shifted(int v, int n, int ox, int oy, int oz) {
  int k = v % n;
  int t = v / n;
  int j = t % n;
  int i = t / n;
  return (wrapi(i + ox, n) * n + wrapi(j + oy, n)) * n + wrapi(k + oz, n);
}

// out[v][co] = sum_tap sum_c in[v + s*o_tap][c] K[tap][co][c] (+ bias, relu),
// s = +1 (forward) or -1 (input gradient).  mask (optional): zero where the
// mask tensor (same layout as out) is <= 0.  out32 (optional) FP32 copy.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void conv_mfma_kernel(const bf16 *__restrict__ in, int n,
                                                        const bf16 *__restrict__ K, const float *__restrict__ bias,
                                                        int relu, int sign, const bf16 *__restrict__ mask,
                                                        bf16 *__restrict__ out, float *__restrict__ out32) {
  constexpr int NT = COUT / 16;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nv = n * n * n;
  const int v0 = (blockIdx.x * 4 + wave) * 16;
  if (v0 >= nv) return;
  const int m = lane & 15, kq = lane >> 4;
  f4v acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
  const int vm = v0 + m;
  for (int tap = 0; tap < 27; tap++) {
    const int ox = (tap / 9 - 1) * sign, oy = ((tap / 3) % 3 - 1) * sign, oz = (tap % 3 - 1) * sign;
    const int u = shifted(vm, n, ox, oy, oz);
    const bf16 *xa = in + (size_t)u * CIN + 4 * kq;
    const bf16 *kb = K + (size_t)tap * COUT * CIN + 4 * kq;
#pragma unroll
    for (int c0 = 0; c0 < CIN; c0 += 16) {
      const s4v a = *reinterpret_cast<const s4v *>(xa + c0);
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const s4v b = *reinterpret_cast<const s4v *>(kb + (size_t)(t * 16 + m) * CIN + c0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
      }
    }
  }
  // lane holds rows 4*kq + r (voxels) of column m (output channel) per tile
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int co = t * 16 + m;
    const float bv = bias ? bias[co] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int v = v0 + 4 * kq + r;
      float x = acc[t][r] + bv;
      if (relu && !(x > 0.f)) x = 0.f;
      const size_t o = (size_t)v * COUT + co;
      if (mask && !(bf2f(mask[o]) > 0.f)) x = 0.f;
      out[o] = f2bf(x);
      if (out32) out32[o] = x;
    }
  }
}

// partial[chunk][tap][ci][co] = sum over the chunk's voxels v of
// in[v + o_tap][ci] * dy[v][co]; one block per (tap, chunk), 4 wavefronts
// sharing the (CIN/16) x (COUT/16) output tiles
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void conv_dw_kernel(const bf16 *__restrict__ in, const bf16 *__restrict__ dy,
                                                      int n, int chunk, float *__restrict__ partial) {
  constexpr int TI = CIN / 16, TJ = COUT / 16, NTILE = TI * TJ;
  constexpr int PER = (NTILE + 3) / 4;
  const int tap = blockIdx.x;
  const int ch = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int ox = tap / 9 - 1, oy = (tap / 3) % 3 - 1, oz = tap % 3 - 1;
  const int nv = n * n * n;
  const int vbeg = ch * chunk;
  const int vend = min(nv, vbeg + chunk);
  f4v acc[PER];
#pragma unroll
  for (int p = 0; p < PER; p++) acc[p] = f4v{0.f, 0.f, 0.f, 0.f};
  for (int vb = vbeg; vb < vend; vb += 16) {
    // K index = voxel vb + 4*kq + r
    int u[4], w[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      w[r] = vb + 4 * kq + r;
      u[r] = shifted(w[r], n, ox, oy, oz);
    }
#pragma unroll
    for (int p = 0; p < PER; p++) {
      const int tile = wave + 4 * p;
      if (tile < NTILE) {
        const int ti = tile / TJ, tj = tile % TJ;
        s4v a, b;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          a[r] = (short)in[(size_t)u[r] * CIN + ti * 16 + m];  // A[ci][v]
          b[r] = (short)dy[(size_t)w[r] * COUT + tj * 16 + m];  // B[v][co]
        }
        acc[p] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[p], 0, 0, 0);
      }
    }
  }
  float *dst = partial + ((size_t)ch * 27 + tap) * CIN * COUT;
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int tile = wave + 4 * p;
    if (tile < NTILE) {
      const int ti = tile / TJ, tj = tile % TJ;
#pragma unroll
      for (int r = 0; r < 4; r++) dst[(size_t)(ti * 16 + 4 * kq + r) * COUT + tj * 16 + m] = acc[p][r];
    }
  }
}

// g[e] = sum_c partial[c][e] (fixed order), e < count; optionally only the
// first cin_real input channels are kept (padded input layer)
__global__ void reduce_partials_kernel(const float *__restrict__ partial, int nchunk, size_t stride, size_t count,
                                       float *__restrict__ g) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; c++) s += partial[(size_t)c * stride + e];
  g[e] = s;
}

// db[co] = sum_v dy[v][co]
__global__ void bias_grad_kernel(const bf16 *__restrict__ dy, int nv, int cout, float *__restrict__ db) {
  __shared__ float red[256];
  const int co = blockIdx.x;
  float s = 0.f;
  for (int v = threadIdx.x; v < nv; v += 256) s += bf2f(dy[(size_t)v * cout + co]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[co] = red[0];
}

// average pool 2x2x2: in (n^3 x c) -> out ((n/2)^3 x c)
__global__ void pool_kernel(const bf16 *__restrict__ in, int n, int c, bf16 *__restrict__ out) {
  const int m = n / 2;
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)m * m * m * c) return;
  const int ch = (int)(e % c);
  const int v = (int)(e / c);
  const int k = v % m, j = (v / m) % m, i = v / (m * m);
  float s = 0.f;
  for (int a = 0; a < 8; a++)
    s += bf2f(in[((size_t)((2 * i + (a >> 2)) * n + 2 * j + ((a >> 1) & 1)) * n + 2 * k + (a & 1)) * c + ch]);
  out[e] = f2bf(0.125f * s);
}

// din (n^3 x c) += 1/8 dout(parent), masked by act > 0 (act = the pooled
// tensor's producer output), written as din = mask * (din + ...)
__global__ void pool_bwd_kernel(const bf16 *__restrict__ dout, int n, int c, const bf16 *__restrict__ act,
                                bf16 *__restrict__ din) {
  const int m = n / 2;
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * n * n * c) return;
  const int ch = (int)(e % c);
  const int v = (int)(e / c);
  const int k = v % n, j = (v / n) % n, i = v / (n * n);
  const float g = bf2f(din[e]) + 0.125f * bf2f(dout[((size_t)((i / 2) * m + j / 2) * m + k / 2) * c + ch]);
  din[e] = f2bf(bf2f(act[e]) > 0.f ? g : 0.f);
}

// concat [up(src_m) | skip]: out ((2m)^3 x (cu + cs))
__global__ void up_concat_kernel(const bf16 *__restrict__ src, int m, int cu, const bf16 *__restrict__ skip, int cs,
                                 bf16 *__restrict__ out) {
  const int n = 2 * m, c = cu + cs;
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * n * n * c) return;
  const int ch = (int)(e % c);
  const int v = (int)(e / c);
  if (ch < cu) {
    const int k = v % n, j = (v / n) % n, i = v / (n * n);
    out[e] = src[((size_t)((i / 2) * m + j / 2) * m + k / 2) * cu + ch];
  } else {
    out[e] = skip[(size_t)v * cs + ch - cu];
  }
}

// split the gradient of a concat: gup (m^3 x cu) = sum of the 8 children of
// channels [0, cu); gskip (n^3 x cs) = channels [cu, cu + cs)
__global__ void up_concat_bwd_kernel(const bf16 *__restrict__ g, int m, int cu, int cs, bf16 *__restrict__ gup,
                                     bf16 *__restrict__ gskip) {
  const int n = 2 * m, c = cu + cs;
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nup = (size_t)m * m * m * cu, nsk = (size_t)n * n * n * cs;
  if (e < nup) {
    const int ch = (int)(e % cu);
    const int v = (int)(e / cu);
    const int k = v % m, j = (v / m) % m, i = v / (m * m);
    float s = 0.f;
    for (int a = 0; a < 8; a++)
      s += bf2f(g[((size_t)((2 * i + (a >> 2)) * n + 2 * j + ((a >> 1) & 1)) * n + 2 * k + (a & 1)) * c + ch]);
    gup[e] = f2bf(s);
  } else if (e < nup + nsk) {
    const size_t q = e - nup;
    const int ch = (int)(q % cs);
    const size_t v = q / cs;
    gskip[q] = g[v * c + cu + ch];
  }
}

// loss = mean_v (sum_d e_d y[v][d] - t[v])^2 (FP64 partial per block);
// dy[v][d] = 2 (pred - t) e_d / nv in BF16, padded to 16 channels
__global__ void loss_kernel(const float *__restrict__ y, const float *__restrict__ t, int nv, float e0, float e1,
                            float e2, bf16 *__restrict__ dy, double *__restrict__ partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (int v = blockIdx.x * 256 + threadIdx.x; v < nv; v += gridDim.x * 256) {
    const float pred = e0 * y[(size_t)v * 16] + e1 * y[(size_t)v * 16 + 1] + e2 * y[(size_t)v * 16 + 2];
    const float r = pred - t[v];
    s += (double)r * r;
    const float gscale = 2.0f * r / (float)nv;
    dy[(size_t)v * 16] = f2bf(gscale * e0);
    dy[(size_t)v * 16 + 1] = f2bf(gscale * e1);
    dy[(size_t)v * 16 + 2] = f2bf(gscale * e2);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

}  // namespace

}  // namespace kapsel
