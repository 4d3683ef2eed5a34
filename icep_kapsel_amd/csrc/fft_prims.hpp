// fft_prims.hpp -- the register-level pieces of the line FFTs (radix-2
// DFTs with compile-time indices, the L = R1 * R2 factorisations, twiddle
// access, the XCD-aware tile order), shared by the FFT passes and the x
// passes fused with k-space operators.
#pragma once

#include <utility>

#include "common.hpp"

namespace kapsel {

// ---------------------------------------------------------------- constants
// cos(2 pi m / 16)
__host__ __device__ constexpr double cos16(int m) {
  m &= 15;
  return m == 0    ? 1.0
         : m == 1  ? 0.92387953251128675613
         : m == 2  ? 0.70710678118654752440
         : m == 3  ? 0.38268343236508977173
         : m == 4  ? 0.0
         : m == 5  ? -0.38268343236508977173
         : m == 6  ? -0.70710678118654752440
         : m == 7  ? -0.92387953251128675613
         : m == 8  ? -1.0
         : m == 9  ? -0.92387953251128675613
         : m == 10 ? -0.70710678118654752440
         : m == 11 ? -0.38268343236508977173
         : m == 12 ? 0.0
         : m == 13 ? 0.38268343236508977173
         : m == 14 ? 0.70710678118654752440
                   : 0.92387953251128675613;
}
__host__ __device__ constexpr double sin16(int m) { return cos16(m - 4); }

// b * W_16^M with W_16 = exp(DIR 2 pi i / 16); exact for the axis multiples
template <int DIR, int M, typename T>
__device__ __forceinline__ V2<T> mulw16(V2<T> b) {
  constexpr int m = M & 15;
  if constexpr (m == 0) {
    return b;
  } else if constexpr (m == 4) {
    if constexpr (DIR < 0) return V2<T>{b.im, -b.re};
    else return V2<T>{-b.im, b.re};
  } else if constexpr (m == 8) {
    return V2<T>{-b.re, -b.im};
  } else if constexpr (m == 12) {
    if constexpr (DIR < 0) return V2<T>{-b.im, b.re};
    else return V2<T>{b.im, -b.re};
  } else {
    constexpr T c = (T)cos16(m);
    constexpr T s = (T)(DIR * sin16(m));
    return V2<T>{b.re * c - b.im * s, b.re * s + b.im * c};
  }
}

template <int R, int DIR, typename T, int... K>
__device__ __forceinline__ void dit_combine(const V2<T> *e, const V2<T> *o, V2<T> *out,
                                            std::integer_sequence<int, K...>) {
  ((out[K] = e[K] + mulw16<DIR, K *(16 / R)>(o[K]),
    out[K + R / 2] = e[K] - mulw16<DIR, K *(16 / R)>(o[K])),
   ...);
}

// out[k] = sum_n in[S n] W_R^{nk}, k < R, radix-2 decimation in time in
// registers (all indices compile-time)
template <int R, int S, int DIR, typename T>
__device__ __forceinline__ void dft(const V2<T> *in, V2<T> *out) {
  if constexpr (R == 1) {
    out[0] = in[0];
  } else {
    V2<T> e[R / 2], o[R / 2];
    dft<R / 2, 2 * S, DIR>(in, e);
    dft<R / 2, 2 * S, DIR>(in + S, o);
    dit_combine<R, DIR>(e, o, out, std::make_integer_sequence<int, R / 2>{});
  }
}

// factorisation L = R1 * R2 (R1 >= R2)
template <int L>
struct FS;
template <> struct FS<8> { static constexpr int R1 = 4, R2 = 2; };
template <> struct FS<16> { static constexpr int R1 = 4, R2 = 4; };
template <> struct FS<32> { static constexpr int R1 = 8, R2 = 4; };
template <> struct FS<64> { static constexpr int R1 = 8, R2 = 8; };
template <> struct FS<128> { static constexpr int R1 = 16, R2 = 8; };
template <> struct FS<256> { static constexpr int R1 = 16, R2 = 16; };

template <typename T>
struct FieldPtrs {
  T *p[6];
};
// one constant per field of a launch
template <typename T>
struct FieldConsts {
  T v[6];
};

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs by linear id (blockIdx.y * gridDim.x + blockIdx.x), and each XCD has
// its own L2; remapping gives each XCD one contiguous range of (field, tile)
// ids, so the cache lines that two neighbouring tiles share are fetched once,
// by one L2.  Returns the remapped linear id; field = id / gridDim.x.
__device__ __forceinline__ unsigned xcd_tile(unsigned &field) {
  constexpr unsigned NXCD = 8;
  const unsigned total = gridDim.x * gridDim.y;
  const unsigned lin = blockIdx.y * gridDim.x + blockIdx.x;
  const unsigned q = total / NXCD, r = total % NXCD;
  const unsigned x = lin % NXCD, k = lin / NXCD;
  const unsigned id = x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
  field = id / gridDim.x;
  return id % gridDim.x;
}

// xcd_tile with the other order of the ids: tile-major, the fields of one
// tile at consecutive ids, so that they run on one XCD close together and
// what they read in common (the inputs of a prologue) is fetched once, by
// its L2.  Returns the tile; field = id % gridDim.y.
__device__ __forceinline__ unsigned xcd_tile_by_tile(unsigned &field) {
  constexpr unsigned NXCD = 8;
  const unsigned total = gridDim.x * gridDim.y;
  const unsigned lin = blockIdx.y * gridDim.x + blockIdx.x;
  const unsigned q = total / NXCD, r = total % NXCD;
  const unsigned x = lin % NXCD, k = lin / NXCD;
  const unsigned id = x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
  field = id % gridDim.y;
  return id / gridDim.y;
}

// twiddle W_L^m, table holds the forward sign
template <int DIR, typename T>
__device__ __forceinline__ V2<T> tw_get(const V2<T> *__restrict__ tw, int m) {
  V2<T> w = tw[m];
  if constexpr (DIR > 0) w.im = -w.im;
  return w;
}

// signed wavenumber index of position i on an axis of n points
__device__ __forceinline__ int signed_k(int i, int n) { return i > n / 2 ? i - n : i; }

// ------------------------------------------------------- line transforms
// LineFFT<T, L>: a complex transform of length L by S threads of P points
// each.  Thread s (s < S) holds before the transform the points in_i(s, n)
// and after it the points out_i(s, m) (n, m < P), in v[].  sl: the line's
// LS slots of LDS; twr: the thread's first-stage twiddles W_L^{s k1}
// (forward sign, k1 < P; line_twiddles loads them, once for all the
// transforms of a kernel); pxor: the lane distance between the threads s and
// s ^ 1 of a line (a power of two below 64).  run() starts with a barrier
// (the slots may still be read by the line's previous transform) and holds
// one more.
//
// Four-step (L = R1 R2, R2 | R1): S = R2, P = R1; thread s loads the R1
// points s + R2 n, a radix-R1 DFT and the twiddles W_L^{s k1}, the exchange
// through LDS, R2-point DFTs of the k1 = s (mod R2); the outputs k1 + R1 k2
// are again the points s (mod R2): in_i = out_i = s + R2 n.
template <typename T, int L, bool SPLIT = (L == 128)>
struct LineFFT {
  static constexpr int R1 = FS<L>::R1, R2 = FS<L>::R2, S = R2, P = R1;
  static constexpr int LS = R2 * (R1 + 1) + 1;
  static_assert(R1 % R2 == 0, "four-step ownership needs R2 | R1");
  __device__ static __forceinline__ int in_i(int s, int n) { return s + R2 * n; }
  __device__ static __forceinline__ int out_i(int s, int m) { return s + R2 * m; }
  template <int DIR>
  __device__ static __forceinline__ void run(V2<T> (&v)[P], V2<T> *sl, const V2<T> (&twr)[P], int s,
                                             int /*pxor*/) {
    V2<T> o[R1];
    dft<R1, 1, DIR>(v, o);
#pragma unroll
    for (int k1 = 1; k1 < R1; k1++) o[k1] = cmul(o[k1], DIR < 0 ? twr[k1] : conjv(twr[k1]));
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) sl[s * (R1 + 1) + k1] = o[k1];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < R1 / R2; c++) {
      const int k1 = s + R2 * c;
      V2<T> w[R2], x[R2];
#pragma unroll
      for (int m = 0; m < R2; m++) w[m] = sl[m * (R1 + 1) + k1];
      dft<R2, 1, DIR>(w, x);
#pragma unroll
      for (int k2 = 0; k2 < R2; k2++) v[c + (R1 / R2) * k2] = x[k2];
    }
  }
};

// the twiddles W_L^{s k1}, k1 < P, of thread s's first stage (both forms)
template <typename F, typename T>
__device__ __forceinline__ void line_twiddles(V2<T> (&twr)[F::P], const V2<T> *__restrict__ tw, int s) {
#pragma unroll
  for (int k1 = 0; k1 < F::P; k1++) twr[k1] = tw[s * k1];
}

template <typename T>
__device__ __forceinline__ T shfl_xor_t(T v, int m) {
  return __shfl_xor(v, m);
}

// The value of lane (lane ^ MASK), MASK < 32, for every lane of a wavefront
// (all lanes active): ds_swizzle in bit mode (and-mask 31, or-mask 0,
// xor-mask MASK), one per 32-bit word, no LDS storage.
template <int MASK>
__device__ __forceinline__ float swz_xor(float x) {
  static_assert(MASK > 0 && MASK < 32, "swizzle within 32 lanes");
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(x), 0x1F | (MASK << 10)));
}
template <int MASK>
__device__ __forceinline__ double swz_xor(double x) {
  static_assert(MASK > 0 && MASK < 32, "swizzle within 32 lanes");
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_ds_swizzle((int)(unsigned)b, 0x1F | (MASK << 10));
  const unsigned hi = (unsigned)__builtin_amdgcn_ds_swizzle((int)(unsigned)(b >> 32), 0x1F | (MASK << 10));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int MASK, typename T>
__device__ __forceinline__ V2<T> swz_xor(V2<T> v) {
  return V2<T>{swz_xor<MASK>(v.re), swz_xor<MASK>(v.im)};
}

// L = 128 as 8 x 16 with the radix-16 stage split over lane pairs (SPLIT,
// the default at this length): S = 16, P = 8.  Thread s loads the points s + 16 n, a radix-8 DFT and the twiddles
// W_128^{s k1}, the exchange through LDS; the 16-point DFT of k1 is shared by
// the threads s = 2 k1 + h: each takes the inputs 2a + h (a < 8) to a
// radix-8 DFT Z_h, thread h = 1 multiplies by W_16^k, and one radix-2
// butterfly between the two lanes gives X[k1 + 8 (k + 8 h)] = Z_0[k] +-
// W_16^k Z_1[k].  Half the points (and registers) of the four-step's
// 16 x 8 per thread, twice the threads per line.
template <typename T>
struct LineFFT<T, 128, true> {
  static constexpr int S = 16, P = 8;
  static constexpr int LS = 16 * 9 + 1;
  __device__ static __forceinline__ int in_i(int s, int n) { return s + 16 * n; }
  __device__ static __forceinline__ int out_i(int s, int m) { return (s >> 1) + 8 * m + 64 * (s & 1); }
  template <int DIR, int... K>
  __device__ static __forceinline__ void w16_times(V2<T> (&z)[8], std::integer_sequence<int, K...>) {
    ((z[K] = mulw16<DIR, K>(z[K])), ...);
  }
  template <int DIR>
  __device__ static __forceinline__ void run(V2<T> (&v)[P], V2<T> *sl, const V2<T> (&twr)[P], int s, int pxor) {
    V2<T> o[8];
    dft<8, 1, DIR>(v, o);
#pragma unroll
    for (int k1 = 1; k1 < 8; k1++) o[k1] = cmul(o[k1], DIR < 0 ? twr[k1] : conjv(twr[k1]));
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < 8; k1++) sl[s * 9 + k1] = o[k1];
    __syncthreads();
    const int k1 = s >> 1, h = s & 1;
    V2<T> w[8], z[8];
#pragma unroll
    for (int a = 0; a < 8; a++) w[a] = sl[(2 * a + h) * 9 + k1];
    dft<8, 1, DIR>(w, z);
    if (h) w16_times<DIR>(z, std::make_integer_sequence<int, 8>{});
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const V2<T> r{shfl_xor_t(z[k].re, pxor), shfl_xor_t(z[k].im, pxor)};
      v[k] = h ? r - z[k] : z[k] + r;
    }
  }
};

}  // namespace kapsel
