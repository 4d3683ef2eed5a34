// fft.hip -- 3-D real FFT on the padded Ooura layout for gfx950.
//
// Forward r2c (A2a_k, fft_wrapper.h:681-701): X(k) = sum_x a(x) e^{-ik.x},
// unscaled, half spectrum kz = 0..NZ/2 stored as (re, im) pairs in rows of
// NZ_ = NZ+2 reals.  Inverse c2r (A_k2a, fft_wrapper.h:704-723) scales by
// 1/(NX NY NZ).
//
// Three passes, each one read and one write of the field (the z and y
// passes are one plane pass where a plane fits in LDS, below):
//   z pass  r2c of each row: a length-NZ/2 complex FFT of the even/odd
//           interleaved row, then the real split (c2r: the split first);
//   y pass  complex FFT along y, lines at stride h = NZ/2+1 complex values;
//   x pass  complex FFT along x, lines at stride NY*h.
// Every line FFT of length L = R1*R2 is a four-step transform: each of R2
// threads loads R1 strided points straight from HBM into registers and does
// a radix-R1 DFT, the twiddled partial results cross once through LDS, and
// R2-point DFTs finish in registers and store straight to HBM.  Adjacent
// lines are adjacent in memory (lines of a tile differ in the fastest
// index), so each load/store instruction of a wavefront touches few 64-B
// segments.  Up to six fields are transformed by one launch (grid.y).
#include <hip/hip_ext.h>

#include "fft_prims.hpp"

namespace kapsel {

// ------------------------------------------------------------ c2c line pass
// Line q (0 <= q < Q) starts at complex offset (q / A) * B + (q % A); its
// elements are ES complex values apart.  Lines q of one tile are adjacent.
// Thread t of a tile works on line t % LPB as its thread s = t / LPB (so s
// and s ^ 1 are LPB lanes apart).  out: where the transformed lines go (out
// = fp transforms in place; each line is read whole before the barrier and
// written after it)
template <typename T, int L, int DIR, int LPB>
__global__ __launch_bounds__((LPB * LineFFT<T, L>::S)) void fft_c2c_lines_io(FieldPtrs<T> fp, FieldPtrs<T> out,
                                                                           long long Q, long long A, long long B,
                                                                           long long ES, const V2<T> *__restrict__ tw) {
  using F = LineFFT<T, L>;
  __shared__ V2<T> sm[LPB * F::LS];
  unsigned field;
  const unsigned tile = xcd_tile(field);
  const T *data = fp.p[field];
  T *dst = out.p[field];
  const int t = threadIdx.x;
  const int l = t % LPB;
  const int s = t / LPB;
  const long long q = (long long)tile * LPB + l;
  const bool valid = q < Q;
  const long long base = valid ? (q / A) * B + (q % A) : 0;
  V2<T> v[F::P], twr[F::P];
#pragma unroll
  for (int n = 0; n < F::P; n++) v[n] = ld2(data + 2 * (base + (long long)F::in_i(s, n) * ES));
  line_twiddles<F>(twr, tw, s);
  F::template run<DIR>(v, sm + l * F::LS, twr, s, LPB);
  if (valid) {
#pragma unroll
    for (int m = 0; m < F::P; m++) st2(dst + 2 * (base + (long long)F::out_i(s, m) * ES), v[m]);
  }
}

// --------------------------------------------------------------- z passes
// Row r (NX*NY rows, pitch nz_ reals).  z[n] = a[2n] + i a[2n+1], n < M.
// The plain passes read and write each row in the transform's order (TPL
// threads per row, 8 bytes each); the r2c pass that forms its input from
// several fields (fft_z_r2c_pro) instead reads and writes its block's tile
// of rows 16 bytes per lane in address order, through the rows' LDS slots.
// (Measured at 256^3 FP32: staging costs the one-input passes 10-15 %, and
// saves the three-input prologue about 30 %.)

// the row stride (complex values) of a row's LDS slot in the z passes
template <int NZ>
constexpr int z_row_stride() {
  return FS<NZ / 2>::R2 * (FS<NZ / 2>::R1 + 1) + 1;
}

// The r2c of a row: thread n2 of the row's TPL = R2 threads holds v[n1] =
// z[R2 n1 + n2]; the M-point complex transform (R1 in registers, R2 across
// the row's threads through its LDS slot sr), then the split step, whose
// outputs X[k], k = n2 + TPL j (and k = M for n2 = 0) go to out(k, X[k])
// while the slot is still read.  Starts with a barrier.
template <typename T, int NZ, typename Out>
__device__ __forceinline__ void z_r2c_core(V2<T> (&v)[FS<NZ / 2>::R1], V2<T> *__restrict__ sr, int n2,
                                           const V2<T> *__restrict__ twm, const V2<T> *__restrict__ zrot,
                                           Out out) {
  constexpr int M = NZ / 2;
  constexpr int R1 = FS<M>::R1, R2 = FS<M>::R2, TPL = R2;
  V2<T> o[R1];
  dft<R1, 1, -1>(v, o);
#pragma unroll
  for (int k1 = 1; k1 < R1; k1++) o[k1] = cmul(o[k1], tw_get<-1>(twm, n2 * k1));
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < R1; k1++) sr[n2 * (R1 + 1) + k1] = o[k1];
  __syncthreads();
  V2<T> zz[R1 / TPL][R2];
#pragma unroll
  for (int c = 0; c < R1 / TPL; c++) {
    const int k1 = n2 + TPL * c;
    V2<T> w[R2];
#pragma unroll
    for (int m = 0; m < R2; m++) w[m] = sr[m * (R1 + 1) + k1];
    dft<R2, 1, -1>(w, zz[c]);
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < R1 / TPL; c++) {
    const int k1 = n2 + TPL * c;
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) sr[k1 + R1 * k2] = zz[c][k2];
  }
  __syncthreads();
  // X[k] = E + W^k O,  E = (Z[k] + conj Z[M-k]) / 2,  O = (Z[k] - conj Z[M-k]) / 2i
  auto split = [&](int k) {
    V2<T> zk = sr[k & (M - 1)];
    V2<T> zm = conjv(sr[(M - k) & (M - 1)]);
    V2<T> e = scal((T)0.5, zk + zm);
    V2<T> d = zk - zm;
    V2<T> od = {(T)0.5 * d.im, (T)-0.5 * d.re};
    return e + cmul(zrot[k], od);
  };
#pragma unroll
  for (int j = 0; j < M / TPL; j++) out(n2 + TPL * j, split(n2 + TPL * j));
  if (n2 == 0) out(M, split(M));
}

template <typename T, int NZ, int RPB>
__global__ __launch_bounds__(RPB *FS<NZ / 2>::R2) void fft_z_r2c(FieldPtrs<T> fp, long long rows, int nz_,
                                                                  const V2<T> *__restrict__ twm,
                                                                  const V2<T> *__restrict__ zrot) {
  constexpr int M = NZ / 2;
  constexpr int R1 = FS<M>::R1, R2 = FS<M>::R2, TPL = R2;
  constexpr int RS = z_row_stride<NZ>();
  static_assert(RS >= M + 1, "row stride");
  __shared__ V2<T> sm[RPB * RS];
  unsigned field;
  const unsigned tile = xcd_tile(field);
  const int t = threadIdx.x;
  const int n2 = t % TPL;
  const int r = t / TPL;
  const long long row = (long long)tile * RPB + r;
  const bool valid = row < rows;
  // row 0 for an invalid row: every load is in bounds and none is
  // predicated; such a row's results are never stored
  T *__restrict__ rp = fp.p[field] + (valid ? row : 0) * nz_;
  V2<T> v[R1];
#pragma unroll
  for (int n1 = 0; n1 < R1; n1++) v[n1] = ld2(rp + 2 * (R2 * n1 + n2));
  z_r2c_core<T, NZ>(v, sm + r * RS, n2, twm, zrot, [&](int k, V2<T> x) {
    if (valid) st2(rp + 2 * k, x);
  });
}

// 16 bytes of complex values
template <typename T>
struct alignas(16) Q16 {
  V2<T> c[16 / (2 * sizeof(T))];
};

// Tile access in address order: complex value q of the block's tile of
// nrow rows (row q / H, column q % H, H = nz_ / 2) has the LDS slot place
// sm[row * RS + column].  f(q) is called for the first value q of every
// 16-byte group (a group may span two rows).
template <typename T, int NZ, int RPB, int THREADS, typename F>
__device__ __forceinline__ void z_tile_groups(int nrow, F f) {
  constexpr int CPG = 16 / (2 * (int)sizeof(T));  // complex values per group
  constexpr int H = NZ / 2 + 1;
  constexpr int NG = RPB * H / CPG;                 // groups in a full tile
  static_assert(RPB * H % CPG == 0, "tile of whole groups");
  const int ng = nrow * H / CPG;
#pragma unroll
  for (int i = 0; i < (NG + THREADS - 1) / THREADS; i++) {
    const int gi = (int)threadIdx.x + THREADS * i;
    if (gi < ng) f(gi * CPG);
  }
}

// fft_z_r2c of a field formed on the fly as the tile is read: out_f =
// a_f + s (b_f + e_f) c_f (a_f NULL: no a term; c_f NULL: a_f + s (b_f +
// e_f); bit f of dsq: (b_f - c_f)(b_f + c_f) in place of the product), the
// prologue of fft_plane_r2c_pro for grids whose planes do not fit in LDS.
// out_f may be one of field f's own inputs (a block reads its tile before
// it stores it), but no other field's input.  The fields of one tile run
// at consecutive ids (xcd_tile_by_tile), so tiles they read in common come
// from HBM once.
template <typename T, int NZ, int RPB>
__global__ __launch_bounds__(RPB *FS<NZ / 2>::R2) void fft_z_r2c_pro(FieldPtrs<T> out, FieldPtrs<T> a,
                                                                      FieldPtrs<T> b, FieldPtrs<T> c,
                                                                      FieldConsts<T> e, T s, unsigned dsq,
                                                                      long long rows, int nz_,
                                                                      const V2<T> *__restrict__ twm,
                                                                      const V2<T> *__restrict__ zrot) {
  constexpr int M = NZ / 2, R1 = FS<M>::R1, R2 = FS<M>::R2, TPL = R2, THREADS = RPB * TPL;
  constexpr int RS = z_row_stride<NZ>(), H = M + 1, CPG = 16 / (2 * (int)sizeof(T));
  __shared__ V2<T> sm[RPB * RS];
  unsigned field;
  const unsigned tile = xcd_tile_by_tile(field);
  const long long row0 = (long long)tile * RPB;
  const int nrow = rows - row0 < RPB ? (int)(rows - row0) : RPB;
  const size_t to = (size_t)row0 * nz_;
  const T *ga = a.p[field] ? a.p[field] + to : nullptr, *gb = b.p[field] + to;
  const T *gc = c.p[field] ? c.p[field] + to : nullptr;
  const bool diff = (dsq >> field) & 1;
  const T ef = e.v[field];
  z_tile_groups<T, NZ, RPB, THREADS>(nrow, [&](int q) {
    Q16<T> vb = *reinterpret_cast<const Q16<T> *>(gb + 2 * (size_t)q);
    Q16<T> vc{}, va{};
    if (gc) vc = *reinterpret_cast<const Q16<T> *>(gc + 2 * (size_t)q);
    if (ga) va = *reinterpret_cast<const Q16<T> *>(ga + 2 * (size_t)q);
#pragma unroll
    for (int h = 0; h < CPG; h++) {
      V2<T> w = V2<T>{vb.c[h].re + ef, vb.c[h].im + ef};
      if (gc) {
        const V2<T> y = vc.c[h];
        w = diff ? V2<T>{(w.re - y.re) * (w.re + y.re), (w.im - y.im) * (w.im + y.im)}
                 : V2<T>{w.re * y.re, w.im * y.im};
      }
      w = ga ? V2<T>{va.c[h].re + s * w.re, va.c[h].im + s * w.im} : V2<T>{s * w.re, s * w.im};
      const int row = (q + h) / H;
      sm[row * RS + (q + h - row * H)] = w;
    }
  });
  __syncthreads();
  const int t = threadIdx.x, n2 = t % TPL;
  V2<T> *sr = sm + (t / TPL) * RS;
  V2<T> v[R1], x[M / TPL + 1];
#pragma unroll
  for (int n1 = 0; n1 < R1; n1++) v[n1] = sr[R2 * n1 + n2];
  z_r2c_core<T, NZ>(v, sr, n2, twm, zrot, [&](int k, V2<T> y) { x[k == M ? M / TPL : (k - n2) / TPL] = y; });
  __syncthreads();
#pragma unroll
  for (int j = 0; j < M / TPL; j++) sr[n2 + TPL * j] = x[j];
  if (n2 == 0) sr[M] = x[M / TPL];
  __syncthreads();
  T *__restrict__ dst = out.p[field] + to;
  z_tile_groups<T, NZ, RPB, THREADS>(nrow, [&](int q) {
    Q16<T> g;
#pragma unroll
    for (int h = 0; h < CPG; h++) {
      const int row = (q + h) / H;
      g.c[h] = sm[row * RS + (q + h - row * H)];
    }
    *reinterpret_cast<Q16<T> *>(dst + 2 * (size_t)q) = g;
  });
}

template <typename T, int NZ, int RPB>
__global__ __launch_bounds__(RPB *FS<NZ / 2>::R2) void fft_z_c2r(FieldPtrs<T> fp, long long rows, int nz_,
                                                                  const V2<T> *__restrict__ twm,
                                                                  const V2<T> *__restrict__ zrot, T scale) {
  constexpr int M = NZ / 2;
  constexpr int R1 = FS<M>::R1, R2 = FS<M>::R2, TPL = R2;
  constexpr int RS = z_row_stride<NZ>();
  __shared__ V2<T> sm[RPB * RS];
  unsigned field;
  const unsigned tile = xcd_tile(field);
  T *__restrict__ data = fp.p[field];
  const int t = threadIdx.x;
  const int n2 = t % TPL;
  const int r = t / TPL;
  const long long row = (long long)tile * RPB + r;
  const bool valid = row < rows;
  T *__restrict__ rp = data + (valid ? row : 0) * nz_;
  V2<T> *sr = sm + r * RS;

  // all loads of the row first (rp points at row 0 for an invalid row, so
  // every load is in bounds), then the LDS stores: no load waits on a store
  constexpr int NJ = M / TPL;
  V2<T> xin[NJ];
#pragma unroll
  for (int j = 0; j < NJ; j++) xin[j] = ld2(rp + 2 * (n2 + TPL * j));
  V2<T> xlast = ld2(rp + 2 * M);
  if (n2 == 0) xin[0].im = 0;  // c2r treats the kz = 0, NZ/2 entries as real
#pragma unroll
  for (int j = 0; j < NJ; j++) sr[n2 + TPL * j] = xin[j];
  if (n2 == 0) sr[M] = V2<T>{xlast.re, (T)0};
  __syncthreads();
  V2<T> v[R1], o[R1];
#pragma unroll
  for (int n1 = 0; n1 < R1; n1++) {
    const int n = R2 * n1 + n2;
    V2<T> xn = sr[n], xm = conjv(sr[M - n]);
    V2<T> e = xn + xm;                                  // 2 E[n]
    V2<T> od = cmul(xn - xm, conjv(zrot[n]));           // 2 O[n]
    v[n1] = V2<T>{e.re - od.im, e.im + od.re};          // E + i O
  }
  dft<R1, 1, +1>(v, o);
#pragma unroll
  for (int k1 = 1; k1 < R1; k1++) o[k1] = cmul(o[k1], tw_get<+1>(twm, n2 * k1));
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < R1; k1++) sr[n2 * (R1 + 1) + k1] = o[k1];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < R1 / TPL; c++) {
    const int k1 = n2 + TPL * c;
    V2<T> w[R2], x[R2];
#pragma unroll
    for (int m = 0; m < R2; m++) w[m] = sr[m * (R1 + 1) + k1];
    dft<R2, 1, +1>(w, x);
    if (valid) {
#pragma unroll
      for (int k2 = 0; k2 < R2; k2++) st2(rp + 2 * (k1 + R1 * k2), scal(scale, x[k2]));
    }
  }
  if (valid && n2 == 0) {
    rp[NZ] = 0;
    rp[NZ + 1] = 0;
  }
}

// ------------------------------------------------------------ plane passes
// The z and y transforms of one x-plane in one workgroup.  Plane i is the
// NY rows of H = NZ/2+1 complex values at complex offset i*NY*H, contiguous
// in memory; it is read once, the row (z) and column (y) transforms run out
// of LDS, and it is written once: one read and one write of the field for
// two of the three axes.  A transform that comes first reads the plane from
// global memory in its first stage, one that comes last writes it there in
// its last stage; only the hand-over between the row and the column
// transforms goes through the LDS plane.  Used where the plane fits
// (gfx950: 160 KB of LDS per CU; at 128^3 FP32 two workgroups share a CU).
//
// Rows: the ZR2 lanes of a row sit in one wavefront (lane = row * ZR2 + n2;
// RPW rows per wavefront), so a row's four-step exchange is ordered by the
// wavefront itself and needs no workgroup barrier.  LDS rows are HP complex
// values apart: room for the exchange's padded slots, and rows 2 ZR2 words
// apart in the banks, so the 16 lanes of a quarter-wavefront (16 / ZR2 rows)
// meet 32 different banks.
constexpr int row_pitch(int need, int r2) {
  int h = need;
  while (h % 16 != r2 % 16) h++;
  return h;
}

template <typename T, int NY, int NZ, int THREADS>
struct PlaneCfg {
  static constexpr int M = NZ / 2, H = NZ / 2 + 1;
  static constexpr int ZR1 = FS<M>::R1, ZR2 = FS<M>::R2;
  static constexpr int RPW = 64 / ZR2;     // rows per wavefront
  static constexpr int NW = THREADS / 64;  // wavefronts
  static constexpr int YR1 = FS<NY>::R1, YR2 = FS<NY>::R2;
  static constexpr int YCOLS = THREADS / YR2;  // columns per round
  static constexpr int XS = ZR2 * (ZR1 + 1);  // a row's exchange slots
  static constexpr int HP = row_pitch(H > XS ? H : XS, ZR2);
  static constexpr int PLANE = NY * H;    // complex values of a plane in global memory
  static constexpr int LPLANE = NY * HP;  // and in LDS
  static constexpr int PK = 2 * NY;         // the packed column's staging slots
  static constexpr size_t LDS_BYTES = sizeof(V2<T>) * (LPLANE + NZ + 1 + NY + PK);
  static constexpr int WGS_PER_CU = 2 * LDS_BYTES <= 160 * 1024 ? 2 : 1;
  static constexpr int WAVES_PER_SIMD = (WGS_PER_CU * NW + 3) / 4;
  static_assert(64 % ZR2 == 0 && THREADS % 64 == 0 && NY % RPW == 0, "row tiling");
  static_assert(THREADS % YR2 == 0, "column tiling");
};

// the twiddle tables of a plane pass, copied to LDS
template <typename T, int NY, int NZ>
struct PlaneTw {
  V2<T> twm[NZ / 2], zrot[NZ / 2 + 1], twy[NY];
};
// (YS: the y table holds YS NY points, every YS-th is taken)
template <typename T, int NY, int NZ, int THREADS, int YS = 1>
__device__ __forceinline__ void plane_tw_load(PlaneTw<T, NY, NZ> &tw, const V2<T> *__restrict__ twm,
                                              const V2<T> *__restrict__ zrot, const V2<T> *__restrict__ twy) {
  for (int e = threadIdx.x; e < NZ / 2; e += THREADS) tw.twm[e] = twm[e];
  for (int e = threadIdx.x; e < NZ / 2 + 1; e += THREADS) tw.zrot[e] = zrot[e];
  for (int e = threadIdx.x; e < NY; e += THREADS) tw.twy[e] = twy[YS * e];
}
// the same in two halves, for a workgroup of at least NZ/2 + 1 and NY
// threads (one entry of each table per thread): the loads into registers,
// then, after the kernel's other loads are issued, the stores to LDS
template <typename T>
struct PlaneTwRegs {
  V2<T> m, z, y;
};
template <typename T, int NY, int NZ, int THREADS, int YS = 1>
__device__ __forceinline__ PlaneTwRegs<T> plane_tw_fetch(const V2<T> *__restrict__ twm,
                                                         const V2<T> *__restrict__ zrot,
                                                         const V2<T> *__restrict__ twy) {
  static_assert(THREADS >= NZ / 2 + 1 && THREADS >= NY, "one table entry per thread");
  const int e = threadIdx.x;
  PlaneTwRegs<T> r;
  if (e < NZ / 2) r.m = twm[e];
  if (e < NZ / 2 + 1) r.z = zrot[e];
  if (e < NY) r.y = twy[YS * e];
  return r;
}
template <typename T, int NY, int NZ>
__device__ __forceinline__ void plane_tw_put(PlaneTw<T, NY, NZ> &tw, const PlaneTwRegs<T> &r) {
  const int e = threadIdx.x;
  if (e < NZ / 2) tw.twm[e] = r.m;
  if (e < NZ / 2 + 1) tw.zrot[e] = r.z;
  if (e < NY) tw.twy[e] = r.y;
}
// plane_tw_load with the loads of issue() issued between the table loads
// and their LDS stores (where one entry per thread covers the tables)
template <typename T, int NY, int NZ, int THREADS, typename Issue>
__device__ __forceinline__ void plane_tw_load_around(PlaneTw<T, NY, NZ> &tw, const V2<T> *__restrict__ twm,
                                                     const V2<T> *__restrict__ zrot, const V2<T> *__restrict__ twy,
                                                     Issue issue) {
  if constexpr (THREADS >= NZ / 2 + 1 && THREADS >= NY) {
    const PlaneTwRegs<T> r = plane_tw_fetch<T, NY, NZ, THREADS>(twm, zrot, twy);
    issue();
    plane_tw_put<T, NY, NZ>(tw, r);
  } else {
    issue();
    plane_tw_load<T, NY, NZ, THREADS>(tw, twm, zrot, twy);
  }
}

// Every lane's LDS accesses before this point precede every lane's after
// it: the LDS operations of one wavefront are performed in issue order, and
// the fences keep the compiler from moving accesses across.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Four-step DFT of one row of length R1*R2 held by its R2 lanes: lane n2
// holds v[n1] = z[n2 + R2 n1]; out(n1, Z) receives Z[n2 + R2 n1] of the
// transform.  The exchange uses the row's padded slots 0 .. R2 (R1+1) - 1,
// and every lane has read its slots before out() is called.
template <typename T, int R1, int R2, int DIR, typename Out>
__device__ __forceinline__ void row_four_step(V2<T> *row, V2<T> (&v)[R1], const V2<T> *tw, int n2, Out out) {
  V2<T> o[R1];
  dft<R1, 1, DIR>(v, o);
#pragma unroll
  for (int k1 = 1; k1 < R1; k1++) o[k1] = cmul(o[k1], tw_get<DIR>(tw, n2 * k1));
  wave_lds_sync();  // every lane's earlier reads of the row are done
#pragma unroll
  for (int k1 = 0; k1 < R1; k1++) row[n2 * (R1 + 1) + k1] = o[k1];
  wave_lds_sync();
  V2<T> res[R1];
#pragma unroll
  for (int c = 0; c < R1 / R2; c++) {
    const int k1 = n2 + R2 * c;
    V2<T> w[R2], x[R2];
#pragma unroll
    for (int m = 0; m < R2; m++) w[m] = row[m * (R1 + 1) + k1];
    dft<R2, 1, DIR>(w, x);
    // output k1 + R1 k2 = n2 + R2 (c + (R1/R2) k2)
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) res[c + (R1 / R2) * k2] = x[k2];
  }
  wave_lds_sync();  // every lane's exchange reads are done
#pragma unroll
  for (int n1 = 0; n1 < R1; n1++) out(n1, res[n1]);
}

// z r2c of one row: v[n1] = z[n2 + R2 n1] = a[2n] + i a[2n+1] (the packed
// reals); a length-M complex FFT, then X[k] = E + W^k O with E = (Z[k] +
// conj Z[M-k]) / 2, O = (Z[k] - conj Z[M-k]) / 2i, k = 0..M-1, into row[k],
// except that row[0] holds the two real values (X[0], X[M]) = (Re Z[0] +
// Im Z[0], Re Z[0] - Im Z[0]): the kz = 0 and kz = M columns of the plane
// travel as one complex column
template <typename T, int M, int R1, int R2>
__device__ __forceinline__ void row_r2c(V2<T> *row, V2<T> (&v)[R1], const V2<T> *twm, const V2<T> *zrot, int n2) {
  row_four_step<T, R1, R2, -1>(row, v, twm, n2, [&](int n1, V2<T> z) { row[n2 + R2 * n1] = z; });
  wave_lds_sync();
  constexpr int KPT = M / R2;  // split outputs per lane
  V2<T> zk[KPT], zm[KPT];
#pragma unroll
  for (int q = 0; q < KPT; q++) {
    const int k = n2 + R2 * q;
    zk[q] = row[k];
    zm[q] = conjv(row[(M - k) & (M - 1)]);
  }
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < KPT; q++) {
    const int k = n2 + R2 * q;
    const V2<T> e = scal((T)0.5, zk[q] + zm[q]);
    const V2<T> d = zk[q] - zm[q];
    const V2<T> od = {(T)0.5 * d.im, (T)-0.5 * d.re};
    V2<T> x = e + cmul(zrot[k], od);
    if (q == 0 && n2 == 0) x = V2<T>{zk[0].re + zk[0].im, zk[0].re - zk[0].im};
    row[k] = x;
  }
}

// z c2r of one row (scaled): X[k] = row[k], k = 1..M-1, and the real X[0],
// X[M] packed in row[0] (as row_r2c leaves them); out(n1, z) receives
// z[n2 + R2 n1] = (a[2n], a[2n+1])
template <typename T, int M, int R1, int R2, typename Out>
__device__ __forceinline__ void row_c2r(V2<T> *row, const V2<T> *twm, const V2<T> *zrot, T scale, int n2, Out out) {
  V2<T> v[R1];
#pragma unroll
  for (int n1 = 0; n1 < R1; n1++) {
    const int n = R2 * n1 + n2;
    V2<T> xn = row[n];
    V2<T> xm = row[(M - n) & (M - 1)];
    if (n1 == 0 && n2 == 0) {  // n = 0: X[0] = row[0].re, X[M] = row[0].im
      xm = V2<T>{xn.im, 0};
      xn.im = 0;
    }
    xm = conjv(xm);
    const V2<T> e = xn + xm;                         // 2 E[n]
    const V2<T> od = cmul(xn - xm, conjv(zrot[n]));  // 2 O[n]
    v[n1] = V2<T>{e.re - od.im, e.im + od.re};       // E + i O
  }
  row_four_step<T, R1, R2, +1>(row, v, twm, n2, [&](int n1, V2<T> z) { out(n1, scal(scale, z)); });
}

// The kz = 0 and kz = M columns of a plane in global memory (at g, the
// plane's first real) folded into one complex column, left in pk[0, NY):
// c[j] = S(X_0)[j] + i S(X_M)[j], S(X)[j] = (X[j] + conj X[-j]) / 2.  For a
// real field both columns are Hermitian in y (S(X) = X), their inverse
// transforms are real, and the inverse of c carries them as its real and
// imaginary parts; S makes that exact for any input, with the real parts of
// the two inverse transforms as the result (what dropping their imaginary
// parts gives).  Wavefront 0 does it through pk (2 NY contiguous slots, so
// its 64 lanes meet distinct banks); a workgroup barrier must follow.
// (plane_pack_fetch: the loads of wavefront 0; plane_pack_fold: the rest,
// so that a kernel can issue its other loads in between)
template <typename T, int NY>
struct PlanePackRegs {
  static constexpr int NJ = (NY + 63) / 64;
  V2<T> a[NJ], b[NJ];
};
template <typename T, int NY, int NZ, int THREADS>
__device__ __forceinline__ void plane_pack_fetch(PlanePackRegs<T, NY> &r, const T *__restrict__ g) {
  using C = PlaneCfg<T, NY, NZ, THREADS>;
  constexpr int H = C::H, M = C::M, NJ = PlanePackRegs<T, NY>::NJ;
  if (threadIdx.x >= 64) return;
#pragma unroll
  for (int q = 0; q < NJ; q++) {
    const int j = threadIdx.x + 64 * q;
    if (j < NY) {
      r.a[q] = ld2(g + 2 * (size_t)j * H);
      r.b[q] = ld2(g + 2 * ((size_t)j * H + M));
    }
  }
}
template <typename T, int NY, int NZ, int THREADS>
__device__ __forceinline__ void plane_pack_fold(V2<T> *pk, const PlanePackRegs<T, NY> &r) {
  constexpr int NJ = PlanePackRegs<T, NY>::NJ;
  if (threadIdx.x >= 64) return;
  const V2<T> *a = r.a, *b = r.b;
#pragma unroll
  for (int q = 0; q < NJ; q++) {
    const int j = threadIdx.x + 64 * q;
    if (j < NY) {
      pk[j] = a[q];
      pk[NY + j] = b[q];
    }
  }
  wave_lds_sync();
  V2<T> c[NJ];
#pragma unroll
  for (int q = 0; q < NJ; q++) {
    const int j = threadIdx.x + 64 * q;
    if (j < NY) {
      const int jm = (NY - j) & (NY - 1);
      const V2<T> sa = scal((T)0.5, a[q] + conjv(pk[jm]));
      const V2<T> sb = scal((T)0.5, b[q] + conjv(pk[NY + jm]));
      c[q] = V2<T>{sa.re - sb.im, sa.im + sb.re};
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int q = 0; q < NJ; q++) {
    const int j = threadIdx.x + 64 * q;
    if (j < NY) pk[j] = c[q];
  }
}
template <typename T, int NY, int NZ, int THREADS>
__device__ __forceinline__ void plane_pack_in(V2<T> *pk, const T *__restrict__ g) {
  PlanePackRegs<T, NY> r;
  plane_pack_fetch<T, NY, NZ, THREADS>(r, g);
  plane_pack_fold<T, NY, NZ, THREADS>(pk, r);
}

// y transform of every column of the plane in P (in place), direction DIR.
// Column 0 is the packed pair of the kz = 0 and kz = M columns (M columns
// in all).  GIN: the columns are read from the plane in global memory at g
// instead of from P (column 0 from pk, where plane_pack_in put it); GOUT:
// the results are written to g instead of P, and the packed column, left in
// pk, is parted into the columns 0 and M of g after the last round: X_0 =
// (C[j] + conj C[-j]) / 2, X_M = (C[j] - conj C[-j]) / 2i.  Consecutive
// lanes take consecutive columns, so the lanes of a wavefront touch
// consecutive LDS words and, in global memory, one contiguous run.
// first(): called by every thread once, in the first round after the
// loads from global memory (GIN) are issued and before anything is read
// from LDS -- for a kernel's set-up (LDS stores, a barrier) to wait for
// those loads together with its own.
struct PlaneNoSetup {
  __device__ void operator()() const {}
};
template <typename T, int NY, int NZ, int THREADS, int DIR, bool GIN = false, bool GOUT = false,
          typename First = PlaneNoSetup>
__device__ __forceinline__ void plane_columns(V2<T> *P, V2<T> *pk, const V2<T> *__restrict__ twy, T *g = nullptr,
                                              First first = First()) {
  using C = PlaneCfg<T, NY, NZ, THREADS>;
  constexpr int H = C::H, HP = C::HP, M = C::M, R1 = C::YR1, R2 = C::YR2;
  const int t = threadIdx.x;
#pragma unroll 1  // one round's registers at a time
  for (int c0 = 0; c0 < M; c0 += C::YCOLS) {
    // a last round of at most 64 / R2 columns gives each column R2 adjacent
    // lanes of wavefront 0, so that the other wavefronts skip its arithmetic
    const bool few = M - c0 < C::YCOLS && (M - c0) * R2 <= 64;
    const int cl = few ? t / R2 : t % C::YCOLS, n2 = few ? t % R2 : t / C::YCOLS;
    const int kz = c0 + cl;
    const bool valid = kz < M;
    const bool packed = kz == 0;
    V2<T> *col = P + (valid ? kz : 0);
    T *gcol = g + 2 * (valid ? kz : 0);
    V2<T> o[R1], v[R1];
    if (GIN && valid && !packed) {
#pragma unroll
      for (int n1 = 0; n1 < R1; n1++) v[n1] = ld2(gcol + 2 * (R2 * n1 + n2) * H);
    }
    if (c0 == 0) first();
    if (valid) {
      if (GIN && packed) {
#pragma unroll
        for (int n1 = 0; n1 < R1; n1++) v[n1] = pk[R2 * n1 + n2];
      } else if (!GIN) {
#pragma unroll
        for (int n1 = 0; n1 < R1; n1++) v[n1] = col[(R2 * n1 + n2) * HP];
      }
      dft<R1, 1, DIR>(v, o);
#pragma unroll
      for (int k1 = 1; k1 < R1; k1++) o[k1] = cmul(o[k1], tw_get<DIR>(twy, n2 * k1));
    }
    __syncthreads();  // every column of the round is in registers
    if (valid) {
#pragma unroll
      for (int k1 = 0; k1 < R1; k1++) col[(n2 * R1 + k1) * HP] = o[k1];
    }
    __syncthreads();
    if (valid) {
      // the slots read for k1 are the slots written for k1 (same thread)
#pragma unroll
      for (int c = 0; c < R1 / R2; c++) {
        const int k1 = n2 + R2 * c;
        V2<T> w[R2], x[R2];
#pragma unroll
        for (int m = 0; m < R2; m++) w[m] = col[(m * R1 + k1) * HP];
        dft<R2, 1, DIR>(w, x);
        if (GOUT && packed) {
#pragma unroll
          for (int k2 = 0; k2 < R2; k2++) pk[k1 + R1 * k2] = x[k2];
        } else if (GOUT) {
#pragma unroll
          for (int k2 = 0; k2 < R2; k2++) st2(gcol + 2 * (k1 + R1 * k2) * H, x[k2]);
        } else {
#pragma unroll
          for (int k2 = 0; k2 < R2; k2++) col[(k1 + R1 * k2) * HP] = x[k2];
        }
      }
    }
    __syncthreads();
  }
  if constexpr (GOUT) {
    if (t < 64) {
      for (int j = t; j < NY; j += 64) {
        const int jm = (NY - j) & (NY - 1);
        const V2<T> c = pk[j], d = conjv(pk[jm]);
        const V2<T> df = c - d;
        st2(g + 2 * (size_t)j * H, scal((T)0.5, c + d));
        st2(g + 2 * ((size_t)j * H + M), V2<T>{(T)0.5 * df.im, (T)-0.5 * df.re});
      }
    }
  }
}

// The rows of lane t: (t / 64) RPW + (t % 64) / ZR2, then every NW RPW rows
// further.  NY is a multiple of RPW, so all lanes of a wavefront run the
// loop alike.
#define KAPSEL_PLANE_ROWS(C)                        \
  const int n2 = threadIdx.x % 64 % C::ZR2;         \
  for (int j = threadIdx.x / 64 * C::RPW + threadIdx.x % 64 / C::ZR2; j < NY; j += C::NW * C::RPW)

// the register budget of a plane pass: as many wavefronts per SIMD as the
// workgroups its LDS lets share a CU (at most two) bring
template <typename T, int NY, int NZ, int THREADS>
constexpr int plane_waves_per_simd() {
  return PlaneCfg<T, NY, NZ, THREADS>::WAVES_PER_SIMD;
}
#define KAPSEL_PLANE_BOUNDS __launch_bounds__(THREADS, (plane_waves_per_simd<T, NY, NZ, THREADS>()))

// The set-up of a plane pass that starts with the y transforms of a plane
// in global memory (at g), and those transforms.  All units of a launch
// start together, so loads waited for one after the other would cost a
// memory latency each: the loads of the packed columns and of the
// twiddles, whose LDS stores wait for them, and those of the first round
// of columns are all issued before the first wait.
template <typename T, int NY, int NZ, int THREADS, int DIR, int YS = 1>
__device__ __forceinline__ void plane_prologue_columns(V2<T> *P, V2<T> *pk, PlaneTw<T, NY, NZ> &tw,
                                                       const V2<T> *__restrict__ twm, const V2<T> *__restrict__ zrot,
                                                       const V2<T> *__restrict__ twy, T *g) {
  if constexpr (THREADS >= NZ / 2 + 1 && THREADS >= NY) {
    PlanePackRegs<T, NY> pr;
    plane_pack_fetch<T, NY, NZ, THREADS>(pr, g);
    const PlaneTwRegs<T> twr = plane_tw_fetch<T, NY, NZ, THREADS, YS>(twm, zrot, twy);
    plane_columns<T, NY, NZ, THREADS, DIR, true, false>(P, pk, tw.twy, g, [&] {
      plane_pack_fold<T, NY, NZ, THREADS>(pk, pr);
      plane_tw_put<T, NY, NZ>(tw, twr);
      __syncthreads();
    });
  } else {
    plane_pack_in<T, NY, NZ, THREADS>(pk, g);
    plane_tw_load<T, NY, NZ, THREADS, YS>(tw, twm, zrot, twy);
    __syncthreads();
    plane_columns<T, NY, NZ, THREADS, DIR, true, false>(P, pk, tw.twy, g);
  }
}

template <typename T, int NY, int NZ, int THREADS>
__global__ KAPSEL_PLANE_BOUNDS void fft_plane_r2c(FieldPtrs<T> fp, const V2<T> *__restrict__ twm,
                                                           const V2<T> *__restrict__ zrot,
                                                           const V2<T> *__restrict__ twy) {
  using C = PlaneCfg<T, NY, NZ, THREADS>;
  __shared__ __align__(16) V2<T> P[C::LPLANE];
  __shared__ __align__(16) V2<T> pk[C::PK];
  __shared__ PlaneTw<T, NY, NZ> tw;
  unsigned field;
  const unsigned plane = xcd_tile(field);
  T *__restrict__ gp = fp.p[field] + (size_t)plane * 2 * C::PLANE;
  const int n2 = threadIdx.x % 64 % C::ZR2, j0 = threadIdx.x / 64 * C::RPW + threadIdx.x % 64 / C::ZR2;
  V2<T> v[C::ZR1];
  auto row_in = [&](int j) {
    const T *grow = gp + (size_t)j * 2 * C::H;
#pragma unroll
    for (int n1 = 0; n1 < C::ZR1; n1++) v[n1] = ld2(grow + 2 * (C::ZR2 * n1 + n2));
  };
  // the first row's loads go out with the twiddles'
  plane_tw_load_around<T, NY, NZ, THREADS>(tw, twm, zrot, twy, [&] {
    if (j0 < NY) row_in(j0);
  });
  __syncthreads();
  for (int j = j0; j < NY; j += C::NW * C::RPW) {
    if (j != j0) row_in(j);
    row_r2c<T, C::M, C::ZR1, C::ZR2>(P + j * C::HP, v, tw.twm, tw.zrot, n2);
  }
  __syncthreads();
  plane_columns<T, NY, NZ, THREADS, -1, false, true>(P, pk, tw.twy, gp);
}

// fft_plane_r2c of a field formed on the fly: out_f = a_f + s (b_f + e_f)
// c_f (a_f NULL: no a term; c_f NULL: a_f + s (b_f + e_f); bit f of dsq:
// (b_f - c_f)(b_f + c_f) in place of the product), read pairwise where the
// row stage reads its input -- the pointwise pass before a forward transform
// (f = u_p - phi u of Make_f_particle_dt_nonsole, u + f dt/rho of the
// dielectric force, the products of U2advection_k, the Coulomb force
// rho (E - grad psi)) without its own read and write of the fields.  out_f
// may be one of field f's own inputs (the plane is read before it is
// written), but no other field's input.  The fields of one plane
// run at consecutive ids (xcd_tile_by_tile), so the planes they read in
// common (phi; the velocity components of the products) come from HBM once.
template <typename T, int NY, int NZ, int THREADS>
__global__ KAPSEL_PLANE_BOUNDS void fft_plane_r2c_pro(FieldPtrs<T> out, FieldPtrs<T> a, FieldPtrs<T> b,
                                                               FieldPtrs<T> c, FieldConsts<T> e, T s, unsigned dsq,
                                                               const V2<T> *__restrict__ twm,
                                                               const V2<T> *__restrict__ zrot,
                                                               const V2<T> *__restrict__ twy) {
  using C = PlaneCfg<T, NY, NZ, THREADS>;
  __shared__ __align__(16) V2<T> P[C::LPLANE];
  __shared__ __align__(16) V2<T> pk[C::PK];
  __shared__ PlaneTw<T, NY, NZ> tw;
  unsigned field;
  const unsigned plane = xcd_tile_by_tile(field);
  const size_t base = (size_t)plane * 2 * C::PLANE;
  T *__restrict__ gp = out.p[field] + base;
  const T *ga = a.p[field] ? a.p[field] + base : nullptr, *gb = b.p[field] + base;
  const T *gc = c.p[field] ? c.p[field] + base : nullptr;
  const bool diff = (dsq >> field) & 1;
  const T ef = e.v[field];
  const int n2 = threadIdx.x % 64 % C::ZR2, j0 = threadIdx.x / 64 * C::RPW + threadIdx.x % 64 / C::ZR2;
  V2<T> v[C::ZR1];
  auto row_in = [&](int j) {
    const size_t ro = (size_t)j * 2 * C::H;
#pragma unroll
    for (int n1 = 0; n1 < C::ZR1; n1++) {
      const size_t o = ro + 2 * (C::ZR2 * n1 + n2);
      V2<T> vb = ld2(gb + o);
      vb = V2<T>{vb.re + ef, vb.im + ef};
      if (gc) {
        const V2<T> vc = ld2(gc + o);
        vb = diff ? V2<T>{(vb.re - vc.re) * (vb.re + vc.re), (vb.im - vc.im) * (vb.im + vc.im)}
                  : V2<T>{vb.re * vc.re, vb.im * vc.im};
      }
      if (ga) {
        const V2<T> va = ld2(ga + o);
        v[n1] = V2<T>{va.re + s * vb.re, va.im + s * vb.im};
      } else {
        v[n1] = V2<T>{s * vb.re, s * vb.im};
      }
    }
  };
  // the first row's loads go out with the twiddles'
  plane_tw_load_around<T, NY, NZ, THREADS>(tw, twm, zrot, twy, [&] {
    if (j0 < NY) row_in(j0);
  });
  __syncthreads();
  for (int j = j0; j < NY; j += C::NW * C::RPW) {
    if (j != j0) row_in(j);
    row_r2c<T, C::M, C::ZR1, C::ZR2>(P + j * C::HP, v, tw.twm, tw.zrot, n2);
  }
  __syncthreads();
  plane_columns<T, NY, NZ, THREADS, -1, false, true>(P, pk, tw.twy, gp);
}

template <typename T, int NY, int NZ, int THREADS>
__global__ KAPSEL_PLANE_BOUNDS void fft_plane_c2r(FieldPtrs<T> fp, const V2<T> *__restrict__ twm,
                                                           const V2<T> *__restrict__ zrot,
                                                           const V2<T> *__restrict__ twy, T scale) {
  using C = PlaneCfg<T, NY, NZ, THREADS>;
  __shared__ __align__(16) V2<T> P[C::LPLANE];
  __shared__ __align__(16) V2<T> pk[C::PK];
  __shared__ PlaneTw<T, NY, NZ> tw;
  unsigned field;
  const unsigned plane = xcd_tile(field);
  T *__restrict__ gp = fp.p[field] + (size_t)plane * 2 * C::PLANE;
  plane_prologue_columns<T, NY, NZ, THREADS, +1>(P, pk, tw, twm, zrot, twy, gp);
  KAPSEL_PLANE_ROWS(C) {
    T *grow = gp + (size_t)j * 2 * C::H;
    row_c2r<T, C::M, C::ZR1, C::ZR2>(P + j * C::HP, tw.twm, tw.zrot, scale, n2,
                                     [&](int n1, V2<T> z) { st2(grow + 2 * (C::ZR2 * n1 + n2), z); });
    if (n2 == 0) st2(grow + 2 * C::M, V2<T>{0, 0});  // the row padding
  }
}

// The middle of the dielectric operator on one plane of the staggered
// gradient field d (fields 0, 1, 2 = x, y, z): inverse z and y transforms,
// times the face permittivity (a real field of the same layout), forward z
// and y transforms -- one read and one write where two plane passes and a
// pointwise pass would make three of each.  A row goes from its inverse z
// transform through the face product into its forward z transform in the
// registers of its own lanes.
template <typename T, int NY, int NZ, int THREADS>
__global__ KAPSEL_PLANE_BOUNDS void fft_plane_face(FieldPtrs<T> fp, const V2<T> *__restrict__ twm,
                                                            const V2<T> *__restrict__ zrot,
                                                            const V2<T> *__restrict__ twy, T scale,
                                                            FieldPtrs<T> face) {
  using C = PlaneCfg<T, NY, NZ, THREADS>;
  __shared__ __align__(16) V2<T> P[C::LPLANE];
  __shared__ __align__(16) V2<T> pk[C::PK];
  __shared__ PlaneTw<T, NY, NZ> tw;
  unsigned field;
  const unsigned plane = xcd_tile(field);
  T *__restrict__ gp = fp.p[field] + (size_t)plane * 2 * C::PLANE;
  const T *fpl = face.p[field] + (size_t)plane * 2 * C::PLANE;
  plane_prologue_columns<T, NY, NZ, THREADS, +1>(P, pk, tw, twm, zrot, twy, gp);
  KAPSEL_PLANE_ROWS(C) {
    const T *frow = fpl + (size_t)j * 2 * C::H;
    V2<T> fv[C::ZR1], v[C::ZR1];
#pragma unroll
    for (int n1 = 0; n1 < C::ZR1; n1++) fv[n1] = ld2(frow + 2 * (C::ZR2 * n1 + n2));
    V2<T> *row = P + j * C::HP;
    row_c2r<T, C::M, C::ZR1, C::ZR2>(row, tw.twm, tw.zrot, scale, n2, [&](int n1, V2<T> z) {
      v[n1] = V2<T>{z.re * fv[n1].re, z.im * fv[n1].im};
    });
    row_r2c<T, C::M, C::ZR1, C::ZR2>(row, v, tw.twm, tw.zrot, n2);
  }
  __syncthreads();
  plane_columns<T, NY, NZ, THREADS, -1, false, true>(P, pk, tw.twy, gp);
}

// fft_plane_face on the y-split gradient fields that fft_x_grad_dec
// writes: half h of plane i (rows hJ .. hJ + J - 1, J = NY/2, contiguous)
// holds the J-point spectra whose inverse y transforms are the real rows
// 2m + h.  One workgroup per (field, plane, half): inverse z and J-point y
// transforms, times the face permittivity of row 2m + h, forward z and
// J-point y transforms, the result (E for the even rows, O for the odd)
// back into the half; fft_x_div_dec combines E(ky) + W^ky O(ky) and
// E(ky) - W^ky O(ky) into rows ky and ky + J.  Twice the units of
// fft_plane_face at half the size each (768 at 128^3: three per CU), the
// two halves of a plane at consecutive ids.
template <typename T, int NY, int NZ, int THREADS>
constexpr int half_waves_per_simd() {
  using C = PlaneCfg<T, NY / 2, NZ, THREADS>;
  constexpr int wgs = C::LDS_BYTES * 3 <= 160 * 1024 ? 3 : C::LDS_BYTES * 2 <= 160 * 1024 ? 2 : 1;
  return (wgs * C::NW + 3) / 4;
}
template <typename T, int NY, int NZ, int THREADS>
__global__ __launch_bounds__(THREADS, (half_waves_per_simd<T, NY, NZ, THREADS>())) void fft_half_face(
    FieldPtrs<T> fp, const V2<T> *__restrict__ twm, const V2<T> *__restrict__ zrot, const V2<T> *__restrict__ twy,
    T scale, FieldPtrs<T> face) {
  constexpr int J = NY / 2;
  using C = PlaneCfg<T, J, NZ, THREADS>;
  __shared__ __align__(16) V2<T> P[C::LPLANE];
  __shared__ __align__(16) V2<T> pk[C::PK];
  __shared__ PlaneTw<T, J, NZ> tw;
  unsigned field;
  const unsigned u = xcd_tile(field);
  const unsigned plane = u >> 1, hf = u & 1;
  T *__restrict__ gp = fp.p[field] + (size_t)plane * 2 * (NY * C::H) + (size_t)hf * 2 * C::PLANE;
  const T *fpl = face.p[field] + (size_t)plane * 2 * (NY * C::H);
  plane_prologue_columns<T, J, NZ, THREADS, +1, 2>(P, pk, tw, twm, zrot, twy, gp);
  const int n2 = threadIdx.x % 64 % C::ZR2;
  for (int m = threadIdx.x / 64 * C::RPW + threadIdx.x % 64 / C::ZR2; m < J; m += C::NW * C::RPW) {
    // the face row is loaded here: issued with the columns it would slow
    // their arrival, which the first transforms wait for
    const T *frow = fpl + (size_t)(2 * m + hf) * 2 * C::H;
    V2<T> fv[C::ZR1], v[C::ZR1];
#pragma unroll
    for (int n1 = 0; n1 < C::ZR1; n1++) fv[n1] = ld2(frow + 2 * (C::ZR2 * n1 + n2));
    V2<T> *row = P + m * C::HP;
    row_c2r<T, C::M, C::ZR1, C::ZR2>(row, tw.twm, tw.zrot, scale, n2, [&](int n1, V2<T> z) {
      v[n1] = V2<T>{z.re * fv[n1].re, z.im * fv[n1].im};
    });
    row_r2c<T, C::M, C::ZR1, C::ZR2>(row, v, tw.twm, tw.zrot, n2);
  }
  __syncthreads();
  plane_columns<T, J, NZ, THREADS, -1, false, true>(P, pk, tw.twy, gp);
}

// One z row of each of the three staggered-gradient spectra (after the x
// and y inverse passes): c2r along z, times the face permittivity row, r2c
// along z -- the middle of the dielectric operator where no plane pass
// applies.  The row stays in LDS between the two transforms.
template <typename T, int NZ, int RPB>
__global__ __launch_bounds__(RPB *FS<NZ / 2>::R2) void fft_z_face(FieldPtrs<T> fp, long long rows, int nz_,
                                                                   const V2<T> *__restrict__ twm,
                                                                   const V2<T> *__restrict__ zrot, T scale,
                                                                   FieldPtrs<T> face) {
  constexpr int M = NZ / 2;
  constexpr int R1 = FS<M>::R1, R2 = FS<M>::R2, TPL = R2;
  constexpr int RS = R2 * (R1 + 1) + 1;
  static_assert(RS >= M + 1, "row stride");
  __shared__ V2<T> sm[RPB * RS];
  unsigned field;
  const unsigned tile = xcd_tile(field);
  const int t = threadIdx.x;
  const int n2 = t % TPL;
  const int r = t / TPL;
  const long long row = (long long)tile * RPB + r;
  const bool valid = row < rows;
  T *__restrict__ rp = fp.p[field] + (valid ? row : 0) * nz_;
  const T *__restrict__ fr = face.p[field] + (valid ? row : 0) * nz_;
  V2<T> *sr = sm + r * RS;

  // c2r (the kz = 0 and kz = NZ/2 entries are taken as real)
  for (int k = n2; k <= M; k += TPL) {
    V2<T> x = valid ? ld2(rp + 2 * k) : V2<T>{0, 0};
    if (k == 0 || k == M) x.im = 0;
    sr[k] = x;
  }
  __syncthreads();
  V2<T> v[R1], o[R1];
#pragma unroll
  for (int n1 = 0; n1 < R1; n1++) {
    const int n = R2 * n1 + n2;
    V2<T> xn = sr[n], xm = conjv(sr[M - n]);
    V2<T> e = xn + xm;
    V2<T> od = cmul(xn - xm, conjv(zrot[n]));
    v[n1] = V2<T>{e.re - od.im, e.im + od.re};
  }
  dft<R1, 1, +1>(v, o);
#pragma unroll
  for (int k1 = 1; k1 < R1; k1++) o[k1] = cmul(o[k1], tw_get<+1>(twm, n2 * k1));
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < R1; k1++) sr[n2 * (R1 + 1) + k1] = o[k1];
  __syncthreads();
  V2<T> zz[R1 / TPL][R2];
#pragma unroll
  for (int c = 0; c < R1 / TPL; c++) {
    const int k1 = n2 + TPL * c;
    V2<T> w[R2], x[R2];
#pragma unroll
    for (int m = 0; m < R2; m++) w[m] = sr[m * (R1 + 1) + k1];
    dft<R2, 1, +1>(w, x);
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) {
      // reals 2n, 2n+1 of the row, times the face permittivity
      const int n = k1 + R1 * k2;
      const V2<T> sx = scal(scale, x[k2]);
      zz[c][k2] = valid ? V2<T>{sx.re * fr[2 * n], sx.im * fr[2 * n + 1]} : V2<T>{0, 0};
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < R1 / TPL; c++) {
    const int k1 = n2 + TPL * c;
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) sr[k1 + R1 * k2] = zz[c][k2];
  }
  __syncthreads();
  // r2c of the product
#pragma unroll
  for (int n1 = 0; n1 < R1; n1++) v[n1] = sr[R2 * n1 + n2];
  dft<R1, 1, -1>(v, o);
#pragma unroll
  for (int k1 = 1; k1 < R1; k1++) o[k1] = cmul(o[k1], tw_get<-1>(twm, n2 * k1));
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < R1; k1++) sr[n2 * (R1 + 1) + k1] = o[k1];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < R1 / TPL; c++) {
    const int k1 = n2 + TPL * c;
    V2<T> w[R2];
#pragma unroll
    for (int m = 0; m < R2; m++) w[m] = sr[m * (R1 + 1) + k1];
    dft<R2, 1, -1>(w, zz[c]);
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < R1 / TPL; c++) {
    const int k1 = n2 + TPL * c;
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) sr[k1 + R1 * k2] = zz[c][k2];
  }
  __syncthreads();
  if (valid) {
    for (int k = n2; k <= M; k += TPL) {
      V2<T> zk = sr[k & (M - 1)];
      V2<T> zm = conjv(sr[(M - k) & (M - 1)]);
      V2<T> e = scal((T)0.5, zk + zm);
      V2<T> d = zk - zm;
      V2<T> od = {(T)0.5 * d.im, (T)-0.5 * d.re};
      st2(rp + 2 * k, e + cmul(zrot[k], od));
    }
  }
}

// ---------------------------------------- x passes fused with the CG algebra
// The dielectric operator applied to a CG direction p in k-space: the first
// pass (x, inverse) also forms p = z + beta p and the staggered gradient
// symbols; the last pass (x, forward) also forms the staggered divergence and
// <p, A p>.  x line q = j*h + kz: elements q + i*plane, i = 0..NX-1.

template <int THREADS>
__device__ __forceinline__ double block_sum_fft(double v) {
  __shared__ double red[THREADS / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < THREADS / 64; w++) s += red[w];
  return s;
}

// The P points of a line's row of p = zs z + beta p (update; stored), of p =
// z - sub (sub; stored) or of z (neither), at the offsets off(n): every
// load is issued before the first store (a store to p may alias a later
// point's load for all the compiler knows, which would otherwise hold each
// point's loads until the store before it is issued).  Zero for an invalid
// line.
template <typename T, int P, typename Off>
__device__ __forceinline__ void p_update_row(V2<T> (&x)[P], const T *z, T *p, const float *sub, bool update, T zs,
                                             T beta, bool valid, Off off) {
  if (!valid) {
#pragma unroll
    for (int n = 0; n < P; n++) x[n] = V2<T>{0, 0};
    return;
  }
  if (update) {
    V2<T> b[P];
#pragma unroll
    for (int n = 0; n < P; n++) x[n] = ld2(z + 2 * off(n));
#pragma unroll
    for (int n = 0; n < P; n++) b[n] = ld2(p + 2 * off(n));
#pragma unroll
    for (int n = 0; n < P; n++) x[n] = scal(zs, x[n]) + scal(beta, b[n]);
  } else if (sub) {
    V2<float> d[P];
#pragma unroll
    for (int n = 0; n < P; n++) x[n] = ld2(z + 2 * off(n));
#pragma unroll
    for (int n = 0; n < P; n++) d[n] = ld2(sub + 2 * off(n));
#pragma unroll
    for (int n = 0; n < P; n++) x[n] = V2<T>{x[n].re - (T)d[n].re, x[n].im - (T)d[n].im};
  } else {
#pragma unroll
    for (int n = 0; n < P; n++) x[n] = ld2(z + 2 * off(n));
    return;
  }
#pragma unroll
  for (int n = 0; n < P; n++) st2(p + 2 * off(n), x[n]);
}

// p = zscale z + sc[beta] p (stored), then g_d = IFFT_x(i k_d e^{i k_d dx/2}
// p), d = x, y, z (thread t of a tile: line t % LPB, its thread s = t /
// LPB); zscale 2 with beta -1 is the linear extrapolation of a start in
// time; with sub (and no beta): p = z - sub (stored), the refinement's x -=
// d before its residual
template <typename T, int L, int LPB>
__global__ __launch_bounds__((LPB * LineFFT<T, L>::S)) void fft_x_grad(
    const T *z, T *p, const double *__restrict__ sc, int beta_slot, T zscale, FieldPtrs<T> g3, long long Q,
    long long plane,
    int h, int ny, double wx, double wy, double wz, const V2<T> *__restrict__ shx, const V2<T> *__restrict__ shy,
    const V2<T> *__restrict__ shz, const V2<T> *__restrict__ tw, const float *__restrict__ sub) {
  using F = LineFFT<T, L>;
  constexpr int P = F::P;
  __shared__ V2<T> sm[LPB * F::LS];
  unsigned field;
  const unsigned tile = xcd_tile(field);
  const int t = threadIdx.x;
  const int l = t % LPB, s = t / LPB;
  const long long q = (long long)tile * LPB + l;
  const bool valid = q < Q;
  const long long base = valid ? q : 0;
  const int j = (int)(base / h), kz = (int)(base % h);
  const bool update = beta_slot >= 0;  // else the gradient of z itself, p untouched
  const T beta = update ? (T)sc[beta_slot] : (T)0;
  V2<T> pv[P];
  p_update_row<T, P>(pv, z, p, sub, update, zscale, beta, valid,
                     [&](int n) { return base + (long long)F::in_i(s, n) * plane; });
  V2<T> *sl = sm + l * F::LS;
  V2<T> twr[P];
  line_twiddles<F>(twr, tw, s);
  // G_x = IFFT(i k_x e^{i k_x dx/2} p); the y and z symbols are constant
  // along the line: G_y = c_y IFFT(p), G_z = c_z IFFT(p) (two transforms)
  V2<T> v[P];
#pragma unroll
  for (int n = 0; n < P; n++) {
    const int i = F::in_i(s, n);
    const T w = (T)(signed_k(i, L) * wx);
    const V2<T> e = shx[i];
    v[n] = cmul(pv[n], V2<T>{-w * e.im, w * e.re});
  }
  F::template run<+1>(v, sl, twr, s, LPB);
  if (valid) {  // (in flight while IFFT(p) runs)
#pragma unroll
    for (int m = 0; m < P; m++) st2(g3.p[0] + 2 * (base + (long long)F::out_i(s, m) * plane), v[m]);
  }
  F::template run<+1>(pv, sl, twr, s, LPB);
  const T ky = (T)(signed_k(j, ny) * wy), kzw = (T)(kz * wz);
  const V2<T> ey = shy[j], ez = shz[kz];
  const V2<T> cy{-ky * ey.im, ky * ey.re}, cz{-kzw * ez.im, kzw * ez.re};
  if (valid) {
#pragma unroll
    for (int m = 0; m < P; m++) {
      const long long off = 2 * (base + (long long)F::out_i(s, m) * plane);
      st2(g3.p[1] + off, cmul(pv[m], cy));
      st2(g3.p[2] + off, cmul(pv[m], cz));
    }
  }
}

// the line transforms of fft_x_div and fft_x_div_dec: the pair split at 128
// points in FP64 only (in FP32 the four-step 16 x 8 measured faster for
// these kernels; fft_x_grad_dec is faster with the split in both)
template <typename T, int L>
using DivFFT = LineFFT<T, L, L == 128 && sizeof(T) == 8>;

// fft_x_grad with the gradients left y-split for fft_half_face: with G_d(j)
// the gradient spectrum of row j (ky index) and J = NY/2, row j' < J
// receives G_d(j') + G_d(j' + J) and row J + j' receives (G_d(j') -
// G_d(j' + J)) W^{-j'} (W = e^{-2 pi i / NY}): the inputs of the J-point
// inverse y transforms of the even and of the odd real rows (one radix-2
// step of the y transform, decimation in frequency).  A tile's LPB line
// slots are LPB/2 row pairs (j', kz), (j' + J, kz): slot l < LPB/2 takes
// row j', slot l + LPB/2 row j' + J (Q = J h pairs).  Each line loads,
// updates and stores its own row p_r and transforms it twice: X_r =
// IFFT_x(p_r) and Y_r = IFFT_x(i k_x e^{i k_x dx/2} p_r).  The y and z
// symbols are constant along a line, so by linearity G_x(r) = Y_r, G_y(r) =
// c_y(r) X_r and G_z(r) = c_z X_r; the two lines of a pair (their threads
// with the same s are LPB/2 lanes apart, in one 32-lane half of a
// wavefront) swap X and Y by a lane swizzle and form the sum or the
// twiddled difference -- two transforms per line where forming the inputs
// first would take three.  32-bit offsets (a field holds fewer than 2^31
// values).
template <typename T, int L, int LPB>
__global__ __launch_bounds__((LPB * LineFFT<T, L>::S)) void fft_x_grad_dec(
    const T *z, T *p, const double *__restrict__ sc, int beta_slot, T zscale, FieldPtrs<T> g3, unsigned Q,
    unsigned plane,
    int h, int ny, double wx, double wy, double wz, const V2<T> *__restrict__ shx, const V2<T> *__restrict__ shy,
    const V2<T> *__restrict__ shz, const V2<T> *__restrict__ tw, const float *__restrict__ sub,
    const V2<T> *__restrict__ twy) {
  using F = LineFFT<T, L>;
  constexpr int P = F::P, LPT = LPB / 2;
  static_assert(64 % LPB == 0 && LPT < 32, "the two lines of a pair share a 32-lane half of a wavefront");
  __shared__ V2<T> sm[LPB * F::LS];
  unsigned field;
  const unsigned tile = xcd_tile(field);
  const int t = threadIdx.x;
  const int l = t % LPB, s = t / LPB;
  const int mb = l / LPT;  // 0: row j' (the sum), 1: row j' + J (the twiddled difference)
  const unsigned q = tile * LPT + l % LPT;
  const bool valid = q < Q;
  const unsigned base = valid ? q : 0, half = (unsigned)(ny / 2) * h, own = base + mb * half;
  const int j = (int)(base / h), kz = (int)(base % h);
  const bool update = beta_slot >= 0;
  const T beta = update ? (T)sc[beta_slot] : (T)0;
  V2<T> *sl = sm + l * F::LS;
  V2<T> x[P], y[P];
  p_update_row<T, P>(x, z, p, sub, update, zscale, beta, valid, [&](int n) { return own + F::in_i(s, n) * plane; });
  // the x gradient symbol i k_x e^{i k_x dx/2} per input point
#pragma unroll
  for (int n = 0; n < P; n++) {
    const int i = F::in_i(s, n);
    const V2<T> e = shx[i];
    const T w = (T)(signed_k(i, L) * wx);
    y[n] = cmul(x[n], V2<T>{-w * e.im, w * e.re});
  }
  V2<T> twr[P];
  line_twiddles<F>(twr, tw, s);
  V2<T> wj = tw_get<+1>(twy, j);
  if (!mb) wj = V2<T>{1, 0};
  const T sg = mb ? (T)-1 : (T)1;
  F::template run<+1>(y, sl, twr, s, LPB);
  F::template run<+1>(x, sl, twr, s, LPB);
  T *__restrict__ gx = g3.p[0], *__restrict__ gy = g3.p[1], *__restrict__ gz = g3.p[2];
#pragma unroll
  for (int m = 0; m < P; m++) {
    const V2<T> yo = swz_xor<LPT>(y[m]);
    const V2<T> y0 = mb ? yo : y[m], y1 = mb ? y[m] : yo;
    if (valid) st2(gx + 2 * (own + F::out_i(s, m) * plane), cmul(y0 + scal(sg, y1), wj));
  }
  // the y symbols of the pair's rows j' and j' + J, the z symbol
  const T k0 = (T)(signed_k(j, ny) * wy), k1 = (T)(signed_k(j + ny / 2, ny) * wy);
  const V2<T> e0 = shy[j], e1 = shy[j + ny / 2], ez = shz[kz];
  const T kzw = (T)(kz * wz);
  const V2<T> c0{-k0 * e0.im, k0 * e0.re}, c1{-k1 * e1.im, k1 * e1.re}, cz{-kzw * ez.im, kzw * ez.re};
#pragma unroll
  for (int m = 0; m < P; m++) {
    const V2<T> xo = swz_xor<LPT>(x[m]);
    const V2<T> x0 = mb ? xo : x[m], x1 = mb ? x[m] : xo;
    const V2<T> vy = cmul(cmul(x0, c0) + scal(sg, cmul(x1, c1)), wj);
    const V2<T> vz = cmul(cmul(x0 + scal(sg, x1), wj), cz);
    if (valid) {
      const unsigned off = 2 * (own + F::out_i(s, m) * plane);
      st2(gy + off, vy);
      st2(gz + off, vz);
    }
  }
}

// The staggered gradients of two FP64 half spectra a, b (grid.y: 0 = a,
// 1 = b) as six FP32 fields after their inverse x transform:
// out[3f + d] = IFFT_x(i k_d e^{i k_d dx/2} in_f).  The x product is formed
// in FP64 and rounded once to FP32 (the value grad_shift2_f64_kernel
// stores); the y and z symbols are constant along the line, so those two
// gradients are the FP32 transform of the rounded input times the rounded
// symbol (two transforms per line); the transforms run in FP32.
template <int L, int LPB>
__global__ __launch_bounds__((LPB * LineFFT<float, L>::S)) void fft_x_grad2_f64(
    const double *__restrict__ a, const double *__restrict__ b, FieldPtrs<float> out, long long Q, long long plane,
    int h, int ny, double wx, double wy, double wz, const V2<double> *__restrict__ shx,
    const V2<double> *__restrict__ shy, const V2<double> *__restrict__ shz, const V2<float> *__restrict__ tw) {
  using F = LineFFT<float, L>;
  constexpr int P = F::P;
  __shared__ V2<float> sm[LPB * F::LS];
  unsigned field;
  const unsigned tile = xcd_tile(field);
  const double *in = field == 0 ? a : b;
  const int t = threadIdx.x;
  const int l = t % LPB, s = t / LPB;
  const long long q = (long long)tile * LPB + l;
  const bool valid = q < Q;
  const long long base = valid ? q : 0;
  const int j = (int)(base / h), kz = (int)(base % h);
  V2<double> xv[P];
#pragma unroll
  for (int n = 0; n < P; n++) xv[n] = ld2(in + 2 * (base + (long long)F::in_i(s, n) * plane));
  V2<float> *sl = sm + l * F::LS;
  V2<float> twr[P];
  line_twiddles<F>(twr, tw, s);
  V2<float> v[P], x[P];
#pragma unroll
  for (int n = 0; n < P; n++) {
    const int i = F::in_i(s, n);
    const double w = signed_k(i, L) * wx;
    const V2<double> gv = cmul(V2<double>{-w * xv[n].im, w * xv[n].re}, shx[i]);
    v[n] = V2<float>{(float)gv.re, (float)gv.im};
    x[n] = V2<float>{(float)xv[n].re, (float)xv[n].im};
  }
  F::template run<+1>(v, sl, twr, s, LPB);
  if (valid) {  // (in flight while the second transform runs)
#pragma unroll
    for (int m = 0; m < P; m++) st2(out.p[3 * field] + 2 * (base + (long long)F::out_i(s, m) * plane), v[m]);
  }
  F::template run<+1>(x, sl, twr, s, LPB);
  const double ky = signed_k(j, ny) * wy, kzw = kz * wz;
  const V2<double> ey = shy[j], ez = shz[kz];
  const V2<float> cy{(float)(-ky * ey.im), (float)(ky * ey.re)}, cz{(float)(-kzw * ez.im), (float)(kzw * ez.re)};
  if (valid) {
#pragma unroll
    for (int m = 0; m < P; m++) {
      const long long off = 2 * (base + (long long)F::out_i(s, m) * plane);
      st2(out.p[3 * field + 1] + off, cmul(x[m], cy));
      st2(out.p[3 * field + 2] + off, cmul(x[m], cz));
    }
  }
}

// w_d = FFT_x(g_d), ap = i sum_d k_d e^{-i k_d dx/2} w_d (stored);
// partial[tile] = sum over the tile of w(kz) Re(conj(p) ap), w = 1 on the
// kz = 0 and NZ/2 planes and 2 elsewhere (Parseval, without N^-3)
// RES: the residual form -- p is b, ap receives r = (div) - b, and the
// partials sum |r|^2 (the refinement's A x - b)
// (x_div's line transforms: DivFFT, above)
template <typename T, int L, int LPB, bool RES = false>
__global__ __launch_bounds__((LPB * DivFFT<T, L>::S)) void fft_x_div(
    FieldPtrs<T> g3, const T *__restrict__ p, T *__restrict__ ap, double *__restrict__ partial, long long Q,
    long long plane, int h, int ny, int nz, double wx, double wy, double wz, const V2<T> *__restrict__ shx,
    const V2<T> *__restrict__ shy, const V2<T> *__restrict__ shz, const V2<T> *__restrict__ tw) {
  using F = DivFFT<T, L>;
  constexpr int P = F::P;
  __shared__ V2<T> sm[LPB * F::LS];
  unsigned field;
  const unsigned tile = xcd_tile(field);
  const int t = threadIdx.x;
  const int l = t % LPB, s = t / LPB;
  const long long q = (long long)tile * LPB + l;
  const bool valid = q < Q;
  const long long base = valid ? q : 0;
  const int j = (int)(base / h), kz = (int)(base % h);
  V2<T> *sl = sm + l * F::LS;
  // The y and z divergence symbols q_d conj(e^{i k_d dx/2}) are constant
  // along the line, so their terms are summed before one transform, which
  // runs while the x input loads (base is 0 for an invalid line, so every
  // load is in bounds).
  V2<T> vy[P], vz[P];
#pragma unroll
  for (int n = 0; n < P; n++) vy[n] = ld2(g3.p[1] + 2 * (base + (long long)F::in_i(s, n) * plane));
#pragma unroll
  for (int n = 0; n < P; n++) vz[n] = ld2(g3.p[2] + 2 * (base + (long long)F::in_i(s, n) * plane));
  V2<T> twr[P];
  line_twiddles<F>(twr, tw, s);
  const V2<T> fy = scal((T)(signed_k(j, ny) * wy), conjv(shy[j]));
  const V2<T> fz = scal((T)(kz * wz), conjv(shz[kz]));
  V2<T> acc[P];
#pragma unroll
  for (int n = 0; n < P; n++) acc[n] = valid ? cmul(vy[n], fy) + cmul(vz[n], fz) : V2<T>{0, 0};
  V2<T> vx[P];
#pragma unroll
  for (int n = 0; n < P; n++) vx[n] = ld2(g3.p[0] + 2 * (base + (long long)F::in_i(s, n) * plane));
  F::template run<-1>(acc, sl, twr, s, LPB);
#pragma unroll
  for (int n = 0; n < P; n++) vx[n] = valid ? vx[n] : V2<T>{0, 0};
  F::template run<-1>(vx, sl, twr, s, LPB);
#pragma unroll
  for (int m = 0; m < P; m++) {
    const int i = F::out_i(s, m);
    acc[m] = acc[m] + cmul(vx[m], scal((T)(signed_k(i, L) * wx), conjv(shx[i])));
  }
  const double wk = (kz == 0 || 2 * kz == nz) ? 1.0 : 2.0;
  double sum = 0.0;
  if (valid) {
#pragma unroll
    for (int m = 0; m < P; m++) {
      const long long off = base + (long long)F::out_i(s, m) * plane;
      const V2<T> av{-acc[m].im, acc[m].re};
      const V2<T> pp = ld2(p + 2 * off);
      if constexpr (RES) {
        const V2<T> rv = av - pp;
        st2(ap + 2 * off, rv);
        sum += wk * ((double)rv.re * rv.re + (double)rv.im * rv.im);
      } else {
        st2(ap + 2 * off, av);
        sum += wk * ((double)pp.re * av.re + (double)pp.im * av.im);
      }
    }
  }
  sum = block_sum_fft<LPB * F::S>(sum);
  if (threadIdx.x == 0) partial[tile] = sum;
}

// fft_x_div on the y-split output of fft_half_face (rows j' < J = NY/2: E,
// rows J + j': O): a tile's line slots are row pairs as in fft_x_grad_dec;
// the line of row j' loads E, the line of row j' + J loads O, they swap
// them by a lane swizzle and form E + W^{j'} O and E - W^{j'} O (W =
// e^{-2 pi i / NY}), the spectra of rows j' and j' + J, before their
// forward x transforms.  The y and z divergence symbols are constant along
// a line, so their terms are summed before one transform (FFT(v_y) f_y +
// FFT(v_z) f_z = FFT(f_y v_y + f_z v_z)): two transforms per line, the y-z
// one running while the x input loads.  32-bit offsets.
template <typename T, int L, int LPB, bool RES = false>
__global__ __launch_bounds__((LPB * DivFFT<T, L>::S), 2) void fft_x_div_dec(
    FieldPtrs<T> g3, const T *__restrict__ p, T *__restrict__ ap, double *__restrict__ partial, unsigned Q,
    unsigned plane, int h, int ny, int nz, double wx, double wy, double wz, const V2<T> *__restrict__ shx,
    const V2<T> *__restrict__ shy, const V2<T> *__restrict__ shz, const V2<T> *__restrict__ tw,
    const V2<T> *__restrict__ twy) {
  using F = DivFFT<T, L>;
  constexpr int P = F::P, LPT = LPB / 2;
  static_assert(64 % LPB == 0 && LPT < 32, "the two lines of a pair share a 32-lane half of a wavefront");
  __shared__ V2<T> sm[LPB * F::LS];
  unsigned field;
  const unsigned tile = xcd_tile(field);
  const int t = threadIdx.x;
  const int l = t % LPB, s = t / LPB;
  const int mb = l / LPT;
  const unsigned q = tile * LPT + l % LPT;
  const bool valid = q < Q;  // the same for both lines of a pair
  const unsigned base = valid ? q : 0, half = (unsigned)(ny / 2) * h, own = base + mb * half;
  const int jp = (int)(base / h), kz = (int)(base % h);
  const int j = jp + mb * (ny / 2);  // this line's row (ky index)
  const V2<T> wj = twy[jp];
  const T sg = mb ? (T)-1 : (T)1;
  V2<T> *sl = sm + l * F::LS;
  // this line's row from its own half a and the other line's (base is 0
  // for an invalid pair, so every load is in bounds)
  auto row_of = [&](const V2<T> a) {
    const V2<T> o = swz_xor<LPT>(a);
    const V2<T> e = mb ? o : a, od = mb ? a : o;
    return valid ? e + scal(sg, cmul(od, wj)) : V2<T>{0, 0};
  };
  V2<T> vy[P], vz[P];
#pragma unroll
  for (int n = 0; n < P; n++) vy[n] = ld2(g3.p[1] + 2 * (own + F::in_i(s, n) * plane));
#pragma unroll
  for (int n = 0; n < P; n++) vz[n] = ld2(g3.p[2] + 2 * (own + F::in_i(s, n) * plane));
  V2<T> twr[P];
  line_twiddles<F>(twr, tw, s);
  // the divergence symbols q_d conj(e^{i k_d dx/2})
  const V2<T> fy = scal((T)(signed_k(j, ny) * wy), conjv(shy[j]));
  const V2<T> fz = scal((T)(kz * wz), conjv(shz[kz]));
  V2<T> acc[P];
#pragma unroll
  for (int n = 0; n < P; n++) acc[n] = cmul(row_of(vy[n]), fy) + cmul(row_of(vz[n]), fz);
  V2<T> vx[P];
#pragma unroll
  for (int n = 0; n < P; n++) vx[n] = ld2(g3.p[0] + 2 * (own + F::in_i(s, n) * plane));
  F::template run<-1>(acc, sl, twr, s, LPB);
#pragma unroll
  for (int n = 0; n < P; n++) vx[n] = row_of(vx[n]);
  F::template run<-1>(vx, sl, twr, s, LPB);
#pragma unroll
  for (int m = 0; m < P; m++) {
    const int i = F::out_i(s, m);
    acc[m] = acc[m] + cmul(vx[m], scal((T)(signed_k(i, L) * wx), conjv(shx[i])));
  }
  const double wk = (kz == 0 || 2 * kz == nz) ? 1.0 : 2.0;
  double sum = 0.0;
  if (valid) {
#pragma unroll
    for (int m = 0; m < P; m++) {
      const unsigned off = own + F::out_i(s, m) * plane;
      const V2<T> av{-acc[m].im, acc[m].re};
      const V2<T> pp = ld2(p + 2 * off);
      if constexpr (RES) {
        const V2<T> rv = av - pp;
        st2(ap + 2 * off, rv);
        sum += wk * ((double)rv.re * rv.re + (double)rv.im * rv.im);
      } else {
        st2(ap + 2 * off, av);
        sum += wk * ((double)pp.re * av.re + (double)pp.im * av.im);
      }
    }
  }
  sum = block_sum_fft<LPB * F::S>(sum);
  if (threadIdx.x == 0) partial[tile] = sum;
}

// ------------------------------------------------------------------- host
namespace {

template <typename T>
void *make_table(int n, int count) {
  std::vector<V2<T>> h(count);
  for (int m = 0; m < count; m++) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
    h[m] = V2<T>{(T)cosl(a), (T)sinl(a)};
  }
  void *d = nullptr;
  KAPSEL_HIP(hipMalloc(&d, sizeof(V2<T>) * count));
  KAPSEL_HIP(hipMemcpy(d, h.data(), sizeof(V2<T>) * count, hipMemcpyHostToDevice));
  return d;
}

// exp(i dx wave k / 2) for the signed wavenumbers of one axis (Shift_x/y/z
// of Init_K, fft_wrapper.cxx:83-85); count = n (x, y) or n/2+1 (z)
template <typename T>
void *make_shift(int n, int count) {
  std::vector<V2<T>> h(count);
  for (int i = 0; i < count; i++) {
    int k = (count == n && i > n / 2) ? i - n : i;
    long double a = 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
    h[i] = V2<T>{(T)cosl(a), (T)sinl(a)};
  }
  void *d = nullptr;
  KAPSEL_HIP(hipMalloc(&d, sizeof(V2<T>) * count));
  KAPSEL_HIP(hipMemcpy(d, h.data(), sizeof(V2<T>) * count, hipMemcpyHostToDevice));
  return d;
}

template <typename T>
void init_tw_t(kapsel_ctx *ctx) {
  const Grid &g = ctx->g;
  Twiddles &tw = ctx->tw[prec_of<T>()];
  tw.x = make_table<T>(g.nx, g.nx);
  tw.y = make_table<T>(g.ny, g.ny);
  tw.zh = make_table<T>(g.nz / 2, g.nz / 2);
  tw.zrot = make_table<T>(g.nz, g.nz / 2 + 1);
  tw.shx = make_shift<T>(g.nx, g.nx);
  tw.shy = make_shift<T>(g.ny, g.ny);
  tw.shz = make_shift<T>(g.nz, g.nz / 2 + 1);
}

constexpr int lines_per_block(size_t elem) { return elem == 8 ? 16 : 32; }
// the y-split x passes of the operator (8 row pairs per workgroup)
constexpr int dec_lines_per_block(size_t) { return 16; }

template <typename T, int L, int DIR>
void launch_lines(kapsel_ctx *ctx, int cls, const FieldPtrs<T> &fp, const FieldPtrs<T> &out, int nf, long long Q,
                  long long A, long long B, long long ES, const void *tw) {
  constexpr int LPB = lines_per_block(sizeof(T));
  dim3 grid((unsigned)((Q + LPB - 1) / LPB), nf);
  LaunchTimer lt(ctx, cls, 2.0 * nf * ctx->g.f * sizeof(T));
  hipLaunchKernelGGL((fft_c2c_lines_io<T, L, DIR, LPB>), grid, dim3(LPB * LineFFT<T, L>::S), 0, ctx->stream, fp,
                     out, Q, A, B, ES, (const V2<T> *)tw);
  KAPSEL_LAUNCH_CHECK();
}

// out: nullptr transforms in place
template <typename T, int DIR>
void lines_dispatch(kapsel_ctx *ctx, int cls, int L, const FieldPtrs<T> &fp, int nf, long long Q, long long A,
                    long long B, long long ES, const void *tw, const FieldPtrs<T> *out = nullptr) {
  const FieldPtrs<T> &o = out ? *out : fp;
  switch (L) {
    case 16: launch_lines<T, 16, DIR>(ctx, cls, fp, o, nf, Q, A, B, ES, tw); break;
    case 32: launch_lines<T, 32, DIR>(ctx, cls, fp, o, nf, Q, A, B, ES, tw); break;
    case 64: launch_lines<T, 64, DIR>(ctx, cls, fp, o, nf, Q, A, B, ES, tw); break;
    case 128: launch_lines<T, 128, DIR>(ctx, cls, fp, o, nf, Q, A, B, ES, tw); break;
    case 256: launch_lines<T, 256, DIR>(ctx, cls, fp, o, nf, Q, A, B, ES, tw); break;
    default: set_error("unsupported FFT length %d", L); throw Error{KAPSEL_ERR_ARG};
  }
}

constexpr int rows_per_block(size_t elem) { return elem == 8 ? 16 : 32; }

// fft_z_r2c_pro reads and writes a block's rows 16 bytes at a time
template <typename T>
void check_z_align(const FieldPtrs<T> &fp, int nf) {
  for (int i = 0; i < nf; i++)
    if (fp.p[i] && ((uintptr_t)fp.p[i] & 15u)) {
      set_error("field pointer %p is not 16-byte aligned", (void *)fp.p[i]);
      throw Error{KAPSEL_ERR_ARG};
    }
}

template <typename T, int NZ>
void launch_z(kapsel_ctx *ctx, const FieldPtrs<T> &fp, int nf, bool fwd) {
  constexpr int RPB = rows_per_block(sizeof(T));
  const Grid &g = ctx->g;
  long long rows = (long long)g.nx * g.ny;
  dim3 grid((unsigned)((rows + RPB - 1) / RPB), nf);
  dim3 block(RPB * FS<NZ / 2>::R2);
  LaunchTimer lt(ctx, fwd ? KAPSEL_K_FFT_Z_R2C : KAPSEL_K_FFT_Z_C2R, 2.0 * nf * g.f * sizeof(T));
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  if (fwd) {
    hipLaunchKernelGGL((fft_z_r2c<T, NZ, RPB>), grid, block, 0, ctx->stream, fp, rows, g.nz_,
                       (const V2<T> *)tw.zh, (const V2<T> *)tw.zrot);
  } else {
    T scale = (T)(1.0 / ((double)g.nx * g.ny * g.nz));
    hipLaunchKernelGGL((fft_z_c2r<T, NZ, RPB>), grid, block, 0, ctx->stream, fp, rows, g.nz_,
                       (const V2<T> *)tw.zh, (const V2<T> *)tw.zrot, scale);
  }
  KAPSEL_LAUNCH_CHECK();
}

template <typename T>
void z_dispatch(kapsel_ctx *ctx, const FieldPtrs<T> &fp, int nf, bool fwd) {
  switch (ctx->g.nz) {
    case 16: launch_z<T, 16>(ctx, fp, nf, fwd); break;
    case 32: launch_z<T, 32>(ctx, fp, nf, fwd); break;
    case 64: launch_z<T, 64>(ctx, fp, nf, fwd); break;
    case 128: launch_z<T, 128>(ctx, fp, nf, fwd); break;
    case 256: launch_z<T, 256>(ctx, fp, nf, fwd); break;
    default: set_error("unsupported NZ %d", ctx->g.nz); throw Error{KAPSEL_ERR_ARG};
  }
}

template <typename T, int N, int THREADS>
void launch_plane(kapsel_ctx *ctx, const FieldPtrs<T> &fp, int nf, bool fwd) {
  const Grid &g = ctx->g;
  static_assert(PlaneCfg<T, N, N, THREADS>::LDS_BYTES <= 160 * 1024, "plane does not fit LDS");
  dim3 grid((unsigned)g.nx, nf);
  LaunchTimer lt(ctx, fwd ? KAPSEL_K_FFT_Z_R2C : KAPSEL_K_FFT_Z_C2R, 2.0 * nf * g.f * sizeof(T));
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  if (fwd) {
    hipLaunchKernelGGL((fft_plane_r2c<T, N, N, THREADS>), grid, dim3(THREADS), 0, ctx->stream, fp,
                       (const V2<T> *)tw.zh, (const V2<T> *)tw.zrot, (const V2<T> *)tw.y);
  } else {
    T scale = (T)(1.0 / ((double)g.nx * g.ny * g.nz));
    hipLaunchKernelGGL((fft_plane_c2r<T, N, N, THREADS>), grid, dim3(THREADS), 0, ctx->stream, fp,
                       (const V2<T> *)tw.zh, (const V2<T> *)tw.zrot, (const V2<T> *)tw.y, scale);
  }
  KAPSEL_LAUNCH_CHECK();
}

// whether the z and y transforms run as one plane pass: NY = NZ = 32, 64,
// 128 (at 128 the padded FP64 plane, 147 KB, leaves one workgroup per CU and
// still beats the separate z and y passes)
template <typename T>
bool plane_ok(const Grid &g) {
  return plane_ok_grid(g);
}

// the z and y transforms as one plane pass where plane_ok; false: use the z
// and y line passes
template <typename T>
bool plane_dispatch(kapsel_ctx *ctx, const FieldPtrs<T> &fp, int nf, bool fwd) {
  const Grid &g = ctx->g;
  if (g.ny != g.nz) return false;
  switch (g.nz) {
    // workgroup sizes: every column of the plane in one round
    case 32: launch_plane<T, 32, 128>(ctx, fp, nf, fwd); return true;
    case 64: launch_plane<T, 64, 256>(ctx, fp, nf, fwd); return true;
    case 128:
      // FP64: the padded plane (147 KB) leaves one workgroup per CU
      launch_plane<T, 128, 512>(ctx, fp, nf, fwd);
      return true;
    default: return false;
  }
}

template <typename T, int N, int THREADS>
void launch_plane_face(kapsel_ctx *ctx, const FieldPtrs<T> &fp, const FieldPtrs<T> &face) {
  const Grid &g = ctx->g;
  dim3 grid((unsigned)g.nx, 3);
  LaunchTimer lt(ctx, KAPSEL_K_FFT_PLANE_FACE, 2.0 * 3 * g.f * sizeof(T) + 3.0 * g.f * sizeof(T));
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  T scale = (T)(1.0 / ((double)g.nx * g.ny * g.nz));
  hipLaunchKernelGGL((fft_plane_face<T, N, N, THREADS>), grid, dim3(THREADS), 0, ctx->stream, fp,
                     (const V2<T> *)tw.zh, (const V2<T> *)tw.zrot, (const V2<T> *)tw.y, scale, face);
  KAPSEL_LAUNCH_CHECK();
}

template <typename T, int NZ>
void launch_z_face(kapsel_ctx *ctx, const FieldPtrs<T> &fp, const FieldPtrs<T> &face) {
  constexpr int RPB = rows_per_block(sizeof(T));
  const Grid &g = ctx->g;
  long long rows = (long long)g.nx * g.ny;
  dim3 grid((unsigned)((rows + RPB - 1) / RPB), 3);
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 2.0 * 3 * g.f * sizeof(T) + 3.0 * g.f * sizeof(T));
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  T scale = (T)(1.0 / ((double)g.nx * g.ny * g.nz));
  hipLaunchKernelGGL((fft_z_face<T, NZ, RPB>), grid, dim3(RPB * FS<NZ / 2>::R2), 0, ctx->stream, fp, rows, g.nz_,
                     (const V2<T> *)tw.zh, (const V2<T> *)tw.zrot, scale, face);
  KAPSEL_LAUNCH_CHECK();
}

template <typename T>
void z_face_dispatch(kapsel_ctx *ctx, const FieldPtrs<T> &fp, const FieldPtrs<T> &face) {
  switch (ctx->g.nz) {
    case 16: launch_z_face<T, 16>(ctx, fp, face); break;
    case 32: launch_z_face<T, 32>(ctx, fp, face); break;
    case 64: launch_z_face<T, 64>(ctx, fp, face); break;
    case 128: launch_z_face<T, 128>(ctx, fp, face); break;
    case 256: launch_z_face<T, 256>(ctx, fp, face); break;
    default: set_error("unsupported NZ %d", ctx->g.nz); throw Error{KAPSEL_ERR_ARG};
  }
}

template <typename T>
FieldPtrs<T> pack(void *const *fields, int nf) {
  FieldPtrs<T> fp{};
  for (int i = 0; i < nf; i++) fp.p[i] = (T *)fields[i];
  return fp;
}

}  // namespace

template <typename T>
void fft_r2c_t(kapsel_ctx *ctx, void *const *fields, int nf) {
  const Grid &g = ctx->g;
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  FieldPtrs<T> fp = pack<T>(fields, nf);
  if (!plane_dispatch<T>(ctx, fp, nf, true)) {
    z_dispatch<T>(ctx, fp, nf, true);
    // y lines: q = i*h + kz -> offset i*NY*h + kz, stride h
    lines_dispatch<T, -1>(ctx, KAPSEL_K_FFT_LINES_FWD, g.ny, fp, nf, (long long)g.nx * g.h, g.h,
                          (long long)g.ny * g.h, g.h, tw.y);
  }
  // x lines: q = j*h + kz -> offset q, stride NY*h
  long long plane = (long long)g.ny * g.h;
  lines_dispatch<T, -1>(ctx, KAPSEL_K_FFT_LINES_FWD, g.nx, fp, nf, plane, plane, 0, plane, tw.x);
}

// out = a + s (b + e) c (as fft_plane_r2c_pro / fft_z_r2c_pro) transformed
// forward: the product formed in the first pass (the plane pass, or the z
// pass followed by the y lines), then the x lines.  e NULL: all e_f = 0.
// Returns false (nothing queued) for an NZ without a z pass.
template <typename T>
bool fft_r2c_pro_t(kapsel_ctx *ctx, void *const *out, void *const *a, void *const *b, void *const *c, double s,
                   int nf, unsigned dsq, const double *e) {
  const Grid &g = ctx->g;
  const bool plane = plane_ok<T>(g);
  if (!plane && !(g.nz == 16 || g.nz == 32 || g.nz == 64 || g.nz == 128 || g.nz == 256)) return false;
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  FieldPtrs<T> fo = pack<T>(out, nf), fb = pack<T>(b, nf);
  FieldPtrs<T> fa = a ? pack<T>(a, nf) : FieldPtrs<T>{};
  FieldPtrs<T> fc = c ? pack<T>(c, nf) : FieldPtrs<T>{};
  FieldConsts<T> fe{};
  for (int i = 0; i < nf; i++) fe.v[i] = e ? (T)e[i] : (T)0;
  {
    LaunchTimer lt(ctx, KAPSEL_K_FFT_Z_R2C, (double)((a ? 1 : 0) + (c ? 3 : 2)) * nf * g.f * sizeof(T));
    if (plane) {
      dim3 grid((unsigned)g.nx, nf);
      switch (g.nz) {
#define KAPSEL_PRO(N, TH)                                                                                 \
  case N:                                                                                                 \
    hipLaunchKernelGGL((fft_plane_r2c_pro<T, N, N, TH>), grid, dim3(TH), 0, ctx->stream, fo, fa, fb, fc, fe, \
                       (T)s, dsq, (const V2<T> *)tw.zh, (const V2<T> *)tw.zrot, (const V2<T> *)tw.y);     \
    break;
        KAPSEL_PRO(32, 128)
        KAPSEL_PRO(64, 256)
        KAPSEL_PRO(128, 512)
#undef KAPSEL_PRO
        default: return false;
      }
    } else {
      check_z_align(fo, nf);
      check_z_align(fa, nf);
      check_z_align(fb, nf);
      check_z_align(fc, nf);
      constexpr int RPB = rows_per_block(sizeof(T));
      const long long rows = (long long)g.nx * g.ny;
      dim3 grid((unsigned)((rows + RPB - 1) / RPB), nf);
      switch (g.nz) {
#define KAPSEL_PRO(N)                                                                                           \
  case N:                                                                                                       \
    hipLaunchKernelGGL((fft_z_r2c_pro<T, N, RPB>), grid, dim3(RPB * FS<N / 2>::R2), 0, ctx->stream, fo, fa, fb, \
                       fc, fe, (T)s, dsq, rows, g.nz_, (const V2<T> *)tw.zh, (const V2<T> *)tw.zrot);          \
    break;
        KAPSEL_PRO(16)
        KAPSEL_PRO(32)
        KAPSEL_PRO(64)
        KAPSEL_PRO(128)
        KAPSEL_PRO(256)
#undef KAPSEL_PRO
      }
    }
    KAPSEL_LAUNCH_CHECK();
  }
  if (!plane)
    lines_dispatch<T, -1>(ctx, KAPSEL_K_FFT_LINES_FWD, g.ny, fo, nf, (long long)g.nx * g.h, g.h,
                          (long long)g.ny * g.h, g.h, tw.y);
  const long long pl = (long long)g.ny * g.h;
  lines_dispatch<T, -1>(ctx, KAPSEL_K_FFT_LINES_FWD, g.nx, fo, nf, pl, pl, 0, pl, tw.x);
  return true;
}
template bool fft_r2c_pro_t<float>(kapsel_ctx *, void *const *, void *const *, void *const *, void *const *, double,
                                   int, unsigned, const double *);
template bool fft_r2c_pro_t<double>(kapsel_ctx *, void *const *, void *const *, void *const *, void *const *,
                                    double, int, unsigned, const double *);


template <typename T>
void fft_c2r_t(kapsel_ctx *ctx, void *const *fields, int nf, void *const *src) {
  const Grid &g = ctx->g;
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  FieldPtrs<T> fp = pack<T>(fields, nf);
  long long plane = (long long)g.ny * g.h;
  if (src) {
    FieldPtrs<T> sp = pack<T>(src, nf);
    lines_dispatch<T, +1>(ctx, KAPSEL_K_FFT_LINES_INV, g.nx, sp, nf, plane, plane, 0, plane, tw.x, &fp);
  } else {
    lines_dispatch<T, +1>(ctx, KAPSEL_K_FFT_LINES_INV, g.nx, fp, nf, plane, plane, 0, plane, tw.x);
  }
  if (!plane_dispatch<T>(ctx, fp, nf, false)) {
    lines_dispatch<T, +1>(ctx, KAPSEL_K_FFT_LINES_INV, g.ny, fp, nf, (long long)g.nx * g.h, g.h,
                          (long long)g.ny * g.h, g.h, tw.y);
    z_dispatch<T>(ctx, fp, nf, false);
  }
}

// fft_c2r_t after its first (x, inverse) pass: the y and z inverse passes
template <typename T>
void fft_c2r_yz_t(kapsel_ctx *ctx, void *const *fields, int nf) {
  const Grid &g = ctx->g;
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  FieldPtrs<T> fp = pack<T>(fields, nf);
  if (!plane_dispatch<T>(ctx, fp, nf, false)) {
    lines_dispatch<T, +1>(ctx, KAPSEL_K_FFT_LINES_INV, g.ny, fp, nf, (long long)g.nx * g.h, g.h,
                          (long long)g.ny * g.h, g.h, tw.y);
    z_dispatch<T>(ctx, fp, nf, false);
  }
}
template void fft_c2r_yz_t<float>(kapsel_ctx *, void *const *, int);
template void fft_c2r_yz_t<double>(kapsel_ctx *, void *const *, int);

template <typename T>
void fft_face_t(kapsel_ctx *ctx, void *const *g3, void *const *face3) {
  const Grid &g = ctx->g;
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  FieldPtrs<T> fp = pack<T>(g3, 3), fc = pack<T>(face3, 3);
  const long long plane = (long long)g.ny * g.h;
  lines_dispatch<T, +1>(ctx, KAPSEL_K_FFT_LINES_INV, g.nx, fp, 3, plane, plane, 0, plane, tw.x);
  bool done = false;
  if (plane_ok<T>(g)) {
    switch (g.nz) {
      case 32: launch_plane_face<T, 32, 128>(ctx, fp, fc); done = true; break;
      case 64: launch_plane_face<T, 64, 256>(ctx, fp, fc); done = true; break;
      case 128: launch_plane_face<T, 128, 512>(ctx, fp, fc); done = true; break;
      default: break;
    }
  }
  if (!done) {
    lines_dispatch<T, +1>(ctx, KAPSEL_K_FFT_LINES_INV, g.ny, fp, 3, (long long)g.nx * g.h, g.h,
                          (long long)g.ny * g.h, g.h, tw.y);
    z_face_dispatch<T>(ctx, fp, fc);
    lines_dispatch<T, -1>(ctx, KAPSEL_K_FFT_LINES_FWD, g.ny, fp, 3, (long long)g.nx * g.h, g.h,
                          (long long)g.ny * g.h, g.h, tw.y);
  }
  lines_dispatch<T, -1>(ctx, KAPSEL_K_FFT_LINES_FWD, g.nx, fp, 3, plane, plane, 0, plane, tw.x);
}

template <typename T, int N, int THREADS>
void launch_half_face(kapsel_ctx *ctx, const FieldPtrs<T> &fp, const FieldPtrs<T> &face) {
  const Grid &g = ctx->g;
  static_assert(PlaneCfg<T, N / 2, N, THREADS>::LDS_BYTES <= 160 * 1024, "half plane does not fit LDS");
  dim3 grid(2u * (unsigned)g.nx, 3);
  // (a sampled launch carries its timing events itself: their interval is
  // the kernel's run, without the dispatch gap before it)
  LaunchTimer lt(ctx, KAPSEL_K_FFT_PLANE_FACE, 2.0 * 3 * g.f * sizeof(T) + 3.0 * g.f * sizeof(T), true);
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  T scale = (T)(1.0 / ((double)g.nx * g.ny * g.nz));
  hipExtLaunchKernelGGL((fft_half_face<T, N, N, THREADS>), grid, dim3(THREADS), 0, ctx->stream, lt.start, lt.stop, 0,
                        fp, (const V2<T> *)tw.zh, (const V2<T> *)tw.zrot, (const V2<T> *)tw.y, scale, face);
  KAPSEL_LAUNCH_CHECK();
}

// whether the operator's x passes and plane round trip use the y-split
// gradient fields (fft_x_grad_dec, fft_half_face, fft_x_div_dec): wherever the
// fused x passes meet a plane pass
template <typename T>
bool dec_ok(const Grid &g) {
  return plane_ok<T>(g) && (g.nx == 64 || g.nx == 128);
}

// the middle of fft_face_t: everything but the two x passes (on the y-split
// fields where dec_ok)
template <typename T>
void fft_face_mid_t(kapsel_ctx *ctx, void *const *g3, void *const *face3) {
  const Grid &g = ctx->g;
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  FieldPtrs<T> fp = pack<T>(g3, 3), fc = pack<T>(face3, 3);
  if (dec_ok<T>(g)) {
    switch (g.nz) {
      case 32: launch_half_face<T, 32, 64>(ctx, fp, fc); return;
      case 64: launch_half_face<T, 64, 128>(ctx, fp, fc); return;
      case 128: launch_half_face<T, 128, 512>(ctx, fp, fc); return;
      default: break;
    }
  }
  if (plane_ok<T>(g)) {
    switch (g.nz) {
      case 32: launch_plane_face<T, 32, 128>(ctx, fp, fc); return;
      case 64: launch_plane_face<T, 64, 256>(ctx, fp, fc); return;
      case 128: launch_plane_face<T, 128, 512>(ctx, fp, fc); return;
      default: break;
    }
  }
  lines_dispatch<T, +1>(ctx, KAPSEL_K_FFT_LINES_INV, g.ny, fp, 3, (long long)g.nx * g.h, g.h, (long long)g.ny * g.h,
                        g.h, tw.y);
  z_face_dispatch<T>(ctx, fp, fc);
  lines_dispatch<T, -1>(ctx, KAPSEL_K_FFT_LINES_FWD, g.ny, fp, 3, (long long)g.nx * g.h, g.h, (long long)g.ny * g.h,
                        g.h, tw.y);
}

template <typename T, int L>
void launch_x_grad(kapsel_ctx *ctx, const T *z, T *p, const double *sc, int beta_slot, void *const *g3,
                   const float *sub, double zscale) {
  const Grid &g = ctx->g;
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  LaunchTimer lt(ctx, KAPSEL_K_FFT_X_GRAD, 6.0 * g.f * sizeof(T));
  if (dec_ok<T>(g)) {
    constexpr int LPB = dec_lines_per_block(sizeof(T));
    constexpr int LPT = LPB / 2;  // row pairs per tile
    const unsigned Q = (unsigned)(g.ny / 2) * g.h;
    hipLaunchKernelGGL((fft_x_grad_dec<T, L, LPB>), dim3((Q + LPT - 1) / LPT), dim3(LPB * LineFFT<T, L>::S), 0,
                       ctx->stream, z, p, sc, beta_slot, (T)zscale, pack<T>(g3, 3), Q, (unsigned)(g.ny * g.h),
                       g.h, g.ny,
                       g.wave_x, g.wave_y, g.wave_z, (const V2<T> *)tw.shx, (const V2<T> *)tw.shy,
                       (const V2<T> *)tw.shz, (const V2<T> *)tw.x, sub, (const V2<T> *)tw.y);
  } else {
    constexpr int LPB = lines_per_block(sizeof(T));
    const long long Q = (long long)g.ny * g.h;
    hipLaunchKernelGGL((fft_x_grad<T, L, LPB>), dim3((unsigned)((Q + LPB - 1) / LPB)), dim3(LPB * LineFFT<T, L>::S),
                       0, ctx->stream, z, p, sc, beta_slot, (T)zscale, pack<T>(g3, 3), Q, Q, g.h, g.ny, g.wave_x,
                       g.wave_y, g.wave_z, (const V2<T> *)tw.shx, (const V2<T> *)tw.shy, (const V2<T> *)tw.shz,
                       (const V2<T> *)tw.x, sub);
  }
  KAPSEL_LAUNCH_CHECK();
}

// a divergence pass writes one partial sum per workgroup into a region of
// max_partials slots (the buffer's next region holds other sums): refuse a
// launch of more workgroups than that
void check_partial_slots(int nb, int max_partials) {
  if (nb > max_partials) {
    set_error("the divergence pass would write %d partial sums into %d slots", nb, max_partials);
    throw Error{KAPSEL_ERR_ARG};
  }
}

template <typename T, int L, bool RES>
int launch_x_div(kapsel_ctx *ctx, void *const *g3, const T *p, T *ap, double *partial, int max_partials) {
  constexpr int LPB = lines_per_block(sizeof(T));
  const Grid &g = ctx->g;
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  LaunchTimer lt(ctx, KAPSEL_K_FFT_X_DIV, 5.0 * g.f * sizeof(T));
  unsigned nb;
  if (dec_ok<T>(g)) {
    constexpr int LPBD = dec_lines_per_block(sizeof(T)), LPT = LPBD / 2;
    const unsigned Q = (unsigned)(g.ny / 2) * g.h;
    nb = (Q + LPT - 1) / LPT;
    check_partial_slots((int)nb, max_partials);
    hipLaunchKernelGGL((fft_x_div_dec<T, L, LPBD, RES>), dim3(nb), dim3(LPBD * DivFFT<T, L>::S), 0, ctx->stream,
                       pack<T>(g3, 3), p, ap, partial, Q, (unsigned)(g.ny * g.h), g.h, g.ny, g.nz, g.wave_x,
                       g.wave_y, g.wave_z, (const V2<T> *)tw.shx, (const V2<T> *)tw.shy, (const V2<T> *)tw.shz,
                       (const V2<T> *)tw.x, (const V2<T> *)tw.y);
  } else {
    const long long Q = (long long)g.ny * g.h;
    nb = (unsigned)((Q + LPB - 1) / LPB);
    check_partial_slots((int)nb, max_partials);
    hipLaunchKernelGGL((fft_x_div<T, L, LPB, RES>), dim3(nb), dim3(LPB * DivFFT<T, L>::S), 0, ctx->stream,
                       pack<T>(g3, 3), p, ap, partial, Q, Q, g.h, g.ny, g.nz, g.wave_x, g.wave_y, g.wave_z,
                       (const V2<T> *)tw.shx, (const V2<T> *)tw.shy, (const V2<T> *)tw.shz, (const V2<T> *)tw.x);
  }
  KAPSEL_LAUNCH_CHECK();
  return (int)nb;
}

template <typename T>
bool fft_x_grad_t(kapsel_ctx *ctx, const T *z, T *p, const double *sc, int beta_slot, void *const *g3,
                  const float *sub, double zscale) {
  switch (ctx->g.nx) {
    case 64: launch_x_grad<T, 64>(ctx, z, p, sc, beta_slot, g3, sub, zscale); return true;
    case 128: launch_x_grad<T, 128>(ctx, z, p, sc, beta_slot, g3, sub, zscale); return true;
    default: return false;
  }
}

template <int L>
void launch_x_grad2_f64(kapsel_ctx *ctx, const double *a, const double *b, void *const *out6) {
  constexpr int LPB = lines_per_block(sizeof(float));
  const Grid &g = ctx->g;
  const Twiddles &t64 = ctx->tw[prec_of<double>()];
  const long long Q = (long long)g.ny * g.h;
  LaunchTimer lt(ctx, KAPSEL_K_FFT_X_GRAD, 2.0 * g.f * sizeof(double) + 6.0 * g.f * sizeof(float));
  hipLaunchKernelGGL((fft_x_grad2_f64<L, LPB>), dim3((unsigned)((Q + LPB - 1) / LPB), 2),
                     dim3(LPB * LineFFT<float, L>::S), 0, ctx->stream, a, b, pack<float>(out6, 6), Q, Q, g.h, g.ny,
                     g.wave_x, g.wave_y, g.wave_z, (const V2<double> *)t64.shx, (const V2<double> *)t64.shy,
                     (const V2<double> *)t64.shz, (const V2<float> *)ctx->tw[prec_of<float>()].x);
  KAPSEL_LAUNCH_CHECK();
}

template <typename T>
int fft_x_div_t(kapsel_ctx *ctx, void *const *g3, const T *p, T *ap, double *partial, int max_partials) {
  switch (ctx->g.nx) {
    case 64: return launch_x_div<T, 64, false>(ctx, g3, p, ap, partial, max_partials);
    case 128: return launch_x_div<T, 128, false>(ctx, g3, p, ap, partial, max_partials);
    default: return 0;
  }
}

template <typename T>
int fft_x_res_t(kapsel_ctx *ctx, void *const *g3, const T *b, T *r, double *partial, int max_partials) {
  switch (ctx->g.nx) {
    case 64: return launch_x_div<T, 64, true>(ctx, g3, b, r, partial, max_partials);
    case 128: return launch_x_div<T, 128, true>(ctx, g3, b, r, partial, max_partials);
    default: return 0;
  }
}

// the staggered gradients of the FP64 spectra a and b as six real FP32
// fields (out6: a's three, then b's): the x pass fused with the gradient
// symbols where NX has one (64, 128), then the y and z passes of all six in
// one launch; false: nothing queued
bool fft_grad2_f64_to_f32(kapsel_ctx *ctx, const double *a, const double *b, void *const *out6) {
  const Grid &g = ctx->g;
  switch (g.nx) {
    case 64: launch_x_grad2_f64<64>(ctx, a, b, out6); break;
    case 128: launch_x_grad2_f64<128>(ctx, a, b, out6); break;
    default: return false;
  }
  FieldPtrs<float> fp = pack<float>(out6, 6);
  if (!plane_dispatch<float>(ctx, fp, 6, false)) {
    lines_dispatch<float, +1>(ctx, KAPSEL_K_FFT_LINES_INV, g.ny, fp, 6, (long long)g.nx * g.h, g.h,
                              (long long)g.ny * g.h, g.h, ctx->tw[prec_of<float>()].y);
    z_dispatch<float>(ctx, fp, 6, false);
  }
  return true;
}

template void fft_face_mid_t<double>(kapsel_ctx *, void *const *, void *const *);
template int fft_x_res_t<double>(kapsel_ctx *, void *const *, const double *, double *, double *, int);
template void fft_face_mid_t<float>(kapsel_ctx *, void *const *, void *const *);
template bool fft_x_grad_t<float>(kapsel_ctx *, const float *, float *, const double *, int, void *const *,
                                  const float *, double);
template bool fft_x_grad_t<double>(kapsel_ctx *, const double *, double *, const double *, int, void *const *,
                                   const float *, double);
template int fft_x_div_t<float>(kapsel_ctx *, void *const *, const float *, float *, double *, int);
template int fft_x_div_t<double>(kapsel_ctx *, void *const *, const double *, double *, double *, int);
template void fft_face_t<double>(kapsel_ctx *, void *const *, void *const *);
template void fft_face_t<float>(kapsel_ctx *, void *const *, void *const *);
template void fft_r2c_t<double>(kapsel_ctx *, void *const *, int);
template void fft_r2c_t<float>(kapsel_ctx *, void *const *, int);
template void fft_c2r_t<double>(kapsel_ctx *, void *const *, int, void *const *);
template void fft_c2r_t<float>(kapsel_ctx *, void *const *, int, void *const *);

// tables of both precisions: an FP32 context refines its Poisson solve in FP64
void init_twiddles(kapsel_ctx *ctx) {
  init_tw_t<double>(ctx);
  init_tw_t<float>(ctx);
}

void free_twiddles(kapsel_ctx *ctx) {
  for (Twiddles &tw : ctx->tw) {
    for (void *p : {tw.x, tw.y, tw.zh, tw.zrot, tw.shx, tw.shy, tw.shz})
      if (p) (void)hipFree(p);
    tw = Twiddles{};
  }
}

void fft_r2c(kapsel_ctx *ctx, void *const *fields, int nf) {
  if (ctx->prec == KAPSEL_FP64) fft_r2c_t<double>(ctx, fields, nf);
  else fft_r2c_t<float>(ctx, fields, nf);
}

void fft_c2r(kapsel_ctx *ctx, void *const *fields, int nf, void *const *src) {
  if (ctx->prec == KAPSEL_FP64) fft_c2r_t<double>(ctx, fields, nf, src);
  else fft_c2r_t<float>(ctx, fields, nf, src);
}

}  // namespace kapsel
