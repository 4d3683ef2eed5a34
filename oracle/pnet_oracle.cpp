/*
 * pnet_oracle.cpp -- FP64 CPU forward/backward of the Poisson warm-start
 * network (the learned part of the dielectric solve; it has no counterpart
 * in the reference).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the GPU training kernels.
 *
 * Network (periodic 3x3x3 convolutions, channels-last activations
 * a[v*C + c], v = (i*N + j)*N + k on an N^3 cube, N divisible by 4):
 *   e1a conv 2->16 act, e1b 16->16 act (s1)        N^3
 *   pool avg 2x2x2
 *   e2a 16->32 act, e2b 32->32 act (s2)            (N/2)^3
 *   pool
 *   e3a 32->32 act, e3b 32->32 act                 (N/4)^3
 *   up nearest x2, concat [up | s2] (64)
 *   d2 64->32 act                                  (N/2)^3
 *   up, concat [up | s1] (48)
 *   d1 48->16 act                                  N^3
 *   out 16->3 linear                               N^3
 * act = leaky ReLU, x > 0 ? x : LEAK x with LEAK = 0.1 (a plain ReLU let
 * whole layers die in training: in the fluid, 97 % of the box, the input is
 * exactly zero and a layer's output there is its activated bias).
 * Weights per layer: W[tap][cin][cout] then b[cout], tap = ((ox+1)*3 + (oy+1))*3
 * + (oz+1); out[v][co] = b[co] + sum_tap sum_ci in[v + o][ci] W[tap][ci][co].
 * Loss: mean_v (sum_d e_d y[v][d] - t[v])^2 for a unit field direction e.
 *
 * Pinning: finite differences of the loss (tests/test_oracle_pnet.py);
 * pooling/upsampling adjointness; translation equivariance.
 */
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double LEAK = 0.1;

struct Layer {
  int cin, cout, level;  // level 0: N, 1: N/2, 2: N/4
  int relu;
  size_t woff, boff;
};

const int NL = 9;
Layer L[NL];
size_t NPARAM = 0;

void setup() {
  if (NPARAM) return;
  const int spec[NL][4] = {{2, 16, 0, 1},  {16, 16, 0, 1}, {16, 32, 1, 1}, {32, 32, 1, 1}, {32, 32, 2, 1},
                           {32, 32, 2, 1}, {64, 32, 1, 1}, {48, 16, 0, 1}, {16, 3, 0, 0}};
  size_t off = 0;
  for (int l = 0; l < NL; l++) {
    L[l].cin = spec[l][0];
    L[l].cout = spec[l][1];
    L[l].level = spec[l][2];
    L[l].relu = spec[l][3];
    L[l].woff = off;
    off += (size_t)27 * L[l].cin * L[l].cout;
    L[l].boff = off;
    off += L[l].cout;
  }
  NPARAM = off;
}

inline size_t vid(int n, int i, int j, int k) {
  i = (i + n) % n;
  j = (j + n) % n;
  k = (k + n) % n;
  return ((size_t)i * n + j) * n + k;
}

/* out (n^3 x cout) = conv(in (n^3 x cin), W, b), optional leaky ReLU */
void conv_fwd(int n, const double *in, int cin, const double *W, const double *b, int cout, double *out, int relu) {
#pragma omp parallel for
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) {
        double acc[64];
        for (int co = 0; co < cout; co++) acc[co] = b[co];
        for (int ox = -1; ox <= 1; ox++)
          for (int oy = -1; oy <= 1; oy++)
            for (int oz = -1; oz <= 1; oz++) {
              const int tap = ((ox + 1) * 3 + (oy + 1)) * 3 + (oz + 1);
              const double *x = in + vid(n, i + ox, j + oy, k + oz) * cin;
              const double *w = W + (size_t)tap * cin * cout;
              for (int ci = 0; ci < cin; ci++)
                for (int co = 0; co < cout; co++) acc[co] += x[ci] * w[ci * cout + co];
            }
        double *o = out + vid(n, i, j, k) * cout;
        for (int co = 0; co < cout; co++) o[co] = relu ? (acc[co] > 0 ? acc[co] : LEAK * acc[co]) : acc[co];
      }
}

/* given dy (pre-activation gradient), accumulate dW, db, and dx (if non-null) */
void conv_bwd(int n, const double *in, int cin, const double *W, int cout, const double *dy, double *dW, double *db,
              double *dx) {
  const size_t nv = (size_t)n * n * n;
  for (size_t v = 0; v < nv; v++)
    for (int co = 0; co < cout; co++) db[co] += dy[v * cout + co];
#pragma omp parallel for
  for (int tap = 0; tap < 27; tap++) {
    const int ox = tap / 9 - 1, oy = (tap / 3) % 3 - 1, oz = tap % 3 - 1;
    double *dw = dW + (size_t)tap * cin * cout;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) {
          const double *x = in + vid(n, i + ox, j + oy, k + oz) * cin;
          const double *g = dy + vid(n, i, j, k) * cout;
          for (int ci = 0; ci < cin; ci++)
            for (int co = 0; co < cout; co++) dw[ci * cout + co] += x[ci] * g[co];
        }
  }
  if (!dx) return;
  /* dx[v + o][ci] += dy[v][co] W[tap][ci][co]  <=>  dx[u][ci] = sum_tap dy[u - o] W */
#pragma omp parallel for
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) {
        double acc[64] = {0};
        for (int ox = -1; ox <= 1; ox++)
          for (int oy = -1; oy <= 1; oy++)
            for (int oz = -1; oz <= 1; oz++) {
              const int tap = ((ox + 1) * 3 + (oy + 1)) * 3 + (oz + 1);
              const double *g = dy + vid(n, i - ox, j - oy, k - oz) * cout;
              const double *w = W + (size_t)tap * cin * cout;
              for (int ci = 0; ci < cin; ci++) {
                double s = 0;
                for (int co = 0; co < cout; co++) s += g[co] * w[ci * cout + co];
                acc[ci] += s;
              }
            }
        double *d = dx + vid(n, i, j, k) * cin;
        for (int ci = 0; ci < cin; ci++) d[ci] += acc[ci];
      }
}

/* average pool 2x2x2: in (n^3 x c) -> out ((n/2)^3 x c) */
void pool_fwd(int n, const double *in, int c, double *out, int cstride_out, int coff_out) {
  const int m = n / 2;
  for (int i = 0; i < m; i++)
    for (int j = 0; j < m; j++)
      for (int k = 0; k < m; k++)
        for (int ch = 0; ch < c; ch++) {
          double s = 0;
          for (int a = 0; a < 8; a++) s += in[vid(n, 2 * i + (a >> 2), 2 * j + ((a >> 1) & 1), 2 * k + (a & 1)) * c + ch];
          out[vid(m, i, j, k) * cstride_out + coff_out + ch] = 0.125 * s;
        }
}
void pool_bwd(int n, const double *dout, int c, double *din) {
  const int m = n / 2;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++)
        for (int ch = 0; ch < c; ch++) din[vid(n, i, j, k) * c + ch] += 0.125 * dout[vid(m, i / 2, j / 2, k / 2) * c + ch];
}
/* nearest upsample x2 into channels [coff, coff + c) of a (2m)^3 x cstride tensor */
void up_fwd(int m, const double *in, int c, double *out, int cstride, int coff) {
  const int n = 2 * m;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++)
        for (int ch = 0; ch < c; ch++) out[vid(n, i, j, k) * cstride + coff + ch] = in[vid(m, i / 2, j / 2, k / 2) * c + ch];
}
void up_bwd(int m, const double *dout, int cstride, int coff, int c, double *din) {
  const int n = 2 * m;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++)
        for (int ch = 0; ch < c; ch++) din[vid(m, i / 2, j / 2, k / 2) * c + ch] += dout[vid(n, i, j, k) * cstride + coff + ch];
}

struct Net {
  int n;
  std::vector<double> x0, a1, s1, p1, a2, s2, p2, a3, a4, c2, d2, c1, d1, y;
  void run(const double *P, const double *input) {
    setup();
    const int n0 = n, n1 = n / 2, n2 = n / 4;
    const size_t v0 = (size_t)n0 * n0 * n0, v1 = (size_t)n1 * n1 * n1, v2 = (size_t)n2 * n2 * n2;
    x0.assign(input, input + v0 * 2);
    a1.assign(v0 * 16, 0); s1.assign(v0 * 16, 0); p1.assign(v1 * 16, 0);
    a2.assign(v1 * 32, 0); s2.assign(v1 * 32, 0); p2.assign(v2 * 32, 0);
    a3.assign(v2 * 32, 0); a4.assign(v2 * 32, 0);
    c2.assign(v1 * 64, 0); d2.assign(v1 * 32, 0);
    c1.assign(v0 * 48, 0); d1.assign(v0 * 16, 0); y.assign(v0 * 3, 0);
    auto cv = [&](int l, int nn, const double *in, double *out) {
      conv_fwd(nn, in, L[l].cin, P + L[l].woff, P + L[l].boff, L[l].cout, out, L[l].relu);
    };
    cv(0, n0, x0.data(), a1.data());
    cv(1, n0, a1.data(), s1.data());
    pool_fwd(n0, s1.data(), 16, p1.data(), 16, 0);
    cv(2, n1, p1.data(), a2.data());
    cv(3, n1, a2.data(), s2.data());
    pool_fwd(n1, s2.data(), 32, p2.data(), 32, 0);
    cv(4, n2, p2.data(), a3.data());
    cv(5, n2, a3.data(), a4.data());
    up_fwd(n2, a4.data(), 32, c2.data(), 64, 0);
    for (size_t v = 0; v < v1; v++)
      for (int ch = 0; ch < 32; ch++) c2[v * 64 + 32 + ch] = s2[v * 32 + ch];
    cv(6, n1, c2.data(), d2.data());
    up_fwd(n1, d2.data(), 32, c1.data(), 48, 0);
    for (size_t v = 0; v < v0; v++)
      for (int ch = 0; ch < 16; ch++) c1[v * 48 + 32 + ch] = s1[v * 16 + ch];
    cv(7, n0, c1.data(), d1.data());
    cv(8, n0, d1.data(), y.data());
  }
};

/* gradient through the activation, from the sign of its output */
void relu_mask(std::vector<double> &g, const std::vector<double> &act) {
  for (size_t i = 0; i < g.size(); i++)
    if (!(act[i] > 0)) g[i] *= LEAK;
}

}  // namespace

extern "C" {

long pnet_param_count(void) {
  setup();
  return (long)NPARAM;
}

/* layer table: 9 x {cin, cout, level, relu, woff, boff} */
void pnet_layers(long *out) {
  setup();
  for (int l = 0; l < NL; l++) {
    out[6 * l + 0] = L[l].cin;
    out[6 * l + 1] = L[l].cout;
    out[6 * l + 2] = L[l].level;
    out[6 * l + 3] = L[l].relu;
    out[6 * l + 4] = (long)L[l].woff;
    out[6 * l + 5] = (long)L[l].boff;
  }
}

/* y (n^3 x 3) = net(input (n^3 x 2)) */
void pnet_forward(int n, const double *P, const double *input, double *y) {
  Net net;
  net.n = n;
  net.run(P, input);
  memcpy(y, net.y.data(), net.y.size() * sizeof(double));
}

/* loss and dL/dP for target t (n^3) and unit direction e[3] */
double pnet_loss_grad(int n, const double *P, const double *input, const double *e, const double *t, double *grad) {
  setup();
  Net net;
  net.n = n;
  net.run(P, input);
  const int n0 = n, n1 = n / 2, n2 = n / 4;
  const size_t v0 = (size_t)n0 * n0 * n0, v1 = (size_t)n1 * n1 * n1;
  memset(grad, 0, NPARAM * sizeof(double));
  double loss = 0.0;
  std::vector<double> gy(v0 * 3);
  for (size_t v = 0; v < v0; v++) {
    double pred = e[0] * net.y[v * 3] + e[1] * net.y[v * 3 + 1] + e[2] * net.y[v * 3 + 2];
    double r = pred - t[v];
    loss += r * r;
    for (int d = 0; d < 3; d++) gy[v * 3 + d] = 2.0 * r * e[d] / (double)v0;
  }
  loss /= (double)v0;
  auto bw = [&](int l, int nn, const std::vector<double> &in, std::vector<double> &dy, std::vector<double> *dx) {
    conv_bwd(nn, in.data(), L[l].cin, P + L[l].woff, L[l].cout, dy.data(), grad + L[l].woff, grad + L[l].boff,
             dx ? dx->data() : nullptr);
  };
  std::vector<double> gd1(v0 * 16, 0.0), gc1(v0 * 48, 0.0);
  bw(8, n0, net.d1, gy, &gd1);
  relu_mask(gd1, net.d1);
  bw(7, n0, net.c1, gd1, &gc1);
  std::vector<double> gd2(v1 * 32, 0.0), gs1(v0 * 16, 0.0);
  up_bwd(n1, gc1.data(), 48, 0, 32, gd2.data());
  for (size_t v = 0; v < v0; v++)
    for (int ch = 0; ch < 16; ch++) gs1[v * 16 + ch] += gc1[v * 48 + 32 + ch];
  relu_mask(gd2, net.d2);
  std::vector<double> gc2(v1 * 64, 0.0);
  bw(6, n1, net.c2, gd2, &gc2);
  std::vector<double> ga4((size_t)n2 * n2 * n2 * 32, 0.0), gs2(v1 * 32, 0.0);
  up_bwd(n2, gc2.data(), 64, 0, 32, ga4.data());
  for (size_t v = 0; v < v1; v++)
    for (int ch = 0; ch < 32; ch++) gs2[v * 32 + ch] += gc2[v * 64 + 32 + ch];
  relu_mask(ga4, net.a4);
  std::vector<double> ga3(ga4.size(), 0.0);
  bw(5, n2, net.a3, ga4, &ga3);
  relu_mask(ga3, net.a3);
  std::vector<double> gp2(ga4.size(), 0.0);
  bw(4, n2, net.p2, ga3, &gp2);
  pool_bwd(n1, gp2.data(), 32, gs2.data());
  relu_mask(gs2, net.s2);
  std::vector<double> ga2(v1 * 32, 0.0);
  bw(3, n1, net.a2, gs2, &ga2);
  relu_mask(ga2, net.a2);
  std::vector<double> gp1(v1 * 16, 0.0);
  bw(2, n1, net.p1, ga2, &gp1);
  pool_bwd(n0, gp1.data(), 16, gs1.data());
  relu_mask(gs1, net.s1);
  std::vector<double> ga1(v0 * 16, 0.0);
  bw(1, n0, net.a1, gs1, &ga1);
  relu_mask(ga1, net.a1);
  bw(0, n0, net.x0, ga1, nullptr);
  return loss;
}

}  // extern "C"
