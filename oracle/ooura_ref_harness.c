/*
 * ooura_ref_harness.c -- calls the reference's own Ooura 3-D real FFT
 * (fftsg3d.c rdft3d + rdft3dsort, compiled from the reference sources where
 * they lie) exactly the way the reference's A2a_k / A_k2a do
 * (fft_wrapper.h:681-723, Ooura branch; work-array sizes of Init_fft_ooura,
 * fft_wrapper.cxx:89-99).
 *
 * TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ by oracle/Makefile and
 * used by the tests to pin the oracle's FFT and as golden-vector source.
 */
#include <math.h>
#include <stdlib.h>

void rdft3d(int n1, int n2, int n3, int isgn, double ***a, double *t, int *ip, double *w);
void rdft3dsort(int n1, int n2, int n3, int isgn, double ***a);

typedef struct {
  double ***view;
  double *t, *w;
  int *ip;
} plan_t;

static void plan_make(plan_t *p, int nx, int ny, int nz, double *a) {
  int nz_ = 2 * (nz / 2 + 1);
  int nt = nx > ny ? nx : ny;
  int n = nt > nz / 2 ? nt : nz / 2;
  p->ip = (int *)calloc(2 + (int)sqrt((double)n + 0.5), sizeof(int));
  p->t = (double *)calloc(8 * nt, sizeof(double));
  p->w = (double *)calloc(n / 2 + nz / 4, sizeof(double));
  p->ip[0] = 0;
  p->view = (double ***)malloc(sizeof(double **) * nx);
  for (int i = 0; i < nx; i++) {
    p->view[i] = (double **)malloc(sizeof(double *) * ny);
    for (int j = 0; j < ny; j++) p->view[i][j] = a + ((size_t)i * ny + j) * nz_;
  }
}

static void plan_free(plan_t *p, int nx) {
  for (int i = 0; i < nx; i++) free(p->view[i]);
  free(p->view);
  free(p->t);
  free(p->w);
  free(p->ip);
}

/* A2a_k, Ooura branch: rdft3d(+1), rdft3dsort(+1), then complex conjugate */
void ref_a2a_k(int nx, int ny, int nz, double *a) {
  plan_t p;
  plan_make(&p, nx, ny, nz, a);
  rdft3d(nx, ny, nz, 1, p.view, p.t, p.ip, p.w);
  rdft3dsort(nx, ny, nz, 1, p.view);
  size_t nc = (size_t)nx * ny * (nz / 2 + 1);
  for (size_t i = 0; i < nc; i++) a[2 * i + 1] = -a[2 * i + 1];
  plan_free(&p, nx);
}

/* A_k2a, Ooura branch: scale 2/(NX NY NZ) and conjugate, rdft3dsort(-1), rdft3d(-1) */
void ref_a_k2a(int nx, int ny, int nz, double *a) {
  plan_t p;
  plan_make(&p, nx, ny, nz, a);
  const double scale = 2.0 / ((double)nx * ny * nz);
  size_t nc = (size_t)nx * ny * (nz / 2 + 1);
  for (size_t i = 0; i < nc; i++) {
    a[2 * i] = scale * a[2 * i];
    a[2 * i + 1] = -scale * a[2 * i + 1];
  }
  rdft3dsort(nx, ny, nz, -1, p.view);
  rdft3d(nx, ny, nz, -1, p.view, p.t, p.ip, p.w);
  plan_free(&p, nx);
}
