/*
 * CPU oracle of PoissonDirect (solver.js:105-165) -- TEST INFRASTRUCTURE ONLY.
 *
 * Solves rhoC(z) lap_h(phi) + d/dz(rhoW(z) dphi/dz) = rhs with phi = 0 just
 * outside the lateral faces (DST-II in x and y) and dphi/dz = 0 at ground and
 * lid, exactly as the reference: forward x then y sine transforms, one Thomas
 * solve per (l, m) with shift lambda_x[l] + lambda_y[m], inverse y then x.
 *
 *   orc_solve_f64     FP64 everywhere (the FP64 oracle of the solve).
 *   orc_solve_f32ref  the reference's arithmetic: FP32 matrices/eigenvalues/zop
 *                     rounded from FP64 values (Float32Array stores), every
 *                     transform output accumulated in FP32 in increasing k
 *                     order (solver.js:32-58), FP32 Thomas (solver.js:86-101).
 *
 * Layout: field[x + nx (y + ny z)] on the (padded) grid.
 * Pinned by tests: scipy.fft.dst(type=2, norm='ortho'), closed-form
 * eigenvalues, a manufactured solution and a sparse-LU solve.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

void orc_dst_matrix_f64(int N, double* q, double* qi) {
  for (int i = 0; i < N; i++)
    for (int l = 0; l < N; l++) {
      double c = sqrt((l == N - 1 ? 1.0 : 2.0) / N);
      double v = c * sin(M_PI * (l + 1) * (i + 0.5) / N);
      q[(size_t)i * N + l] = v;
      qi[(size_t)l * N + i] = v;
    }
}

void orc_dst_eig_f64(int N, double dx, double* lam) {
  for (int l = 0; l < N; l++) {
    double s = sin(M_PI * (l + 1) / (2.0 * N));
    lam[l] = -4.0 / (dx * dx) * s * s;
  }
}

/* x pass: out[col][l] = sum_i M[l][i] in[col][i], MT[i][l] = M[l][i];
 * y pass: out[k][m][i] = sum_j M[m][j] in[k][j][i].
 * Each output accumulates in increasing i (j) order. */
#define DEF_PASSES(T, SFX)                                                                           \
  static void xpass_##SFX(int nx, int ny, int nz, const T* MT, const T* in, T* out, int nth) {     \
    long rows = (long)ny * nz;                                                                     \
    _Pragma("omp parallel for num_threads(nth) schedule(static)") for (long r = 0; r < rows; r++) { \
      const T* f = in + (size_t)r * nx;                                                            \
      T* o = out + (size_t)r * nx;                                                                 \
      for (int l = 0; l < nx; l++) o[l] = 0;                                                       \
      for (int i = 0; i < nx; i++) {                                                               \
        T fi = f[i];                                                                               \
        const T* m = MT + (size_t)i * nx;                                                          \
        for (int l = 0; l < nx; l++) o[l] += m[l] * fi;                                            \
      }                                                                                            \
    }                                                                                              \
  }                                                                                                \
  static void ypass_##SFX(int nx, int ny, int nz, const T* M, const T* in, T* out, int nth) {      \
    long planes = (long)nz * ny;                                                                   \
    _Pragma("omp parallel for num_threads(nth) schedule(static)") for (long pm = 0; pm < planes; pm++) { \
      long k = pm / ny, m = pm % ny;                                                               \
      const T* pl = in + (size_t)k * nx * ny;                                                      \
      T* o = out + (size_t)k * nx * ny + (size_t)m * nx;                                           \
      for (int i = 0; i < nx; i++) o[i] = 0;                                                       \
      for (int j = 0; j < ny; j++) {                                                               \
        T w = M[(size_t)m * ny + j];                                                               \
        const T* row = pl + (size_t)j * nx;                                                        \
        for (int i = 0; i < nx; i++) o[i] += w * row[i];                                           \
      }                                                                                            \
    }                                                                                              \
  }

DEF_PASSES(double, d)
DEF_PASSES(float, f)

/* zop per level: (lo, up, diag, rhoC) (solver.js:139-146) */
static void zop_d(int nz, double dz, const double* rhoC, const double* rhoW, double* zo) {
  double a = 1.0 / (dz * dz);
  for (int k = 0; k < nz; k++) {
    double lo = k > 0 ? rhoW[k] * a : 0.0, up = k < nz - 1 ? rhoW[k + 1] * a : 0.0;
    zo[4 * k] = lo; zo[4 * k + 1] = up; zo[4 * k + 2] = -(lo + up); zo[4 * k + 3] = rhoC[k];
  }
}

void orc_solve_f64(int nx, int ny, int nz, double dx, double dz, const double* rhoC, const double* rhoW,
                   const double* rhs, double* phi, int nth) {
  size_t cells = (size_t)nx * ny * nz;
  double *qx = malloc(sizeof(double) * nx * nx), *qix = malloc(sizeof(double) * nx * nx);
  double *qy = malloc(sizeof(double) * ny * ny), *qiy = malloc(sizeof(double) * ny * ny);
  double *lam = malloc(sizeof(double) * (nx + ny)), *zo = malloc(sizeof(double) * 4 * nz);
  double *tmp = malloc(sizeof(double) * cells), *hat = malloc(sizeof(double) * cells);
  if (nth < 1) nth = 1;
  orc_dst_matrix_f64(nx, qx, qix);
  orc_dst_matrix_f64(ny, qy, qiy);
  orc_dst_eig_f64(nx, dx, lam);
  orc_dst_eig_f64(ny, dx, lam + nx);
  zop_d(nz, dz, rhoC, rhoW, zo);
  /* forward x: M = qi (M[l][i] = v_l(i)), so MT[i][l] = q[i][l] */
  xpass_d(nx, ny, nz, qx, rhs, tmp, nth);
  /* forward y: M[m][j] = v_m(j) = qi */
  ypass_d(nx, ny, nz, qiy, tmp, hat, nth);
  /* Thomas per (l, m) */
  long modes = (long)nx * ny;
#pragma omp parallel num_threads(nth)
  {
    double* cp = malloc(sizeof(double) * nz);
#pragma omp for schedule(static)
    for (long b = 0; b < modes; b++) {
      int l = (int)(b % nx), m = (int)(b / nx);
      double shift = lam[l] + lam[nx + m];
      double cprev = 0, dprev = 0;
      for (int k = 0; k < nz; k++) {
        const double* o = zo + 4 * k;
        size_t idx = (size_t)b + (size_t)modes * k;
        double den = o[2] + shift * o[3] - o[0] * cprev;
        cprev = o[1] / den;
        dprev = (hat[idx] - o[0] * dprev) / den;
        cp[k] = cprev;
        hat[idx] = dprev;
      }
      double above = hat[(size_t)b + (size_t)modes * (nz - 1)];
      for (int k = nz - 2; k >= 0; k--) {
        size_t idx = (size_t)b + (size_t)modes * k;
        above = hat[idx] - cp[k] * above;
        hat[idx] = above;
      }
    }
    free(cp);
  }
  /* inverse y: M[j][m] = v_m(j) = q */
  ypass_d(nx, ny, nz, qy, hat, tmp, nth);
  /* inverse x: M[i][l] = v_l(i) = q, MT[l][i] = qi */
  xpass_d(nx, ny, nz, qix, tmp, phi, nth);
  free(qx); free(qix); free(qy); free(qiy); free(lam); free(zo); free(tmp); free(hat);
}

void orc_solve_f32ref(int nx, int ny, int nz, double dx, double dz, const double* rhoC, const double* rhoW,
                      const float* rhs, float* phi, int nth) {
  size_t cells = (size_t)nx * ny * nz;
  double *qx = malloc(sizeof(double) * nx * nx), *qix = malloc(sizeof(double) * nx * nx);
  double *qy = malloc(sizeof(double) * ny * ny), *qiy = malloc(sizeof(double) * ny * ny);
  double *lamd = malloc(sizeof(double) * (nx + ny)), *zod = malloc(sizeof(double) * 4 * nz);
  float *fqx = malloc(sizeof(float) * nx * nx), *fqix = malloc(sizeof(float) * nx * nx);
  float *fqy = malloc(sizeof(float) * ny * ny), *fqiy = malloc(sizeof(float) * ny * ny);
  float *lam = malloc(sizeof(float) * (nx + ny)), *zo = malloc(sizeof(float) * 4 * nz);
  float *tmp = malloc(sizeof(float) * cells), *hat = malloc(sizeof(float) * cells), *scr = malloc(sizeof(float) * cells);
  if (nth < 1) nth = 1;
  orc_dst_matrix_f64(nx, qx, qix);
  orc_dst_matrix_f64(ny, qy, qiy);
  orc_dst_eig_f64(nx, dx, lamd);
  orc_dst_eig_f64(ny, dx, lamd + nx);
  zop_d(nz, dz, rhoC, rhoW, zod);
  for (int i = 0; i < nx * nx; i++) { fqx[i] = (float)qx[i]; fqix[i] = (float)qix[i]; }
  for (int i = 0; i < ny * ny; i++) { fqy[i] = (float)qy[i]; fqiy[i] = (float)qiy[i]; }
  for (int i = 0; i < nx + ny; i++) lam[i] = (float)lamd[i];
  for (int i = 0; i < 4 * nz; i++) zo[i] = (float)zod[i];
  xpass_f(nx, ny, nz, fqx, rhs, tmp, nth);
  ypass_f(nx, ny, nz, fqiy, tmp, hat, nth);
  long modes = (long)nx * ny;
#pragma omp parallel for num_threads(nth) schedule(static)
  for (long b = 0; b < modes; b++) {
    int l = (int)(b % nx), m = (int)(b / nx);
    float shift = lam[l] + lam[nx + m];
    float cprev = 0, dprev = 0;
    for (int k = 0; k < nz; k++) {
      const float* o = zo + 4 * k;
      size_t idx = (size_t)b + (size_t)modes * k;
      float den = o[2] + shift * o[3] - o[0] * cprev;
      cprev = o[1] / den;
      dprev = (hat[idx] - o[0] * dprev) / den;
      scr[idx] = cprev;
      hat[idx] = dprev;
    }
    float above = hat[(size_t)b + (size_t)modes * (nz - 1)];
    for (int k = nz - 2; k >= 0; k--) {
      size_t idx = (size_t)b + (size_t)modes * k;
      above = hat[idx] - scr[idx] * above;
      hat[idx] = above;
    }
  }
  ypass_f(nx, ny, nz, fqy, hat, tmp, nth);
  xpass_f(nx, ny, nz, fqix, tmp, phi, nth);
  free(qx); free(qix); free(qy); free(qiy); free(lamd); free(zod);
  free(fqx); free(fqix); free(fqy); free(fqiy); free(lam); free(zo); free(tmp); free(hat); free(scr);
}
