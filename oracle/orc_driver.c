/*
 * CPU oracle of one whole model step -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Calls the kernels in the order of RegionalModel.encStep (model.js:1442-1476)
 * with the buffer roles of makeBindGroups (model.js:1248-1274): the FP32
 * kernels with the FP32 reference-arithmetic solve (orc_step_f32), or the FP64
 * kernels with the FP64 solve (orc_step_f64, the FP64 oracle of the step).
 * orc_step_f64_split runs the FP64 step in two parts around the solve (part 0:
 * every kernel up to div, which writes rhs; part 1: proj from phi), so a caller
 * can supply phi from another FP64 solve of the same rhs.
 */
#include <stdlib.h>

#include "oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void orc_step_f32(const float* U, const float* BS, const float* BD, const float* CI, int run_lsf, int parity,
                  float* V2, float* Q12, float* Q22, float* Vt, float* Q1t, float* Q2t, float* aux, float* phT,
                  float* lsfA, float* lsfB, float* ciMap, float* rhs, float* phi, int nthreads) {
  int nx = (int)U[0], ny = (int)U[1], nz = (int)U[2];
  int nxp = (int)U[37], nyp = (int)U[38];
  size_t nv = (size_t)(nx + 1) * (ny + 1) * (nz + 1) * 4, nq = (size_t)nx * ny * nz * 4;
  float *Vc = V2 + (parity ? nv : 0), *Vn = V2 + (parity ? 0 : nv);
  float *Q1a = Q12, *Q1b = Q12 + nq, *Q2a = Q22, *Q2b = Q22 + nq;
  double* rc = malloc(sizeof(double) * nz);
  double* rw = malloc(sizeof(double) * (nz + 1));
  for (int k = 0; k < nz; k++) rc[k] = BS[k * 8 + 2];
  for (int k = 0; k <= nz; k++) rw[k] = BS[k * 8 + 3];
  orc_set_threads(nthreads);
  if (run_lsf) orc_lsf_f32(U, BS, BD, lsfA, lsfB, phT);
  orc_ci_f32(U, BS, CI, ciMap);
  orc_nut_f32(U, BS, Vc, Q1a, aux);
  orc_advv1_f32(U, BS, Vc, Vt);
  orc_advv2_f32(U, BS, Vc, Vt, Vn);
  orc_advq1_f32(U, BS, Vc, Q1a, Q2a, Q1t, Q2t);
  orc_advq2_f32(U, BS, Vc, Q1a, Q2a, Q1t, Q2t, Q1b, Q2b);
  orc_phys_f32(U, BS, Vn, Q1b, Q2b, aux, lsfA, lsfB, phT, ciMap, Vc, Q1a, Q2a);
  orc_div_f32(U, BS, Vc, rhs);
  orc_solve_f32ref(nxp, nyp, nz, (double)U[3], (double)U[4], rc, rw, rhs, phi, nthreads);
  orc_proj_f32(U, BS, Vc, phi, Vn);
  free(rc);
  free(rw);
}

void orc_step_f64(const float* U, const float* BS, const float* BD, const float* CI, int run_lsf, int parity,
                  double* V2, double* Q12, double* Q22, double* Vt, double* Q1t, double* Q2t, double* aux, double* phT,
                  double* lsfA, double* lsfB, double* ciMap, double* rhs, double* phi, int nthreads) {
  int nx = (int)U[0], ny = (int)U[1], nz = (int)U[2];
  int nxp = (int)U[37], nyp = (int)U[38];
  size_t nv = (size_t)(nx + 1) * (ny + 1) * (nz + 1) * 4, nq = (size_t)nx * ny * nz * 4;
  double *Vc = V2 + (parity ? nv : 0), *Vn = V2 + (parity ? 0 : nv);
  double *Q1a = Q12, *Q1b = Q12 + nq, *Q2a = Q22, *Q2b = Q22 + nq;
  double* rc = malloc(sizeof(double) * nz);
  double* rw = malloc(sizeof(double) * (nz + 1));
  for (int k = 0; k < nz; k++) rc[k] = BS[k * 8 + 2];
  for (int k = 0; k <= nz; k++) rw[k] = BS[k * 8 + 3];
  orc_set_threads(nthreads);
  if (run_lsf) orc_lsf_f64(U, BS, BD, lsfA, lsfB, phT);
  orc_ci_f64(U, BS, CI, ciMap);
  orc_nut_f64(U, BS, Vc, Q1a, aux);
  orc_advv1_f64(U, BS, Vc, Vt);
  orc_advv2_f64(U, BS, Vc, Vt, Vn);
  orc_advq1_f64(U, BS, Vc, Q1a, Q2a, Q1t, Q2t);
  orc_advq2_f64(U, BS, Vc, Q1a, Q2a, Q1t, Q2t, Q1b, Q2b);
  orc_phys_f64(U, BS, Vn, Q1b, Q2b, aux, lsfA, lsfB, phT, ciMap, Vc, Q1a, Q2a);
  orc_div_f64(U, BS, Vc, rhs);
  orc_solve_f64(nxp, nyp, nz, (double)U[3], (double)U[4], rc, rw, rhs, phi, nthreads);
  orc_proj_f64(U, BS, Vc, phi, Vn);
  free(rc);
  free(rw);
}

void orc_step_f64_split(const float* U, const float* BS, const float* BD, const float* CI, int run_lsf, int parity,
                        double* V2, double* Q12, double* Q22, double* Vt, double* Q1t, double* Q2t, double* aux,
                        double* phT, double* lsfA, double* lsfB, double* ciMap, double* rhs, double* phi, int nthreads,
                        int part) {
  int nx = (int)U[0], ny = (int)U[1], nz = (int)U[2];
  size_t nv = (size_t)(nx + 1) * (ny + 1) * (nz + 1) * 4, nq = (size_t)nx * ny * nz * 4;
  double *Vc = V2 + (parity ? nv : 0), *Vn = V2 + (parity ? 0 : nv);
  double *Q1a = Q12, *Q1b = Q12 + nq, *Q2a = Q22, *Q2b = Q22 + nq;
  orc_set_threads(nthreads);
  if (part == 0) {
    if (run_lsf) orc_lsf_f64(U, BS, BD, lsfA, lsfB, phT);
    orc_ci_f64(U, BS, CI, ciMap);
    orc_nut_f64(U, BS, Vc, Q1a, aux);
    orc_advv1_f64(U, BS, Vc, Vt);
    orc_advv2_f64(U, BS, Vc, Vt, Vn);
    orc_advq1_f64(U, BS, Vc, Q1a, Q2a, Q1t, Q2t);
    orc_advq2_f64(U, BS, Vc, Q1a, Q2a, Q1t, Q2t, Q1b, Q2b);
    orc_phys_f64(U, BS, Vn, Q1b, Q2b, aux, lsfA, lsfB, phT, ciMap, Vc, Q1a, Q2a);
    orc_div_f64(U, BS, Vc, rhs);
  } else {
    orc_proj_f64(U, BS, Vc, phi, Vn);
  }
}
