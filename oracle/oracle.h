/*
 * CPU oracle of OutbreakLab's model time step -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's WGSL kernels (model.js, solver.js,
 * noise.js), written to check the HIP path against.  Nothing in the product
 * links or calls it: only tests/, the smoke check and the CPU-baseline leg of
 * bench.py load liboracle.so, and only as the checker.
 *
 * The same source (orc_step.c) is compiled twice: with REAL = float (suffix
 * _f32, the reference's FP32 arithmetic, no FMA contraction) and with
 * REAL = double (suffix _f64, the FP64 oracle of the same step).  Fields are
 * passed as REAL arrays in the reference's texel layouts:
 *   V        (nx+1)(ny+1)(nz+1) x 4   u,v,w,theta   index x+(nx+1)(y+(ny+1)z)
 *   Q1, Q2   nx*ny*nz x 4             qv,qc,qi,qr / qs,qg,Nr,spare
 *   aux      nx*ny*nz x 2             nu_h, nu_v
 *   phT      nx*ny*nz                 target hydrostatic phi
 *   lsfA/B, ciMap, nx*ny x 4          2-D column maps
 *   rhs, phi nxp*nyp*nz               padded pressure grid
 * The uniform block U is always the reference's 76-float block (19 vec4f,
 * MODEL_FIELDS, model.js:23-37) and BS the (nz+1)*8 base-state table
 * (sounding.js:139-197), both float as on the GPU.
 *
 * Parity status per function (what pins it) is listed in the header of
 * orc_step.c.
 */
#ifndef OBL_ORACLE_H
#define OBL_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define ORC_DECL(R, S)                                                                              \
  void orc_lsf_##S(const float* U, const float* BS, const float* BD, R* lsfA, R* lsfB, R* phT);     \
  void orc_ci_##S(const float* U, const float* BS, const float* CI, R* ciMap);                      \
  void orc_init_##S(const float* U, const float* BS, const R* lsfA, const R* lsfB, R* V, R* Q1,     \
                    R* Q2, int seed);                                                               \
  void orc_nut_##S(const float* U, const float* BS, const R* V, const R* Q1, R* aux);               \
  void orc_advv1_##S(const float* U, const float* BS, const R* V, R* Vt);                           \
  void orc_advv2_##S(const float* U, const float* BS, const R* V, const R* Vt, R* Vout);            \
  void orc_advq1_##S(const float* U, const float* BS, const R* V, const R* Q1, const R* Q2, R* Q1t,  \
                     R* Q2t);                                                                       \
  void orc_advq2_##S(const float* U, const float* BS, const R* V, const R* Q1, const R* Q2,         \
                     const R* Q1t, const R* Q2t, R* Q1o, R* Q2o);                                   \
  void orc_phys_##S(const float* U, const float* BS, const R* Vin, const R* Q1in, const R* Q2in,    \
                    const R* aux, const R* lsfA, const R* lsfB, const R* phT, const R* ciMap,       \
                    R* Vout, R* Q1o, R* Q2o);                                                       \
  void orc_micro_point_##S(const float* U, const float* BS, int k, R* st);                          \
  void orc_fall_##S(const float* U, double q, double n, double rho, double* out);                \
  void orc_div_##S(const float* U, const float* BS, const R* V, R* rhs);                            \
  void orc_proj_##S(const float* U, const float* BS, const R* V, const R* phi, R* Vout);            \
  void orc_bubble_##S(const float* U, R* V, double x, double y, double z, double rh, double rz,      \
                      double dth);                                                                  \
  void orc_maps_##S(const float* U, const float* BS, const R* V, const R* Q1, const R* Q2,          \
                    const R* phi, R* maps);                                                         \
  /* hurricane mode (model.js:88-119, 284-389, 392-576, 579-907, 948-1026): sfc = nx*ny x 4 float    \
     (land fraction, SST K, 0, 0) or NULL, VX = the vortex table nz*nr x 4 float or NULL, Vval = the \
     winds the V advection carries, phi = the padded solve */                                        \
  void orc_init2_##S(const float* U, const float* BS, const R* lsfA, const R* lsfB, const float* VX, \
                     R* V, R* Q1, R* Q2, int seed);                                                 \
  void orc_nut2_##S(const float* U, const float* BS, const R* V, const R* Q1, const float* sfc,      \
                    R* aux);                                                                        \
  void orc_advv1_2_##S(const float* U, const float* BS, const R* V, const R* Vval, R* Vt);          \
  void orc_advv2_2_##S(const float* U, const float* BS, const R* V, const R* Vval, const R* Vt,     \
                       R* Vout);                                                                    \
  void orc_phys2_##S(const float* U, const float* BS, const R* Vin, const R* Q1in, const R* Q2in,   \
                     const R* aux, const R* lsfA, const R* lsfB, const R* phT, const R* ciMap,      \
                     const float* sfc, const R* phi, R* Vout, R* Q1o, R* Q2o);                      \
  void orc_maps2_##S(const float* U, const float* BS, const R* V, const R* Q1, const R* Q2,         \
                     const R* phi, const float* sfc, R* maps);                                      \
  void orc_prep_##S(const float* U, const float* BS, const R* V, const R* phi, R* Vp);              \
  void orc_filt_##S(const float* U, const float* BS, const R* V, const R* Q1, const R* Q2, R* Vo,   \
                    R* Q1o, R* Q2o);

ORC_DECL(float, f32)
ORC_DECL(double, f64)

/* Virtual radar cut (radar.js:41-268), FP32: both passes of one cut, and the beam geometry of a gate. */
void orc_radar_cut(const float* ru, const float* BS, const float* V, const float* Q1, const float* Q2, const float* aux,
                   const float* tor, uint32_t* out, float* kdp);
/* listed gates' moment values before rounding and decision margins, and the values at a given geometry (the
   GPU's): rg = n (radial, gate) pairs, gg = null or n x 7 floats, out[n][24] */
void orc_radar_gates(const float* ru, const float* BS, const float* V, const float* Q1, const float* Q2,
                     const float* aux, const float* tor, int n, const int* rg, const float* gg, float* out);
void orc_radar_geom(const float* ru, int ri, int gi, float* hs_xy);

/* Noise hash (noise.js:5-9), integer-exact. */
uint32_t orc_pcg(uint32_t v);
uint32_t orc_hash3i(int32_t x, int32_t y, int32_t z);

/* PoissonDirect (solver.js:105-165).
 *   f64:    FP64 matrices, FP64 GEMMs and Thomas (the FP64 oracle of the solve).
 *   f32ref: the reference arithmetic -- FP32 matrices rounded from FP64, each
 *           GEMM output accumulated sequentially in k in FP32 (solver.js:32-58),
 *           FP32 Thomas with c' stored (solver.js:86-101).
 * rhoC[nz], rhoW[nz+1] as in setDensity (solver.js:139-146). */
void orc_solve_f64(int nx, int ny, int nz, double dx, double dz, const double* rhoC,
                   const double* rhoW, const double* rhs, double* phi, int nthreads);
void orc_solve_f32ref(int nx, int ny, int nz, double dx, double dz, const double* rhoC,
                      const double* rhoW, const float* rhs, float* phi, int nthreads);
/* orthonormal DST-II matrices: q[i*N+l] = v_l(i), qi[l*N+i] = v_l(i) (solver.js:115-123) */
void orc_dst_matrix_f64(int N, double* q, double* qi);
/* horizontal eigenvalues lam[l] = -4/dx^2 sin^2(pi(l+1)/(2N)) (solver.js:129-131) */
void orc_dst_eig_f64(int N, double dx, double* lam);

/* One whole model step in the order of RegionalModel.encStep (model.js:1442-1476).
 * State: V[2], Q1[2], Q2[2] as two consecutive copies; the state enters in
 * V[parity], Q1[0], Q2[0] and leaves in V[1-parity], Q1[0], Q2[0]. */
void orc_step_f32(const float* U, const float* BS, const float* BD, const float* CI, int run_lsf,
                  int parity, float* V2, float* Q12, float* Q22, float* Vt, float* Q1t, float* Q2t,
                  float* aux, float* phT, float* lsfA, float* lsfB, float* ciMap, float* rhs,
                  float* phi, int nthreads);
/* The same step with the FP64 kernels and the FP64 solve (every array double). */
void orc_step_f64(const float* U, const float* BS, const float* BD, const float* CI, int run_lsf,
                  int parity, double* V2, double* Q12, double* Q22, double* Vt, double* Q1t, double* Q2t,
                  double* aux, double* phT, double* lsfA, double* lsfB, double* ciMap, double* rhs,
                  double* phi, int nthreads);
/* the FP64 step in two parts around the solve: part 0 runs every kernel up to div (rhs), part 1 the projection
 * from phi (orc_driver.c) */
void orc_step_f64_split(const float* U, const float* BS, const float* BD, const float* CI, int run_lsf, int parity,
                        double* V2, double* Q12, double* Q22, double* Vt, double* Q1t, double* Q2t, double* aux,
                        double* phT, double* lsfA, double* lsfB, double* ciMap, double* rhs, double* phi, int nthreads,
                        int part);

#ifdef __cplusplus
}
#endif
#endif
