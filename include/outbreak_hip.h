/*
 * liboutbreak_hip -- C ABI of the MI355X (gfx950) model time step.
 *
 * The reference has no FFI: its step API is the JavaScript class pair
 * RegionalModel (model.js:1156-1496) and PoissonDirect (solver.js:105-165)
 * over WebGPU objects. Each entry point below replaces one of their methods
 * (cited per function); a Node N-API addon and a Python ctypes layer rebuild
 * the two classes on top of it with the reference's names and semantics.
 *
 * Conventions
 *   - Every call returns an int status: 0 = OK, < 0 = error. Nothing throws
 *     across the ABI. The last error text of a handle is kept until the next
 *     failing call on it (obl_last_error); failures without a handle set the
 *     process-wide text (obl_global_error).
 *   - Host arrays are copied during the call (the reference's GPU.write /
 *     queue.writeBuffer semantics, util.js:190).
 *   - GPU work is enqueued on the handle's own HIP stream and runs in order,
 *     like the reference's single queue; obl_sync waits for it
 *     (onSubmittedWorkDone, app.js:1704).
 *   - Layouts are the reference's texel layouts (x fastest):
 *       V   float4 (nx+1)(ny+1)(nz+1)    u, v, w, theta     (two copies, parity)
 *       Q1  float4 nx ny nz              qv, qc, qi, qr     (two copies)
 *       Q2  float4 nx ny nz              qs, qg, Nr, spare  (two copies)
 *       aux float2 nx ny nz              nu_h, nu_v
 *       phi, rhs float (nx+2p)(ny+2p)nz  the padded pressure grid
 *     The uniform block is the reference's 76-float MODEL_FIELDS block
 *     (model.js:23-37), packed by the host exactly like writeUniforms
 *     (model.js:1391-1425).
 */
#ifndef OUTBREAK_HIP_H
#define OUTBREAK_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OBL_ABI_VERSION 2
#define OBL_UNIFORM_FLOATS 76
#define OBL_BOUND_FLOATS 148 /* one boundary record, model.js:1371-1381 */
#define OBL_MAX_BOUNDS 8     /* MAX_BOUNDS, model.js:38 */
#define OBL_BOUND_POINTS 32  /* points per boundary polyline, model.js:1373-1380 */
#define OBL_CI_FLOATS 12     /* one CI tower, model.js:1383-1389 */
#define OBL_VORTEX_MAX_RADII 2048 /* radii of the vortex table, VX_MAXR model.js:38 */
#define OBL_RADAR_SCRATCH_FLOATS (720 * 1192) /* obl_radar_cut_into's KDP scratch: radials x Doppler gates */
#define OBL_MAX_DEVICES 64
#define OBL_MAX_CI 256       /* MAX_CI, model.js:38 */

typedef struct obl_model obl_model;
typedef struct obl_poisson obl_poisson;

/* Grid of RegionalModel.allocTextures (model.js:1202-1238): dz = 20000 / nz,
 * pad = gridPad(dx) (model.js:1122) cells of solve padding per side. */
typedef struct {
  int nx, ny, nz;
  double dx, dz;
  int pad;
} obl_grid;

/* Field ids for obl_upload / obl_download / obl_field_ptr */
enum {
  OBL_FIELD_V = 0,    /* copy = parity 0/1 */
  OBL_FIELD_Q1 = 1,   /* copy 0/1 */
  OBL_FIELD_Q2 = 2,   /* copy 0/1 */
  OBL_FIELD_VT = 3,
  OBL_FIELD_Q1T = 4,
  OBL_FIELD_Q2T = 5,
  OBL_FIELD_AUX = 6,
  OBL_FIELD_PHT = 7,
  OBL_FIELD_LSFA = 8,
  OBL_FIELD_LSFB = 9,
  OBL_FIELD_CIMAP = 10,
  OBL_FIELD_PHI = 11,
  OBL_FIELD_RHS = 12,
  OBL_FIELD_BS = 13,
  OBL_FIELD_SFC = 14, /* hurricane mode: sea surface map, nx*ny float4 (land fraction, SST K, 0, 0) */
  OBL_FIELD_VP = 15   /* hurricane mode: the winds the advection carries (WGSL_M_PREP), as V */
};

/* obl_step flags */
#define OBL_STEP_LSF 1u   /* run the large-scale forcing this step (model.js:1461) */
#define OBL_STEP_FILT 4u  /* hurricane mode: run the 2-dx filter before the step (model.js:1447-1455) */
#define OBL_STEP_JUDGE 2u /* with a surrogate attached: also run the plain microphysics from the same
                             inputs and add the difference to the judge statistics (not part of the
                             step's result; no reference counterpart) */

/* Solver precision (obl_set_solver_mode / obl_poisson_solve) */
enum {
  OBL_GEMM_FP32 = 0,   /* FP32 MFMA sine transforms */
  OBL_GEMM_BF16X3 = 1, /* split hi/lo BF16 products, FP32 accumulate (about 16 bits) */
  OBL_GEMM_BF16X6 = 2  /* three BF16 pieces and the six products down to order 2^-16: FP32 accuracy */
};
enum {
  OBL_THOMAS_FP32 = 0,  /* the reference's FP32 Thomas (solver.js:86-101) */
  OBL_THOMAS_FP64 = 1,  /* FP64 arithmetic for every (l, m) */
  OBL_THOMAS_MIXED = 2  /* FP64 where the vertical system's condition estimate 4 / (|lx + ly| dz^2) exceeds
                           fp64_cond (model default 64); the other modes in FP32 through one reciprocal per level */
};

/* ---------------- errors and version ---------------- */
int obl_abi_version(void);
const char* obl_last_error(const obl_model* h);
const char* obl_global_error(void);

/* ---------------- RegionalModel ---------------- */
/* allocate (model.js:1186-1238): all fields, the padded solve and its matrices.
 * Fails with the reference's size message when a field would not fit. */
int obl_create(const obl_grid* g, obl_model** out);
/* the device bytes obl_create allocates for a grid (gridBytes, model.js:1125-1128, for this layout); host only */
double obl_grid_bytes(const obl_grid* g);
/* destroy (model.js:1239-1247) */
void obl_destroy(obl_model* h);
/* setSounding -> bs + solver.setDensity (model.js:1276-1288): B holds (nz+1)*8 floats */
int obl_set_base_state(obl_model* h, const float* B, int nz);
/* setBoundaries (model.js:1371-1381): count records of 148 floats; a record's first float is its point
 * count, 0..32 (larger counts are rejected) */
int obl_set_boundaries(obl_model* h, const float* f, int count);
/* setCI (model.js:1383-1389): count towers of 12 floats */
int obl_set_ci(obl_model* h, const float* f, int count);
/* setSurface (model.js:1298-1303): surf = nx*ny float4 (land fraction, SST K, 0, 0) per column, the
 * sea that hurricane mode (hur.x = 1 in the uniform block) reads; NULL = outbreak mode (no sea).
 * bytes is the size of the caller's buffer, which must hold nx*ny*16 bytes. */
int obl_set_surface(obl_model* h, const float* surf, size_t bytes);
/* setPrepField (model.js:1290-1296): allocate (on) or free the Vp field of hurricane runs; while it
 * exists every step first runs WGSL_M_PREP and the wind advection carries Vp's values. */
int obl_set_prep_field(obl_model* h, int on);
/* setVortex (model.js:1318-1365), its device half: the table the init kernel reads, nz x nr float4
 * (tangential wind m/s, theta' K, 0, 0) at radii i * dr, built on the host; used by obl_reset when
 * vtx2.z = 1, with vtx = (x, y, dr, nr). NULL / nr 0 drops it. bytes is the size of the caller's
 * buffer, which must hold nz*nr*16 bytes. */
int obl_set_vortex(obl_model* h, const float* table, int nr, size_t bytes);
/* setVortex's table builder (model.js:1318-1365), host code only (no GPU, no handle): the balanced warm-core
 * vortex v(r, z) = V(r) F(z) on the base state B ((nz+1)*8 floats), spec = (vmax m/s, rmw m, r0 m, ztop m),
 * mid = 1 for a mid-level vortex. Writes tab (nz x nr float4: tangential wind, theta', 0, 0; tab_floats is its
 * capacity, nz * 2048 * 4 always suffices) and info = (dr m, nr, central pressure deficit hPa, least core
 * d theta / dz K/km, warm-core peak K). Returns 0, -1 for bad arguments, -2 if tab is too small. */
int obl_vortex_table(int nz, double dz, double dx, double lat, const float* B, const double* spec, int mid,
                     float* tab, size_t tab_floats, double* info);
/* reset (model.js:1427-1440): lsf + init with the given uniform block (written
 * for dt = 1), clears phi/rhs, copies V0 -> V1. seed 0 = the reference's noise. */
int obl_reset(obl_model* h, const float* uniforms, int seed);
/* encStep (model.js:1442-1476): enqueue one step with this uniform block; the
 * state moves from V[parity] to V[1-parity]. The caller keeps time/steps/parity
 * as the reference does and passes parity explicitly. */
int obl_step(obl_model* h, const float* uniforms, int parity, unsigned flags);
/* requestMaps (model.js:1478-1494) with the map kernel (model.js:1028-1100): enqueue the 2-D maps of
 * the current state (V[parity] -- the copy the next obl_step would read --, Q1[0], Q2[0] and phi of the
 * last solve) and their copy into one of two pinned host slots. *slot receives 0 or 1, or -1 when both
 * slots still await collection (the reference returns null then). The maps are m0..m3, each nx*ny
 * float4 (x fastest), stored one after the other (16*nx*ny floats), as the reference's map buffer:
 *   m0 = (composite dBZ, dBZ at ~1 km, lowest-level theta - theta0(0), zeta at ~1 km 1/s)
 *   m1 = (max w below 12 km, 2-5 km updraft helicity m^2/s^2, lowest-level convergence 1e-3/s,
 *         lowest-level wind speed)
 *   m2 = (u, v at the lowest level, u, v at ~1 km)
 *   m3 = (surface pressure perturbation hPa, 10.7 um brightness temperature K, ~10-m wind, zeta at ~5 km) */
int obl_maps_request(obl_model* h, const float* uniforms, int parity, int* slot);
/* Wait for a requested slot's maps and copy them (16*nx*ny floats; dst_floats at least that) to dst;
 * the slot is free again. */
int obl_maps_collect(obl_model* h, int slot, float* dst, size_t dst_floats);

/* ---------------- virtual WSR-88D radars ----------------
 * The reference's RadarSim (radar.js:271-413) keeps the VCP layout, scan schedule and uniform
 * packing on the host; these calls are its GPU side (radar.js:41-268, 274-275, 318-328). A radar owns a volume buffer
 * of `bytes` (8-bit Archive II codes, 4 gates per 32-bit word, each cut's moments radial-major at
 * the word offsets its uniform block names), a KDP scratch of `kdp_floats` (>= radials x Doppler
 * gates of any cut) and two pinned readback slots. Errors are reported on the model's handle.
 * The model owns its radars: obl_destroy frees any not yet destroyed, after which their handles
 * are dead. */
typedef struct obl_radar obl_radar;
int obl_radar_create(obl_model* h, size_t bytes, size_t kdp_floats, obl_radar** out);
void obl_radar_destroy(obl_radar* r);
/* setTornadoes (radar.js:282-287): up to 64 sub-grid tornadoes of 8 floats (x, y, core radius, Vmax,
 * debris top, debris dBZ, debris 0-1, 0), model metres; shared by every radar of the model */
int obl_radar_set_tornadoes(obl_model* h, const float* f, int n);
/* encCut (radar.js:369-390): enqueue one cut from its RADAR_FIELDS block (44 floats, radar.js:37-39)
 * on the current state (V[parity], Q1[0], Q2[0], aux); PHIDP's pass runs when the moment mask holds
 * it. The block must name this model's dimensions and fit the radar's buffers. */
int obl_radar_cut(obl_radar* r, const float* ru, int parity);
/* One radar cut, as obl_radar_cut, into a volume buffer of the caller's (out_dev: device memory of out_bytes,
 * the cut's moments at the block's word offsets), allocating nothing: no hipMalloc, hipFree or allocating
 * library call, so callers that cannot allocate (code under HIP graph capture, serving loops) can render cuts.
 * No counterpart in the reference. The KDP scratch between its two kernels is a fixed device array of
 * OBL_RADAR_SCRATCH_FLOATS floats per GPU (720 radials x 1192 Doppler gates, every VCP's largest cut), held until
 * the process ends; a cut needing more is rejected. Every call on one GPU shares that array: calls that overlap
 * on one GPU (two streams, two threads) race. Enqueued on the model's stream; no timing (obl_profile) is taken. */
int obl_radar_cut_into(obl_model* h, const float* ru, int parity, uint32_t* out_dev, size_t out_bytes);
/* Diagnostic, no counterpart in the reference: the geometry of listed gates of a cut (ru as obl_radar_cut) as the
 * cut kernel computes it, for comparing the GPU's FP32 geometry with a CPU evaluation of the same formulas.
 * rg = n (radial, gate) pairs (host), out = n x 7 floats (host): model x, y (m), the radial unit vector (ex, ey),
 * the beam-centre height and the heights 0.4 beam widths below and above it (m above the model's ground).
 * Allocates and frees two small device buffers; synchronous on the model's stream. */
int obl_radar_geometry(obl_model* h, const float* ru, int n, const int* rg, float* out);
/* the volume's copy into a pinned slot (the reference's copyBufferToBuffer into read[slot]):
 * *slot = 0 / 1, or -1 while both slots await collection */
int obl_radar_readback(obl_radar* r, int* slot);
/* wait for a slot and copy its `bytes` out into dst (dst_bytes >= bytes); the slot is free again */
int obl_radar_collect(obl_radar* r, int slot, void* dst, size_t dst_bytes);
/* Warm bubble, theta += dth cos^2(pi r/2) (no reference counterpart: the configs
 * start storms with bubbles instead of CI towers), on both V copies. */
int obl_add_bubble(obl_model* h, double x, double y, double z, double rh, double rz, double dth);
/* GPU.write / writeTexture and readback copies (util.js:190; model.js:1478-1494) */
int obl_upload(obl_model* h, int field, int copy, const void* src, size_t bytes, size_t offset);
int obl_download(obl_model* h, int field, int copy, void* dst, size_t bytes, size_t offset);
/* Device pointer of a field, for in-process consumers (radar.js binds m.V[par], m.Q1[0], ...) */
int obl_field_ptr(obl_model* h, int field, int copy, void** ptr, size_t* bytes);
/* onSubmittedWorkDone (app.js:1704) */
int obl_sync(obl_model* h);
/* Precision of the model's pressure solve (defaults: three-piece BF16 GEMM at FP32 accuracy, mixed Thomas) */
int obl_set_solver_mode(obl_model* h, int gemm_mode, int thomas_mode, double fp64_cond);
/* Per-kernel HIP-event timing of subsequent obl_step calls (0 = off). */
int obl_profile(obl_model* h, int on);
/* Average device time per launch (ms) of each timed kernel since obl_profile(h, 1);
 * names are written as a '\n'-separated list. Returns the number of kernels. */
int obl_profile_read(obl_model* h, char* names, size_t names_len, double* ms, int* launches, int max);

/* ---------------- microphysics surrogate: teacher sampling ----------------
 * No reference counterpart (the reference has no surrogate). While enabled, each obl_step
 * appends sampled (input, output) pairs of the microphysics block (model.js:775-899) to a
 * device buffer of `capacity` rows of 24 floats: T, p0, rho0, qv, qc, qr, qi, qs, qg, Nr, Ni,
 * w, dz, dt, then the tendencies of the block's process part (everything before its closing
 * saturation adjustment, model.js:775-874) dT, dqv, dqc, dqi, dqr, dqs, dqg, dNr, dNi, then the
 * level. Points holding more than 1e-9 kg/kg of condensate are kept with probability p_active,
 * the rest with p_clear.
 * capacity 0 disables sampling; enabling again clears the buffer. */
int obl_dump_enable(obl_model* h, unsigned capacity, float p_active, float p_clear);
/* rows written so far (<= capacity); copies up to max_rows rows to out (may be NULL) and
 * returns the device buffer in *dev_buf (may be NULL) */
int obl_dump_read(obl_model* h, float* out, size_t max_rows, size_t* rows, void** dev_buf);

/* ---------------- microphysics surrogate network ----------------
 * No reference counterpart. A 16 -> 128 -> 128 -> 128 -> 16 MLP (GELU) whose 9 used outputs are
 * the tendencies of the microphysics block's process part, scaled by out_scale, except output 7
 * (Nr), which is log((Nr + dNr + 1) / (Nr + 1)); inputs are the 14 values of a sampled row
 * (obl_dump_*), mapped to features on the device. FP32 master parameters, Adam. */
typedef struct obl_mlp obl_mlp;
int obl_mlp_create(obl_mlp** out);
void obl_mlp_destroy(obl_mlp* m);
const char* obl_mlp_last_error(const obl_mlp* m);
int obl_mlp_nparams(void);
int obl_mlp_set_params(obl_mlp* m, const float* p, int n);
int obl_mlp_get_params(obl_mlp* m, float* p, int n);
/* out_scale[9]: tendency scales (targets are tendency / scale); in_lo/in_hi[14]: the training
 * envelope (NULL = unbounded); points outside it use the plain scheme at inference */
int obl_mlp_set_scales(obl_mlp* m, const float* out_scale, const float* in_lo, const float* in_hi);
/* a species absent before the block (below frac * its output scale) is changed at inference only
 * when its predicted tendency exceeds frac * its output scale (default 0.02) */
int obl_mlp_set_dead_zone(obl_mlp* m, float frac);
/* only points holding at least qmin kg/kg of condensate (qc + qi + qr + qs + qg) are handed to the network; the
 * others run the plain block (default 1e-9, every point with condensate) */
int obl_mlp_set_hand_min(obl_mlp* m, float qmin);
/* copies nrows host rows (stride floats each) into device memory the network owns (replacing any
 * earlier copy); obl_mlp_train with rows_dev == NULL trains on them */
int obl_mlp_load_rows(obl_mlp* m, const float* rows_host, size_t nrows, int stride);
/* `steps` Adam steps on batches of `batch` rows drawn (hashed with `seed`) from nrows device rows
 * (rows_dev == NULL: the rows of obl_mlp_load_rows, nrows and stride ignored);
 * quant != 0 rounds weights, features and activations to BF16 in the forward pass as inference does */
int obl_mlp_train(obl_mlp* m, const float* rows_dev, size_t nrows, int stride, int steps, int batch, float lr,
                  unsigned seed, int quant, float* mean_loss);
/* exact evaluation of n host rows: mean loss, and optionally the gradient of the mean loss
 * (obl_mlp_nparams floats) and the raw outputs (n x 16) */
int obl_mlp_eval(obl_mlp* m, const float* rows_host, int n, int stride, int quant, float* loss, float* grads, float* y);
/* the inference form the model step uses (BF16 MFMA, FP32 accumulation): raw outputs (n x 16,
 * before output scaling) for n host rows of >= 14 inputs */
int obl_mlp_infer(obl_mlp* m, const float* rows_host, int n, int stride, float* y);

/* ---------------- microphysics surrogate in the step ----------------
 * No reference counterpart. With a network attached, each obl_step hands the points that hold more
 * than 1e-9 kg/kg of condensate and whose inputs lie in the network's envelope to the network; the rest run the plain
 * block. After inference: a dead zone for absent species, positivity, total-water conservation,
 * latent-heat-consistent temperature, Thompson Nr bounds; a non-finite prediction falls back to
 * the plain block. `capacity` bounds the points per step (0: half the
 * cells); points beyond it run the plain block. net = NULL detaches. The network must stay alive
 * while attached, and retraining it races with steps still in flight (obl_sync first). */
int obl_set_surrogate(obl_model* h, obl_mlp* net, unsigned capacity);
typedef struct {
  double steps;      /* steps run with the surrogate */
  double points;     /* points handed to the network, summed over steps */
  double fallbacks;  /* of which fell back to the plain block */
  double overflow;   /* candidates beyond the capacity (ran the plain block) */
  double judged;     /* steps run with OBL_STEP_JUDGE */
  /* over judged steps, surrogate (s) against plain (p) from the same inputs, all cells:
   * sum (s - p)^2 and sum p^2 for qv qc qi qr qs qg Nr theta */
  double d2[8], r2[8];
  /* column condensate sum_k rho0 (qc + qi + qr + qs + qg) dz (kg m^-2), per column */
  double cc_d2, cc_r2, cc_maxdiff, cc_max;
  /* composite reflectivity (dBZ, model.js:1042-1050), both floored at 0 dBZ (echo weaker than
   * that counts as none), over columns above 0 dBZ in either */
  double dbz_d2, dbz_cols, dbz_maxdiff;
  /* surface precipitation accumulated over judged steps (kg m^-2), per column */
  double pr_d2, pr_r2, pr_sum_s, pr_sum_p;
} obl_surr_stats;
int obl_surrogate_stats(obl_model* h, obl_surr_stats* out, int reset);

/* ---------------- PoissonDirect ---------------- */
/* constructor (solver.js:106-137): nx, ny even; dz, dx in metres */
int obl_poisson_create(int nx, int ny, int nz, double dx, double dz, obl_poisson** out);
void obl_poisson_destroy(obl_poisson* p);
/* setDensity (solver.js:139-146) */
int obl_poisson_set_density(obl_poisson* p, const double* rhoC, const double* rhoW);
/* enc (solver.js:155-163): device pointers rhs -> phi, enqueued on stream (0 = the
 * solver's own stream) */
int obl_poisson_solve(obl_poisson* p, const float* rhs_dev, float* phi_dev, int gemm_mode, int thomas_mode,
                      double fp64_cond, void* stream);
/* host-memory convenience: copies in, solves, copies out, synchronises */
int obl_poisson_solve_host(obl_poisson* p, const float* rhs, float* phi, int gemm_mode, int thomas_mode,
                           double fp64_cond);
const char* obl_poisson_last_error(const obl_poisson* p);
/* Per-pass device time (ms) of the last solve when timing is on */
int obl_poisson_profile(obl_poisson* p, int on);
int obl_poisson_profile_read(obl_poisson* p, double* ms5);

#ifdef __cplusplus
}
#endif
#endif
