// C ABI of the gfx950 model step (declared in include/outbreak_hip.h).
//
// Owns the device fields of one RegionalModel, its HIP stream and its
// PoissonDirect, and records the step's launches in the reference's order
// (RegionalModel.encStep, model.js:1442-1476).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/outbreak_hip.h"
#include "obl_kernels.h"
#include "obl_radar.h"
#include "obl_mlp.h"

using namespace obl;

static thread_local std::string g_err;

static int set_global(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}

// ---------------------------------------------------------------- timing
struct Timer {
  bool on = false;
  std::vector<std::string> names;
  std::vector<double> total;
  std::vector<int> count;
  struct Pend {
    int id;
    hipEvent_t a, b;
  };
  std::vector<Pend> pend;
  std::vector<hipEvent_t> pool;
  hipEvent_t ev() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
  }
  int id_of(const char* n) {
    for (size_t i = 0; i < names.size(); i++)
      if (names[i] == n) return (int)i;
    names.push_back(n);
    total.push_back(0.0);
    count.push_back(0);
    return (int)names.size() - 1;
  }
  void begin(hipStream_t s, const char* n, hipEvent_t* a) {
    if (!on) return;
    (void)n;
    *a = ev();
    hipEventRecord(*a, s);
  }
  void end(hipStream_t s, const char* n, hipEvent_t a) {
    if (!on) return;
    hipEvent_t b = ev();
    hipEventRecord(b, s);
    pend.push_back({id_of(n), a, b});
  }
  void collect() {
    for (auto& p : pend) {
      hipEventSynchronize(p.b);
      float ms = 0.f;
      hipEventElapsedTime(&ms, p.a, p.b);
      total[p.id] += ms;
      count[p.id] += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pend.clear();
  }
  void reset() {
    collect();
    for (auto& t : total) t = 0.0;
    for (auto& c : count) c = 0;
  }
  ~Timer() {
    for (auto& p : pend) {
      hipEventDestroy(p.a);
      hipEventDestroy(p.b);
    }
    for (auto e : pool) hipEventDestroy(e);
  }
};
#define TIMED(T, S, NAME, CALL)   \
  do {                            \
    hipEvent_t _a = nullptr;      \
    (T).begin((S), (NAME), &_a);  \
    CALL;                         \
    (T).end((S), (NAME), _a);     \
  } while (0)

// ---------------------------------------------------------------- PoissonDirect
struct AxisMats {
  int N = 0, Nh = 0, Npad = 0, Kp = 0;
  float* W[4] = {nullptr, nullptr, nullptr, nullptr};  // We, Wo, WeT, WoT
  // multi-piece BF16 form: tiled split matrices [P - 2][forward, inverse], each [2 parities][P][bRB][KB] tiles;
  // NF = column fragments per wave (64 NF columns per block and parity), tilesN column tiles
  int NF = 0, tilesN = 0, bRB = 0, KB = 0;
  uint16_t* Wm[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
};

// Multi-piece BF16 operand geometry (see k_solve.hip): 32-row blocks of a pass's data operand
static int mx_rb(int rows) { return (rows + 32 * MX_MF - 1) / (32 * MX_MF) * MX_MF; }
static size_t mx_bytes(int P, int RB, int KB) { return (size_t)2 * P * RB * (KB + 1) * 1024; }
// bf16 offset of (row, k) in one piece plane of a tiled operand (host copy of mx_off)
static size_t mx_off_host(int row, int k, int KB) {
  const int ri = row & 31, kc = (k >> 3) & 1;
  const int chunk = 2 * ri + (kc ^ (((ri >> 2) ^ (ri >> 3)) & 1));
  return ((size_t)(row >> 5) * (KB + 1) + (k >> 4)) * 512 + chunk * 8 + (k & 7);
}

// round-to-nearest-even FP32 -> BF16 bits
static uint16_t bf16_bits(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float bf16_value(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

struct obl_poisson {
  int nx = 0, ny = 0, nz = 0;
  double dx = 0, dz = 0;
  hipStream_t stream = nullptr;
  AxisMats ax, ay;
  float *tmp = nullptr, *hat = nullptr, *cp32 = nullptr, *lam32 = nullptr;
  double *cp64 = nullptr, *dp64 = nullptr, *lam64 = nullptr, *zop64 = nullptr;
  float4* zop32 = nullptr;
  float *hrhs = nullptr, *hphi = nullptr;  // device staging for solve_host
  // multi-piece BF16 chain: the tiled split operands of the XF, YF, YI and XI passes (sized for 3 pieces,
  // zeroed once: nothing writes their padding) and their 32-row block counts
  uint16_t* mx[4] = {nullptr, nullptr, nullptr, nullptr};
  int mxRB[4] = {0, 0, 0, 0};
  // the modes the Thomas sweep solves in FP64 (capacity nx ny), built for one (Thomas mode, shift bound) at a time
  int2* modes64 = nullptr;
  int n64 = 0, list_mode = -1;
  double list_lo = -1.0;
  std::vector<double> lam64h;  // the permuted eigenvalues, x then y
  std::vector<int2> hmodes;
  std::string err;
  Timer timer;
  int fail(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return -1;
  }
};

static int hip_ok(hipError_t e, std::string* err, const char* what) {
  if (e == hipSuccess) return 0;
  char buf[512];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  if (err) *err = buf;
  else g_err = buf;
  return -1;
}

// orthonormal DST-II v_l(i) = c_l sin(pi (l+1)(i+1/2)/N), c_l = sqrt(2/N) (sqrt(1/N) for
// l = N-1), computed in FP64 and rounded to FP32 (solver.js:115-123), split into the
// folded even/odd half matrices.
static int build_axis(AxisMats& A, int N, std::string* err) {
  A.N = N;
  A.Nh = N / 2;
  A.Npad = (A.Nh + 63) / 64 * 64;
  A.Kp = (A.Nh + 15) / 16 * 16;
  size_t n = (size_t)A.Npad * A.Kp;
  std::vector<float> h[4];
  for (int t = 0; t < 4; t++) h[t].assign(n, 0.0f);
  auto v = [&](int l, int i) {
    double c = sqrt((l == N - 1 ? 1.0 : 2.0) / N);
    return c * sin(M_PI * (l + 1) * (i + 0.5) / N);
  };
  for (int s = 0; s < A.Nh; s++)
    for (int i = 0; i < A.Nh; i++) {
      float ve = (float)v(2 * s, i), vo = (float)v(2 * s + 1, i);
      h[0][(size_t)s * A.Kp + i] = ve;  // We[s][i]
      h[1][(size_t)s * A.Kp + i] = vo;  // Wo[s][i]
      h[2][(size_t)i * A.Kp + s] = ve;  // WeT[i][s]
      h[3][(size_t)i * A.Kp + s] = vo;  // WoT[i][s]
    }
  for (int t = 0; t < 4; t++) {
    if (hip_ok(hipMalloc(&A.W[t], n * sizeof(float)), err, "hipMalloc(sine matrix)")) return -1;
    if (hip_ok(hipMemcpy(A.W[t], h[t].data(), n * sizeof(float), hipMemcpyHostToDevice), err, "hipMemcpy(sine matrix)"))
      return -1;
  }
  // multi-piece copies, in the tiled layout: the block width 64 NF (NF = 2, 3 or 4) that pads the N / 2 columns
  // least, the wider one on a tie (fewer fragment loads per MFMA: with NF = 2 the loads saturate the CU's load
  // path); forward rows n = s (mode), k = i; inverse rows n = i, k = s. Pieces split from the FP64 values.
  int best = 1 << 30;
  for (int nf = 4; nf >= 2; nf--) {
    const int w = 64 * nf, pad = (A.Nh + w - 1) / w * w;
    if (pad < best) {
      best = pad;
      A.NF = nf;
    }
  }
  A.tilesN = (A.Nh + 64 * A.NF - 1) / (64 * A.NF);
  A.bRB = A.tilesN * 2 * A.NF;
  A.KB = (A.Nh + 31) / 32 * 2;  // even: the pass kernels take k-blocks two at a time
  const size_t plane = (size_t)A.bRB * (A.KB + 1) * 512;
  for (int P = 2; P <= 3; P++) {
    std::vector<uint16_t> hm[2];
    for (int inv = 0; inv < 2; inv++) hm[inv].assign(2 * P * plane, 0);
    for (int par = 0; par < 2; par++)
      for (int s = 0; s < A.Nh; s++)
        for (int i = 0; i < A.Nh; i++) {
          const double val = v(2 * s + par, i);
          uint16_t pc[3];
          double r = val;
          for (int q = 0; q < P; q++) {
            pc[q] = bf16_bits((float)r);
            r -= (double)bf16_value(pc[q]);
          }
          for (int q = 0; q < P; q++) {
            hm[0][(size_t)(par * P + q) * plane + mx_off_host(s, i, A.KB)] = pc[q];
            hm[1][(size_t)(par * P + q) * plane + mx_off_host(i, s, A.KB)] = pc[q];
          }
        }
    for (int inv = 0; inv < 2; inv++) {
      const size_t bytes = hm[inv].size() * 2;
      if (hip_ok(hipMalloc(&A.Wm[P - 2][inv], bytes), err, "hipMalloc(tiled sine matrix)")) return -1;
      if (hip_ok(hipMemcpy(A.Wm[P - 2][inv], hm[inv].data(), bytes, hipMemcpyHostToDevice), err,
                 "hipMemcpy(tiled sine matrix)"))
        return -1;
    }
  }
  return 0;
}

static void free_axis(AxisMats& A) {
  for (int t = 0; t < 4; t++)
    if (A.W[t]) hipFree(A.W[t]);
  for (int P = 0; P < 2; P++)
    for (int inv = 0; inv < 2; inv++)
      if (A.Wm[P][inv]) hipFree(A.Wm[P][inv]);
}

static DstArgs dst_args(const float* in, float* out, const AxisMats& M, int inv, int rows, int inner) {
  DstArgs a;
  a.in = in;
  a.out = out;
  a.W0 = M.W[inv ? 2 : 0];
  a.W1 = M.W[inv ? 3 : 1];
  a.N = M.N;
  a.Nh = M.Nh;
  a.rows = rows;
  a.inner = inner;
  a.Kp = M.Kp;
  return a;
}

// permuted eigenvalues: spectral slot s < N/2 holds mode 2s, slot N/2 + s holds mode 2s+1
static void perm_eig(int N, double dx, double* out) {
  int Nh = N / 2;
  for (int s = 0; s < N; s++) {
    int l = s < Nh ? 2 * s : 2 * (s - Nh) + 1;
    double sn = sin(M_PI * (l + 1) / (2.0 * N));
    out[s] = -4.0 / (dx * dx) * sn * sn;
  }
}

extern "C" int obl_poisson_set_density(obl_poisson* p, const double* rhoC, const double* rhoW) {
  if (!p || !rhoC || !rhoW) return set_global("obl_poisson_set_density: null argument");
  int nz = p->nz;
  double a = 1.0 / (p->dz * p->dz);
  std::vector<double> z64(4 * (size_t)nz);
  std::vector<float4> z32(nz);
  for (int k = 0; k < nz; k++) {
    double lo = k > 0 ? rhoW[k] * a : 0.0, up = k < nz - 1 ? rhoW[k + 1] * a : 0.0;
    z64[4 * k] = lo; z64[4 * k + 1] = up; z64[4 * k + 2] = -(lo + up); z64[4 * k + 3] = rhoC[k];
    z32[k] = make_float4((float)lo, (float)up, (float)(-(lo + up)), (float)rhoC[k]);
  }
  if (hip_ok(hipMemcpyAsync(p->zop64, z64.data(), z64.size() * sizeof(double), hipMemcpyHostToDevice, p->stream), &p->err,
             "setDensity"))
    return -1;
  if (hip_ok(hipMemcpyAsync(p->zop32, z32.data(), z32.size() * sizeof(float4), hipMemcpyHostToDevice, p->stream), &p->err,
             "setDensity"))
    return -1;
  return hip_ok(hipStreamSynchronize(p->stream), &p->err, "setDensity");
}

extern "C" void obl_poisson_destroy(obl_poisson* p) {
  if (!p) return;
  if (p->stream) hipStreamSynchronize(p->stream);
  free_axis(p->ax);
  free_axis(p->ay);
  void* bufs[] = {p->tmp,   p->hat,   p->cp32,  p->lam32, p->cp64,  p->dp64,  p->lam64, p->zop64,
                  p->zop32, p->hrhs,  p->hphi,  p->mx[0], p->mx[1], p->mx[2], p->mx[3], p->modes64};
  for (void* b : bufs)
    if (b) hipFree(b);
  if (p->stream) hipStreamDestroy(p->stream);
  delete p;
}

extern "C" int obl_poisson_create(int nx, int ny, int nz, double dx, double dz, obl_poisson** out) {
  if (!out) return set_global("obl_poisson_create: null output");
  *out = nullptr;
  if (nx < 8 || ny < 8 || nz < 2 || nx % 8 || ny % 8)
    return set_global("PoissonDirect: nx and ny must be multiples of 8 (>= 8) and nz >= 2, got %d x %d x %d", nx, ny, nz);
  if (!(dx > 0) || !(dz > 0)) return set_global("PoissonDirect: dx and dz must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_global("PoissonDirect: no ROCm GPU is available");
  obl_poisson* p = new obl_poisson();
  p->nx = nx; p->ny = ny; p->nz = nz; p->dx = dx; p->dz = dz;
  std::string e;
  size_t cells = (size_t)nx * ny * nz;
  if (hip_ok(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), &e, "hipStreamCreate") ||
      build_axis(p->ax, nx, &e) || build_axis(p->ay, ny, &e) ||
      hip_ok(hipMalloc(&p->tmp, cells * 4), &e, "hipMalloc(tmp)") ||
      hip_ok(hipMalloc(&p->hat, cells * 4), &e, "hipMalloc(hat)") ||
      hip_ok(hipMalloc(&p->cp32, cells * 4), &e, "hipMalloc(cp)") ||
      hip_ok(hipMalloc(&p->cp64, cells * 8), &e, "hipMalloc(cp64)") ||
      hip_ok(hipMalloc(&p->dp64, cells * 8), &e, "hipMalloc(dp64)") ||
      hip_ok(hipMalloc(&p->lam32, (nx + ny) * 4), &e, "hipMalloc(lam)") ||
      hip_ok(hipMalloc(&p->lam64, (nx + ny) * 8), &e, "hipMalloc(lam64)") ||
      hip_ok(hipMalloc(&p->zop32, nz * 16), &e, "hipMalloc(zop)") ||
      hip_ok(hipMalloc(&p->zop64, nz * 32), &e, "hipMalloc(zop64)") ||
      hip_ok(hipMalloc(&p->modes64, (size_t)nx * ny * sizeof(int2)), &e, "hipMalloc(modes64)")) {
    obl_poisson_destroy(p);
    return set_global("PoissonDirect(%d, %d, %d): %s", nx, ny, nz, e.c_str());
  }
  // multi-piece operands: XF rows are tiles of 32 MX_MF rows holding 16 MX_MF (z, y) pairs; YF and YI rows
  // (z, sx); XI rows (z, y)
  p->mxRB[0] = MX_MF * ((nz * (ny / 2) + 16 * MX_MF - 1) / (16 * MX_MF));
  p->mxRB[1] = p->mxRB[2] = mx_rb(nz * nx);
  p->mxRB[3] = mx_rb(nz * ny);
  const int mxKB[4] = {p->ax.KB, p->ay.KB, p->ay.KB, p->ax.KB};
  for (int b = 0; b < 4; b++) {
    const size_t bytes = mx_bytes(3, p->mxRB[b], mxKB[b]);
    if (hip_ok(hipMalloc(&p->mx[b], bytes), &e, "hipMalloc(split operand)") ||
        hip_ok(hipMemset(p->mx[b], 0, bytes), &e, "hipMemset(split operand)")) {
      obl_poisson_destroy(p);
      return set_global("PoissonDirect(%d, %d, %d): %s", nx, ny, nz, e.c_str());
    }
  }
  std::vector<double> l64(nx + ny);
  perm_eig(nx, dx, l64.data());
  perm_eig(ny, dx, l64.data() + nx);
  std::vector<float> l32(nx + ny);
  for (int i = 0; i < nx + ny; i++) l32[i] = (float)l64[i];
  hipMemcpy(p->lam64, l64.data(), l64.size() * 8, hipMemcpyHostToDevice);
  p->lam64h = l64;
  hipMemcpy(p->lam32, l32.data(), l32.size() * 4, hipMemcpyHostToDevice);
  std::vector<double> ones(nz + 1, 1.0);
  if (obl_poisson_set_density(p, ones.data(), ones.data())) {
    std::string m = p->err;
    obl_poisson_destroy(p);
    return set_global("%s", m.c_str());
  }
  *out = p;
  return 0;
}

// The FP64 modes of a Thomas mode and shift bound (mode 2: |lx + ly| < shift_lo, the sum in double as the FP32
// kernel tests it; mode 1: every mode), uploaded when they change (a set-up step, not part of a steady run of
// solves: it waits for the copy).
static int thomas_modes(obl_poisson* p, int thomas, double shift_lo, hipStream_t s) {
  if (thomas == OBL_THOMAS_FP32 || p->nz > 104) {
    p->n64 = 0;
    return 0;
  }
  if (thomas == p->list_mode && (thomas == OBL_THOMAS_FP64 || shift_lo == p->list_lo)) return 0;
  const int nx = p->nx, ny = p->ny;
  p->hmodes.clear();
  for (int sx = 0; sx < nx; sx++)
    for (int sy = 0; sy < ny; sy++)
      if (thomas == OBL_THOMAS_FP64 || fabs(p->lam64h[sx] + p->lam64h[nx + sy]) < shift_lo)
        p->hmodes.push_back(make_int2(sx, sy));
  p->n64 = (int)p->hmodes.size();
  if (p->n64 > 0) {
    if (hip_ok(hipMemcpyAsync(p->modes64, p->hmodes.data(), p->hmodes.size() * sizeof(int2), hipMemcpyHostToDevice, s),
               &p->err, "Thomas FP64 modes") ||
        hip_ok(hipStreamSynchronize(s), &p->err, "Thomas FP64 modes"))
      return -1;
  }
  p->list_mode = thomas;
  p->list_lo = shift_lo;
  return 0;
}

// folded: the x forward pass's operand is already written (launch_div_fold, the model step's divergence)
static int poisson_enc(obl_poisson* p, const float* rhs, float* phi, int gemm, int thomas, double cond,
                       hipStream_t s, bool folded = false) {
  if (gemm != OBL_GEMM_FP32 && gemm != OBL_GEMM_BF16X3 && gemm != OBL_GEMM_BF16X6)
    return p->fail("unknown gemm mode %d", gemm);
  if (thomas < 0 || thomas > 2) return p->fail("unknown Thomas mode %d", thomas);
  int nx = p->nx, ny = p->ny, nz = p->nz;
  if (gemm == OBL_GEMM_BF16X3 || gemm == OBL_GEMM_BF16X6) {
    const int P = gemm == OBL_GEMM_BF16X6 ? 3 : 2;
    const AxisMats &X = p->ax, &Y = p->ay;
    if (!folded) TIMED(p->timer, s, "ps.fold", launch_mx_fold_x(s, P, rhs, p->mx[0], nx, ny, nz, p->mxRB[0], X.KB));
    MxArgs m;
    m.Nx = nx; m.Ny = ny; m.nz = nz;
    // forward x: rows (z, y) pairs; the epilogue folds y into the YF operand
    m.A = p->mx[0]; m.B = X.Wm[P - 2][0]; m.aRB = p->mxRB[0]; m.bRB = X.bRB; m.KB = X.KB;
    m.Mvalid = nz * (ny / 2); m.tilesN = X.tilesN; m.outT = p->mx[1]; m.oRB = p->mxRB[1]; m.oKB = Y.KB; m.outF = nullptr;
    TIMED(p->timer, s, "ps.xf", launch_dst_mx(s, DST_XF, P, X.NF, m));
    // forward y: rows (z, sx) -> spectral array [z][sx][sy]
    m.A = p->mx[1]; m.B = Y.Wm[P - 2][0]; m.aRB = p->mxRB[1]; m.bRB = Y.bRB; m.KB = Y.KB;
    m.Mvalid = nz * nx; m.tilesN = Y.tilesN; m.outT = nullptr; m.outF = p->hat;
    TIMED(p->timer, s, "ps.yf", launch_dst_mx(s, DST_YF, P, Y.NF, m));
    ThomasArgs t;
    t.hat = p->hat; t.cp32 = p->cp32; t.cp64 = p->cp64; t.dp64 = p->dp64;
    t.lam32 = p->lam32; t.lam64 = p->lam64; t.zop32 = p->zop32; t.zop64 = p->zop64;
    t.nx = nx; t.ny = ny; t.nz = nz; t.mode = thomas;
    t.shift_lo = cond > 0 ? 4.0 / (cond * p->dz * p->dz) : 0.0;
    t.layout = 1; t.outT = p->mx[2]; t.P = P; t.oKB = Y.KB;
    if (thomas_modes(p, thomas, t.shift_lo, s)) return -1;
    t.modes64 = p->modes64; t.n64 = p->n64;
    TIMED(p->timer, s, "ps.tri", launch_thomas(s, t));
    // inverse y: rows (z, sx); the epilogue unfolds y into the XI operand, rows (z, y)
    m.A = p->mx[2]; m.B = Y.Wm[P - 2][1]; m.aRB = p->mxRB[2]; m.bRB = Y.bRB; m.KB = Y.KB;
    m.Mvalid = nz * nx; m.tilesN = Y.tilesN; m.outT = p->mx[3]; m.oRB = p->mxRB[3]; m.oKB = X.KB; m.outF = nullptr;
    TIMED(p->timer, s, "ps.yi", launch_dst_mx(s, DST_YI, P, Y.NF, m));
    // inverse x -> phi
    m.A = p->mx[3]; m.B = X.Wm[P - 2][1]; m.aRB = p->mxRB[3]; m.bRB = X.bRB; m.KB = X.KB;
    m.Mvalid = nz * ny; m.tilesN = X.tilesN; m.outT = nullptr; m.outF = phi;
    TIMED(p->timer, s, "ps.xi", launch_dst_mx(s, DST_XI, P, X.NF, m));
    return hip_ok(hipGetLastError(), &p->err, "PoissonDirect.enc launch");
  }
  DstArgs a;
  // forward x: rhs -> tmp
  a = dst_args(rhs, p->tmp, p->ax, 0, ny * nz, nx);
  TIMED(p->timer, s, "ps.xf", launch_dst(s, DST_XF, a));
  // forward y: tmp -> hat
  a = dst_args(p->tmp, p->hat, p->ay, 0, nz, nx);
  TIMED(p->timer, s, "ps.yf", launch_dst(s, DST_YF, a));
  ThomasArgs t;
  t.hat = p->hat; t.cp32 = p->cp32; t.cp64 = p->cp64; t.dp64 = p->dp64;
  t.lam32 = p->lam32; t.lam64 = p->lam64; t.zop32 = p->zop32; t.zop64 = p->zop64;
  t.nx = nx; t.ny = ny; t.nz = nz; t.mode = thomas;
  // FP64 where the estimated condition 4 / (|shift| dz^2) exceeds `cond`
  t.shift_lo = cond > 0 ? 4.0 / (cond * p->dz * p->dz) : 0.0;
  t.layout = 0; t.outT = nullptr; t.P = 0; t.oKB = 0;
  if (thomas_modes(p, thomas, t.shift_lo, s)) return -1;
  t.modes64 = p->modes64; t.n64 = p->n64;
  TIMED(p->timer, s, "ps.tri", launch_thomas(s, t));
  // inverse y: hat -> tmp
  a = dst_args(p->hat, p->tmp, p->ay, 1, nz, nx);
  TIMED(p->timer, s, "ps.yi", launch_dst(s, DST_YI, a));
  // inverse x: tmp -> phi
  a = dst_args(p->tmp, phi, p->ax, 1, ny * nz, nx);
  TIMED(p->timer, s, "ps.xi", launch_dst(s, DST_XI, a));
  return hip_ok(hipGetLastError(), &p->err, "PoissonDirect.enc launch");
}

extern "C" int obl_poisson_solve(obl_poisson* p, const float* rhs_dev, float* phi_dev, int gemm_mode, int thomas_mode,
                                 double fp64_cond, void* stream) {
  if (!p) return set_global("obl_poisson_solve: null handle");
  if (!rhs_dev || !phi_dev) return p->fail("obl_poisson_solve: null buffer");
  hipStream_t s = stream ? (hipStream_t)stream : p->stream;
  return poisson_enc(p, rhs_dev, phi_dev, gemm_mode, thomas_mode, fp64_cond, s);
}

extern "C" int obl_poisson_solve_host(obl_poisson* p, const float* rhs, float* phi, int gemm_mode, int thomas_mode,
                                      double fp64_cond) {
  if (!p) return set_global("obl_poisson_solve_host: null handle");
  size_t bytes = (size_t)p->nx * p->ny * p->nz * 4;
  if (!p->hrhs && (hip_ok(hipMalloc(&p->hrhs, bytes), &p->err, "hipMalloc") || hip_ok(hipMalloc(&p->hphi, bytes), &p->err, "hipMalloc")))
    return -1;
  if (hip_ok(hipMemcpyAsync(p->hrhs, rhs, bytes, hipMemcpyHostToDevice, p->stream), &p->err, "copy rhs")) return -1;
  if (poisson_enc(p, p->hrhs, p->hphi, gemm_mode, thomas_mode, fp64_cond, p->stream)) return -1;
  if (hip_ok(hipMemcpyAsync(phi, p->hphi, bytes, hipMemcpyDeviceToHost, p->stream), &p->err, "copy phi")) return -1;
  return hip_ok(hipStreamSynchronize(p->stream), &p->err, "solve");
}

extern "C" const char* obl_poisson_last_error(const obl_poisson* p) { return p ? p->err.c_str() : g_err.c_str(); }

extern "C" int obl_poisson_profile(obl_poisson* p, int on) {
  if (!p) return set_global("null handle");
  hipStreamSynchronize(p->stream);
  p->timer.reset();
  p->timer.on = on != 0;
  return 0;
}

extern "C" int obl_poisson_profile_read(obl_poisson* p, double* ms5) {
  if (!p) return set_global("null handle");
  p->timer.collect();
  const char* nm[5] = {"ps.xf", "ps.yf", "ps.tri", "ps.yi", "ps.xi"};
  for (int i = 0; i < 5; i++) {
    int id = p->timer.id_of(nm[i]);
    ms5[i] = p->timer.count[id] ? p->timer.total[id] / p->timer.count[id] : 0.0;
  }
  return 0;
}

// ---------------------------------------------------------------- RegionalModel
struct obl_model {
  int nx = 0, ny = 0, nz = 0, pad = 0, nxp = 0, nyp = 0;
  int dev = 0;  // the device the model's fields live on (current at obl_create)
  double dx = 0, dz = 0;
  hipStream_t stream = nullptr;
  float4 *V[2] = {nullptr, nullptr}, *Vt = nullptr;
  float4 *Q1[2] = {nullptr, nullptr}, *Q2[2] = {nullptr, nullptr}, *Q1t = nullptr, *Q2t = nullptr;
  float4 *lsfA = nullptr, *lsfB = nullptr, *ciMap = nullptr, *BS = nullptr;
  float2* aux = nullptr;
  float *phT = nullptr, *phi = nullptr, *rhs = nullptr, *BD = nullptr, *CI = nullptr;
  obl_poisson* solver = nullptr;
  int gemm_mode = OBL_GEMM_BF16X6, thomas_mode = OBL_THOMAS_MIXED;
  double fp64_cond = 64.0;
  // microphysics teacher sampling (surrogate training data)
  float* dump_buf = nullptr;
  unsigned* dump_count = nullptr;
  unsigned dump_cap = 0, dump_salt = 0;
  float dump_pa = 0.f, dump_pc = 0.f;
  bool dump_on = false;
  // microphysics surrogate: attached network, point records, counters
  obl_mlp* net = nullptr;
  float* srec = nullptr;
  unsigned* scount = nullptr;       // [0] points handed this step, [1] fallbacks this step
  unsigned scap = 0;
  unsigned long long* stally = nullptr;  // steps, points, fallbacks, overflow
  // judge: the plain step's outputs, accumulated precipitation, sums
  float4 *jV = nullptr, *jQ1 = nullptr, *jQ2 = nullptr;
  double *jprs = nullptr, *jprp = nullptr, *jsums = nullptr;  // jsums: JUDGE_NSUM + 4 precip totals
  unsigned long long* jmax = nullptr;
  double judged = 0.0;
  int sblocks = 256;  // persistent blocks of the surrogate kernel (one per CU)
  // requestMaps: one device map buffer, two pinned host slots with their copy events
  float4* mapd = nullptr;
  float* maph[2] = {nullptr, nullptr};
  hipEvent_t mapev[2] = {nullptr, nullptr};
  bool mapbusy[2] = {false, false};
  // virtual radars: the sub-grid tornado list shared by every radar of the model
  float4* tor = nullptr;
  int ntor = 0;
  std::vector<obl_radar*> radars;  // owned: freed with the model
  // hurricane mode
  float4* sfc = nullptr;  // sea surface map, nx*ny
  float4* Vp = nullptr;   // the winds the advection carries (prep field)
  float4* vx = nullptr;   // initial vortex table, nz*vx_nr
  int vx_nr = 0;
  std::string err;
  Timer timer;
  int fail(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return -1;
  }
  size_t vcount() const { return (size_t)(nx + 1) * (ny + 1) * (nz + 1); }
  size_t ccount() const { return (size_t)nx * ny * nz; }
  size_t pcount() const { return (size_t)nxp * nyp * nz; }
};

extern "C" int obl_abi_version(void) { return OBL_ABI_VERSION; }
extern "C" const char* obl_last_error(const obl_model* h) { return h ? h->err.c_str() : g_err.c_str(); }
extern "C" const char* obl_global_error(void) { return g_err.c_str(); }

extern "C" void obl_destroy(obl_model* h) {
  if (!h) return;
  if (h->stream) hipStreamSynchronize(h->stream);
  void* bufs[] = {h->V[0], h->V[1], h->Vt, h->Q1[0], h->Q1[1], h->Q2[0], h->Q2[1], h->Q1t, h->Q2t, h->lsfA,
                  h->lsfB, h->ciMap, h->BS, h->aux, h->phT, h->phi, h->rhs, h->BD, h->CI, h->dump_buf, h->dump_count,
                  h->srec, h->scount, h->stally, h->jV, h->jQ1, h->jQ2, h->jprs, h->jprp, h->jsums, h->jmax};
  for (void* b : bufs)
    if (b) hipFree(b);
  while (!h->radars.empty()) obl_radar_destroy(h->radars.back());
  for (void* b : {(void*)h->sfc, (void*)h->Vp, (void*)h->vx})
    if (b) hipFree(b);
  if (h->mapd) hipFree(h->mapd);
  if (h->tor) hipFree(h->tor);
  for (int q = 0; q < 2; q++) {
    if (h->maph[q]) hipHostFree(h->maph[q]);
    if (h->mapev[q]) hipEventDestroy(h->mapev[q]);
  }
  if (h->solver) obl_poisson_destroy(h->solver);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

// device bytes a grid's fields take: V x 3, Q x 6, aux, phT; the padded solve's phi, rhs, tmp, hat and FP32 / FP64
// Thomas scratch; the 2-D maps (for the out-of-memory message, the counterpart of gridBytes, model.js:1125-1128)
static double fields_bytes(int nx, int ny, int nz, int pad) {
  const double v = (double)(nx + 1) * (ny + 1) * (nz + 1), c = (double)nx * ny * nz;
  const double pc = (double)(nx + 2 * pad) * (ny + 2 * pad) * nz;
  return v * 16 * 3 + c * (16 * 6 + 8 + 4) + pc * (4 * 5 + 16) + (double)nx * ny * 48;
}
static double grid_bytes(const obl_model* h) { return fields_bytes(h->nx, h->ny, h->nz, h->pad); }
extern "C" double obl_grid_bytes(const obl_grid* g) { return g ? fields_bytes(g->nx, g->ny, g->nz, g->pad) : 0.0; }

extern "C" int obl_create(const obl_grid* g, obl_model** out) {
  if (!g || !out) return set_global("obl_create: null argument");
  *out = nullptr;
  if (g->nx < 8 || g->ny < 8 || g->nz < 2 || g->nx % 8 || g->ny % 8)
    return set_global("allocate: nx and ny must be multiples of 8 and nz >= 2 (got %d x %d x %d)", g->nx, g->ny, g->nz);
  if (g->pad < 0 || g->pad % 4) return set_global("allocate: the solve padding must be a non-negative multiple of 4");
  // the stencil kernels address fields with 32-bit texel offsets (24-bit factors)
  if ((double)(g->nx + 1) * (g->ny + 1) * (g->nz + 1) >= (double)(1u << 28) || (double)(g->ny + 1) * (g->nz + 1) >= (double)(1u << 24) ||
      (double)(g->nx + 1 + 2 * g->pad) * (g->ny + 1 + 2 * g->pad) >= (double)(1u << 24) ||
      (double)(g->ny + 2 * g->pad) * g->nz >= (double)(1u << 24))
    return set_global("allocate: %d x %d x %d exceeds 2^28 texels per field (4 GiB of float4)", g->nx, g->ny, g->nz);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_global("allocate: no ROCm GPU is available");
  obl_model* h = new obl_model();
  hipGetDevice(&h->dev);
  h->nx = g->nx; h->ny = g->ny; h->nz = g->nz; h->dx = g->dx; h->dz = g->dz; h->pad = g->pad;
  h->nxp = h->nx + 2 * h->pad; h->nyp = h->ny + 2 * h->pad;
  size_t freeb = 0, totalb = 0;
  hipMemGetInfo(&freeb, &totalb);
  double need = grid_bytes(h);
  if (freeb && need > (double)freeb) {
    delete h;
    return set_global("not enough GPU memory for this grid (it needs about %.1f GB, %.1f GB free)", need / 1e9,
                      freeb / 1e9);
  }
  std::string e;
  size_t vb = h->vcount() * 16, cb = h->ccount() * 16, pb = h->pcount() * 4, mb = (size_t)h->nx * h->ny * 16;
  int bad = hip_ok(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), &e, "hipStreamCreate") ||
            hip_ok(hipMalloc(&h->V[0], vb), &e, "hipMalloc(V0)") || hip_ok(hipMalloc(&h->V[1], vb), &e, "hipMalloc(V1)") ||
            hip_ok(hipMalloc(&h->Vt, vb), &e, "hipMalloc(Vt)") ||
            hip_ok(hipMalloc(&h->Q1[0], cb), &e, "hipMalloc(Q1a)") || hip_ok(hipMalloc(&h->Q1[1], cb), &e, "hipMalloc(Q1b)") ||
            hip_ok(hipMalloc(&h->Q2[0], cb), &e, "hipMalloc(Q2a)") || hip_ok(hipMalloc(&h->Q2[1], cb), &e, "hipMalloc(Q2b)") ||
            hip_ok(hipMalloc(&h->Q1t, cb), &e, "hipMalloc(Q1t)") || hip_ok(hipMalloc(&h->Q2t, cb), &e, "hipMalloc(Q2t)") ||
            hip_ok(hipMalloc(&h->aux, h->ccount() * 8), &e, "hipMalloc(aux)") ||
            hip_ok(hipMalloc(&h->phT, h->ccount() * 4), &e, "hipMalloc(phT)") ||
            hip_ok(hipMalloc(&h->lsfA, mb), &e, "hipMalloc(lsfA)") || hip_ok(hipMalloc(&h->lsfB, mb), &e, "hipMalloc(lsfB)") ||
            hip_ok(hipMalloc(&h->ciMap, mb), &e, "hipMalloc(ciMap)") ||
            hip_ok(hipMalloc(&h->BS, (size_t)(h->nz + 1) * 32), &e, "hipMalloc(bs)") ||
            hip_ok(hipMalloc(&h->BD, OBL_MAX_BOUNDS * OBL_BOUND_FLOATS * 4), &e, "hipMalloc(bounds)") ||
            hip_ok(hipMalloc(&h->CI, OBL_MAX_CI * OBL_CI_FLOATS * 4), &e, "hipMalloc(ci)") ||
            hip_ok(hipMalloc(&h->phi, pb), &e, "hipMalloc(phi)") || hip_ok(hipMalloc(&h->rhs, pb), &e, "hipMalloc(rhs)");
  if (bad) {
    obl_destroy(h);
    return set_global("not enough GPU memory for this grid (it needs about %.1f GB): %s", need / 1e9, e.c_str());
  }
  if (obl_poisson_create(h->nxp, h->nyp, h->nz, h->dx, h->dz, &h->solver)) {
    std::string m = g_err;
    obl_destroy(h);
    return set_global("%s", m.c_str());
  }
  // fields start cleared (WebGPU zero-fills new textures and buffers)
  void* bufs[] = {h->V[0], h->V[1], h->Vt, h->Q1[0], h->Q1[1], h->Q2[0], h->Q2[1], h->Q1t, h->Q2t,
                  h->lsfA, h->lsfB, h->ciMap, h->BS, h->aux, h->phT, h->phi, h->rhs, h->BD, h->CI};
  size_t szs[] = {vb, vb, vb, cb, cb, cb, cb, cb, cb, mb, mb, mb, (size_t)(h->nz + 1) * 32, h->ccount() * 8,
                  h->ccount() * 4, pb, pb, OBL_MAX_BOUNDS * OBL_BOUND_FLOATS * 4, OBL_MAX_CI * OBL_CI_FLOATS * 4};
  for (size_t i = 0; i < sizeof(szs) / sizeof(szs[0]); i++) hipMemsetAsync(bufs[i], 0, szs[i], h->stream);
  if (hip_ok(hipStreamSynchronize(h->stream), &e, "clear fields")) {
    obl_destroy(h);
    return set_global("%s", e.c_str());
  }
  *out = h;
  return 0;
}

static int check_u(obl_model* h, const float* u) {
  if (!u) return h->fail("uniform block is null");
  if ((int)u[0] != h->nx || (int)u[1] != h->ny || (int)u[2] != h->nz)
    return h->fail("uniform dims (%g, %g, %g) do not match the grid %d x %d x %d", u[0], u[1], u[2], h->nx, h->ny, h->nz);
  if ((int)u[36] != h->pad || (int)u[37] != h->nxp || (int)u[38] != h->nyp)
    return h->fail("uniform pad block (%g, %g, %g) does not match pad %d (%d x %d)", u[36], u[37], u[38], h->pad, h->nxp,
                   h->nyp);
  if (u[44] > 0.5f && !h->sfc) return h->fail("hurricane mode (hur.x = 1) needs a sea surface: setSurface first");
  if (u[75] > 0.5f && !(u[44] > 0.5f)) return h->fail("midpoint trajectories (cn.w = 1) are a hurricane-mode setting");
  if (!(u[8] > 0.0f)) return h->fail("dt must be positive");
  // the kernels loop over these counts in fixed-size device tables
  if (!(u[30] >= 0.0f && u[30] <= (float)OBL_MAX_BOUNDS))
    return h->fail("uniform cor.z (boundary count %g) must lie in 0..%d", u[30], OBL_MAX_BOUNDS);
  if (!(u[32] >= 0.0f && u[32] <= (float)OBL_MAX_CI))
    return h->fail("uniform ci.x (tower count %g) must lie in 0..%d", u[32], OBL_MAX_CI);
  return 0;
}

static MU to_mu(const float* u) {
  MU m;
  memcpy(&m, u, sizeof(MU));
  return m;
}

extern "C" int obl_set_base_state(obl_model* h, const float* B, int nz) {
  if (!h) return set_global("null handle");
  if (!B || nz != h->nz) return h->fail("base state must hold (nz + 1) * 8 floats for nz = %d", h->nz);
  if (hip_ok(hipMemcpyAsync(h->BS, B, (size_t)(nz + 1) * 32, hipMemcpyHostToDevice, h->stream), &h->err, "setSounding"))
    return -1;
  std::vector<double> rc(nz), rw(nz + 1);
  for (int k = 0; k < nz; k++) rc[k] = B[k * 8 + 2];
  for (int k = 0; k <= nz; k++) rw[k] = B[k * 8 + 3];
  if (hip_ok(hipStreamSynchronize(h->stream), &h->err, "setSounding")) return -1;
  if (obl_poisson_set_density(h->solver, rc.data(), rw.data())) return h->fail("%s", h->solver->err.c_str());
  return 0;
}

extern "C" int obl_set_boundaries(obl_model* h, const float* f, int count) {
  if (!h) return set_global("null handle");
  if (count < 0 || count > OBL_MAX_BOUNDS || (count && !f)) return h->fail("setBoundaries: 0..%d records", OBL_MAX_BOUNDS);
  for (int i = 0; i < count; i++) {
    const float np = f[(size_t)i * OBL_BOUND_FLOATS];
    if (!(np >= 0.0f && np <= (float)OBL_BOUND_POINTS))
      return h->fail("setBoundaries: record %d has %g points (0..%d)", i, np, OBL_BOUND_POINTS);
  }
  std::vector<float> buf((size_t)OBL_MAX_BOUNDS * OBL_BOUND_FLOATS, 0.0f);
  if (count) memcpy(buf.data(), f, (size_t)count * OBL_BOUND_FLOATS * 4);
  if (hip_ok(hipMemcpyAsync(h->BD, buf.data(), buf.size() * 4, hipMemcpyHostToDevice, h->stream), &h->err, "setBoundaries"))
    return -1;
  return hip_ok(hipStreamSynchronize(h->stream), &h->err, "setBoundaries");
}

extern "C" int obl_set_ci(obl_model* h, const float* f, int count) {
  if (!h) return set_global("null handle");
  if (count < 0 || count > OBL_MAX_CI || (count && !f)) return h->fail("setCI: 0..%d towers", OBL_MAX_CI);
  if (count == 0) return 0;
  if (hip_ok(hipMemcpyAsync(h->CI, f, (size_t)count * OBL_CI_FLOATS * 4, hipMemcpyHostToDevice, h->stream), &h->err, "setCI"))
    return -1;
  return hip_ok(hipStreamSynchronize(h->stream), &h->err, "setCI");
}

extern "C" int obl_set_surface(obl_model* h, const float* surf, size_t bytes) {
  if (!h) return set_global("null handle");
  if (!surf) {
    if (h->sfc) {
      hipStreamSynchronize(h->stream);
      hipFree(h->sfc);
      h->sfc = nullptr;
    }
    return 0;
  }
  const size_t b = (size_t)h->nx * h->ny * 16;
  if (bytes < b) return h->fail("setSurface: %zu bytes given, %zu (nx*ny float4) needed", bytes, b);
  if (!h->sfc && hip_ok(hipMalloc(&h->sfc, b), &h->err, "hipMalloc(surface)")) return -1;
  if (hip_ok(hipMemcpyAsync(h->sfc, surf, b, hipMemcpyHostToDevice, h->stream), &h->err, "setSurface")) return -1;
  return hip_ok(hipStreamSynchronize(h->stream), &h->err, "setSurface");
}

extern "C" int obl_set_prep_field(obl_model* h, int on) {
  if (!h) return set_global("null handle");
  if (!on) {
    if (h->Vp) {
      hipStreamSynchronize(h->stream);
      hipFree(h->Vp);
      h->Vp = nullptr;
    }
    return 0;
  }
  if (h->Vp) return 0;
  if (hip_ok(hipMalloc(&h->Vp, h->vcount() * 16), &h->err, "hipMalloc(prep field)")) return -1;
  return hip_ok(hipMemsetAsync(h->Vp, 0, h->vcount() * 16, h->stream), &h->err, "prep field clear");
}

extern "C" int obl_set_vortex(obl_model* h, const float* table, int nr, size_t bytes) {
  if (!h) return set_global("null handle");
  if (h->vx) {
    hipStreamSynchronize(h->stream);
    hipFree(h->vx);
    h->vx = nullptr;
    h->vx_nr = 0;
  }
  if (!table || nr == 0) return 0;
  if (nr < 2 || nr > 65536) return h->fail("setVortex: the table needs 2..65536 radii (got %d)", nr);
  const size_t b = (size_t)h->nz * nr * 16;
  if (bytes < b) return h->fail("setVortex: %zu bytes given, %zu (nz*nr float4) needed", bytes, b);
  if (hip_ok(hipMalloc(&h->vx, b), &h->err, "hipMalloc(vortex)")) return -1;
  h->vx_nr = nr;
  if (hip_ok(hipMemcpyAsync(h->vx, table, b, hipMemcpyHostToDevice, h->stream), &h->err, "setVortex")) return -1;
  return hip_ok(hipStreamSynchronize(h->stream), &h->err, "setVortex");
}

static Dims dims_of(const obl_model* h, const MU& u) {
  Dims d;
  d.nx = h->nx; d.ny = h->ny; d.nz = h->nz;
  d.dx = u.v[U_DIMS].w; d.dz = u.v[U_DZV].x; d.dt = u.v[U_STEP].x;
  // (the kernels make their own in dims(u); the host uses the sizes)
  d.kx = d.dt / d.dx; d.kz = d.dt / d.dz;
  d.rdx = 1.0f / d.dx; d.rdz = 1.0f / d.dz; d.r2dx = 0.5f / d.dx; d.r2dz = 0.5f / d.dz;
  d.rdt = 1.0f / d.dt;
  return d;
}

extern "C" int obl_reset(obl_model* h, const float* uniforms, int seed) {
  if (!h) return set_global("null handle");
  if (check_u(h, uniforms)) return -1;
  MU u = to_mu(uniforms);
  Dims d = dims_of(h, u);
  hipStream_t s = h->stream;
  launch_lsf(s, u, d, h->BS, h->BD, h->lsfA, h->lsfB, h->phT);
  const float4* VX = nullptr;
  if (u.v[U_VTX2].z > 0.5f) {
    if (!h->vx || (int)u.v[U_VTX].w != h->vx_nr)
      return h->fail("reset: vtx2.z = 1 needs the vortex table of %g radii (setVortex; %d loaded)", u.v[U_VTX].w, h->vx_nr);
    VX = h->vx;
  }
  launch_init(s, u, d, h->BS, h->lsfA, h->lsfB, VX, h->V[0], h->Q1[0], h->Q2[0], seed);
  hipMemsetAsync(h->phi, 0, h->pcount() * 4, s);
  hipMemsetAsync(h->rhs, 0, h->pcount() * 4, s);
  hipMemcpyAsync(h->V[1], h->V[0], h->vcount() * 16, hipMemcpyDeviceToDevice, s);
  if (hip_ok(hipGetLastError(), &h->err, "reset")) return -1;
  return hip_ok(hipStreamSynchronize(s), &h->err, "reset");
}

extern "C" int obl_add_bubble(obl_model* h, double x, double y, double z, double rh, double rz, double dth) {
  if (!h) return set_global("null handle");
  if (!(rh > 0) || !(rz > 0)) return h->fail("bubble radii must be positive");
  MU u;
  memset(&u, 0, sizeof u);
  u.v[U_DIMS] = make_float4((float)h->nx, (float)h->ny, (float)h->nz, (float)h->dx);
  u.v[U_DZV].x = (float)h->dz;
  Dims d = dims_of(h, u);
  launch_bubble(h->stream, u, d, h->V[0], h->V[1], x, y, z, rh, rz, dth);
  return hip_ok(hipGetLastError(), &h->err, "bubble");
}

// judge buffers, allocated at the first judged step
static int surr_judge_alloc(obl_model* h) {
  if (h->jV) return 0;
  size_t cols = (size_t)h->nx * h->ny;
  if (hip_ok(hipMalloc(&h->jV, h->vcount() * 16), &h->err, "hipMalloc(judge V)") ||
      hip_ok(hipMalloc(&h->jQ1, h->ccount() * 16), &h->err, "hipMalloc(judge Q1)") ||
      hip_ok(hipMalloc(&h->jQ2, h->ccount() * 16), &h->err, "hipMalloc(judge Q2)") ||
      hip_ok(hipMalloc(&h->jprs, cols * 8), &h->err, "hipMalloc(judge precip)") ||
      hip_ok(hipMalloc(&h->jprp, cols * 8), &h->err, "hipMalloc(judge precip)") ||
      hip_ok(hipMemsetAsync(h->jprs, 0, cols * 8, h->stream), &h->err, "clear") ||
      hip_ok(hipMemsetAsync(h->jprp, 0, cols * 8, h->stream), &h->err, "clear"))
    return -1;
  return 0;
}

extern "C" int obl_set_surrogate(obl_model* h, obl_mlp* net, unsigned capacity) {
  if (!h) return set_global("null handle");
  if (hip_ok(hipStreamSynchronize(h->stream), &h->err, "sync")) return -1;
  void* old[] = {h->srec, h->scount, h->stally, h->jV, h->jQ1, h->jQ2, h->jprs, h->jprp, h->jsums, h->jmax};
  for (void* b : old)
    if (b) hipFree(b);
  h->srec = nullptr; h->scount = nullptr; h->stally = nullptr;
  h->jV = h->jQ1 = h->jQ2 = nullptr;
  h->jprs = h->jprp = h->jsums = nullptr;
  h->jmax = nullptr;
  h->judged = 0.0;
  h->net = nullptr;
  if (!net) return 0;
  size_t cap = capacity ? (size_t)capacity : h->ccount() / 2;
  if (cap > 0xffffffffull) cap = 0xffffffffull;
  if (hip_ok(hipMalloc(&h->srec, cap * SURR_REC * 4), &h->err, "hipMalloc(surrogate records)") ||
      hip_ok(hipMalloc(&h->scount, 2 * sizeof(unsigned)), &h->err, "hipMalloc(surrogate counters)") ||
      hip_ok(hipMalloc(&h->stally, 4 * sizeof(unsigned long long)), &h->err, "hipMalloc(surrogate tally)") ||
      hip_ok(hipMalloc(&h->jsums, (JUDGE_NSUM + 4) * sizeof(double)), &h->err, "hipMalloc(judge sums)") ||
      hip_ok(hipMalloc(&h->jmax, JUDGE_NMAX * sizeof(unsigned long long)), &h->err, "hipMalloc(judge max)") ||
      hip_ok(hipMemset(h->scount, 0, 2 * sizeof(unsigned)), &h->err, "clear") ||
      hip_ok(hipMemset(h->stally, 0, 4 * sizeof(unsigned long long)), &h->err, "clear") ||
      hip_ok(hipMemset(h->jsums, 0, (JUDGE_NSUM + 4) * sizeof(double)), &h->err, "clear") ||
      hip_ok(hipMemset(h->jmax, 0, JUDGE_NMAX * sizeof(unsigned long long)), &h->err, "clear")) {
    obl_set_surrogate(h, nullptr, 0);
    return -1;
  }
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      cus > 0)
    h->sblocks = cus;
  h->scap = (unsigned)cap;
  h->net = net;
  return 0;
}

extern "C" int obl_surrogate_stats(obl_model* h, obl_surr_stats* out, int reset) {
  if (!h) return set_global("null handle");
  if (!out) return h->fail("null output");
  memset(out, 0, sizeof *out);
  if (!h->net) return h->fail("no surrogate attached");
  hipStream_t s = h->stream;
  unsigned long long tally[4] = {0, 0, 0, 0}, mx[JUDGE_NMAX] = {0, 0, 0};
  double sums[JUDGE_NSUM + 4];
  memset(sums, 0, sizeof sums);
  if (h->jprs) {
    hipMemsetAsync(h->jsums + JUDGE_NSUM, 0, 4 * sizeof(double), s);
    launch_judge_precip(s, h->jprs, h->jprp, (size_t)h->nx * h->ny, h->jsums + JUDGE_NSUM);
  }
  if (hip_ok(hipMemcpyAsync(tally, h->stally, sizeof tally, hipMemcpyDeviceToHost, s), &h->err, "stats") ||
      hip_ok(hipMemcpyAsync(sums, h->jsums, sizeof sums, hipMemcpyDeviceToHost, s), &h->err, "stats") ||
      hip_ok(hipMemcpyAsync(mx, h->jmax, sizeof mx, hipMemcpyDeviceToHost, s), &h->err, "stats") ||
      hip_ok(hipStreamSynchronize(s), &h->err, "stats"))
    return -1;
  out->steps = (double)tally[0];
  out->points = (double)tally[1];
  out->fallbacks = (double)tally[2];
  out->overflow = (double)tally[3];
  out->judged = h->judged;
  for (int c = 0; c < 8; c++) {
    out->d2[c] = sums[c];
    out->r2[c] = sums[8 + c];
  }
  out->cc_d2 = sums[16];
  out->cc_r2 = sums[17];
  out->dbz_d2 = sums[18];
  out->dbz_cols = sums[19];
  double m[JUDGE_NMAX];
  memcpy(m, mx, sizeof m);
  out->cc_maxdiff = m[0];
  out->cc_max = m[1];
  out->dbz_maxdiff = m[2];
  out->pr_d2 = sums[JUDGE_NSUM];
  out->pr_r2 = sums[JUDGE_NSUM + 1];
  out->pr_sum_s = sums[JUDGE_NSUM + 2];
  out->pr_sum_p = sums[JUDGE_NSUM + 3];
  if (reset) {
    size_t cols = (size_t)h->nx * h->ny;
    hipMemsetAsync(h->stally, 0, 4 * sizeof(unsigned long long), s);
    hipMemsetAsync(h->jsums, 0, (JUDGE_NSUM + 4) * sizeof(double), s);
    hipMemsetAsync(h->jmax, 0, JUDGE_NMAX * sizeof(unsigned long long), s);
    if (h->jprs) {
      hipMemsetAsync(h->jprs, 0, cols * 8, s);
      hipMemsetAsync(h->jprp, 0, cols * 8, s);
    }
    h->judged = 0.0;
    if (hip_ok(hipStreamSynchronize(s), &h->err, "stats reset")) return -1;
  }
  return 0;
}

// ---------------------------------------------------------------- requestMaps
extern "C" int obl_maps_request(obl_model* h, const float* uniforms, int parity, int* slot) {
  if (!h) return set_global("null handle");
  if (!slot) return h->fail("requestMaps: null slot");
  *slot = -1;
  if (check_u(h, uniforms)) return -1;
  if (parity != 0 && parity != 1) return h->fail("parity must be 0 or 1");
  const int q = h->mapbusy[0] ? (h->mapbusy[1] ? -1 : 1) : 0;
  if (q < 0) return 0;  // both slots await collection (the reference returns null)
  const size_t bytes = (size_t)h->nx * h->ny * 64;
  if (!h->mapd && hip_ok(hipMalloc(&h->mapd, bytes), &h->err, "hipMalloc(maps)")) return -1;
  if (!h->maph[q] && hip_ok(hipHostMalloc((void**)&h->maph[q], bytes, hipHostMallocDefault), &h->err, "hipHostMalloc(maps)"))
    return -1;
  if (!h->mapev[q] && hip_ok(hipEventCreateWithFlags(&h->mapev[q], hipEventDisableTiming), &h->err, "hipEventCreate(maps)"))
    return -1;
  MU u = to_mu(uniforms);
  Dims d = dims_of(h, u);
  hipStream_t s = h->stream;
  TIMED(h->timer, s, "m.map", launch_maps(s, u, d, h->BS, h->V[parity], h->Q1[0], h->Q2[0], h->phi, h->sfc, h->mapd));
  if (hip_ok(hipMemcpyAsync(h->maph[q], h->mapd, bytes, hipMemcpyDeviceToHost, s), &h->err, "requestMaps copy") ||
      hip_ok(hipEventRecord(h->mapev[q], s), &h->err, "requestMaps event"))
    return -1;
  h->mapbusy[q] = true;
  *slot = q;
  return 0;
}

extern "C" int obl_maps_collect(obl_model* h, int slot, float* dst, size_t dst_floats) {
  if (!h) return set_global("null handle");
  if ((slot != 0 && slot != 1) || !h->mapbusy[slot]) return h->fail("collectMaps: slot %d holds no requested maps", slot);
  if (!dst || dst_floats < (size_t)h->nx * h->ny * 16)
    return h->fail("collectMaps: the destination holds %zu of %zu floats", dst_floats, (size_t)h->nx * h->ny * 16);
  if (hip_ok(hipEventSynchronize(h->mapev[slot]), &h->err, "collectMaps")) return -1;
  memcpy(dst, h->maph[slot], (size_t)h->nx * h->ny * 64);
  h->mapbusy[slot] = false;
  return 0;
}

// ---------------------------------------------------------------- virtual radars
struct obl_radar {
  obl_model* m = nullptr;
  uint32_t* out = nullptr;
  size_t bytes = 0;
  float* kdp = nullptr;
  size_t kdp_floats = 0;
  uint8_t* host[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool busy[2] = {false, false};
};

extern "C" int obl_radar_create(obl_model* h, size_t bytes, size_t kdp_floats, obl_radar** out) {
  if (!h || !out) return set_global("obl_radar_create: null argument");
  *out = nullptr;
  if (bytes == 0 || bytes % 4 || kdp_floats == 0) return h->fail("radar: volume bytes must be a positive multiple of 4");
  obl_radar* r = new obl_radar();
  r->m = h;
  r->bytes = bytes;
  r->kdp_floats = kdp_floats;
  if (hip_ok(hipMalloc(&r->out, bytes), &h->err, "hipMalloc(radar volume)") ||
      hip_ok(hipMemsetAsync(r->out, 0, bytes, h->stream), &h->err, "radar clear") ||
      hip_ok(hipMalloc(&r->kdp, kdp_floats * 4), &h->err, "hipMalloc(radar kdp)")) {
    obl_radar_destroy(r);
    return -1;
  }
  for (int q = 0; q < 2; q++)
    if (hip_ok(hipHostMalloc((void**)&r->host[q], bytes, hipHostMallocDefault), &h->err, "hipHostMalloc(radar)") ||
        hip_ok(hipEventCreateWithFlags(&r->ev[q], hipEventDisableTiming), &h->err, "hipEventCreate(radar)")) {
      obl_radar_destroy(r);
      return -1;
    }
  if (!h->tor) {
    if (hip_ok(hipMalloc(&h->tor, RADAR_MAX_TOR * 2 * sizeof(float4)), &h->err, "hipMalloc(tornadoes)") ||
        hip_ok(hipMemsetAsync(h->tor, 0, RADAR_MAX_TOR * 2 * sizeof(float4), h->stream), &h->err, "tornado clear")) {
      obl_radar_destroy(r);
      return -1;
    }
  }
  h->radars.push_back(r);
  *out = r;
  return 0;
}

extern "C" void obl_radar_destroy(obl_radar* r) {
  if (!r) return;
  if (r->m) {
    if (r->m->stream) hipStreamSynchronize(r->m->stream);
    auto& v = r->m->radars;
    for (size_t i = 0; i < v.size(); i++)
      if (v[i] == r) {
        v.erase(v.begin() + (long)i);
        break;
      }
  }
  if (r->out) hipFree(r->out);
  if (r->kdp) hipFree(r->kdp);
  for (int q = 0; q < 2; q++) {
    if (r->host[q]) hipHostFree(r->host[q]);
    if (r->ev[q]) hipEventDestroy(r->ev[q]);
  }
  delete r;
}

extern "C" int obl_radar_set_tornadoes(obl_model* h, const float* f, int n) {
  if (!h) return set_global("null handle");
  if (n < 0 || n > RADAR_MAX_TOR || (n > 0 && !f)) return h->fail("setTornadoes: 0..%d tornadoes of 8 floats", RADAR_MAX_TOR);
  if (!h->tor && hip_ok(hipMalloc(&h->tor, RADAR_MAX_TOR * 2 * sizeof(float4)), &h->err, "hipMalloc(tornadoes)")) return -1;
  if (n > 0 && hip_ok(hipMemcpyAsync(h->tor, f, (size_t)n * 8 * 4, hipMemcpyHostToDevice, h->stream), &h->err, "setTornadoes"))
    return -1;
  h->ntor = n;
  return hip_ok(hipStreamSynchronize(h->stream), &h->err, "setTornadoes");
}

// The cut's uniform block checked against the model and against a volume of `words` 32-bit words and a KDP
// scratch of `kdp_floats` floats (the counts that the kernels index with).
static int check_cut(obl_model* h, const RadarU& R, size_t words, size_t kdp_floats, size_t volume_bytes) {
  if ((int)R.mdl.x != h->nx || (int)R.mdl.y != h->ny || (int)R.mdl.z != h->nz)
    return h->fail("radar cut: mdl (%g, %g, %g) is not the model's %d x %d x %d", R.mdl.x, R.mdl.y, R.mdl.z, h->nx, h->ny, h->nz);
  const long nrad = (long)R.cut.w, ngr = (long)R.gates.x, ngd = (long)R.gates.y;
  const unsigned mask = (unsigned)R.offB.w;
  if (nrad < 1 || ngr < 4 || ngr % 4 || ngd < 0 || ngd % 4 || ngd > ngr || R.phys.w < 0 || (int)R.phys.w > h->ntor)
    return h->fail("radar cut: %ld radials, %ld / %ld gates, %g tornadoes (%d set)", nrad, ngr, ngd, R.phys.w, h->ntor);
  if ((size_t)(nrad * ngd) > kdp_floats)
    return h->fail("radar cut: %ld radials x %ld Doppler gates exceed the %zu-float KDP scratch", nrad, ngd, kdp_floats);
  auto fits = [&](float off, long n) { return off >= 0 && (size_t)off + (size_t)n <= words; };
  bool ok = fits(R.offA.x, nrad * ngr / 4);
  if (mask & 2u) ok = ok && fits(R.offA.y, nrad * ngd / 4);
  if (mask & 4u) ok = ok && fits(R.offA.z, nrad * ngd / 4);
  if (mask & 8u) ok = ok && fits(R.offA.w, nrad * ngd / 4);
  if (mask & 16u) ok = ok && fits(R.offB.x, nrad * ngd / 4);
  if (mask & 32u) ok = ok && fits(R.offB.y, nrad * ngd / 4);
  if (!ok) return h->fail("radar cut: the cut's moments do not fit the %zu-byte volume", volume_bytes);
  return 0;
}

extern "C" int obl_radar_cut(obl_radar* r, const float* ru, int parity) {
  if (!r) return set_global("null radar");
  obl_model* h = r->m;
  if (!ru) return h->fail("radar cut: null uniform block");
  if (parity != 0 && parity != 1) return h->fail("parity must be 0 or 1");
  RadarU R;
  memcpy(&R, ru, sizeof(R));
  // the block must describe this model and stay inside this radar's buffers
  if (check_cut(h, R, r->bytes / 4, r->kdp_floats, r->bytes)) return -1;
  RadarArgs a{h->BS, h->V[parity], h->Q1[0], h->Q2[0], h->aux, h->tor, r->out, r->kdp};
  TIMED(h->timer, h->stream, "r.cut", launch_radar_cut(h->stream, R, a));
  return hip_ok(hipGetLastError(), &h->err, "radar cut launch");
}

// The allocation-free radar cut: its KDP scratch (nrad x Doppler gates floats between the cut kernel and the PHIDP
// kernel, fully written by the first before the second reads it) is this file-scope device array, of which each
// GPU holds its own copy until the process ends. Host code never passes the array's name to a kernel: its device
// address is looked up once per device (hipGetSymbolAddress, with that device current) and kept here.
static __device__ __align__(256) float scratch[OBL_RADAR_SCRATCH_FLOATS];
static float* g_scratch_dev[OBL_MAX_DEVICES] = {nullptr};
static std::mutex g_scratch_mu;

static float* radar_scratch(obl_model* h) {
  if (h->dev < 0 || h->dev >= OBL_MAX_DEVICES) {
    h->fail("radar cut: device %d is beyond the %d the library serves", h->dev, OBL_MAX_DEVICES);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[h->dev]) {
    int cur = 0;
    hipGetDevice(&cur);
    if (cur != h->dev && hip_ok(hipSetDevice(h->dev), &h->err, "radar scratch: hipSetDevice")) return nullptr;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != h->dev) hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      h->fail("radar scratch: hipGetSymbolAddress: %s", hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_dev[h->dev] = (float*)p;
  }
  return g_scratch_dev[h->dev];
}

extern "C" int obl_radar_cut_into(obl_model* h, const float* ru, int parity, uint32_t* out_dev, size_t out_bytes) {
  if (!h) return set_global("null handle");
  if (!ru || !out_dev) return h->fail("radar cut: null uniform block or volume");
  if (parity != 0 && parity != 1) return h->fail("parity must be 0 or 1");
  if (out_bytes % 4) return h->fail("radar cut: the volume must be whole 32-bit words");
  RadarU R;
  memcpy(&R, ru, sizeof(R));
  if (check_cut(h, R, out_bytes / 4, OBL_RADAR_SCRATCH_FLOATS, out_bytes)) return -1;
  float* kdp = radar_scratch(h);
  if (!kdp) return -1;
  if ((int)R.phys.w > 0 && !h->tor) return h->fail("radar cut: tornadoes without a tornado table (setTornadoes)");
  RadarArgs a{h->BS, h->V[parity], h->Q1[0], h->Q2[0], h->aux, h->tor, out_dev, kdp};
  launch_radar_cut(h->stream, R, a);
  return hip_ok(hipGetLastError(), &h->err, "radar cut launch");
}

extern "C" int obl_radar_geometry(obl_model* h, const float* ru, int n, const int* rg, float* out) {
  if (!h) return set_global("null handle");
  if (!ru || n < 0 || (n > 0 && (!rg || !out))) return h->fail("radar geometry: null uniform block, gates or output");
  RadarU R;
  memcpy(&R, ru, sizeof(R));
  const long nrad = (long)R.cut.w, ngr = (long)R.gates.x;
  for (int q = 0; q < n; q++)
    if (rg[2 * q] < 0 || rg[2 * q] >= nrad || rg[2 * q + 1] < 0 || rg[2 * q + 1] >= ngr)
      return h->fail("radar geometry: gate %d (%d, %d) is outside the cut's %ld x %ld", q, rg[2 * q], rg[2 * q + 1],
                     nrad, ngr);
  if (n == 0) return 0;
  int* drg = nullptr;
  float* dout = nullptr;
  int rc = 0;
  if (hip_ok(hipMalloc(&drg, (size_t)n * 2 * sizeof(int)), &h->err, "radar geometry: hipMalloc") ||
      hip_ok(hipMalloc(&dout, (size_t)n * 7 * sizeof(float)), &h->err, "radar geometry: hipMalloc") ||
      hip_ok(hipMemcpyAsync(drg, rg, (size_t)n * 2 * sizeof(int), hipMemcpyHostToDevice, h->stream), &h->err,
             "radar geometry: upload")) {
    rc = -1;
  } else {
    launch_radar_geom(h->stream, R, n, drg, dout);
    rc = hip_ok(hipGetLastError(), &h->err, "radar geometry launch") ||
                 hip_ok(hipMemcpyAsync(out, dout, (size_t)n * 7 * sizeof(float), hipMemcpyDeviceToHost, h->stream),
                        &h->err, "radar geometry: download") ||
                 hip_ok(hipStreamSynchronize(h->stream), &h->err, "radar geometry: sync")
             ? -1
             : 0;
  }
  if (drg) hipFree(drg);
  if (dout) hipFree(dout);
  return rc;
}

extern "C" int obl_radar_readback(obl_radar* r, int* slot) {
  if (!r) return set_global("null radar");
  obl_model* h = r->m;
  if (!slot) return h->fail("radar readback: null slot");
  *slot = r->busy[0] ? (r->busy[1] ? -1 : 1) : 0;
  if (*slot < 0) return 0;
  const int q = *slot;
  if (hip_ok(hipMemcpyAsync(r->host[q], r->out, r->bytes, hipMemcpyDeviceToHost, h->stream), &h->err, "radar readback") ||
      hip_ok(hipEventRecord(r->ev[q], h->stream), &h->err, "radar readback event"))
    return -1;
  r->busy[q] = true;
  return 0;
}

extern "C" int obl_radar_collect(obl_radar* r, int slot, void* dst, size_t dst_bytes) {
  if (!r) return set_global("null radar");
  obl_model* h = r->m;
  if ((slot != 0 && slot != 1) || !r->busy[slot]) return h->fail("radar collect: slot %d holds no volume", slot);
  if (!dst || dst_bytes < r->bytes) return h->fail("radar collect: the destination holds %zu of %zu bytes", dst_bytes, r->bytes);
  if (hip_ok(hipEventSynchronize(r->ev[slot]), &h->err, "radar collect")) return -1;
  memcpy(dst, r->host[slot], r->bytes);
  r->busy[slot] = false;
  return 0;
}

// The hurricane-mode step (model.js:1442-1476 with this.hur): the 2-dx filter every 4th step (state ->
// the spare copies -> back), PREP while the prep field exists, then the step with the hurricane kernels.
// The surrogate, its judge and teacher sampling are outbreak-mode tools and are refused here.
// The divergence before the solve: with the multi-piece BF16 transforms (and a padded grid the fused kernel
// takes), one kernel also writes the x forward pass's operand, and the solve skips its fold; returns whether it did.
static bool step_div(obl_model* h, const MU& u, const Dims& d, const float4* Vc, Timer& T, hipStream_t s) {
  obl_poisson* p = h->solver;
  if (h->gemm_mode != OBL_GEMM_FP32 && div_fold_fits(p->nx, p->ny)) {
    const int P = h->gemm_mode == OBL_GEMM_BF16X6 ? 3 : 2;
    TIMED(T, s, "m.divf", launch_div_fold(s, u, d, h->BS, Vc, h->rhs, P, p->mx[0], p->mxRB[0], p->ax.KB));
    return true;
  }
  TIMED(T, s, "m.div", launch_div(s, u, d, h->BS, Vc, h->rhs));
  return false;
}

static int step_hurricane(obl_model* h, const MU& u, const Dims& d, int parity, unsigned flags) {
  hipStream_t s = h->stream;
  Timer& T = h->timer;
  float4 *Vc = h->V[parity], *Vn = h->V[1 - parity];
  if (h->net || h->dump_on || (flags & OBL_STEP_JUDGE))
    return h->fail("hurricane mode runs the plain microphysics: detach the surrogate and teacher sampling");
  if (flags & OBL_STEP_FILT) {
    TIMED(T, s, "m.filt", launch_filt(s, u, d, Vc, h->Q1[0], h->Q2[0], Vn, h->Q1[1], h->Q2[1]));
    if (hip_ok(hipMemcpyAsync(Vc, Vn, h->vcount() * 16, hipMemcpyDeviceToDevice, s), &h->err, "filter copy") ||
        hip_ok(hipMemcpyAsync(h->Q1[0], h->Q1[1], h->ccount() * 16, hipMemcpyDeviceToDevice, s), &h->err, "filter copy") ||
        hip_ok(hipMemcpyAsync(h->Q2[0], h->Q2[1], h->ccount() * 16, hipMemcpyDeviceToDevice, s), &h->err, "filter copy"))
      return -1;
  }
  const float4* Vval = Vc;
  if (h->Vp) {
    TIMED(T, s, "m.prep", launch_prep(s, u, d, Vc, h->phi, h->Vp));
    Vval = h->Vp;
  }
  if (flags & OBL_STEP_LSF) TIMED(T, s, "m.lsf", launch_lsf(s, u, d, h->BS, h->BD, h->lsfA, h->lsfB, h->phT));
  TIMED(T, s, "m.ci", launch_ci(s, u, d, h->CI, h->ciMap));
  TIMED(T, s, "m.nut", launch_nut(s, u, d, h->BS, Vc, h->Q1[0], h->sfc, h->aux));
  TIMED(T, s, "m.advV1", launch_advv1_hur(s, u, d, Vc, Vval, h->Vt));
  TIMED(T, s, "m.advV2", launch_advv2_hur(s, u, d, Vc, Vval, h->Vt, Vn));
  TIMED(T, s, "m.advQ1", launch_advq1_hur(s, u, d, h->BS, Vc, h->Q1[0], h->Q2[0], h->Q1t, h->Q2t));
  TIMED(T, s, "m.advQ2", launch_advq2_hur(s, u, d, h->BS, Vc, h->Q1[0], h->Q2[0], h->Q1t, h->Q2t, h->Q1[1], h->Q2[1]));
  HurArgs hu = {h->sfc, h->phi, Vn};
  TIMED(T, s, "m.phys", launch_phys_hur(s, u, d, h->BS, Vn, h->Q1[1], h->Q2[1], h->aux, h->lsfA, h->lsfB, h->phT,
                                        h->ciMap, Vc, h->Q1[0], h->Q2[0], hu));
  const bool folded = step_div(h, u, d, Vc, T, s);
  h->solver->timer.on = T.on;
  if (poisson_enc(h->solver, h->rhs, h->phi, h->gemm_mode, h->thomas_mode, h->fp64_cond, s, folded))
    return h->fail("%s", h->solver->err.c_str());
  TIMED(T, s, "m.proj", launch_proj(s, u, d, Vc, h->phi, Vn));
  return hip_ok(hipGetLastError(), &h->err, "encStep");
}

extern "C" int obl_step(obl_model* h, const float* uniforms, int parity, unsigned flags) {
  if (!h) return set_global("null handle");
  if (check_u(h, uniforms)) return -1;
  if (parity != 0 && parity != 1) return h->fail("parity must be 0 or 1");
  MU u = to_mu(uniforms);
  Dims d = dims_of(h, u);
  hipStream_t s = h->stream;
  float4 *Vc = h->V[parity], *Vn = h->V[1 - parity];
  Timer& T = h->timer;
  if (hur_on(u)) return step_hurricane(h, u, d, parity, flags);
  if (flags & OBL_STEP_FILT) return h->fail("the 2-dx filter (OBL_STEP_FILT) is a hurricane-mode step");
  if (flags & OBL_STEP_LSF) TIMED(T, s, "m.lsf", launch_lsf(s, u, d, h->BS, h->BD, h->lsfA, h->lsfB, h->phT));
  TIMED(T, s, "m.ci", launch_ci(s, u, d, h->CI, h->ciMap));
  // nut and the wind advection (into Vn, aux), and the moisture advection (into Q1[1], Q2[1]), in one launch
  TIMED(T, s, "m.adv", launch_adv(s, u, d, h->BS, Vc, h->Q1[0], h->Q2[0], Vn, h->aux, h->Q1[1], h->Q2[1]));
  if (h->net) {
    MlpView mv;
    mlp_view(h->net, &mv);
    if (flags & OBL_STEP_JUDGE) {
      // the plain step from the same inputs, for the judge (not part of the surrogate step)
      if (surr_judge_alloc(h)) return -1;
      launch_phys(s, u, d, h->BS, Vn, h->Q1[1], h->Q2[1], h->aux, h->lsfA, h->lsfB, h->phT, h->ciMap, h->jV, h->jQ1,
                  h->jQ2);
    }
    SurrPhys sp;
    sp.rec = h->srec;
    sp.count = h->scount;
    sp.cap = h->scap;
    for (int i = 0; i < 14; i++) {
      sp.lo[i] = mv.lo[i];
      sp.hi[i] = mv.hi[i];
    }
    sp.qmin = mv.hand_min;
    SurrApply sa;
    sa.rec = h->srec;
    sa.count = h->scount;
    sa.cap = h->scap;
    sa.frag = mv.frag;
    sa.bias = mv.bias;
    for (int j = 0; j < 9; j++) {
      sa.out_scale[j] = mv.out_scale[j];
      sa.dead[j] = mv.dead_frac * mv.out_scale[j];
    }
    sa.Vout = Vc;
    sa.Q1o = h->Q1[0];
    sa.Q2o = h->Q2[0];
    sa.nfallback = h->scount + 1;
    TIMED(T, s, "m.phys", launch_phys_surr(s, u, d, h->BS, Vn, h->Q1[1], h->Q2[1], h->aux, h->lsfA, h->lsfB, h->phT,
                                           h->ciMap, Vc, h->Q1[0], h->Q2[0], sp));
    TIMED(T, s, "m.surr", launch_surr_apply(s, u, sa, h->sblocks));
    if (flags & OBL_STEP_JUDGE) {
      JudgeArgs ja = {h->BS, Vc, h->Q1[0], h->Q2[0], h->jV, h->jQ1, h->jQ2, h->jprs, h->jprp, h->jsums, h->jmax};
      launch_judge_cols(s, u, d, ja);
      h->judged += 1.0;
    }
    launch_surr_tally(s, h->scount, h->scap, h->stally);
  } else if (h->dump_on) {
    MicroDump md = {h->dump_buf, h->dump_count, h->dump_cap, 0x9e3779b9u * (++h->dump_salt), h->dump_pa, h->dump_pc};
    TIMED(T, s, "m.phys", launch_phys_dump(s, u, d, h->BS, Vn, h->Q1[1], h->Q2[1], h->aux, h->lsfA, h->lsfB, h->phT,
                                           h->ciMap, Vc, h->Q1[0], h->Q2[0], md));
  } else {
    TIMED(T, s, "m.phys", launch_phys(s, u, d, h->BS, Vn, h->Q1[1], h->Q2[1], h->aux, h->lsfA, h->lsfB, h->phT, h->ciMap,
                                      Vc, h->Q1[0], h->Q2[0]));
  }
  const bool folded = step_div(h, u, d, Vc, T, s);
  h->solver->timer.on = T.on;
  if (poisson_enc(h->solver, h->rhs, h->phi, h->gemm_mode, h->thomas_mode, h->fp64_cond, s, folded))
    return h->fail("%s", h->solver->err.c_str());
  TIMED(T, s, "m.proj", launch_proj(s, u, d, Vc, h->phi, Vn));
  return hip_ok(hipGetLastError(), &h->err, "encStep");
}

static int field_info(obl_model* h, int field, int copy, void** ptr, size_t* bytes) {
  size_t vb = h->vcount() * 16, cb = h->ccount() * 16, pb = h->pcount() * 4, mb = (size_t)h->nx * h->ny * 16;
  int c2 = copy == 1 ? 1 : 0;
  if ((field == OBL_FIELD_V || field == OBL_FIELD_Q1 || field == OBL_FIELD_Q2) && copy != 0 && copy != 1)
    return h->fail("copy must be 0 or 1 for field %d", field);
  switch (field) {
    case OBL_FIELD_V: *ptr = h->V[c2]; *bytes = vb; return 0;
    case OBL_FIELD_Q1: *ptr = h->Q1[c2]; *bytes = cb; return 0;
    case OBL_FIELD_Q2: *ptr = h->Q2[c2]; *bytes = cb; return 0;
    case OBL_FIELD_VT: *ptr = h->Vt; *bytes = vb; return 0;
    case OBL_FIELD_Q1T: *ptr = h->Q1t; *bytes = cb; return 0;
    case OBL_FIELD_Q2T: *ptr = h->Q2t; *bytes = cb; return 0;
    case OBL_FIELD_AUX: *ptr = h->aux; *bytes = h->ccount() * 8; return 0;
    case OBL_FIELD_PHT: *ptr = h->phT; *bytes = h->ccount() * 4; return 0;
    case OBL_FIELD_LSFA: *ptr = h->lsfA; *bytes = mb; return 0;
    case OBL_FIELD_LSFB: *ptr = h->lsfB; *bytes = mb; return 0;
    case OBL_FIELD_CIMAP: *ptr = h->ciMap; *bytes = mb; return 0;
    case OBL_FIELD_PHI: *ptr = h->phi; *bytes = pb; return 0;
    case OBL_FIELD_RHS: *ptr = h->rhs; *bytes = pb; return 0;
    case OBL_FIELD_BS: *ptr = h->BS; *bytes = (size_t)(h->nz + 1) * 32; return 0;
    case OBL_FIELD_SFC:
      if (!h->sfc) return h->fail("no sea surface (setSurface)");
      *ptr = h->sfc; *bytes = mb; return 0;
    case OBL_FIELD_VP:
      if (!h->Vp) return h->fail("no prep field (setPrepField)");
      *ptr = h->Vp; *bytes = vb; return 0;
  }
  return h->fail("unknown field id %d", field);
}

extern "C" int obl_field_ptr(obl_model* h, int field, int copy, void** ptr, size_t* bytes) {
  if (!h) return set_global("null handle");
  if (!ptr || !bytes) return h->fail("null output");
  return field_info(h, field, copy, ptr, bytes);
}

extern "C" int obl_upload(obl_model* h, int field, int copy, const void* src, size_t bytes, size_t offset) {
  if (!h) return set_global("null handle");
  void* p;
  size_t n;
  if (field_info(h, field, copy, &p, &n)) return -1;
  if (!src || offset > n || bytes > n - offset) return h->fail("upload of %zu bytes at %zu exceeds the field (%zu)", bytes, offset, n);
  if (hip_ok(hipMemcpyAsync((char*)p + offset, src, bytes, hipMemcpyHostToDevice, h->stream), &h->err, "upload")) return -1;
  return hip_ok(hipStreamSynchronize(h->stream), &h->err, "upload");
}

extern "C" int obl_download(obl_model* h, int field, int copy, void* dst, size_t bytes, size_t offset) {
  if (!h) return set_global("null handle");
  void* p;
  size_t n;
  if (field_info(h, field, copy, &p, &n)) return -1;
  if (!dst || offset > n || bytes > n - offset) return h->fail("download of %zu bytes at %zu exceeds the field (%zu)", bytes, offset, n);
  if (hip_ok(hipMemcpyAsync(dst, (const char*)p + offset, bytes, hipMemcpyDeviceToHost, h->stream), &h->err, "download"))
    return -1;
  return hip_ok(hipStreamSynchronize(h->stream), &h->err, "download");
}

extern "C" int obl_sync(obl_model* h) {
  if (!h) return set_global("null handle");
  return hip_ok(hipStreamSynchronize(h->stream), &h->err, "sync");
}

extern "C" int obl_set_solver_mode(obl_model* h, int gemm_mode, int thomas_mode, double fp64_cond) {
  if (!h) return set_global("null handle");
  if (gemm_mode != OBL_GEMM_FP32 && gemm_mode != OBL_GEMM_BF16X3 && gemm_mode != OBL_GEMM_BF16X6)
    return h->fail("unknown gemm mode %d", gemm_mode);
  if (thomas_mode < 0 || thomas_mode > 2) return h->fail("unknown Thomas mode %d", thomas_mode);
  h->gemm_mode = gemm_mode;
  h->thomas_mode = thomas_mode;
  h->fp64_cond = fp64_cond;
  return 0;
}

extern "C" int obl_profile(obl_model* h, int on) {
  if (!h) return set_global("null handle");
  hipStreamSynchronize(h->stream);
  h->timer.reset();
  h->solver->timer.reset();
  h->timer.on = on != 0;
  h->solver->timer.on = on != 0;
  return 0;
}

extern "C" int obl_profile_read(obl_model* h, char* names, size_t names_len, double* ms, int* launches, int max) {
  if (!h) return set_global("null handle");
  hipStreamSynchronize(h->stream);
  h->timer.collect();
  h->solver->timer.collect();
  std::string nm;
  int n = 0;
  auto add = [&](Timer& T) {
    for (size_t i = 0; i < T.names.size() && n < max; i++) {
      if (!T.count[i]) continue;
      nm += T.names[i];
      nm += '\n';
      if (ms) ms[n] = T.total[i] / T.count[i];
      if (launches) launches[n] = T.count[i];
      n++;
    }
  };
  add(h->timer);
  add(h->solver->timer);
  if (names && names_len) {
    size_t c = nm.size() < names_len - 1 ? nm.size() : names_len - 1;
    memcpy(names, nm.data(), c);
    names[c] = 0;
  }
  return n;
}

// ---------------------------------------------------------------- surrogate teacher sampling
extern "C" int obl_dump_enable(obl_model* h, unsigned capacity, float p_active, float p_clear) {
  if (!h) return set_global("null handle");
  if (capacity == 0) {
    h->dump_on = false;
    return 0;
  }
  if (capacity != h->dump_cap) {
    if (h->dump_buf) hipFree(h->dump_buf);
    h->dump_buf = nullptr;
    if (hip_ok(hipMalloc(&h->dump_buf, (size_t)capacity * OBL_DUMP_STRIDE * 4), &h->err, "hipMalloc(dump)")) return -1;
    h->dump_cap = capacity;
  }
  if (!h->dump_count && hip_ok(hipMalloc(&h->dump_count, sizeof(unsigned)), &h->err, "hipMalloc(dump count)")) return -1;
  if (hip_ok(hipMemsetAsync(h->dump_count, 0, sizeof(unsigned), h->stream), &h->err, "dump reset")) return -1;
  h->dump_pa = p_active;
  h->dump_pc = p_clear;
  h->dump_on = true;
  return hip_ok(hipStreamSynchronize(h->stream), &h->err, "dump enable");
}

extern "C" int obl_dump_read(obl_model* h, float* out, size_t max_rows, size_t* rows, void** dev_buf) {
  if (!h) return set_global("null handle");
  if (!h->dump_count) return h->fail("sampling is not enabled");
  unsigned n = 0;
  if (hip_ok(hipMemcpyAsync(&n, h->dump_count, sizeof n, hipMemcpyDeviceToHost, h->stream), &h->err, "dump count") ||
      hip_ok(hipStreamSynchronize(h->stream), &h->err, "dump count"))
    return -1;
  size_t m = n < h->dump_cap ? n : h->dump_cap;
  if (rows) *rows = m;
  if (dev_buf) *dev_buf = h->dump_buf;
  if (out && max_rows) {
    size_t c = m < max_rows ? m : max_rows;
    if (hip_ok(hipMemcpy(out, h->dump_buf, c * OBL_DUMP_STRIDE * 4, hipMemcpyDeviceToHost), &h->err, "dump read")) return -1;
  }
  return 0;
}
