// C ABI of the microphysics surrogate network (include/outbreak_hip.h, obl_mlp_*):
// parameters on the device, GPU training with Adam, and exact evaluation for checks.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/outbreak_hip.h"
#include "obl_mlp.h"
#include "obl_mlp_infer.h"

using namespace obl;

struct obl_mlp {
  hipStream_t stream = nullptr;
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr, *loss = nullptr;
  float *Gpart = nullptr, *lpart = nullptr;  // per-block gradient and loss partials
  int part_blocks = 0;
  uint16_t* frag = nullptr;  // inference form: packed BF16 weights and FP32 biases
  float* bias = nullptr;
  float* rows = nullptr;  // rows of obl_mlp_load_rows
  size_t nrows = 0;
  int stride = 0;
  float out_scale[MLP_OUT];
  float in_lo[MLP_NIN], in_hi[MLP_NIN];
  int adam_t = 0;
  float dead_frac = 0.02f;
  float hand_min = 1e-9f;
  std::string err;
  int fail(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return -1;
  }
  int ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    return fail("%s: %s", what, hipGetErrorString(e));
  }
};

static std::string g_mlp_err;

// refresh the inference form from the master parameters (after every change of them)
static int repack(obl_mlp* m) {
  launch_mlp_pack(m->stream, m->P, m->frag, m->bias);
  if (m->ok(hipGetLastError(), "pack")) return -1;
  return m->ok(hipStreamSynchronize(m->stream), "pack");
}

namespace obl {
int mlp_view(obl_mlp* m, MlpView* v) {
  if (!m || !v) return -1;
  v->frag = m->frag;
  v->bias = m->bias;
  for (int j = 0; j < MLP_NOUT; j++) v->out_scale[j] = m->out_scale[j];
  for (int i = 0; i < MLP_NIN; i++) {
    v->lo[i] = m->in_lo[i];
    v->hi[i] = m->in_hi[i];
  }
  v->dead_frac = m->dead_frac;
  v->hand_min = m->hand_min;
  return 0;
}
}  // namespace obl

extern "C" const char* obl_mlp_last_error(const obl_mlp* m) { return m ? m->err.c_str() : g_mlp_err.c_str(); }

extern "C" void obl_mlp_destroy(obl_mlp* m) {
  if (!m) return;
  if (m->stream) hipStreamSynchronize(m->stream);
  void* b[] = {m->P, m->G, m->M, m->V, m->loss, m->rows, m->frag, m->bias, m->Gpart, m->lpart};
  for (void* p : b)
    if (p) hipFree(p);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
}

extern "C" int obl_mlp_create(obl_mlp** out) {
  if (!out) {
    g_mlp_err = "null output";
    return -1;
  }
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    g_mlp_err = "surrogate: no ROCm GPU is available";
    return -1;
  }
  obl_mlp* m = new obl_mlp();
  for (int j = 0; j < MLP_OUT; j++) m->out_scale[j] = 1.0f;
  for (int i = 0; i < MLP_NIN; i++) { m->in_lo[i] = -3.0e38f; m->in_hi[i] = 3.0e38f; }
  size_t n = (size_t)MLP_NPARAM * 4;
  if (m->ok(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "stream") || m->ok(hipMalloc(&m->P, n), "P") ||
      m->ok(hipMalloc(&m->G, n), "G") || m->ok(hipMalloc(&m->M, n), "M") || m->ok(hipMalloc(&m->V, n), "V") ||
      m->ok(hipMalloc(&m->loss, 4), "loss") || m->ok(hipMalloc(&m->frag, (size_t)MLP_PACK_HALVES * 2), "frag") ||
      m->ok(hipMalloc(&m->bias, (size_t)MLP_BIAS_FLOATS * 4), "bias") ||
      m->ok(hipMemsetAsync(m->P, 0, n, m->stream), "clear") ||
      m->ok(hipMemsetAsync(m->G, 0, n, m->stream), "clear") || m->ok(hipMemsetAsync(m->M, 0, n, m->stream), "clear") ||
      m->ok(hipMemsetAsync(m->V, 0, n, m->stream), "clear") || m->ok(hipStreamSynchronize(m->stream), "sync") ||
      repack(m)) {
    g_mlp_err = m->err;
    obl_mlp_destroy(m);
    return -1;
  }
  *out = m;
  return 0;
}

extern "C" int obl_mlp_nparams(void) { return MLP_NPARAM; }

extern "C" int obl_mlp_set_params(obl_mlp* m, const float* p, int n) {
  if (!m || !p || n != MLP_NPARAM) return m ? m->fail("expected %d parameters", MLP_NPARAM) : -1;
  m->adam_t = 0;
  size_t b = (size_t)MLP_NPARAM * 4;
  if (m->ok(hipMemcpyAsync(m->P, p, b, hipMemcpyHostToDevice, m->stream), "set params") ||
      m->ok(hipMemsetAsync(m->M, 0, b, m->stream), "clear") || m->ok(hipMemsetAsync(m->V, 0, b, m->stream), "clear") ||
      m->ok(hipMemsetAsync(m->G, 0, b, m->stream), "clear"))
    return -1;
  if (m->ok(hipStreamSynchronize(m->stream), "set params")) return -1;
  return repack(m);
}

extern "C" int obl_mlp_get_params(obl_mlp* m, float* p, int n) {
  if (!m || !p || n != MLP_NPARAM) return m ? m->fail("expected %d parameters", MLP_NPARAM) : -1;
  if (m->ok(hipMemcpyAsync(p, m->P, (size_t)MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->stream), "get params")) return -1;
  return m->ok(hipStreamSynchronize(m->stream), "get params");
}

extern "C" int obl_mlp_set_scales(obl_mlp* m, const float* out_scale, const float* in_lo, const float* in_hi) {
  if (!m || !out_scale) return m ? m->fail("null scales") : -1;
  for (int j = 0; j < MLP_OUT; j++) m->out_scale[j] = j < MLP_NOUT ? out_scale[j] : 1.0f;
  for (int i = 0; i < MLP_NIN; i++) {
    m->in_lo[i] = in_lo ? in_lo[i] : -3.0e38f;
    m->in_hi[i] = in_hi ? in_hi[i] : 3.0e38f;
  }
  return 0;
}

extern "C" int obl_mlp_set_dead_zone(obl_mlp* m, float frac) {
  if (!m) return -1;
  if (!(frac >= 0.0f && frac < 1.0f)) return m->fail("dead zone fraction must be in [0, 1)");
  m->dead_frac = frac;
  return 0;
}

extern "C" int obl_mlp_set_hand_min(obl_mlp* m, float qmin) {
  if (!m) return -1;
  if (!(qmin >= 0.0f && qmin < 1.0f)) return m->fail("handing threshold must be in [0, 1) kg/kg");
  m->hand_min = qmin;
  return 0;
}

// partial buffers for a batch (grown on demand)
static int ensure_parts(obl_mlp* m, int batch) {
  int nb = mlp_grad_blocks(batch);
  if (nb <= m->part_blocks) return 0;
  if (m->Gpart || m->lpart) {
    hipStreamSynchronize(m->stream);
    hipFree(m->Gpart);
    hipFree(m->lpart);
    m->Gpart = m->lpart = nullptr;
    m->part_blocks = 0;
  }
  if (m->ok(hipMalloc(&m->Gpart, (size_t)nb * MLP_NPARAM * 4), "gradient partials") ||
      m->ok(hipMalloc(&m->lpart, (size_t)nb * 4), "loss partials"))
    return -1;
  m->part_blocks = nb;
  return 0;
}

static void fill_args(obl_mlp* m, MlpTrainArgs& a, const float* rows, size_t nrows, int stride, int batch, int quant) {
  memset(&a, 0, sizeof a);
  a.P = m->P;
  a.rows = rows;
  a.stride = stride;
  a.nrows = nrows;
  a.batch = batch;
  for (int j = 0; j < MLP_OUT; j++) a.out_inv_scale[j] = 1.0f / m->out_scale[j];
  a.loss_scale = 1.0f / ((float)batch * (float)MLP_NOUT);
  a.lpart = m->lpart;
  a.quant = quant;
}

extern "C" int obl_mlp_load_rows(obl_mlp* m, const float* rows_host, size_t nrows, int stride) {
  if (!m) return -1;
  if (!rows_host || nrows == 0 || stride < MLP_NIN + MLP_NOUT)
    return m->fail("load_rows: need rows (stride >= %d)", MLP_NIN + MLP_NOUT);
  if (m->rows) {
    hipStreamSynchronize(m->stream);
    hipFree(m->rows);
    m->rows = nullptr;
    m->nrows = 0;
  }
  size_t b = nrows * (size_t)stride * 4;
  if (m->ok(hipMalloc(&m->rows, b), "rows")) return -1;
  if (m->ok(hipMemcpyAsync(m->rows, rows_host, b, hipMemcpyHostToDevice, m->stream), "rows") ||
      m->ok(hipStreamSynchronize(m->stream), "rows"))
    return -1;
  m->nrows = nrows;
  m->stride = stride;
  return 0;
}

extern "C" int obl_mlp_train(obl_mlp* m, const float* rows_dev, size_t nrows, int stride, int steps, int batch, float lr,
                             unsigned seed, int quant, float* mean_loss) {
  if (!m) return -1;
  if (!rows_dev) {
    if (!m->rows) return m->fail("train: no device rows given and none loaded");
    rows_dev = m->rows;
    nrows = m->nrows;
    stride = m->stride;
  }
  if (!rows_dev || nrows == 0 || stride < MLP_NIN + MLP_NOUT || batch < 1 || steps < 1)
    return m->fail("train: need rows (stride >= %d), batch >= 1, steps >= 1", MLP_NIN + MLP_NOUT);
  if (ensure_parts(m, batch)) return -1;
  MlpTrainArgs a;
  fill_args(m, a, rows_dev, nrows, stride, batch, quant);
  a.G = m->Gpart;
  if (m->ok(hipMemsetAsync(m->loss, 0, 4, m->stream), "loss")) return -1;
  for (int s = 0; s < steps; s++) {
    a.seed = seed * 2654435761u + (unsigned)s;
    launch_mlp_grad(m->stream, a, m->G, m->loss);
    launch_adam(m->stream, m->P, m->G, m->M, m->V, MLP_NPARAM, lr, 0.9f, 0.999f, 1e-8f, ++m->adam_t);
  }
  if (m->ok(hipGetLastError(), "train launch")) return -1;
  float l = 0.f;
  if (m->ok(hipMemcpyAsync(&l, m->loss, 4, hipMemcpyDeviceToHost, m->stream), "loss") ||
      m->ok(hipStreamSynchronize(m->stream), "train") || repack(m))
    return -1;
  if (mean_loss) *mean_loss = l / ((float)steps * (float)batch * (float)MLP_NOUT);
  return 0;
}

extern "C" int obl_mlp_infer(obl_mlp* m, const float* rows_host, int n, int stride, float* y) {
  if (!m) return -1;
  if (!rows_host || !y || n < 1 || stride < MLP_NIN) return m->fail("infer: need n >= 1 rows of >= %d floats", MLP_NIN);
  float *drows = nullptr, *dy = nullptr;
  size_t rb = (size_t)n * stride * 4, yb = (size_t)n * MLP_OUT * 4;
  if (m->ok(hipMalloc(&drows, rb), "rows") || m->ok(hipMalloc(&dy, yb), "y")) {
    hipFree(drows);
    hipFree(dy);
    return -1;
  }
  int rc = m->ok(hipMemcpyAsync(drows, rows_host, rb, hipMemcpyHostToDevice, m->stream), "rows");
  if (!rc) {
    launch_mlp_eval(m->stream, m->frag, m->bias, drows, stride, n, dy);
    rc = m->ok(hipGetLastError(), "infer launch");
  }
  if (!rc) rc = m->ok(hipMemcpyAsync(y, dy, yb, hipMemcpyDeviceToHost, m->stream), "y");
  if (!rc) rc = m->ok(hipStreamSynchronize(m->stream), "infer");
  hipStreamSynchronize(m->stream);
  hipFree(drows);
  hipFree(dy);
  return rc;
}

extern "C" int obl_mlp_eval(obl_mlp* m, const float* rows_host, int n, int stride, int quant, float* loss, float* grads,
                            float* y) {
  if (!m) return -1;
  if (!rows_host || n < 1 || stride < MLP_NIN + MLP_NOUT) return m->fail("eval: need n >= 1 rows");
  float *drows = nullptr, *dy = nullptr;
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  int* didx = nullptr;
  size_t rb = (size_t)n * stride * 4;
  if (m->ok(hipMalloc(&drows, rb), "rows") || m->ok(hipMalloc(&didx, (size_t)n * 4), "idx") ||
      m->ok(hipMalloc(&dy, (size_t)n * MLP_OUT * 4), "y")) {
    hipFree(drows); hipFree(didx); hipFree(dy);
    return -1;
  }
  hipMemcpyAsync(drows, rows_host, rb, hipMemcpyHostToDevice, m->stream);
  hipMemcpyAsync(didx, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice, m->stream);
  hipMemsetAsync(m->loss, 0, 4, m->stream);
  if (ensure_parts(m, n)) {
    hipFree(drows); hipFree(didx); hipFree(dy);
    return -1;
  }
  MlpTrainArgs a;
  fill_args(m, a, drows, (size_t)n, stride, n, quant);
  a.idx = didx;
  a.G = grads ? m->Gpart : nullptr;
  a.y_out = dy;
  launch_mlp_grad(m->stream, a, m->G, m->loss);
  float l = 0.f;
  int rc = m->ok(hipGetLastError(), "eval launch");
  if (!rc) rc = m->ok(hipMemcpyAsync(&l, m->loss, 4, hipMemcpyDeviceToHost, m->stream), "loss");
  if (!rc && grads) rc = m->ok(hipMemcpyAsync(grads, m->G, (size_t)MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->stream), "grads");
  if (!rc && y) rc = m->ok(hipMemcpyAsync(y, dy, (size_t)n * MLP_OUT * 4, hipMemcpyDeviceToHost, m->stream), "y");
  if (!rc) rc = m->ok(hipStreamSynchronize(m->stream), "eval");
  hipStreamSynchronize(m->stream);
  hipFree(drows);
  hipFree(didx);
  hipFree(dy);
  if (loss) *loss = l / ((float)n * (float)MLP_NOUT);
  return rc;
}

