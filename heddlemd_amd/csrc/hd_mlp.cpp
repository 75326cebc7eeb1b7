// Host side of the SHAKE multiplier predictor: parameters, Adam state and
// the training loop over device-resident (features, multiplier) rows,
// either copied in from host arrays or read in place from a context's
// SHAKE step log.  The network and the kernels are in k_mlp.hip; the
// reference has no learned component.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/heddle.h"
#include "hd_kernels.h"

// Scratch of hd_mlp_gradient_noalloc: the per-workgroup gradient partials
// of at most HD_MLP_NOALLOC_ROWS rows, one copy per GPU, held from the
// library's first use of that GPU until the process ends.  Host code never
// passes this symbol itself to a kernel (on the host it names a stand-in,
// not device memory): each device's address is resolved once with
// hipGetSymbolAddress and kept in scratch_dev below.
#define HD_MLP_NOALLOC_BLOCKS (HD_MLP_NOALLOC_ROWS / HD_MLP_SAMPLES)
static __device__ __align__(256) float scratch[HD_MLP_NOALLOC_BLOCKS * HD_MLP_PSTRIDE];

struct hd_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  float* params = nullptr;  // HD_MLP_NPARAM
  float *m = nullptr, *v = nullptr, *grad = nullptr;  // HD_MLP_NTRAIN each
  double* loss = nullptr;
  float* part = nullptr;  // gradient partials
  unsigned part_blocks = 0;
  float *x = nullptr, *t = nullptr, *y = nullptr;  // staging for host arrays
  uint64_t stage_rows = 0;
  long adam_step = 0;
  float* scratch_dev = nullptr;  // this device's copy of `scratch`
};

namespace {

int mfail(hd_mlp* m, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (m) m->err = buf;
  return code;
}

#define MCHK(m, expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
      return mfail((m), HD_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                   __FILE__, __LINE__);                                                    \
  } while (0)

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Device address of `scratch` on the current device, resolved once per
// device (the first predictor created there) and cached.
#define HD_MLP_MAX_DEVICES 64
static float* g_scratch_dev[HD_MLP_MAX_DEVICES] = {nullptr};

hipError_t scratch_address(int device, float** out) {
  if (device < 0 || device >= HD_MLP_MAX_DEVICES) return hipErrorInvalidDevice;
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = (float*)p;
  }
  *out = g_scratch_dev[device];
  return hipSuccess;
}

uint64_t gcd64(uint64_t a, uint64_t b) {
  while (b) {
    uint64_t r = a % b;
    a = b;
    b = r;
  }
  return a;
}

int ensure_partials(hd_mlp* m, unsigned blocks) {
  if (blocks <= m->part_blocks) return HD_OK;
  if (m->part) (void)hipFree(m->part);
  m->part = nullptr;
  m->part_blocks = 0;
  MCHK(m, hipMalloc((void**)&m->part, (size_t)blocks * HD_MLP_PSTRIDE * sizeof(float)));
  m->part_blocks = blocks;
  return HD_OK;
}

int ensure_stage(hd_mlp* m, uint64_t rows) {
  if (rows <= m->stage_rows) return HD_OK;
  (void)hipFree(m->x);
  (void)hipFree(m->t);
  (void)hipFree(m->y);
  m->x = m->t = m->y = nullptr;
  m->stage_rows = 0;
  MCHK(m, hipMalloc((void**)&m->x, rows * HD_MLP_IN * sizeof(float)));
  MCHK(m, hipMalloc((void**)&m->t, rows * sizeof(float)));
  MCHK(m, hipMalloc((void**)&m->y, rows * sizeof(float)));
  m->stage_rows = rows;
  return HD_OK;
}

int stage_in(hd_mlp* m, const float* x, const float* t, uint64_t n) {
  int r = ensure_stage(m, n);
  if (r) return r;
  MCHK(m, hipMemcpyAsync(m->x, x, n * HD_MLP_IN * sizeof(float), hipMemcpyHostToDevice,
                         m->stream));
  if (t)
    MCHK(m, hipMemcpyAsync(m->t, t, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
  return HD_OK;
}

float scale_of(hd_mlp* m, int* rc) {
  float s = 0.f;
  hipError_t e = hipMemcpyAsync(&s, m->params + HD_MLP_SCALE, sizeof(float),
                                hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  *rc = e == hipSuccess ? HD_OK
                        : mfail(m, HD_ERR_HIP, "reading the scale: %s", hipGetErrorString(e));
  return s;
}

// Adam over mini-batches of `batch` rows in a fresh pseudo-random order per
// epoch: row(q) = (mul q + add) mod n with gcd(mul, n) = 1, a permutation.
int train_rows(hd_mlp* m, const float* x, const float* t, uint64_t n, int epochs, int batch,
               float lr, uint64_t seed, double* loss_out) {
  if (n == 0 || n >= (1ull << 32)) return mfail(m, HD_ERR_ARG, "training rows %llu out of range",
                                                (unsigned long long)n);
  if (epochs < 1 || batch < HD_MLP_SAMPLES || batch > (1 << 20) || batch % HD_MLP_SAMPLES ||
      !(lr > 0.f))
    return mfail(m, HD_ERR_ARG, "bad training arguments (epochs %d, batch %d, lr %g)", epochs,
                 batch, (double)lr);
  int r = ensure_partials(m, (unsigned)(batch / HD_MLP_SAMPLES));
  if (r) return r;
  float s = scale_of(m, &r);
  if (r) return r;
  if (!(s > 0.f)) hd_launch_mlp_scale(m->stream, x, n, m->params);
  const float b1 = 0.9f, b2 = 0.999f;
  uint64_t rng = seed;
  for (int e = 0; e < epochs; ++e) {
    uint64_t mul = 1, add = 0;
    if (n > 1) {
      do {
        mul = (splitmix64(rng) >> 33) | 1ull;  // odd, below 2^31
      } while (gcd64(mul, n) != 1);
      add = splitmix64(rng) % n;
    }
    MCHK(m, hipMemsetAsync(m->loss, 0, sizeof(double), m->stream));
    for (uint64_t first = 0; first < n; first += (uint64_t)batch) {
      unsigned cnt = (unsigned)std::min<uint64_t>((uint64_t)batch, n - first);
      unsigned nb = hd_launch_mlp_grad(m->stream, m->params, x, t, n, first, cnt, mul, add,
                                       1.0f / (float)cnt, m->part);
      ++m->adam_step;
      HdAdam a;
      a.b1 = b1;
      a.b2 = b2;
      a.eps = 1e-8f;
      double bc1 = 1.0 - std::pow((double)b1, (double)m->adam_step);
      double bc2 = 1.0 - std::pow((double)b2, (double)m->adam_step);
      a.lr_t = (float)(lr * std::sqrt(bc2) / bc1);
      a.m = m->m;
      a.v = m->v;
      hd_launch_mlp_reduce(m->stream, m->part, nb, m->params, nullptr, &a, m->loss);
    }
    MCHK(m, hipGetLastError());
    double l = 0;
    MCHK(m, hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    if (loss_out) loss_out[e] = l / (double)n;
  }
  return HD_OK;
}

}  // namespace

int hd_mlp_param_count(void) { return HD_MLP_NPARAM; }

int hd_mlp_create(int device, uint64_t seed, hd_mlp** out) {
  if (!out) return HD_ERR_ARG;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return HD_ERR_HIP;
  hd_mlp* m = new hd_mlp;
  m->device = device;
  auto bad = [&](hipError_t e) {
    if (e == hipSuccess) return false;
    hd_mlp_destroy(m);
    return true;
  };
  if (bad(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking))) return HD_ERR_HIP;
  if (bad(hipMalloc((void**)&m->params, HD_MLP_NPARAM * sizeof(float)))) return HD_ERR_HIP;
  if (bad(hipMalloc((void**)&m->m, HD_MLP_NTRAIN * sizeof(float)))) return HD_ERR_HIP;
  if (bad(hipMalloc((void**)&m->v, HD_MLP_NTRAIN * sizeof(float)))) return HD_ERR_HIP;
  if (bad(hipMalloc((void**)&m->grad, HD_MLP_NTRAIN * sizeof(float)))) return HD_ERR_HIP;
  if (bad(hipMalloc((void**)&m->loss, sizeof(double)))) return HD_ERR_HIP;
  if (bad(scratch_address(device, &m->scratch_dev))) return HD_ERR_HIP;
  // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) hidden weights, zero biases,
  // zero output layer (the untrained predictor returns the first-order
  // multiplier); scale 0 = derive it from the first training data
  std::vector<float> p(HD_MLP_NPARAM, 0.f);
  uint64_t rng = seed;
  auto uni = [&](float a) {
    double u = (double)(splitmix64(rng) >> 11) * (1.0 / 9007199254740992.0);
    return (float)((2.0 * u - 1.0) * a);
  };
  for (int k = 0; k < HD_MLP_H * HD_MLP_IN; ++k) p[HD_MLP_W1 + k] = uni(1.0f / std::sqrt((float)HD_MLP_IN));
  for (int k = 0; k < HD_MLP_H * HD_MLP_H; ++k) p[HD_MLP_W2 + k] = uni(1.0f / std::sqrt((float)HD_MLP_H));
  int r = hd_mlp_set_params(m, p.data());
  if (r) {
    hd_mlp_destroy(m);
    return r;
  }
  *out = m;
  return HD_OK;
}

void hd_mlp_destroy(hd_mlp* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* p : {(void*)m->params, (void*)m->m, (void*)m->v, (void*)m->grad, (void*)m->loss,
                  (void*)m->part, (void*)m->x, (void*)m->t, (void*)m->y})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

const char* hd_mlp_last_error(const hd_mlp* m) { return m ? m->err.c_str() : "null predictor"; }

int hd_mlp_set_params(hd_mlp* m, const float* params) {
  if (!m || !params) return HD_ERR_ARG;
  for (int k = 0; k < HD_MLP_NPARAM; ++k)
    if (!std::isfinite(params[k])) return mfail(m, HD_ERR_ARG, "parameter %d not finite", k);
  MCHK(m, hipSetDevice(m->device));
  MCHK(m, hipMemcpyAsync(m->params, params, HD_MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice,
                         m->stream));
  MCHK(m, hipMemsetAsync(m->m, 0, HD_MLP_NTRAIN * sizeof(float), m->stream));
  MCHK(m, hipMemsetAsync(m->v, 0, HD_MLP_NTRAIN * sizeof(float), m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  m->adam_step = 0;
  return HD_OK;
}

int hd_mlp_get_params(hd_mlp* m, float* params) {
  if (!m || !params) return HD_ERR_ARG;
  MCHK(m, hipSetDevice(m->device));
  MCHK(m, hipMemcpyAsync(params, m->params, HD_MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost,
                         m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return HD_OK;
}

int hd_mlp_predict(hd_mlp* m, const float* x, uint64_t n, float* nu) {
  if (!m || (n && (!x || !nu))) return HD_ERR_ARG;
  if (n >= (1ull << 32)) return mfail(m, HD_ERR_ARG, "too many rows");
  if (n == 0) return HD_OK;
  MCHK(m, hipSetDevice(m->device));
  int r = stage_in(m, x, nullptr, n);
  if (r) return r;
  hd_launch_mlp_predict(m->stream, m->params, m->x, (unsigned)n, m->y);
  MCHK(m, hipGetLastError());
  MCHK(m, hipMemcpyAsync(nu, m->y, n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return HD_OK;
}

int hd_mlp_gradient(hd_mlp* m, const float* x, const float* t, uint64_t n, float* grad,
                    double* loss) {
  if (!m || !x || !t || !grad || n == 0) return HD_ERR_ARG;
  if (n > (1ull << 24)) return mfail(m, HD_ERR_ARG, "gradient over at most 2^24 rows");
  MCHK(m, hipSetDevice(m->device));
  int r = stage_in(m, x, t, n);
  if (r) return r;
  if ((r = ensure_partials(m, (unsigned)((n + HD_MLP_SAMPLES - 1) / HD_MLP_SAMPLES)))) return r;
  MCHK(m, hipMemsetAsync(m->loss, 0, sizeof(double), m->stream));
  unsigned nb = hd_launch_mlp_grad(m->stream, m->params, m->x, m->t, n, 0, (unsigned)n, 1, 0,
                                   1.0f / (float)n, m->part);
  hd_launch_mlp_reduce(m->stream, m->part, nb, m->params, m->grad, nullptr, m->loss);
  MCHK(m, hipGetLastError());
  double l = 0;
  MCHK(m, hipMemcpyAsync(grad, m->grad, HD_MLP_NTRAIN * sizeof(float), hipMemcpyDeviceToHost,
                         m->stream));
  MCHK(m, hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  if (loss) *loss = l / (double)n;
  return HD_OK;
}

int hd_mlp_gradient_noalloc(hd_mlp* m, const float* x, const float* t, uint64_t n, float* grad,
                            double* loss) {
  if (!m || !x || !t || !grad || !loss || n == 0) return HD_ERR_ARG;
  if (n > HD_MLP_NOALLOC_ROWS)
    return mfail(m, HD_ERR_ARG, "hd_mlp_gradient_noalloc takes at most %d rows (got %llu)",
                 HD_MLP_NOALLOC_ROWS, (unsigned long long)n);
  if (!m->scratch_dev) return mfail(m, HD_ERR_STATE, "scratch address not resolved");
  MCHK(m, hipSetDevice(m->device));
  // the reduction adds the loss sum into *loss: cleared by a kernel (no
  // memset node, so the call stays capturable)
  hd_launch_fill_f32(m->stream, (float*)loss, 0.f, 2);
  unsigned nb = hd_launch_mlp_grad(m->stream, m->params, x, t, n, 0, (unsigned)n, 1, 0,
                                   1.0f / (float)n, m->scratch_dev);
  hd_launch_mlp_reduce(m->stream, m->scratch_dev, nb, m->params, grad, nullptr, loss);
  MCHK(m, hipGetLastError());
  MCHK(m, hipStreamSynchronize(m->stream));
  return HD_OK;
}

int hd_mlp_train(hd_mlp* m, const float* x, const float* t, uint64_t n, int epochs, int batch,
                 float lr, uint64_t seed, double* loss) {
  if (!m || !x || !t) return HD_ERR_ARG;
  MCHK(m, hipSetDevice(m->device));
  int r = stage_in(m, x, t, n);
  if (r) return r;
  return train_rows(m, m->x, m->t, n, epochs, batch, lr, seed, loss);
}

int hd_mlp_train_log(hd_mlp* m, hd_ctx* c, int epochs, int batch, float lr, uint64_t seed,
                     double* loss) {
  if (!m || !c) return HD_ERR_ARG;
  int dev = -1;
  const float *x = nullptr, *t = nullptr;
  uint64_t rows = 0;
  int r = hd_ctx_shake_log_view(c, &dev, &x, &t, &rows);
  if (r) return mfail(m, r, "context: %s", hd_last_error(c));
  if (dev != m->device)
    return mfail(m, HD_ERR_ARG, "log on device %d, predictor on device %d", dev, m->device);
  MCHK(m, hipSetDevice(m->device));
  return train_rows(m, x, t, rows, epochs, batch, lr, seed, loss);
}
