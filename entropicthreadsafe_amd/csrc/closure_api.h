// closure_api.h -- host side of the learned subgrid closure (the tslb_closure_*
// entry points of include/tslb.h), included at the end of tslb.hip.
#pragma once

struct tslb_closure {
  int device = 0;
  hipStream_t stream = nullptr;
  float *th = nullptr;                   // FP32 master parameters (natural layout)
  bf16x8 *wf = nullptr, *wb = nullptr;   // packed BF16 operands (forward, backward)
  float *grad = nullptr, *m1 = nullptr, *m2 = nullptr;  // gradient, Adam moments
  int adam_t = 0;
  size_t cap = 0, ns = 0;                // sample capacity and count
  float *feat = nullptr;                 // inputs, SoA 10 x cap
  float *aux = nullptr;                  // rho_bar, tau[6]: SoA 7 x cap
  size_t tcap = 0;                       // training scratch: batch columns
  int ldb = 0;
  __bf16 *tbuf = nullptr;                // H0 | H1 | H2 | H3 | D1 | D2 | D3 | D4
  int *idx = nullptr;
  double *loss = nullptr;
  float *coarse = nullptr;               // box-filter output, SoA 5 x ccap
  size_t ccap = 0;
  size_t bytes = 0;
};

namespace {

constexpr int kTrainRows = 16 + 3 * 64 + 3 * 64 + 16;

int clo_alloc(tslb_closure *m, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(TSLB_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                  hipGetErrorString(e));
  }
  m->bytes += bytes;
  return TSLB_OK;
}

int check_model(tslb_closure *m) {
  if (!m) return fail(TSLB_E_ARG, "null closure handle");
  HIPCK(hipSetDevice(m->device));
  return TSLB_OK;
}

__global__ __launch_bounds__(256) void k_sample_idx(int *__restrict__ idx, int nb, long long ns, uint32_t seed,
                                                    uint32_t step) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nb) return;
  // splitmix64 of (seed, step, i)
  uint64_t x = (uint64_t)seed * 0x9E3779B97F4A7C15ull + (uint64_t)step * 0xD1B54A32D192ED03ull + (uint64_t)i;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  idx[i] = (int)(x % (uint64_t)ns);
}
__global__ __launch_bounds__(256) void k_iota(int *__restrict__ idx, int nb, int first) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nb) idx[i] = first + i;
}

void pack_weights(tslb_closure *m, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(((kFwdBlocks + kBwdBlocks) * 64 + 255) / 256), dim3(256), 0, st, m->th,
                     m->wf, m->wb);
}

unsigned grid_for_tiles(long long tiles) {
  long long b = (tiles + 3) / 4;
  if (b > 2048) b = 2048;
  return (unsigned)(b < 1 ? 1 : b);
}

MlpTrainBufs train_bufs(tslb_closure *m) {
  MlpTrainBufs tb;
  __bf16 *b = m->tbuf;
  const size_t r = (size_t)m->ldb;
  tb.h0 = b;
  tb.h1 = b + 16 * r;
  tb.h2 = tb.h1 + 64 * r;
  tb.h3 = tb.h2 + 64 * r;
  tb.d1 = tb.h3 + 64 * r;
  tb.d2 = tb.d1 + 64 * r;
  tb.d3 = tb.d2 + 64 * r;
  tb.d4 = tb.d3 + 64 * r;
  return tb;
}

int ensure_train(tslb_closure *m, size_t nb) {
  if (nb <= m->tcap) return TSLB_OK;
  (void)hipFree(m->tbuf);
  (void)hipFree(m->idx);
  m->tbuf = nullptr;
  m->idx = nullptr;
  m->tcap = 0;
  const size_t ldb = (nb + 31) / 32 * 32;
  if (ldb > (size_t)INT32_MAX / kTrainRows) return fail(TSLB_E_ARG, "batch too large");
  int rc;
  if ((rc = clo_alloc(m, (void **)&m->tbuf, (size_t)kTrainRows * ldb * sizeof(__bf16))) ||
      (rc = clo_alloc(m, (void **)&m->idx, ldb * sizeof(int))))
    return rc;
  m->tcap = nb;
  m->ldb = (int)ldb;
  return TSLB_OK;
}

// batches above this many samples are processed in chunks
constexpr int64_t kMaxChunk = 1 << 20;

void zero_grad(tslb_closure *m) {
  hipLaunchKernelGGL(k_fill, dim3((kMlpNParam + 255) / 256), dim3(256), 0, m->stream, m->grad, 0.f, kMlpNParam);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, m->stream, (float *)m->loss, 0.f, 2);
}

// adds the gradient (and FP64 loss) of the samples idx[0..nb) to m->grad,
// m->loss; ntot is the batch size the mean is taken over
int run_grad(tslb_closure *m, int nb, int64_t ntot, float t_scale) {
  hipStream_t st = m->stream;
  const float inv = (float)(1.0 / ((double)t_scale * t_scale * (double)ntot));
  const MlpTrainBufs tb = train_bufs(m);
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(grid_for_tiles((nb + 15) / 16)), dim3(256), 0, st, m->feat,
                     (long long)m->cap, m->aux, (long long)m->cap, m->idx, nb, m->ldb, m->wf, m->wb, m->th, inv,
                     tb, m->loss);
  const int ksl = 512;  // columns per weight-gradient slice (multiple of 32)
  const int nslice = (nb + ksl - 1) / ksl;
  const int waves = kWgradJobs * nslice;
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((waves + 3) / 4), dim3(256), 0, st, tb, nb, m->ldb, ksl, m->grad);
  HIPCK(hipGetLastError());
  return TSLB_OK;
}

// closure fields of every node of c from its current state (before a collision)
int launch_closure(tslb_ctx *c) {
  tslb_closure *m = c->model;
  const Params &p = c->kp;
  const long long n = (long long)c->n;
  with_storage(c, [&](auto tag) {
    using S = decltype(tag);
    StateView<S> v{(const S *)c->buf[c->cur], (long long)c->fs};
    hipLaunchKernelGGL(k_features<StateView<S>>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, v,
                       p, (int)TSLB_U, p.ncomp == 2 ? (int)TSLB_PHI : -1, c->cfeat, n, 0LL);
    return 0;
  });
  hipLaunchKernelGGL(k_mlp_infer, dim3(grid_for_tiles((n + 15) / 16)), dim3(256), 0, c->stream, c->cfeat, n, n,
                     m->wf, m->th, c->clo, (long long)c->fs);
  HIPCK(hipGetLastError());
  return TSLB_OK;
}

int ensure_closure_fields(tslb_ctx *c) {
  int rc;
  if (!c->clo && (rc = dev_alloc(c, (void **)&c->clo, 7 * c->fs * sizeof(float)))) return rc;
  if (!c->cfeat && (rc = dev_alloc(c, (void **)&c->cfeat, 10 * c->n * sizeof(float)))) return rc;
  return TSLB_OK;
}

int closure_allowed(tslb_ctx *c) {
  if (c->prm.ncomp == TSLB_COLOUR) return fail(TSLB_E_ARG, "the learned closure supports ncomp 1 and 2");
  if (c->jet) return fail(TSLB_E_STATE, "the learned closure is not combined with the jet boundary");
  return TSLB_OK;
}

}  // namespace

extern "C" {

int tslb_closure_create(tslb_closure **out, int64_t max_samples) {
  if (!out) return fail(TSLB_E_ARG, "null output handle");
  *out = nullptr;
  if (max_samples < 0) return fail(TSLB_E_ARG, "max_samples < 0");
  tslb_closure *m = new tslb_closure();
  HIPCK(hipGetDevice(&m->device));
  int rc = TSLB_OK;
  const size_t cap = (size_t)max_samples;
  if ((rc = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess
                ? TSLB_OK
                : fail(TSLB_E_HIP, "hipStreamCreate failed")) ||
      (rc = clo_alloc(m, (void **)&m->th, kMlpNParam * sizeof(float))) ||
      (rc = clo_alloc(m, (void **)&m->wf, kFwdBlocks * 64 * sizeof(bf16x8))) ||
      (rc = clo_alloc(m, (void **)&m->wb, kBwdBlocks * 64 * sizeof(bf16x8))) ||
      (rc = clo_alloc(m, (void **)&m->grad, kMlpNParam * sizeof(float))) ||
      (rc = clo_alloc(m, (void **)&m->m1, kMlpNParam * sizeof(float))) ||
      (rc = clo_alloc(m, (void **)&m->m2, kMlpNParam * sizeof(float))) ||
      (rc = clo_alloc(m, (void **)&m->loss, 2 * sizeof(double))) ||
      (cap && (rc = clo_alloc(m, (void **)&m->feat, 10 * cap * sizeof(float)))) ||
      (cap && (rc = clo_alloc(m, (void **)&m->aux, kAux * cap * sizeof(float))))) {
    tslb_closure_destroy(m);
    return rc;
  }
  m->cap = cap;
  for (float *b : {m->th, m->grad, m->m1, m->m2})
    hipLaunchKernelGGL(k_fill, dim3((kMlpNParam + 255) / 256), dim3(256), 0, m->stream, b, 0.f, kMlpNParam);
  pack_weights(m, m->stream);
  HIPCK(hipStreamSynchronize(m->stream));
  *out = m;
  return TSLB_OK;
}

int tslb_closure_destroy(tslb_closure *m) {
  if (!m) return TSLB_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void *b : {(void *)m->th, (void *)m->wf, (void *)m->wb, (void *)m->grad, (void *)m->m1, (void *)m->m2,
                  (void *)m->loss, (void *)m->feat, (void *)m->aux, (void *)m->tbuf, (void *)m->idx,
                  (void *)m->coarse})
    (void)hipFree(b);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return TSLB_OK;
}

int tslb_closure_set_params(tslb_closure *m, const float *theta, int n) {
  int rc = check_model(m);
  if (rc) return rc;
  if (!theta || n != kMlpNParam) return fail(TSLB_E_ARG, "expected TSLB_CLOSURE_NPARAM parameters");
  HIPCK(hipMemcpyAsync(m->th, theta, kMlpNParam * sizeof(float), hipMemcpyHostToDevice, m->stream));
  for (float *b : {m->m1, m->m2})
    hipLaunchKernelGGL(k_fill, dim3((kMlpNParam + 255) / 256), dim3(256), 0, m->stream, b, 0.f, kMlpNParam);
  m->adam_t = 0;
  pack_weights(m, m->stream);
  HIPCK(hipStreamSynchronize(m->stream));
  return TSLB_OK;
}

int tslb_closure_get_params(tslb_closure *m, float *theta, int n) {
  int rc = check_model(m);
  if (rc) return rc;
  if (!theta || n != kMlpNParam) return fail(TSLB_E_ARG, "expected TSLB_CLOSURE_NPARAM parameters");
  HIPCK(hipMemcpyAsync(theta, m->th, kMlpNParam * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  HIPCK(hipStreamSynchronize(m->stream));
  return TSLB_OK;
}

int tslb_closure_add_samples(tslb_closure *m, tslb_ctx *c, int factor) {
  int rc = check_model(m);
  if (rc) return rc;
  if ((rc = check_ctx(c))) return rc;
  if (c->prm.device != m->device) return fail(TSLB_E_ARG, "closure and lattice are on different devices");
  if (c->prm.ncomp == TSLB_COLOUR) return fail(TSLB_E_ARG, "samples are taken from ncomp 1 or 2 lattices");
  const int r = factor;
  if (r < 1 || c->prm.nx % r || c->prm.ny % r || c->prm.nz % r)
    return fail(TSLB_E_ARG, "factor must divide nx, ny and nz");
  const size_t nc = c->n / ((size_t)r * r * r);
  if (m->ns + nc > m->cap) return fail(TSLB_E_ARG, "sample capacity exceeded");
  if (nc > m->ccap) {
    (void)hipFree(m->coarse);
    m->coarse = nullptr;
    m->ccap = 0;
    if ((rc = clo_alloc(m, (void **)&m->coarse, 5 * nc * sizeof(float)))) return rc;
    m->ccap = nc;
  }
  Params pc = c->kp;
  pc.nx /= r;
  pc.ny /= r;
  pc.nz /= r;
  const unsigned g = (unsigned)((nc + 255) / 256);
  with_storage(c, [&](auto tag) {
    using S = decltype(tag);
    StateView<S> v{(const S *)c->buf[c->cur], (long long)c->fs};
    hipLaunchKernelGGL(k_box_filter<S>, dim3(g), dim3(256), 0, c->stream, v, c->kp, r, m->coarse, (long long)nc,
                       m->aux, (long long)m->cap, (long long)m->ns);
    return 0;
  });
  PlainView cv{m->coarse, (long long)nc};
  hipLaunchKernelGGL(k_features<PlainView>, dim3(g), dim3(256), 0, c->stream, cv, pc, 1, c->prm.ncomp == 2 ? 4 : -1,
                     m->feat, (long long)m->cap, (long long)m->ns);
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(c->stream));
  m->ns += nc;
  return TSLB_OK;
}

int tslb_closure_num_samples(tslb_closure *m, int64_t *ns) {
  int rc = check_model(m);
  if (rc) return rc;
  if (!ns) return fail(TSLB_E_ARG, "null output");
  *ns = (int64_t)m->ns;
  return TSLB_OK;
}

int tslb_closure_clear_samples(tslb_closure *m) {
  int rc = check_model(m);
  if (rc) return rc;
  m->ns = 0;
  return TSLB_OK;
}

int tslb_closure_get_samples(tslb_closure *m, float *feat, float *aux) {
  int rc = check_model(m);
  if (rc) return rc;
  if (!feat || !aux) return fail(TSLB_E_ARG, "null output");
  if (m->ns) {
    HIPCK(hipMemcpy2DAsync(feat, m->ns * sizeof(float), m->feat, m->cap * sizeof(float), m->ns * sizeof(float), 10,
                           hipMemcpyDeviceToHost, m->stream));
    HIPCK(hipMemcpy2DAsync(aux, m->ns * sizeof(float), m->aux, m->cap * sizeof(float), m->ns * sizeof(float), kAux,
                           hipMemcpyDeviceToHost, m->stream));
  }
  HIPCK(hipStreamSynchronize(m->stream));
  return TSLB_OK;
}

int tslb_closure_set_samples(tslb_closure *m, const float *feat, const float *aux, int64_t ns) {
  int rc = check_model(m);
  if (rc) return rc;
  if (!feat || !aux || ns < 0 || (size_t)ns > m->cap) return fail(TSLB_E_ARG, "bad sample arrays or count");
  if (ns) {
    HIPCK(hipMemcpy2DAsync(m->feat, m->cap * sizeof(float), feat, (size_t)ns * sizeof(float), ns * sizeof(float), 10,
                           hipMemcpyHostToDevice, m->stream));
    HIPCK(hipMemcpy2DAsync(m->aux, m->cap * sizeof(float), aux, (size_t)ns * sizeof(float), ns * sizeof(float), kAux,
                           hipMemcpyHostToDevice, m->stream));
  }
  HIPCK(hipStreamSynchronize(m->stream));
  m->ns = (size_t)ns;
  return TSLB_OK;
}

int tslb_closure_forward(tslb_closure *m, float *out) {
  int rc = check_model(m);
  if (rc) return rc;
  if (!out) return fail(TSLB_E_ARG, "null output");
  if (!m->ns) return TSLB_OK;
  float *d = nullptr;
  if ((rc = clo_alloc(m, (void **)&d, kMlpNO * m->ns * sizeof(float)))) return rc;
  hipLaunchKernelGGL(k_mlp_infer, dim3(grid_for_tiles(((long long)m->ns + 15) / 16)), dim3(256), 0, m->stream,
                     m->feat, (long long)m->cap, (long long)m->ns, m->wf, m->th, d, (long long)m->ns);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, kMlpNO * m->ns * sizeof(float), hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  (void)hipFree(d);
  m->bytes -= kMlpNO * m->ns * sizeof(float);
  if (e != hipSuccess) return fail(TSLB_E_HIP, std::string("closure forward failed: ") + hipGetErrorString(e));
  return TSLB_OK;
}

int tslb_closure_loss_grad(tslb_closure *m, const int32_t *idx, int64_t nb, float t_scale, double *loss,
                           float *grad) {
  int rc = check_model(m);
  if (rc) return rc;
  if (!loss) return fail(TSLB_E_ARG, "null loss output");
  if (!idx) nb = (int64_t)m->ns;
  if (nb <= 0 || nb > INT32_MAX) return fail(TSLB_E_ARG, "empty or oversized batch");
  if (!(t_scale > 0.f)) return fail(TSLB_E_ARG, "t_scale must be positive");
  if (idx) {
    for (int64_t i = 0; i < nb; i++)
      if (idx[i] < 0 || (size_t)idx[i] >= m->ns) return fail(TSLB_E_ARG, "sample index out of range");
  }
  const int64_t chunk = nb < kMaxChunk ? nb : kMaxChunk;
  if ((rc = ensure_train(m, (size_t)chunk))) return rc;
  zero_grad(m);
  for (int64_t c0 = 0; c0 < nb; c0 += chunk) {
    const int len = (int)(nb - c0 < chunk ? nb - c0 : chunk);
    if (idx) {
      HIPCK(hipMemcpyAsync(m->idx, idx + c0, len * sizeof(int32_t), hipMemcpyHostToDevice, m->stream));
    } else {
      hipLaunchKernelGGL(k_iota, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, m->stream, m->idx, len,
                         (int)c0);
    }
    if ((rc = run_grad(m, len, nb, t_scale))) return rc;
    // the next chunk's indices overwrite m->idx once this chunk is done
    if (idx) HIPCK(hipStreamSynchronize(m->stream));
  }
  HIPCK(hipMemcpyAsync(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (grad) HIPCK(hipMemcpyAsync(grad, m->grad, kMlpNParam * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  HIPCK(hipStreamSynchronize(m->stream));
  return TSLB_OK;
}

int tslb_closure_train(tslb_closure *m, int steps, int batch, float lr, float t_scale, uint32_t seed,
                       double *last_loss) {
  int rc = check_model(m);
  if (rc) return rc;
  if (steps < 0 || batch <= 0 || batch > kMaxChunk) return fail(TSLB_E_ARG, "steps < 0 or batch outside 1..2^20");
  if (!m->ns) return fail(TSLB_E_STATE, "no samples");
  if (!(t_scale > 0.f) || !(lr > 0.f)) return fail(TSLB_E_ARG, "lr and t_scale must be positive");
  if ((rc = ensure_train(m, (size_t)batch))) return rc;
  for (int s = 0; s < steps; s++) {
    hipLaunchKernelGGL(k_sample_idx, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, m->stream, m->idx, batch,
                       (long long)m->ns, seed, (uint32_t)(m->adam_t + 1));
    zero_grad(m);
    if ((rc = run_grad(m, batch, batch, t_scale))) return rc;
    m->adam_t += 1;
    hipLaunchKernelGGL(k_mlp_adam, dim3((kOffIn + 255) / 256), dim3(256), 0, m->stream, m->th, m->grad, m->m1, m->m2,
                       lr, 0.9f, 0.999f, 1e-8f, m->adam_t);
    pack_weights(m, m->stream);
  }
  HIPCK(hipGetLastError());
  if (last_loss) {
    *last_loss = 0.0;
    if (steps) HIPCK(hipMemcpyAsync(last_loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  }
  HIPCK(hipStreamSynchronize(m->stream));
  return TSLB_OK;
}

int tslb_set_closure(tslb_ctx *c, tslb_closure *m) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!m) {
    c->model = nullptr;
    c->clo_mode = 0;
    return TSLB_OK;
  }
  if (m->device != c->prm.device) return fail(TSLB_E_ARG, "closure and lattice are on different devices");
  if ((rc = closure_allowed(c)) || (rc = ensure_closure_fields(c))) return rc;
  // the model's parameters must be in place before the lattice stream reads them
  HIPCK(hipStreamSynchronize(m->stream));
  c->model = m;
  c->clo_mode = 2;
  return TSLB_OK;
}

int tslb_set_closure_fields(tslb_ctx *c, const float *clo) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!clo) {
    c->model = nullptr;
    c->clo_mode = 0;
    return TSLB_OK;
  }
  if ((rc = closure_allowed(c)) || (rc = ensure_closure_fields(c))) return rc;
  HIPCK(hipMemcpy2DAsync(c->clo, c->fs * sizeof(float), clo, c->n * sizeof(float), c->n * sizeof(float), 7,
                         hipMemcpyHostToDevice, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  c->model = nullptr;
  c->clo_mode = 1;
  return TSLB_OK;
}

int tslb_download_closure(tslb_ctx *c, float *clo) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!clo) return fail(TSLB_E_ARG, "null output");
  if (!c->clo_mode) return fail(TSLB_E_STATE, "no closure attached");
  if (c->clo_mode == 2 && (rc = launch_closure(c))) return rc;
  HIPCK(hipMemcpy2DAsync(clo, c->n * sizeof(float), c->clo, c->fs * sizeof(float), c->n * sizeof(float), 7,
                         hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return TSLB_OK;
}

}  // extern "C"
