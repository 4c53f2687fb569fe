// Context, device buffers, scene containers, profiling and error reporting.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <atomic>

#include "nts_internal.h"

static thread_local char g_err[1024] = "";

int nts_fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char* nts_last_error(void) { return g_err; }
extern "C" int nts_version(void) { return 1; }

int nts_dev_alloc(nts_ctx* ctx, void** p, uint64_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  if (ctx->hbm_budget && ctx->allocated + bytes > ctx->hbm_budget)
    return nts_fail(NTS_E_NOMEM, "HBM budget exceeded: %llu + %llu > %llu", (unsigned long long)ctx->allocated,
                    (unsigned long long)bytes, (unsigned long long)ctx->hbm_budget);
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return nts_fail(NTS_E_NOMEM, "hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
  }
  ctx->allocated += bytes;
  return NTS_OK;
}

void nts_dev_free(nts_ctx* ctx, void* p, uint64_t bytes) {
  if (!p) return;
  if (bytes == 0) bytes = 256;
  (void)hipFree(p);
  ctx->allocated -= bytes;
}

int nts_scratch(nts_ctx* ctx, uint64_t bytes, void** out) {
  if (bytes > ctx->scratch_bytes) {
    if (ctx->scratch) {
      NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
      nts_dev_free(ctx, ctx->scratch, ctx->scratch_bytes);
      ctx->scratch = nullptr;
      ctx->scratch_bytes = 0;
    }
    uint64_t want = bytes + bytes / 4 + (1u << 20);
    int rc = nts_dev_alloc(ctx, &ctx->scratch, want);
    if (rc) return rc;
    ctx->scratch_bytes = want;
  }
  *out = ctx->scratch;
  return NTS_OK;
}

// ---- profiling ---------------------------------------------------------------
int nts_prof_begin(nts_ctx* ctx, const char* name, PendingEvent* pe) {
  if (!ctx->prof) return 0;
  int idx = -1;
  for (size_t i = 0; i < ctx->prof_entries.size(); ++i)
    if (ctx->prof_entries[i].name == name) { idx = (int)i; break; }
  if (idx < 0) {
    ctx->prof_entries.push_back(ProfEntry{name, 0, 0.0});
    idx = (int)ctx->prof_entries.size() - 1;
  }
  pe->entry = idx;
  for (int k = 0; k < 2; ++k) {
    hipEvent_t ev;
    if (!ctx->event_pool.empty()) {
      ev = ctx->event_pool.back();
      ctx->event_pool.pop_back();
    } else if (hipEventCreate(&ev) != hipSuccess) {
      return 0;
    }
    (k == 0 ? pe->start : pe->stop) = ev;
  }
  (void)hipEventRecord(pe->start, ctx->stream);
  return 1;
}

void nts_prof_end(nts_ctx* ctx, PendingEvent* pe) {
  (void)hipEventRecord(pe->stop, ctx->stream);
  ctx->pending.push_back(*pe);
  if (ctx->pending.size() > 4096) nts_prof_resolve(ctx);
}

int nts_prof_resolve(nts_ctx* ctx) {
  if (ctx->pending.empty()) return NTS_OK;
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (auto& pe : ctx->pending) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, pe.start, pe.stop);
    ctx->prof_entries[pe.entry].launches += 1;
    ctx->prof_entries[pe.entry].total_ms += ms;
    ctx->event_pool.push_back(pe.start);
    ctx->event_pool.push_back(pe.stop);
  }
  ctx->pending.clear();
  return NTS_OK;
}

extern "C" int nts_prof_enable(nts_ctx* ctx, int on) {
  NTS_CHECK_ARG(ctx);
  if (!on) {
    int rc = nts_prof_resolve(ctx);
    if (rc) return rc;
  }
  ctx->prof = on != 0;
  return NTS_OK;
}

extern "C" int nts_prof_reset(nts_ctx* ctx) {
  NTS_CHECK_ARG(ctx);
  int rc = nts_prof_resolve(ctx);
  if (rc) return rc;
  ctx->prof_entries.clear();
  return NTS_OK;
}

extern "C" int nts_prof_query(nts_ctx* ctx, const char* prefix, double* total_ms, uint64_t* launches) {
  NTS_CHECK_ARG(ctx && prefix);
  int rc = nts_prof_resolve(ctx);
  if (rc) return rc;
  double t = 0.0;
  uint64_t l = 0;
  const size_t pl = strlen(prefix);
  for (auto& e : ctx->prof_entries)
    if (e.name.compare(0, pl, prefix) == 0) {
      t += e.total_ms;
      l += e.launches;
    }
  if (total_ms) *total_ms = t;
  if (launches) *launches = l;
  return NTS_OK;
}

extern "C" int nts_prof_report(nts_ctx* ctx, char* buf, size_t cap) {
  NTS_CHECK_ARG(ctx && buf && cap > 0);
  int rc = nts_prof_resolve(ctx);
  if (rc) return rc;
  size_t off = 0;
  buf[0] = 0;
  for (auto& e : ctx->prof_entries) {
    int w = snprintf(buf + off, cap - off, "%s %llu %.6f\n", e.name.c_str(), (unsigned long long)e.launches,
                     e.total_ms);
    if (w < 0 || (size_t)w >= cap - off) break;
    off += (size_t)w;
  }
  return NTS_OK;
}

// ---- context -----------------------------------------------------------------
extern "C" int nts_ctx_create(int device, uint64_t hbm_budget_bytes, nts_ctx** out) {
  NTS_CHECK_ARG(out);
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return nts_fail(NTS_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  }
  if (device < 0 || device >= ndev) return nts_fail(NTS_E_INVALID, "device %d out of range (%d)", device, ndev);
  NTS_HIP_TRY(hipSetDevice(device));
  nts_ctx* ctx = new nts_ctx();
  ctx->device = device;
  ctx->hbm_budget = hbm_budget_bytes;
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return nts_fail(NTS_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  e = hipHostMalloc((void**)&ctx->host_word, 64, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return nts_fail(NTS_E_HIP, "hipHostMalloc: %s", hipGetErrorString(e));
  }
  *out = ctx;
  return NTS_OK;
}

extern "C" int nts_ctx_destroy(nts_ctx* ctx) {
  if (!ctx) return NTS_OK;
  (void)hipStreamSynchronize(ctx->stream);
  nts_prof_resolve(ctx);
  for (auto ev : ctx->event_pool) (void)hipEventDestroy(ev);
  if (ctx->scratch) nts_dev_free(ctx, ctx->scratch, ctx->scratch_bytes);
  if (ctx->host_word) (void)hipHostFree(ctx->host_word);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return NTS_OK;
}

extern "C" int nts_sync(nts_ctx* ctx) {
  NTS_CHECK_ARG(ctx);
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return NTS_OK;
}

// ---- fill ------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t value, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = value;
}

int nts_fill_u32(nts_ctx* ctx, void* p, uint32_t value, uint64_t n_words) {
  if (n_words == 0) return NTS_OK;
  uint32_t grid = nts_div_up(n_words, 256);
  if (grid > 4096) grid = 4096;
  NTS_LAUNCH(ctx, "fill_u32", k_fill_u32, dim3(grid), dim3(256), 0, (uint32_t*)p, value, n_words);
  return NTS_OK;
}

// ---- buffers -----------------------------------------------------------------------
extern "C" int nts_buffer_create(nts_ctx* ctx, uint64_t bytes, nts_buffer** out) {
  NTS_CHECK_ARG(ctx && out);
  *out = nullptr;
  nts_buffer* b = new nts_buffer();
  const uint64_t padded = (bytes + 255) & ~255ull;
  int rc = nts_dev_alloc(ctx, &b->ptr, padded);
  if (rc) {
    delete b;
    return rc;
  }
  b->bytes = bytes;
  rc = nts_fill_u32(ctx, b->ptr, 0, padded / 4);
  if (rc) return rc;
  *out = b;
  return NTS_OK;
}

extern "C" int nts_buffer_destroy(nts_ctx* ctx, nts_buffer* buf) {
  if (!ctx || !buf) return NTS_OK;
  (void)hipStreamSynchronize(ctx->stream);
  nts_dev_free(ctx, buf->ptr, (buf->bytes + 255) & ~255ull);
  delete buf;
  return NTS_OK;
}

extern "C" uint64_t nts_buffer_size(const nts_buffer* buf) { return buf ? buf->bytes : 0; }

extern "C" int nts_buffer_upload(nts_ctx* ctx, nts_buffer* buf, uint64_t offset, const void* host, uint64_t bytes) {
  NTS_CHECK_ARG(ctx && buf && (host || bytes == 0));
  NTS_CHECK_ARG(offset + bytes <= buf->bytes);
  if (bytes == 0) return NTS_OK;
  NTS_HIP_TRY(hipMemcpyAsync((char*)buf->ptr + offset, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return NTS_OK;
}

extern "C" int nts_buffer_download(nts_ctx* ctx, const nts_buffer* buf, uint64_t offset, void* host, uint64_t bytes) {
  NTS_CHECK_ARG(ctx && buf && (host || bytes == 0));
  NTS_CHECK_ARG(offset + bytes <= buf->bytes);
  if (bytes == 0) return NTS_OK;
  NTS_HIP_TRY(hipMemcpyAsync(host, (const char*)buf->ptr + offset, bytes, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return NTS_OK;
}

extern "C" int nts_buffer_zero(nts_ctx* ctx, nts_buffer* buf) {
  NTS_CHECK_ARG(ctx && buf);
  return nts_fill_u32(ctx, buf->ptr, 0, ((buf->bytes + 255) & ~255ull) / 4);
}

// ---- scene ---------------------------------------------------------------------------
static const int kAttrPlane[6] = {P_POS, P_ROT, P_SCALE, P_LOGIT, P_DC, P_REST};
static const int kAttrWidth[6] = {3, 4, 3, 1, 3, 45};

uint64_t nts_scene_next_gen() {
  static std::atomic<uint64_t> next{1};
  return next.fetch_add(1);
}

extern "C" int nts_scene_create(nts_ctx* ctx, uint32_t capacity, uint32_t flags, nts_scene** out) {
  NTS_CHECK_ARG(ctx && out && capacity > 0 && (flags & ~NTS_SCENE_TRAINABLE) == 0);
  *out = nullptr;
  if (!(flags & NTS_SCENE_TRAINABLE)) return nts_scene_alloc_params(ctx, capacity, out);
  nts_scene* s = new nts_scene();
  // pad each plane to a multiple of 64 floats (256 B) so every plane starts aligned
  const uint64_t cap = ((uint64_t)capacity + 63) & ~63ull;
  const uint64_t plane = cap * 4;
  s->bytes = plane * (4ull * P_COUNT + 1);
  int rc = nts_dev_alloc(ctx, &s->base, s->bytes);
  if (rc) {
    delete s;
    return rc;
  }
  s->cap = (uint32_t)cap;
  s->gen = nts_scene_next_gen();
  s->trainable = true;
  s->params = (float*)s->base;
  s->grads = s->params + cap * P_COUNT;
  s->m = s->grads + cap * P_COUNT;
  s->v = s->m + cap * P_COUNT;
  s->grad_acc = s->v + cap * P_COUNT;
  rc = nts_fill_u32(ctx, s->base, 0, s->bytes / 4);
  if (rc) {
    nts_dev_free(ctx, s->base, s->bytes);
    delete s;
    return rc;
  }
  *out = s;
  return NTS_OK;
}

// A scene of parameter planes only (no gradients, moments or grad_acc), for
// render-only residents.
int nts_scene_alloc_params(nts_ctx* ctx, uint32_t capacity, nts_scene** out) {
  *out = nullptr;
  nts_scene* s = new nts_scene();
  const uint64_t cap = ((uint64_t)(capacity ? capacity : 1) + 63) & ~63ull;
  s->bytes = cap * 4 * P_COUNT;
  int rc = nts_dev_alloc(ctx, &s->base, s->bytes);
  if (rc) {
    delete s;
    return rc;
  }
  s->cap = (uint32_t)cap;
  s->gen = nts_scene_next_gen();
  s->trainable = false;
  s->params = (float*)s->base;
  *out = s;
  return NTS_OK;
}

extern "C" int nts_scene_destroy(nts_ctx* ctx, nts_scene* scene) {
  if (!ctx || !scene) return NTS_OK;
  (void)hipStreamSynchronize(ctx->stream);
  nts_dev_free(ctx, scene->base, scene->bytes);
  delete scene;
  return NTS_OK;
}

extern "C" uint32_t nts_scene_count(const nts_scene* scene) { return scene ? scene->n : 0; }

// Host SoA (per-splat interleaved attributes) <-> device planes.
int nts_upload_planes(nts_ctx* ctx, const nts_scene* s, float* dst_planes, const nts_splats_host* h) {
  const uint32_t n = h->n;
  const float* src[6] = {h->pos, h->rot, h->log_scale, h->logit, h->sh_dc, h->sh_rest};
  std::vector<float> staging((size_t)P_COUNT * n);
  for (int a = 0; a < 6; ++a) {
    if (!src[a]) return nts_fail(NTS_E_INVALID, "attribute %d pointer is NULL", a);
    const int w = kAttrWidth[a];
    for (int k = 0; k < w; ++k) {
      float* dst = staging.data() + (size_t)(kAttrPlane[a] + k) * n;
      const float* sp = src[a] + k;
      for (uint32_t i = 0; i < n; ++i) dst[i] = sp[(size_t)i * w];
    }
  }
  for (int p = 0; p < P_COUNT; ++p)
    NTS_HIP_TRY(hipMemcpyAsync(dst_planes + (size_t)p * s->cap, staging.data() + (size_t)p * n, (size_t)n * 4,
                               hipMemcpyHostToDevice, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return NTS_OK;
}

static int download_planes(nts_ctx* ctx, const nts_scene* s, const float* src_planes, nts_splats_host* h) {
  const uint32_t n = s->n;
  std::vector<float> staging((size_t)P_COUNT * n);
  for (int p = 0; p < P_COUNT; ++p)
    NTS_HIP_TRY(hipMemcpyAsync(staging.data() + (size_t)p * n, src_planes + (size_t)p * s->cap, (size_t)n * 4,
                               hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  float* dst[6] = {h->pos, h->rot, h->log_scale, h->logit, h->sh_dc, h->sh_rest};
  for (int a = 0; a < 6; ++a) {
    if (!dst[a]) return nts_fail(NTS_E_INVALID, "attribute %d pointer is NULL", a);
    const int w = kAttrWidth[a];
    for (int k = 0; k < w; ++k) {
      const float* sp = staging.data() + (size_t)(kAttrPlane[a] + k) * n;
      float* dp = dst[a] + k;
      for (uint32_t i = 0; i < n; ++i) dp[(size_t)i * w] = sp[i];
    }
  }
  h->n = n;
  return NTS_OK;
}

extern "C" int nts_scene_upload(nts_ctx* ctx, nts_scene* scene, const nts_splats_host* host) {
  NTS_CHECK_ARG(ctx && scene && host);
  if (host->n > scene->cap)
    return nts_fail(NTS_E_CAPACITY, "scene capacity %u < %u splats", scene->cap, host->n);
  if (scene->trainable) {
    int rc = nts_fill_u32(ctx, scene->grads, 0, (uint64_t)scene->cap * (3 * P_COUNT + 1));
    if (rc) return rc;
  }
  scene->n = host->n;
  return nts_upload_planes(ctx, scene, scene->params, host);
}

extern "C" int nts_scene_sync_params(nts_ctx* ctx, nts_scene* scene, const nts_splats_host* host) {
  NTS_CHECK_ARG(ctx && scene && host);
  if (host->n != scene->n)
    return nts_fail(NTS_E_INVALID, "sync_params needs the scene's splat count (%u), got %u", scene->n, host->n);
  return nts_upload_planes(ctx, scene, scene->params, host);
}

extern "C" int nts_scene_download(nts_ctx* ctx, const nts_scene* scene, nts_splats_host* host) {
  NTS_CHECK_ARG(ctx && scene && host);
  return download_planes(ctx, scene, scene->params, host);
}

extern "C" int nts_scene_download_grads(nts_ctx* ctx, const nts_scene* scene, nts_splats_host* grads) {
  NTS_CHECK_ARG(ctx && scene && grads);
  NTS_NEED_TRAINABLE(scene);
  return download_planes(ctx, scene, scene->grads, grads);
}

extern "C" int nts_scene_upload_grads(nts_ctx* ctx, nts_scene* scene, const nts_splats_host* grads) {
  NTS_CHECK_ARG(ctx && scene && grads && grads->n == scene->n);
  NTS_NEED_TRAINABLE(scene);
  return nts_upload_planes(ctx, scene, scene->grads, grads);
}

extern "C" int nts_scene_zero_grads(nts_ctx* ctx, nts_scene* scene) {
  NTS_CHECK_ARG(ctx && scene);
  NTS_NEED_TRAINABLE(scene);
  return nts_fill_u32(ctx, scene->grads, 0, (uint64_t)scene->cap * P_COUNT);
}

extern "C" int nts_scene_download_moments(nts_ctx* ctx, const nts_scene* scene, nts_splats_host* m,
                                          nts_splats_host* v) {
  NTS_CHECK_ARG(ctx && scene && m && v);
  NTS_NEED_TRAINABLE(scene);
  int rc = download_planes(ctx, scene, scene->m, m);
  if (rc) return rc;
  return download_planes(ctx, scene, scene->v, v);
}

extern "C" int nts_scene_upload_moments(nts_ctx* ctx, nts_scene* scene, const nts_splats_host* m,
                                        const nts_splats_host* v) {
  NTS_CHECK_ARG(ctx && scene && m && v && m->n == scene->n && v->n == scene->n);
  NTS_NEED_TRAINABLE(scene);
  int rc = nts_upload_planes(ctx, scene, scene->m, m);
  if (rc) return rc;
  return nts_upload_planes(ctx, scene, scene->v, v);
}

extern "C" int nts_scene_download_grad_acc(nts_ctx* ctx, const nts_scene* scene, float* host) {
  NTS_CHECK_ARG(ctx && scene && host);
  NTS_NEED_TRAINABLE(scene);
  if (scene->n == 0) return NTS_OK;
  NTS_HIP_TRY(hipMemcpyAsync(host, scene->grad_acc, (size_t)scene->n * 4, hipMemcpyDeviceToHost, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return NTS_OK;
}

extern "C" int nts_scene_upload_grad_acc(nts_ctx* ctx, nts_scene* scene, const float* host) {
  NTS_CHECK_ARG(ctx && scene && host);
  NTS_NEED_TRAINABLE(scene);
  if (scene->n == 0) return NTS_OK;
  NTS_HIP_TRY(hipMemcpyAsync(scene->grad_acc, host, (size_t)scene->n * 4, hipMemcpyHostToDevice, ctx->stream));
  NTS_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return NTS_OK;
}

// ---- camera ----------------------------------------------------------------------------
// raster.rs:79-99 with glam's right-handed look-at: rows s, u, -f of the
// view rotation; translation (-eye.s, -eye.u, eye.f); focal H / (2 tan(fov/2)).
extern "C" int nts_camera_from_pose(const float eye[3], const float target[3], const float up[3], float fov_y,
                                    uint32_t width, uint32_t height, uint32_t n_splats, nts_camera* out) {
  NTS_CHECK_ARG(eye && target && up && out && width > 0 && height > 0);
  float f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
  float il = 1.0f / sqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (int i = 0; i < 3; ++i) f[i] = f[i] * il;
  float s[3] = {f[1] * up[2] - up[1] * f[2], f[2] * up[0] - up[2] * f[0], f[0] * up[1] - up[0] * f[1]};
  il = 1.0f / sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  for (int i = 0; i < 3; ++i) s[i] = s[i] * il;
  const float u[3] = {s[1] * f[2] - f[1] * s[2], s[2] * f[0] - f[2] * s[0], s[0] * f[1] - f[0] * s[1]};
  for (int c = 0; c < 3; ++c) {
    out->view[c * 4 + 0] = s[c];
    out->view[c * 4 + 1] = u[c];
    out->view[c * 4 + 2] = -f[c];
    out->view[c * 4 + 3] = 0.0f;
  }
  out->view[12] = -(eye[0] * s[0] + eye[1] * s[1] + eye[2] * s[2]);
  out->view[13] = -(eye[0] * u[0] + eye[1] * u[1] + eye[2] * u[2]);
  out->view[14] = eye[0] * f[0] + eye[1] * f[1] + eye[2] * f[2];
  out->view[15] = 1.0f;
  for (int i = 0; i < 3; ++i) out->cam_pos[i] = eye[i];
  out->near_ = 0.01f;
  out->focal = (float)height / (2.0f * tanf(fov_y * 0.5f));
  out->width = (float)width;
  out->height = (float)height;
  out->n_splats = n_splats;
  return NTS_OK;
}
