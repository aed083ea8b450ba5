// emd_core.hip -- context, device memory, kernel timing and the System image.
#include <mutex>

#include "emd_internal.hpp"

// -------------------------------------------------------------- errors
emd_status emd_set_error(emd_ctx *ctx, emd_status code, const char *fmt, ...) {
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

extern "C" const char *emd_last_error(const emd_ctx *ctx) { return ctx ? ctx->err : "null context"; }
extern "C" const char *emd_version(void) { return "examinimd-amd 0.1 (gfx950)"; }

// ------------------------------------------------------------ scan helper
__global__ void __launch_bounds__(1024) k_scan_blocks(int *bs, int nb, int *total) {
  __shared__ int s_w[16];
  __shared__ int s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  const int wid = threadIdx.x / EMD_WAVE;
  for (int base = 0; base < nb; base += 1024) {
    int i = base + threadIdx.x;
    int v = (i < nb) ? bs[i] : 0;
    int inc = emd_wave_incl_scan(v);
    if (__lane_id() == EMD_WAVE - 1) s_w[wid] = inc;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wid; w++) woff += s_w[w];
    int carry = s_carry;
    if (i < nb) bs[i] = carry + woff + inc - v;
    __syncthreads();
    if (threadIdx.x == 1023) s_carry = carry + woff + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) total[0] = s_carry;
}

// ------------------------------------------------------------- context
extern "C" emd_status emd_create(int device, emd_ctx **out) {
  if (!out) return EMD_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return EMD_ERR_HIP;
  if (device < 0 || device >= ndev) return EMD_ERR_ARG;
  emd_ctx *ctx = new emd_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete ctx; return EMD_ERR_HIP; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return EMD_ERR_HIP;
  }
  bool ok = hipMalloc(&ctx->d_small, 64 * sizeof(int)) == hipSuccess &&
            hipMalloc(&ctx->d_dsmall, 64 * sizeof(double)) == hipSuccess &&
            hipHostMalloc(&ctx->h_small, 64 * sizeof(int), hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc(&ctx->h_dsmall, 64 * sizeof(double), hipHostMallocDefault) == hipSuccess &&
            hipMalloc(&ctx->d_ljtab, EMD_LJ_TABLE_BYTES) == hipSuccess &&
            hipHostMalloc(&ctx->h_ljtab, EMD_LJ_TABLE_BYTES, hipHostMallocDefault) == hipSuccess &&
            hipEventCreateWithFlags(&ctx->ljtab_ev, hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    emd_destroy(ctx);
    return EMD_ERR_ALLOC;
  }
  hipMemset(ctx->d_small, 0, 64 * sizeof(int));
  // the allocation-free entry points' scratch: this device's copy
  if (emd_lj_scratch_setup(ctx) != EMD_OK) {
    emd_destroy(ctx);
    return EMD_ERR_HIP;
  }
  // lanes per atom of the LJ kernel: EMD_LJ_LANES (both precisions) or
  // EMD_LJ_LANES_FP64 / EMD_LJ_LANES_MIXED; 0 or unset = built-in default
  if (const char *e = getenv("EMD_LJ_LANES")) ctx->lj_lanes[0] = ctx->lj_lanes[1] = atoi(e);
  if (const char *e = getenv("EMD_LJ_LANES_FP64")) ctx->lj_lanes[0] = atoi(e);
  if (const char *e = getenv("EMD_LJ_LANES_MIXED")) ctx->lj_lanes[1] = atoi(e);
  for (int p = 0; p < 2; p++) {
    const int k = ctx->lj_lanes[p];
    if (!(k == 0 || k == 1 || k == 2 || k == 4 || k == 8 || k == 16)) ctx->lj_lanes[p] = 0;
  }
  // ghost force fold of a small halo (<= 16384 ghosts) in one launch (1) or
  // one launch per phase (0)
  if (const char *e = getenv("EMD_FOLD_ONE")) ctx->fold_one = atoi(e) != 0;
  // host waits for counts: poll an event (1) or hipStreamSynchronize (0)
  if (const char *e = getenv("EMD_SPIN_WAIT")) ctx->spin_wait = atoi(e) != 0;
  // exchange_halo of a small system (<= 16384 local + ghost atoms) in one
  // launch (1) or as six scans with a host sync each (0)
  if (const char *e = getenv("EMD_HALO_ONE")) ctx->halo_one = atoi(e) != 0;
  // binning of a small range (<= 7168 atoms, <= 4096 bins) in one launch
  if (const char *e = getenv("EMD_BIN_ONE")) ctx->bin_one = atoi(e) != 0;
  if (const char *e = getenv("EMD_LJ_PRIO")) {
    const int v = atoi(e);
    ctx->lj_prio = (v >= 0 && v <= 2) ? v : -1;
  }
  *out = ctx;
  return EMD_OK;
}

extern "C" emd_status emd_destroy(emd_ctx *ctx) {
  if (!ctx) return EMD_OK;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  for (auto &r : ctx->prof_pending) { hipEventDestroy(r.start); hipEventDestroy(r.stop); }
  for (auto e : ctx->prof_pool) hipEventDestroy(e);
  if (ctx->d_small) hipFree(ctx->d_small);
  if (ctx->d_dsmall) hipFree(ctx->d_dsmall);
  if (ctx->h_small) hipHostFree(ctx->h_small);
  if (ctx->h_dsmall) hipHostFree(ctx->h_dsmall);
  if (ctx->d_scan) hipFree(ctx->d_scan);
  if (ctx->d_tmp) hipFree(ctx->d_tmp);
  if (ctx->d_tmp2) hipFree(ctx->d_tmp2);
  if (ctx->d_ljtab) hipFree(ctx->d_ljtab);
  if (ctx->d_ljbig) hipFree(ctx->d_ljbig);
  if (ctx->h_ljtab) hipHostFree(ctx->h_ljtab);
  if (ctx->ljtab_ev) hipEventDestroy(ctx->ljtab_ev);
  if (ctx->wait_ev) hipEventDestroy(ctx->wait_ev);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return EMD_OK;
}

emd_status emd_wait(emd_ctx *ctx) {
  if (!ctx->spin_wait) {
    EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EMD_OK;
  }
  if (!ctx->wait_ev) EMD_HIP(ctx, hipEventCreateWithFlags(&ctx->wait_ev, hipEventDisableTiming));
  EMD_HIP(ctx, hipEventRecord(ctx->wait_ev, ctx->stream));
  hipError_t e;
  while ((e = hipEventQuery(ctx->wait_ev)) == hipErrorNotReady) {
  }
  EMD_HIP(ctx, e);
  return EMD_OK;
}

extern "C" emd_status emd_synchronize(emd_ctx *ctx) {
  if (!ctx) return EMD_ERR_ARG;
  return emd_wait(ctx);
}

extern "C" void *emd_get_stream(emd_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

// --------------------------------------------------------------- memory
emd_status emd_ensure(emd_ctx *ctx, void **p, size_t *cap, size_t bytes) {
  if (bytes <= *cap && *p) return EMD_OK;
  if (*p) {
    EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    EMD_HIP(ctx, hipFree(*p));
  }
  *p = nullptr;
  *cap = 0;
  size_t want = bytes + bytes / 8 + 256;
  if (hipMalloc(p, want) != hipSuccess)
    return emd_set_error(ctx, EMD_ERR_ALLOC, "hipMalloc of %zu bytes failed", want);
  *cap = want;
  return EMD_OK;
}

emd_status emd_ensure_int(emd_ctx *ctx, int **p, int *cap, long count) {
  if (count <= *cap && *p) return EMD_OK;
  size_t c = 0;
  void *q = *p;
  if (q) {
    EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    EMD_HIP(ctx, hipFree(q));
  }
  *p = nullptr;
  *cap = 0;
  long want = count + count / 8 + 64;
  if (hipMalloc(&q, (size_t)want * sizeof(int)) != hipSuccess)
    return emd_set_error(ctx, EMD_ERR_ALLOC, "hipMalloc of %ld ints failed", want);
  (void)c;
  *p = (int *)q;
  *cap = (int)want;
  return EMD_OK;
}

extern "C" emd_status emd_malloc(emd_ctx *ctx, void **ptr, size_t bytes) {
  if (!ctx || !ptr) return EMD_ERR_ARG;
  if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess)
    return emd_set_error(ctx, EMD_ERR_ALLOC, "hipMalloc of %zu bytes failed", bytes);
  return EMD_OK;
}

extern "C" emd_status emd_free(emd_ctx *ctx, void *ptr) {
  if (!ctx) return EMD_ERR_ARG;
  if (ptr) {
    EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    EMD_HIP(ctx, hipFree(ptr));
  }
  return EMD_OK;
}

extern "C" emd_status emd_memcpy_h2d(emd_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return EMD_ERR_ARG;
  if (!bytes) return EMD_OK;
  EMD_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EMD_OK;
}

extern "C" emd_status emd_memcpy_d2h(emd_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return EMD_ERR_ARG;
  if (!bytes) return EMD_OK;
  EMD_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EMD_OK;
}

extern "C" emd_status emd_memcpy_d2d(emd_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return EMD_ERR_ARG;
  if (!bytes) return EMD_OK;
  EMD_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return EMD_OK;
}

// -------------------------------------------------------------- timing
static int prof_id(emd_ctx *ctx, const char *name) {
  auto it = ctx->prof_ids.find(name);
  if (it != ctx->prof_ids.end()) return it->second;
  int id = (int)ctx->prof_names.size();
  ctx->prof_names.push_back(name);
  ctx->prof_ids[name] = id;
  ctx->prof_ms.push_back(0.0);
  ctx->prof_count.push_back(0);
  return id;
}

static hipEvent_t prof_event(emd_ctx *ctx) {
  if (!ctx->prof_pool.empty()) {
    hipEvent_t e = ctx->prof_pool.back();
    ctx->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void emd_prof_begin(emd_ctx *ctx, const char *name, emd_prof_rec *rec) {
  rec->name = -1;
  if (!ctx->profile) return;
  if (!ctx->prof_select.empty()) {
    bool hit = false;
    for (const auto &sel : ctx->prof_select) hit = hit || sel == name;
    if (!hit) return;
  }
  rec->name = prof_id(ctx, name);
  rec->start = prof_event(ctx);
  rec->stop = prof_event(ctx);
  if (rec->start) hipEventRecord(rec->start, ctx->stream);
}

void emd_prof_end(emd_ctx *ctx, emd_prof_rec *rec) {
  if (rec->name < 0 || !rec->start || !rec->stop) return;
  hipEventRecord(rec->stop, ctx->stream);
  ctx->prof_pending.push_back(*rec);
}

static void prof_collect(emd_ctx *ctx) {
  if (ctx->prof_pending.empty()) return;
  hipStreamSynchronize(ctx->stream);
  for (auto &r : ctx->prof_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.start, r.stop) == hipSuccess) {
      ctx->prof_ms[r.name] += ms;
      ctx->prof_count[r.name] += 1;
    }
    ctx->prof_pool.push_back(r.start);
    ctx->prof_pool.push_back(r.stop);
  }
  ctx->prof_pending.clear();
}

extern "C" emd_status emd_profile_enable(emd_ctx *ctx, int on) {
  if (!ctx) return EMD_ERR_ARG;
  prof_collect(ctx);
  ctx->profile = on != 0;
  return EMD_OK;
}

extern "C" emd_status emd_profile_select(emd_ctx *ctx, const char *names) {
  if (!ctx) return EMD_ERR_ARG;
  prof_collect(ctx);
  ctx->prof_select.clear();
  if (!names) return EMD_OK;
  std::string cur;
  for (const char *p = names;; p++) {
    if (*p == ',' || *p == 0) {
      if (!cur.empty()) ctx->prof_select.push_back(cur);
      cur.clear();
      if (*p == 0) break;
    } else if (*p != ' ') {
      cur += *p;
    }
  }
  return EMD_OK;
}

extern "C" emd_status emd_profile_reset(emd_ctx *ctx) {
  if (!ctx) return EMD_ERR_ARG;
  prof_collect(ctx);
  for (auto &m : ctx->prof_ms) m = 0.0;
  for (auto &c : ctx->prof_count) c = 0;
  return EMD_OK;
}

extern "C" emd_status emd_profile_get(emd_ctx *ctx, const char *name, double *ms, long *n) {
  if (!ctx || !name) return EMD_ERR_ARG;
  prof_collect(ctx);
  auto it = ctx->prof_ids.find(name);
  if (ms) *ms = it == ctx->prof_ids.end() ? 0.0 : ctx->prof_ms[it->second];
  if (n) *n = it == ctx->prof_ids.end() ? 0 : ctx->prof_count[it->second];
  return EMD_OK;
}

extern "C" size_t emd_profile_report(emd_ctx *ctx, char *buf, size_t len) {
  if (!ctx) return 0;
  prof_collect(ctx);
  std::string s;
  char line[256];
  for (size_t i = 0; i < ctx->prof_names.size(); i++) {
    snprintf(line, sizeof(line), "%s %ld %.6f\n", ctx->prof_names[i].c_str(), ctx->prof_count[i],
             ctx->prof_ms[i]);
    s += line;
  }
  if (buf && len) {
    size_t n = s.size() < len - 1 ? s.size() : len - 1;
    memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return s.size() + 1;
}

// --------------------------------------------------------------- system
template <class T>
static emd_status grow_one(emd_ctx *ctx, T **p, int old_n, int new_n, int width) {
  T *q = nullptr;
  size_t nb = (size_t)new_n * width * sizeof(T);
  if (hipMalloc(&q, nb ? nb : 1) != hipSuccess)
    return emd_set_error(ctx, EMD_ERR_ALLOC, "system grow: hipMalloc of %zu bytes failed", nb);
  size_t ob = (size_t)old_n * width * sizeof(T);
  if (*p && ob) EMD_HIP(ctx, hipMemcpyAsync(q, *p, ob, hipMemcpyDeviceToDevice, ctx->stream));
  if (nb > ob) EMD_HIP(ctx, hipMemsetAsync((char *)q + ob, 0, nb - ob, ctx->stream));
  if (*p) {
    EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    EMD_HIP(ctx, hipFree(*p));
  }
  *p = q;
  return EMD_OK;
}

extern "C" emd_status emd_system_grow(emd_ctx *ctx, emd_system *s, int n) {
  if (!ctx || !s || n < 0) return EMD_ERR_ARG;
  if (n <= s->n_max && s->x) return EMD_OK;
  int old = s->x ? s->n_max : 0;
  EMD_TRY(grow_one(ctx, &s->x, old, n, 3));
  EMD_TRY(grow_one(ctx, &s->v, old, n, 3));
  EMD_TRY(grow_one(ctx, &s->f, old, n, 3));
  EMD_TRY(grow_one(ctx, &s->f32, old, n, 3));
  EMD_TRY(grow_one(ctx, &s->type, old, n, 1));
  EMD_TRY(grow_one(ctx, &s->id, old, n, 1));
  EMD_TRY(grow_one(ctx, &s->q, old, n, 1));
  if (!s->mass) {
    EMD_HIP(ctx, hipMalloc(&s->mass, EMD_MAX_TYPES * sizeof(double)));
    EMD_HIP(ctx, hipMemsetAsync(s->mass, 0, EMD_MAX_TYPES * sizeof(double), ctx->stream));
  }
  s->n_max = n;
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EMD_OK;
}

extern "C" emd_status emd_system_free(emd_ctx *ctx, emd_system *s) {
  if (!ctx || !s) return EMD_ERR_ARG;
  hipStreamSynchronize(ctx->stream);
  void *ptrs[] = {s->x, s->v, s->f, s->f32, s->type, s->id, s->q, s->mass};
  for (void *p : ptrs)
    if (p) hipFree(p);
  s->x = s->v = s->f = s->q = s->mass = nullptr;
  s->f32 = nullptr;
  s->type = s->id = nullptr;
  s->n_max = 0;
  return EMD_OK;
}

extern "C" emd_status emd_system_upload(emd_ctx *ctx, emd_system *s, int n, const double *x,
                                        const double *v, const int *type, const int *id,
                                        const double *q) {
  if (!ctx || !s || n < 0) return EMD_ERR_ARG;
  EMD_TRY(emd_system_grow(ctx, s, n));
  hipStream_t st = ctx->stream;
  if (x) EMD_HIP(ctx, hipMemcpyAsync(s->x, x, (size_t)n * 24, hipMemcpyHostToDevice, st));
  if (v) EMD_HIP(ctx, hipMemcpyAsync(s->v, v, (size_t)n * 24, hipMemcpyHostToDevice, st));
  if (type) EMD_HIP(ctx, hipMemcpyAsync(s->type, type, (size_t)n * 4, hipMemcpyHostToDevice, st));
  if (id) EMD_HIP(ctx, hipMemcpyAsync(s->id, id, (size_t)n * 4, hipMemcpyHostToDevice, st));
  if (q) EMD_HIP(ctx, hipMemcpyAsync(s->q, q, (size_t)n * 8, hipMemcpyHostToDevice, st));
  EMD_HIP(ctx, hipStreamSynchronize(st));
  return EMD_OK;
}

extern "C" emd_status emd_system_download(emd_ctx *ctx, const emd_system *s, int n, double *x,
                                          double *v, double *f, int *type, int *id, double *q) {
  if (!ctx || !s || n < 0 || n > s->n_max) return EMD_ERR_ARG;
  hipStream_t st = ctx->stream;
  if (x) EMD_HIP(ctx, hipMemcpyAsync(x, s->x, (size_t)n * 24, hipMemcpyDeviceToHost, st));
  if (v) EMD_HIP(ctx, hipMemcpyAsync(v, s->v, (size_t)n * 24, hipMemcpyDeviceToHost, st));
  if (f) EMD_HIP(ctx, hipMemcpyAsync(f, s->f, (size_t)n * 24, hipMemcpyDeviceToHost, st));
  if (type) EMD_HIP(ctx, hipMemcpyAsync(type, s->type, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  if (id) EMD_HIP(ctx, hipMemcpyAsync(id, s->id, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  if (q) EMD_HIP(ctx, hipMemcpyAsync(q, s->q, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  EMD_HIP(ctx, hipStreamSynchronize(st));
  return EMD_OK;
}

// ------------------------------------------------- reference View layouts
// element (i, d) of a [n][3] double array: LayoutRight a[3 i + d],
// LayoutLeft a[i + d ld]
__device__ __forceinline__ size_t vec3_at(int layout, long ld, long i, int d) {
  return layout == EMD_LAYOUT_LEFT ? (size_t)i + (size_t)d * (size_t)ld : 3 * (size_t)i + d;
}

// one thread per element, threads ordered by the destination's storage so
// the stores are coalesced (the loads are strided when the layouts differ)
__global__ void __launch_bounds__(256)
k_vec3_copy(double *__restrict__ dst, int dl, long dld, const double *__restrict__ src, int sl,
            long sld, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  long i;
  int d;
  if (dl == EMD_LAYOUT_LEFT) { d = (int)(t / n); i = t - (long)d * n; }
  else { i = t / 3; d = (int)(t - 3 * i); }
  dst[vec3_at(dl, dld, i, d)] = src[vec3_at(sl, sld, i, d)];
}

static emd_status vec3_copy(emd_ctx *ctx, double *dst, int dl, long dld, const double *src, int sl,
                            long sld, int n) {
  if (n <= 0) return EMD_OK;
  EMD_ARG(ctx, (dl == EMD_LAYOUT_RIGHT || dl == EMD_LAYOUT_LEFT) &&
                   (sl == EMD_LAYOUT_RIGHT || sl == EMD_LAYOUT_LEFT), "unknown layout");
  EMD_ARG(ctx, (dl != EMD_LAYOUT_LEFT || dld >= n) && (sl != EMD_LAYOUT_LEFT || sld >= n),
          "LayoutLeft leading dimension below the atom count");
  if (dl == EMD_LAYOUT_RIGHT && sl == EMD_LAYOUT_RIGHT) {
    EMD_HIP(ctx, hipMemcpyAsync(dst, src, (size_t)n * 24, hipMemcpyDeviceToDevice, ctx->stream));
    return EMD_OK;
  }
  EMD_LAUNCH(ctx, "layout_copy", k_vec3_copy, emd_blocks(3l * n, 256), 256, 0, dst, dl, dld, src,
             sl, sld, (long)n);
  return EMD_OK;
}

extern "C" emd_status emd_system_import(emd_ctx *ctx, emd_system *s, int n, emd_vec3_view x,
                                        emd_vec3_view v, emd_vec3_view f, const int *type,
                                        const int *id, const double *q) {
  if (!ctx || !s || n < 0) return EMD_ERR_ARG;
  EMD_TRY(emd_system_grow(ctx, s, n));
  hipStream_t st = ctx->stream;
  if (x.data) EMD_TRY(vec3_copy(ctx, s->x, EMD_LAYOUT_RIGHT, 0, x.data, x.layout, x.ld, n));
  if (v.data) EMD_TRY(vec3_copy(ctx, s->v, EMD_LAYOUT_RIGHT, 0, v.data, v.layout, v.ld, n));
  if (f.data) EMD_TRY(vec3_copy(ctx, s->f, EMD_LAYOUT_RIGHT, 0, f.data, f.layout, f.ld, n));
  if (type) EMD_HIP(ctx, hipMemcpyAsync(s->type, type, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  if (id) EMD_HIP(ctx, hipMemcpyAsync(s->id, id, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  if (q) EMD_HIP(ctx, hipMemcpyAsync(s->q, q, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
  return EMD_OK;
}

extern "C" emd_status emd_system_export(emd_ctx *ctx, const emd_system *s, int n,
                                        emd_vec3_view x, emd_vec3_view v, emd_vec3_view f) {
  if (!ctx || !s || n < 0 || n > s->n_max) return EMD_ERR_ARG;
  if (x.data) EMD_TRY(vec3_copy(ctx, x.data, x.layout, x.ld, s->x, EMD_LAYOUT_RIGHT, 0, n));
  if (v.data) EMD_TRY(vec3_copy(ctx, v.data, v.layout, v.ld, s->v, EMD_LAYOUT_RIGHT, 0, n));
  if (f.data) EMD_TRY(vec3_copy(ctx, f.data, f.layout, f.ld, s->f, EMD_LAYOUT_RIGHT, 0, n));
  return EMD_OK;
}

// one context per device per process, created on first use
extern "C" emd_status emd_ctx_shared(int device, emd_ctx **out) {
  static std::mutex lock;
  static emd_ctx *shared[64] = {nullptr};
  if (!out || device < 0 || device >= 64) return EMD_ERR_ARG;
  std::lock_guard<std::mutex> g(lock);
  if (!shared[device]) {
    const emd_status st = emd_create(device, &shared[device]);
    if (st != EMD_OK) return st;
  }
  *out = shared[device];
  return EMD_OK;
}

extern "C" emd_status emd_system_set_mass(emd_ctx *ctx, emd_system *s, int ntypes,
                                          const double *mass) {
  if (!ctx || !s || ntypes < 1 || ntypes > EMD_MAX_TYPES || !mass) return EMD_ERR_ARG;
  if (!s->mass) EMD_HIP(ctx, hipMalloc(&s->mass, EMD_MAX_TYPES * sizeof(double)));
  EMD_HIP(ctx, hipMemcpyAsync(s->mass, mass, ntypes * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  s->ntypes = ntypes;
  return EMD_OK;
}

emd_status emd_grow_keep_int(emd_ctx *ctx, int **p, int *cap, long count, long keep) {
  if (count <= *cap && *p) return EMD_OK;
  long want = count + count / 4 + 64;
  int *q = nullptr;
  if (hipMalloc(&q, (size_t)want * sizeof(int)) != hipSuccess)
    return emd_set_error(ctx, EMD_ERR_ALLOC, "hipMalloc of %ld ints failed", want);
  if (*p && keep > 0)
    EMD_HIP(ctx, hipMemcpyAsync(q, *p, (size_t)keep * sizeof(int), hipMemcpyDeviceToDevice,
                                ctx->stream));
  if (*p) {
    EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    EMD_HIP(ctx, hipFree(*p));
  }
  *p = q;
  *cap = (int)want;
  return EMD_OK;
}
