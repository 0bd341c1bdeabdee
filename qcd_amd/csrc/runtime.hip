// runtime.hip -- context, errors, lattices, fields, host<->device transfer,
// precision conversion and the synthetic-input generator.
#include "internal.h"
#include "generator.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <map>
#include <thread>

namespace qcdhip {

static thread_local char g_err[1024] = "";
static Context g_ctx;

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
  set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
  return 2;
}

Context *ctx() { return &g_ctx; }

std::recursive_mutex &state_mutex() {
  static std::recursive_mutex m;
  return m;
}

int ensure_init() {
  if (g_ctx.stream) return 0;
  return qcd_hip_init(-1) ? 3 : 0;
}

// Caching allocator: freed blocks are kept (up to kPoolCap bytes) and reused
// for requests of the same size.  hipFree synchronises the whole device, so
// freeing inside a serving loop would serialise concurrent requests.  When
// the pool is full the least recently freed blocks are returned to the
// runtime, so a workload that changes shapes keeps the blocks it now uses.
constexpr size_t kPoolCap = (size_t)64 << 30;
struct PoolBlock {
  void *p;
  uint64_t tick;   // when it was freed
};
static std::multimap<size_t, PoolBlock> g_pool;
static size_t g_pool_bytes = 0;
static uint64_t g_pool_tick = 0;

int dev_alloc(void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return 0;
  {
    std::lock_guard<std::recursive_mutex> lock(state_mutex());
    auto it = g_pool.find(bytes);
    if (it != g_pool.end()) {
      *p = it->second.p;
      g_pool.erase(it);
      g_pool_bytes -= bytes;
      g_ctx.bytes_in_use += bytes;
      return 0;
    }
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    // give the cached blocks back to the runtime and try once more
    qcd_hip_release_cache();
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error("out of device memory: requested %zu bytes with %zu held by the library", bytes,
              g_ctx.bytes_in_use.load());
    *p = nullptr;
    return 4;
  }
  g_ctx.bytes_in_use += bytes;
  return 0;
}

void dev_free(void *p, size_t bytes) {
  if (!p) return;
  g_ctx.bytes_in_use -= bytes;
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  if (bytes > kPoolCap) {
    (void)hipFree(p);
    return;
  }
  while (g_pool_bytes + bytes > kPoolCap && !g_pool.empty()) {   // evict the oldest blocks
    auto old = g_pool.begin();
    for (auto it = g_pool.begin(); it != g_pool.end(); ++it)
      if (it->second.tick < old->second.tick) old = it;
    (void)hipFree(old->second.p);
    g_pool_bytes -= old->first;
    g_pool.erase(old);
  }
  g_pool.emplace(bytes, PoolBlock{p, ++g_pool_tick});
  g_pool_bytes += bytes;
}

// Device staging buffer for AoS transfers, grown on demand.
static void *g_stage = nullptr;
static size_t g_stage_bytes = 0;

static int stage(size_t bytes, void **out) {
  if (bytes > g_stage_bytes) {
    if (g_stage) QH_CHECK(hipStreamSynchronize(g_ctx.stream));
    dev_free(g_stage, g_stage_bytes);
    g_stage = nullptr;
    g_stage_bytes = 0;
    QH_TRY(dev_alloc(&g_stage, bytes));
    g_stage_bytes = bytes;
  }
  *out = g_stage;
  return 0;
}

// ------------------------------------------------------------ scratch fields
// Library-owned spinors reused across calls (reallocated when the shape
// changes), so that entry points needing temporaries do not allocate per call.
// A slot outlives the lattices it is used with, so it records the shape it
// was allocated for (plane stride, precision, rhs count) and never looks at
// the lattice it was last lent to, which the caller may have destroyed since.
struct ScratchSlot {
  qcd_hip_field f = nullptr;
  long stride = 0;
  int prec = -1;
  int nrhs = 0;
};
static ScratchSlot g_scratch[8];

int scratch_spinor(int slot, qcd_hip_lattice lat, int prec, int nrhs, qcd_hip_field *out) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  if (slot < 0 || slot >= 8) {
    set_error("scratch slot %d out of range", slot);
    return 1;
  }
  ScratchSlot &sc = g_scratch[slot];
  if (sc.f && (sc.prec != prec || sc.nrhs < nrhs || sc.stride != lat->stride)) {
    qcd_hip_field_destroy(sc.f);
    sc = ScratchSlot();
  }
  if (!sc.f) {
    QH_TRY(qcd_hip_spinor_create(&sc.f, lat, prec, nrhs));
    sc.stride = lat->stride;
    sc.prec = prec;
    sc.nrhs = nrhs;
  }
  sc.f->lat = lat;   // lent for this call; the caller's lattice is alive while it uses the field
  *out = sc.f;
  return 0;
}

// ------------------------------------------------------------ layout kernels
// index of component k of site s in a field of element type T: NC = 12
// spinor, NC = 36 gauge field
template <typename T>
__device__ __forceinline__ size_t elem(int nc, int k, long s, long stride) {
  return nc == 36 ? gauge_index_e<(int)sizeof(T)>(k, (size_t)s, (size_t)stride)
                  : tile_index_e<(int)sizeof(T)>(nc, k, (size_t)s);
}


// AoS site-major (site, k) -> planes (k, site) with precision change.
template <typename Dst>
__device__ __forceinline__ Dst cvt2(double2 v);
template <>
__device__ __forceinline__ double2 cvt2<double2>(double2 v) { return v; }
template <>
__device__ __forceinline__ float2 cvt2<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }
template <>
__device__ __forceinline__ __half2 cvt2<__half2>(double2 v) { return __floats2half2_rn((float)v.x, (float)v.y); }

template <typename Dst, int NC>
__global__ void __launch_bounds__(kBlock) k_aos_to_planes(const double2 *__restrict__ aos, Dst *__restrict__ planes,
                                                          long V, long stride) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
#pragma unroll
  for (int k = 0; k < NC; k++) planes[elem<Dst>(NC, k, s, stride)] = cvt2<Dst>(aos[s * NC + k]);
}

template <typename Src, int NC>
__global__ void __launch_bounds__(kBlock) k_planes_to_aos(const Src *__restrict__ planes, double2 *__restrict__ aos,
                                                          long V, long stride) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
#pragma unroll
  for (int k = 0; k < NC; k++) {
    Src v = planes[elem<Src>(NC, k, s, stride)];
    double2 o;
    if constexpr (sizeof(Src) == 16) {
      o = *(const double2 *)&v;
    } else if constexpr (sizeof(Src) == 8) {
      float2 f = *(const float2 *)&v;
      o = make_double2(f.x, f.y);
    } else {
      float2 f = __half22float2(*(const __half2 *)&v);
      o = make_double2(f.x, f.y);
    }
    aos[s * NC + k] = o;
  }
}

// Half-precision spinor with a per-site scale: norm = max |real| of the site.
__global__ void __launch_bounds__(kBlock) k_aos_to_half_spinor(const double2 *__restrict__ aos, __half2 *__restrict__ planes,
                                                               float *__restrict__ norm, long V, long stride) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  double mx = 0.0;
  double2 v[12];
#pragma unroll
  for (int k = 0; k < 12; k++) {
    v[k] = aos[s * 12 + k];
    mx = fmax(mx, fmax(fabs(v[k].x), fabs(v[k].y)));
  }
  float n = (float)mx;
  if (!(n >= FLT_MIN)) n = 0.f;   // 1/n must stay finite
  float inv = n > 0.f ? 1.0f / n : 0.f;
  norm[s] = n;
#pragma unroll
  for (int k = 0; k < 12; k++) planes[tile_index_e<4>(12, k, s)] = __floats2half2_rn((float)v[k].x * inv, (float)v[k].y * inv);
}

__global__ void __launch_bounds__(kBlock) k_half_spinor_to_aos(const __half2 *__restrict__ planes, const float *__restrict__ norm,
                                                               double2 *__restrict__ aos, long V, long stride) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  float n = norm[s];
#pragma unroll
  for (int k = 0; k < 12; k++) {
    float2 f = __half22float2(planes[tile_index_e<4>(12, k, s)]);
    aos[s * 12 + k] = make_double2((double)f.x * n, (double)f.y * n);
  }
}

// plane-to-plane precision conversion; F16 spinors carry the site norm.
template <typename Src>
__device__ __forceinline__ double2 load_as_double(const Src *p, long i, const float *norm, long s) {
  if constexpr (sizeof(Src) == 16) {
    return *(const double2 *)&p[i];
  } else if constexpr (sizeof(Src) == 8) {
    float2 f = *(const float2 *)&p[i];
    return make_double2(f.x, f.y);
  } else {
    float2 f = __half22float2(*(const __half2 *)&p[i]);
    float n = norm ? norm[s] : 1.0f;
    return make_double2((double)f.x * n, (double)f.y * n);
  }
}

template <typename Src, typename Dst, int NC>
__global__ void __launch_bounds__(kBlock) k_convert_planes(const Src *__restrict__ src, const float *__restrict__ snorm,
                                                           Dst *__restrict__ dst, float *__restrict__ dnorm,
                                                           long V, long stride) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  double2 v[NC];
  double mx = 0.0;
#pragma unroll
  for (int k = 0; k < NC; k++) {
    v[k] = load_as_double(src, elem<Src>(NC, k, s, stride), snorm, s);
    mx = fmax(mx, fmax(fabs(v[k].x), fabs(v[k].y)));
  }
  double inv = 1.0;
  if (dnorm) {
    float n = (float)mx;
    if (!(n >= FLT_MIN)) n = 0.f;   // 1/n must stay finite
    dnorm[s] = n;
    inv = n > 0.f ? 1.0 / (double)n : 0.0;
  }
#pragma unroll
  for (int k = 0; k < NC; k++) dst[elem<Dst>(NC, k, s, stride)] = cvt2<Dst>(make_double2(v[k].x * inv, v[k].y * inv));
}

}  // namespace qcdhip

using namespace qcdhip;

// ================================================================ C ABI
extern "C" {

const char *qcd_hip_last_error(void) { return g_err; }

int qcd_hip_init(int device) {
  if (g_ctx.stream) {
    if (device < 0 || device == g_ctx.device) return 0;
    set_error("qcd_hip_init: already initialised on device %d", g_ctx.device);
    return 1;
  }
  int n = 0;
  QH_CHECK(hipGetDeviceCount(&n));
  if (n <= 0) {
    set_error("qcd_hip_init: no HIP device visible");
    return 1;
  }
  if (device < 0) QH_CHECK(hipGetDevice(&device));
  if (device >= n) {
    set_error("qcd_hip_init: device %d out of range (%d devices)", device, n);
    return 1;
  }
  QH_CHECK(hipSetDevice(device));
  hipDeviceProp_t prop;
  QH_CHECK(hipGetDeviceProperties(&prop, device));
  g_ctx.device = device;
  g_ctx.num_cu = prop.multiProcessorCount;
  QH_CHECK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  size_t pbytes = sizeof(double) * kMaxReduceBlocks * 2 * (kMaxKrylov + 4);
  QH_TRY(dev_alloc((void **)&g_ctx.partials, pbytes));
  QH_TRY(dev_alloc((void **)&g_ctx.scalars, sizeof(double) * 4096));
  QH_CHECK(hipHostMalloc((void **)&g_ctx.host_scalars, sizeof(double) * 4096, hipHostMallocDefault));
  return 0;
}

int qcd_hip_device_name(char *buf, int len) {
  QH_TRY(ensure_init());
  hipDeviceProp_t prop;
  QH_CHECK(hipGetDeviceProperties(&prop, g_ctx.device));
  snprintf(buf, (size_t)len, "%s (%s, %d CUs, %.1f GB)", prop.name, prop.gcnArchName, prop.multiProcessorCount,
           prop.totalGlobalMem / 1e9);
  return 0;
}

int qcd_hip_synchronize(void) {
  QH_TRY(ensure_init());
  QH_CHECK(hipStreamSynchronize(g_ctx.stream));
  return 0;
}

size_t qcd_hip_bytes_in_use(void) { return g_ctx.bytes_in_use.load(); }

void qcd_hip_release_cache(void) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  for (auto &b : g_pool) (void)hipFree(b.second.p);
  g_pool.clear();
  g_pool_bytes = 0;
}

// ------------------------------------------------------------ lattices
int qcd_hip_lattice_create(qcd_hip_lattice *out, const int L[4], const double theta[4]) {
  *out = nullptr;
  for (int d = 0; d < 4; d++)
    if (L[d] < 2) {
      set_error("lattice extent L[%d] = %d must be at least 2", d, L[d]);
      return 1;
    }
  long V = (long)L[0] * L[1] * L[2] * L[3];
  if (V >= (1L << 24)) {
    set_error("lattice of %ld sites exceeds the 2^24-site limit of 32-bit in-field offsets", V);
    return 1;
  }
  if (V % 2) {
    set_error("lattice volume must be even (lib/src/qcd_init.c:45-51)");
    return 1;
  }
  auto *lat = new qcd_hip_lattice_s();
  for (int d = 0; d < 4; d++) {
    lat->L[d] = L[d];
    lat->theta[d] = theta ? theta[d] : 0.0;
    double a = lat->theta[d] / (double)L[d];
    lat->twisted[d] = lat->theta[d] != 0.0;
    lat->ph_re[d] = lat->twisted[d] ? cos(a) : 1.0;
    lat->ph_im[d] = lat->twisted[d] ? sin(a) : 0.0;
  }
  lat->V = V;
  lat->stride = round_up(V, 64);
  *out = lat;
  return 0;
}

void qcd_hip_lattice_destroy(qcd_hip_lattice lat) { delete lat; }

long qcd_hip_lattice_volume(qcd_hip_lattice lat) { return lat ? lat->V : 0; }

// ------------------------------------------------------------ fields
static int field_alloc(qcd_hip_field *out, qcd_hip_lattice lat, int kind, int prec, int nrhs) {
  *out = nullptr;
  QH_TRY(ensure_init());
  if (!lat) {
    set_error("field create: null lattice");
    return 1;
  }
  if (prec < QCD_HIP_F64 || prec > QCD_HIP_F16) {
    set_error("field create: unknown precision %d", prec);
    return 1;
  }
  if (kind == QCD_HIP_SPINOR && nrhs < 1) {
    set_error("spinor create: nrhs must be >= 1 (got %d)", nrhs);
    return 1;
  }
  auto *f = new qcd_hip_field_s();
  f->kind = kind;
  f->prec = prec;
  f->nrhs = kind == QCD_HIP_SPINOR ? nrhs : 1;
  f->lat = lat;
  size_t ncomp = kind == QCD_HIP_SPINOR ? 12 * (size_t)nrhs : 36;
  size_t data_bytes = ncomp * (size_t)lat->stride * spinor_comp_bytes(prec);
  size_t norm_bytes = (kind == QCD_HIP_SPINOR && prec == QCD_HIP_F16) ? sizeof(float) * (size_t)nrhs * lat->stride : 0;
  int rc = dev_alloc(&f->data, data_bytes + norm_bytes);
  if (rc) {
    delete f;
    return rc;
  }
  f->norm = norm_bytes ? (float *)((char *)f->data + data_bytes) : nullptr;
  f->bytes = data_bytes + norm_bytes;
  // complete before returning: the field may next be used on any stream, and
  // the null stream does not order against the library's non-blocking streams
  hipError_t e = hipMemsetAsync(f->data, 0, f->bytes, g_ctx.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_ctx.stream);
  if (e != hipSuccess) {
    dev_free(f->data, f->bytes);
    delete f;
    return hip_fail(e, "hipMemsetAsync", __FILE__, __LINE__);
  }
  *out = f;
  return 0;
}

int qcd_hip_spinor_create(qcd_hip_field *f, qcd_hip_lattice lat, int prec, int nrhs) {
  return field_alloc(f, lat, QCD_HIP_SPINOR, prec, nrhs);
}

int qcd_hip_gauge_create(qcd_hip_field *f, qcd_hip_lattice lat, int prec) {
  return field_alloc(f, lat, QCD_HIP_GAUGE, prec, 1);
}

void qcd_hip_field_destroy(qcd_hip_field f) {
  if (!f) return;
  (void)hipStreamSynchronize(g_ctx.stream);
  dev_free(f->data, f->bytes);
  dev_free(f->hop_links, f->hop_bytes);
  delete f;
}

static void *rhs_ptr(qcd_hip_field f, int rhs) {
  return (char *)f->data + (size_t)rhs * 12 * f->lat->stride * spinor_comp_bytes(f->prec);
}

int qcd_hip_spinor_upload(qcd_hip_field f, int rhs, const double *aos) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  QH_TRY(ensure_init());
  if (!f || f->kind != QCD_HIP_SPINOR || rhs < 0 || rhs >= f->nrhs || !aos) {
    set_error("spinor upload: bad field, rhs or pointer");
    return 1;
  }
  const long V = f->lat->V, st = f->lat->stride;
  void *buf;
  QH_TRY(stage((size_t)V * 192, &buf));
  QH_CHECK(hipMemcpyAsync(buf, aos, (size_t)V * 192, hipMemcpyHostToDevice, g_ctx.stream));
  const double2 *src = (const double2 *)buf;
  if (f->prec == QCD_HIP_F64)
    k_aos_to_planes<double2, 12><<<grid_for(V), kBlock, 0, g_ctx.stream>>>(src, (double2 *)rhs_ptr(f, rhs), V, st);
  else if (f->prec == QCD_HIP_F32)
    k_aos_to_planes<float2, 12><<<grid_for(V), kBlock, 0, g_ctx.stream>>>(src, (float2 *)rhs_ptr(f, rhs), V, st);
  else
    k_aos_to_half_spinor<<<grid_for(V), kBlock, 0, g_ctx.stream>>>(src, (__half2 *)rhs_ptr(f, rhs),
                                                                   f->norm + (size_t)rhs * st, V, st);
  QH_CHECK(hipGetLastError());
  QH_CHECK(hipStreamSynchronize(g_ctx.stream));
  return 0;
}

int qcd_hip_spinor_download(qcd_hip_field f, int rhs, double *aos) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  QH_TRY(ensure_init());
  if (!f || f->kind != QCD_HIP_SPINOR || rhs < 0 || rhs >= f->nrhs || !aos) {
    set_error("spinor download: bad field, rhs or pointer");
    return 1;
  }
  const long V = f->lat->V, st = f->lat->stride;
  void *buf;
  QH_TRY(stage((size_t)V * 192, &buf));
  double2 *dst = (double2 *)buf;
  if (f->prec == QCD_HIP_F64)
    k_planes_to_aos<double2, 12><<<grid_for(V), kBlock, 0, g_ctx.stream>>>((const double2 *)rhs_ptr(f, rhs), dst, V, st);
  else if (f->prec == QCD_HIP_F32)
    k_planes_to_aos<float2, 12><<<grid_for(V), kBlock, 0, g_ctx.stream>>>((const float2 *)rhs_ptr(f, rhs), dst, V, st);
  else
    k_half_spinor_to_aos<<<grid_for(V), kBlock, 0, g_ctx.stream>>>((const __half2 *)rhs_ptr(f, rhs),
                                                                   f->norm + (size_t)rhs * st, dst, V, st);
  QH_CHECK(hipGetLastError());
  QH_CHECK(hipMemcpyAsync(aos, buf, (size_t)V * 192, hipMemcpyDeviceToHost, g_ctx.stream));
  QH_CHECK(hipStreamSynchronize(g_ctx.stream));
  return 0;
}

int qcd_hip_gauge_upload(qcd_hip_field f, const double *aos) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  QH_TRY(ensure_init());
  if (!f || f->kind != QCD_HIP_GAUGE || !aos) {
    set_error("gauge upload: bad field or pointer");
    return 1;
  }
  const long V = f->lat->V, st = f->lat->stride;
  void *buf;
  QH_TRY(stage((size_t)V * 576, &buf));
  QH_CHECK(hipMemcpyAsync(buf, aos, (size_t)V * 576, hipMemcpyHostToDevice, g_ctx.stream));
  const double2 *src = (const double2 *)buf;
  if (f->prec == QCD_HIP_F64)
    k_aos_to_planes<double2, 36><<<grid_for(V), kBlock, 0, g_ctx.stream>>>(src, (double2 *)f->data, V, st);
  else if (f->prec == QCD_HIP_F32)
    k_aos_to_planes<float2, 36><<<grid_for(V), kBlock, 0, g_ctx.stream>>>(src, (float2 *)f->data, V, st);
  else
    k_aos_to_planes<__half2, 36><<<grid_for(V), kBlock, 0, g_ctx.stream>>>(src, (__half2 *)f->data, V, st);
  gauge_written(f);
  QH_CHECK(hipGetLastError());
  QH_CHECK(hipStreamSynchronize(g_ctx.stream));
  return 0;
}

int qcd_hip_gauge_download(qcd_hip_field f, double *aos) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  QH_TRY(ensure_init());
  if (!f || f->kind != QCD_HIP_GAUGE || !aos) {
    set_error("gauge download: bad field or pointer");
    return 1;
  }
  const long V = f->lat->V, st = f->lat->stride;
  void *buf;
  QH_TRY(stage((size_t)V * 576, &buf));
  double2 *dst = (double2 *)buf;
  if (f->prec == QCD_HIP_F64)
    k_planes_to_aos<double2, 36><<<grid_for(V), kBlock, 0, g_ctx.stream>>>((const double2 *)f->data, dst, V, st);
  else if (f->prec == QCD_HIP_F32)
    k_planes_to_aos<float2, 36><<<grid_for(V), kBlock, 0, g_ctx.stream>>>((const float2 *)f->data, dst, V, st);
  else
    k_planes_to_aos<__half2, 36><<<grid_for(V), kBlock, 0, g_ctx.stream>>>((const __half2 *)f->data, dst, V, st);
  QH_CHECK(hipGetLastError());
  QH_CHECK(hipMemcpyAsync(aos, buf, (size_t)V * 576, hipMemcpyDeviceToHost, g_ctx.stream));
  QH_CHECK(hipStreamSynchronize(g_ctx.stream));
  return 0;
}

}  // extern "C"

template <typename Src, typename Dst, int NC>
static void launch_convert(const void *src, const float *sn, void *dst, float *dn, long V, long st) {
  k_convert_planes<Src, Dst, NC><<<grid_for(V), kBlock, 0, g_ctx.stream>>>((const Src *)src, sn, (Dst *)dst, dn, V, st);
}

template <int NC>
static int convert_planes(int sp, const void *src, const float *sn, int dp, void *dst, float *dn, long V, long st) {
  if (sp == QCD_HIP_F64 && dp == QCD_HIP_F64) launch_convert<double2, double2, NC>(src, sn, dst, dn, V, st);
  else if (sp == QCD_HIP_F64 && dp == QCD_HIP_F32) launch_convert<double2, float2, NC>(src, sn, dst, dn, V, st);
  else if (sp == QCD_HIP_F64 && dp == QCD_HIP_F16) launch_convert<double2, __half2, NC>(src, sn, dst, dn, V, st);
  else if (sp == QCD_HIP_F32 && dp == QCD_HIP_F64) launch_convert<float2, double2, NC>(src, sn, dst, dn, V, st);
  else if (sp == QCD_HIP_F32 && dp == QCD_HIP_F32) launch_convert<float2, float2, NC>(src, sn, dst, dn, V, st);
  else if (sp == QCD_HIP_F32 && dp == QCD_HIP_F16) launch_convert<float2, __half2, NC>(src, sn, dst, dn, V, st);
  else if (sp == QCD_HIP_F16 && dp == QCD_HIP_F64) launch_convert<__half2, double2, NC>(src, sn, dst, dn, V, st);
  else if (sp == QCD_HIP_F16 && dp == QCD_HIP_F32) launch_convert<__half2, float2, NC>(src, sn, dst, dn, V, st);
  else launch_convert<__half2, __half2, NC>(src, sn, dst, dn, V, st);
  QH_CHECK(hipGetLastError());
  return 0;
}

extern "C" {

int qcd_hip_field_convert(qcd_hip_field dst, qcd_hip_field src) {
  QH_TRY(ensure_init());
  if (!dst || !src || dst->kind != src->kind || dst->lat->V != src->lat->V || dst->nrhs != src->nrhs) {
    set_error("field convert: fields differ in kind, lattice or nrhs");
    return 1;
  }
  const long V = src->lat->V, st = src->lat->stride;
  if (src->kind == QCD_HIP_GAUGE) {
    gauge_written(dst);
    QH_TRY(convert_planes<36>(src->prec, src->data, nullptr, dst->prec, dst->data, nullptr, V, st));
  } else {
    for (int r = 0; r < src->nrhs; r++) {
      const float *sn = src->norm ? src->norm + (size_t)r * st : nullptr;
      float *dn = dst->norm ? dst->norm + (size_t)r * st : nullptr;
      QH_TRY(convert_planes<12>(src->prec, rhs_ptr(src, r), sn, dst->prec, rhs_ptr(dst, r), dn, V, st));
    }
  }
  QH_CHECK(hipStreamSynchronize(g_ctx.stream));   // the copy may be used on another stream next
  return 0;
}

}  // extern "C"

// ================================================================ generator
// The recipe lives in generator.h (shared with the device kernels below).
namespace {

using qcdgen::C;

template <typename F>
void parallel_for(long n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  long chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    long b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=]() {
      for (long i = b; i < e; i++) f(i);
    });
  }
  for (auto &x : th) x.join();
}

// one lane per link (site, mu), written straight into the tiled F64 layout
__global__ void __launch_bounds__(kBlock) k_gauge_generate(double2 *data, long V, long stride, double eps,
                                                           uint64_t seed) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * V) return;
  double g[18];
  qcdgen::gaussians(g, 18, seed, (uint64_t)i);
  C U[3][3];
  qcdgen::su3_link(U, g, eps);
  const long site = i / 4;
  const int mu = (int)(i % 4);
  for (int k = 0; k < 9; k++)
    data[gauge_index(9 * mu + k, (size_t)site, (size_t)stride)] = make_double2(U[k / 3][k % 3].re, U[k / 3][k % 3].im);
}

// field rhs = delta(site) on component (spin, col), zero elsewhere
__global__ void __launch_bounds__(kBlock) k_point_source(double2 *rhs0, long V, long site, int comp) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  for (int k = 0; k < 12; k++)
    rhs0[tile_index(12, k, (size_t)s)] = make_double2(s == site && k == comp ? 1.0 : 0.0, 0.0);
}

}  // namespace

namespace qcdhip {

int gauge_generate_device(qcd_hip_field g, double eps, uint64_t seed, hipStream_t st) {
  if (!g || g->kind != QCD_HIP_GAUGE || g->prec != QCD_HIP_F64) {
    set_error("gauge_generate: need an FP64 gauge field");
    return 1;
  }
  const long V = g->lat->V;
  gauge_written(g);
  k_gauge_generate<<<grid_for(4 * V), kBlock, 0, st>>>((double2 *)g->data, V, g->lat->stride, eps, seed);
  QH_CHECK(hipGetLastError());
  return 0;
}

int point_source_device(qcd_hip_field f, int rhs, long site, int spin, int col, hipStream_t st) {
  if (!f || f->kind != QCD_HIP_SPINOR || f->prec != QCD_HIP_F64 || rhs < 0 || rhs >= f->nrhs || site < 0 ||
      site >= f->lat->V || spin < 0 || spin > 3 || col < 0 || col > 2) {
    set_error("point_source: bad field, rhs, site or component");
    return 1;
  }
  const long V = f->lat->V;
  k_point_source<<<grid_for(V), kBlock, 0, st>>>((double2 *)f->data + (size_t)rhs * 12 * f->lat->stride, V, site,
                                                 3 * spin + col);
  QH_CHECK(hipGetLastError());
  return 0;
}

}  // namespace qcdhip

extern "C" {

int qcd_hip_gauge_generate_host(double *aos, long V, double eps, uint64_t seed) {
  if (!aos || V <= 0) {
    qcdhip::set_error("gauge generate: bad pointer or volume");
    return 1;
  }
  parallel_for(4 * V, [=](long i) {
    double g[18];
    qcdgen::gaussians(g, 18, seed, (uint64_t)i);
    qcdgen::su3_link(reinterpret_cast<C(*)[3]>(aos + 18 * i), g, eps);
  });
  return 0;
}

int qcd_hip_spinor_gaussian_host(double *aos, long V, uint64_t seed) {
  if (!aos || V <= 0) {
    qcdhip::set_error("spinor generate: bad pointer or volume");
    return 1;
  }
  parallel_for(V, [=](long i) { qcdgen::gaussians(aos + 24 * i, 24, seed ^ qcdgen::kSpinorSeedSalt, (uint64_t)i); });
  return 0;
}

int qcd_hip_gauge_generate(qcd_hip_field g, double eps, uint64_t seed) {
  QH_TRY(qcdhip::ensure_init());
  QH_TRY(qcdhip::gauge_generate_device(g, eps, seed, qcdhip::ctx()->stream));
  QH_CHECK(hipStreamSynchronize(qcdhip::ctx()->stream));
  return 0;
}

int qcd_hip_point_source(qcd_hip_field f, int rhs, const int site[4], int spin, int col) {
  QH_TRY(qcdhip::ensure_init());
  if (!f || !site) {
    qcdhip::set_error("point_source: null argument");
    return 1;
  }
  const int *L = f->lat->L;
  for (int d = 0; d < 4; d++)
    if (site[d] < 0 || site[d] >= L[d]) {
      qcdhip::set_error("point_source: coordinate %d out of range", d);
      return 1;
    }
  const long l = site[0] + (long)L[0] * (site[1] + (long)L[1] * (site[2] + (long)L[2] * site[3]));
  QH_TRY(qcdhip::point_source_device(f, rhs, l, spin, col, qcdhip::ctx()->stream));
  QH_CHECK(hipStreamSynchronize(qcdhip::ctx()->stream));
  return 0;
}

}  // extern "C"
