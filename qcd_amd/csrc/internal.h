// internal.h -- shared definitions of the MI355X (gfx950) lattice-QCD library.
//
// Device data layout: tiles of 64 sites (one wavefront), complex components
// of a tile stored one after another (array of structures of arrays).  A
// wave reading component k of its 64 sites reads one contiguous segment, and
// all components of a tile sit in one small region (DRAM page locality).
//   spinor   comp k = 3*spin + colour; per right-hand side
//            [stride/64][12][64] complex; F64 double2, F32 float2, F16 __half2
//            + float norm [nrhs][stride] (value = half * norm, norm = max |re,im|)
//   gauge    one array per direction mu: [4][stride/64][9][64] complex with
//            k = 3*row + col (F16 needs no norm: SU(3) entries are bounded by 1)
// Sites are numbered as on the host: l = t + T*(x + X*(y + Y*z)).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <cfloat>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "qcd_hip.h"

namespace qcdhip {

constexpr int kBlock = 256;          // threads per workgroup for site kernels
constexpr int kNumXcd = 8;           // MI355X: 8 accelerator complex dies
constexpr int kMaxReduceBlocks = 1024;
constexpr int kMaxKrylov = 64;       // largest GCR restart length

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define QH_CHECK(call)                                                          \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) return ::qcdhip::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

#define QH_TRY(call)                 \
  do {                               \
    int rc_ = (call);                \
    if (rc_ != 0) return rc_;        \
  } while (0)

// ---------------------------------------------------------------- context
struct Context {
  int device = -1;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  std::atomic<size_t> bytes_in_use{0}; // device bytes held by library objects
  // reduction scratch shared by the BLAS of this context
  double *partials = nullptr;         // [kMaxReduceBlocks * 2 * (kMaxKrylov + 4)]
  double *scalars = nullptr;          // small device scalar area
  double *host_scalars = nullptr;     // pinned mirror for reads
};

Context *ctx();                       // current context (initialised on first use)
// Serialises the library's shared host-side state (staging buffer, scratch
// fields, tuning table) when several host threads drive separate streams.
std::recursive_mutex &state_mutex();
int ensure_init();
int dev_alloc(void **p, size_t bytes);
void dev_free(void *p, size_t bytes);
// a library call wrote new links into gauge field f
inline void gauge_written(qcd_hip_field f);

}  // namespace qcdhip

// ---------------------------------------------------------------- objects
struct qcd_hip_lattice_s {
  int L[4];
  long V;            // sites
  long stride;       // plane stride in sites (V rounded up to 64)
  double theta[4];
  double ph_re[4];   // e^{+i theta_mu / L_mu}
  double ph_im[4];
  int twisted[4];    // theta_mu != 0
};

struct qcd_hip_field_s {
  int kind;          // QCD_HIP_SPINOR or QCD_HIP_GAUGE
  int prec;          // QCD_HIP_F64 / F32 / F16
  int nrhs;          // spinor: number of right-hand sides
  qcd_hip_lattice lat;
  void *data;
  float *norm;       // F16 spinors only
  size_t bytes;      // total device bytes (data + norm)
  // Gauge fields: every library write bumps version.  The matrix-core hops
  // keep the phased links in operand order (mma.hip) built from version
  // hop_version with the phases hop_ph; rebuilt when either changed.
  unsigned version = 0;
  void *hop_links = nullptr;
  size_t hop_bytes = 0;
  unsigned hop_version = 0;
  double hop_ph[4][2] = {};
};

namespace qcdhip {

inline void gauge_written(qcd_hip_field f) {
  if (f) f->version++;
}

// --------------------------------------------------------- small helpers
inline long round_up(long x, long m) { return (x + m - 1) / m * m; }
inline int grid_for(long n, int block = kBlock) { return (int)((n + block - 1) / block); }
inline size_t spinor_comp_bytes(int prec) { return prec == QCD_HIP_F64 ? 16 : (prec == QCD_HIP_F32 ? 8 : 4); }

constexpr int kTile = 64;   // sites per layout tile
// element index of component k (of nc) of site s inside one tiled array
__host__ __device__ __forceinline__ size_t tile_index(int nc, int k, size_t s) {
  return ((s / kTile) * (size_t)nc + (size_t)k) * kTile + (s % kTile);
}
// element index of gauge component k = 9*mu + 3*row + col of site s
__host__ __device__ __forceinline__ size_t gauge_index(int k, size_t s, size_t stride) {
  return (size_t)(k / 9) * 9 * stride + tile_index(9, k % 9, s);
}
// Reduced-precision fields group their components inside a tile: G
// consecutive components of a site are adjacent, so a lane reads a group
// with one 16-byte (or, for FP16 links, 12-byte) load: tile
// [nc/G groups][64 sites][G].  G = 4 for FP16 spinors (nc = 12), 3 for FP16
// links (nc = 9), 2 for FP32 spinors; FP32 links and FP64 fields keep one
// plane per component (G = 1).  ESZ is the element size in bytes.
__host__ __device__ constexpr int tile_group(int esz, int nc) {
  return esz == 4 ? (nc == 12 ? 4 : 3) : (esz == 8 && nc == 12) ? 2 : 1;
}
template <int ESZ>
__host__ __device__ __forceinline__ size_t tile_index_e(int nc, int k, size_t s) {
  const int G = tile_group(ESZ, nc);
  if (G == 1) return tile_index(nc, k, s);
  return ((s / kTile) * (size_t)nc + (size_t)(k / G) * G) * kTile + (s % kTile) * G + (size_t)(k % G);
}
template <int ESZ>
__host__ __device__ __forceinline__ size_t gauge_index_e(int k, size_t s, size_t stride) {
  return (size_t)(k / 9) * 9 * stride + tile_index_e<ESZ>(9, k % 9, s);
}

// Hardware dispatches workgroup i to XCD i % 8.  Remap so each XCD owns one
// contiguous range of logical blocks: neighbouring sites then share an L2.
__device__ __forceinline__ unsigned xcd_block(unsigned bid, unsigned nblk) {
  unsigned q = nblk / kNumXcd, r = nblk % kNumXcd;
  unsigned xcd = bid % kNumXcd, local = bid / kNumXcd;
  return xcd * q + (xcd < r ? xcd : r) + local;
}

}  // namespace qcdhip
