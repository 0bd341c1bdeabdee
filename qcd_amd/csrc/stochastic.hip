// stochastic.hip -- Z4 noise sources (lib/src/qcd_stochastic.c:8-58) and the
// random-number interface of the reference (lib/src/qcd_rng.c:1-33).
//
// The reference draws from GSL's ranlux389 (gsl_rng_uniform_int(r, 4)); GSL
// is not part of this library, so the stream here is the splitmix64
// finaliser of a 64-bit counter keyed by the process seed: reproducible for
// a seed, statistically the same noise, different numbers.  The host drop-in
// qcd_z2Vector consumes the stream in the reference's loop order; the
// device entry point qcd_hip_z4_source keys each (site, component) draw by
// its index instead, so its vectors do not depend on launch order.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "blas_common.h"
#include "generator.h"
#include "qcd.h"

namespace {

struct HostStream {
  uint64_t key = 0, n = 0;
  bool init = false;
};
HostStream g_rng;

// the four noise values, indexed as in qcd_stochastic.c:10-14
const double kZ4[4][2] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}, {0.0, -1.0}};

}  // namespace

namespace qcdhip {

__global__ void __launch_bounds__(kBlock) k_z4_source(double2 *__restrict__ v, long V, int T, int t_slice,
                                                      uint64_t key) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  const bool on = t_slice < 0 || (int)(s % T) == t_slice;
#pragma unroll
  for (int k = 0; k < 12; k++) {
    double2 z = make_double2(0.0, 0.0);
    if (on) {
      const unsigned r = (unsigned)(qcdgen::mix64(key + (uint64_t)(12 * s + k) * 0x632BE59BD9B4E019ULL) >> 62);
      z = make_double2(r == 0 ? 1.0 : r == 2 ? -1.0 : 0.0, r == 1 ? 1.0 : r == 3 ? -1.0 : 0.0);
    }
    v[tile_index(12, k, (size_t)s)] = z;
  }
}

}  // namespace qcdhip

using namespace qcdhip;

extern "C" {

// qcd_rng.c:5-17: the process seed is seed for rank 0 and the myid-th value
// of the C library's rand() after srand(seed) otherwise
void qcd_rng_init(unsigned int seed, int myid) {
  unsigned int process_seed = seed;
  srand(seed);
  for (int i = 0; i < myid; i++) process_seed = (unsigned int)rand();
  g_rng.key = qcdgen::mix64((uint64_t)process_seed) ^ 0x5A0F1E2D3C4B6978ULL;
  g_rng.n = 0;
  g_rng.init = true;
}

void qcd_rng_finalize(void) { g_rng = HostStream(); }

unsigned long int qcd_get_rand_4(void) {
  if (!g_rng.init) qcd_rng_init(0u, 0);
  const uint64_t r = qcdgen::mix64(g_rng.key + (g_rng.n++) * 0x632BE59BD9B4E019ULL);
  return (unsigned long int)(r >> 62);
}

// qcd_stochastic.c:8-58 (one rank: every slice is local)
void qcd_z2Vector(qcd_vector *vec, int t_slice) {
  if (!vec || !vec->initialized) {
    fprintf(stderr, "Error in qcd_z2Vector! Vector must be properly initialized\n");
    return;
  }
  const qcd_uint_2 *L = vec->geo->lL;
  std::memset(&vec->D[0][0][0].re, 0, 12 * vec->geo->lV * sizeof(qcd_complex_16));
  auto fill = [&](int t) {
    for (int z = 0; z < L[3]; z++)
      for (int y = 0; y < L[2]; y++)
        for (int x = 0; x < L[1]; x++) {
          const qcd_uint_8 iv = qcd_LEXIC(t, x, y, z, L);
          for (int sp = 0; sp < 4; sp++)
            for (int col = 0; col < 3; col++) {
              const unsigned long r = qcd_get_rand_4();
              vec->D[iv][sp][col].re = kZ4[r][0];
              vec->D[iv][sp][col].im = kZ4[r][1];
            }
        }
  };
  if (t_slice >= 0) {
    printf("time dilution \n");
    if (t_slice < L[0]) fill(t_slice);
  } else {
    printf("full in time \n");
    for (int t = 0; t < L[0]; t++) fill(t);
  }
}

int qcd_hip_z4_source(qcd_hip_field f, int rhs, int t_slice, uint64_t seed) {
  QH_TRY(ensure_init());
  if (!f || f->kind != QCD_HIP_SPINOR || f->prec != QCD_HIP_F64 || rhs < 0 || rhs >= f->nrhs ||
      t_slice >= f->lat->L[0]) {
    set_error("z4_source: need an FP64 spinor, rhs in range and t_slice < T");
    return 1;
  }
  const long V = f->lat->V;
  const uint64_t key = qcdgen::mix64(seed) ^ 0x3C6EF372FE94F82BULL;
  hipStream_t st = ctx()->stream;
  k_z4_source<<<grid_for(V), kBlock, 0, st>>>((double2 *)f->data + (size_t)rhs * 12 * f->lat->stride, V, f->lat->L[0],
                                               t_slice, key);
  QH_CHECK(hipGetLastError());
  QH_CHECK(hipStreamSynchronize(st));
  return 0;
}

}  // extern "C"
