// dslash.hip -- twisted-mass Wilson Dirac operator on gfx950.
//
// One lane per site, 256-lane workgroups remapped so that each XCD sweeps one
// contiguous range of the lattice (neighbour reuse stays in that XCD's L2).
// Neighbours are computed arithmetically (no index tables: the reference's
// plus/minus tables, lib/src/qcd_init.c:110-124, cost 64 B/site of traffic).
// Each hop is spin-projected to a half spinor before the SU(3) multiply
// (lib/src/qcd_wilson.c:190-260): 2 colour vectors per link instead of 4.
//
//  D psi(x) = (m+4 + i tm g5) psi(x)
//    - 1/2 sum_mu [ e^{+i th_mu/L_mu} U_mu(x) (1-g_mu) psi(x+mu)
//                 + e^{-i th_mu/L_mu} U_mu(x-mu)^+ (1+g_mu) psi(x-mu) ]
//
// Epilogues: D, Q = i g5 D (lib/src/qcd_wilson.c:805-812), and the residuals
// b - D x, b - Q x used by the solver, optionally with |out|^2 block partials.
#include "dslash_common.h"

namespace qcdhip {

// OCC: 0 = compiler's register allocation, W > 0 = at least W waves per SIMD.
template <int PS, int PG, int EPI, bool RED, int OCC>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1, OCC > 0 ? OCC : 8)))
k_dslash(DslashArgs a, SpinorIn<PS> in, GaugeIn<PG> U, SpinorOut<PS> out, SpinorIn<PS> b,
         double *__restrict__ partials) {
  using R = typename Prec<PS>::R;
  const Site st = site_of_lane(a);
  const unsigned s = st.s;
  double red = 0.0;
  if (st.valid) {
    Cx<R> acc[4][3];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int c = 0; c < 3; c++) acc[i][c] = Cx<R>{R(0), R(0)};

    hop_dir<0>(acc, a, st, in, U);
    hop_dir<1>(acc, a, st, in, U);
    hop_dir<2>(acc, a, st, in, U);
    hop_dir<3>(acc, a, st, in, U);

    Cx<R> psi[12];
    in.load(s, psi);
    Cx<R> o[12];
    const R mp4 = R(a.mp4), tm = R(a.tm);
#pragma unroll
    for (int sp = 0; sp < 4; sp++) {
      const R g = sp < 2 ? tm : -tm;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        Cx<R> v = psi[3 * sp + c];
        Cx<R> d{mp4 * v.re - g * v.im, mp4 * v.im + g * v.re};
        o[3 * sp + c] = Cx<R>{d.re - R(0.5) * acc[sp][c].re, d.im - R(0.5) * acc[sp][c].im};
      }
    }
    apply_epilogue<EPI>(o, b, s);
    if (RED) {
#pragma unroll
      for (int k = 0; k < 12; k++) red += (double)o[k].re * (double)o[k].re + (double)o[k].im * (double)o[k].im;
    }
    out.store(s, o);
  }
  if (RED) block_reduce_store(red, partials, blockIdx.x);
}

template <int PS, int PG, int EPI, bool RED, int OCC>
static int launch_occ(const DslashArgs &a, int grid, qcd_hip_field out, int ro, qcd_hip_field in, int ri,
                      qcd_hip_field g, qcd_hip_field b, int rb, double *partials, hipStream_t stream) {
  SpinorIn<PS> bi = b ? SpinorIn<PS>::of(b, rb) : SpinorIn<PS>::of(in, ri);
  k_dslash<PS, PG, EPI, RED, OCC><<<grid, kBlock, 0, stream>>>(a, SpinorIn<PS>::of(in, ri), GaugeIn<PG>::of(g),
                                                               SpinorOut<PS>::of(out, ro), bi, partials);
  QH_CHECK(hipGetLastError());
  return 0;
}

template <int PS, int PG, int EPI, bool RED>
static int launch_one(const DslashArgs &a, int occ, int grid, qcd_hip_field out, int ro, qcd_hip_field in, int ri,
                      qcd_hip_field g, qcd_hip_field b, int rb, double *partials, hipStream_t stream) {
  switch (occ) {
    case 3: return launch_occ<PS, PG, EPI, RED, 3>(a, grid, out, ro, in, ri, g, b, rb, partials, stream);
    case 4: return launch_occ<PS, PG, EPI, RED, 4>(a, grid, out, ro, in, ri, g, b, rb, partials, stream);
    default: return launch_occ<PS, PG, EPI, RED, 0>(a, grid, out, ro, in, ri, g, b, rb, partials, stream);
  }
}

template <int PS, int PG>
static int launch_epi(int epi, bool red, const DslashArgs &a, int occ, int grid, qcd_hip_field out, int ro,
                      qcd_hip_field in, int ri, qcd_hip_field g, qcd_hip_field b, int rb, double *partials,
                      hipStream_t stream) {
#define QH_L(E, RD) launch_one<PS, PG, E, RD>(a, occ, grid, out, ro, in, ri, g, b, rb, partials, stream)
  switch (epi * 2 + (red ? 1 : 0)) {
    case 0: return QH_L(EPI_D, false);
    case 1: return QH_L(EPI_D, true);
    case 2: return QH_L(EPI_Q, false);
    case 3: return QH_L(EPI_Q, true);
    case 4: return QH_L(EPI_RES_D, false);
    case 5: return QH_L(EPI_RES_D, true);
    case 6: return QH_L(EPI_RES_Q, false);
    default: return QH_L(EPI_RES_Q, true);
  }
#undef QH_L
}

DslashArgs make_args(qcd_hip_lattice lat, double m, double tm) {
  DslashArgs a;
  std::memset(&a, 0, sizeof(a));
  a.V = (unsigned)lat->V;
  a.stride = lat->stride;
  for (int d = 0; d < 4; d++) {
    a.L[d] = (unsigned)lat->L[d];
    a.ph_re[d] = lat->ph_re[d];
    a.ph_im[d] = lat->ph_im[d];
    a.twisted[d] = lat->twisted[d];
  }
  a.mp4 = m + 4.0;
  a.tm = tm;
  a.map = MAP_XCD;
  return a;
}

// Fill the traversal fields of a; returns the grid size.
int set_traversal(DslashArgs &a, int map) {
  if (map == MAP_SLAB && a.L[2] >= (unsigned)kNumXcd) {
    a.map = MAP_SLAB;
    const unsigned P = a.L[0] * a.L[1];
    unsigned most = 0;
    for (int k = 0; k <= kNumXcd; k++) a.slab_y0[k] = (unsigned)((unsigned long)k * a.L[2] / kNumXcd);
    for (int k = 0; k < kNumXcd; k++) {
      const unsigned long n = (unsigned long)P * (a.slab_y0[k + 1] - a.slab_y0[k]) * a.L[3];
      most = std::max(most, (unsigned)((n + kBlock - 1) / kBlock));
    }
    a.slab_blocks = most;
    return (int)(most * kNumXcd);
  }
  a.map = MAP_XCD;
  return grid_for(a.V);
}

// ------------------------------------------------------------ tuning
// Best (traversal, occupancy) per (lattice extents, precisions), chosen by
// timing every candidate on first use.  All candidates compute bit-identical
// results (same per-site arithmetic); only speed differs.
struct TuneKey {
  int L[4], ps, pg;
  bool operator==(const TuneKey &o) const { return !std::memcmp(this, &o, sizeof(*this)); }
};
struct TuneVal {
  int map, occ;
  float ms;
};
static std::vector<std::pair<TuneKey, TuneVal>> g_tune;
static int g_autotune = -1;   // -1: read QCD_HIP_AUTOTUNE once

static const TuneVal *tune_lookup(const TuneKey &k) {
  for (auto &e : g_tune)
    if (e.first == k) return &e.second;
  return nullptr;
}

static TuneKey key_of(qcd_hip_field in, qcd_hip_field g) {
  TuneKey k;
  std::memset(&k, 0, sizeof(k));
  for (int d = 0; d < 4; d++) k.L[d] = in->lat->L[d];
  k.ps = in->prec;
  k.pg = g->prec;
  return k;
}

static int launch_cfg(int epi, bool red, int map, int occ, qcd_hip_field out, int ro, qcd_hip_field in, int ri,
                      qcd_hip_field g, qcd_hip_field b, int rb, double m, double tm, double *partials,
                      hipStream_t stream, int *grid_out) {
  DslashArgs a = make_args(in->lat, m, tm);
  const int grid = set_traversal(a, map);
  if (grid_out) *grid_out = grid;
  const int ps = in->prec, pg = g->prec;
  if (ps == QCD_HIP_F64 && pg == QCD_HIP_F64)
    return launch_epi<QCD_HIP_F64, QCD_HIP_F64>(epi, red, a, occ, grid, out, ro, in, ri, g, b, rb, partials, stream);
  if (ps == QCD_HIP_F32 && pg == QCD_HIP_F32)
    return launch_epi<QCD_HIP_F32, QCD_HIP_F32>(epi, red, a, occ, grid, out, ro, in, ri, g, b, rb, partials, stream);
  if (ps == QCD_HIP_F32 && pg == QCD_HIP_F16)
    return launch_epi<QCD_HIP_F32, QCD_HIP_F16>(epi, red, a, occ, grid, out, ro, in, ri, g, b, rb, partials, stream);
  if (ps == QCD_HIP_F16 && pg == QCD_HIP_F16)
    return launch_epi<QCD_HIP_F16, QCD_HIP_F16>(epi, red, a, occ, grid, out, ro, in, ri, g, b, rb, partials, stream);
  set_error("dslash: unsupported precision pair (spinor %d, gauge %d)", ps, pg);
  return 1;
}

static const int kMaps[] = {MAP_XCD, MAP_SLAB};
static const int kOccs[] = {0, 3, 4};

int dslash_tune(qcd_hip_field out, qcd_hip_field in, qcd_hip_field g, int force) {
  const TuneKey key = key_of(in, g);
  if (!force && tune_lookup(key)) return 0;
  hipStream_t st = ctx()->stream;
  hipEvent_t e0, e1;
  QH_CHECK(hipEventCreate(&e0));
  QH_CHECK(hipEventCreate(&e1));
  TuneVal best{MAP_XCD, 0, 1e30f};
  const int reps = in->lat->V > (1L << 22) ? 3 : 10;
  for (int map : kMaps) {
    if (map == MAP_SLAB && in->lat->L[2] < kNumXcd) continue;
    for (int occ : kOccs) {
      for (int w = 0; w < 2; w++)
        QH_TRY(launch_cfg(EPI_D, false, map, occ, out, 0, in, 0, g, nullptr, 0, 0.0, 0.0, nullptr, st, nullptr));
      QH_CHECK(hipEventRecord(e0, st));
      for (int r = 0; r < reps; r++)
        QH_TRY(launch_cfg(EPI_D, false, map, occ, out, 0, in, 0, g, nullptr, 0, 0.0, 0.0, nullptr, st, nullptr));
      QH_CHECK(hipEventRecord(e1, st));
      QH_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      QH_CHECK(hipEventElapsedTime(&ms, e0, e1));
      ms /= reps;
      if (ms < best.ms) best = TuneVal{map, occ, ms};
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (auto &e : g_tune)
    if (e.first == key) {
      e.second = best;
      return 0;
    }
  g_tune.push_back({key, best});
  return 0;
}

static bool autotune_enabled() {
  if (g_autotune < 0) {
    const char *e = getenv("QCD_HIP_AUTOTUNE");
    g_autotune = (e && e[0] == '0') ? 0 : 1;
  }
  return g_autotune == 1;
}

int dslash_check(qcd_hip_field out, qcd_hip_field in, qcd_hip_field g) {
  if (!out || !in || !g || out->kind != QCD_HIP_SPINOR || in->kind != QCD_HIP_SPINOR || g->kind != QCD_HIP_GAUGE) {
    set_error("dslash: expected (spinor out, spinor in, gauge)");
    return 1;
  }
  if (out->lat->V != in->lat->V || g->lat->V != in->lat->V || out->lat->stride != in->lat->stride) {
    set_error("dslash: fields live on different lattices");
    return 1;
  }
  for (int d = 0; d < 4; d++)
    if (out->lat->L[d] != in->lat->L[d] || g->lat->L[d] != in->lat->L[d]) {
      set_error("dslash: lattice extents differ");
      return 1;
    }
  if (out->prec != in->prec) {
    set_error("dslash: in and out precisions differ");
    return 1;
  }
  if (out == in) {
    set_error("dslash: out must not alias in");
    return 1;
  }
  if (in->lat->V >= (1L << 31)) {
    set_error("dslash: lattice too large for 32-bit site indices");
    return 1;
  }
  return 0;
}

// Single-rhs launch used by the solver and the public entry points.  Uses
// the tuned configuration of this (lattice, precision) if there is one.
int dslash_launch(int epi, bool red, qcd_hip_field out, int ro, qcd_hip_field in, int ri, qcd_hip_field g,
                  qcd_hip_field b, int rb, double m, double tm, double *partials, hipStream_t stream, int *grid_out) {
  const TuneVal *t = tune_lookup(key_of(in, g));
  const int map = t ? t->map : MAP_SLAB, occ = t ? t->occ : 0;
  return launch_cfg(epi, red, map, occ, out, ro, in, ri, g, b, rb, m, tm, partials, stream, grid_out);
}

// Tune (once per key) before the first use, outside any graph capture.
int dslash_prepare(qcd_hip_field out, qcd_hip_field in, qcd_hip_field g) {
  if (!autotune_enabled()) return 0;
  return dslash_tune(out, in, g, 0);
}

int dslash_grid(qcd_hip_field in, qcd_hip_field g) {
  const TuneVal *t = tune_lookup(key_of(in, g));
  DslashArgs a = make_args(in->lat, 0.0, 0.0);
  return set_traversal(a, t ? t->map : MAP_SLAB);
}

}  // namespace qcdhip

using namespace qcdhip;

extern "C" {

int qcd_hip_dslash(qcd_hip_field out, qcd_hip_field in, qcd_hip_field gauge, double m, double tm, int op) {
  QH_TRY(ensure_init());
  QH_TRY(dslash_check(out, in, gauge));
  if (out->nrhs != in->nrhs) {
    set_error("dslash: in has %d rhs, out has %d", in->nrhs, out->nrhs);
    return 1;
  }
  const int epi = op == QCD_HIP_OP_Q ? EPI_Q : EPI_D;
  QH_TRY(dslash_prepare(out, in, gauge));
  for (int r = 0; r < in->nrhs; r++)
    QH_TRY(dslash_launch(epi, false, out, r, in, r, gauge, nullptr, 0, m, tm, nullptr, ctx()->stream, nullptr));
  QH_CHECK(hipStreamSynchronize(ctx()->stream));
  return 0;
}

int qcd_hip_dslash_time(qcd_hip_field out, qcd_hip_field in, qcd_hip_field gauge, double m, double tm, int op,
                        int warmup, int iters, double *avg_ms) {
  QH_TRY(ensure_init());
  QH_TRY(dslash_check(out, in, gauge));
  if (iters < 1 || warmup < 0 || !avg_ms) {
    set_error("dslash_time: iters must be >= 1");
    return 1;
  }
  hipStream_t st = ctx()->stream;
  const int epi = op == QCD_HIP_OP_Q ? EPI_Q : EPI_D;
  QH_TRY(dslash_prepare(out, in, gauge));
  for (int i = 0; i < warmup; i++)
    for (int r = 0; r < in->nrhs; r++) QH_TRY(dslash_launch(epi, false, out, r, in, r, gauge, nullptr, 0, m, tm, nullptr, st, nullptr));
  hipEvent_t e0, e1;
  QH_CHECK(hipEventCreate(&e0));
  QH_CHECK(hipEventCreate(&e1));
  QH_CHECK(hipEventRecord(e0, st));
  for (int i = 0; i < iters; i++)
    for (int r = 0; r < in->nrhs; r++) QH_TRY(dslash_launch(epi, false, out, r, in, r, gauge, nullptr, 0, m, tm, nullptr, st, nullptr));
  QH_CHECK(hipEventRecord(e1, st));
  QH_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  QH_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / iters;
  return 0;
}

// Time every traversal / occupancy candidate on these fields and keep the
// fastest for their (lattice, precision).  Fills tuned[0..2] = map, occ, ms.
int qcd_hip_dslash_tune(qcd_hip_field out, qcd_hip_field in, qcd_hip_field gauge, int tuned[2], double *ms) {
  QH_TRY(ensure_init());
  QH_TRY(dslash_check(out, in, gauge));
  QH_TRY(dslash_tune(out, in, gauge, 1));
  const TuneVal *t = tune_lookup(key_of(in, gauge));
  if (tuned) {
    tuned[0] = t->map;
    tuned[1] = t->occ;
  }
  if (ms) *ms = t->ms;
  return 0;
}

// Force a configuration (map = 0 contiguous-per-XCD, 1 = y-slab per XCD;
// occ = 0 default, 3 or 4 minimum waves per SIMD) for these fields' key.
int qcd_hip_dslash_set_config(qcd_hip_field in, qcd_hip_field gauge, int map, int occ) {
  QH_TRY(ensure_init());
  if (!in || !gauge || (map != 0 && map != 1) || (occ != 0 && occ != 3 && occ != 4)) {
    set_error("dslash_set_config: bad arguments");
    return 1;
  }
  const TuneKey k = key_of(in, gauge);
  for (auto &e : g_tune)
    if (e.first == k) {
      e.second = TuneVal{map, occ, 0.f};
      return 0;
    }
  g_tune.push_back({k, TuneVal{map, occ, 0.f}});
  return 0;
}

}  // extern "C"
