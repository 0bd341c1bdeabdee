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

template <int PS, int PG, int EPI, bool RED>
__global__ void __launch_bounds__(kBlock) k_dslash(DslashArgs a, SpinorIn<PS> in, GaugeIn<PG> U, SpinorOut<PS> out,
                                                   SpinorIn<PS> b, double *__restrict__ partials) {
  using R = typename Prec<PS>::R;
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const unsigned s = lb * kBlock + threadIdx.x;
  double red = 0.0;
  if (s < a.V) {
    Cx<R> acc[4][3];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int c = 0; c < 3; c++) acc[i][c] = Cx<R>{R(0), R(0)};

    Site st = site_coords(s, a);
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
  if (RED) block_reduce_store(red, partials, lb);
}

template <int PS, int PG, int EPI, bool RED>
static int launch_one(const DslashArgs &a, qcd_hip_field out, int ro, qcd_hip_field in, int ri, qcd_hip_field g,
                      qcd_hip_field b, int rb, double *partials, hipStream_t stream) {
  const int grid = grid_for(a.V);
  SpinorIn<PS> bi = b ? SpinorIn<PS>::of(b, rb) : SpinorIn<PS>::of(in, ri);
  k_dslash<PS, PG, EPI, RED><<<grid, kBlock, 0, stream>>>(a, SpinorIn<PS>::of(in, ri), GaugeIn<PG>::of(g),
                                                          SpinorOut<PS>::of(out, ro), bi, partials);
  QH_CHECK(hipGetLastError());
  return 0;
}

template <int PS, int PG>
static int launch_epi(int epi, bool red, const DslashArgs &a, qcd_hip_field out, int ro, qcd_hip_field in, int ri,
                      qcd_hip_field g, qcd_hip_field b, int rb, double *partials, hipStream_t stream) {
  switch (epi * 2 + (red ? 1 : 0)) {
    case 0: return launch_one<PS, PG, EPI_D, false>(a, out, ro, in, ri, g, b, rb, partials, stream);
    case 1: return launch_one<PS, PG, EPI_D, true>(a, out, ro, in, ri, g, b, rb, partials, stream);
    case 2: return launch_one<PS, PG, EPI_Q, false>(a, out, ro, in, ri, g, b, rb, partials, stream);
    case 3: return launch_one<PS, PG, EPI_Q, true>(a, out, ro, in, ri, g, b, rb, partials, stream);
    case 4: return launch_one<PS, PG, EPI_RES_D, false>(a, out, ro, in, ri, g, b, rb, partials, stream);
    case 5: return launch_one<PS, PG, EPI_RES_D, true>(a, out, ro, in, ri, g, b, rb, partials, stream);
    case 6: return launch_one<PS, PG, EPI_RES_Q, false>(a, out, ro, in, ri, g, b, rb, partials, stream);
    default: return launch_one<PS, PG, EPI_RES_Q, true>(a, out, ro, in, ri, g, b, rb, partials, stream);
  }
}

DslashArgs make_args(qcd_hip_lattice lat, double m, double tm) {
  DslashArgs a;
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
  return a;
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

// Single-rhs launch used by the solver and the public entry points.
int dslash_launch(int epi, bool red, qcd_hip_field out, int ro, qcd_hip_field in, int ri, qcd_hip_field g,
                  qcd_hip_field b, int rb, double m, double tm, double *partials, hipStream_t stream) {
  DslashArgs a = make_args(in->lat, m, tm);
  const int ps = in->prec, pg = g->prec;
  if (ps == QCD_HIP_F64 && pg == QCD_HIP_F64)
    return launch_epi<QCD_HIP_F64, QCD_HIP_F64>(epi, red, a, out, ro, in, ri, g, b, rb, partials, stream);
  if (ps == QCD_HIP_F32 && pg == QCD_HIP_F32)
    return launch_epi<QCD_HIP_F32, QCD_HIP_F32>(epi, red, a, out, ro, in, ri, g, b, rb, partials, stream);
  if (ps == QCD_HIP_F32 && pg == QCD_HIP_F16)
    return launch_epi<QCD_HIP_F32, QCD_HIP_F16>(epi, red, a, out, ro, in, ri, g, b, rb, partials, stream);
  if (ps == QCD_HIP_F16 && pg == QCD_HIP_F16)
    return launch_epi<QCD_HIP_F16, QCD_HIP_F16>(epi, red, a, out, ro, in, ri, g, b, rb, partials, stream);
  set_error("dslash: unsupported precision pair (spinor %d, gauge %d)", ps, pg);
  return 1;
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
  for (int r = 0; r < in->nrhs; r++)
    QH_TRY(dslash_launch(epi, false, out, r, in, r, gauge, nullptr, 0, m, tm, nullptr, ctx()->stream));
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
  for (int i = 0; i < warmup; i++)
    for (int r = 0; r < in->nrhs; r++) QH_TRY(dslash_launch(epi, false, out, r, in, r, gauge, nullptr, 0, m, tm, nullptr, st));
  hipEvent_t e0, e1;
  QH_CHECK(hipEventCreate(&e0));
  QH_CHECK(hipEventCreate(&e1));
  QH_CHECK(hipEventRecord(e0, st));
  for (int i = 0; i < iters; i++)
    for (int r = 0; r < in->nrhs; r++) QH_TRY(dslash_launch(epi, false, out, r, in, r, gauge, nullptr, 0, m, tm, nullptr, st));
  QH_CHECK(hipEventRecord(e1, st));
  QH_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  QH_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / iters;
  return 0;
}

}  // extern "C"
