// smear.hip -- 3D Gaussian (Wuppertal) smearing on gfx950.
//
// One iteration on time-slice t (lib/src/qcd_smearing.c:25-107):
//   v(x) <- ( v(x) + alpha * sum_{nu=1..3} [U_nu(x) v(x+nu) + U_nu(x-nu)^+ v(x-nu)] ) / (1 + 6 alpha)
// The hop stays inside the time-slice, so an iteration on one slice reads and
// writes only that slice: V/T sites instead of the reference's whole-volume
// zero / add passes (lib/src/qcd_smearing.c:45, :102).  Iterations ping-pong
// between the field and a slice-sized-or-larger scratch field; the scale and
// add of the reference are fused into the hop kernel.
#include "blas_common.h"

namespace qcdhip {

template <int P>
__global__ void __launch_bounds__(kBlock) k_gauss(DslashArgs a, SpinorIn<P> in, GaugeIn<P> U, SpinorOut<P> out,
                                                  double alpha, double nrm, int t0, unsigned nsites) {
  using R = typename Prec<P>::R;
  const unsigned j = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (j >= nsites) return;
  const unsigned s = t0 >= 0 ? (unsigned)t0 + a.L[0] * j : j;
  const Site st = site_coords(s, a);
  Cx<R> acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = Cx<R>{R(0), R(0)};
  Cx<R> psi[12], u[9], y[3];
  auto step = [&](unsigned sf, unsigned sb, int nu) {
    Cx<R> up[12], ud[12];
    in.load(sf, psi);
    U.load(nu, s, u);
#pragma unroll
    for (int sp = 0; sp < 4; sp++) {
      su3_mul<false>(y, u, &psi[3 * sp]);
#pragma unroll
      for (int c = 0; c < 3; c++) up[3 * sp + c] = y[c];
    }
    in.load(sb, psi);
    U.load(nu, sb, u);
#pragma unroll
    for (int sp = 0; sp < 4; sp++) {
      su3_mul<true>(y, u, &psi[3 * sp]);
#pragma unroll
      for (int c = 0; c < 3; c++) ud[3 * sp + c] = y[c];
    }
#pragma unroll
    for (int k = 0; k < 12; k++) {
      acc[k].re += up[k].re + ud[k].re;
      acc[k].im += up[k].im + ud[k].im;
    }
  };
  step(neighbour<1, +1>(st, a), neighbour<1, -1>(st, a), 1);
  step(neighbour<2, +1>(st, a), neighbour<2, -1>(st, a), 2);
  step(neighbour<3, +1>(st, a), neighbour<3, -1>(st, a), 3);
  in.load(s, psi);
  const R al = (R)alpha, nr = (R)nrm;
#pragma unroll
  for (int k = 0; k < 12; k++)
    psi[k] = Cx<R>{(psi[k].re + acc[k].re * al) * nr, (psi[k].im + acc[k].im * al) * nr};
  out.store(s, psi);
}

template <int P>
static int gauss_launch(const DslashArgs &a, qcd_hip_field in, int ri, qcd_hip_field g, qcd_hip_field out, int ro,
                        double alpha, int t0, hipStream_t st) {
  const unsigned n = t0 >= 0 ? a.V / a.L[0] : a.V;
  k_gauss<P><<<grid_for(n), kBlock, 0, st>>>(a, SpinorIn<P>::of(in, ri), GaugeIn<P>::of(g), SpinorOut<P>::of(out, ro),
                                             alpha, 1.0 / (1.0 + 6.0 * alpha), t0, n);
  QH_CHECK(hipGetLastError());
  return 0;
}

int gauss_iterate(qcd_hip_field v, qcd_hip_field g, qcd_hip_field tmp, double alpha, int t0, int niter,
                  hipStream_t st) {
  DslashArgs a = make_args(v->lat, 0.0, 0.0);
  for (int r = 0; r < v->nrhs; r++) {
    for (int it = 0; it < niter; it++) {
      const bool from_v = (it % 2) == 0;
      qcd_hip_field in = from_v ? v : tmp, out = from_v ? tmp : v;
      const int ri = from_v ? r : 0, ro = from_v ? 0 : r;
      if (v->prec == QCD_HIP_F64) QH_TRY(gauss_launch<QCD_HIP_F64>(a, in, ri, g, out, ro, alpha, t0, st));
      else if (v->prec == QCD_HIP_F32) QH_TRY(gauss_launch<QCD_HIP_F32>(a, in, ri, g, out, ro, alpha, t0, st));
      else QH_TRY(gauss_launch<QCD_HIP_F16>(a, in, ri, g, out, ro, alpha, t0, st));
    }
    if (niter % 2) QH_TRY(lincomb_launch(v, r, 1.0, 0.0, tmp, 0, 0.0, 0.0, nullptr, 0, t0, st));
  }
  return 0;
}

}  // namespace qcdhip

using namespace qcdhip;

extern "C" {

int qcd_hip_gauss_smear(qcd_hip_field v, qcd_hip_field gauge, double alpha, int t, int niter) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());   // shared scratch field and stream
  QH_TRY(ensure_init());
  if (!v || !gauge || v->kind != QCD_HIP_SPINOR || gauge->kind != QCD_HIP_GAUGE || v->lat->V != gauge->lat->V) {
    qcdhip::set_error("gauss_smear: expected a spinor and a gauge field on one lattice");
    return 1;
  }
  if (v->prec != gauge->prec) {
    qcdhip::set_error("gauss_smear: spinor and gauge precision differ");
    return 1;
  }
  if (t >= v->lat->L[0] || niter < 0) {
    qcdhip::set_error("gauss_smear: time-slice %d outside 0..%d or negative iteration count", t, v->lat->L[0] - 1);
    return 1;
  }
  if (niter == 0) return 0;
  qcd_hip_field tmp = nullptr;
  QH_TRY(scratch_spinor(0, v->lat, v->prec, 1, &tmp));
  QH_TRY(gauss_iterate(v, gauge, tmp, alpha, t < 0 ? -1 : t, niter, ctx()->stream));
  QH_CHECK(hipStreamSynchronize(ctx()->stream));
  return 0;
}

}  // extern "C"
