// solver.hip -- the propagator solve of invert.c on the GPU.
//
// Outer loop (invert.c:255-294): true residual r = b - A x in FP64, stop when
// ||r|| / ||b|| <= tol, otherwise solve A e = r with GCR(n) and add e to x.
// Inner GCR(n) (invert.c:20-113, algorithm gcr.m:3-48) in the inner precision
// (FP64 for the plain solver, FP32 or FP16 storage for the mixed one) on the
// operator A = Q = i g5 D (as invert.c) or A = D.  For A = Q the system solved
// is Q x = i g5 b, i.e. D x = b, so the residual is ||b - D x|| either way.
//
// GCR iteration k (w_k = A res_k, chi_k = w_k / |w_k|, unnormalised w kept):
//   multi-dot:   d_j = <w_j, w_k>, j < k                    (reads k+1 vectors)
//   scalars:     a[k][j] = -s_j d_j, coefficient -s_j^2 d_j  (1 workgroup)
//   multi-axpy:  w_k += sum_j coef_j w_j, with |w_k|^2 and <w_k, src>
//   scalars:     s_k = 1/|w_k|, c_k = s_k <w_k, src>, invert.c's recurrence
//                a[k][j] <- a[k][j] a[j][j] + sum_{l=j+1}^{k-1} a[k][l] a[l][j]
//   axpy:        res_{k+1} = res_k - c_k s_k w_k
// After n iterations one multi-axpy forms x = sum_j beta_j res_j with
// beta_j = sum_{k>=j} c_k a[k][j] (the reference updates x every iteration).
// All coefficients live in device memory: a cycle needs no host round trip.
// Orthogonalisation is classical Gram-Schmidt in one fused pass (the
// reference's modified Gram-Schmidt, invert.c:68-73, is the same projection in
// exact arithmetic); the outer FP64 residual absorbs the difference.
#include "blas_common.h"

#include <chrono>
#include <cmath>

namespace qcdhip {

constexpr int kN = kMaxKrylov;

struct GcrScal {
  double2 a[kN * kN];
  double2 c[kN];
  double s[kN];
  double2 coef[kN];
  double2 beta[kN];
  double2 alpha;
  double rnorm;
  double rscale;
  int bad;
};

__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Sum nvals doubles over the workgroup; lane 0 of the group writes
// out[v * nb + slot] for v < nvals.  All lanes must call it.
template <int NV>
__device__ __forceinline__ void group_sum(double (&v)[NV], double *out, unsigned nb, unsigned slot) {
  __shared__ double sh[NV][kBlock / 64];
#pragma unroll
  for (int i = 0; i < NV; i++)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; i++) sh[i][wid] = v[i];
  __syncthreads();
  if (threadIdx.x < NV) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) t += sh[threadIdx.x][w];
    out[(size_t)threadIdx.x * nb + slot] = t;
  }
  __syncthreads();
}

// d_j = <w_j, w_k> for j < k; partials [2k][nb]
template <int P>
__global__ void __launch_bounds__(kBlock) k_multidot(SpinorSet<P> W, int k, unsigned V, double *part) {
  using R = typename Prec<P>::R;
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  const bool ok = s < V;
  Cx<R> wk[12];
  if (ok) W.load(k, s, wk);
  for (int j = 0; j < k; j++) {
    double v[2] = {0.0, 0.0};
    if (ok) {
      Cx<R> wj[12];
      W.load(j, s, wj);
#pragma unroll
      for (int c = 0; c < 12; c++) {
        v[0] += (double)wj[c].re * (double)wk[c].re + (double)wj[c].im * (double)wk[c].im;
        v[1] += (double)wj[c].re * (double)wk[c].im - (double)wj[c].im * (double)wk[c].re;
      }
    }
    group_sum<2>(v, part + (size_t)2 * j * gridDim.x, gridDim.x, blockIdx.x);
  }
}

// w_k += sum_j coef_j w_j; partials [3][nb] = |w_k|^2, Re/Im <w_k, src>
template <int P>
__global__ void __launch_bounds__(kBlock) k_multiaxpy_norm(SpinorSet<P> W, int k, const GcrScal *sc, SpinorSet<P> S,
                                                           int js, unsigned V, double *part) {
  using R = typename Prec<P>::R;
  __shared__ double2 cf[kN];
  for (int j = threadIdx.x; j < k; j += kBlock) cf[j] = sc->coef[j];
  __syncthreads();
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (s < V) {
    Cx<R> wk[12];
    W.load(k, s, wk);
    if (k > 0) {
      Cx<double> acc[12];
#pragma unroll
      for (int c = 0; c < 12; c++) acc[c] = Cx<double>{(double)wk[c].re, (double)wk[c].im};
      for (int j = 0; j < k; j++) {
        Cx<R> wj[12];
        W.load(j, s, wj);
        const double2 a = cf[j];
#pragma unroll
        for (int c = 0; c < 12; c++) {
          acc[c].re += a.x * (double)wj[c].re - a.y * (double)wj[c].im;
          acc[c].im += a.x * (double)wj[c].im + a.y * (double)wj[c].re;
        }
      }
#pragma unroll
      for (int c = 0; c < 12; c++) wk[c] = Cx<R>{(R)acc[c].re, (R)acc[c].im};
      W.store(k, s, wk);
      if constexpr (P == QCD_HIP_F16) W.load(k, s, wk);   // norms of the stored (scaled) vector
    }
    Cx<R> sv[12];
    S.load(js, s, sv);
#pragma unroll
    for (int c = 0; c < 12; c++) {
      const double wr = wk[c].re, wi = wk[c].im;
      v[0] += wr * wr + wi * wi;
      v[1] += wr * (double)sv[c].re + wi * (double)sv[c].im;
      v[2] += wr * (double)sv[c].im - wi * (double)sv[c].re;
    }
  }
  group_sum<3>(v, part, gridDim.x, blockIdx.x);
}

// res_{k+1} = res_k + alpha w_k
template <int P>
__global__ void __launch_bounds__(kBlock) k_res_update(SpinorSet<P> Rs, SpinorSet<P> W, int k, const GcrScal *sc,
                                                       unsigned V) {
  using R = typename Prec<P>::R;
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  const double2 al = sc->alpha;
  Cx<R> r[12], w[12];
  Rs.load(k, s, r);
  W.load(k, s, w);
#pragma unroll
  for (int c = 0; c < 12; c++) {
    const double wr = w[c].re, wi = w[c].im;
    r[c] = Cx<R>{(R)((double)r[c].re + al.x * wr - al.y * wi), (R)((double)r[c].im + al.x * wi + al.y * wr)};
  }
  Rs.store(k + 1, s, r);
}

// e = sum_j beta_j res_j
template <int P>
__global__ void __launch_bounds__(kBlock) k_solution(SpinorSet<P> Rs, int n, const GcrScal *sc, SpinorSet<P> E, int je,
                                                     unsigned V) {
  using R = typename Prec<P>::R;
  __shared__ double2 bt[kN];
  for (int j = threadIdx.x; j < n; j += kBlock) bt[j] = sc->beta[j];
  __syncthreads();
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  Cx<double> acc[12];
#pragma unroll
  for (int c = 0; c < 12; c++) acc[c] = Cx<double>{0.0, 0.0};
  for (int j = 0; j < n; j++) {
    Cx<R> r[12];
    Rs.load(j, s, r);
    const double2 b = bt[j];
#pragma unroll
    for (int c = 0; c < 12; c++) {
      acc[c].re += b.x * (double)r[c].re - b.y * (double)r[c].im;
      acc[c].im += b.x * (double)r[c].im + b.y * (double)r[c].re;
    }
  }
  Cx<R> o[12];
#pragma unroll
  for (int c = 0; c < 12; c++) o[c] = Cx<R>{(R)acc[c].re, (R)acc[c].im};
  E.store(je, s, o);
}

// Sum a [nv][nb] partial array (one wave per value, fixed order) into sums.
__device__ void wave_sums(const double *part, int nv, unsigned nb, double *sums) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int v = wid; v < nv; v += nw) {
    double t = 0.0;
    for (unsigned i = lane; i < nb; i += 64) t += part[(size_t)v * nb + i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0) sums[v] = t;
  }
}

// after the multi-dot of iteration k
__global__ void __launch_bounds__(1024) k_scal_dots(GcrScal *sc, const double *part, unsigned nb, int k, int n) {
  __shared__ double sums[2 * kN];
  wave_sums(part, 2 * k, nb, sums);
  __syncthreads();
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    const double sj = sc->s[j];
    const double2 d = make_double2(sums[2 * j], sums[2 * j + 1]);
    sc->a[k * n + j] = make_double2(-sj * d.x, -sj * d.y);
    sc->coef[j] = make_double2(-sj * sj * d.x, -sj * sj * d.y);
  }
}

// after the multi-axpy of iteration k: normalisation, c_k and the
// coefficient recurrence of invert.c:83-94
__global__ void __launch_bounds__(1024) k_scal_norm(GcrScal *sc, const double *part, unsigned nb, int k, int n) {
  __shared__ double sums[3];
  wave_sums(part, 3, nb, sums);
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double sk = 1.0 / sqrt(sums[0]);
  if (!(sk == sk) || isinf(sk)) sc->bad = 1;   // Krylov space is sick (invert.c:75-79)
  sc->s[k] = sk;
  sc->a[k * n + k] = make_double2(sk, 0.0);
  const double2 ck = make_double2(sk * sums[1], sk * sums[2]);
  sc->c[k] = ck;
  for (int j = 0; j < k; j++) {
    double2 &a = sc->a[k * n + j];
    a = make_double2(a.x * sk, a.y * sk);
  }
  for (int j = 0; j < k; j++) {
    double2 t = sc->a[k * n + j];
    const double ajj = sc->a[j * n + j].x;
    t = make_double2(t.x * ajj, t.y * ajj);
    for (int l = j + 1; l < k; l++) {
      const double2 p = cmul2(sc->a[k * n + l], sc->a[l * n + j]);
      t = make_double2(t.x + p.x, t.y + p.y);
    }
    sc->a[k * n + j] = t;
  }
  sc->alpha = make_double2(-ck.x * sk, -ck.y * sk);
}

// beta_j = sum_{k >= j} c_k a[k][j]; keep_q1 drops the k = 0 term as
// invert.c:50-53 does
__global__ void k_scal_beta(GcrScal *sc, int n, int keep_q1) {
  const int j = threadIdx.x;
  if (j >= n) return;
  double2 b = make_double2(0.0, 0.0);
  for (int k = j; k < n; k++) {
    if (k == 0 && keep_q1) continue;
    const double2 p = cmul2(sc->c[k], sc->a[k * n + j]);
    b = make_double2(b.x + p.x, b.y + p.y);
  }
  sc->beta[j] = b;
}

// outer residual norm from Dslash partials
__global__ void __launch_bounds__(1024) k_scal_rnorm(GcrScal *sc, const double *part, unsigned nb) {
  __shared__ double sums[1];
  wave_sums(part, 1, nb, sums);
  __syncthreads();
  if (threadIdx.x == 0) {
    sc->rnorm = sqrt(sums[0]);
    sc->rscale = sc->rnorm > 0.0 ? 1.0 / sc->rnorm : 0.0;
  }
}

// res_0 = r / |r| in the inner precision
template <int P>
__global__ void __launch_bounds__(kBlock) k_prepare_src(SpinorIn<QCD_HIP_F64> r, const GcrScal *sc, SpinorSet<P> Rs,
                                                        unsigned V) {
  using R = typename Prec<P>::R;
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  const double f = sc->rscale;
  Cx<double> v[12];
  r.load(s, v);
  Cx<R> o[12];
#pragma unroll
  for (int c = 0; c < 12; c++) o[c] = Cx<R>{(R)(v[c].re * f), (R)(v[c].im * f)};
  Rs.store(0, s, o);
}

// x += |r| e (FP64 accumulation of the inner correction)
template <int P>
__global__ void __launch_bounds__(kBlock) k_accumulate(SpinorIn<QCD_HIP_F64> xin, SpinorOut<QCD_HIP_F64> xout,
                                                       SpinorSet<P> E, int je, const GcrScal *sc, unsigned V) {
  using R = typename Prec<P>::R;
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  const double f = sc->rnorm;
  Cx<double> x[12];
  Cx<R> e[12];
  xin.load(s, x);
  E.load(je, s, e);
#pragma unroll
  for (int c = 0; c < 12; c++) x[c] = Cx<double>{x[c].re + f * (double)e[c].re, x[c].im + f * (double)e[c].im};
  xout.store(s, x);
}

// out = i g5 in (spins 0,1 times i, spins 2,3 times -i)
__global__ void __launch_bounds__(kBlock) k_ig5(SpinorIn<QCD_HIP_F64> in, SpinorOut<QCD_HIP_F64> out, unsigned V) {
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  Cx<double> v[12];
  in.load(s, v);
#pragma unroll
  for (int c = 0; c < 6; c++) v[c] = ipow<1>(v[c]);
#pragma unroll
  for (int c = 6; c < 12; c++) v[c] = ipow<3>(v[c]);
  out.store(s, v);
}

}  // namespace qcdhip

using namespace qcdhip;

// ============================================================ solver object
struct qcd_hip_solver_s {
  qcd_hip_lattice lat;
  qcd_hip_solver_params prm;
  qcd_hip_field g64 = nullptr;      // caller's FP64 gauge field (not owned)
  qcd_hip_field gin = nullptr;      // inner-precision copy (owned unless F64)
  qcd_hip_field Rs = nullptr;       // n search directions res_j (inner precision)
  qcd_hip_field Ws = nullptr;       // n vectors w_j = A res_j (inner precision)
  qcd_hip_field E = nullptr;        // inner correction e (inner precision)
  qcd_hip_field r64 = nullptr;      // FP64 residual
  qcd_hip_field bq = nullptr;       // FP64 right-hand side of the solved system
  GcrScal *sc = nullptr;
  double *part = nullptr;
  size_t part_bytes = 0;
  unsigned nb = 0;                   // site-kernel workgroups
};

namespace {

template <int P>
int gcr_cycle(qcd_hip_solver_s *S, double kappa, double mu, hipStream_t st) {
  const int n = S->prm.n_restart;
  const unsigned V = (unsigned)S->lat->V, nb = S->nb;
  const int epi = S->prm.op == QCD_HIP_OP_Q ? EPI_Q : EPI_D;
  const double m = 1.0 / (2.0 * kappa) - 4.0;
  SpinorSet<P> Rs = SpinorSet<P>::of(S->Rs), Ws = SpinorSet<P>::of(S->Ws), E = SpinorSet<P>::of(S->E);
  for (int k = 0; k < n; k++) {
    QH_TRY(dslash_launch(epi, false, S->Ws, k, S->Rs, k, S->gin, nullptr, 0, m, mu, nullptr, st, nullptr));
    if (k > 0) {
      k_multidot<P><<<nb, kBlock, 0, st>>>(Ws, k, V, S->part);
      k_scal_dots<<<1, 1024, 0, st>>>(S->sc, S->part, nb, k, n);
    }
    k_multiaxpy_norm<P><<<nb, kBlock, 0, st>>>(Ws, k, S->sc, Rs, 0, V, S->part);
    k_scal_norm<<<1, 1024, 0, st>>>(S->sc, S->part, nb, k, n);
    if (k + 1 < n) k_res_update<P><<<nb, kBlock, 0, st>>>(Rs, Ws, k, S->sc, V);
  }
  k_scal_beta<<<1, kN, 0, st>>>(S->sc, n, S->prm.keep_q1);
  k_solution<P><<<nb, kBlock, 0, st>>>(Rs, n, S->sc, E, 0, V);
  QH_CHECK(hipGetLastError());
  return 0;
}

int inner_cycle(qcd_hip_solver_s *S, double kappa, double mu, hipStream_t st) {
  switch (S->prm.inner_prec) {
    case QCD_HIP_F64: return gcr_cycle<QCD_HIP_F64>(S, kappa, mu, st);
    case QCD_HIP_F32: return gcr_cycle<QCD_HIP_F32>(S, kappa, mu, st);
    default: return gcr_cycle<QCD_HIP_F16>(S, kappa, mu, st);
  }
}

template <int P>
void launch_prepare(qcd_hip_solver_s *S, hipStream_t st) {
  k_prepare_src<P><<<S->nb, kBlock, 0, st>>>(SpinorIn<QCD_HIP_F64>::of(S->r64, 0), S->sc, SpinorSet<P>::of(S->Rs),
                                             (unsigned)S->lat->V);
}

template <int P>
void launch_accumulate(qcd_hip_solver_s *S, qcd_hip_field x, int rx, hipStream_t st) {
  k_accumulate<P><<<S->nb, kBlock, 0, st>>>(SpinorIn<QCD_HIP_F64>::of(x, rx), SpinorOut<QCD_HIP_F64>::of(x, rx),
                                            SpinorSet<P>::of(S->E), 0, S->sc, (unsigned)S->lat->V);
}

// r = bq - A x with |r|^2 into sc->rnorm; returns |r| on the host
int true_residual(qcd_hip_solver_s *S, qcd_hip_field x, int rx, double kappa, double mu, hipStream_t st,
                  double *rnorm) {
  const int epi = S->prm.op == QCD_HIP_OP_Q ? EPI_RES_Q : EPI_RES_D;
  int grid = 0;
  QH_TRY(dslash_launch(epi, true, S->r64, 0, x, rx, S->g64, S->bq, 0, 1.0 / (2.0 * kappa) - 4.0, mu, S->part, st,
                       &grid));
  k_scal_rnorm<<<1, 1024, 0, st>>>(S->sc, S->part, (unsigned)grid);
  QH_CHECK(hipGetLastError());
  Context *c = ctx();
  QH_CHECK(hipMemcpyAsync(c->host_scalars, &S->sc->rnorm, sizeof(double), hipMemcpyDeviceToHost, st));
  QH_CHECK(hipStreamSynchronize(st));
  *rnorm = c->host_scalars[0];
  return 0;
}

}  // namespace

extern "C" {

void qcd_hip_solver_default_params(qcd_hip_solver_params *p) {
  std::memset(p, 0, sizeof(*p));
  p->n_restart = 32;
  p->max_outer = 10000;
  p->tol = 1e-10;
  p->inner_prec = QCD_HIP_F64;
  p->op = QCD_HIP_OP_Q;
  p->keep_q1 = 0;
}

int qcd_hip_solver_create(qcd_hip_solver *out, qcd_hip_lattice lat, const qcd_hip_solver_params *prm) {
  *out = nullptr;
  QH_TRY(ensure_init());
  if (!lat || !prm || prm->n_restart < 1 || prm->n_restart > kN || prm->tol <= 0.0 || prm->max_outer < 1 ||
      prm->inner_prec < QCD_HIP_F64 || prm->inner_prec > QCD_HIP_F16 ||
      (prm->op != QCD_HIP_OP_Q && prm->op != QCD_HIP_OP_D)) {
    set_error("solver_create: bad parameters (n_restart 1..%d, tol > 0, max_outer >= 1)", kN);
    return 1;
  }
  auto *S = new qcd_hip_solver_s();
  S->lat = lat;
  S->prm = *prm;
  S->nb = (unsigned)grid_for(lat->V);
  int rc = qcd_hip_spinor_create(&S->Rs, lat, prm->inner_prec, prm->n_restart);
  if (!rc) rc = qcd_hip_spinor_create(&S->Ws, lat, prm->inner_prec, prm->n_restart);
  if (!rc) rc = qcd_hip_spinor_create(&S->E, lat, prm->inner_prec, 1);
  if (!rc) rc = qcd_hip_spinor_create(&S->r64, lat, QCD_HIP_F64, 1);
  if (!rc) rc = qcd_hip_spinor_create(&S->bq, lat, QCD_HIP_F64, 1);
  if (!rc) rc = dev_alloc((void **)&S->sc, sizeof(GcrScal));
  // partials: 2 (n-1) per site workgroup (multi-dot) or one per Dslash workgroup
  S->part_bytes = sizeof(double) * (size_t)std::max<size_t>(2 * (size_t)prm->n_restart + 4, 8) * (S->nb + 64);
  if (!rc) rc = dev_alloc((void **)&S->part, S->part_bytes);
  if (rc) {
    qcd_hip_solver_destroy(S);
    return rc;
  }
  *out = S;
  return 0;
}

void qcd_hip_solver_destroy(qcd_hip_solver S) {
  if (!S) return;
  qcd_hip_field_destroy(S->Rs);
  qcd_hip_field_destroy(S->Ws);
  qcd_hip_field_destroy(S->E);
  qcd_hip_field_destroy(S->r64);
  qcd_hip_field_destroy(S->bq);
  if (S->gin && S->gin != S->g64) qcd_hip_field_destroy(S->gin);
  (void)hipStreamSynchronize(ctx()->stream);
  dev_free(S->sc, sizeof(GcrScal));
  dev_free(S->part, S->part_bytes);
  delete S;
}

int qcd_hip_solver_set_gauge(qcd_hip_solver S, qcd_hip_field g64) {
  QH_TRY(ensure_init());
  if (!S || !g64 || g64->kind != QCD_HIP_GAUGE || g64->prec != QCD_HIP_F64 || g64->lat->V != S->lat->V) {
    set_error("solver_set_gauge: need an FP64 gauge field on the solver's lattice");
    return 1;
  }
  S->g64 = g64;
  if (S->prm.inner_prec == QCD_HIP_F64) {
    if (S->gin && S->gin != g64) qcd_hip_field_destroy(S->gin);
    S->gin = g64;
    return 0;
  }
  if (!S->gin || S->gin == g64) {
    S->gin = nullptr;
    QH_TRY(qcd_hip_gauge_create(&S->gin, S->lat, S->prm.inner_prec));
  }
  QH_TRY(qcd_hip_field_convert(S->gin, g64));
  return 0;
}

int qcd_hip_solver_solve(qcd_hip_solver S, qcd_hip_field x, qcd_hip_field b, double kappa, double mu,
                         qcd_hip_solver_stats *stats) {
  QH_TRY(ensure_init());
  if (!S || !S->g64) {
    set_error("solver_solve: no gauge field set (qcd_hip_solver_set_gauge)");
    return 1;
  }
  if (!x || !b || x->kind != QCD_HIP_SPINOR || b->kind != QCD_HIP_SPINOR || x->prec != QCD_HIP_F64 ||
      b->prec != QCD_HIP_F64 || x->nrhs != b->nrhs || x->lat->V != S->lat->V || b->lat->V != S->lat->V) {
    set_error("solver_solve: x and b must be FP64 spinors with equal rhs counts on the solver's lattice");
    return 1;
  }
  if (kappa <= 0.0) {
    set_error("solver_solve: kappa must be positive");
    return 1;
  }
  hipStream_t st = ctx()->stream;
  const auto t0 = std::chrono::steady_clock::now();
  // Tune the inner and outer operators before the first timed use.
  QH_TRY(dslash_prepare(S->Ws, S->Rs, S->gin));
  QH_TRY(dslash_prepare(S->r64, x, S->g64));
  const unsigned V = (unsigned)S->lat->V;
  int worst = 0;
  qcd_hip_solver_stats tot;
  std::memset(&tot, 0, sizeof(tot));
  for (int r = 0; r < b->nrhs; r++) {
    // system right-hand side: b, or i g5 b when solving with Q
    if (S->prm.op == QCD_HIP_OP_Q)
      k_ig5<<<S->nb, kBlock, 0, st>>>(SpinorIn<QCD_HIP_F64>::of(b, r), SpinorOut<QCD_HIP_F64>::of(S->bq, 0), V);
    else
      QH_TRY(lincomb_launch(S->bq, 0, 1.0, 0.0, b, r, 0.0, 0.0, nullptr, 0, -1, st));
    QH_TRY(lincomb_launch(x, r, 0.0, 0.0, x, r, 0.0, 0.0, nullptr, 0, -1, st));   // x = 0
    QH_CHECK(hipMemsetAsync(S->sc, 0, sizeof(GcrScal), st));
    double bnorm = 0.0, rn = 0.0;
    QH_TRY(true_residual(S, x, r, kappa, mu, st, &bnorm));
    int outer = 0, bad = 0;
    double rel = bnorm > 0.0 ? 1.0 : 0.0;
    while (rel > S->prm.tol && outer < S->prm.max_outer) {
      if (S->prm.inner_prec == QCD_HIP_F64) launch_prepare<QCD_HIP_F64>(S, st);
      else if (S->prm.inner_prec == QCD_HIP_F32) launch_prepare<QCD_HIP_F32>(S, st);
      else launch_prepare<QCD_HIP_F16>(S, st);
      QH_TRY(inner_cycle(S, kappa, mu, st));
      if (S->prm.inner_prec == QCD_HIP_F64) launch_accumulate<QCD_HIP_F64>(S, x, r, st);
      else if (S->prm.inner_prec == QCD_HIP_F32) launch_accumulate<QCD_HIP_F32>(S, x, r, st);
      else launch_accumulate<QCD_HIP_F16>(S, x, r, st);
      QH_CHECK(hipGetLastError());
      QH_TRY(true_residual(S, x, r, kappa, mu, st, &rn));
      outer++;
      rel = rn / bnorm;
      int flag = 0;
      QH_CHECK(hipMemcpy(&flag, &S->sc->bad, sizeof(int), hipMemcpyDeviceToHost));
      if (flag || !(rel == rel)) {
        bad = 1;
        break;
      }
    }
    tot.outer += outer;
    tot.inner_iters += outer * S->prm.n_restart;
    tot.relres = std::max(tot.relres, rel);
    if (bad) worst = std::max(worst, 5);
    else if (rel > S->prm.tol) worst = std::max(worst, 6);
  }
  tot.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = tot;
  if (worst == 5) set_error("solver: Krylov space breakdown (NaN or zero norm)");
  if (worst == 6) set_error("solver: not converged to %g within %d outer iterations (reached %g)", S->prm.tol,
                            S->prm.max_outer, tot.relres);
  return worst;
}

// ||D x - b|| / ||b|| in FP64 with m = 1/(2 kappa) - 4 (b_minus_Dx.c:152-167)
int qcd_hip_b_minus_Dx(qcd_hip_field x, int rx, qcd_hip_field b, int rb, qcd_hip_field g, double kappa, double mu,
                       double *rel) {
  QH_TRY(ensure_init());
  if (!x || !b || !g || x->prec != QCD_HIP_F64 || b->prec != QCD_HIP_F64 || g->prec != QCD_HIP_F64 || !rel ||
      rx < 0 || rx >= x->nrhs || rb < 0 || rb >= b->nrhs) {
    set_error("b_minus_Dx: needs FP64 x, b and gauge field");
    return 1;
  }
  QH_TRY(dslash_check(x, b, g));
  qcd_hip_field r = nullptr;
  QH_TRY(scratch_spinor(1, x->lat, QCD_HIP_F64, 1, &r));
  hipStream_t st = ctx()->stream;
  QH_TRY(dslash_prepare(r, x, g));
  int grid = 0;
  Context *c = ctx();
  const size_t need = sizeof(double) * (size_t)(dslash_grid(x, g) + 64);
  double *part = nullptr;
  QH_TRY(dev_alloc((void **)&part, need));
  int rc = dslash_launch(EPI_RES_D, true, r, 0, x, rx, g, b, rb, 1.0 / (2.0 * kappa) - 4.0, mu, part, st, &grid);
  if (!rc) rc = finalize(part, grid, 1, c->scalars, st);
  double num = 0.0, den = 0.0;
  if (!rc) rc = hipMemcpyAsync(&num, c->scalars, sizeof(double), hipMemcpyDeviceToHost, st) ? 2 : 0;
  if (!rc) rc = hipStreamSynchronize(st) ? 2 : 0;
  dev_free(part, need);
  QH_TRY(rc);
  QH_TRY(qcd_hip_norm2(b, rb, &den));
  *rel = std::sqrt(num) / std::sqrt(den);
  return 0;
}

// One-shot convenience: create a solver, solve every rhs of b, destroy it.
int qcd_hip_invert(qcd_hip_field x, qcd_hip_field b, qcd_hip_field g64, double kappa, double mu,
                   const qcd_hip_solver_params *prm, qcd_hip_solver_stats *stats) {
  qcd_hip_solver S = nullptr;
  qcd_hip_solver_params def;
  if (!prm) {
    qcd_hip_solver_default_params(&def);
    prm = &def;
  }
  if (!b) {
    set_error("invert: null right-hand side");
    return 1;
  }
  QH_TRY(qcd_hip_solver_create(&S, b->lat, prm));
  int rc = qcd_hip_solver_set_gauge(S, g64);
  if (!rc) rc = qcd_hip_solver_solve(S, x, b, kappa, mu, stats);
  qcd_hip_solver_destroy(S);
  return rc;
}

}  // extern "C"
