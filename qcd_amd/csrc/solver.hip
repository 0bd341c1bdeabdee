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
// Right-hand sides are solved in batches of prm.batch that advance in
// lockstep: every kernel takes a second grid dimension over the batch, and
// right-hand sides that have converged drop out of the launches.
// Orthogonalisation is classical Gram-Schmidt in one fused pass (the
// reference's modified Gram-Schmidt, invert.c:68-73, is the same projection in
// exact arithmetic); the outer FP64 residual absorbs the difference.
#include "eo_common.h"

#include <chrono>
#include <cmath>

namespace qcdhip {

// eo.hip
int eo_split(qcd_hip_field full, int f0, qcd_hip_field ev, qcd_hip_field od, const int *ids, int count,
             hipStream_t st);
int eo_accumulate(qcd_hip_field x, int f0, qcd_hip_field ev, qcd_hip_field od, const int *ids, int count,
                  hipStream_t st);
int eo_minv(qcd_hip_lattice full, qcd_hip_field in, qcd_hip_field out, double m, double tm, const int *ids, int count,
            hipStream_t st);
int eo_scale64(qcd_hip_lattice full, qcd_hip_field in, qcd_hip_field out, const double *scale, size_t scale_stride,
               const int *ids, int count, hipStream_t st);
qcd_hip_lattice eo_half_lattice(qcd_hip_lattice full);

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

// Batched kernels: grid row y works on right-hand-side slot ids[y]; the
// Krylov vector j of slot b is vector b*n + j of the Rs / Ws fields; row y
// writes its partial sums at part + y*rowsz, laid out [value][workgroup].

// d_j = <w_j, w_k> for j < k, in chunks of kDotChunk basis vectors: their
// loads are in flight together and the workgroup reduces the chunk's sums
// at once (each d_j is summed exactly as it would be alone)
constexpr int kDotChunk = 4;
template <int P>
__global__ void __launch_bounds__(kBlock) k_multidot(SpinorSet<P> W, int n, int k, const int *ids, unsigned V,
                                                     double *part, size_t rowsz) {
  using R = typename Prec<P>::R;
  const int b = ids[blockIdx.y];
  double *pr = part + blockIdx.y * rowsz;
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  const bool ok = s < V;
  Cx<R> wk[12];
  if (ok) W.load(b * n + k, s, wk);
  for (int j0 = 0; j0 < k; j0 += kDotChunk) {
    double v[2 * kDotChunk];
#pragma unroll
    for (int i = 0; i < 2 * kDotChunk; i++) v[i] = 0.0;
    if (ok) {
#pragma unroll
      for (int i = 0; i < kDotChunk; i++) {
        if (j0 + i < k) {
          Cx<R> wj[12];
          W.load(b * n + j0 + i, s, wj);
#pragma unroll
          for (int c = 0; c < 12; c++) {
            v[2 * i] += (double)wj[c].re * (double)wk[c].re + (double)wj[c].im * (double)wk[c].im;
            v[2 * i + 1] += (double)wj[c].re * (double)wk[c].im - (double)wj[c].im * (double)wk[c].re;
          }
        }
      }
    }
    // sums of j >= k are written too (zeros) but never read
    group_sum<2 * kDotChunk>(v, pr + (size_t)2 * j0 * gridDim.x, gridDim.x, blockIdx.x);
  }
}

// w_k += sum_j coef_j w_j; partials: |w_k|^2, Re/Im <w_k, src = res_0>
template <int P>
__global__ void __launch_bounds__(kBlock) k_multiaxpy_norm(SpinorSet<P> W, SpinorSet<P> Rs, int n, int k,
                                                           const GcrScal *sc, const int *ids, unsigned V, double *part,
                                                           size_t rowsz) {
  using R = typename Prec<P>::R;
  const int b = ids[blockIdx.y];
  __shared__ double2 cf[kN];
  for (int j = threadIdx.x; j < k; j += kBlock) cf[j] = sc[b].coef[j];
  __syncthreads();
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (s < V) {
    Cx<R> wk[12];
    W.load(b * n + k, s, wk);
    if (k > 0) {
      Cx<double> acc[12];
#pragma unroll
      for (int c = 0; c < 12; c++) acc[c] = Cx<double>{(double)wk[c].re, (double)wk[c].im};
#pragma unroll 4
      for (int j = 0; j < k; j++) {
        Cx<R> wj[12];
        W.load(b * n + j, s, wj);
        const double2 a = cf[j];
#pragma unroll
        for (int c = 0; c < 12; c++) {
          acc[c].re += a.x * (double)wj[c].re - a.y * (double)wj[c].im;
          acc[c].im += a.x * (double)wj[c].im + a.y * (double)wj[c].re;
        }
      }
#pragma unroll
      for (int c = 0; c < 12; c++) wk[c] = Cx<R>{(R)acc[c].re, (R)acc[c].im};
      W.store_back(b * n + k, s, wk);   // the norms below are those of the stored (rounded) vector
    }
    Cx<R> sv[12];
    Rs.load(b * n, s, sv);
#pragma unroll
    for (int c = 0; c < 12; c++) {
      const double wr = wk[c].re, wi = wk[c].im;
      v[0] += wr * wr + wi * wi;
      v[1] += wr * (double)sv[c].re + wi * (double)sv[c].im;
      v[2] += wr * (double)sv[c].im - wi * (double)sv[c].re;
    }
  }
  group_sum<3>(v, part + blockIdx.y * rowsz, gridDim.x, blockIdx.x);
}

// res_{k+1} = res_k + alpha w_k
template <int P>
__global__ void __launch_bounds__(kBlock) k_res_update(SpinorSet<P> Rs, SpinorSet<P> W, int n, int k,
                                                       const GcrScal *sc, const int *ids, unsigned V) {
  using R = typename Prec<P>::R;
  const int b = ids[blockIdx.y];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  const double2 al = sc[b].alpha;
  Cx<R> r[12], w[12];
  Rs.load(b * n + k, s, r);
  W.load(b * n + k, s, w);
#pragma unroll
  for (int c = 0; c < 12; c++) {
    const double wr = w[c].re, wi = w[c].im;
    r[c] = Cx<R>{(R)((double)r[c].re + al.x * wr - al.y * wi), (R)((double)r[c].im + al.x * wi + al.y * wr)};
  }
  Rs.store(b * n + k + 1, s, r);
}

// e_b = sum_j beta_j res_j, times |r_b| (the norm the inner source was
// divided by) when scaled
template <int P>
__global__ void __launch_bounds__(kBlock) k_solution(SpinorSet<P> Rs, int n, const GcrScal *sc, SpinorSet<P> E,
                                                     const int *ids, unsigned V, int scaled) {
  using R = typename Prec<P>::R;
  const int b = ids[blockIdx.y];
  __shared__ double2 bt[kN];
  const double f = scaled ? sc[b].rnorm : 1.0;
  for (int j = threadIdx.x; j < n; j += kBlock) bt[j] = make_double2(sc[b].beta[j].x * f, sc[b].beta[j].y * f);
  __syncthreads();
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  Cx<double> acc[12];
#pragma unroll
  for (int c = 0; c < 12; c++) acc[c] = Cx<double>{0.0, 0.0};
#pragma unroll 4
  for (int j = 0; j < n; j++) {
    Cx<R> r[12];
    Rs.load(b * n + j, s, r);
    const double2 bb = bt[j];
#pragma unroll
    for (int c = 0; c < 12; c++) {
      acc[c].re += bb.x * (double)r[c].re - bb.y * (double)r[c].im;
      acc[c].im += bb.x * (double)r[c].im + bb.y * (double)r[c].re;
    }
  }
  Cx<R> o[12];
#pragma unroll
  for (int c = 0; c < 12; c++) o[c] = Cx<R>{(R)acc[c].re, (R)acc[c].im};
  E.store(b, s, o);
}

// Sum a [nv][nb] partial array (one wave per value, an order fixed by nb)
// into sums.  Each lane keeps four running sums over partials 256 apart, so
// its loads do not wait on one chain of additions (rows of 20 000 partials
// on 48^3x96).
__device__ void wave_sums(const double *part, int nv, unsigned nb, double *sums) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int v = wid; v < nv; v += nw) {
    const double *p = part + (size_t)v * nb;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
    unsigned i = lane;
    for (; i + 192 < nb; i += 256) {
      t0 += p[i];
      t1 += p[i + 64];
      t2 += p[i + 128];
      t3 += p[i + 192];
    }
    for (; i < nb; i += 64) t0 += p[i];
    double t = (t0 + t1) + (t2 + t3);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0) sums[v] = t;
  }
}

// after the multi-dot of iteration k (one workgroup per slot)
__global__ void __launch_bounds__(1024) k_scal_dots(GcrScal *scs, const double *part, size_t rowsz, unsigned nb, int k,
                                                    int n, const int *ids) {
  __shared__ double sums[2 * kN];
  GcrScal *sc = scs + ids[blockIdx.x];
  wave_sums(part + blockIdx.x * rowsz, 2 * k, nb, sums);
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
// Row k is scaled by s_k first; then every a[k][j] is updated from the scaled
// row and the finished rows l < k, so the j are independent: thread j does
// one of them, with rows 0..k-1 staged in LDS (packed lower triangle).
__global__ void __launch_bounds__(1024) k_scal_norm(GcrScal *scs, const double *part, size_t rowsz, unsigned nb, int k,
                                                    int n, const int *ids) {
  __shared__ double sums[3];
  __shared__ double2 rowk[kN];
  __shared__ double diag[kN];
  __shared__ double2 tri[kN * (kN - 1) / 2];   // a[l][j], j < l < k, at l (l - 1) / 2 + j
  GcrScal *sc = scs + ids[blockIdx.x];
  wave_sums(part + blockIdx.x * rowsz, 3, nb, sums);
  const int tid = threadIdx.x;
  for (int base = 0; base < k; base += 1024 / 64) {
    const int l = base + (tid >> 6), j = tid & 63;
    if (l < k && j < l) tri[l * (l - 1) / 2 + j] = sc->a[l * n + j];
  }
  if (tid < k) diag[tid] = sc->a[tid * n + tid].x;
  __syncthreads();
  const double sk = 1.0 / sqrt(sums[0]);
  if (tid < k) {
    const double2 a = sc->a[k * n + tid];
    rowk[tid] = make_double2(a.x * sk, a.y * sk);
  }
  __syncthreads();
  if (tid < k) {
    const int j = tid;
    double2 t = rowk[j];
    t = make_double2(t.x * diag[j], t.y * diag[j]);
    for (int l = j + 1; l < k; l++) {
      const double2 p = cmul2(rowk[l], tri[l * (l - 1) / 2 + j]);
      t = make_double2(t.x + p.x, t.y + p.y);
    }
    sc->a[k * n + j] = t;
  }
  if (tid == 0) {
    if (!(sk == sk) || isinf(sk)) sc->bad = 1;   // Krylov space is sick (invert.c:75-79)
    sc->s[k] = sk;
    sc->a[k * n + k] = make_double2(sk, 0.0);
    const double2 ck = make_double2(sk * sums[1], sk * sums[2]);
    sc->c[k] = ck;
    sc->alpha = make_double2(-ck.x * sk, -ck.y * sk);
  }
}

// beta_j = sum_{k >= j} c_k a[k][j]; keep_q1 drops the k = 0 term as
// invert.c:50-53 does
__global__ void k_scal_beta(GcrScal *scs, int n, int keep_q1, const int *ids) {
  GcrScal *sc = scs + ids[blockIdx.x];
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

// outer residual norms from the Dslash partials (row x of the launch)
__global__ void __launch_bounds__(1024) k_scal_rnorm(GcrScal *scs, const double *part, unsigned nb, const int *ids,
                                                     double *rn) {
  __shared__ double sums[1];
  const int b = ids[blockIdx.x];
  wave_sums(part + (size_t)blockIdx.x * nb, 1, nb, sums);
  __syncthreads();
  if (threadIdx.x == 0) {
    GcrScal *sc = scs + b;
    sc->rnorm = sqrt(sums[0]);
    sc->rscale = sc->rnorm > 0.0 ? 1.0 / sc->rnorm : 0.0;
    rn[b] = sc->bad ? __builtin_nan("") : sc->rnorm;   // a breakdown reaches the host as NaN
  }
}

// res_0 = r / |r| in the inner precision P (r in precision PI)
template <int PI, int P>
__global__ void __launch_bounds__(kBlock) k_prepare_src(SpinorSet<PI> r, const GcrScal *sc, SpinorSet<P> Rs, int n,
                                                        const int *ids, unsigned V) {
  using R = typename Prec<P>::R;
  using RI = typename Prec<PI>::R;
  const int b = ids[blockIdx.y];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  const double f = sc[b].rscale;
  Cx<RI> v[12];
  r.load(b, s, v);
  Cx<R> o[12];
#pragma unroll
  for (int c = 0; c < 12; c++) o[c] = Cx<R>{(R)((double)v[c].re * f), (R)((double)v[c].im * f)};
  Rs.store(b * n, s, o);
}

// x_{x0+b} += |r_b| e_b (FP64 accumulation of the inner correction)
template <int P>
__global__ void __launch_bounds__(kBlock) k_accumulate(SpinorSet<QCD_HIP_F64> X, int x0, SpinorSet<P> E,
                                                       const GcrScal *sc, const int *ids, unsigned V) {
  using R = typename Prec<P>::R;
  const int b = ids[blockIdx.y];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  const double f = sc[b].rnorm;
  Cx<double> x[12];
  Cx<R> e[12];
  X.load(x0 + b, s, x);
  E.load(b, s, e);
#pragma unroll
  for (int c = 0; c < 12; c++) x[c] = Cx<double>{x[c].re + f * (double)e[c].re, x[c].im + f * (double)e[c].im};
  X.store(x0 + b, s, x);
}

// system right-hand side of slot b (i g5 b for Q, b for D) and x = 0
__global__ void __launch_bounds__(kBlock) k_start(SpinorSet<QCD_HIP_F64> B, SpinorSet<QCD_HIP_F64> X, int r0,
                                                  SpinorSet<QCD_HIP_F64> Bq, int use_q, unsigned V) {
  const int b = blockIdx.y;
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  Cx<double> v[12];
  B.load(r0 + b, s, v);
  if (use_q) {
#pragma unroll
    for (int c = 0; c < 6; c++) v[c] = ipow<1>(v[c]);
#pragma unroll
    for (int c = 6; c < 12; c++) v[c] = ipow<3>(v[c]);
  }
  Bq.store(b, s, v);
#pragma unroll
  for (int c = 0; c < 12; c++) v[c] = Cx<double>{0.0, 0.0};
  X.store(r0 + b, s, v);
}

}  // namespace qcdhip

using namespace qcdhip;

// ============================================================ solver object
struct qcd_hip_solver_s {
  qcd_hip_lattice lat;
  qcd_hip_solver_params prm;
  int B = 1;                        // right-hand sides per batch
  hipStream_t stream = nullptr;     // the solver's own stream
  qcd_hip_field g64 = nullptr;      // caller's FP64 gauge field (not owned)
  qcd_hip_field gin = nullptr;      // inner-precision copy (owned unless F64)
  qcd_hip_field Rs = nullptr;       // n*B search directions res_j (inner precision)
  qcd_hip_field Ws = nullptr;       // n*B vectors w_j = A res_j (inner precision)
  qcd_hip_field E = nullptr;        // B inner corrections (inner precision)
  qcd_hip_field r64 = nullptr;      // B FP64 residuals (not allocated with split_res)
  // FP16-inner even-odd: the true residual is stored straight into the FP16
  // halves reh, roh and th = M^-1 reh, never as FP64
  bool split_res = false;
  qcd_hip_field bq = nullptr;       // B FP64 right-hand sides of the solved system
  GcrScal *sc = nullptr;            // B coefficient sets
  double *rn = nullptr;             // B residual norms (device)
  double *rn_host = nullptr;        // pinned mirror
  int *ids = nullptr;               // active slots (device)
  int *ids_host = nullptr;          // pinned mirror
  double *part = nullptr;
  size_t part_bytes = 0, rowsz = 0;
  unsigned nb = 0;                   // site-kernel workgroups (full lattice)
  // even-odd preconditioning (prm.eo): the inner GCR runs on the Schur
  // complement D^ over odd-site half vectors
  bool use_q = true;                 // outer system Q x = i g5 b (else D x = b)
  qcd_hip_lattice half = nullptr;    // half lattice of the inner vectors (owned)
  unsigned Vi = 0, nbi = 0;          // inner vector sites and site-kernel workgroups
  qcd_hip_field te = nullptr;        // EO: B inner even temporaries of the matvec
  // EO: B half vectors each for the outer step: even and odd residual halves,
  // M^-1 temporaries, the odd correction; FP64 for an FP64 inner solve, the
  // inner precision otherwise
  qcd_hip_field reh = nullptr, roh = nullptr, th = nullptr, eoh = nullptr;
  // the captured GCR cycle (prm.graph) and the launch arguments it was
  // captured with; every buffer it touches belongs to the solver
  hipGraphExec_t cycle = nullptr;
  int cycle_count = 0;
  double cycle_kappa = 0.0, cycle_mu = 0.0;
  void drop_cycle() {
    if (!cycle) return;
    (void)hipStreamSynchronize(stream);   // no replay may still be running
    (void)hipGraphExecDestroy(cycle);
    cycle = nullptr;
  }
};

namespace {

template <int P>
int gcr_cycle(qcd_hip_solver_s *S, int count, double kappa, double mu, hipStream_t st) {
  const int n = S->prm.n_restart;
  const unsigned V = S->Vi, nb = S->nbi;
  const int epi = S->use_q ? EPI_Q : EPI_D;
  const double m = 1.0 / (2.0 * kappa) - 4.0;
  SpinorSet<P> Rs = SpinorSet<P>::of(S->Rs), Ws = SpinorSet<P>::of(S->Ws), E = SpinorSet<P>::of(S->E);
  const dim3 g2(nb, (unsigned)count);
  for (int k = 0; k < n; k++) {
    if (S->prm.eo) {
      // w_k = D^ res_k = M res_k - H_oe (M^-1 H_eo res_k)
      QH_TRY(eo_launch(EO_MINV, false, S->te, 0, 1, S->Rs, k, n, S->gin, nullptr, 0, 0, S->ids, count, m, mu, 0,
                       nullptr, st, nullptr));
      QH_TRY(eo_launch(EO_SCHUR, false, S->Ws, k, n, S->te, 0, 1, S->gin, S->Rs, k, n, S->ids, count, m, mu, 1,
                       nullptr, st, nullptr));
    } else {
      QH_TRY(dslash_launch_batch(epi, false, S->Ws, k, n, S->Rs, k, n, S->gin, nullptr, 0, 0, S->ids, count, m, mu,
                                 nullptr, st, nullptr));
    }
    if (k > 0) {
      k_multidot<P><<<g2, kBlock, 0, st>>>(Ws, n, k, S->ids, V, S->part, S->rowsz);
      k_scal_dots<<<count, 1024, 0, st>>>(S->sc, S->part, S->rowsz, nb, k, n, S->ids);
    }
    k_multiaxpy_norm<P><<<g2, kBlock, 0, st>>>(Ws, Rs, n, k, S->sc, S->ids, V, S->part, S->rowsz);
    k_scal_norm<<<count, 1024, 0, st>>>(S->sc, S->part, S->rowsz, nb, k, n, S->ids);
    if (k + 1 < n) k_res_update<P><<<g2, kBlock, 0, st>>>(Rs, Ws, n, k, S->sc, S->ids, V);
  }
  k_scal_beta<<<count, kN, 0, st>>>(S->sc, n, S->prm.keep_q1, S->ids);
  // mixed even-odd solves take the correction scaled by |r^_o| (the outer
  // step adds it as it is); the others scale it where they add it
  const int scaled = S->prm.eo && S->prm.inner_prec != QCD_HIP_F64;
  k_solution<P><<<g2, kBlock, 0, st>>>(Rs, n, S->sc, E, S->ids, V, scaled);
  QH_CHECK(hipGetLastError());
  return 0;
}

// One GCR cycle on stream st.  With prm.graph the launches of gcr_cycle are
// captured into a graph once per (active rhs count, kappa, mu) and the graph
// is replayed: the kernels read the active-slot list and every coefficient
// from device memory, so a replay is the same work as the eager launches.
// The cycle records no memset or copy nodes.
template <int P>
int run_cycle(qcd_hip_solver_s *S, int count, double kappa, double mu, hipStream_t st) {
  if (!S->prm.graph) return gcr_cycle<P>(S, count, kappa, mu, st);
  if (!S->cycle || S->cycle_count != count || S->cycle_kappa != kappa || S->cycle_mu != mu) {
    S->drop_cycle();
    if (S->prm.eo) {
      // the cycle's first matvec once, eagerly: hop launch shapes seen for the
      // first time are tuned here, which cannot happen under capture (both
      // results are recomputed by the cycle)
      const double m = 1.0 / (2.0 * kappa) - 4.0;
      const int n = S->prm.n_restart;
      QH_TRY(eo_launch(EO_MINV, false, S->te, 0, 1, S->Rs, 0, n, S->gin, nullptr, 0, 0, S->ids, count, m, mu, 0,
                       nullptr, st, nullptr));
      QH_TRY(eo_launch(EO_SCHUR, false, S->Ws, 0, n, S->te, 0, 1, S->gin, S->Rs, 0, n, S->ids, count, m, mu, 1,
                       nullptr, st, nullptr));
    }
    QH_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    const int rc = gcr_cycle<P>(S, count, kappa, mu, st);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    QH_CHECK(e);
    const hipError_t ei = hipGraphInstantiate(&S->cycle, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) S->cycle = nullptr;
    QH_CHECK(ei);
    S->cycle_count = count;
    S->cycle_kappa = kappa;
    S->cycle_mu = mu;
  }
  QH_CHECK(hipGraphLaunch(S->cycle, st));
  return 0;
}

template <int P>
int outer_step(qcd_hip_solver_s *S, qcd_hip_field x, int x0, int count, double kappa, double mu, hipStream_t st) {
  const dim3 g2(S->nbi, (unsigned)count);
  const unsigned V = S->Vi;
  if (!S->prm.eo) {
    k_prepare_src<QCD_HIP_F64, P><<<g2, kBlock, 0, st>>>(SpinorSet<QCD_HIP_F64>::of(S->r64), S->sc,
                                                         SpinorSet<P>::of(S->Rs), S->prm.n_restart, S->ids, V);
    QH_TRY(run_cycle<P>(S, count, kappa, mu, st));
    k_accumulate<P><<<g2, kBlock, 0, st>>>(SpinorSet<QCD_HIP_F64>::of(x), x0, SpinorSet<P>::of(S->E), S->sc, S->ids,
                                           V);
    QH_CHECK(hipGetLastError());
    return 0;
  }
  // Even-odd: D e = r  <=>  D^ e_o = r_o - H_oe M^-1 r_e,  e_e = M^-1 (r_e - H_eo e_o)
  const double m = 1.0 / (2.0 * kappa) - 4.0;
  const size_t sstride = sizeof(GcrScal) / sizeof(double);
  int grid = 0;
  if constexpr (P == QCD_HIP_F64) {
    QH_TRY(eo_split(S->r64, 0, S->reh, S->roh, S->ids, count, st));
    QH_TRY(eo_minv(S->lat, S->reh, S->th, m, mu, S->ids, count, st));
    QH_TRY(eo_launch(EO_SUB, true, S->eoh, 0, 1, S->th, 0, 1, S->g64, S->roh, 0, 1, S->ids, count, m, mu, 1, S->part,
                     st, &grid));
    k_scal_rnorm<<<count, 1024, 0, st>>>(S->sc, S->part, (unsigned)grid, S->ids, S->rn);   // |r^_o| for the scaling
    k_prepare_src<QCD_HIP_F64, P><<<g2, kBlock, 0, st>>>(SpinorSet<QCD_HIP_F64>::of(S->eoh), S->sc,
                                                         SpinorSet<P>::of(S->Rs), S->prm.n_restart, S->ids, V);
    QH_CHECK(hipGetLastError());
    QH_TRY(run_cycle<P>(S, count, kappa, mu, st));
    QH_TRY(eo_scale64(S->lat, S->E, S->eoh, &S->sc[0].rnorm, sstride, S->ids, count, st));
    QH_TRY(eo_launch(EO_SUB_MINV, false, S->th, 0, 1, S->eoh, 0, 1, S->g64, S->reh, 0, 1, S->ids, count, m, mu, 0,
                     nullptr, st, nullptr));
    QH_TRY(eo_accumulate(x, x0, S->th, S->eoh, S->ids, count, st));
    return 0;
  }
  // Mixed precision: only the true residual, the accumulation into x and the
  // reductions are FP64.  The halves of the residual (reh, roh and th =
  // M^-1 reh; FP16: written by the true residual itself), the inner source
  // r^_o and the even correction e_e are in the inner precision, with the
  // inner-precision links: their rounding only perturbs the correction of this
  // outer step, which the next FP64 true residual measures.
  if (!S->split_res) QH_TRY(eo_split_minv(S->r64, 0, S->reh, S->roh, S->th, m, mu, S->ids, count, st));
  QH_TRY(eo_launch(EO_SUB, true, S->eoh, 0, 1, S->th, 0, 1, S->gin, S->roh, 0, 1, S->ids, count, m, mu, 1, S->part,
                   st, &grid));
  k_scal_rnorm<<<count, 1024, 0, st>>>(S->sc, S->part, (unsigned)grid, S->ids, S->rn);   // |r^_o| for the scaling
  k_prepare_src<P, P><<<g2, kBlock, 0, st>>>(SpinorSet<P>::of(S->eoh), S->sc, SpinorSet<P>::of(S->Rs),
                                             S->prm.n_restart, S->ids, V);
  QH_CHECK(hipGetLastError());
  QH_TRY(run_cycle<P>(S, count, kappa, mu, st));
  QH_TRY(eo_launch(EO_SUB_MINV, false, S->th, 0, 1, S->E, 0, 1, S->gin, S->reh, 0, 1, S->ids, count, m, mu, 0,
                   nullptr, st, nullptr));
  QH_TRY(eo_accumulate_p(x, x0, S->th, S->E, S->ids, count, st));
  return 0;
}

// r_b = bq_b - A x_{x0+b} for the active slots; residual norms to rn_host
int true_residual(qcd_hip_solver_s *S, qcd_hip_field x, int x0, int count, double kappa, double mu, hipStream_t st) {
  const int epi = S->use_q ? EPI_RES_Q : EPI_RES_D;
  const double m = 1.0 / (2.0 * kappa) - 4.0;
  int grid = 0;
  if (S->split_res)
    QH_TRY(dslash_residual_eo_split(x, x0, 1, S->g64, S->bq, 0, 1, S->ids, count, m, mu, S->reh, S->roh, S->th,
                                    S->part, st, &grid));
  else
    QH_TRY(dslash_launch_batch(epi, true, S->r64, 0, 1, x, x0, 1, S->g64, S->bq, 0, 1, S->ids, count, m, mu, S->part,
                               st, &grid));
  k_scal_rnorm<<<count, 1024, 0, st>>>(S->sc, S->part, (unsigned)grid, S->ids, S->rn);
  QH_CHECK(hipGetLastError());
  QH_CHECK(hipMemcpyAsync(S->rn_host, S->rn, sizeof(double) * S->B, hipMemcpyDeviceToHost, st));
  QH_CHECK(hipStreamSynchronize(st));
  return 0;
}

int upload_ids(qcd_hip_solver_s *S, const std::vector<int> &ids, hipStream_t st) {
  for (size_t i = 0; i < ids.size(); i++) S->ids_host[i] = ids[i];
  QH_CHECK(hipMemcpyAsync(S->ids, S->ids_host, sizeof(int) * ids.size(), hipMemcpyHostToDevice, st));
  QH_CHECK(hipStreamSynchronize(st));   // the pinned mirror is reused next step
  return 0;
}

// Solve rhs [x0, x0 + cnt) of b into x.
int solve_batch(qcd_hip_solver_s *S, qcd_hip_field x, qcd_hip_field b, int x0, int cnt, double kappa, double mu,
                hipStream_t st, qcd_hip_solver_stats *tot, int *worst) {
  const unsigned V = (unsigned)S->lat->V;
  k_start<<<dim3(S->nb, (unsigned)cnt), kBlock, 0, st>>>(SpinorSet<QCD_HIP_F64>::of(b), SpinorSet<QCD_HIP_F64>::of(x),
                                                         x0, SpinorSet<QCD_HIP_F64>::of(S->bq), S->use_q, V);
  QH_CHECK(hipGetLastError());
  QH_CHECK(hipMemsetAsync(S->sc, 0, sizeof(GcrScal) * S->B, st));
  std::vector<int> act(cnt);
  for (int i = 0; i < cnt; i++) act[i] = i;
  QH_TRY(upload_ids(S, act, st));
  QH_TRY(true_residual(S, x, x0, cnt, kappa, mu, st));
  std::vector<double> bnorm(cnt), rel(cnt);
  std::vector<int> outer(cnt, 0);
  std::vector<GcrScal> host_sc;
  for (int i = 0; i < cnt; i++) {
    bnorm[i] = S->rn_host[i];
    rel[i] = bnorm[i] > 0.0 ? 1.0 : 0.0;
  }
  int bad = 0;
  std::vector<int> uploaded = act;   // the list the device holds
  for (int it = 0; it < S->prm.max_outer; it++) {
    act.clear();
    for (int i = 0; i < cnt; i++)
      if (rel[i] > S->prm.tol) act.push_back(i);
    if (act.empty()) break;
    if (act != uploaded) {   // most outer steps keep every rhs active: no upload, no host round trip
      QH_TRY(upload_ids(S, act, st));
      uploaded = act;
    }
    const int na = (int)act.size();
    switch (S->prm.inner_prec) {
      case QCD_HIP_F64: QH_TRY(outer_step<QCD_HIP_F64>(S, x, x0, na, kappa, mu, st)); break;
      case QCD_HIP_F32: QH_TRY(outer_step<QCD_HIP_F32>(S, x, x0, na, kappa, mu, st)); break;
      default: QH_TRY(outer_step<QCD_HIP_F16>(S, x, x0, na, kappa, mu, st)); break;
    }
    QH_TRY(true_residual(S, x, x0, na, kappa, mu, st));
    for (int i : act) {
      outer[i]++;
      rel[i] = S->rn_host[i] / bnorm[i];
      if (!(rel[i] == rel[i])) bad = 1;
    }
    if (bad) break;
  }
  for (int i = 0; i < cnt; i++) {
    tot->outer += outer[i];
    tot->inner_iters += outer[i] * S->prm.n_restart;
    tot->relres = std::max(tot->relres, rel[i]);
    if (!bad && rel[i] > S->prm.tol) *worst = std::max(*worst, 6);
  }
  if (bad) *worst = 5;
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
  p->batch = 1;
  p->eo = 0;
  p->graph = 1;
}

}  // extern "C"

namespace qcdhip {
// Device bytes qcd_hip_solver_create and qcd_hip_solver_set_gauge allocate
// for a lattice of extents L with parameters prm: the same field shapes and
// reduction rows, without allocating anything.
size_t solver_device_bytes(const int L[4], const qcd_hip_solver_params *prm) {
  qcd_hip_lattice_s lat{};
  for (int d = 0; d < 4; d++) lat.L[d] = L[d];
  lat.V = (long)L[0] * L[1] * L[2] * L[3];
  lat.stride = round_up(lat.V, kTile);
  auto field = [](long stride, int prec, int nrhs) {
    return (size_t)12 * nrhs * stride * spinor_comp_bytes(prec) + (prec == QCD_HIP_F16 ? (size_t)4 * nrhs * stride : 0);
  };
  const int n = prm->n_restart, B = prm->batch, ip = prm->inner_prec;
  const long V = lat.V;
  const long Vi = prm->eo ? V / 2 : V;
  const long sti = round_up(Vi, kTile);
  size_t b = 0;
  b += 2 * field(sti, ip, n * B) + field(sti, ip, B);                  // Rs, Ws, E
  if (prm->eo) b += 5 * field(sti, ip, B);                             // te, reh, roh, th, eoh
  const bool split = prm->eo && ip == QCD_HIP_F16;                     // no r64 (split residual)
  b += (split ? 1 : 2) * field(lat.stride, QCD_HIP_F64, B);            // r64, bq
  b += sizeof(GcrScal) * B + sizeof(double) * B + sizeof(int) * B;      // sc, rn, ids
  const size_t nb = (size_t)grid_for(V);
  const size_t rowmax = ((size_t)2 * (n + kDotChunk) + 4) * (nb + 64);
  b += sizeof(double) * std::max(rowmax, (size_t)dslash_grid_bound(&lat)) * B;   // partials
  if (ip != QCD_HIP_F64) b += (size_t)36 * lat.stride * spinor_comp_bytes(ip);   // inner-precision links
  if (ip == QCD_HIP_F16) b += (size_t)8 * 36 * lat.V;                            // matrix-core hop links
  return b;
}
}  // namespace qcdhip

extern "C" {

size_t qcd_hip_solver_bytes(const int L[4], const qcd_hip_solver_params *prm) {
  if (!L || !prm || prm->n_restart < 1 || prm->batch < 1 || L[0] < 1 || L[1] < 1 || L[2] < 1 || L[3] < 1) return 0;
  return solver_device_bytes(L, prm);
}

int qcd_hip_solver_create(qcd_hip_solver *out, qcd_hip_lattice lat, const qcd_hip_solver_params *prm) {
  *out = nullptr;
  QH_TRY(ensure_init());
  if (!lat || !prm || prm->n_restart < 1 || prm->n_restart > kN || prm->tol <= 0.0 || prm->max_outer < 1 ||
      prm->inner_prec < QCD_HIP_F64 || prm->inner_prec > QCD_HIP_F16 || prm->batch < 1 || prm->batch > 1024 ||
      (prm->op != QCD_HIP_OP_Q && prm->op != QCD_HIP_OP_D) || (prm->eo != 0 && prm->eo != 1) ||
      (prm->graph != 0 && prm->graph != 1)) {
    set_error("solver_create: bad parameters (n_restart 1..%d, tol > 0, max_outer >= 1, batch 1..1024, eo 0/1, "
              "graph 0/1)", kN);
    return 1;
  }
  if (prm->eo && !eo_supported(lat)) {
    set_error("solver_create: even-odd preconditioning needs every lattice extent even");
    return 1;
  }
  auto *S = new qcd_hip_solver_s();
  S->lat = lat;
  S->prm = *prm;
  if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess) {
    delete S;
    set_error("solver_create: cannot create a stream");
    return 2;
  }
  S->B = prm->batch;
  S->nb = (unsigned)grid_for(lat->V);
  S->use_q = prm->op == QCD_HIP_OP_Q && !prm->eo;   // the EO inner system is D's Schur complement
  qcd_hip_lattice ilat = lat;
  if (prm->eo) ilat = S->half = eo_half_lattice(lat);
  S->Vi = (unsigned)ilat->V;
  S->nbi = (unsigned)grid_for(ilat->V);
  const int n = prm->n_restart, B = S->B;
  int rc = qcd_hip_spinor_create(&S->Rs, ilat, prm->inner_prec, n * B);
  if (!rc) rc = qcd_hip_spinor_create(&S->Ws, ilat, prm->inner_prec, n * B);
  if (!rc) rc = qcd_hip_spinor_create(&S->E, ilat, prm->inner_prec, B);
  if (!rc && prm->eo) rc = qcd_hip_spinor_create(&S->te, ilat, prm->inner_prec, B);
  for (qcd_hip_field *f : {&S->reh, &S->roh, &S->th, &S->eoh})
    if (!rc && prm->eo) rc = qcd_hip_spinor_create(f, ilat, prm->inner_prec, B);
  S->split_res = prm->eo && prm->inner_prec == QCD_HIP_F16;
  if (!rc && !S->split_res) rc = qcd_hip_spinor_create(&S->r64, lat, QCD_HIP_F64, B);
  if (!rc) rc = qcd_hip_spinor_create(&S->bq, lat, QCD_HIP_F64, B);
  if (!rc) rc = dev_alloc((void **)&S->sc, sizeof(GcrScal) * B);
  if (!rc) rc = dev_alloc((void **)&S->rn, sizeof(double) * B);
  if (!rc) rc = dev_alloc((void **)&S->ids, sizeof(int) * B);
  if (!rc && hipHostMalloc((void **)&S->rn_host, sizeof(double) * B, hipHostMallocDefault) != hipSuccess) rc = 4;
  if (!rc && hipHostMalloc((void **)&S->ids_host, sizeof(int) * B, hipHostMallocDefault) != hipSuccess) rc = 4;
  // per batch row: 2(n-1) multi-dot sums per site workgroup, or one sum per
  // Dslash workgroup (a z-march grid never exceeds site workgroups + T*Z)
  // (the multi-dot writes whole chunks: up to 2 (n + kDotChunk) sums)
  const size_t rowmax = ((size_t)2 * (n + kDotChunk) + 4) * ((size_t)S->nb + 64);
  const size_t dsl = (size_t)dslash_grid_bound(lat);
  S->rowsz = std::max(rowmax, dsl);
  S->part_bytes = sizeof(double) * S->rowsz * B;
  if (!rc) rc = dev_alloc((void **)&S->part, S->part_bytes);
  if (rc) {
    if (rc == 4) set_error("solver_create: out of memory for %d rhs x %d Krylov vectors", B, n);
    qcd_hip_solver_destroy(S);
    return rc;
  }
  *out = S;
  return 0;
}

void qcd_hip_solver_destroy(qcd_hip_solver S) {
  if (!S) return;
  S->drop_cycle();
  qcd_hip_field_destroy(S->Rs);
  qcd_hip_field_destroy(S->Ws);
  qcd_hip_field_destroy(S->E);
  qcd_hip_field_destroy(S->r64);
  qcd_hip_field_destroy(S->bq);
  for (qcd_hip_field f : {S->te, S->reh, S->roh, S->th, S->eoh}) qcd_hip_field_destroy(f);
  if (S->gin && S->gin != S->g64) qcd_hip_field_destroy(S->gin);
  if (S->stream) {
    (void)hipStreamSynchronize(S->stream);
    (void)hipStreamDestroy(S->stream);
  }
  dev_free(S->sc, sizeof(GcrScal) * S->B);
  dev_free(S->rn, sizeof(double) * S->B);
  dev_free(S->ids, sizeof(int) * S->B);
  dev_free(S->part, S->part_bytes);
  if (S->rn_host) (void)hipHostFree(S->rn_host);
  if (S->ids_host) (void)hipHostFree(S->ids_host);
  delete S->half;
  delete S;
}

int qcd_hip_solver_set_gauge(qcd_hip_solver S, qcd_hip_field g64) {
  QH_TRY(ensure_init());
  if (!S || !g64 || g64->kind != QCD_HIP_GAUGE || g64->prec != QCD_HIP_F64 || g64->lat->V != S->lat->V) {
    set_error("solver_set_gauge: need an FP64 gauge field on the solver's lattice");
    return 1;
  }
  S->g64 = g64;
  S->drop_cycle();   // the captured cycle holds the previous field's pointers
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
  if (S->prm.inner_prec == QCD_HIP_F16) {
    // the matrix-core hops read phased links prepared once per gauge field
    QH_TRY(mma_prepare(S->gin, S->lat, ctx()->stream));
    QH_CHECK(hipStreamSynchronize(ctx()->stream));
  }
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
  if (x == b) {
    set_error("solver_solve: x must not alias b");
    return 1;
  }
  if (kappa <= 0.0) {
    set_error("solver_solve: kappa must be positive");
    return 1;
  }
  hipStream_t st = S->stream;
  const auto t0 = std::chrono::steady_clock::now();
  // Tune the inner and outer operators before the first timed use.
  if (!S->prm.eo) QH_TRY(dslash_prepare(S->Ws, S->Rs, S->gin));
  // (scratch output of the tuning launches: r64, or bq, which each batch
  // fills afterwards)
  QH_TRY(dslash_prepare(S->r64 ? S->r64 : S->bq, x, S->g64));
  qcd_hip_solver_stats tot;
  std::memset(&tot, 0, sizeof(tot));
  int worst = 0;
  for (int r0 = 0; r0 < b->nrhs && worst != 5; r0 += S->B) {
    const int cnt = std::min(S->B, b->nrhs - r0);
    QH_TRY(solve_batch(S, x, b, r0, cnt, kappa, mu, st, &tot, &worst));
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
  std::lock_guard<std::recursive_mutex> lock(state_mutex());   // shared scratch field and stream
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

}  // extern "C"

// ------------------------------------------------------------ allocation-free residual
// qcd_hip_b_minus_Dx_noalloc keeps its workgroup partial sums and the two
// finished sums in this array instead of library allocations: at most
// kNoAllocParts Dslash workgroups (every launch configuration of a 2^24-site
// lattice fits), kMaxReduceBlocks norm partials and 2 results, 2.0 MiB per
// GPU, held until the process ends.  In host code `scratch` only names the
// symbol; kernels get its device address from noalloc_scratch().
namespace {
constexpr size_t kNoAllocParts = (size_t)1 << 18;
constexpr size_t kNoAllocNorm = kNoAllocParts;                      // offset of the norm partials
constexpr size_t kNoAllocOut = kNoAllocParts + kMaxReduceBlocks;    // offset of the two sums
constexpr size_t kNoAllocDoubles = kNoAllocOut + 2;
__device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices];

// device address of `scratch` on the current device (looked up once per device)
int noalloc_scratch(double **out) {
  int dev = 0;
  QH_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("b_minus_Dx_noalloc: device %d out of range", dev);
    return 1;
  }
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    QH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" {

int qcd_hip_b_minus_Dx_noalloc(qcd_hip_field x, int rx, qcd_hip_field b, int rb, qcd_hip_field g, double kappa,
                               double mu, qcd_hip_field r, int rr, double *rel) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());   // the library stream
  if (!ctx()->stream) {
    set_error("b_minus_Dx_noalloc: the library is not initialised (create the fields first)");
    return 1;
  }
  if (!x || !b || !g || !r || x->prec != QCD_HIP_F64 || b->prec != QCD_HIP_F64 || g->prec != QCD_HIP_F64 ||
      r->prec != QCD_HIP_F64 || r->kind != QCD_HIP_SPINOR || !rel || rx < 0 || rx >= x->nrhs || rb < 0 ||
      rb >= b->nrhs || rr < 0 || rr >= r->nrhs) {
    set_error("b_minus_Dx_noalloc: needs FP64 x, b, r and gauge field and valid rhs indices");
    return 1;
  }
  QH_TRY(dslash_check(r, x, g));
  QH_TRY(dslash_check(x, b, g));
  const int grid = dslash_grid(x, g);
  if (grid < 1 || (size_t)grid > kNoAllocParts) {
    set_error("b_minus_Dx_noalloc: %d Dslash workgroups exceed the %zu partial sums of the static scratch", grid,
              kNoAllocParts);
    return 1;
  }
  double *sp = nullptr;
  QH_TRY(noalloc_scratch(&sp));
  hipStream_t st = ctx()->stream;
  int used = 0, nb = 0;
  // r = b - D x with |r|^2 workgroup partials (every workgroup writes its slot)
  QH_TRY(dslash_launch(EPI_RES_D, true, r, rr, x, rx, g, b, rb, 1.0 / (2.0 * kappa) - 4.0, mu, sp, st, &used));
  if (used != grid) {
    set_error("b_minus_Dx_noalloc: launch used %d workgroups, expected %d", used, grid);
    return 1;
  }
  QH_TRY(norm2_launch(b, rb, sp + kNoAllocNorm, &nb, st));
  QH_TRY(finalize(sp, used, 1, sp + kNoAllocOut, st));
  QH_TRY(finalize(sp + kNoAllocNorm, nb, 1, sp + kNoAllocOut + 1, st));
  double h[2] = {0.0, 0.0};
  QH_CHECK(hipMemcpyAsync(h, sp + kNoAllocOut, sizeof(h), hipMemcpyDeviceToHost, st));
  QH_CHECK(hipStreamSynchronize(st));
  *rel = std::sqrt(h[0]) / std::sqrt(h[1]);
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
