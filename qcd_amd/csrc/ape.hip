// ape.hip -- 3D APE smearing of the gauge field with SU(3) projection, stout
// smearing (3d or 4d), and the plaquette.
//
// One step (lib/src/qcd_smearing.c:205-273), for the spatial directions
// mu = 1, 2, 3 (temporal links are copied unchanged):
//   U'_mu(x) = P_SU3[ U_mu(x) + alpha sum_{nu = 1..3, nu != mu}
//                 ( U_nu(x) U_mu(x+nu) U_nu(x+mu)^+ + U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu) ) ]
// P_SU3(M) is the unitary polar factor of M, M (M^+ M)^-1/2, times
// exp(-i arg(det M) / 3) (lib/src/qcd_blas.c:940-1364, which diagonalises
// M^+ M by Cardano's formula).  The polar factor is unique for non-singular
// M; it is computed here by Newton's iteration X <- (g X + X^-+ / g) / 2 with
// Frobenius-norm scaling g, which converges quadratically from X = M.
// One lane per (site, direction); FP64 throughout.  The gauge field is read
// in the device layout (one array per direction of [tile][9][64] links).
#include "blas_common.h"

namespace qcdhip {

using C = Cx<double>;

__device__ __forceinline__ C cm(C a, C b) { return C{fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)}; }
__device__ __forceinline__ C cj(C a) { return C{a.re, -a.im}; }

// c = a b (DA: a^+, DB: b^+)
template <bool DA, bool DB>
__device__ __forceinline__ void mm3(C c[9], const C a[9], const C b[9]) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C s{0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const C x = DA ? cj(a[3 * k + i]) : a[3 * i + k];
        const C y = DB ? cj(b[3 * j + k]) : b[3 * k + j];
        const C p = cm(x, y);
        s = C{s.re + p.re, s.im + p.im};
      }
      c[3 * i + j] = s;
    }
}

__device__ __forceinline__ C det3(const C a[9]) {
  C d{0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 3; c++) {   // expansion along row 0, cyclic minors carry their sign
    const C t = cm(a[3 + (c + 1) % 3], a[6 + (c + 2) % 3]);
    const C s = cm(a[3 + (c + 2) % 3], a[6 + (c + 1) % 3]);
    d = C{d.re + a[c].re * (t.re - s.re) - a[c].im * (t.im - s.im), d.im + a[c].re * (t.im - s.im) + a[c].im * (t.re - s.re)};
  }
  return d;
}

// SU(3) projection in place
__device__ void project_su3(C m[9]) {
  C x[9];
#pragma unroll
  for (int k = 0; k < 9; k++) x[k] = m[k];
  for (int it = 0; it < 100; it++) {
    // y = (x^-1)^+ = conj(cofactor(x)) / conj(det x)
    const C d = det3(x);
    const double rdd = 1.0 / (d.re * d.re + d.im * d.im);
    C y[9];
    double nx = 0.0, ny = 0.0;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        const C t = cm(x[3 * r1 + c1], x[3 * r2 + c2]);
        const C s = cm(x[3 * r1 + c2], x[3 * r2 + c1]);
        const C cof{t.re - s.re, t.im - s.im};
        // cof / det, then conjugate
        const C q{(cof.re * d.re + cof.im * d.im) * rdd, (cof.im * d.re - cof.re * d.im) * rdd};
        y[3 * r + c] = cj(q);
        nx += x[3 * r + c].re * x[3 * r + c].re + x[3 * r + c].im * x[3 * r + c].im;
        ny += q.re * q.re + q.im * q.im;
      }
    const double g = it < 8 ? sqrt(sqrt(ny / nx)) : 1.0, rg = 1.0 / g;
    double diff = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const C n{0.5 * (g * x[k].re + y[k].re * rg), 0.5 * (g * x[k].im + y[k].im * rg)};
      diff += (n.re - x[k].re) * (n.re - x[k].re) + (n.im - x[k].im) * (n.im - x[k].im);
      x[k] = n;
    }
    if (diff <= 1e-30) break;   // a fixed point of the scaled step is unitary
  }
  const C d = det3(m);
  const double ph = atan2(d.im, d.re) / 3.0;
  const C e{cos(ph), -sin(ph)};
#pragma unroll
  for (int k = 0; k < 9; k++) m[k] = cm(x[k], e);
}

// Plaquette sum of one site over the 6 planes mu < nu (qcd_observables.c:15-36):
// Re tr[U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+], one workgroup partial each.
__global__ void __launch_bounds__(kBlock) k_plaquette(DslashArgs a, GaugeIn<QCD_HIP_F64> U, double *partials) {
  const unsigned s = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  double sum = 0.0;
  if (s < a.V) {
    const Site st = site_coords(s, a);
    unsigned fwd[4];
    fwd[0] = neighbour<0, +1>(st, a);
    fwd[1] = neighbour<1, +1>(st, a);
    fwd[2] = neighbour<2, +1>(st, a);
    fwd[3] = neighbour<3, +1>(st, a);
    for (int mu = 0; mu < 3; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        C A[9], B[9], T[9], P[9];
        U.load(mu, s, A);
        U.load(nu, fwd[mu], B);
        mm3<false, false>(T, A, B);
        U.load(mu, fwd[nu], A);
        mm3<false, true>(P, T, A);
        U.load(nu, s, A);
        // Re tr(P A^+) = sum_ij Re(P_ij conj(A_ij))
#pragma unroll
        for (int k = 0; k < 9; k++) sum += P[k].re * A[k].re + P[k].im * A[k].im;
      }
  }
  block_reduce_store(sum, partials, blockIdx.x);
}

// ------------------------------------------------------------ stout
// sin(w)/w, with the reference's series below |w| = 0.05 (qcd_smearing.c:279-290)
__device__ __forceinline__ double stout_xi0(double w) {
  if (fabs(w) < 0.05) {
    const double t = 1.0 - w * w / 42.0;
    const double t1 = 1.0 - w * w / 20.0 * t;
    return 1.0 - w * w / 6.0 * t1;
  }
  return sin(w) / w;
}

// m <- exp(i m) for Hermitian traceless m: f0 + f1 Q + f2 Q^2 with the
// Cayley-Hamilton coefficients of Peardon & Morningstar (hep-lat/0311018
// eqs. 19, 29-34; qcd_smearing.c:293-370).  Where the reference divides 0 by
// 0 (tr Q^2 / 2 below 1e-24, e.g. the unit gauge field) this uses
// 1 + iQ - Q^2/2, exact to double precision there; where rounding puts the
// argument of acos above 1 (two equal eigenvalues) it is clamped to 1.  The
// reference returns NaN in both cases.
__device__ void stout_exp(C m[9]) {
  C m2[9], m3[9];
  mm3<false, false>(m2, m, m);
  mm3<false, false>(m3, m2, m);
  double c0 = (m3[0].re + m3[4].re + m3[8].re) / 3.0;
  const double c1 = (m2[0].re + m2[4].re + m2[8].re) / 2.0;
  if (c1 < 1e-24) {
#pragma unroll
    for (int k = 0; k < 9; k++)
      m[k] = C{(k % 4 == 0 ? 1.0 : 0.0) - m[k].im - 0.5 * m2[k].re, m[k].re - 0.5 * m2[k].im};
    return;
  }
  const bool neg = c0 < 0.0;
  if (neg) c0 = -c0;
  const double c0max = 2.0 * pow(c1 / 3.0, 1.5);
  const double theta = acos(c0 < c0max ? c0 / c0max : 1.0);   // degenerate spectra: rounding can exceed 1
  const double u = sqrt(c1 / 3.0) * cos(theta / 3.0);
  const double w = sqrt(c1) * sin(theta / 3.0);
  const double xi = stout_xi0(w), cw = cos(w);
  const C e2iu{cos(2 * u), sin(2 * u)}, emiu{cos(u), -sin(u)};
  const C h0 = [&] {
    const C a = cm(C{u * u - w * w, 0.0}, e2iu), b = cm(emiu, C{8 * u * u * cw, 2 * u * (3 * u * u + w * w) * xi});
    return C{a.re + b.re, a.im + b.im};
  }();
  const C h1 = [&] {
    const C a = cm(C{2 * u, 0.0}, e2iu), b = cm(emiu, C{2 * u * cw, -(3 * u * u - w * w) * xi});
    return C{a.re - b.re, a.im - b.im};
  }();
  const C h2 = [&] {
    const C b = cm(emiu, C{cw, 3 * u * xi});
    return C{e2iu.re - b.re, e2iu.im - b.im};
  }();
  const double fac = 9 * u * u - w * w;
  // eq. (34): f_j(-c0) = (-1)^j conj(f_j(c0))
  const C f0 = neg ? C{h0.re / fac, -h0.im / fac} : C{h0.re / fac, h0.im / fac};
  const C f1 = neg ? C{-h1.re / fac, h1.im / fac} : C{h1.re / fac, h1.im / fac};
  const C f2 = neg ? C{h2.re / fac, -h2.im / fac} : C{h2.re / fac, h2.im / fac};
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const C a = cm(f1, m[k]), b = cm(f2, m2[k]);
    C r{a.re + b.re, a.im + b.im};
    if (k % 4 == 0) r = C{r.re + f0.re, r.im + f0.im};
    m[k] = r;
  }
}

template <int MU>
__device__ __forceinline__ unsigned fwd_of(const Site &st, const DslashArgs &a) {
  return neighbour<MU, +1>(st, a);
}
__device__ __forceinline__ unsigned fwd_dir(unsigned mu, const Site &st, const DslashArgs &a) {
  return mu == 0 ? fwd_of<0>(st, a) : mu == 1 ? fwd_of<1>(st, a) : mu == 2 ? fwd_of<2>(st, a) : fwd_of<3>(st, a);
}
__device__ __forceinline__ unsigned bwd_dir(unsigned mu, const Site &st, const DslashArgs &a) {
  return mu == 0 ? neighbour<0, -1>(st, a) : mu == 1 ? neighbour<1, -1>(st, a)
                                           : mu == 2 ? neighbour<2, -1>(st, a) : neighbour<3, -1>(st, a);
}

// c += a b (DA: a^+, DB: b^+); each element's product is summed from zero
// and then added, the rounding of computing a b first
template <bool DA, bool DB>
__device__ __forceinline__ void mm3_add(C c[9], const C a[9], const C b[9]) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C s{0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const C x = DA ? cj(a[3 * k + i]) : a[3 * i + k];
        const C y = DB ? cj(b[3 * j + k]) : b[3 * k + j];
        const C p = cm(x, y);
        s = C{s.re + p.re, s.im + p.im};
      }
      c[3 * i + j] = C{c[3 * i + j].re + s.re, c[3 * i + j].im + s.im};
    }
}

// One APE step for the spatial directions, one lane per (site, direction)
// with the direction mu = 1 + blockIdx.y uniform over the workgroup, so the
// staple loops do not diverge.  Each staple's three links are loaded as its
// products need them, which bounds the live matrices to the staple sum and
// two factors.  The temporal links are copied by the caller.
__global__ void __launch_bounds__(kBlock) k_ape3d(DslashArgs a, GaugeIn<QCD_HIP_F64> U, double2 *__restrict__ out,
                                                  size_t stride, double alpha) {
  const unsigned s = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (s >= a.V) return;
  const unsigned mu = 1 + blockIdx.y;
  const Site st = site_coords(s, a);
  C S[9];
#pragma unroll
  for (int k = 0; k < 9; k++) S[k] = C{0.0, 0.0};
  // forward staples U_nu(x) U_mu(x+nu) U_nu(x+mu)^+, then backward ones
  // U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu) (the reference's order)
  const unsigned sf = fwd_dir(mu, st, a);
#pragma unroll 1
  for (unsigned nu = 1; nu < 4; nu++) {
    if (nu == mu) continue;
    C T[9];
    {
      C A[9], B[9];
      U.load(nu, s, A);
      U.load(mu, fwd_dir(nu, st, a), B);
      mm3<false, false>(T, A, B);
    }
    C W[9];
    U.load(nu, sf, W);
    mm3_add<false, true>(S, T, W);
  }
#pragma unroll 1
  for (unsigned nu = 1; nu < 4; nu++) {
    if (nu == mu) continue;
    const unsigned sn = bwd_dir(nu, st, a);
    C T[9];
    {
      C B[9], W[9];
      U.load(mu, sn, B);
      U.load(nu, fwd_dir(mu, site_coords(sn, a), a), W);
      mm3<false, false>(T, B, W);
    }
    C A[9];
    U.load(nu, sn, A);
    mm3_add<true, false>(S, A, T);
  }
  C m[9];
  U.load(mu, s, m);
#pragma unroll
  for (int k = 0; k < 9; k++) m[k] = C{m[k].re + alpha * S[k].re, m[k].im + alpha * S[k].im};
  project_su3(m);
  double2 *o = out + (size_t)mu * 9 * stride + (size_t)(s / kTile) * 9 * kTile + s % kTile;
#pragma unroll
  for (int k = 0; k < 9; k++) o[(size_t)k * kTile] = make_double2(m[k].re, m[k].im);
}

// One stout step (qcd_smearing.c:374-476), one lane per (site, direction):
// C = sum over nu != mu (nu from START) of the backward staple then the
// forward one, W = C U^+, A = W - W^+ minus tr(A)/3, Q = -i rho/2 A,
// U' = exp(iQ) U.  The direction is START + blockIdx.y, uniform over the
// workgroup; directions below START (the temporal one for 3d stout) are
// copied by the caller.
template <unsigned START>
__global__ void __launch_bounds__(kBlock) k_stout(DslashArgs a, GaugeIn<QCD_HIP_F64> U, double2 *__restrict__ out,
                                                  size_t stride, double rho) {
  const unsigned s = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (s >= a.V) return;
  const unsigned mu = START + blockIdx.y;   // uniform over the workgroup
  const Site st = site_coords(s, a);
  C S[9];
#pragma unroll
  for (int k = 0; k < 9; k++) S[k] = C{0.0, 0.0};
  const unsigned sf = fwd_dir(mu, st, a);
#pragma unroll 1
  for (unsigned nu = START; nu < 4; nu++) {
    if (nu == mu) continue;
    // backward: U_nu(x-nu)^+ [U_mu(x-nu) U_nu(x-nu+mu)]
    const unsigned sb = bwd_dir(nu, st, a);
    {
      C T[9];
      {
        C B[9], W[9];
        U.load(mu, sb, B);
        U.load(nu, fwd_dir(mu, site_coords(sb, a), a), W);
        mm3<false, false>(T, B, W);
      }
      C A[9];
      U.load(nu, sb, A);
      mm3_add<true, false>(S, A, T);
    }
    // forward: [U_nu(x) U_mu(x+nu)] U_nu(x+mu)^+
    {
      C T[9];
      {
        C A[9], B[9];
        U.load(nu, s, A);
        U.load(mu, fwd_dir(nu, st, a), B);
        mm3<false, false>(T, A, B);
      }
      C W[9];
      U.load(nu, sf, W);
      mm3_add<false, true>(S, T, W);
    }
  }
  C m[9], W[9];
  U.load(mu, s, m);
  mm3<false, true>(W, S, m);   // W = C U^+
  C q[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) q[3 * r + c] = C{W[3 * r + c].re - W[3 * c + r].re, W[3 * r + c].im + W[3 * c + r].im};
  const C tr{(q[0].re + q[4].re + q[8].re) / 3.0, (q[0].im + q[4].im + q[8].im) / 3.0};
#pragma unroll
  for (int r = 0; r < 3; r++) q[4 * r] = C{q[4 * r].re - tr.re, q[4 * r].im - tr.im};
#pragma unroll
  for (int k = 0; k < 9; k++) q[k] = C{0.5 * rho * q[k].im, -0.5 * rho * q[k].re};   // Q = -i rho/2 A
  stout_exp(q);
  mm3<false, false>(W, q, m);
  double2 *o = out + (size_t)mu * 9 * stride + (size_t)(s / kTile) * 9 * kTile + s % kTile;
#pragma unroll
  for (int k = 0; k < 9; k++) o[(size_t)k * kTile] = make_double2(W[k].re, W[k].im);
}

int stout_step(qcd_hip_field out, qcd_hip_field in, double rho, int ndim, hipStream_t st) {
  const DslashArgs a = make_args(in->lat, 0.0, 0.0);
  const int grid = grid_for(in->lat->V);
  if (ndim == 3) {
    // temporal links copied: direction 0 is the first 9 stride links of the layout
    QH_CHECK(hipMemcpyAsync(out->data, in->data, (size_t)9 * in->lat->stride * sizeof(double2),
                            hipMemcpyDeviceToDevice, st));
    k_stout<1><<<dim3(grid, 3), kBlock, 0, st>>>(a, GaugeIn<QCD_HIP_F64>::of(in), (double2 *)out->data,
                                                 (size_t)in->lat->stride, rho);
  } else {
    k_stout<0><<<dim3(grid, 4), kBlock, 0, st>>>(a, GaugeIn<QCD_HIP_F64>::of(in), (double2 *)out->data,
                                                 (size_t)in->lat->stride, rho);
  }
  QH_CHECK(hipGetLastError());
  return 0;
}

int ape_step(qcd_hip_field out, qcd_hip_field in, double alpha, hipStream_t st) {
  const DslashArgs a = make_args(in->lat, 0.0, 0.0);
  // temporal links: direction 0 is the first 9 stride links of the layout
  QH_CHECK(hipMemcpyAsync(out->data, in->data, (size_t)9 * in->lat->stride * sizeof(double2),
                          hipMemcpyDeviceToDevice, st));
  k_ape3d<<<dim3(grid_for(in->lat->V), 3), kBlock, 0, st>>>(a, GaugeIn<QCD_HIP_F64>::of(in), (double2 *)out->data,
                                                           (size_t)in->lat->stride, alpha);
  QH_CHECK(hipGetLastError());
  return 0;
}

// every link (all four directions) projected onto SU(3) in place
// (lib/src/qcd_blas.c:541-939): one lane per (site, direction)
__global__ void __launch_bounds__(kBlock) k_project_links(GaugeIn<QCD_HIP_F64> U, double2 *__restrict__ data,
                                                          unsigned V, size_t stride) {
  const unsigned s = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (s >= V) return;
  const unsigned mu = blockIdx.y;
  C m[9];
  U.load(mu, s, m);
  project_su3(m);
  double2 *o = data + (size_t)mu * 9 * stride + (size_t)(s / kTile) * 9 * kTile + s % kTile;
#pragma unroll
  for (int k = 0; k < 9; k++) o[(size_t)k * kTile] = make_double2(m[k].re, m[k].im);
}
}  // namespace qcdhip

using namespace qcdhip;

extern "C" {

int qcd_hip_project_su3(qcd_hip_field g) {
  QH_TRY(ensure_init());
  if (!g || g->kind != QCD_HIP_GAUGE || g->prec != QCD_HIP_F64) {
    set_error("project_su3: expected an FP64 gauge field");
    return 1;
  }
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  hipStream_t st = ctx()->stream;
  gauge_written(g);
  k_project_links<<<dim3(grid_for(g->lat->V), 4), kBlock, 0, st>>>(GaugeIn<QCD_HIP_F64>::of(g), (double2 *)g->data,
                                                                   (unsigned)g->lat->V, (size_t)g->lat->stride);
  QH_CHECK(hipGetLastError());
  QH_CHECK(hipStreamSynchronize(st));
  return 0;
}

int qcd_hip_ape_smear(qcd_hip_field out, qcd_hip_field in, double alpha, int nsteps) {
  QH_TRY(ensure_init());
  if (!out || !in || out == in || out->kind != QCD_HIP_GAUGE || in->kind != QCD_HIP_GAUGE ||
      out->prec != QCD_HIP_F64 || in->prec != QCD_HIP_F64 || out->lat->V != in->lat->V || nsteps < 0) {
    set_error("ape_smear: expected two distinct FP64 gauge fields on one lattice and nsteps >= 0");
    return 1;
  }
  for (int d = 0; d < 4; d++)
    if (out->lat->L[d] != in->lat->L[d]) {
      set_error("ape_smear: lattice extents differ");
      return 1;
    }
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  gauge_written(out);
  hipStream_t st = ctx()->stream;
  if (nsteps == 0) {
    QH_CHECK(hipMemcpyAsync(out->data, in->data, out->bytes, hipMemcpyDeviceToDevice, st));
    QH_CHECK(hipStreamSynchronize(st));
    return 0;
  }
  qcd_hip_field tmp = nullptr;
  if (nsteps > 1) QH_TRY(qcd_hip_gauge_create(&tmp, in->lat, QCD_HIP_F64));
  // ping-pong so that the last step writes out
  qcd_hip_field src = in;
  int rc = 0;
  for (int k = 0; k < nsteps && !rc; k++) {
    qcd_hip_field dst = ((nsteps - 1 - k) % 2 == 0) ? out : tmp;
    rc = ape_step(dst, src, alpha, st);
    src = dst;
  }
  if (!rc && hipStreamSynchronize(st) != hipSuccess) {
    set_error("ape_smear: %s", hipGetErrorString(hipGetLastError()));
    rc = 2;
  }
  qcd_hip_field_destroy(tmp);
  return rc;
}

int qcd_hip_stout_smear(qcd_hip_field out, qcd_hip_field in, double rho, int ndim, int nsteps) {
  QH_TRY(ensure_init());
  if (!out || !in || out == in || out->kind != QCD_HIP_GAUGE || in->kind != QCD_HIP_GAUGE ||
      out->prec != QCD_HIP_F64 || in->prec != QCD_HIP_F64 || out->lat->V != in->lat->V || nsteps < 0) {
    set_error("stout_smear: expected two distinct FP64 gauge fields on one lattice and nsteps >= 0");
    return 1;
  }
  if (ndim != 3 && ndim != 4) {
    set_error("stout_smear: ndim should be either 3 or 4");
    return 1;
  }
  for (int d = 0; d < 4; d++)
    if (out->lat->L[d] != in->lat->L[d]) {
      set_error("stout_smear: lattice extents differ");
      return 1;
    }
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  gauge_written(out);
  hipStream_t st = ctx()->stream;
  if (nsteps == 0) {
    QH_CHECK(hipMemcpyAsync(out->data, in->data, out->bytes, hipMemcpyDeviceToDevice, st));
    QH_CHECK(hipStreamSynchronize(st));
    return 0;
  }
  qcd_hip_field tmp = nullptr;
  if (nsteps > 1) QH_TRY(qcd_hip_gauge_create(&tmp, in->lat, QCD_HIP_F64));
  qcd_hip_field src = in;
  int rc = 0;
  for (int k = 0; k < nsteps && !rc; k++) {
    qcd_hip_field dst = ((nsteps - 1 - k) % 2 == 0) ? out : tmp;
    rc = stout_step(dst, src, rho, ndim, st);
    src = dst;
  }
  if (!rc && hipStreamSynchronize(st) != hipSuccess) {
    set_error("stout_smear: %s", hipGetErrorString(hipGetLastError()));
    rc = 2;
  }
  qcd_hip_field_destroy(tmp);
  return rc;
}

int qcd_hip_plaquette(qcd_hip_field gauge, double *plaq) {
  QH_TRY(ensure_init());
  if (!gauge || !plaq || gauge->kind != QCD_HIP_GAUGE || gauge->prec != QCD_HIP_F64) {
    set_error("plaquette: expected an FP64 gauge field");
    return 1;
  }
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  hipStream_t st = ctx()->stream;
  const DslashArgs a = make_args(gauge->lat, 0.0, 0.0);
  const int grid = grid_for(gauge->lat->V);
  double *part = nullptr;
  QH_TRY(dev_alloc((void **)&part, sizeof(double) * (size_t)grid));
  k_plaquette<<<grid, kBlock, 0, st>>>(a, GaugeIn<QCD_HIP_F64>::of(gauge), part);
  int rc = hipGetLastError() == hipSuccess ? 0 : 2;
  if (!rc) rc = finalize(part, grid, 1, ctx()->scalars, st);
  double sum = 0.0;
  if (!rc) rc = hipMemcpyAsync(&sum, ctx()->scalars, sizeof(double), hipMemcpyDeviceToHost, st) ? 2 : 0;
  if (!rc) rc = hipStreamSynchronize(st) ? 2 : 0;
  dev_free(part, sizeof(double) * (size_t)grid);
  if (rc) {
    set_error("plaquette: kernel failed");
    return rc;
  }
  *plaq = sum / ((double)gauge->lat->V * 3.0 * 6.0);
  return 0;
}

}  // extern "C"
