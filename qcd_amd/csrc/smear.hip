// smear.hip -- 3D Gaussian (Wuppertal) smearing on gfx950.
//
// One iteration on time-slice t (lib/src/qcd_smearing.c:25-107):
//   v(x) <- ( v(x) + alpha * sum_{nu=1..3} [U_nu(x) v(x+nu) + U_nu(x-nu)^+ v(x-nu)] ) / (1 + 6 alpha)
// The hop stays inside the time-slice, so an iteration on one slice reads and
// writes only that slice: V/T sites instead of the reference's whole-volume
// zero / add passes (lib/src/qcd_smearing.c:45, :102).  The scale and add of
// the reference are fused into the hop kernel.
//
// Layout of the iterations.  Sites of one time-slice lie T apart in the field
// (l = t + T(x + X(y + Yz))), so a 64-site tile of the field holds at most
// ceil(64/T) of them and a wavefront's loads of one component touch as many
// cache lines as it has lanes.  The iterations therefore run on two compact
// buffers A and B that hold the slice in (x, y, z) order, r = x + X(y + Yz):
//   gather v -> A;  A -> B -> A ...;  the last iteration writes v itself,
// so the strided field is read once and written once per call whatever the
// iteration count.  The slice's spatial links are copied once per call into a
// compact gauge field in the same order, so every iteration reads links and
// spinors with whole-tile loads.  Every launch covers all right-hand sides,
// with the right-hand side the fastest block index so the rows of one block
// of sites run together on one XCD and share its links through that XCD's L2.
// The whole-volume form (t < 0) is the same with T slices per pass; its
// compact order is the field's own.
#include <type_traits>
#include "blas_common.h"

namespace qcdhip {

struct GaussArgs {
  unsigned T, X, Y, Z;   // lattice extents
  unsigned tn;           // time values per pass: 1 (one slice) or T (whole volume)
  unsigned t0;           // the slice when tn == 1
  unsigned n;            // sites per right-hand side: tn X Y Z
  unsigned count;        // right-hand sides in the launch
  unsigned ncb;          // k_gauss: workgroups of columns per z chunk
  unsigned zlen;         // k_gauss: z steps per lane
};

// lane -> (row, compact site j); j = tt + tn r with r = x + X(y + Yz)
struct GaussSite {
  unsigned row, tt, r, x, y, z;
  bool valid;
};

__device__ __forceinline__ GaussSite gauss_site(const GaussArgs &g) {
  GaussSite q;
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  q.row = lb % g.count;
  const unsigned j = (lb / g.count) * kBlock + threadIdx.x;
  q.valid = j < g.n;
  unsigned r = j;
  q.tt = 0;
  if (g.tn > 1) {
    r = j / g.tn;
    q.tt = j - r * g.tn;
  }
  q.r = r;
  const unsigned r2 = r / g.X;
  q.x = r - r2 * g.X;
  q.z = r2 / g.Y;
  q.y = r2 - q.z * g.Y;
  return q;
}

// site of slice point r3 in the compact buffers and in the field
__device__ __forceinline__ unsigned compact_of(const GaussArgs &g, const GaussSite &q, unsigned r3) {
  return q.tt + g.tn * r3;
}
__device__ __forceinline__ unsigned field_of(const GaussArgs &g, const GaussSite &q, unsigned r3) {
  return (g.tn > 1 ? q.tt : g.t0) + g.T * r3;
}

// One iteration: input from a compact buffer, output to a compact buffer
// (OUT_FIELD false) or to the field (OUT_FIELD true).  U holds the links in
// compact order.
//
// z-march (as the Dslash's): a lane owns one (t, x, y) column of one rhs and
// steps through zlen planes.  It carries the centre spinor of the next step
// (loaded as this step's forward z neighbour) and the backward z hop
// U_z(z-1)^+ psi(z-1) (made while it held psi(z-1) and U_z(z-1)), so a step
// reads only the z+1 plane, the x and y neighbours (the neighbouring lanes'
// centres, read in the same step) and the links.  Hops are added into acc
// one at a time and pinned there.  acc = psi / alpha + sum of hops and
// out = alpha nrm acc; alpha = 0 never reaches the kernel (the step is the
// identity then).  Rows are the fastest block index: the rows of a column
// block read its links through one L2.
template <int P, bool OUT_FIELD>
__global__ void __launch_bounds__(kBlock) k_gauss(GaussArgs g, SpinorIn<P> in, GaugeIn<P> U, SpinorOut<P> out,
                                                  double alpha, double nrm) {
  using R = typename Prec<P>::R;
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const unsigned row = lb % g.count, rest = lb / g.count;
  const unsigned zc = rest / g.ncb, col = (rest - zc * g.ncb) * kBlock + threadIdx.x;
  const unsigned PC = g.tn * g.X * g.Y;   // columns per plane
  const unsigned z0 = zc * g.zlen;
  if (col >= PC || z0 >= g.Z) return;
  const unsigned z1 = min(z0 + g.zlen, g.Z);
  GaussSite q;
  q.row = row;
  q.tt = col % g.tn;
  {
    const unsigned rxy = col / g.tn;
    q.y = rxy / g.X;
    q.x = rxy - q.y * g.X;
  }
  q.valid = true;
  const SpinorIn<P> src = in.advanced(row);
  const SpinorOut<P> dst = out.advanced(row);
  const unsigned XY = g.X * g.Y;
  const R ia = (R)(1.0 / alpha), f = (R)(alpha * nrm);
  auto pin = [&](Cx<R> v[12]) {
#pragma unroll
    for (int k = 0; k < 12; k++) asm volatile("" : "+v"(v[k].re), "+v"(v[k].im));
    __builtin_amdgcn_sched_barrier(0);
  };
  // acc += W psi(rr) with W = U_nu(c) (forward) or U_nu(rr)^+ (backward)
  auto hop = [&](Cx<R> acc[12], unsigned rr, unsigned c, int nu, auto dag) {
    Cx<R> ps[12], u[9], y[3];
    const unsigned cc = compact_of(g, q, rr);
    src.load(cc, ps);
    U.load(nu, decltype(dag)::value ? cc : c, u);
#pragma unroll
    for (int sp = 0; sp < 4; sp++) {
      su3_mul<decltype(dag)::value>(y, u, &ps[3 * sp]);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        acc[3 * sp + k].re += y[k].re;
        acc[3 * sp + k].im += y[k].im;
      }
    }
    pin(acc);
  };
  // the carried state: centre psi(z) and bz = U_z(z-1)^+ psi(z-1)
  Cx<R> cen[12], bz[12];
  {
    const unsigned zm = z0 == 0 ? g.Z - 1 : z0 - 1;
    const unsigned rm = q.x + g.X * (q.y + g.Y * zm);
#pragma unroll
    for (int k = 0; k < 12; k++) bz[k] = Cx<R>{R(0), R(0)};
    hop(bz, rm, 0u, 3, std::true_type{});
    src.load(compact_of(g, q, q.x + g.X * (q.y + g.Y * z0)), cen);
  }
  for (unsigned z = z0; z < z1; z++) {
    q.z = z;
    const unsigned r = q.x + g.X * (q.y + g.Y * z);
    q.r = r;
    const unsigned c = compact_of(g, q, r);
    Cx<R> acc[12];
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = Cx<R>{cen[k].re * ia + bz[k].re, cen[k].im * ia + bz[k].im};
    pin(acc);
    hop(acc, q.x == g.X - 1 ? r - (g.X - 1) : r + 1, c, 1, std::false_type{});
    hop(acc, q.x == 0 ? r + (g.X - 1) : r - 1, c, 1, std::true_type{});
    hop(acc, q.y == g.Y - 1 ? r - (g.Y - 1) * g.X : r + g.X, c, 2, std::false_type{});
    hop(acc, q.y == 0 ? r + (g.Y - 1) * g.X : r - g.X, c, 2, std::true_type{});
    // last, the z direction: the next step's backward hop from this centre
    // and U_z here (the centre dies with it), then the forward hop, whose
    // spinor is the next step's centre
    {
      Cx<R> u[9], y[3];
      U.load(3, c, u);
#pragma unroll
      for (int sp = 0; sp < 4; sp++) {
        su3_mul<true>(y, u, &cen[3 * sp]);
#pragma unroll
        for (int k = 0; k < 3; k++) bz[3 * sp + k] = y[k];
      }
      pin(bz);
      const unsigned rf = q.z == g.Z - 1 ? r - (g.Z - 1) * XY : r + XY;
      src.load(compact_of(g, q, rf), cen);
#pragma unroll
      for (int sp = 0; sp < 4; sp++) {
        su3_mul<false>(y, u, &cen[3 * sp]);
#pragma unroll
        for (int k = 0; k < 3; k++) {
          acc[3 * sp + k].re += y[k].re;
          acc[3 * sp + k].im += y[k].im;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = Cx<R>{acc[k].re * f, acc[k].im * f};
    dst.store(OUT_FIELD ? field_of(g, q, r) : c, acc);
  }
}

// Stored values (and FP16 norms) of the slice's sites: field -> compact buffer.
// src and dst point at right-hand side 0; rows lie 12 sstride / 12 dstride
// elements apart.
template <int P>
__global__ void __launch_bounds__(kBlock) k_gauss_gather(GaussArgs g, const typename Prec<P>::T *src,
                                                         const float *snorm, size_t sstride,
                                                         typename Prec<P>::T *dst, float *dnorm, size_t dstride) {
  using T = typename Prec<P>::T;
  constexpr int E = (int)sizeof(T);
  const GaussSite q = gauss_site(g);
  if (!q.valid) return;
  const size_t s = field_of(g, q, q.r), c = compact_of(g, q, q.r);
  const T *a = src + (size_t)q.row * 12 * sstride;
  T *b = dst + (size_t)q.row * 12 * dstride;
#pragma unroll
  for (int k = 0; k < 12; k++) b[tile_index_e<E>(12, k, c)] = a[tile_index_e<E>(12, k, s)];
  if (snorm) dnorm[(size_t)q.row * dstride + c] = snorm[(size_t)q.row * sstride + s];
}

// Stored values of the slice's links in directions 1..3: field -> compact
// gauge field (one lane per slice site).
template <int P>
__global__ void __launch_bounds__(kBlock) k_gauss_gather_links(GaussArgs g, const typename Prec<P>::T *src,
                                                               size_t sstride, typename Prec<P>::T *dst,
                                                               size_t dstride) {
  using T = typename Prec<P>::T;
  constexpr int E = (int)sizeof(T);
  const unsigned r = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (r >= g.X * g.Y * g.Z) return;
  const size_t s = g.t0 + (size_t)g.T * r;
#pragma unroll
  for (int k = 9; k < 36; k++) dst[gauge_index_e<E>(k, r, dstride)] = src[gauge_index_e<E>(k, s, sstride)];
}

// z steps per lane of k_gauss: four z chunks (at least 2 steps).  On a 48^3
// slice with 12 rhs every length from 2 to 24 measured 201-226 us per
// iteration (48: 328).  QCD_HIP_GAUSS_ZLEN overrides it.
static unsigned gauss_zlen(unsigned Z) {
  static const int env = [] {
    const char *e = std::getenv("QCD_HIP_GAUSS_ZLEN");
    return e ? std::atoi(e) : 0;
  }();
  const unsigned r = env > 0 ? (unsigned)env : std::max(2u, (Z + 3) / 4);
  return r > Z ? Z : r;
}

// Rows of one pass: two compact buffers of `rows` right-hand sides each stay
// within 4 GiB (at least one row); QCD_HIP_GAUSS_ROWS lowers the cap.
static int gauss_rows(long row_bytes, int nrhs) {
  const long cap = 4L << 30;
  long rows = cap / (2 * row_bytes);
  if (const char *e = std::getenv("QCD_HIP_GAUSS_ROWS")) {
    const long r = std::atol(e);
    if (r > 0 && r < rows) rows = r;
  }
  if (rows < 1) rows = 1;
  return (int)(rows < nrhs ? rows : nrhs);
}

// lattice of the compact buffers and the compact links of one slice, kept
// for the life of the process (guarded by state_mutex)
static qcd_hip_lattice_s g_slice_lat;
static qcd_hip_field g_slice_links = nullptr;
static long g_slice_links_stride = 0;   // stride g_slice_links was allocated for

template <int P>
static int gauss_run(qcd_hip_field v, qcd_hip_field gauge, qcd_hip_field buf, int rows, double alpha, int t0,
                     int niter, hipStream_t st) {
  using T = typename Prec<P>::T;
  const qcd_hip_lattice L = v->lat;
  GaussArgs g;
  g.T = L->L[0];
  g.X = L->L[1];
  g.Y = L->L[2];
  g.Z = L->L[3];
  g.tn = t0 >= 0 ? 1u : g.T;
  g.t0 = t0 >= 0 ? (unsigned)t0 : 0u;
  g.n = g.tn * g.X * g.Y * g.Z;
  g.count = 0;   // set per pass below
  const double nrm = 1.0 / (1.0 + 6.0 * alpha);
  const size_t vst = (size_t)L->stride, bst = (size_t)buf->lat->stride;
  qcd_hip_field links = gauge;
  if (t0 >= 0) {
    // g_slice_lat is rewritten on every call, so the allocation's own stride
    // is kept apart from it
    if (g_slice_links && (g_slice_links->prec != v->prec || g_slice_links_stride != buf->lat->stride)) {
      qcd_hip_field_destroy(g_slice_links);
      g_slice_links = nullptr;
    }
    if (!g_slice_links) {
      QH_TRY(qcd_hip_gauge_create(&g_slice_links, buf->lat, v->prec));
      g_slice_links_stride = buf->lat->stride;
    }
    g_slice_links->lat = buf->lat;
    links = g_slice_links;
    k_gauss_gather_links<P><<<grid_for((long)g.X * g.Y * g.Z), kBlock, 0, st>>>(
        g, (const T *)gauge->data, vst, (T *)links->data, bst);
    QH_CHECK(hipGetLastError());
  }
  const GaugeIn<P> U = GaugeIn<P>::of(links);
  const unsigned nb = (unsigned)grid_for(g.n);
  g.zlen = gauss_zlen(g.Z);
  g.ncb = (unsigned)grid_for((long)g.tn * g.X * g.Y);
  const unsigned nzc = (g.Z + g.zlen - 1) / g.zlen;
  for (int r0 = 0; r0 < v->nrhs; r0 += rows) {
    g.count = (unsigned)std::min(rows, v->nrhs - r0);
    const unsigned grid = nb * g.count, zgrid = g.ncb * nzc * g.count;
    const SpinorOut<P> vout = SpinorOut<P>::of(v, r0);
    const SpinorIn<P> ain = SpinorIn<P>::of(buf, 0), bin = SpinorIn<P>::of(buf, rows);
    const SpinorOut<P> aout = SpinorOut<P>::of(buf, 0), bout = SpinorOut<P>::of(buf, rows);
    k_gauss_gather<P><<<grid, kBlock, 0, st>>>(g, (const T *)v->data + (size_t)r0 * 12 * vst,
                                               v->norm ? v->norm + (size_t)r0 * vst : nullptr, vst, (T *)buf->data,
                                               buf->norm, bst);
    QH_CHECK(hipGetLastError());
    for (int it = 0; it < niter; it++) {
      const bool from_a = (it % 2) == 0;
      const SpinorIn<P> &src = from_a ? ain : bin;
      if (it == niter - 1) k_gauss<P, true><<<zgrid, kBlock, 0, st>>>(g, src, U, vout, alpha, nrm);
      else k_gauss<P, false><<<zgrid, kBlock, 0, st>>>(g, src, U, from_a ? bout : aout, alpha, nrm);
      QH_CHECK(hipGetLastError());
    }
  }
  return 0;
}

int gauss_smear_field(qcd_hip_field v, qcd_hip_field gauge, double alpha, int t0, int niter, hipStream_t st) {
  const qcd_hip_lattice L = v->lat;
  qcd_hip_lattice bl = L;
  if (t0 >= 0) {
    qcd_hip_lattice_s &s = g_slice_lat;
    s = qcd_hip_lattice_s();
    s.L[0] = 1;
    for (int d = 1; d < 4; d++) s.L[d] = L->L[d];
    s.V = (long)L->L[1] * L->L[2] * L->L[3];
    s.stride = round_up(s.V, 64);
    for (int d = 0; d < 4; d++) s.ph_re[d] = 1.0;
    bl = &g_slice_lat;
  }
  const long row_bytes = 12L * bl->stride * (long)spinor_comp_bytes(v->prec) + (v->prec == QCD_HIP_F16 ? 4L * bl->stride : 0);
  const int rows = gauss_rows(row_bytes, v->nrhs);
  qcd_hip_field buf = nullptr;
  QH_TRY(scratch_spinor(0, bl, v->prec, 2 * rows, &buf));
  if (v->prec == QCD_HIP_F64) return gauss_run<QCD_HIP_F64>(v, gauge, buf, rows, alpha, t0, niter, st);
  if (v->prec == QCD_HIP_F32) return gauss_run<QCD_HIP_F32>(v, gauge, buf, rows, alpha, t0, niter, st);
  return gauss_run<QCD_HIP_F16>(v, gauge, buf, rows, alpha, t0, niter, st);
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
  // alpha = 0: every step is psi / (1 + 0) = psi exactly
  if (niter == 0 || alpha == 0.0) return 0;
  QH_TRY(gauss_smear_field(v, gauge, alpha, t < 0 ? -1 : t, niter, ctx()->stream));
  QH_CHECK(hipStreamSynchronize(ctx()->stream));
  return 0;
}

}  // extern "C"
