// blas.hip -- vector BLAS and deterministic FP64 reductions on spinor fields.
//
// Reductions: a fixed grid (depending only on the lattice volume) writes one
// partial per workgroup; a single workgroup then sums the partials in a fixed
// order.  Results are therefore bitwise reproducible run to run and do not
// depend on what else shares the GPU (lib/src/qcd_blas.c:456-491 sums
// sequentially; the difference is rounding of order 1e-16 * sqrt(V)).
#include "blas_common.h"

namespace qcdhip {

template <int P>
__global__ void __launch_bounds__(kBlock) k_norm2(SpinorIn<P> v, unsigned V, double *partials) {
  using R = typename Prec<P>::R;
  double acc = 0.0;
  for (unsigned s = blockIdx.x * kBlock + threadIdx.x; s < V; s += gridDim.x * kBlock) {
    Cx<R> x[12];
    v.load(s, x);
#pragma unroll
    for (int k = 0; k < 12; k++) acc += (double)x[k].re * (double)x[k].re + (double)x[k].im * (double)x[k].im;
  }
  block_reduce_store(acc, partials, blockIdx.x);
}

template <int P>
__global__ void __launch_bounds__(kBlock) k_dot(SpinorIn<P> a, SpinorIn<P> b, unsigned V, double *partials) {
  using R = typename Prec<P>::R;
  double re = 0.0, im = 0.0;
  for (unsigned s = blockIdx.x * kBlock + threadIdx.x; s < V; s += gridDim.x * kBlock) {
    Cx<R> x[12], y[12];
    a.load(s, x);
    b.load(s, y);
#pragma unroll
    for (int k = 0; k < 12; k++) {
      re += (double)x[k].re * (double)y[k].re + (double)x[k].im * (double)y[k].im;
      im += (double)x[k].re * (double)y[k].im - (double)x[k].im * (double)y[k].re;
    }
  }
  block_reduce_store(re, partials, blockIdx.x);
  __syncthreads();
  block_reduce_store(im, partials + gridDim.x, blockIdx.x);
}

// y = a x + y
template <int P>
__global__ void __launch_bounds__(kBlock) k_axpy(SpinorIn<P> x, SpinorIn<P> yin, SpinorOut<P> y, double are,
                                                 double aim, unsigned V) {
  using R = typename Prec<P>::R;
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  Cx<R> xv[12], yv[12];
  x.load(s, xv);
  yin.load(s, yv);
  const R ar = (R)are, ai = (R)aim;
#pragma unroll
  for (int k = 0; k < 12; k++)
    yv[k] = Cx<R>{yv[k].re + ar * xv[k].re - ai * xv[k].im, yv[k].im + ar * xv[k].im + ai * xv[k].re};
  y.store(s, yv);
}

// out = a x + b y on every site, or only on time-slice t0 when t0 >= 0
// (other sites of out are left untouched); y may be absent (b = 0).
template <int P>
__global__ void __launch_bounds__(kBlock) k_lincomb(SpinorIn<P> x, SpinorIn<P> y, bool has_y, SpinorOut<P> out,
                                                    double are, double aim, double bre, double bim, unsigned V,
                                                    int t0, unsigned L0) {
  using R = typename Prec<P>::R;
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  if (t0 >= 0 && (int)(s % L0) != t0) return;
  Cx<R> xv[12], yv[12];
  x.load(s, xv);
  if (has_y) y.load(s, yv);
  const R ar = (R)are, ai = (R)aim, br = (R)bre, bi = (R)bim;
#pragma unroll
  for (int k = 0; k < 12; k++) {
    Cx<R> o{ar * xv[k].re - ai * xv[k].im, ar * xv[k].im + ai * xv[k].re};
    if (has_y) {
      o.re += br * yv[k].re - bi * yv[k].im;
      o.im += br * yv[k].im + bi * yv[k].re;
    }
    xv[k] = o;
  }
  out.store(s, xv);
}

// Sum n_sums interleaved partial arrays (each nparts long) into out[0..n_sums).
__global__ void __launch_bounds__(kBlock) k_finalize(const double *partials, int nparts, int n_sums, double *out) {
  __shared__ double sh[kBlock];
  for (int j = 0; j < n_sums; j++) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kBlock) v += partials[(size_t)j * nparts + i];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = sh[0];
    __syncthreads();
  }
}

int reduce_grid(long V) {
  int g = grid_for(V);
  return g < kMaxReduceBlocks ? g : kMaxReduceBlocks;
}

int finalize(const double *partials, int nparts, int n_sums, double *out, hipStream_t st) {
  k_finalize<<<1, kBlock, 0, st>>>(partials, nparts, n_sums, out);
  QH_CHECK(hipGetLastError());
  return 0;
}

static int read_scalars(int n, double *host) {
  Context *c = ctx();
  QH_CHECK(hipMemcpyAsync(c->host_scalars, c->scalars, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  QH_CHECK(hipStreamSynchronize(c->stream));
  std::memcpy(host, c->host_scalars, sizeof(double) * n);
  return 0;
}

static int check_rhs(qcd_hip_field f, int r, const char *who) {
  if (!f || f->kind != QCD_HIP_SPINOR || r < 0 || r >= f->nrhs) {
    set_error("%s: bad spinor field or rhs index", who);
    return 1;
  }
  return 0;
}

int lincomb_launch(qcd_hip_field out, int ro, double are, double aim, qcd_hip_field x, int rx, double bre, double bim,
                   qcd_hip_field y, int ry, int t0, hipStream_t st) {
  QH_TRY(check_rhs(out, ro, "lincomb"));
  QH_TRY(check_rhs(x, rx, "lincomb"));
  if (y) QH_TRY(check_rhs(y, ry, "lincomb"));
  if (x->prec != out->prec || x->lat->V != out->lat->V || (y && (y->prec != out->prec || y->lat->V != out->lat->V))) {
    set_error("lincomb: fields differ in precision or lattice");
    return 1;
  }
  if (t0 >= out->lat->L[0]) {
    set_error("lincomb: time-slice %d outside 0..%d", t0, out->lat->L[0] - 1);
    return 1;
  }
  const unsigned V = (unsigned)out->lat->V, L0 = (unsigned)out->lat->L[0];
  const bool hy = y != nullptr;
  qcd_hip_field yy = hy ? y : x;
  const int ryy = hy ? ry : rx;
  switch (out->prec) {
    case QCD_HIP_F64:
      k_lincomb<QCD_HIP_F64><<<grid_for(V), kBlock, 0, st>>>(SpinorIn<QCD_HIP_F64>::of(x, rx), SpinorIn<QCD_HIP_F64>::of(yy, ryy), hy, SpinorOut<QCD_HIP_F64>::of(out, ro), are, aim, bre, bim, V, t0, L0);
      break;
    case QCD_HIP_F32:
      k_lincomb<QCD_HIP_F32><<<grid_for(V), kBlock, 0, st>>>(SpinorIn<QCD_HIP_F32>::of(x, rx), SpinorIn<QCD_HIP_F32>::of(yy, ryy), hy, SpinorOut<QCD_HIP_F32>::of(out, ro), are, aim, bre, bim, V, t0, L0);
      break;
    default:
      k_lincomb<QCD_HIP_F16><<<grid_for(V), kBlock, 0, st>>>(SpinorIn<QCD_HIP_F16>::of(x, rx), SpinorIn<QCD_HIP_F16>::of(yy, ryy), hy, SpinorOut<QCD_HIP_F16>::of(out, ro), are, aim, bre, bim, V, t0, L0);
      break;
  }
  QH_CHECK(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------ propagator columns
// A host propagator is [V][4 mu][4 nu][3 c1][3 c2] complex (include/qcd.h:110-117).
// Its column (nu, c2) is the spinor psi[mu][c1] = P[mu][nu][c1][c2]
// (qcd_copyVectorPropagator, lib/src/qcd_blas.c:182-212); here column
// 3 nu + c2 is rhs r0 + 3 nu + c2 of a spinor field.  Grid (sites, 12
// columns): consecutive lanes take consecutive sites of one column, so the
// tiled field is written (read) in whole segments; each site's 2304-byte
// block of the propagator is read (written) by its 12 column lanes.
template <int P>
__global__ void __launch_bounds__(kBlock) k_prop_gather(const double2 *__restrict__ prop, SpinorSet<P> out, int r0,
                                                        unsigned V) {
  using R = typename Prec<P>::R;
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  const int col = blockIdx.y, nu = col / 3, c2 = col % 3;
  const double2 *p = prop + (size_t)s * 144 + nu * 9 + c2;
  Cx<R> v[12];
#pragma unroll
  for (int mu = 0; mu < 4; mu++)
#pragma unroll
    for (int c1 = 0; c1 < 3; c1++) {
      const double2 d = p[mu * 36 + c1 * 3];
      v[3 * mu + c1] = Cx<R>{(R)d.x, (R)d.y};
    }
  out.store(r0 + col, s, v);
}

template <int P>
__global__ void __launch_bounds__(kBlock) k_prop_scatter(SpinorSet<P> in, int r0, double2 *__restrict__ prop,
                                                         unsigned V) {
  using R = typename Prec<P>::R;
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= V) return;
  const int col = blockIdx.y, nu = col / 3, c2 = col % 3;
  Cx<R> v[12];
  in.load(r0 + col, s, v);
  double2 *p = prop + (size_t)s * 144 + nu * 9 + c2;
#pragma unroll
  for (int mu = 0; mu < 4; mu++)
#pragma unroll
    for (int c1 = 0; c1 < 3; c1++) p[mu * 36 + c1 * 3] = make_double2((double)v[3 * mu + c1].re, (double)v[3 * mu + c1].im);
}

// |v[rhs]|^2 workgroup partials into partials[0 .. *nparts), asynchronous on st
int norm2_launch(qcd_hip_field v, int rhs, double *partials, int *nparts, hipStream_t st) {
  QH_TRY(check_rhs(v, rhs, "norm2"));
  const unsigned V = (unsigned)v->lat->V;
  const int g = reduce_grid(V);
  switch (v->prec) {
    case QCD_HIP_F64: k_norm2<QCD_HIP_F64><<<g, kBlock, 0, st>>>(SpinorIn<QCD_HIP_F64>::of(v, rhs), V, partials); break;
    case QCD_HIP_F32: k_norm2<QCD_HIP_F32><<<g, kBlock, 0, st>>>(SpinorIn<QCD_HIP_F32>::of(v, rhs), V, partials); break;
    default: k_norm2<QCD_HIP_F16><<<g, kBlock, 0, st>>>(SpinorIn<QCD_HIP_F16>::of(v, rhs), V, partials); break;
  }
  QH_CHECK(hipGetLastError());
  *nparts = g;
  return 0;
}

}  // namespace qcdhip

using namespace qcdhip;

extern "C" {

// propagator (host AoS) -> rhs r0 .. r0+11 of f (up = true), or back
static int propagator_copy(qcd_hip_field f, int r0, double *prop, bool up) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());   // library stream
  QH_TRY(ensure_init());
  if (!f || f->kind != QCD_HIP_SPINOR || !prop || r0 < 0 || r0 + 12 > f->nrhs) {
    set_error("propagator %s: need a spinor field with rhs r0 .. r0+11 and a host propagator",
              up ? "upload" : "download");
    return 1;
  }
  const unsigned V = (unsigned)f->lat->V;
  const size_t bytes = (size_t)V * 144 * sizeof(double2);
  double2 *buf = nullptr;
  QH_TRY(dev_alloc((void **)&buf, bytes));
  hipStream_t st = ctx()->stream;
  const dim3 grid((unsigned)grid_for(V), 12);
  int rc = 0;
  if (up && hipMemcpyAsync(buf, prop, bytes, hipMemcpyHostToDevice, st) != hipSuccess) rc = 2;
  if (!rc) {
    switch (f->prec) {
      case QCD_HIP_F64:
        if (up) k_prop_gather<QCD_HIP_F64><<<grid, kBlock, 0, st>>>(buf, SpinorSet<QCD_HIP_F64>::of(f), r0, V);
        else k_prop_scatter<QCD_HIP_F64><<<grid, kBlock, 0, st>>>(SpinorSet<QCD_HIP_F64>::of(f), r0, buf, V);
        break;
      case QCD_HIP_F32:
        if (up) k_prop_gather<QCD_HIP_F32><<<grid, kBlock, 0, st>>>(buf, SpinorSet<QCD_HIP_F32>::of(f), r0, V);
        else k_prop_scatter<QCD_HIP_F32><<<grid, kBlock, 0, st>>>(SpinorSet<QCD_HIP_F32>::of(f), r0, buf, V);
        break;
      default:
        if (up) k_prop_gather<QCD_HIP_F16><<<grid, kBlock, 0, st>>>(buf, SpinorSet<QCD_HIP_F16>::of(f), r0, V);
        else k_prop_scatter<QCD_HIP_F16><<<grid, kBlock, 0, st>>>(SpinorSet<QCD_HIP_F16>::of(f), r0, buf, V);
        break;
    }
    if (hipGetLastError() != hipSuccess) rc = 2;
  }
  if (!rc && !up && hipMemcpyAsync(prop, buf, bytes, hipMemcpyDeviceToHost, st) != hipSuccess) rc = 2;
  if (hipStreamSynchronize(st) != hipSuccess) rc = 2;
  dev_free(buf, bytes);
  if (rc) set_error("propagator %s: device error", up ? "upload" : "download");
  return rc;
}

int qcd_hip_propagator_upload(qcd_hip_field f, int r0, const double *prop) {
  return propagator_copy(f, r0, const_cast<double *>(prop), true);
}

int qcd_hip_propagator_download(qcd_hip_field f, int r0, double *prop) {
  return propagator_copy(f, r0, prop, false);
}

int qcd_hip_norm2(qcd_hip_field v, int rhs, double *out) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());   // shared scratch and stream
  QH_TRY(ensure_init());
  QH_TRY(check_rhs(v, rhs, "norm2"));
  Context *c = ctx();
  const unsigned V = (unsigned)v->lat->V;
  const int g = reduce_grid(V);
  switch (v->prec) {
    case QCD_HIP_F64: k_norm2<QCD_HIP_F64><<<g, kBlock, 0, c->stream>>>(SpinorIn<QCD_HIP_F64>::of(v, rhs), V, c->partials); break;
    case QCD_HIP_F32: k_norm2<QCD_HIP_F32><<<g, kBlock, 0, c->stream>>>(SpinorIn<QCD_HIP_F32>::of(v, rhs), V, c->partials); break;
    default: k_norm2<QCD_HIP_F16><<<g, kBlock, 0, c->stream>>>(SpinorIn<QCD_HIP_F16>::of(v, rhs), V, c->partials); break;
  }
  QH_CHECK(hipGetLastError());
  QH_TRY(finalize(c->partials, g, 1, c->scalars, c->stream));
  return read_scalars(1, out);
}

int qcd_hip_dot(qcd_hip_field a, int ra, qcd_hip_field b, int rb, double out[2]) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());   // shared scratch and stream
  QH_TRY(ensure_init());
  QH_TRY(check_rhs(a, ra, "dot"));
  QH_TRY(check_rhs(b, rb, "dot"));
  if (a->prec != b->prec || a->lat->V != b->lat->V) {
    set_error("dot: fields differ in precision or lattice");
    return 1;
  }
  Context *c = ctx();
  const unsigned V = (unsigned)a->lat->V;
  const int g = reduce_grid(V);
  switch (a->prec) {
    case QCD_HIP_F64: k_dot<QCD_HIP_F64><<<g, kBlock, 0, c->stream>>>(SpinorIn<QCD_HIP_F64>::of(a, ra), SpinorIn<QCD_HIP_F64>::of(b, rb), V, c->partials); break;
    case QCD_HIP_F32: k_dot<QCD_HIP_F32><<<g, kBlock, 0, c->stream>>>(SpinorIn<QCD_HIP_F32>::of(a, ra), SpinorIn<QCD_HIP_F32>::of(b, rb), V, c->partials); break;
    default: k_dot<QCD_HIP_F16><<<g, kBlock, 0, c->stream>>>(SpinorIn<QCD_HIP_F16>::of(a, ra), SpinorIn<QCD_HIP_F16>::of(b, rb), V, c->partials); break;
  }
  QH_CHECK(hipGetLastError());
  QH_TRY(finalize(c->partials, g, 2, c->scalars, c->stream));
  return read_scalars(2, out);
}

int qcd_hip_lincomb(qcd_hip_field out, int ro, double are, double aim, qcd_hip_field x, int rx, double bre,
                    double bim, qcd_hip_field y, int ry, int t0) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());   // shared scratch and stream
  QH_TRY(ensure_init());
  QH_TRY(lincomb_launch(out, ro, are, aim, x, rx, bre, bim, y, ry, t0, ctx()->stream));
  QH_CHECK(hipStreamSynchronize(ctx()->stream));
  return 0;
}

int qcd_hip_axpy(qcd_hip_field y, int ry, double are, double aim, qcd_hip_field x, int rx) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());   // shared scratch and stream
  QH_TRY(ensure_init());
  QH_TRY(check_rhs(y, ry, "axpy"));
  QH_TRY(check_rhs(x, rx, "axpy"));
  if (x->prec != y->prec || x->lat->V != y->lat->V) {
    set_error("axpy: fields differ in precision or lattice");
    return 1;
  }
  Context *c = ctx();
  const unsigned V = (unsigned)y->lat->V;
  switch (y->prec) {
    case QCD_HIP_F64:
      k_axpy<QCD_HIP_F64><<<grid_for(V), kBlock, 0, c->stream>>>(SpinorIn<QCD_HIP_F64>::of(x, rx), SpinorIn<QCD_HIP_F64>::of(y, ry), SpinorOut<QCD_HIP_F64>::of(y, ry), are, aim, V);
      break;
    case QCD_HIP_F32:
      k_axpy<QCD_HIP_F32><<<grid_for(V), kBlock, 0, c->stream>>>(SpinorIn<QCD_HIP_F32>::of(x, rx), SpinorIn<QCD_HIP_F32>::of(y, ry), SpinorOut<QCD_HIP_F32>::of(y, ry), are, aim, V);
      break;
    default:
      k_axpy<QCD_HIP_F16><<<grid_for(V), kBlock, 0, c->stream>>>(SpinorIn<QCD_HIP_F16>::of(x, rx), SpinorIn<QCD_HIP_F16>::of(y, ry), SpinorOut<QCD_HIP_F16>::of(y, ry), are, aim, V);
      break;
  }
  QH_CHECK(hipGetLastError());
  QH_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"
