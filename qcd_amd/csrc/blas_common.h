// blas_common.h -- declarations shared by the BLAS, smearing and solver files.
#pragma once

#include "dslash_common.h"

namespace qcdhip {

// A set of spinors stored back to back (the rhs of one field), addressed by
// index: the kernels of the solver take the whole Krylov basis this way.
template <int P>
struct SpinorSet {
  using T = typename Prec<P>::T;
  using R = typename Prec<P>::R;
  T *base;            // component 0, tile 0 of vector 0
  float *norm;        // F16 only
  size_t vstride;     // elements between vectors
  size_t nstride;     // norms between vectors
  static SpinorSet of(qcd_hip_field f) {
    SpinorSet s;
    s.base = (T *)f->data;
    s.norm = f->norm;
    s.vstride = 12 * (size_t)f->lat->stride;
    s.nstride = (size_t)f->lat->stride;
    return s;
  }
  __device__ __forceinline__ void load(int j, unsigned s, Cx<R> v[12]) const {
    const T *b = base + (size_t)j * vstride;
    const unsigned off = SpinorIn<P>::offset(s);
    if constexpr (P == QCD_HIP_F16) {
      const R n = __builtin_nontemporal_load(&norm[(size_t)j * nstride + s]);
#pragma unroll
      for (int g = 0; g < 3; g++) {
        const u32x4 q = __builtin_nontemporal_load((const u32x4 *)((const char *)b + g * 4 * kTile * 4 + off));
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const Cx<R> c = to_cx<R>(as_half2(q[i]));
          v[4 * g + i] = Cx<R>{c.re * n, c.im * n};
        }
      }
    } else if constexpr (P == QCD_HIP_F32) {
#pragma unroll
      for (int g = 0; g < 6; g++) {
        typedef float f4v_ __attribute__((ext_vector_type(4)));
        const f4v_ q = __builtin_nontemporal_load((const f4v_ *)((const char *)b + g * 2 * kTile * 8 + off));
        v[2 * g] = Cx<R>{q.x, q.y};
        v[2 * g + 1] = Cx<R>{q.z, q.w};
      }
    } else {
#pragma unroll
      for (int k = 0; k < 12; k++) v[k] = to_cx<R>(ld_plane_nt(b + k * kTile, off));
    }
  }
  // store v and leave in v the stored values (FP16: rounded to the stored
  // half form and multiplied back by the site norm), without reading them
  __device__ __forceinline__ void store_back(int j, unsigned s, Cx<R> v[12]) const {
    if constexpr (P == QCD_HIP_F16) {
      T *b = base + (size_t)j * vstride;
      const unsigned off = SpinorIn<P>::offset(s);
      R mx = R(0);
#pragma unroll
      for (int k = 0; k < 12; k++) mx = fmaxf(mx, fmaxf(fabsf(v[k].re), fabsf(v[k].im)));
      if (!(mx >= FLT_MIN)) mx = R(0);
      const R inv = mx > R(0) ? R(1) / mx : R(0);
      __builtin_nontemporal_store((float)mx, &norm[(size_t)j * nstride + s]);
#pragma unroll
      for (int g = 0; g < 3; g++) {
        u32x4 q;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          q[i] = as_u32(from_cx<T>(Cx<R>{v[4 * g + i].re * inv, v[4 * g + i].im * inv}));
          const Cx<R> c = to_cx<R>(as_half2(q[i]));
          v[4 * g + i] = Cx<R>{c.re * mx, c.im * mx};
        }
        __builtin_nontemporal_store(q, (u32x4 *)((char *)b + g * 4 * kTile * 4 + off));
      }
    } else {
      store(j, s, v);
    }
  }
  __device__ __forceinline__ void store(int j, unsigned s, const Cx<R> v[12]) const {
    T *b = base + (size_t)j * vstride;
    const unsigned off = SpinorIn<P>::offset(s);
    if constexpr (P == QCD_HIP_F16) {
      R mx = R(0);
#pragma unroll
      for (int k = 0; k < 12; k++) mx = fmaxf(mx, fmaxf(fabsf(v[k].re), fabsf(v[k].im)));
      // a site whose largest entry is below the smallest normal float is
      // stored as zero: 1/mx would overflow and 0 * inf give NaN
      if (!(mx >= FLT_MIN)) mx = R(0);
      const R inv = mx > R(0) ? R(1) / mx : R(0);
      __builtin_nontemporal_store((float)mx, &norm[(size_t)j * nstride + s]);
#pragma unroll
      for (int g = 0; g < 3; g++) {
        u32x4 q;
#pragma unroll
        for (int i = 0; i < 4; i++)
          q[i] = as_u32(from_cx<T>(Cx<R>{v[4 * g + i].re * inv, v[4 * g + i].im * inv}));
        __builtin_nontemporal_store(q, (u32x4 *)((char *)b + g * 4 * kTile * 4 + off));
      }
    } else if constexpr (P == QCD_HIP_F32) {
      typedef float f4v_ __attribute__((ext_vector_type(4)));
#pragma unroll
      for (int g = 0; g < 6; g++)
        __builtin_nontemporal_store(
            f4v_{(float)v[2 * g].re, (float)v[2 * g].im, (float)v[2 * g + 1].re, (float)v[2 * g + 1].im},
            (f4v_ *)((char *)b + g * 2 * kTile * 8 + off));
    } else {
      typedef double d2v_ __attribute__((ext_vector_type(2)));
#pragma unroll
      for (int k = 0; k < 12; k++)
        __builtin_nontemporal_store(d2v_{(double)v[k].re, (double)v[k].im},
                                    (d2v_ *)((char *)(b + k * kTile) + off));
    }
  }
};

// grid of the reduction kernels: depends only on the volume
int reduce_grid(long V);
// sum n_sums partial arrays of length nparts into out (device), one workgroup
int finalize(const double *partials, int nparts, int n_sums, double *out, hipStream_t st);

// out[ro] = a x[rx] + b y[ry] (all sites, or slice t0 >= 0), asynchronous
int lincomb_launch(qcd_hip_field out, int ro, double are, double aim, qcd_hip_field x, int rx, double bre, double bim,
                   qcd_hip_field y, int ry, int t0, hipStream_t st);
// |v[rhs]|^2 workgroup partials (at most kMaxReduceBlocks) into partials,
// their count in *nparts; asynchronous on st
int norm2_launch(qcd_hip_field v, int rhs, double *partials, int *nparts, hipStream_t st);
// cached scratch spinor number `slot` (resized on demand), owned by the library
int scratch_spinor(int slot, qcd_hip_lattice lat, int prec, int nrhs, qcd_hip_field *out);
// niter >= 1 iterations of Gauss smearing of every rhs of v on slice t0 (all
// slices when t0 < 0), through the library's scratch spinor 0
int gauss_smear_field(qcd_hip_field v, qcd_hip_field g, double alpha, int t0, int niter, hipStream_t st);

// even-odd preconditioning applies (every extent even), eo.hip
bool eo_supported(qcd_hip_lattice lat);

// device bytes a solver with these parameters allocates on a lattice of
// extents L (qcd_hip_solver_create plus the inner-precision links of
// qcd_hip_solver_set_gauge), solver.hip
size_t solver_device_bytes(const int L[4], const qcd_hip_solver_params *prm);

// device-side synthetic inputs (asynchronous on st)
int gauge_generate_device(qcd_hip_field g, double eps, uint64_t seed, hipStream_t st);
int point_source_device(qcd_hip_field f, int rhs, long site, int spin, int col, hipStream_t st);

}  // namespace qcdhip
