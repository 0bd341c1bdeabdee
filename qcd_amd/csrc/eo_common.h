// eo_common.h -- even-odd (Schur complement) definitions shared by the
// vector-ALU hop kernels (eo.hip) and the matrix-core ones (mma.hip).
#pragma once

#include "blas_common.h"

namespace qcdhip {

// EO epilogues: out = M^-1 H in | M b - H in | b - H in | M^-1 (b - H in)
enum { EO_MINV = 0, EO_SCHUR = 1, EO_SUB = 2, EO_SUB_MINV = 3 };

struct EoArgs {
  DslashArgs d;            // full-lattice geometry, phases, mass, batch
  unsigned Vh;             // sites per parity
  unsigned par;            // parity of the output sites
  double mi_re, mi_im;     // 1 / (m + 4 + i tm)
  unsigned rgroup;         // batch rows per group of the hop traversal (eo_row)
};

// Block decomposition of the batched hop: one 1D grid of nblk * count
// workgroups, remapped per XCD.  The batch rows go in groups of rgroup; within
// a group the row is the fastest index, so a group's rows read each site
// block's links together through one XCD's L2, and successive site blocks of
// the group follow, so the z-neighbour planes read a few site blocks later are
// still in that L2.  rgroup = count shares the links most, rgroup = 1 keeps
// one row's spinors most local.
__device__ __forceinline__ BlockRow eo_row_of(unsigned count, unsigned R, unsigned nb) {
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const unsigned full = count / R, span = nb * R;
  if (lb < full * span) {
    const unsigned g = lb / span, w = lb - g * span;
    return BlockRow{g * R + w % R, w / R};
  }
  const unsigned rem = count - full * R, w = lb - full * span;
  return BlockRow{full * R + w % rem, w / rem};
}
__device__ __forceinline__ BlockRow eo_row(const EoArgs &e) { return eo_row_of(e.d.count, e.rgroup, e.d.nblk); }

// the site of parity par stored at half-lattice entry h
__device__ __forceinline__ Site eo_site(unsigned h, unsigned par, const DslashArgs &a) {
  Site st;
  const unsigned l0 = 2 * h;
  const unsigned r = l0 / a.L[0];
  const unsigned t0 = l0 - r * a.L[0];
  const unsigned r2 = r / a.L[1];
  st.x = r - r2 * a.L[1];
  st.z = r2 / a.L[2];
  st.y = r2 - st.z * a.L[2];
  const unsigned odd = (par + st.x + st.y + st.z) & 1u;
  st.t = t0 + odd;
  st.s = l0 + odd;
  st.valid = true;
  return st;
}

// M v (inv = false) or M^-1 v (inv = true) on one site: spins 0, 1 carry
// m + 4 + i tm, spins 2, 3 its conjugate (g5 = diag(1, 1, -1, -1))
template <bool INV, typename R>
__device__ __forceinline__ void eo_mass(Cx<R> v[12], const EoArgs &e) {
  const R are = INV ? (R)e.mi_re : (R)e.d.mp4;
  const R aim = INV ? (R)e.mi_im : (R)e.d.tm;
#pragma unroll
  for (int k = 0; k < 12; k++) {
    const R im = k < 6 ? aim : -aim;
    v[k] = Cx<R>{fma(are, v[k].re, -(im * v[k].im)), fma(are, v[k].im, im * v[k].re)};
  }
}

// o from H in = -1/2 acc by the epilogue EPI (b read at half-site h)
template <int EPI, int PS, typename R>
__device__ __forceinline__ void eo_epilogue(Cx<R> o[12], const Cx<R> acc[4][3], const EoArgs &e,
                                            const SpinorIn<PS> &b, unsigned h) {
  if constexpr (EPI == EO_MINV) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int c = 0; c < 3; c++) o[3 * q + c] = Cx<R>{R(-0.5) * acc[q][c].re, R(-0.5) * acc[q][c].im};
    eo_mass<true>(o, e);
  } else {
    // b is read once: a streaming load leaves L2 to the neighbour planes
    Cx<R> bv[12];
    b.template load<true>(h, bv);
    if constexpr (EPI == EO_SCHUR) eo_mass<false>(bv, e);
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int c = 0; c < 3; c++)
        o[3 * q + c] = Cx<R>{fma(R(0.5), acc[q][c].re, bv[3 * q + c].re), fma(R(0.5), acc[q][c].im, bv[3 * q + c].im)};
    if constexpr (EPI == EO_SUB_MINV) eo_mass<true>(o, e);
  }
}

// Batched EO hop launch (eo.hip); see there.
int eo_launch(int epi, bool red, qcd_hip_field out, int ro, int out_rs, qcd_hip_field in, int ri, int in_rs,
              qcd_hip_field g, qcd_hip_field b, int rb, int b_rs, const int *ids, int count, double m, double tm,
              unsigned par, double *partials, hipStream_t st, int *grid_out);
// The outer step of a mixed even-odd solve in the inner precision (eo.hip):
// FP64 residual -> even half, odd half and M^-1 (even half) (FP32 inner
// vectors; for FP16 the true residual writes the halves,
// dslash_residual_eo_split); x += merged halves in FP64.
int eo_split_minv(qcd_hip_field r, int f0, qcd_hip_field ev, qcd_hip_field od, qcd_hip_field tev, double m, double tm,
                  const int *ids, int count, hipStream_t st);
int eo_accumulate_p(qcd_hip_field x, int f0, qcd_hip_field ev, qcd_hip_field od, const int *ids, int count,
                    hipStream_t st);
// the same hops of an FP16 batch on the matrix cores (mma.hip): EO_MINV and
// EO_SCHUR, any rhs count
int eo_mma_launch(int epi, const EoArgs &e, qcd_hip_field out, int ro, qcd_hip_field in, int ri, qcd_hip_field g,
                  qcd_hip_field b, int rb, hipStream_t st);
// build the matrix-core hop links of FP16 gauge field g for the phases of
// lattice lat before a graph capture uses them (mma.hip)
int mma_prepare(qcd_hip_field g, qcd_hip_lattice lat, hipStream_t st);

}  // namespace qcdhip
