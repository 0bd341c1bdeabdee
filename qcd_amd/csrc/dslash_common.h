// dslash_common.h -- device building blocks shared by the Dslash kernels:
// storage accessors per precision, the spin-projection table, one hopping
// direction, epilogues and the block reduction.
#pragma once

#include "internal.h"

namespace qcdhip {

template <typename R>
struct Cx {
  R re, im;
};

template <int P>
struct Prec;
template <>
struct Prec<QCD_HIP_F64> {
  using T = double2;
  using R = double;
};
template <>
struct Prec<QCD_HIP_F32> {
  using T = float2;
  using R = float;
};
template <>
struct Prec<QCD_HIP_F16> {
  using T = __half2;
  using R = float;
};

template <typename R, typename T>
__device__ __forceinline__ Cx<R> to_cx(T v) {
  if constexpr (sizeof(T) == 4) {
    float2 f = __half22float2(v);
    return Cx<R>{(R)f.x, (R)f.y};
  } else {
    return Cx<R>{(R)v.x, (R)v.y};
  }
}

template <typename T, typename R>
__device__ __forceinline__ T from_cx(Cx<R> v) {
  if constexpr (sizeof(T) == 16) {
    return make_double2((double)v.re, (double)v.im);
  } else if constexpr (sizeof(T) == 8) {
    return make_float2((float)v.re, (float)v.im);
  } else {
    return __floats2half2_rn((float)v.re, (float)v.im);
  }
}

// ------------------------------------------------------------ accessors
// Base pointers travel as kernel arguments (scalar registers); a lane
// addresses every component with one 32-bit byte offset plus the component's
// distance from its base, which the load encodes as an immediate (at most
// 4095 bytes on gfx950).  So an accessor keeps one base per group of kPer
// components (all 12 spinor components of FP16, 8 of FP32, 4 of FP64) and a
// load needs no address arithmetic beyond the per-site offset; one pointer per
// component would not fit the scalar register file with four fields and
// forces spills into vector registers.  Host code guarantees
// 12 * stride * sizeof(complex) < 2^32 (one rhs of a spinor, one direction
// of a gauge field).
template <typename T>
__device__ __forceinline__ T ld_plane(const T *base, unsigned off) {
  return *(const T *)((const char *)base + off);
}
// the same load with the nontemporal (streaming) cache hint
template <typename T>
__device__ __forceinline__ T ld_plane_nt(const T *base, unsigned off) {
  const T *p = (const T *)((const char *)base + off);
  if constexpr (sizeof(T) == 16) {
    typedef double v2d __attribute__((ext_vector_type(2)));
    const v2d v = __builtin_nontemporal_load((const v2d *)p);
    T r;
    r.x = v.x;
    r.y = v.y;
    return r;
  } else {
    typedef float v2f __attribute__((ext_vector_type(2)));
    const v2f v = __builtin_nontemporal_load((const v2f *)p);
    T r;
    r.x = v.x;
    r.y = v.y;
    return r;
  }
}
template <typename T>
__device__ __forceinline__ void st_plane(T *base, unsigned off, T v) {
  *(T *)((char *)base + off) = v;
}

// FP16 fields are grouped (tile_index_e): a spinor site's 12 components are
// 3 groups of 4 (16 bytes each, 1024 bytes apart), a link's 9 are 3 groups
// of 3 (12 bytes each, 768 bytes apart).
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
struct alignas(4) U32x3 {
  unsigned x, y, z;
};
__device__ __forceinline__ __half2 as_half2(unsigned v) { return __builtin_bit_cast(__half2, v); }
__device__ __forceinline__ unsigned as_u32(__half2 v) { return __builtin_bit_cast(unsigned, v); }

// components per base pointer for complex type T, and bases for n components
template <typename T>
constexpr int comp_per_base() { return 4096 / (kTile * (int)sizeof(T)); }
template <typename T, int N>
constexpr int num_bases() { return (N + comp_per_base<T>() - 1) / comp_per_base<T>(); }

template <int P>
struct SpinorIn {
  using T = typename Prec<P>::T;
  using R = typename Prec<P>::R;
  static constexpr int kPer = comp_per_base<T>(), kNb = num_bases<T, 12>();
  const T *p[kNb];       // component kPer*i of tile 0
  const float *norm;
  unsigned vstride;      // elements per vector (advance by whole vectors)
  static SpinorIn of(qcd_hip_field f, int rhs) {
    SpinorIn s;
    const size_t st = (size_t)f->lat->stride;
    const T *b = (const T *)f->data + (size_t)rhs * 12 * st;
    for (int i = 0; i < kNb; i++) s.p[i] = b + i * kPer * kTile;
    s.norm = f->norm ? f->norm + (size_t)rhs * st : nullptr;
    s.vstride = (unsigned)(12 * st);
    return s;
  }
  // the same accessor moved forward by nv vectors of the field
  __device__ __forceinline__ SpinorIn advanced(unsigned nv) const {
    SpinorIn s = *this;
    const size_t e = (size_t)nv * vstride;
#pragma unroll
    for (int i = 0; i < kNb; i++) s.p[i] += e;
    if (norm) s.norm += e / 12;
    return s;
  }
  // component k of tile 0
  __device__ __forceinline__ const T *comp(int k) const { return p[k / kPer] + (k % kPer) * kTile; }
  // components per 16-byte group (tile_index_e): 4 (FP16), 2 (FP32), 1 (FP64)
  static constexpr int kG = 16 / (int)sizeof(T);
  // byte offset of site s from the tile-0 pointer of its first component
  // (group)
  static __device__ __forceinline__ unsigned offset(unsigned s) {
    return ((s / kTile) * (12u * kTile) + (s % kTile) * (unsigned)kG) * (unsigned)sizeof(T);
  }
  // stored values of site s (and its FP16 norm); NT: nontemporal
  // (streaming) loads, for values read once
  template <bool NT = false>
  __device__ __forceinline__ void load_raw(unsigned s, T r[12], float &n) const {
    const unsigned off = offset(s);
    if constexpr (P == QCD_HIP_F16) {
#pragma unroll
      for (int g = 0; g < 3; g++) {
        const u32x4 *a = (const u32x4 *)((const char *)p[0] + g * 4 * kTile * 4 + off);
        u32x4 q;
        if constexpr (NT) q = __builtin_nontemporal_load(a);
        else q = *a;
#pragma unroll
        for (int i = 0; i < 4; i++) r[4 * g + i] = as_half2(q[i]);
      }
      if constexpr (NT) n = __builtin_nontemporal_load((const float *)((const char *)norm + s * 4u));
      else n = ld_plane(norm, s * 4u);
    } else if constexpr (P == QCD_HIP_F32) {
      typedef float f4v_ __attribute__((ext_vector_type(4)));
#pragma unroll
      for (int g = 0; g < 6; g++) {
        // groups 4 and 5 start at the second base (element 8 * 64), so
        // every immediate offset stays below 4096 bytes
        const f4v_ *a = (const f4v_ *)((const char *)p[g / 4] + (g % 4) * 2 * kTile * 8 + off);
        f4v_ q;
        if constexpr (NT) q = __builtin_nontemporal_load(a);
        else q = *a;
        r[2 * g] = make_float2(q.x, q.y);
        r[2 * g + 1] = make_float2(q.z, q.w);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 12; k++) r[k] = NT ? ld_plane_nt(comp(k), off) : ld_plane(comp(k), off);
    }
  }
  static __device__ __forceinline__ void convert(const T r[12], float n, Cx<R> v[12]) {
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = to_cx<R>(r[k]);
    if constexpr (P == QCD_HIP_F16) {
#pragma unroll
      for (int k = 0; k < 12; k++) {
        v[k].re *= (R)n;
        v[k].im *= (R)n;
      }
    }
  }
  template <bool NT = false>
  __device__ __forceinline__ void load(unsigned s, Cx<R> v[12]) const {
    T r[12];
    float n = 0.f;
    load_raw<NT>(s, r, n);
    convert(r, n, v);
  }
};

template <int P>
struct SpinorOut {
  using T = typename Prec<P>::T;
  using R = typename Prec<P>::R;
  static constexpr int kPer = comp_per_base<T>(), kNb = num_bases<T, 12>();
  T *p[kNb];
  float *norm;
  unsigned vstride;
  static SpinorOut of(qcd_hip_field f, int rhs) {
    SpinorOut s;
    const size_t st = (size_t)f->lat->stride;
    T *b = (T *)f->data + (size_t)rhs * 12 * st;
    for (int i = 0; i < kNb; i++) s.p[i] = b + i * kPer * kTile;
    s.norm = f->norm ? f->norm + (size_t)rhs * st : nullptr;
    s.vstride = (unsigned)(12 * st);
    return s;
  }
  __device__ __forceinline__ SpinorOut advanced(unsigned nv) const {
    SpinorOut s = *this;
    const size_t e = (size_t)nv * vstride;
#pragma unroll
    for (int i = 0; i < kNb; i++) s.p[i] += e;
    if (norm) s.norm += e / 12;
    return s;
  }
  __device__ __forceinline__ T *comp(int k) const { return p[k / kPer] + (k % kPer) * kTile; }
  // NT: nontemporal (streaming) stores, for outputs no later kernel of the
  // same pass reads back through L2
  template <bool NT = false>
  __device__ __forceinline__ void store(unsigned s, const Cx<R> v[12]) const {
    const unsigned off = SpinorIn<P>::offset(s);
    if constexpr (P == QCD_HIP_F16) {
      R mx = R(0);
#pragma unroll
      for (int k = 0; k < 12; k++) mx = fmaxf(mx, fmaxf(fabsf(v[k].re), fabsf(v[k].im)));
      // a site whose largest entry is below the smallest normal float is
      // stored as zero: 1/mx would overflow and 0 * inf give NaN
      if (!(mx >= FLT_MIN)) mx = R(0);
      const R inv = mx > R(0) ? R(1) / mx : R(0);
      if constexpr (NT) __builtin_nontemporal_store((float)mx, (float *)((char *)norm + s * 4u));
      else st_plane(norm, s * 4u, (float)mx);
#pragma unroll
      for (int g = 0; g < 3; g++) {
        u32x4 q;
#pragma unroll
        for (int i = 0; i < 4; i++)
          q[i] = as_u32(from_cx<T>(Cx<R>{v[4 * g + i].re * inv, v[4 * g + i].im * inv}));
        u32x4 *d = (u32x4 *)((char *)p[0] + g * 4 * kTile * 4 + off);
        if constexpr (NT) __builtin_nontemporal_store(q, d);
        else *d = q;
      }
    } else if constexpr (P == QCD_HIP_F32) {
#pragma unroll
      for (int g = 0; g < 6; g++) {
        typedef float v4f __attribute__((ext_vector_type(4)));
        const v4f q{(float)v[2 * g].re, (float)v[2 * g].im, (float)v[2 * g + 1].re, (float)v[2 * g + 1].im};
        v4f *d = (v4f *)((char *)p[g / 4] + (g % 4) * 2 * kTile * 8 + off);
        if constexpr (NT) __builtin_nontemporal_store(q, d);
        else *d = q;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 12; k++) {
        typedef double v2d __attribute__((ext_vector_type(2)));
        const v2d q{(double)v[k].re, (double)v[k].im};
        v2d *d = (v2d *)((char *)comp(k) + off);
        if constexpr (NT) __builtin_nontemporal_store(q, d);
        else *d = q;
      }
    }
  }
};

template <int P>
struct GaugeIn {
  using T = typename Prec<P>::T;
  static constexpr int kPer = comp_per_base<T>(), kNb = num_bases<T, 9>();
  const T *p[4 * kNb];   // link component kPer*i of direction mu, tile 0, at p[kNb mu + i]
  static GaugeIn of(qcd_hip_field f) {
    GaugeIn g;
    const size_t st = (size_t)f->lat->stride;
    for (int mu = 0; mu < 4; mu++)
      for (int i = 0; i < kNb; i++) g.p[kNb * mu + i] = (const T *)f->data + (size_t)mu * 9 * st + i * kPer * kTile;
    return g;
  }
  __device__ __forceinline__ const T *comp(int mu, int k) const { return p[kNb * mu + k / kPer] + (k % kPer) * kTile; }
  // stored values of the link of direction mu at site s (NT: streaming
  // cache hint, FP32/FP64)
  template <bool NT = false>
  __device__ __forceinline__ void load_raw(int mu, unsigned s, T r[9]) const {
    if constexpr (P == QCD_HIP_F16) {
      const unsigned off = ((s / kTile) * (9u * kTile) + (s % kTile) * 3u) * 4u;
#pragma unroll
      for (int g = 0; g < 3; g++) {
        const U32x3 q = *(const U32x3 *)((const char *)p[kNb * mu] + g * 3 * kTile * 4 + off);
        r[3 * g] = as_half2(q.x);
        r[3 * g + 1] = as_half2(q.y);
        r[3 * g + 2] = as_half2(q.z);
      }
    } else {
      const unsigned off = ((s / kTile) * (9u * kTile) + (s % kTile)) * (unsigned)sizeof(T);
#pragma unroll
      for (int k = 0; k < 9; k++) r[k] = NT ? ld_plane_nt(comp(mu, k), off) : ld_plane(comp(mu, k), off);
    }
  }
  template <typename R>
  __device__ __forceinline__ void load(int mu, unsigned s, Cx<R> u[9]) const {
    T r[9];
    load_raw(mu, s, r);
#pragma unroll
    for (int k = 0; k < 9; k++) u[k] = to_cx<R>(r[k]);
  }
};

// ------------------------------------------------------------ geometry
// Site traversal: how workgroups map onto lattice sites.
//  MAP_XCD   : logical block b covers sites [256 b, 256 b + 256); logical
//              blocks are spread so each XCD owns one contiguous range.
//  MAP_SLAB  : XCD k owns the y-slab [y0[k], y0[k+1]) and its workgroups sweep
//              it in (t, x, y, z) order, so the z-neighbour of a site was
//              touched T*X*(y1-y0) sites earlier by the same XCD (same L2)
//              instead of T*X*Y sites earlier.
//  MAP_ZMARCH: each lane owns one (t, x, y) column and steps through zlen
//              consecutive z values, carrying the centre spinor and the
//              backward z hop in registers (2.5D blocking).
//  MAP_ZTILE : z-march of a tile of whole t-rows: a workgroup owns xb
//              consecutive x-rows (all t) of one y and stages them, with one
//              halo row on each side, in LDS every step, so t and x
//              neighbours never go through the vector memory path.
//  MAP_ZMARCH_LDS: MAP_ZMARCH whose lanes keep the forward z neighbour,
//              loaded for one step's z hop, in an LDS slot of their own and
//              read it back as the next step's centre spinor instead of
//              loading it again through L2.
//  MAP_ZMARCH2: MAP_ZMARCH for batches (FP32 and FP16): each lane works on
//              two right-hand sides of its column and loads every link once
//              for both.
//  MAP_MMA   : batches of FP16 spinors and links on the matrix cores: a
//              wavefront owns 16 sites x 4 right-hand sides and applies each
//              link to the projected neighbour spinors of its 4 rhs with
//              MFMA (mma.hip).
enum { MAP_XCD = 0, MAP_SLAB = 1, MAP_ZMARCH = 2, MAP_ZTILE = 3, MAP_ZMARCH_LDS = 4, MAP_ZMARCH2 = 5, MAP_MMA = 6 };

struct DslashArgs {
  unsigned V;
  unsigned L[4];
  size_t stride;
  double ph_re[4], ph_im[4];
  int twisted[4];
  double mp4, tm;
  int map;
  unsigned slab_y0[kNumXcd + 1];   // MAP_SLAB: y range of each XCD
  unsigned slab_blocks;            // MAP_SLAB: workgroups per XCD
  unsigned zlen;                   // MAP_ZMARCH: z steps per lane
  unsigned col_blocks;             // MAP_ZMARCH, MAP_ZTILE: workgroups per z chunk
  unsigned xb, nxg, lanes;         // MAP_ZTILE: x-rows per tile, tiles per x line, lanes per workgroup
  // right-hand-side batch: grid row y works on slot = ids ? ids[y] : y, i.e.
  // on vectors in + slot*in_rs, out + slot*out_rs, b + slot*b_rs
  const int *ids;
  unsigned in_rs, out_rs, b_rs;
  unsigned count;                  // batch rows (z-march: interleaved into grid.x, else grid.y)
  unsigned nblk;                   // workgroups per batch row
  unsigned pairs;                  // MAP_ZMARCH2: lanes carry batch rows 2p and 2p+1; pairs = ceil(count / 2)
  unsigned rgroup;                 // z-march batches: rows per group (block_row); 0 = the whole batch
};

template <int PS>
struct RhsView {
  SpinorIn<PS> in, b;
  SpinorOut<PS> out;
};

template <int PS>
__device__ __forceinline__ RhsView<PS> rhs_view(const DslashArgs &a, const SpinorIn<PS> &in, const SpinorOut<PS> &out,
                                                const SpinorIn<PS> &b, unsigned row) {
  const unsigned slot = a.ids ? (unsigned)a.ids[row] : row;
  RhsView<PS> v;
  v.in = in.advanced(slot * a.in_rs);
  v.out = out.advanced(slot * a.out_rs);
  v.b = b.advanced(slot * a.b_rs);
  return v;
}

// Block decomposition of the z-march traversals: one 1D grid of nblk * count
// workgroups, remapped per XCD, with the right-hand side the fastest index,
// so the count right-hand sides of one tile run together on one XCD and share
// its links through that XCD's L2.  row = batch row, blk = tile index.
struct BlockRow {
  unsigned row, blk;
};
__device__ __forceinline__ BlockRow block_row(const DslashArgs &a) {
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const unsigned R = a.rgroup;
  if (R == 0 || R >= a.count) return BlockRow{lb % a.count, lb / a.count};
  // rows in groups of R, the row fastest within a group: a group's site
  // blocks share links through one L2 while enough site blocks of a group
  // are in flight to share their neighbour planes too
  const unsigned nb = a.nblk, full = a.count / R, span = nb * R;
  if (lb < full * span) {
    const unsigned g = lb / span, w = lb - g * span;
    return BlockRow{g * R + w % R, w / R};
  }
  const unsigned rem = a.count - full * R, w = lb - full * span;
  return BlockRow{full * R + w % rem, w / rem};
}

struct Site {
  unsigned s, t, x, y, z;
  bool valid;
};

__device__ __forceinline__ Site site_coords(unsigned s, const DslashArgs &a) {
  Site st;
  st.s = s;
  unsigned r = s / a.L[0];
  st.t = s - r * a.L[0];
  unsigned r2 = r / a.L[1];
  st.x = r - r2 * a.L[1];
  st.z = r2 / a.L[2];
  st.y = r2 - st.z * a.L[2];
  return st;
}

// The site handled by this lane under the traversal a.map.
__device__ __forceinline__ Site site_of_lane(const DslashArgs &a) {
  Site st;
  if (a.map == MAP_SLAB) {
    const unsigned k = blockIdx.x % kNumXcd, local = blockIdx.x / kNumXcd;
    const unsigned y0 = a.slab_y0[k], ny = a.slab_y0[k + 1] - y0;
    const unsigned P = a.L[0] * a.L[1];
    const unsigned j = local * kBlock + threadIdx.x;
    st.valid = ny > 0 && j < P * ny * a.L[3];
    if (!st.valid) {
      st.s = st.t = st.x = st.y = st.z = 0;
      return st;
    }
    const unsigned tx = j % P, r = j / P;
    const unsigned yy = r % ny;
    st.z = r / ny;
    st.y = y0 + yy;
    st.x = tx / a.L[0];
    st.t = tx - st.x * a.L[0];
    st.s = tx + P * (st.y + a.L[2] * st.z);
    return st;
  }
  const unsigned s = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  st.valid = s < a.V;
  if (!st.valid) {
    st.s = st.t = st.x = st.y = st.z = 0;
    return st;
  }
  st = site_coords(s, a);
  st.valid = true;
  return st;
}

// neighbour of st in direction MU: forward (+1) or backward (-1)
template <int MU, int DIR>
__device__ __forceinline__ unsigned neighbour(const Site &st, const DslashArgs &a) {
  unsigned c, len, step;
  if constexpr (MU == 0) {
    c = st.t; len = a.L[0]; step = 1;
  } else if constexpr (MU == 1) {
    c = st.x; len = a.L[1]; step = a.L[0];
  } else if constexpr (MU == 2) {
    c = st.y; len = a.L[2]; step = a.L[0] * a.L[1];
  } else {
    c = st.z; len = a.L[3]; step = a.L[0] * a.L[1] * a.L[2];
  }
  if constexpr (DIR > 0) return (c == len - 1) ? st.s - (len - 1) * step : st.s + step;
  else return (c == 0) ? st.s + (len - 1) * step : st.s - step;
}

// ------------------------------------------------------- projector table
// Chiral basis of lib/src/qcd_gamma.c:46-88.  (1 + S g_mu) psi has upper spins
//   h0 = psi_0 + S i^A psi_P,  h1 = psi_1 + S i^B psi_Q
// and lower spins  psi_2 = S i^C h_R,  psi_3 = S i^D h_W  (qcd_wilson.h:17-387).
template <int MU>
struct Proj;
template <>
struct Proj<0> { static constexpr int A = 2, B = 2, C = 2, D = 2, P = 2, Q = 3, R = 0, W = 1; };
template <>
struct Proj<1> { static constexpr int A = 3, B = 3, C = 1, D = 1, P = 3, Q = 2, R = 1, W = 0; };
template <>
struct Proj<2> { static constexpr int A = 2, B = 0, C = 0, D = 2, P = 3, Q = 2, R = 1, W = 0; };
template <>
struct Proj<3> { static constexpr int A = 3, B = 1, C = 1, D = 3, P = 2, Q = 3, R = 0, W = 1; };

template <int POW, typename R>
__device__ __forceinline__ Cx<R> ipow(Cx<R> v) {
  constexpr int p = POW & 3;
  if constexpr (p == 0) return v;
  else if constexpr (p == 1) return Cx<R>{-v.im, v.re};
  else if constexpr (p == 2) return Cx<R>{-v.re, -v.im};
  else return Cx<R>{v.im, -v.re};
}

template <typename R>
__device__ __forceinline__ Cx<R> cadd(Cx<R> a, Cx<R> b) { return Cx<R>{a.re + b.re, a.im + b.im}; }

// y = U x (DAG = false) or U^+ x (DAG = true) for one colour vector.  Every
// complex multiply-add is written as explicit fused multiply-adds so that all
// kernels (whatever the surrounding code) round identically.
template <bool DAG, typename R>
__device__ __forceinline__ void su3_mul(Cx<R> y[3], const Cx<R> u[9], const Cx<R> x[3]) {
#pragma unroll
  for (int r = 0; r < 3; r++) {
    R re, im;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const Cx<R> m = DAG ? Cx<R>{u[3 * k + r].re, -u[3 * k + r].im} : u[3 * r + k];
      if (k == 0) {
        re = m.re * x[k].re;
        im = m.re * x[k].im;
      } else {
        re = fma(m.re, x[k].re, re);
        im = fma(m.re, x[k].im, im);
      }
      re = fma(-m.im, x[k].im, re);
      im = fma(m.im, x[k].re, im);
    }
    y[r] = Cx<R>{re, im};
  }
}

// v * ph
template <typename R>
__device__ __forceinline__ Cx<R> cmul_ph(Cx<R> v, Cx<R> ph) {
  return Cx<R>{fma(v.re, ph.re, -(v.im * ph.im)), fma(v.re, ph.im, v.im * ph.re)};
}

// -2 (m + 4 + i g tm) v: the diagonal term of one spin component (g = +1 for
// spins 0, 1 and -1 for spins 2, 3)
template <typename R>
__device__ __forceinline__ Cx<R> diag_term(Cx<R> v, R mp4, R gtm) {
  return Cx<R>{R(-2) * fma(mp4, v.re, -(gtm * v.im)), R(-2) * fma(mp4, v.im, gtm * v.re)};
}

// the colour-multiplied, phased half spinor g of one hop (before rebuilding)
template <int MU, int S, bool DAG, typename R>
__device__ __forceinline__ void hop_half(Cx<R> g0[3], Cx<R> g1[3], const Cx<R> psi[12], const Cx<R> u[9], Cx<R> ph) {
  using PJ = Proj<MU>;
  constexpr int sp = S > 0 ? 0 : 2;
  Cx<R> h0[3], h1[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    h0[c] = cadd(psi[c], ipow<PJ::A + sp>(psi[3 * PJ::P + c]));
    h1[c] = cadd(psi[3 + c], ipow<PJ::B + sp>(psi[3 * PJ::Q + c]));
  }
  su3_mul<DAG>(g0, u, h0);
  su3_mul<DAG>(g1, u, h1);
  // Unconditional: an untwisted direction carries the phase (1, 0), which
  // multiplies exactly and avoids a divergent-looking branch that doubles
  // the live registers.
#pragma unroll
  for (int c = 0; c < 3; c++) {
    g0[c] = cmul_ph(g0[c], ph);
    g1[c] = cmul_ph(g1[c], ph);
  }
}

// ------------------------------------------------------------ FP16 hops
// With FP16 spinors and links a hop works on the stored values: the spin
// projection is packed FP16 on the unit-normalised components (|h| <= 2), the
// colour product uses v_dot2_f32_f16 (two FP16 products added to an FP32
// accumulator), and the site norm is folded into the phase factor that
// multiplies the six colour products.  About half the vector ALU work of
// converting the 12 components and 9 links to FP32 first.
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ h2v as_h2v(__half2 a) { return __builtin_bit_cast(h2v, a); }

template <int POW>
__device__ __forceinline__ h2v ipow_h(h2v v) {
  constexpr int p = POW & 3;
  if constexpr (p == 0) return v;
  else if constexpr (p == 1) return h2v{-v.y, v.x};
  else if constexpr (p == 2) return -v;
  else return h2v{v.y, -v.x};
}

// y = U x (DAG = false) or U^+ x (DAG = true) with v_dot2_f32_f16 (two FP16
// products added to an FP32 accumulator) on (re, im) pairs:
//   U x:   re = u . (x.re, -x.im),  im = u . (x.im, x.re)
//   U^+ x: re = u . (x.re,  x.im),  im = u . (x.im, -x.re)
// The negated high half is the instruction's neg_hi source modifier (the
// compiler would build such operands with xors); the swapped pair xs takes
// one instruction per colour (the instruction has no op_sel on gfx950).
// The 18 products form one asm block so their order is fixed: the six sums
// advance together, each accumulator read (as the addend, which needs no
// wait states) six instructions after it was written.  A DOT result read by
// any other VALU instruction needs 3 wait states on CDNA3/4, which the
// compiler cannot insert for an asm block, hence the closing s_nop 2.
template <bool DAG>
__device__ __forceinline__ void su3_mul_h(Cx<float> y[3], const __half2 u[9], const h2v x[3]) {
  h2v m[9], xs[3];
#pragma unroll
  for (int k = 0; k < 3; k++) xs[k] = h2v{x[k].y, x[k].x};
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int k = 0; k < 3; k++) m[3 * r + k] = as_h2v(DAG ? u[3 * k + r] : u[3 * r + k]);
  float o[6];
#define QH_DOT_OPERANDS                                                                                       \
  : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(o[4]), "=&v"(o[5])                            \
  : "v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]), "v"(m[7]), "v"(m[8]),       \
    "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(xs[0]), "v"(xs[1]), "v"(xs[2])
  if constexpr (DAG) {
    asm(
        "v_dot2_f32_f16 %0, %6, %15, 0\n\t"
        "v_dot2_f32_f16 %1, %6, %18, 0 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %2, %9, %15, 0\n\t"
        "v_dot2_f32_f16 %3, %9, %18, 0 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %4, %12, %15, 0\n\t"
        "v_dot2_f32_f16 %5, %12, %18, 0 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %0, %7, %16, %0\n\t"
        "v_dot2_f32_f16 %1, %7, %19, %1 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %2, %10, %16, %2\n\t"
        "v_dot2_f32_f16 %3, %10, %19, %3 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %4, %13, %16, %4\n\t"
        "v_dot2_f32_f16 %5, %13, %19, %5 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %0, %8, %17, %0\n\t"
        "v_dot2_f32_f16 %1, %8, %20, %1 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %2, %11, %17, %2\n\t"
        "v_dot2_f32_f16 %3, %11, %20, %3 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %4, %14, %17, %4\n\t"
        "v_dot2_f32_f16 %5, %14, %20, %5 neg_hi:[0,1,0]\n\t"
        "s_nop 2"
        QH_DOT_OPERANDS);
  } else {
    asm(
        "v_dot2_f32_f16 %0, %6, %15, 0 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %1, %6, %18, 0\n\t"
        "v_dot2_f32_f16 %2, %9, %15, 0 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %3, %9, %18, 0\n\t"
        "v_dot2_f32_f16 %4, %12, %15, 0 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %5, %12, %18, 0\n\t"
        "v_dot2_f32_f16 %0, %7, %16, %0 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %1, %7, %19, %1\n\t"
        "v_dot2_f32_f16 %2, %10, %16, %2 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %3, %10, %19, %3\n\t"
        "v_dot2_f32_f16 %4, %13, %16, %4 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %5, %13, %19, %5\n\t"
        "v_dot2_f32_f16 %0, %8, %17, %0 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %1, %8, %20, %1\n\t"
        "v_dot2_f32_f16 %2, %11, %17, %2 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %3, %11, %20, %3\n\t"
        "v_dot2_f32_f16 %4, %14, %17, %4 neg_hi:[0,1,0]\n\t"
        "v_dot2_f32_f16 %5, %14, %20, %5\n\t"
        "s_nop 2"
        QH_DOT_OPERANDS);
  }
#undef QH_DOT_OPERANDS
#pragma unroll
  for (int r = 0; r < 3; r++) y[r] = Cx<float>{o[2 * r], o[2 * r + 1]};
}

// hop_half from stored values: spinor components r with site norm n (FP16
// storage; n is unused otherwise) and link components u
template <int MU, int S, bool DAG, int PS, int PG, typename R>
__device__ __forceinline__ void hop_half_raw(Cx<R> g0[3], Cx<R> g1[3], const typename Prec<PS>::T r[12], float n,
                                             const typename Prec<PG>::T u[9], Cx<R> ph) {
  if constexpr (PS == QCD_HIP_F16 && PG == QCD_HIP_F16) {
    using PJ = Proj<MU>;
    constexpr int sp = S > 0 ? 0 : 2;
    h2v h0[3], h1[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      h0[c] = as_h2v(r[c]) + ipow_h<PJ::A + sp>(as_h2v(r[3 * PJ::P + c]));
      h1[c] = as_h2v(r[3 + c]) + ipow_h<PJ::B + sp>(as_h2v(r[3 * PJ::Q + c]));
    }
    su3_mul_h<DAG>(g0, u, h0);
    su3_mul_h<DAG>(g1, u, h1);
    const Cx<R> f{n * ph.re, n * ph.im};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      g0[c] = cmul_ph(g0[c], f);
      g1[c] = cmul_ph(g1[c], f);
    }
  } else {
    Cx<R> psi[12], uc[9];
    SpinorIn<PS>::convert(r, n, psi);
#pragma unroll
    for (int k = 0; k < 9; k++) uc[k] = to_cx<R>(u[k]);
    hop_half<MU, S, DAG>(g0, g1, psi, uc, ph);
  }
}

// add a half spinor of (1 + S g_MU) rebuilt to four spins
template <int MU, int S, typename R>
__device__ __forceinline__ void rebuild_add(Cx<R> acc[4][3], const Cx<R> g0[3], const Cx<R> g1[3]) {
  using PJ = Proj<MU>;
  constexpr int sp = S > 0 ? 0 : 2;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    acc[0][c] = cadd(acc[0][c], g0[c]);
    acc[1][c] = cadd(acc[1][c], g1[c]);
    acc[2][c] = cadd(acc[2][c], ipow<PJ::C + sp>(PJ::R == 0 ? g0[c] : g1[c]));
    acc[3][c] = cadd(acc[3][c], ipow<PJ::D + sp>(PJ::W == 0 ? g0[c] : g1[c]));
  }
}

// hop_half of the spinor at site sn of `in` with the link of direction MU at sl
template <int MU, int S, bool DAG, int PS, int PG, typename R>
__device__ __forceinline__ void hop_half_at(Cx<R> g0[3], Cx<R> g1[3], const SpinorIn<PS> &in, unsigned sn,
                                            const GaugeIn<PG> &U, unsigned sl, Cx<R> ph) {
  typename Prec<PS>::T r[12];
  typename Prec<PG>::T u[9];
  float n = 0.f;
  in.load_raw(sn, r, n);
  U.load_raw(MU, sl, u);
  hop_half_raw<MU, S, DAG, PS, PG>(g0, g1, r, n, u, ph);
}

// the same hop added to acc (rebuilt to four spins)
template <int MU, int S, bool DAG, int PS, int PG, typename R>
__device__ __forceinline__ void hop_acc_at(Cx<R> acc[4][3], const SpinorIn<PS> &in, unsigned sn, const GaugeIn<PG> &U,
                                           unsigned sl, Cx<R> ph) {
  Cx<R> g0[3], g1[3];
  hop_half_at<MU, S, DAG>(g0, g1, in, sn, U, sl, ph);
  rebuild_add<MU, S>(acc, g0, g1);
}

// forward hop U_mu(x)(1-g_mu)psi(x+mu) e^{+i th/L} and backward hop
// U_mu(x-mu)^+ (1+g_mu) psi(x-mu) e^{-i th/L}
// FENCE: keep the loads of this direction from being hoisted above the
// previous one (bounds register pressure at the cost of memory parallelism).
template <int MU, int PS, int PG, typename R, bool FENCE = false>
__device__ __forceinline__ void hop_dir(Cx<R> acc[4][3], const DslashArgs &a, const Site &st, const SpinorIn<PS> &in,
                                        const GaugeIn<PG> &U) {
  if constexpr (FENCE) __builtin_amdgcn_sched_barrier(0);
  const Cx<R> php{(R)a.ph_re[MU], (R)a.ph_im[MU]};
  const Cx<R> phm{(R)a.ph_re[MU], -(R)a.ph_im[MU]};
  hop_acc_at<MU, -1, false>(acc, in, neighbour<MU, +1>(st, a), U, st.s, php);
  const unsigned sb = neighbour<MU, -1>(st, a);
  hop_acc_at<MU, +1, true>(acc, in, sb, U, sb, phm);
}

// ------------------------------------------------------------ epilogues
// EPI_RES_D_EO16: the residual b - D x, stored not as one FP64 field but split
// by parity into the FP16 half fields of an even-odd solve's inner
// precision, with M^-1 of the even half as well (eo_split_store)
enum { EPI_D = 0, EPI_Q = 1, EPI_RES_D = 2, EPI_RES_Q = 3, EPI_RES_D_EO16 = 4 };
template <int EPI>
constexpr bool epi_eo_split() {
  return EPI == EPI_RES_D_EO16;
}

template <int EPI, int PS, typename R>
__device__ __forceinline__ void apply_epilogue(Cx<R> o[12], const SpinorIn<PS> &b, unsigned s) {
  if constexpr (EPI == EPI_Q || EPI == EPI_RES_Q) {
#pragma unroll
    for (int k = 0; k < 6; k++) o[k] = ipow<1>(o[k]);
#pragma unroll
    for (int k = 6; k < 12; k++) o[k] = ipow<3>(o[k]);
  }
  if constexpr (EPI == EPI_RES_D || EPI == EPI_RES_Q || epi_eo_split<EPI>()) {
    // b is read once: a streaming load leaves L2 to the neighbour planes
    Cx<R> bv[12];
    b.template load<true>(s, bv);
#pragma unroll
    for (int k = 0; k < 12; k++) o[k] = Cx<R>{bv[k].re - o[k].re, bv[k].im - o[k].im};
  }
}

// The split targets of EPI_RES_D_EO16 (kernel argument; NoSplit for every
// other epilogue): row slot b's residual at site s goes to entry s >> 1
// of vector b of ev (even s) or od (odd s) in precision P, and for even s
// M^-1 of it, M = m + 4 + i tm g5, to tev -- the even-odd split of an outer
// step, from the FP64 values with the same arithmetic as a split kernel that
// reads them back.
struct NoSplit {};
template <int P>
struct EoSplit {
  SpinorOut<P> ev, od, tev;
  double mi_re, mi_im;   // 1 / (m + 4 + i tm)
};
template <int EPI>
struct SplitOf {
  using type = NoSplit;
};
template <>
struct SplitOf<EPI_RES_D_EO16> {
  using type = EoSplit<QCD_HIP_F16>;
};

// the split targets of batch vector slot (the row's views, formed once)
template <int P>
__device__ __forceinline__ EoSplit<P> split_row(const EoSplit<P> &sp, unsigned slot) {
  return EoSplit<P>{sp.ev.advanced(slot), sp.od.advanced(slot), sp.tev.advanced(slot), sp.mi_re, sp.mi_im};
}
__device__ __forceinline__ NoSplit split_row(const NoSplit &sp, unsigned) { return sp; }

// sp: the row's views (split_row)
template <int P>
__device__ __forceinline__ void eo_split_store(const EoSplit<P> &sp, unsigned s, unsigned par, const Cx<double> o[12]) {
  using R = typename Prec<P>::R;
  Cx<R> v[12];
  if (par) {
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = Cx<R>{(R)o[k].re, (R)o[k].im};
    sp.od.template store<true>(s >> 1, v);
  } else {
    const double are = sp.mi_re, aim = sp.mi_im;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      const double im = k < 6 ? aim : -aim;   // g5 = diag(1, 1, -1, -1)
      v[k] = Cx<R>{(R)fma(are, o[k].re, -(im * o[k].im)), (R)fma(are, o[k].im, im * o[k].re)};
    }
    sp.tev.template store<true>(s >> 1, v);
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = Cx<R>{(R)o[k].re, (R)o[k].im};
    sp.ev.template store<true>(s >> 1, v);
  }
}

// the residual of a site stored as the epilogue asks: out, or the even-odd
// split (PS = FP64 only; split: the row's views)
template <int EPI, int PS, typename R>
__device__ __forceinline__ void store_site(const SpinorOut<PS> &out, const typename SplitOf<EPI>::type &split,
                                           const Site &st, const Cx<R> o[12]) {
  if constexpr (epi_eo_split<EPI>()) {
    static_assert(PS == QCD_HIP_F64, "the even-odd split reads the FP64 residual");
    // keep the next z step's loads below the split's conversions and stores
    __builtin_amdgcn_sched_barrier(0);
    eo_split_store(split, st.s, (st.t + st.x + st.y + st.z) & 1u, o);
    __builtin_amdgcn_sched_barrier(0);
  } else {
    out.template store<true>(st.s, o);
  }
}

// ------------------------------------------------------------ reduction
// Sum v over the workgroup (64-wide wavefront shuffles, then LDS) and store
// it at partials[slot].  Every lane of the workgroup must call this.
__device__ __forceinline__ void block_reduce_store(double v, double *partials, unsigned slot) {
  __shared__ double wsum[kBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wsum[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    const int nw = (blockDim.x + 63) / 64;
    for (int w = 0; w < nw; w++) t += wsum[w];
    partials[slot] = t;
  }
}

// matrix-core batch launch (mma.hip): FP16 spinors and links, D or Q
// epilogue, batch rows as in DslashArgs
int dslash_mma_launch(int epi, DslashArgs a, qcd_hip_field out, int ro, qcd_hip_field in, int ri, qcd_hip_field g,
                      qcd_hip_field b, int rb, hipStream_t st);

// launchers defined in dslash.hip
// Batched launch over `count` right-hand sides (see DslashArgs::ids): row y
// applies the operator to vector ri + slot*in_rs of `in`, writes vector
// ro + slot*out_rs of `out`, reads vector rb + slot*b_rs of `b` (residual
// epilogues); with red, |out|^2 partials go to partials[y * grid + x].
int dslash_launch_batch(int epi, bool red, qcd_hip_field out, int ro, int out_rs, qcd_hip_field in, int ri, int in_rs,
                        qcd_hip_field g, qcd_hip_field b, int rb, int b_rs, const int *ids, int count, double m,
                        double tm, double *partials, hipStream_t stream, int *grid_out);
// r = b - D x of a batch of FP64 fields with |r|^2 partials, stored split by
// parity into FP16 half fields ev, od and tev = M^-1 ev (vector ids[y] of each
// for batch row y), M = m + 4 + i tm g5
int dslash_residual_eo_split(qcd_hip_field x, int x0, int x_rs, qcd_hip_field g, qcd_hip_field b, int rb, int b_rs,
                             const int *ids, int count, double m, double tm, qcd_hip_field ev, qcd_hip_field od,
                             qcd_hip_field tev, double *partials, hipStream_t stream, int *grid_out);
DslashArgs make_args(qcd_hip_lattice lat, double m, double tm);
int dslash_check(qcd_hip_field out, qcd_hip_field in, qcd_hip_field g);
// asynchronous single-rhs launch; *grid_out (if given) receives the number of
// workgroups, i.e. of |out|^2 partials written when red is set
int dslash_launch(int epi, bool red, qcd_hip_field out, int ro, qcd_hip_field in, int ri, qcd_hip_field g,
                  qcd_hip_field b, int rb, double m, double tm, double *partials, hipStream_t stream, int *grid_out);
// autotune the configuration for (in, g) once (call outside graph capture)
int dslash_prepare(qcd_hip_field out, qcd_hip_field in, qcd_hip_field g);
// number of workgroups dslash_launch will use for (in, g)
int dslash_grid(qcd_hip_field in, qcd_hip_field g);
// upper bound of that number over every launch configuration
int dslash_grid_bound(qcd_hip_lattice lat);

}  // namespace qcdhip
