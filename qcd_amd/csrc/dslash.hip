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
#include <array>
#include <vector>

#include "dslash_common.h"

namespace qcdhip {

// OCC & 15: 0 = compiler's register allocation, W > 0 = at least W waves per
// SIMD; OCC & 16: scheduling fence between hop directions.
#define QH_WAVES(OCC) __attribute__((amdgpu_waves_per_eu(((OCC) & 15) > 0 ? ((OCC) & 15) : 1, ((OCC) & 15) > 0 ? ((OCC) & 15) : 8)))

// compile-time (direction, projector sign, dagger) of one hop
template <int MU, int S, bool DAG>
struct HopTag {
  static constexpr int mu = MU, s = S;
  static constexpr bool dag = DAG;
};

// acc = -2 (m + 4 + i tm g5) psi + the backward z hop (cz0, cz1) carried
// from the previous z step; psi given by its stored values rc (norm nc)
template <int PS, typename R>
__device__ __forceinline__ void local_terms(Cx<R> acc[4][3], const typename Prec<PS>::T rc[12], float nc,
                                            const Cx<R> cz0[3], const Cx<R> cz1[3], R mp4, R tm) {
  Cx<R> psi[12];
  SpinorIn<PS>::convert(rc, nc, psi);
#pragma unroll
  for (int sp = 0; sp < 4; sp++) {
    const R gg = sp < 2 ? tm : -tm;
#pragma unroll
    for (int c = 0; c < 3; c++) acc[sp][c] = diag_term(psi[3 * sp + c], mp4, gg);
  }
  rebuild_add<3, +1>(acc, cz0, cz1);
}

template <int PS, int PG, int EPI, bool RED, int OCC>
__global__ void __launch_bounds__(kBlock) QH_WAVES(OCC)
k_dslash(DslashArgs a, SpinorIn<PS> in_all, GaugeIn<PG> U, SpinorOut<PS> out_all, SpinorIn<PS> b_all,
         double *__restrict__ partials, typename SplitOf<EPI>::type split) {
  using R = typename Prec<PS>::R;
  const Site st = site_of_lane(a);
  const unsigned s = st.s;
  const RhsView<PS> rv = rhs_view(a, in_all, out_all, b_all, blockIdx.y);
  const SpinorIn<PS> &in = rv.in, &b = rv.b;
  const SpinorOut<PS> &out = rv.out;
  double red = 0.0;
  if (st.valid) {
    // same summation order as k_dslash_zm, so every traversal gives the same
    // bits: acc = -2 diag(psi) + z back + z forward + t, x, y hops; out = -acc/2
    Cx<R> acc[4][3];
    const R mp4 = R(a.mp4), tm = R(a.tm);
    {
      Cx<R> psi[12];
      in.load(s, psi);
#pragma unroll
      for (int sp = 0; sp < 4; sp++) {
        const R gg = sp < 2 ? tm : -tm;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const Cx<R> v = psi[3 * sp + c];
          acc[sp][c] = diag_term(v, mp4, gg);
        }
      }
      const unsigned sb = neighbour<3, -1>(st, a), sf = neighbour<3, +1>(st, a);
      hop_acc_at<3, +1, true>(acc, in, sb, U, sb, Cx<R>{(R)a.ph_re[3], -(R)a.ph_im[3]});
      hop_acc_at<3, -1, false>(acc, in, sf, U, s, Cx<R>{(R)a.ph_re[3], (R)a.ph_im[3]});
    }
    constexpr bool F = (OCC & 16) != 0;
    hop_dir<0, PS, PG, R, F>(acc, a, st, in, U);
    hop_dir<1, PS, PG, R, F>(acc, a, st, in, U);
    hop_dir<2, PS, PG, R, F>(acc, a, st, in, U);
    if constexpr (F) __builtin_amdgcn_sched_barrier(0);
    Cx<R> o[12];
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int c = 0; c < 3; c++) o[3 * sp + c] = Cx<R>{R(-0.5) * acc[sp][c].re, R(-0.5) * acc[sp][c].im};
    apply_epilogue<EPI>(o, b, s);
    if (RED) {
#pragma unroll
      for (int k = 0; k < 12; k++) red += (double)o[k].re * (double)o[k].re + (double)o[k].im * (double)o[k].im;
    }
    store_site<EPI, PS>(out, split_row(split, a.ids ? (unsigned)a.ids[blockIdx.y] : blockIdx.y), st, o);
  }
  if (RED) block_reduce_store(red, partials, blockIdx.y * gridDim.x + blockIdx.x);
}

// z-marching Dslash: lane = one (t, x, y) column, steps z0 .. z0+zlen-1.
// Only the backward z hop U_z(x-z)^+ (1+g_z) psi(x-z) (a colour-multiplied
// half spinor, computed one step earlier) is carried between steps; the
// centre spinor is reloaded (it was this lane's forward z neighbour one step
// before, so it comes from cache).  Each step starts with the local terms
// (mass, both z hops) so the centre spinor dies before the t/x/y hops load.
// acc holds -2 diag(psi) + sum of hops, so out = -1/2 acc.
// LDSC (MAP_ZMARCH_LDS): the forward z neighbour's stored values go to this
// lane's LDS slot and come back as the next step's centre (same bits).
template <int PS, int PG, int EPI, bool RED, int OCC, bool LDSC = false>
__global__ void __launch_bounds__(kBlock) QH_WAVES(OCC)
k_dslash_zm(DslashArgs a, SpinorIn<PS> in_all, GaugeIn<PG> U, SpinorOut<PS> out_all, SpinorIn<PS> b_all,
            double *__restrict__ partials, typename SplitOf<EPI>::type split) {
  using R = typename Prec<PS>::R;
  using T = typename Prec<PS>::T;
  __shared__ T zfwd[LDSC ? 12 : 1][kBlock];
  __shared__ float zfwd_n[LDSC && PS == QCD_HIP_F16 ? kBlock : 1];
  const BlockRow br = block_row(a);
  const RhsView<PS> rv = rhs_view(a, in_all, out_all, b_all, br.row);
  const SpinorIn<PS> &in = rv.in, &b = rv.b;
  const SpinorOut<PS> &out = rv.out;
  const auto srow = split_row(split, a.ids ? (unsigned)a.ids[br.row] : br.row);
  const unsigned lb = br.blk;
  const unsigned cb = lb % a.col_blocks, zc = lb / a.col_blocks;
  const unsigned P3 = a.L[0] * a.L[1] * a.L[2];
  const unsigned col = cb * kBlock + threadIdx.x;
  const unsigned z0 = zc * a.zlen;
  const unsigned z1 = min(z0 + a.zlen, a.L[3]);
  double red = 0.0;
  if (col < P3 && z0 < a.L[3]) {
    Site st;
    st.valid = true;
    st.x = col / a.L[0];
    st.t = col - st.x * a.L[0];
    st.y = st.x / a.L[1];
    st.x -= st.y * a.L[1];
    const Cx<R> php{(R)a.ph_re[3], (R)a.ph_im[3]};
    const Cx<R> phm{(R)a.ph_re[3], -(R)a.ph_im[3]};
    const R mp4 = R(a.mp4), tm = R(a.tm);
    Cx<R> cz0[3], cz1[3];
    {
      const unsigned zm = z0 == 0 ? a.L[3] - 1 : z0 - 1;
      const unsigned sm = col + P3 * zm;
      hop_half_at<3, +1, true>(cz0, cz1, in, sm, U, sm, phm);
    }
    for (unsigned z = z0; z < z1; z++) {
      st.z = z;
      st.s = col + P3 * z;
      const unsigned s = st.s;
      Cx<R> acc[4][3];
      {
        T rc[12];
        typename Prec<PG>::T u[9];
        float nc = 0.f;
        if (LDSC && z > z0) {
#pragma unroll
          for (int k = 0; k < 12; k++) rc[k] = zfwd[k][threadIdx.x];
          if constexpr (PS == QCD_HIP_F16) nc = zfwd_n[threadIdx.x];
        } else {
          in.load_raw(s, rc, nc);
        }
        local_terms<PS>(acc, rc, nc, cz0, cz1, mp4, tm);
        // U_z(s) is read by this lane only (the next step carries the hop it
        // feeds): a streaming load leaves L2 to the spinor planes
        // residual epilogues run over batch rows whose groups share links: a normal load
        U.template load_raw<EPI == EPI_D || EPI == EPI_Q>(3, s, u);
        // the next site's backward z hop: U_z(s)^+ (1 + g_z) psi(s)
        hop_half_raw<3, +1, true, PS, PG>(cz0, cz1, rc, nc, u, phm);
        const unsigned sf = z + 1 == a.L[3] ? col : s + P3;
        Cx<R> g0[3], g1[3];
        float nf = 0.f;
        in.load_raw(sf, rc, nf);
        if constexpr (LDSC) {
#pragma unroll
          for (int k = 0; k < 12; k++) zfwd[k][threadIdx.x] = rc[k];
          if constexpr (PS == QCD_HIP_F16) zfwd_n[threadIdx.x] = nf;
        }
        hop_half_raw<3, -1, false, PS, PG>(g0, g1, rc, nf, u, php);
        rebuild_add<3, -1>(acc, g0, g1);
      }
      constexpr bool F = (OCC & 16) != 0;
      if constexpr (F) __builtin_amdgcn_sched_barrier(0);
      hop_dir<0, PS, PG, R, F>(acc, a, st, in, U);
      hop_dir<1, PS, PG, R, F>(acc, a, st, in, U);
      hop_dir<2, PS, PG, R, F>(acc, a, st, in, U);
      if constexpr (F) __builtin_amdgcn_sched_barrier(0);
      Cx<R> o[12];
#pragma unroll
      for (int sp = 0; sp < 4; sp++)
#pragma unroll
        for (int c = 0; c < 3; c++) o[3 * sp + c] = Cx<R>{R(-0.5) * acc[sp][c].re, R(-0.5) * acc[sp][c].im};
      apply_epilogue<EPI>(o, b, s);
      if (RED) {
#pragma unroll
        for (int k = 0; k < 12; k++) red += (double)o[k].re * (double)o[k].re + (double)o[k].im * (double)o[k].im;
      }
      store_site<EPI, PS>(out, srow, st, o);
    }
  }
  if (RED) block_reduce_store(red, partials, br.row * a.nblk + br.blk);
}

// both hops of direction MU for two right-hand sides: each link is loaded
// once; per right-hand side the arithmetic is that of hop_dir
template <int MU, int PS, int PG, typename R>
__device__ __forceinline__ void hop_dir2(Cx<R> acc0[4][3], Cx<R> acc1[4][3], const DslashArgs &a, const Site &st,
                                         const SpinorIn<PS> &in0, const SpinorIn<PS> &in1, const GaugeIn<PG> &U) {
  const Cx<R> php{(R)a.ph_re[MU], (R)a.ph_im[MU]};
  const Cx<R> phm{(R)a.ph_re[MU], -(R)a.ph_im[MU]};
  typename Prec<PS>::T r[12];
  typename Prec<PG>::T u[9];
  Cx<R> g0[3], g1[3];
  float n = 0.f;
  {
    const unsigned sf = neighbour<MU, +1>(st, a);
    U.load_raw(MU, st.s, u);
    in0.load_raw(sf, r, n);
    hop_half_raw<MU, -1, false, PS, PG>(g0, g1, r, n, u, php);
    rebuild_add<MU, -1>(acc0, g0, g1);
    in1.load_raw(sf, r, n);
    hop_half_raw<MU, -1, false, PS, PG>(g0, g1, r, n, u, php);
    rebuild_add<MU, -1>(acc1, g0, g1);
  }
  {
    const unsigned sb = neighbour<MU, -1>(st, a);
    U.load_raw(MU, sb, u);
    in0.load_raw(sb, r, n);
    hop_half_raw<MU, +1, true, PS, PG>(g0, g1, r, n, u, phm);
    rebuild_add<MU, +1>(acc0, g0, g1);
    in1.load_raw(sb, r, n);
    hop_half_raw<MU, +1, true, PS, PG>(g0, g1, r, n, u, phm);
    rebuild_add<MU, +1>(acc1, g0, g1);
  }
}

// out = -1/2 acc with the D or Q epilogue, stored at site s
template <int EPI, int PS, typename R>
__device__ __forceinline__ void finish_site(const Cx<R> acc[4][3], const SpinorOut<PS> &out, const SpinorIn<PS> &b,
                                            unsigned s) {
  Cx<R> o[12];
#pragma unroll
  for (int sp = 0; sp < 4; sp++)
#pragma unroll
    for (int c = 0; c < 3; c++) o[3 * sp + c] = Cx<R>{R(-0.5) * acc[sp][c].re, R(-0.5) * acc[sp][c].im};
  apply_epilogue<EPI>(o, b, s);
  out.template store<true>(s, o);
}

// MAP_ZMARCH2: the z-march of k_dslash_zm with two right-hand sides per lane
// (batch rows 2p and 2p+1 of pair p; an odd batch's last pair repeats its
// first row and stores it once).  Same per-rhs arithmetic and order as
// k_dslash_zm, so the results are bit-identical.  D and Q epilogues only.
template <int PS, int PG, int EPI, int OCC>
__global__ void __launch_bounds__(kBlock) QH_WAVES(OCC)
k_dslash_zm2(DslashArgs a, SpinorIn<PS> in_all, GaugeIn<PG> U, SpinorOut<PS> out_all, SpinorIn<PS> b_all) {
  using R = typename Prec<PS>::R;
  using T = typename Prec<PS>::T;
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const unsigned p = lb % a.pairs, blk = lb / a.pairs;
  const unsigned r0 = 2 * p, r1 = min(2 * p + 1, a.count - 1);
  const bool two = r1 != r0;
  const RhsView<PS> v0 = rhs_view(a, in_all, out_all, b_all, r0);
  const RhsView<PS> v1 = rhs_view(a, in_all, out_all, b_all, r1);
  const unsigned cb = blk % a.col_blocks, zc = blk / a.col_blocks;
  const unsigned P3 = a.L[0] * a.L[1] * a.L[2];
  const unsigned col = cb * kBlock + threadIdx.x;
  const unsigned z0 = zc * a.zlen;
  const unsigned z1 = min(z0 + a.zlen, a.L[3]);
  if (col >= P3 || z0 >= a.L[3]) return;
  Site st;
  st.valid = true;
  st.x = col / a.L[0];
  st.t = col - st.x * a.L[0];
  st.y = st.x / a.L[1];
  st.x -= st.y * a.L[1];
  const Cx<R> php{(R)a.ph_re[3], (R)a.ph_im[3]};
  const Cx<R> phm{(R)a.ph_re[3], -(R)a.ph_im[3]};
  const R mp4 = R(a.mp4), tm = R(a.tm);
  Cx<R> cz0[2][3], cz1[2][3];
  {
    const unsigned zm = z0 == 0 ? a.L[3] - 1 : z0 - 1;
    const unsigned sm = col + P3 * zm;
    typename Prec<PG>::T u[9];
    T r[12];
    float n = 0.f;
    U.load_raw(3, sm, u);
    v0.in.load_raw(sm, r, n);
    hop_half_raw<3, +1, true, PS, PG>(cz0[0], cz1[0], r, n, u, phm);
    v1.in.load_raw(sm, r, n);
    hop_half_raw<3, +1, true, PS, PG>(cz0[1], cz1[1], r, n, u, phm);
  }
  for (unsigned z = z0; z < z1; z++) {
    st.z = z;
    st.s = col + P3 * z;
    const unsigned s = st.s;
    Cx<R> acc0[4][3], acc1[4][3];
    {
      T rc[12];
      typename Prec<PG>::T u[9];
      float nc = 0.f;
      Cx<R> g0[3], g1[3];
      U.load_raw(3, s, u);
      const unsigned sf = z + 1 == a.L[3] ? col : s + P3;
      v0.in.load_raw(s, rc, nc);
      local_terms<PS>(acc0, rc, nc, cz0[0], cz1[0], mp4, tm);
      hop_half_raw<3, +1, true, PS, PG>(cz0[0], cz1[0], rc, nc, u, phm);
      v0.in.load_raw(sf, rc, nc);
      hop_half_raw<3, -1, false, PS, PG>(g0, g1, rc, nc, u, php);
      rebuild_add<3, -1>(acc0, g0, g1);
      v1.in.load_raw(s, rc, nc);
      local_terms<PS>(acc1, rc, nc, cz0[1], cz1[1], mp4, tm);
      hop_half_raw<3, +1, true, PS, PG>(cz0[1], cz1[1], rc, nc, u, phm);
      v1.in.load_raw(sf, rc, nc);
      hop_half_raw<3, -1, false, PS, PG>(g0, g1, rc, nc, u, php);
      rebuild_add<3, -1>(acc1, g0, g1);
    }
    constexpr bool F = (OCC & 16) != 0;
    if constexpr (F) __builtin_amdgcn_sched_barrier(0);
    hop_dir2<0, PS, PG, R>(acc0, acc1, a, st, v0.in, v1.in, U);
    if constexpr (F) __builtin_amdgcn_sched_barrier(0);
    hop_dir2<1, PS, PG, R>(acc0, acc1, a, st, v0.in, v1.in, U);
    if constexpr (F) __builtin_amdgcn_sched_barrier(0);
    hop_dir2<2, PS, PG, R>(acc0, acc1, a, st, v0.in, v1.in, U);
    if constexpr (F) __builtin_amdgcn_sched_barrier(0);
    finish_site<EPI>(acc0, v0.out, v0.b, s);
    if (two) finish_site<EPI>(acc1, v1.out, v1.b, s);
  }
}

// Bytes of dynamic LDS a MAP_ZTILE workgroup uses: (xb + 2) staged rows of T
// sites, each site its 12 stored components and (FP16) its norm.
inline size_t ztile_lds_bytes(int ps, unsigned T, unsigned xb) {
  const size_t comp = ps == QCD_HIP_F64 ? 16 : ps == QCD_HIP_F32 ? 8 : 4;
  return (size_t)(xb + 2) * T * (12 * comp + (ps == QCD_HIP_F16 ? 4 : 0));
}

// z-marching Dslash over tiles of whole t-rows.  Workgroup (tile g, y, z
// chunk) owns rows x0 .. x0+nr-1 (nr <= xb) of plane y, lane (t, r) one site
// of row x0 + r.  Each step it writes the stored spinors of its rows and of
// the halo rows x0-1 and x0+nr (mod X) to LDS; the t hops (always within the
// row) and the x hops then read LDS, and only the y and z hops and the links
// use the vector memory path.  Same arithmetic and summation order as
// k_dslash_zm: the results are bit-identical.
template <int PS, int PG, int EPI, bool RED, int OCC>
__global__ void __launch_bounds__(kBlock) QH_WAVES(OCC)
k_dslash_zt(DslashArgs a, SpinorIn<PS> in_all, GaugeIn<PG> U, SpinorOut<PS> out_all, SpinorIn<PS> b_all,
            double *__restrict__ partials, typename SplitOf<EPI>::type split) {
  using R = typename Prec<PS>::R;
  using T = typename Prec<PS>::T;
  extern __shared__ __align__(16) unsigned char qh_dyn_lds[];
  const unsigned Lt = a.L[0], Lx = a.L[1], Ly = a.L[2];
  const unsigned nslot = (a.xb + 2) * Lt;
  T *stv = (T *)qh_dyn_lds;                      // [12][nslot]
  float *stn = (float *)(stv + 12 * nslot);     // [nslot] (FP16 norms)
  const BlockRow br = block_row(a);
  const RhsView<PS> rv = rhs_view(a, in_all, out_all, b_all, br.row);
  const SpinorIn<PS> &in = rv.in, &b = rv.b;
  const SpinorOut<PS> &out = rv.out;
  const auto srow = split_row(split, a.ids ? (unsigned)a.ids[br.row] : br.row);
  const unsigned lb = br.blk;
  const unsigned g = lb % a.nxg, r1 = lb / a.nxg;
  const unsigned y = r1 % Ly, zc = r1 / Ly;
  const unsigned x0 = g * a.xb;
  const unsigned nr = min(a.xb, Lx - x0);
  const unsigned tid = threadIdx.x;
  const unsigned t = tid % Lt, xr = tid / Lt;
  const bool active = xr < nr;
  const unsigned P2 = Lt * Lx, P3 = P2 * Ly;
  const unsigned z0 = zc * a.zlen;
  const unsigned z1 = min(z0 + a.zlen, a.L[3]);
  // halo rows: lanes h < 2 Lt (strided) load row x0-1 (h < Lt) or x0+nr (mod X)
  const unsigned xlo = x0 == 0 ? Lx - 1 : x0 - 1;
  const unsigned xhi = x0 + nr == Lx ? 0 : x0 + nr;
  Site st;
  st.valid = active;
  st.t = t;
  st.x = x0 + xr;
  st.y = y;
  const unsigned col = t + Lt * (st.x + Lx * y);
  const Cx<R> php{(R)a.ph_re[3], (R)a.ph_im[3]};
  const Cx<R> phm{(R)a.ph_re[3], -(R)a.ph_im[3]};
  const R mp4 = R(a.mp4), tm = R(a.tm);
  // staged slot of (t', row j): j * Lt + t', rows j = 0 (x0-1), 1..nr, nr+1 (x0+nr)
  const unsigned own = (xr + 1) * Lt + t;
  const unsigned tf = (xr + 1) * Lt + (t + 1 == Lt ? 0 : t + 1);
  const unsigned tb = (xr + 1) * Lt + (t == 0 ? Lt - 1 : t - 1);
  const unsigned xf = (xr + 2) * Lt + t, xbk = xr * Lt + t;
  double red = 0.0;
  Cx<R> cz0[3], cz1[3];
  if (active) {
    const unsigned zm = z0 == 0 ? a.L[3] - 1 : z0 - 1;
    const unsigned sm = col + P3 * zm;
    hop_half_at<3, +1, true>(cz0, cz1, in, sm, U, sm, phm);
  }
  for (unsigned z = z0; z < z1; z++) {   // z0, z1 are uniform over the workgroup
    st.z = z;
    st.s = col + P3 * z;
    const unsigned s = st.s;
    T rc[12];
    float nc = 0.f;
    if (active) in.load_raw(s, rc, nc);
    __syncthreads();   // every lane is done with the previous plane
    if (active) {
#pragma unroll
      for (int k = 0; k < 12; k++) stv[k * nslot + own] = rc[k];
      if constexpr (PS == QCD_HIP_F16) stn[own] = nc;
    }
    for (unsigned h = tid; h < 2 * Lt; h += blockDim.x) {
      const unsigned th = h < Lt ? h : h - Lt;
      const unsigned xh = h < Lt ? xlo : xhi;
      const unsigned slot = (h < Lt ? 0 : nr + 1) * Lt + th;
      T rh[12];
      float nh = 0.f;
      in.load_raw(th + Lt * (xh + Lx * y) + P3 * z, rh, nh);
#pragma unroll
      for (int k = 0; k < 12; k++) stv[k * nslot + slot] = rh[k];
      if constexpr (PS == QCD_HIP_F16) stn[slot] = nh;
    }
    __syncthreads();
    if (active) {
      Cx<R> acc[4][3];
      {
        typename Prec<PG>::T u[9];
        local_terms<PS>(acc, rc, nc, cz0, cz1, mp4, tm);
        // U_z(s) is read by this lane only (the next step carries the hop it
        // feeds): a streaming load leaves L2 to the spinor planes
        // residual epilogues run over batch rows whose groups share links: a normal load
        U.template load_raw<EPI == EPI_D || EPI == EPI_Q>(3, s, u);
        // the next site's backward z hop: U_z(s)^+ (1 + g_z) psi(s)
        hop_half_raw<3, +1, true, PS, PG>(cz0, cz1, rc, nc, u, phm);
        const unsigned sf = z + 1 == a.L[3] ? col : s + P3;
        Cx<R> g0[3], g1[3];
        float nf = 0.f;
        in.load_raw(sf, rc, nf);
        hop_half_raw<3, -1, false, PS, PG>(g0, g1, rc, nf, u, php);
        rebuild_add<3, -1>(acc, g0, g1);
      }
      constexpr bool F = (OCC & 16) != 0;
      // t and x hops from the staged rows
      {
        // one hop of direction MU from staged slot `slot`, link at site sl
        auto staged_hop = [&](auto mu_s_dag, unsigned slot, unsigned sl, Cx<R> ph) {
          constexpr int MU = decltype(mu_s_dag)::mu, S = decltype(mu_s_dag)::s;
          constexpr bool DAG = decltype(mu_s_dag)::dag;
          T r[12];
          typename Prec<PG>::T u[9];
#pragma unroll
          for (int k = 0; k < 12; k++) r[k] = stv[k * nslot + slot];
          U.load_raw(MU, sl, u);
          Cx<R> g0[3], g1[3];
          hop_half_raw<MU, S, DAG, PS, PG>(g0, g1, r, PS == QCD_HIP_F16 ? stn[slot] : 0.f, u, ph);
          rebuild_add<MU, S>(acc, g0, g1);
        };
        if constexpr (F) __builtin_amdgcn_sched_barrier(0);
        staged_hop(HopTag<0, -1, false>{}, tf, s, Cx<R>{(R)a.ph_re[0], (R)a.ph_im[0]});
        staged_hop(HopTag<0, +1, true>{}, tb, neighbour<0, -1>(st, a), Cx<R>{(R)a.ph_re[0], -(R)a.ph_im[0]});
        if constexpr (F) __builtin_amdgcn_sched_barrier(0);
        staged_hop(HopTag<1, -1, false>{}, xf, s, Cx<R>{(R)a.ph_re[1], (R)a.ph_im[1]});
        staged_hop(HopTag<1, +1, true>{}, xbk, neighbour<1, -1>(st, a), Cx<R>{(R)a.ph_re[1], -(R)a.ph_im[1]});
      }
      hop_dir<2, PS, PG, R, F>(acc, a, st, in, U);
      if constexpr (F) __builtin_amdgcn_sched_barrier(0);
      Cx<R> o[12];
#pragma unroll
      for (int sp = 0; sp < 4; sp++)
#pragma unroll
        for (int c = 0; c < 3; c++) o[3 * sp + c] = Cx<R>{R(-0.5) * acc[sp][c].re, R(-0.5) * acc[sp][c].im};
      apply_epilogue<EPI>(o, b, s);
      if (RED) {
#pragma unroll
        for (int k = 0; k < 12; k++) red += (double)o[k].re * (double)o[k].re + (double)o[k].im * (double)o[k].im;
      }
      store_site<EPI, PS>(out, srow, st, o);
    }
  }
  if (RED) block_reduce_store(red, partials, br.row * a.nblk + br.blk);
}

template <int PS, int PG, int EPI, bool RED, int OCC>
static int launch_occ(const DslashArgs &a, int grid, qcd_hip_field out, int ro, qcd_hip_field in, int ri,
                      qcd_hip_field g, qcd_hip_field b, int rb, double *partials, hipStream_t stream,
                      const void *split_host) {
  SpinorIn<PS> bi = b ? SpinorIn<PS>::of(b, rb) : SpinorIn<PS>::of(in, ri);
  typename SplitOf<EPI>::type sp{};
  if constexpr (epi_eo_split<EPI>()) {
    if (!split_host || (a.map != MAP_ZMARCH && a.map != MAP_ZMARCH_LDS && a.map != MAP_ZTILE && a.map != MAP_XCD &&
                        a.map != MAP_SLAB)) {
      set_error("dslash: the even-odd split residual needs its half fields and a one-rhs-per-lane traversal");
      return 1;
    }
    sp = *(const typename SplitOf<EPI>::type *)split_host;
  }
  // z-march traversals interleave the batch rows into grid.x (see block_row)
  if (a.map == MAP_MMA) {
    if constexpr (PS == QCD_HIP_F16 && PG == QCD_HIP_F16 && !RED && (EPI == EPI_D || EPI == EPI_Q)) {
      return dslash_mma_launch(EPI, a, out, ro, in, ri, g, b, rb, stream);
    } else {
      set_error("dslash: the matrix-core traversal needs FP16 spinors and links, D or Q and no reduction");
      return 1;
    }
  }
  const bool zmarch = a.map == MAP_ZMARCH || a.map == MAP_ZTILE || a.map == MAP_ZMARCH_LDS || a.map == MAP_ZMARCH2;
  const dim3 gr(zmarch ? (unsigned)grid * a.count : (unsigned)grid, zmarch ? 1u : (unsigned)a.count);
  if (a.map == MAP_ZMARCH)
    k_dslash_zm<PS, PG, EPI, RED, OCC><<<gr, kBlock, 0, stream>>>(a, SpinorIn<PS>::of(in, ri), GaugeIn<PG>::of(g),
                                                                  SpinorOut<PS>::of(out, ro), bi, partials, sp);
  else if (a.map == MAP_ZMARCH_LDS)
    k_dslash_zm<PS, PG, EPI, RED, OCC, true><<<gr, kBlock, 0, stream>>>(
        a, SpinorIn<PS>::of(in, ri), GaugeIn<PG>::of(g), SpinorOut<PS>::of(out, ro), bi, partials, sp);
  else if (a.map == MAP_ZMARCH2) {
    if constexpr (PS != QCD_HIP_F64 && !RED && (EPI == EPI_D || EPI == EPI_Q))
      k_dslash_zm2<PS, PG, EPI, OCC><<<dim3((unsigned)grid * a.pairs), kBlock, 0, stream>>>(
          a, SpinorIn<PS>::of(in, ri), GaugeIn<PG>::of(g), SpinorOut<PS>::of(out, ro), bi);
    else {
      set_error("dslash: two-rhs traversal needs FP32/FP16, the D or Q epilogue and no reduction");
      return 1;
    }
  }
  else if (a.map == MAP_ZTILE) {
    const size_t lds = ztile_lds_bytes(PS, a.L[0], a.xb);
    static std::atomic<size_t> lds_set{0};   // per instantiation: raise the dynamic LDS limit once
    if (lds > 65536 && lds > lds_set.load()) {
      QH_CHECK(hipFuncSetAttribute((const void *)k_dslash_zt<PS, PG, EPI, RED, OCC>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      lds_set.store(lds);
    }
    k_dslash_zt<PS, PG, EPI, RED, OCC><<<gr, a.lanes, lds, stream>>>(a, SpinorIn<PS>::of(in, ri), GaugeIn<PG>::of(g),
                                                                     SpinorOut<PS>::of(out, ro), bi, partials, sp);
  }
  else
    k_dslash<PS, PG, EPI, RED, OCC><<<gr, kBlock, 0, stream>>>(a, SpinorIn<PS>::of(in, ri), GaugeIn<PG>::of(g),
                                                               SpinorOut<PS>::of(out, ro), bi, partials, sp);
  QH_CHECK(hipGetLastError());
  return 0;
}

template <int PS, int PG, int EPI, bool RED>
static int launch_one(const DslashArgs &a, int occ, int grid, qcd_hip_field out, int ro, qcd_hip_field in, int ri,
                      qcd_hip_field g, qcd_hip_field b, int rb, double *partials, hipStream_t stream,
                      const void *split_host) {
#define QH_O(N) launch_occ<PS, PG, EPI, RED, N>(a, grid, out, ro, in, ri, g, b, rb, partials, stream, split_host)
  switch (occ) {
    case 2: return QH_O(2);
    case 3: return QH_O(3);
    case 16: return QH_O(16);
    default: return QH_O(0);
  }
#undef QH_O
}

template <int PS, int PG>
static int launch_epi(int epi, bool red, const DslashArgs &a, int occ, int grid, qcd_hip_field out, int ro,
                      qcd_hip_field in, int ri, qcd_hip_field g, qcd_hip_field b, int rb, double *partials,
                      hipStream_t stream, const void *split_host = nullptr) {
#define QH_L(E, RD) launch_one<PS, PG, E, RD>(a, occ, grid, out, ro, in, ri, g, b, rb, partials, stream, split_host)
  switch (epi * 2 + (red ? 1 : 0)) {
    case 0: return QH_L(EPI_D, false);
    case 1: return QH_L(EPI_D, true);
    case 2: return QH_L(EPI_Q, false);
    case 3: return QH_L(EPI_Q, true);
    case 4: return QH_L(EPI_RES_D, false);
    case 5: return QH_L(EPI_RES_D, true);
    case 6: return QH_L(EPI_RES_Q, false);
    case 7: return QH_L(EPI_RES_Q, true);
    case 9:   // the even-odd split residual: FP64 fields, with the |r|^2 reduction
      if constexpr (PS == QCD_HIP_F64 && PG == QCD_HIP_F64) return QH_L(EPI_RES_D_EO16, true);
      break;
    default: break;
  }
#undef QH_L
  set_error("dslash: unsupported epilogue %d (reduction %d) for spinor precision %d", epi, (int)red, PS);
  return 1;
}

DslashArgs make_args(qcd_hip_lattice lat, double m, double tm) {
  DslashArgs a;
  std::memset(&a, 0, sizeof(a));
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
  a.map = MAP_XCD;
  a.ids = nullptr;
  a.in_rs = a.out_rs = a.b_rs = 0;
  a.count = 1;
  return a;
}

// Fill the traversal fields of a; returns the grid size.  For MAP_ZMARCH the
// low byte of `map` is 2 and the z steps per lane are map >> 8.
int set_traversal(DslashArgs &a, int map, int ps) {
  if ((map & 0xff) == MAP_MMA) {
    // 16-site blocks; the launcher derives its grid from them and the batch
    a.map = MAP_MMA;
    a.zlen = 1;
    return (int)((a.V + 15) / 16);
  }
  if ((map & 0xff) == MAP_ZTILE) {
    // rows per tile: the most whole t-rows in one workgroup, preferring a lane
    // count that fills whole wavefronts; falls back to MAP_ZMARCH when no tile
    // of whole rows fits a workgroup and the LDS
    const unsigned Lt = a.L[0], Lx = a.L[1];
    unsigned xb = std::min(Lx, (unsigned)kBlock / std::max(Lt, 1u));
    for (unsigned r = xb; r >= 1; r--)
      if ((Lt * r) % 64 == 0) {
        xb = r;
        break;
      }
    if (xb >= 1 && Lt <= (unsigned)kBlock && ztile_lds_bytes(ps, Lt, xb) <= 150 * 1024) {
      const unsigned zlen = std::max(1u, std::min((unsigned)(map >> 8), a.L[3]));
      a.map = MAP_ZTILE;
      a.zlen = zlen;
      a.xb = xb;
      a.nxg = (Lx + xb - 1) / xb;
      a.lanes = (Lt * xb + 63) / 64 * 64;
      a.col_blocks = a.nxg * a.L[2];
      const unsigned nz = (a.L[3] + zlen - 1) / zlen;
      return (int)(a.col_blocks * nz);
    }
    map = MAP_ZMARCH | (map & ~0xff);
  }
  if ((map & 0xff) == MAP_ZMARCH || (map & 0xff) == MAP_ZMARCH_LDS || (map & 0xff) == MAP_ZMARCH2) {
    const unsigned zlen = std::max(1u, std::min((unsigned)(map >> 8), a.L[3]));
    const unsigned P3 = a.L[0] * a.L[1] * a.L[2];
    a.map = map & 0xff;
    a.zlen = zlen;
    a.col_blocks = (P3 + kBlock - 1) / kBlock;
    const unsigned nz = (a.L[3] + zlen - 1) / zlen;
    return (int)(a.col_blocks * nz);
  }
  if (map == MAP_SLAB && a.L[2] >= (unsigned)kNumXcd) {
    a.map = MAP_SLAB;
    const unsigned P = a.L[0] * a.L[1];
    unsigned most = 0;
    for (int k = 0; k <= kNumXcd; k++) a.slab_y0[k] = (unsigned)((unsigned long)k * a.L[2] / kNumXcd);
    for (int k = 0; k < kNumXcd; k++) {
      const unsigned long n = (unsigned long)P * (a.slab_y0[k + 1] - a.slab_y0[k]) * a.L[3];
      most = std::max(most, (unsigned)((n + kBlock - 1) / kBlock));
    }
    a.slab_blocks = most;
    return (int)(most * kNumXcd);
  }
  a.map = MAP_XCD;
  return grid_for(a.V);
}

// ------------------------------------------------------------ tuning
// Best (traversal, occupancy) per (lattice extents, precisions), chosen by
// timing every candidate on first use.  All candidates compute bit-identical
// results (same per-site arithmetic); only speed differs.
struct TuneKey {
  int L[4], ps, pg;
  bool operator==(const TuneKey &o) const { return !std::memcmp(this, &o, sizeof(*this)); }
};
struct TuneVal {
  int map, occ;
  float ms;
  int bmap = 0, bocc = 0;   // launches over two or more rhs (0: use map, occ)
};
static std::vector<std::pair<TuneKey, TuneVal>> g_tune;
static int g_autotune = -1;   // -1: read QCD_HIP_AUTOTUNE once

// copy of the tuned entry for k (the table may grow under other threads)
static bool tune_get(const TuneKey &k, TuneVal *out) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  for (auto &e : g_tune)
    if (e.first == k) {
      if (out) *out = e.second;
      return true;
    }
  return false;
}

static TuneKey key_of(qcd_hip_field in, qcd_hip_field g) {
  TuneKey k;
  std::memset(&k, 0, sizeof(k));
  for (int d = 0; d < 4; d++) k.L[d] = in->lat->L[d];
  k.ps = in->prec;
  k.pg = g->prec;
  return k;
}

struct Batch {
  const int *ids;
  int count, in_rs, out_rs, b_rs;
};
static const Batch kOne{nullptr, 1, 0, 0, 0};

static int launch_cfg(int epi, bool red, int map, int occ, qcd_hip_field out, int ro, qcd_hip_field in, int ri,
                      qcd_hip_field g, qcd_hip_field b, int rb, double m, double tm, double *partials,
                      hipStream_t stream, int *grid_out, const Batch &bt = kOne, const void *split = nullptr) {
  DslashArgs a = make_args(in->lat, m, tm);
  const int grid = set_traversal(a, map, in->prec);
  a.ids = bt.ids;
  a.count = (unsigned)bt.count;
  a.in_rs = (unsigned)bt.in_rs;
  a.out_rs = (unsigned)bt.out_rs;
  a.b_rs = (unsigned)bt.b_rs;
  a.nblk = (unsigned)grid;
  a.pairs = (a.count + 1) / 2;
  {
    // batch rows in groups of 4 (block_row): measured at 24^3x48 against 1, 2,
    // 8 and the whole batch, 4 was the best or within 2% (FP64 12/48 rhs -3%,
    // FP32 16 rhs -10% against the whole batch).  QCD_HIP_DSLASH_RGROUP
    // overrides it (0 = the whole batch).
    static const long env = [] {
      const char *e = getenv("QCD_HIP_DSLASH_RGROUP");
      return e ? std::atol(e) : -1L;
    }();
    a.rgroup = env >= 0 ? (unsigned)env : 4u;
  }
  if (a.map == MAP_ZMARCH2 && (a.count < 2 || in->prec == QCD_HIP_F64 || red || (epi != EPI_D && epi != EPI_Q))) {
    // the two-rhs traversal does not apply: same z-march, one rhs per lane
    a.map = MAP_ZMARCH;
  }
  if (a.map == MAP_MMA && (in->prec != QCD_HIP_F16 || g->prec != QCD_HIP_F16 || red || (epi != EPI_D && epi != EPI_Q))) {
    // the matrix-core traversal does not apply: the z-march of 4 steps
    const int gz = set_traversal(a, MAP_ZMARCH | (4 << 8), in->prec);
    a.nblk = (unsigned)gz;
    if (grid_out) *grid_out = gz;
    const int ps = in->prec, pg = g->prec;
    if (ps == QCD_HIP_F64 && pg == QCD_HIP_F64)
      return launch_epi<QCD_HIP_F64, QCD_HIP_F64>(epi, red, a, occ, gz, out, ro, in, ri, g, b, rb, partials, stream, split);
    if (ps == QCD_HIP_F32 && pg == QCD_HIP_F32)
      return launch_epi<QCD_HIP_F32, QCD_HIP_F32>(epi, red, a, occ, gz, out, ro, in, ri, g, b, rb, partials, stream, split);
    if (ps == QCD_HIP_F32 && pg == QCD_HIP_F16)
      return launch_epi<QCD_HIP_F32, QCD_HIP_F16>(epi, red, a, occ, gz, out, ro, in, ri, g, b, rb, partials, stream, split);
    return launch_epi<QCD_HIP_F16, QCD_HIP_F16>(epi, red, a, occ, gz, out, ro, in, ri, g, b, rb, partials, stream, split);
  }
  if (grid_out) *grid_out = grid;
  const int ps = in->prec, pg = g->prec;
  if (ps == QCD_HIP_F64 && pg == QCD_HIP_F64)
    return launch_epi<QCD_HIP_F64, QCD_HIP_F64>(epi, red, a, occ, grid, out, ro, in, ri, g, b, rb, partials, stream, split);
  if (ps == QCD_HIP_F32 && pg == QCD_HIP_F32)
    return launch_epi<QCD_HIP_F32, QCD_HIP_F32>(epi, red, a, occ, grid, out, ro, in, ri, g, b, rb, partials, stream, split);
  if (ps == QCD_HIP_F32 && pg == QCD_HIP_F16)
    return launch_epi<QCD_HIP_F32, QCD_HIP_F16>(epi, red, a, occ, grid, out, ro, in, ri, g, b, rb, partials, stream, split);
  if (ps == QCD_HIP_F16 && pg == QCD_HIP_F16)
    return launch_epi<QCD_HIP_F16, QCD_HIP_F16>(epi, red, a, occ, grid, out, ro, in, ri, g, b, rb, partials, stream, split);
  set_error("dslash: unsupported precision pair (spinor %d, gauge %d)", ps, pg);
  return 1;
}

// untuned default: z-march of 6 steps per lane
static const int kDefaultMap = MAP_ZMARCH | (6 << 8);
// Candidates: z-march lengths (1 = one site per lane), without and with the
// LDS-staged (t, x) neighbourhood.  The per-site traversals MAP_XCD / MAP_SLAB stay selectable but are
// never faster: their straight-line hop sequence lets the compiler hoist every
// load and spill.
static const int kMaps[] = {MAP_ZMARCH | (1 << 8), MAP_ZMARCH | (2 << 8), MAP_ZMARCH | (3 << 8),
                            MAP_ZMARCH | (4 << 8), MAP_ZMARCH | (6 << 8), MAP_ZMARCH | (8 << 8),
                            MAP_ZMARCH | (12 << 8), MAP_ZMARCH | (16 << 8),
                            MAP_ZTILE | (1 << 8),  MAP_ZTILE | (2 << 8),  MAP_ZTILE | (3 << 8),
                            MAP_ZTILE | (4 << 8),  MAP_ZTILE | (6 << 8),  MAP_ZTILE | (8 << 8),
                            MAP_ZTILE | (12 << 8), MAP_ZTILE | (16 << 8),
                            MAP_ZMARCH_LDS | (2 << 8), MAP_ZMARCH_LDS | (3 << 8), MAP_ZMARCH_LDS | (4 << 8),
                            MAP_ZMARCH_LDS | (6 << 8), MAP_ZMARCH_LDS | (8 << 8), MAP_ZMARCH_LDS | (12 << 8),
                            MAP_ZMARCH_LDS | (16 << 8)};
static const int kOccs[] = {0, 2, 3, 16};

int dslash_tune(qcd_hip_field out, qcd_hip_field in, qcd_hip_field g, int force) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  const TuneKey key = key_of(in, g);
  if (!force && tune_get(key, nullptr)) return 0;
  hipStream_t st = ctx()->stream;
  hipEvent_t e0, e1;
  QH_CHECK(hipEventCreate(&e0));
  QH_CHECK(hipEventCreate(&e1));
  TuneVal best{MAP_XCD, 0, 1e30f};
  const int reps = in->lat->V > (1L << 22) ? 3 : 10;
  for (int map : kMaps) {
    if (map == MAP_SLAB && in->lat->L[2] < kNumXcd) continue;
    for (int occ : kOccs) {
      for (int w = 0; w < 2; w++)
        QH_TRY(launch_cfg(EPI_D, false, map, occ, out, 0, in, 0, g, nullptr, 0, 0.0, 0.0, nullptr, st, nullptr));
      QH_CHECK(hipEventRecord(e0, st));
      for (int r = 0; r < reps; r++)
        QH_TRY(launch_cfg(EPI_D, false, map, occ, out, 0, in, 0, g, nullptr, 0, 0.0, 0.0, nullptr, st, nullptr));
      QH_CHECK(hipEventRecord(e1, st));
      QH_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      QH_CHECK(hipEventElapsedTime(&ms, e0, e1));
      ms /= reps;
      if (getenv("QCD_HIP_TUNE_VERBOSE"))
        fprintf(stderr, "tune L=%dx%dx%dx%d ps=%d pg=%d map=%d zlen=%d occ=%d: %.2f us\n", key.L[0], key.L[1], key.L[2],
                key.L[3], key.ps, key.pg, map & 0xff, map >> 8, occ, ms * 1e3);
      if (ms < best.ms) best = TuneVal{map, occ, ms};
    }
  }
  // Batches (fields with several rhs): the best one-rhs traversal against
  // other traversals, timed over the first min(nrhs, 48) rhs (at 24^3x48 the
  // matrix cores trail at 16 rhs and lead at 48, so a choice made on fewer rhs
  // than the batch can be wrong).  FP32/FP16: two rhs per lane and the matrix
  // cores.  FP64 (one rhs per lane): other z-march lengths and row tiles; over
  // 48 rhs at 24^3x48 the z-march of 8 steps takes 5.8 ms against 6.8 for the
  // one-rhs choice (LDS carry, 12 steps).
  const int cnt = std::min(std::min(in->nrhs, out->nrhs), 48);
  if (cnt >= 2) {
    const Batch bt{nullptr, cnt, 1, 1, 1};
    float bbest = 1e30f;
    auto time_batch = [&](int map, int occ, float *ms) -> int {
      for (int w = 0; w < 2; w++)
        QH_TRY(launch_cfg(EPI_D, false, map, occ, out, 0, in, 0, g, nullptr, 0, 0.0, 0.0, nullptr, st, nullptr, bt));
      QH_CHECK(hipEventRecord(e0, st));
      for (int r = 0; r < reps; r++)
        QH_TRY(launch_cfg(EPI_D, false, map, occ, out, 0, in, 0, g, nullptr, 0, 0.0, 0.0, nullptr, st, nullptr, bt));
      QH_CHECK(hipEventRecord(e1, st));
      QH_CHECK(hipEventSynchronize(e1));
      QH_CHECK(hipEventElapsedTime(ms, e0, e1));
      *ms /= reps;
      return 0;
    };
    QH_TRY(time_batch(best.map, best.occ, &bbest));
    const float one = bbest;
    if (in->prec == QCD_HIP_F64) {
      for (int map : {MAP_ZMARCH | (4 << 8), MAP_ZMARCH | (6 << 8), MAP_ZMARCH | (8 << 8), MAP_ZMARCH | (12 << 8),
                      MAP_ZMARCH_LDS | (8 << 8), MAP_ZTILE | (4 << 8), MAP_ZTILE | (8 << 8)})
        for (int occ : {0, 16}) {
          if (map == best.map && occ == best.occ) continue;
          float ms = 0.f;
          QH_TRY(time_batch(map, occ, &ms));
          if (getenv("QCD_HIP_TUNE_VERBOSE"))
            fprintf(stderr, "tune L=%dx%dx%dx%d ps=%d pg=%d %d rhs map=%d zlen=%d occ=%d: %.2f us (one-rhs %.2f)\n",
                    key.L[0], key.L[1], key.L[2], key.L[3], key.ps, key.pg, cnt, map & 0xff, map >> 8, occ, ms * 1e3,
                    one * 1e3);
          if (ms < bbest) {
            bbest = ms;
            best.bmap = map;
            best.bocc = occ;
          }
        }
    }
    if (in->prec == QCD_HIP_F16 && g->prec == QCD_HIP_F16) {
      float ms = 0.f;
      QH_TRY(time_batch(MAP_MMA, 0, &ms));
      if (getenv("QCD_HIP_TUNE_VERBOSE"))
        fprintf(stderr, "tune L=%dx%dx%dx%d ps=%d pg=%d %d rhs matrix cores: %.2f us (one-rhs %.2f)\n", key.L[0],
                key.L[1], key.L[2], key.L[3], key.ps, key.pg, cnt, ms * 1e3, one * 1e3);
      if (ms < bbest) {
        bbest = ms;
        best.bmap = MAP_MMA;
        best.bocc = 0;
      }
    }
    if (in->prec != QCD_HIP_F64)
      for (int z : {2, 3, 4, 6, 8})
        for (int occ : {0, 16}) {
          float ms = 0.f;
          QH_TRY(time_batch(MAP_ZMARCH2 | (z << 8), occ, &ms));
          if (getenv("QCD_HIP_TUNE_VERBOSE"))
            fprintf(stderr,
                    "tune L=%dx%dx%dx%d ps=%d pg=%d %d rhs two-rhs lanes zlen=%d occ=%d: %.2f us (one-rhs %.2f)\n",
                    key.L[0], key.L[1], key.L[2], key.L[3], key.ps, key.pg, cnt, z, occ, ms * 1e3, one * 1e3);
          if (ms < bbest) {
            bbest = ms;
            best.bmap = MAP_ZMARCH2 | (z << 8);
            best.bocc = occ;
          }
        }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (auto &e : g_tune)
    if (e.first == key) {
      e.second = best;
      return 0;
    }
  g_tune.push_back({key, best});
  return 0;
}

static bool autotune_enabled() {
  if (g_autotune < 0) {
    const char *e = getenv("QCD_HIP_AUTOTUNE");
    g_autotune = (e && e[0] == '0') ? 0 : 1;
  }
  return g_autotune == 1;
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
  if (in->lat->stride * 12L * 16L >= (1L << 32)) {
    set_error("dslash: lattice too large for 32-bit offsets (V must be below 2^24)");
    return 1;
  }
  return 0;
}

// Single-rhs launch used by the solver and the public entry points.  Uses
// the tuned configuration of this (lattice, precision) if there is one.
int dslash_launch(int epi, bool red, qcd_hip_field out, int ro, qcd_hip_field in, int ri, qcd_hip_field g,
                  qcd_hip_field b, int rb, double m, double tm, double *partials, hipStream_t stream, int *grid_out) {
  TuneVal t;
  const bool have = tune_get(key_of(in, g), &t);
  const int map = have ? t.map : kDefaultMap, occ = have ? t.occ : 0;
  return launch_cfg(epi, red, map, occ, out, ro, in, ri, g, b, rb, m, tm, partials, stream, grid_out);
}

int dslash_launch_batch(int epi, bool red, qcd_hip_field out, int ro, int out_rs, qcd_hip_field in, int ri, int in_rs,
                        qcd_hip_field g, qcd_hip_field b, int rb, int b_rs, const int *ids, int count, double m,
                        double tm, double *partials, hipStream_t stream, int *grid_out) {
  if (count < 1) return 0;
  TuneVal t;
  const bool have = tune_get(key_of(in, g), &t);
  // the batch traversal applies where it was timed: FP64 batches were tuned on
  // D without reduction (the residual epilogues need more registers and keep
  // the one-rhs choice); FP32/FP16 residuals fall back inside launch_cfg
  const bool timed = in->prec != QCD_HIP_F64 || (!red && (epi == EPI_D || epi == EPI_Q));
  const bool two = have && count >= 2 && t.bmap != 0 && timed;
  const int map = two ? t.bmap : have ? t.map : kDefaultMap, occ = two ? t.bocc : have ? t.occ : 0;
  const Batch bt{ids, count, in_rs, out_rs, b_rs};
  return launch_cfg(epi, red, map, occ, out, ro, in, ri, g, b, rb, m, tm, partials, stream, grid_out, bt);
}

// r = b - D x of a batch (FP64 fields, as dslash_launch_batch with EPI_RES_D
// and the |r|^2 partials), stored split by parity into the FP16 half fields
// ev, od and tev = M^-1 ev (vector ids[y] of each for row y) instead of an
// FP64 field: the FP64 residual never goes to memory.  (An FP32 form of the
// same epilogue spilled registers in these kernels: FP32 inner solves keep the
// FP64 residual and a split kernel.)
int dslash_residual_eo_split(qcd_hip_field x, int x0, int x_rs, qcd_hip_field g, qcd_hip_field b, int rb, int b_rs,
                             const int *ids, int count, double m, double tm, qcd_hip_field ev, qcd_hip_field od,
                             qcd_hip_field tev, double *partials, hipStream_t stream, int *grid_out) {
  if (count < 1) return 0;
  if (x->prec != QCD_HIP_F64 || g->prec != QCD_HIP_F64 || b->prec != QCD_HIP_F64 || ev->prec != QCD_HIP_F16 ||
      od->prec != QCD_HIP_F16 || tev->prec != QCD_HIP_F16) {
    set_error("dslash: the split residual takes FP64 x, b and gauge fields and FP16 half fields");
    return 1;
  }
  TuneVal t;
  const bool have = tune_get(key_of(x, g), &t);
  // the one-rhs choice of D to start with (the FP64 batch choice was timed on
  // D and measured slower here: row tiles of 8 steps, 6.7 against 6.3 ms over
  // 48 rhs of 24^3x48)
  int map = have ? t.map : kDefaultMap, occ = have ? t.occ : 0;
  const Batch bt{ids, count, x_rs, 0, b_rs};
  const double a4 = m + 4.0, den = a4 * a4 + tm * tm;   // M^-1 as make_eo_args forms it
  const EoSplit<QCD_HIP_F16> sp{SpinorOut<QCD_HIP_F16>::of(ev, 0), SpinorOut<QCD_HIP_F16>::of(od, 0),
                                SpinorOut<QCD_HIP_F16>::of(tev, 0), a4 / den, -tm / den};
  if (count >= 2 && autotune_enabled()) {
    // This epilogue's register load differs from D's, and so does its best
    // traversal: the first batched call per lattice outside graph capture
    // times a few candidates on the call itself (each launch rewrites only the
    // outputs and partials; every candidate's grid is within
    // dslash_grid_bound).  24^3x48, 48 rhs: LDS-carry z-march of 12 steps
    // 5.8 ms against 6.4 for row tiles of 8, two choices the one-rhs D tuner
    // finds within 1% of each other.
    std::lock_guard<std::recursive_mutex> lock(state_mutex());
    static std::vector<std::pair<std::array<int, 4>, std::pair<int, int>>> tuned;
    const std::array<int, 4> key{(int)x->lat->L[0], (int)x->lat->L[1], (int)x->lat->L[2], (int)x->lat->L[3]};
    bool found = false;
    for (const auto &e : tuned)
      if (e.first == key) {
        map = e.second.first;
        occ = e.second.second;
        found = true;
      }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    QH_CHECK(hipStreamIsCapturing(stream, &cs));
    if (!found && cs == hipStreamCaptureStatusNone) {
      hipEvent_t e0, e1;
      QH_CHECK(hipEventCreate(&e0));
      QH_CHECK(hipEventCreate(&e1));
      float best = 1e30f;
      int bmap = map, bocc = occ, rc = 0;
      const int cands[][2] = {{map, occ},
                              {MAP_ZMARCH_LDS | (12 << 8), 16}, {MAP_ZMARCH_LDS | (8 << 8), 16},
                              {MAP_ZMARCH | (8 << 8), 16},      {MAP_ZMARCH | (4 << 8), 16},
                              {MAP_ZTILE | (8 << 8), 16},       {MAP_ZTILE | (4 << 8), 16}};
      for (const auto &c : cands) {
        rc = launch_cfg(EPI_RES_D_EO16, true, c[0], c[1], x, 0, x, x0, g, b, rb, m, tm, partials, stream, nullptr,
                        bt, &sp);
        if (!rc && hipEventRecord(e0, stream) != hipSuccess) rc = 2;
        for (int r = 0; r < 2 && !rc; r++)
          rc = launch_cfg(EPI_RES_D_EO16, true, c[0], c[1], x, 0, x, x0, g, b, rb, m, tm, partials, stream, nullptr,
                          bt, &sp);
        float ms = 0.f;
        if (!rc && (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
          rc = 2;
        if (rc) break;
        if (getenv("QCD_HIP_TUNE_VERBOSE"))
          fprintf(stderr, "tune split residual L=%dx%dx%dx%d %d rhs map=%d zlen=%d occ=%d: %.2f us\n", key[0], key[1],
                  key[2], key[3], count, c[0] & 0xff, c[0] >> 8, c[1], ms * 500.f);
        if (ms < best) {
          best = ms;
          bmap = c[0];
          bocc = c[1];
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      if (rc) {
        if (rc == 2) set_error("dslash: timing the split residual failed: %s", hipGetErrorString(hipGetLastError()));
        return rc;
      }
      map = bmap;
      occ = bocc;
      tuned.push_back({key, {map, occ}});
    }
  }
  return launch_cfg(EPI_RES_D_EO16, true, map, occ, x, 0, x, x0, g, b, rb, m, tm, partials, stream, grid_out, bt, &sp);
}

// Tune (once per key) before the first use, outside any graph capture.
int dslash_prepare(qcd_hip_field out, qcd_hip_field in, qcd_hip_field g) {
  if (!autotune_enabled()) return 0;
  return dslash_tune(out, in, g, 0);
}

// an upper bound of the Dslash grid over every traversal candidate
int dslash_grid_bound(qcd_hip_lattice lat) {
  int most = 0;
  for (int map : {(int)MAP_XCD, (int)MAP_SLAB, MAP_ZMARCH | (1 << 8), MAP_ZTILE | (1 << 8)})   // z-march of 1 has the most groups
    for (int ps : {QCD_HIP_F64, QCD_HIP_F32, QCD_HIP_F16}) {
      DslashArgs a = make_args(lat, 0.0, 0.0);
      most = std::max(most, set_traversal(a, map, ps));
    }
  return most;
}

int dslash_grid(qcd_hip_field in, qcd_hip_field g) {
  TuneVal t;
  const bool have = tune_get(key_of(in, g), &t);
  DslashArgs a = make_args(in->lat, 0.0, 0.0);
  return set_traversal(a, have ? t.map : kDefaultMap, in->prec);
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
  QH_TRY(dslash_prepare(out, in, gauge));
  QH_TRY(dslash_launch_batch(epi, false, out, 0, 1, in, 0, 1, gauge, nullptr, 0, 0, nullptr, in->nrhs, m, tm, nullptr,
                             ctx()->stream, nullptr));
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
  QH_TRY(dslash_prepare(out, in, gauge));
  for (int i = 0; i < warmup; i++)
    QH_TRY(dslash_launch_batch(epi, false, out, 0, 1, in, 0, 1, gauge, nullptr, 0, 0, nullptr, in->nrhs, m, tm, nullptr,
                               st, nullptr));
  hipEvent_t e0, e1;
  QH_CHECK(hipEventCreate(&e0));
  QH_CHECK(hipEventCreate(&e1));
  QH_CHECK(hipEventRecord(e0, st));
  for (int i = 0; i < iters; i++)
    QH_TRY(dslash_launch_batch(epi, false, out, 0, 1, in, 0, 1, gauge, nullptr, 0, 0, nullptr, in->nrhs, m, tm, nullptr,
                               st, nullptr));
  QH_CHECK(hipEventRecord(e1, st));
  QH_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  QH_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / iters;
  return 0;
}

// Time every traversal / occupancy candidate on these fields and keep the
// fastest for their (lattice, precision).  Fills tuned[0..2] = map, occ, ms.
int qcd_hip_dslash_tune(qcd_hip_field out, qcd_hip_field in, qcd_hip_field gauge, int tuned[2], double *ms) {
  QH_TRY(ensure_init());
  QH_TRY(dslash_check(out, in, gauge));
  QH_TRY(dslash_tune(out, in, gauge, 1));
  TuneVal t{0, 0, 0.f};
  tune_get(key_of(in, gauge), &t);
  if (tuned) {
    tuned[0] = t.map;
    tuned[1] = t.occ;
  }
  if (ms) *ms = t.ms;
  return 0;
}

// The traversal candidates the autotuner times (before the two-rhs batch
// candidates): fills up to max entries of maps, returns how many there are.
int qcd_hip_dslash_candidates(int *maps, int max) {
  const int n = (int)(sizeof(kMaps) / sizeof(kMaps[0]));
  for (int i = 0; maps && i < n && i < max; i++) maps[i] = kMaps[i];
  return n;
}

// Force a configuration for these fields' key: map = 0 contiguous per XCD,
// 1 y-slab per XCD, 2 | (n << 8) z-march of n steps, 3 | (n << 8) z-march of
// n steps with LDS-staged t and x neighbours, 4 | (n << 8) z-march with the
// forward z neighbour carried in LDS, 5 | (n << 8) z-march with two rhs per
// lane (batches of FP32/FP16 D or Q; otherwise the one-rhs z-march); occ = 0
// compiler default, 2 or 3 minimum waves per SIMD, 16 scheduling fence
// between hop directions.
int qcd_hip_dslash_set_config(qcd_hip_field in, qcd_hip_field gauge, int map, int occ) {
  QH_TRY(ensure_init());
  const int m0 = map & 0xff;
  if (!in || !gauge || m0 < 0 || m0 > 6 || (m0 >= 2 && m0 <= 5 && (map >> 8) < 1) ||
      (occ != 0 && occ != 2 && occ != 3 && occ != 16)) {
    set_error("dslash_set_config: bad arguments");
    return 1;
  }
  const TuneKey k = key_of(in, gauge);
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  for (auto &e : g_tune)
    if (e.first == k) {
      e.second = TuneVal{map, occ, 0.f};
      return 0;
    }
  g_tune.push_back({k, TuneVal{map, occ, 0.f}});   // forced: batches use the same traversal
  return 0;
}

}  // extern "C"
