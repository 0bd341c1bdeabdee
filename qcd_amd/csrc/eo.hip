// eo.hip -- even-odd (Schur complement) form of the twisted-mass Wilson
// operator, used by the inner solve of the propagator solver.
//
// With sites split by parity p = (t + x + y + z) mod 2 and D = M + H, where
// M = (m + 4) + i tm g5 is site-diagonal and the hopping term H only couples
// opposite parities,
//   D = [[M, H_eo], [H_oe, M]],   D^ = M - H_oe M^-1 H_eo   (on odd sites).
// D x = b is solved as D^ x_o = b_o - H_oe M^-1 b_e, x_e = M^-1 (b_e - H_eo x_o).
// Half-lattice fields hold the sites of one parity: site l of parity p is
// entry l >> 1 (every extent even, so the t-rows alternate parities).
// The hopping sum uses the projector/link/phase arithmetic of the full Dslash.
#include "eo_common.h"

#include <atomic>
#include <cstdio>
#include <tuple>
#include <map>
#include <type_traits>

namespace qcdhip {

// both hops of direction MU into acc, the neighbour spinors read from the
// half-lattice field of the other parity
template <int MU, int PS, int PG, typename R>
__device__ __forceinline__ void eo_hop_dir(Cx<R> acc[4][3], const DslashArgs &a, const Site &st,
                                           const SpinorIn<PS> &in, const GaugeIn<PG> &U) {
  const Cx<R> php{(R)a.ph_re[MU], (R)a.ph_im[MU]};
  const Cx<R> phm{(R)a.ph_re[MU], -(R)a.ph_im[MU]};
  hop_acc_at<MU, -1, false>(acc, in, neighbour<MU, +1>(st, a) >> 1, U, st.s, php);
  const unsigned sb = neighbour<MU, -1>(st, a);
  hop_acc_at<MU, +1, true>(acc, in, sb >> 1, U, sb, phm);
}

// Everything that feeds acc is computed before this point, and no later load
// is scheduled above it (FP64 and FP32; FP16 has no register pressure).
template <int PS, typename R>
__device__ __forceinline__ void eo_order_point(Cx<R> acc[4][3]) {
  if constexpr (PS != QCD_HIP_F16) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int c = 0; c < 3; c++) asm volatile("" : "+v"(acc[q][c].re), "+v"(acc[q][c].im));
    __builtin_amdgcn_sched_barrier(0);
  }
}

// both hops of direction MU for two right-hand sides: links loaded once
template <int MU, int PS, int PG, typename R>
__device__ __forceinline__ void eo_hop_dir2(Cx<R> acc0[4][3], Cx<R> acc1[4][3], const DslashArgs &a, const Site &st,
                                            const SpinorIn<PS> &in0, const SpinorIn<PS> &in1, const GaugeIn<PG> &U) {
  const Cx<R> php{(R)a.ph_re[MU], (R)a.ph_im[MU]};
  const Cx<R> phm{(R)a.ph_re[MU], -(R)a.ph_im[MU]};
  typename Prec<PS>::T r[12];
  typename Prec<PG>::T u[9];
  Cx<R> g0[3], g1[3];
  float n = 0.f;
  {
    const unsigned sf = neighbour<MU, +1>(st, a) >> 1;
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
    in0.load_raw(sb >> 1, r, n);
    hop_half_raw<MU, +1, true, PS, PG>(g0, g1, r, n, u, phm);
    rebuild_add<MU, +1>(acc0, g0, g1);
    in1.load_raw(sb >> 1, r, n);
    hop_half_raw<MU, +1, true, PS, PG>(g0, g1, r, n, u, phm);
    rebuild_add<MU, +1>(acc1, g0, g1);
  }
}

// k_eo_hop with two right-hand sides per lane (batch rows 2p and 2p+1 of pair
// p; an odd batch's last pair repeats its first row and stores it once), for
// FP32/FP16 batches with the M^-1 H or M b - H epilogue.  Per rhs the same
// arithmetic in the same order as k_eo_hop: bit-identical results.
template <int PS, int PG, int EPI>
__global__ void __launch_bounds__(kBlock) k_eo_hop2(EoArgs e, SpinorIn<PS> in_all, GaugeIn<PG> U,
                                                    SpinorOut<PS> out_all, SpinorIn<PS> b_all) {
  using R = typename Prec<PS>::R;
  const DslashArgs &a = e.d;
  const unsigned pairs = (a.count + 1) / 2;
  const BlockRow br = eo_row_of(pairs, max(1u, e.rgroup / 2), a.nblk);
  const unsigned r0 = 2 * br.row, r1 = min(2 * br.row + 1, a.count - 1);
  const RhsView<PS> v0 = rhs_view(a, in_all, out_all, b_all, r0);
  const RhsView<PS> v1 = rhs_view(a, in_all, out_all, b_all, r1);
  const unsigned h = br.blk * kBlock + threadIdx.x;
  if (h >= e.Vh) return;
  const Site st = eo_site(h, e.par, a);
  Cx<R> acc0[4][3], acc1[4][3];
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int c = 0; c < 3; c++) acc0[q][c] = acc1[q][c] = Cx<R>{R(0), R(0)};
  eo_hop_dir2<0, PS, PG, R>(acc0, acc1, a, st, v0.in, v1.in, U);
  eo_order_point<PS>(acc0);
  eo_order_point<PS>(acc1);
  eo_hop_dir2<1, PS, PG, R>(acc0, acc1, a, st, v0.in, v1.in, U);
  eo_order_point<PS>(acc0);
  eo_order_point<PS>(acc1);
  eo_hop_dir2<2, PS, PG, R>(acc0, acc1, a, st, v0.in, v1.in, U);
  eo_order_point<PS>(acc0);
  eo_order_point<PS>(acc1);
  eo_hop_dir2<3, PS, PG, R>(acc0, acc1, a, st, v0.in, v1.in, U);
  Cx<R> o[12];
  eo_epilogue<EPI>(o, acc0, e, v0.b, h);
  v0.out.template store<true>(h, o);
  if (r1 != r0) {
    eo_epilogue<EPI>(o, acc1, e, v1.b, h);
    v1.out.template store<true>(h, o);
  }
}

// one lane per output site; workgroups map to (batch row, site block) by eo_row
template <int PS, int PG, int EPI, bool RED>
__global__ void __launch_bounds__(kBlock) k_eo_hop(EoArgs e, SpinorIn<PS> in_all, GaugeIn<PG> U,
                                                   SpinorOut<PS> out_all, SpinorIn<PS> b_all,
                                                   double *__restrict__ partials) {
  using R = typename Prec<PS>::R;
  const DslashArgs &a = e.d;
  const BlockRow br = eo_row(e);
  const RhsView<PS> rv = rhs_view(a, in_all, out_all, b_all, br.row);
  const unsigned h = br.blk * kBlock + threadIdx.x;
  double red = 0.0;
  if (h < e.Vh) {
    const Site st = eo_site(h, e.par, a);
    Cx<R> acc[4][3];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int c = 0; c < 3; c++) acc[q][c] = Cx<R>{R(0), R(0)};
    // One direction at a time: the instruction selector would otherwise issue
    // the loads of all eight hops first and leave the arithmetic for the end
    // (FP64: 512 VGPRs plus spills, FP32: 298).  Pinning acc after each
    // direction and a scheduling fence keep each direction's loads and
    // arithmetic together.
    eo_hop_dir<0, PS, PG, R>(acc, a, st, rv.in, U);
    eo_order_point<PS>(acc);
    eo_hop_dir<1, PS, PG, R>(acc, a, st, rv.in, U);
    eo_order_point<PS>(acc);
    eo_hop_dir<2, PS, PG, R>(acc, a, st, rv.in, U);
    eo_order_point<PS>(acc);
    eo_hop_dir<3, PS, PG, R>(acc, a, st, rv.in, U);
    Cx<R> o[12];
    eo_epilogue<EPI>(o, acc, e, rv.b, h);
    if (RED) {
#pragma unroll
      for (int k = 0; k < 12; k++) red += (double)o[k].re * (double)o[k].re + (double)o[k].im * (double)o[k].im;
    }
    rv.out.template store<true>(h, o);
  }
  if (RED) block_reduce_store(red, partials, br.row * a.nblk + br.blk);
}

// acc of one or two rows pinned in registers and a scheduling fence: the
// loads of the next hop direction are not hoisted above this point (every
// precision: the z-march carries a ring of hops on top of the accumulators)
template <typename R>
__device__ __forceinline__ void zm_order_point(Cx<R> acc[4][3]) {
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int c = 0; c < 3; c++) asm volatile("" : "+v"(acc[q][c].re), "+v"(acc[q][c].im));
  __builtin_amdgcn_sched_barrier(0);
}
template <typename R>
__device__ __forceinline__ void zm_order_point(Cx<R> acc0[4][3], Cx<R> acc1[4][3]) {
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      asm volatile("" : "+v"(acc0[q][c].re), "+v"(acc0[q][c].im));
      asm volatile("" : "+v"(acc1[q][c].re), "+v"(acc1[q][c].im));
    }
  __builtin_amdgcn_sched_barrier(0);
}

// z-marching EO hop: a lane owns one half column j (the t-pair 2j', 2j'+1 of
// one (x, y)) of NR batch rows and steps through planes z0 .. z1-1.  Its output
// site at plane z has t = 2j' + odd(z), odd(z) = (par + x + y + z) & 1, and
// half-lattice entry j + PP z.  The z neighbours of that site are the input
// half column j at planes z -+ 1, and the forward one, (t, x, y, z + 1), is also
// the backward z neighbour of the lane's output site at z + 2.  So when the lane
// loads it for the forward hop it also forms the backward hop of step z + 2,
// U_z(t, x, y, z + 1)^+ (1 + g_z) psi, and carries it two steps (a ring of two
// colour-multiplied half spinors per row).  Every input spinor of a z chunk is
// then read once by its own lane for both z hops, and the t, x, y neighbours
// at plane z are read by the lanes of nearby columns at the same step.
// Each hop is computed from the same values with the same arithmetic as
// k_eo_hop / k_eo_hop2 and added in the same order (t, x, y, z; forward, then
// backward), so the results are bit-identical.  NR = 2: rows 2p and 2p + 1 of
// pair p, each link loaded once for both (an odd batch's last pair repeats
// its row and stores it once).
template <int PS, int PG, int EPI, bool RED, int NR>
__global__ void __launch_bounds__(kBlock) k_eo_hop_zm(EoArgs e, SpinorIn<PS> in_all, GaugeIn<PG> U,
                                                      SpinorOut<PS> out_all, SpinorIn<PS> b_all,
                                                      double *__restrict__ partials) {
  static_assert(NR == 1 || NR == 2, "one or two rows per lane");
  using R = typename Prec<PS>::R;
  const DslashArgs &a = e.d;
  const unsigned nrow = NR == 2 ? (a.count + 1) / 2 : a.count;
  const BlockRow br = eo_row_of(nrow, NR == 2 ? max(1u, e.rgroup / 2) : e.rgroup, a.nblk);
  const unsigned r0 = NR * br.row, r1 = NR == 2 ? min(r0 + 1, a.count - 1) : r0;
  RhsView<PS> v[NR] = {rhs_view(a, in_all, out_all, b_all, r0)};
  if constexpr (NR == 2) v[1] = rhs_view(a, in_all, out_all, b_all, r1);
  const unsigned cb = br.blk % a.col_blocks, zc = br.blk / a.col_blocks;
  const unsigned P3 = a.L[0] * a.L[1] * a.L[2], PP = P3 / 2, Lz = a.L[3];
  const unsigned j = cb * kBlock + threadIdx.x;
  const unsigned z0 = zc * a.zlen, z1 = min(z0 + a.zlen, Lz);
  double red = 0.0;
  if (j < PP && z0 < Lz) {
    Site st;
    st.valid = true;
    const unsigned l0 = 2 * j;
    const unsigned rr = l0 / a.L[0];
    const unsigned t0 = l0 - rr * a.L[0];
    st.y = rr / a.L[1];
    st.x = rr - st.y * a.L[1];
    const Cx<R> php{(R)a.ph_re[3], (R)a.ph_im[3]};
    const Cx<R> phm{(R)a.ph_re[3], -(R)a.ph_im[3]};
    // cz[n][q]: backward z hop (two half spinors) of row n's output site at a
    // plane z with (z - z0) % 2 == q; a step uses its slot and refills it with
    // the hop of z + 2, so the slot index is fixed per unrolled step
    Cx<R> cz[NR][2][2][3];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const unsigned zq = z0 + q;   // output plane; its backward neighbour is at zq - 1
      const unsigned odd = (e.par + st.x + st.y + zq) & 1u;
      const unsigned sb = l0 + odd + P3 * (zq == 0 ? Lz - 1 : zq - 1);   // zq <= Lz: in range
      typename Prec<PG>::T u[9];
      U.load_raw(3, sb, u);
#pragma unroll
      for (int n = 0; n < NR; n++) {
        typename Prec<PS>::T r[12];
        float nn = 0.f;
        v[n].in.load_raw(sb >> 1, r, nn);
        hop_half_raw<3, +1, true, PS, PG>(cz[n][q][0], cz[n][q][1], r, nn, u, phm);
      }
    }
    auto step = [&](const unsigned z, auto slot) {
      constexpr int S = decltype(slot)::value;
      const unsigned odd = (e.par + st.x + st.y + z) & 1u;
      st.t = t0 + odd;
      st.z = z;
      st.s = l0 + odd + P3 * z;
      const unsigned h = j + PP * z;
      Cx<R> acc[NR][4][3];
#pragma unroll
      for (int n = 0; n < NR; n++)
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
          for (int c = 0; c < 3; c++) acc[n][q][c] = Cx<R>{R(0), R(0)};
      if constexpr (NR == 1) {
        eo_hop_dir<0, PS, PG, R>(acc[0], a, st, v[0].in, U);
        zm_order_point(acc[0]);
        eo_hop_dir<1, PS, PG, R>(acc[0], a, st, v[0].in, U);
        zm_order_point(acc[0]);
        eo_hop_dir<2, PS, PG, R>(acc[0], a, st, v[0].in, U);
        zm_order_point(acc[0]);
      } else {
        eo_hop_dir2<0, PS, PG, R>(acc[0], acc[1], a, st, v[0].in, v[1].in, U);
        zm_order_point(acc[0], acc[1]);
        eo_hop_dir2<1, PS, PG, R>(acc[0], acc[1], a, st, v[0].in, v[1].in, U);
        zm_order_point(acc[0], acc[1]);
        eo_hop_dir2<2, PS, PG, R>(acc[0], acc[1], a, st, v[0].in, v[1].in, U);
        zm_order_point(acc[0], acc[1]);
      }
      // z: forward hop from (t, x, y, z + 1), whose backward hop for the output
      // site at z + 2 is formed from the same stored values; then the carried
      // backward hop of this step
      const bool ahead = z + 2 < z1;   // uniform over the workgroup
      const unsigned sf = z + 1 == Lz ? l0 + odd : st.s + P3;
      typename Prec<PG>::T u[9], u2[9];
      U.load_raw(3, st.s, u);
      if (ahead) U.load_raw(3, sf, u2);
#pragma unroll
      for (int n = 0; n < NR; n++) {
        typename Prec<PS>::T r[12];
        float nn = 0.f;
        v[n].in.load_raw(sf >> 1, r, nn);
        Cx<R> g0[3], g1[3];
        hop_half_raw<3, -1, false, PS, PG>(g0, g1, r, nn, u, php);
        rebuild_add<3, -1>(acc[n], g0, g1);
        rebuild_add<3, +1>(acc[n], cz[n][S][0], cz[n][S][1]);
        if (ahead) hop_half_raw<3, +1, true, PS, PG>(cz[n][S][0], cz[n][S][1], r, nn, u2, phm);
      }
#pragma unroll
      for (int n = 0; n < NR; n++) {
        if (n == 1 && r1 == r0) break;
        Cx<R> o[12];
        eo_epilogue<EPI>(o, acc[n], e, v[n].b, h);
        if (RED) {
#pragma unroll
          for (int k = 0; k < 12; k++) red += (double)o[k].re * (double)o[k].re + (double)o[k].im * (double)o[k].im;
        }
        v[n].out.template store<true>(h, o);
      }
    };
    for (unsigned z = z0; z < z1; z += 2) {
      step(z, std::integral_constant<int, 0>{});
      if (z + 1 < z1) step(z + 1, std::integral_constant<int, 1>{});
    }
  }
  if (RED) block_reduce_store(red, partials, br.row * a.nblk + br.blk);
}

// Parity of full-lattice site s; its half-lattice entry is s >> 1 of the
// field of that parity.
__device__ __forceinline__ unsigned site_parity(unsigned s, const DslashArgs &a) {
  const Site st = site_coords(s, a);
  return (st.t + st.x + st.y + st.z) & 1u;
}

// One lane per FP64 element of the full vector: lane j holds component
// (j % 768) / 64 of site (j / 768) * 64 + j % 64, so every lane issues one
// 16-byte load or store per vector and a wavefront covers one contiguous
// 1 KB plane of a tile (a copy of this form streams at 6.1 TB/s on MI355X,
// against 5.35 TB/s for 12 components per lane).
struct EoElem {
  unsigned s, k;
  bool valid;
};
__device__ __forceinline__ EoElem eo_elem(const EoArgs &e) {
  const unsigned j = blockIdx.x * kBlock + threadIdx.x;
  EoElem q;
  q.s = (j / (12u * kTile)) * kTile + j % kTile;
  q.k = (j % (12u * kTile)) / kTile;
  q.valid = q.s < e.d.V;
  return q;
}
// element index of component k of site s in an FP64 vector
__device__ __forceinline__ size_t eo_at(unsigned k, unsigned s) {
  return (size_t)(s / kTile) * 12 * kTile + (size_t)k * kTile + s % kTile;
}

// full FP64 vector f0 + ids[row] -> half vectors ev[slot], od[slot]
__global__ void __launch_bounds__(kBlock) k_eo_split(EoArgs e, SpinorSet<QCD_HIP_F64> full, int f0,
                                                     SpinorSet<QCD_HIP_F64> ev, SpinorSet<QCD_HIP_F64> od,
                                                     const int *ids) {
  const int b = ids[blockIdx.y];
  const EoElem q = eo_elem(e);
  if (!q.valid) return;
  const SpinorSet<QCD_HIP_F64> &h = site_parity(q.s, e.d) ? od : ev;
  h.base[(size_t)b * h.vstride + eo_at(q.k, q.s >> 1)] = full.base[(size_t)(f0 + b) * full.vstride + eo_at(q.k, q.s)];
}

// x[f0 + slot] += (ev[slot] on even sites, od[slot] on odd sites)
__global__ void __launch_bounds__(kBlock) k_eo_accumulate(EoArgs e, SpinorSet<QCD_HIP_F64> x, int f0,
                                                          SpinorSet<QCD_HIP_F64> ev, SpinorSet<QCD_HIP_F64> od,
                                                          const int *ids) {
  const int b = ids[blockIdx.y];
  const EoElem q = eo_elem(e);
  if (!q.valid) return;
  const SpinorSet<QCD_HIP_F64> &h = site_parity(q.s, e.d) ? od : ev;
  double2 &xv = x.base[(size_t)(f0 + b) * x.vstride + eo_at(q.k, q.s)];
  const double2 d = h.base[(size_t)b * h.vstride + eo_at(q.k, q.s >> 1)];
  xv = make_double2(xv.x + d.x, xv.y + d.y);
}

// out[slot] = M^-1 in[slot] (FP64 half vectors)
__global__ void __launch_bounds__(kBlock) k_eo_minv(EoArgs e, SpinorSet<QCD_HIP_F64> in, SpinorSet<QCD_HIP_F64> out,
                                                    const int *ids) {
  const int b = ids[blockIdx.y];
  const unsigned h = blockIdx.x * kBlock + threadIdx.x;
  if (h >= e.Vh) return;
  Cx<double> v[12];
  in.load(b, h, v);
  eo_mass<true>(v, e);
  out.store(b, h, v);
}

// out[slot] = f_slot * in[slot] with f_slot = scale[slot] (FP64 result)
template <int P>
__global__ void __launch_bounds__(kBlock) k_eo_scale64(EoArgs e, SpinorSet<P> in, SpinorSet<QCD_HIP_F64> out,
                                                       const double *scale, size_t scale_stride, const int *ids) {
  using R = typename Prec<P>::R;
  const int b = ids[blockIdx.y];
  const unsigned h = blockIdx.x * kBlock + threadIdx.x;
  if (h >= e.Vh) return;
  const double f = scale[(size_t)b * scale_stride];
  Cx<R> v[12];
  in.load(b, h, v);
  Cx<double> o[12];
#pragma unroll
  for (int k = 0; k < 12; k++) o[k] = Cx<double>{f * (double)v[k].re, f * (double)v[k].im};
  out.store(b, h, o);
}

// Inner-precision outer step of a mixed even-odd solve with FP32 inner
// vectors, one lane per site: the FP64 residual f0 + slot of the full lattice
// -> its even half ev, its odd half od and tev = M^-1 ev, all three in
// precision P (FP16 solves get the halves from the residual kernel itself)
template <int P>
__global__ void __launch_bounds__(kBlock) k_eo_split_minv(EoArgs e, SpinorSet<QCD_HIP_F64> full, int f0,
                                                          SpinorSet<P> ev, SpinorSet<P> od, SpinorSet<P> tev,
                                                          const int *ids) {
  using R = typename Prec<P>::R;
  const int b = ids[blockIdx.y];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= e.d.V) return;
  Cx<double> v[12];
  full.load(f0 + b, s, v);
  Cx<R> o[12];
#pragma unroll
  for (int k = 0; k < 12; k++) o[k] = Cx<R>{(R)v[k].re, (R)v[k].im};
  if (site_parity(s, e.d)) {
    od.store(b, s >> 1, o);
    return;
  }
  ev.store(b, s >> 1, o);
  eo_mass<true>(v, e);
#pragma unroll
  for (int k = 0; k < 12; k++) o[k] = Cx<R>{(R)v[k].re, (R)v[k].im};
  tev.store(b, s >> 1, o);
}

// x[f0 + slot] += (ev[slot] on even sites, od[slot] on odd sites), the halves
// in precision P, one lane per site, FP64 accumulation
template <int P>
__global__ void __launch_bounds__(kBlock) k_eo_accumulate_p(EoArgs e, SpinorSet<QCD_HIP_F64> x, int f0,
                                                            SpinorSet<P> ev, SpinorSet<P> od, const int *ids) {
  using R = typename Prec<P>::R;
  const int b = ids[blockIdx.y];
  const unsigned s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= e.d.V) return;
  Cx<R> d[12];
  if (site_parity(s, e.d)) od.load(b, s >> 1, d);
  else ev.load(b, s >> 1, d);
  Cx<double> v[12];
  x.load(f0 + b, s, v);
#pragma unroll
  for (int k = 0; k < 12; k++) v[k] = Cx<double>{v[k].re + (double)d[k].re, v[k].im + (double)d[k].im};
  x.store(f0 + b, s, v);
}

// Rows per group of the hop traversal.  4 was within 5% of the best group
// size for every case measured (24^3x48 and 16^3x32; FP64, FP32, FP16; 1, 12
// and 48 rows); 1 and the whole batch were up to 14% slower.
// QCD_HIP_EO_RGROUP overrides it.
static unsigned eo_row_group(qcd_hip_field in, int count) {
  static const long env = [] {
    const char *s = getenv("QCD_HIP_EO_RGROUP");
    return s ? std::atol(s) : 0L;
  }();
  long r = env;
  if (r <= 0) r = 4;
  return (unsigned)std::min<long>(std::max(1L, r), count);
}

EoArgs make_eo_args(qcd_hip_lattice full, double m, double tm, unsigned par) {
  EoArgs e;
  std::memset(&e, 0, sizeof(e));
  e.d = make_args(full, m, tm);
  e.Vh = (unsigned)(full->V / 2);
  e.par = par;
  const double a = m + 4.0, den = a * a + tm * tm;
  e.mi_re = a / den;
  e.mi_im = -tm / den;
  return e;
}

bool eo_supported(qcd_hip_lattice lat) {
  for (int d = 0; d < 4; d++)
    if (lat->L[d] % 2) return false;
  return true;
}

qcd_hip_lattice eo_half_lattice(qcd_hip_lattice full) {
  auto *h = new qcd_hip_lattice_s(*full);
  h->L[0] = full->L[0] / 2;
  h->V = full->V / 2;
  h->stride = round_up(h->V, 64);
  return h;
}

// FP32/FP16 batches of the M^-1 H and M b - H hops run two rhs per lane
// (k_eo_hop2) unless QCD_HIP_EO_TWO=0.
static bool eo_two_rhs() {
  static const bool on = [] {
    const char *s = getenv("QCD_HIP_EO_TWO");
    return !(s && s[0] == '0');
  }();
  return on;
}

// FP16 batches of at least this many rhs run the M^-1 H and M b - H hops on
// the matrix cores (k_eo_hop_mma); QCD_HIP_EO_MMA sets the threshold (0 =
// never).  Default 16: at 24^3x48 the Schur operator takes 43.6 against 44.0
// us per rhs at 16 rhs, 38.6 against 40.3 at 48, and loses below 16 (a group
// of 4 batch rows per wavefront idles lanes on small batches).  A rhs gets the
// same bits in any batch of one kernel: alone and in a batch of 16 or more it
// comes from different kernels.
static int eo_mma_min_rhs() {
  static const int n = [] {
    const char *s = getenv("QCD_HIP_EO_MMA");
    if (!s) return 16;
    const int v = std::atoi(s);
    return v <= 0 ? 1 << 30 : v;
  }();
  return n;
}

// z steps per lane of the z-marching hop (k_eo_hop_zm); 1 runs the one-site
// kernels, 0 lets eo_zlen choose.  QCD_HIP_EO_ZLEN sets the start value,
// qcd_hip_eo_set_zlen changes it.
static std::atomic<int> &eo_zlen_setting() {
  static std::atomic<int> v([] {
    const char *s = getenv("QCD_HIP_EO_ZLEN");
    return s ? std::max(0, std::atoi(s)) : 0;
  }());
  return v;
}

// z for a chunk of zc planes, or 1 when the z-march grid would have more
// workgroups per row than the one-site kernels (the reduction buffers are
// sized for those)
static unsigned eo_zlen_fits(const EoArgs &e, unsigned zc) {
  const unsigned Lz = e.d.L[3];
  const unsigned z = std::min(zc, Lz);
  const unsigned PP = e.d.L[0] * e.d.L[1] * e.d.L[2] / 2;
  if ((size_t)((PP + kBlock - 1) / kBlock) * ((Lz + z - 1) / z) > (size_t)grid_for(e.Vh)) return 1;
  return z;
}

// The z chunk when it cannot be timed (under graph capture), or as set.
// From Schur-operator timings: the z-march gains once a launch has about
// 2^17 (row, half-site) items (24^3x48: one FP64 rhs -24% at 6 steps, one FP32
// rhs -18% at 8, 12 FP32 / FP16 rhs -12% / -3.5% at 8; 48^3x96: one FP64 rhs
// -24% and 12 FP16 rhs -15% at 4-12); with fewer the shorter grid loses
// (16^3x32, one FP64 rhs: +6% at 2 steps); short z extents gain only at 2
// steps (16^3x32, 12 FP16 rhs: -2%).
static unsigned eo_zlen(const EoArgs &e, unsigned rows) {
  const int set = eo_zlen_setting().load();
  unsigned zc = 1;
  if (set > 0) {
    zc = (unsigned)set;
  } else {
    const double items = (double)rows * e.Vh;
    if (items < 131072.0) zc = 1;
    else if (e.d.L[3] <= 16) zc = 2;
    else zc = (unsigned)std::min(8.0, std::max(2.0, std::floor(items / 65536.0)));
  }
  return eo_zlen_fits(e, zc);
}

template <int PS, int PG, int NR>
static int eo_zm_launch(int epi, bool red, EoArgs e, unsigned zlen, unsigned rows, const SpinorIn<PS> &si,
                        const GaugeIn<PG> &gu, const SpinorOut<PS> &so, const SpinorIn<PS> &sb, double *partials,
                        hipStream_t st) {
  const unsigned PP = e.d.L[0] * e.d.L[1] * e.d.L[2] / 2;
  e.d.col_blocks = (PP + kBlock - 1) / kBlock;
  e.d.zlen = zlen;
  e.d.nblk = e.d.col_blocks * ((e.d.L[3] + zlen - 1) / zlen);
  const dim3 gr(e.d.nblk * rows);
#define QH_Z(E, RD) k_eo_hop_zm<PS, PG, E, RD, NR><<<gr, kBlock, 0, st>>>(e, si, gu, so, sb, partials)
  switch (epi * 2 + (red ? 1 : 0)) {
    case 0: QH_Z(EO_MINV, false); break;
    case 2: QH_Z(EO_SCHUR, false); break;
    case 4: QH_Z(EO_SUB, false); break;
    case 5: QH_Z(EO_SUB, true); break;
    case 6: QH_Z(EO_SUB_MINV, false); break;
    default: set_error("eo: unsupported epilogue %d (reduction %d)", epi, (int)red); return 1;
  }
#undef QH_Z
  QH_CHECK(hipGetLastError());
  return 0;
}

// one launch of the hop with z chunks of zlen planes (1: one site per lane);
// two: two rows per lane, rows: rows or row pairs
template <int PS, int PG>
static int eo_launch_z(int epi, bool red, const EoArgs &e, unsigned zlen, bool two, unsigned rows, dim3 gr,
                       const SpinorIn<PS> &si, const GaugeIn<PG> &gu, const SpinorOut<PS> &so, const SpinorIn<PS> &sb,
                       double *partials, hipStream_t st) {
  if (zlen > 1) {
    if constexpr (PS != QCD_HIP_F64) {
      if (two) return eo_zm_launch<PS, PG, 2>(epi, red, e, zlen, rows, si, gu, so, sb, partials, st);
    }
    return eo_zm_launch<PS, PG, 1>(epi, red, e, zlen, rows, si, gu, so, sb, partials, st);
  }
  if constexpr (PS != QCD_HIP_F64) {
    if (two) {
      const dim3 g2(e.d.nblk * rows);
      if (epi == EO_MINV) k_eo_hop2<PS, PG, EO_MINV><<<g2, kBlock, 0, st>>>(e, si, gu, so, sb);
      else k_eo_hop2<PS, PG, EO_SCHUR><<<g2, kBlock, 0, st>>>(e, si, gu, so, sb);
      QH_CHECK(hipGetLastError());
      return 0;
    }
  }
#define QH_E(E, RD) k_eo_hop<PS, PG, E, RD><<<gr, kBlock, 0, st>>>(e, si, gu, so, sb, partials)
  switch (epi * 2 + (red ? 1 : 0)) {
    case 0: QH_E(EO_MINV, false); break;
    case 2: QH_E(EO_SCHUR, false); break;
    case 4: QH_E(EO_SUB, false); break;
    case 5: QH_E(EO_SUB, true); break;
    case 6: QH_E(EO_SUB_MINV, false); break;
    default: set_error("eo: unsupported epilogue %d (reduction %d)", epi, (int)red); return 1;
  }
#undef QH_E
  QH_CHECK(hipGetLastError());
  return 0;
}

// z chunks tuned per (lattice extents, spinor precision, rows per lane, rows
// or row pairs per launch): the first launch of a shape with at least 2^17
// (row, half-site) items that may run twice (not under graph capture, out
// aliasing neither in nor b) times chunks of 1 to 12 planes and keeps the
// fastest; smaller launches keep one site per lane.  All give bit-identical
// results.  Timing rather than a formula: how the grid fills the 256 CUs
// decides (24^3x48, one FP64 rhs: 289 us at 6 steps, 423 at 5, 383 at 1).
struct EoZKey {
  int L[4], ps, nr;
  unsigned rows;
  bool operator<(const EoZKey &o) const {
    return std::tie(L[0], L[1], L[2], L[3], ps, nr, rows) < std::tie(o.L[0], o.L[1], o.L[2], o.L[3], o.ps, o.nr, o.rows);
  }
};
struct EoZVal {
  unsigned zlen, rgroup;
};
static std::map<EoZKey, EoZVal> &eo_ztuned() {
  static std::map<EoZKey, EoZVal> m;
  return m;
}

template <int PS, int PG>
static int eo_launch_pp(int epi, bool red, const EoArgs &e, dim3 gr, qcd_hip_field out, int ro, qcd_hip_field in,
                        int ri, qcd_hip_field g, qcd_hip_field b, int rb, double *partials, hipStream_t st,
                        int *grid_out) {
  const SpinorIn<PS> si = SpinorIn<PS>::of(in, ri);
  const SpinorIn<PS> sb = b ? SpinorIn<PS>::of(b, rb) : si;
  const SpinorOut<PS> so = SpinorOut<PS>::of(out, ro);
  const GaugeIn<PG> gu = GaugeIn<PG>::of(g);
  if constexpr (PS == QCD_HIP_F16 && PG == QCD_HIP_F16) {
    if ((int)e.d.count >= eo_mma_min_rhs() && !red && (epi == EO_MINV || epi == EO_SCHUR))
      return eo_mma_launch(epi, e, out, ro, in, ri, g, b, rb, st);
  }
  const bool two = PS != QCD_HIP_F64 && e.d.count >= 2 && !red && (epi == EO_MINV || epi == EO_SCHUR) &&
                   eo_two_rhs();
  const unsigned rows = two ? (e.d.count + 1) / 2 : e.d.count;
  unsigned zlen = 1;
  EoArgs et = e;   // et.rgroup: tuned with the z chunk
  if (eo_zlen_setting().load() > 0 || (double)rows * e.Vh < 131072.0) {
    zlen = eo_zlen(e, rows);   // as set, or too small a launch to gain from the z-march
  } else {
    std::lock_guard<std::recursive_mutex> lock(state_mutex());
    const EoZKey key{{(int)e.d.L[0], (int)e.d.L[1], (int)e.d.L[2], (int)e.d.L[3]}, PS, two ? 2 : 1, rows};
    auto it = eo_ztuned().find(key);
    if (it != eo_ztuned().end()) {
      zlen = it->second.zlen;
      et.rgroup = it->second.rgroup;
    } else {
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      QH_CHECK(hipStreamIsCapturing(st, &cs));
      if (cs != hipStreamCaptureStatusNone || out == in || out == b) {
        zlen = eo_zlen(e, rows);   // no timing possible: the work-size rule
      } else {
        hipEvent_t e0, e1;
        QH_CHECK(hipEventCreate(&e0));
        QH_CHECK(hipEventCreate(&e1));
        const int reps = (double)rows * e.Vh > 4.0e6 ? 2 : 5;
        float best = 1e30f;
        int rc = 0;
        // batch rows per group (eo_row_of): 4 suits one site per lane; with
        // the z-march larger groups can share the links better (24^3x48, 12
        // rhs: FP64 165 -> 146, FP32 82 -> 75 us per rhs with 12)
        const bool rg_set = getenv("QCD_HIP_EO_RGROUP") != nullptr;   // a set group size is kept
        const unsigned rg4 = rg_set ? e.rgroup : std::min(4u, e.d.count);
        const unsigned rg12 = rg_set ? e.rgroup : std::min(12u, e.d.count);
        for (unsigned z : {1u, 2u, 3u, 4u, 6u, 8u, 12u}) {
          if (z > 1 && eo_zlen_fits(e, z) != z) continue;
          for (unsigned rg : {rg4, rg12}) {
            EoArgs ec = e;
            ec.rgroup = rg;
            rc = eo_launch_z<PS, PG>(epi, red, ec, z, two, rows, gr, si, gu, so, sb, partials, st);
            if (!rc && hipEventRecord(e0, st) != hipSuccess) rc = 2;
            for (int r = 0; r < reps && !rc; r++)
              rc = eo_launch_z<PS, PG>(epi, red, ec, z, two, rows, gr, si, gu, so, sb, partials, st);
            float ms = 0.f;
            if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
              rc = 2;
            if (rc) break;
            if (getenv("QCD_HIP_TUNE_VERBOSE"))
              fprintf(stderr, "eo tune L=%dx%dx%dx%d ps=%d rows/lane=%d rows=%u zlen=%u rgroup=%u: %.2f us\n",
                      key.L[0], key.L[1], key.L[2], key.L[3], PS, key.nr, rows, z, rg, ms * 1e3f / reps);
            if (ms < best) {
              best = ms;
              zlen = z;
              et.rgroup = rg;
            }
            if (rg4 == rg12) break;   // one group size
          }
          if (rc) break;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) {
          if (rc == 2) set_error("eo: timing the z chunks failed: %s", hipGetErrorString(hipGetLastError()));
          return rc;
        }
        eo_ztuned()[key] = EoZVal{zlen, et.rgroup};
      }
    }
  }
  if (zlen > 1 && red && grid_out)
    *grid_out = (int)(((e.d.L[0] * e.d.L[1] * e.d.L[2] / 2 + kBlock - 1) / kBlock) * ((e.d.L[3] + zlen - 1) / zlen));
  return eo_launch_z<PS, PG>(epi, red, et, zlen, two, rows, gr, si, gu, so, sb, partials, st);
}

// Batched EO hop: output sites of parity par; out/in/b are half-lattice
// fields, g the full-lattice gauge field; row y of the batch works on vectors
// ro + ids[y]*out_rs, ri + ids[y]*in_rs, rb + ids[y]*b_rs.  With red, |out|^2
// partials go to partials[y * grid + x] and *grid_out receives grid.
int eo_launch(int epi, bool red, qcd_hip_field out, int ro, int out_rs, qcd_hip_field in, int ri, int in_rs,
              qcd_hip_field g, qcd_hip_field b, int rb, int b_rs, const int *ids, int count, double m, double tm,
              unsigned par, double *partials, hipStream_t st, int *grid_out) {
  EoArgs e = make_eo_args(g->lat, m, tm, par);
  e.d.ids = ids;
  e.d.count = (unsigned)count;
  e.d.in_rs = (unsigned)in_rs;
  e.d.out_rs = (unsigned)out_rs;
  e.d.b_rs = (unsigned)b_rs;
  const int grid = grid_for(e.Vh);
  if (grid_out) *grid_out = grid;   // the z-marching hop sets its own count
  e.d.nblk = (unsigned)grid;
  e.rgroup = eo_row_group(in, count);
  const dim3 gr((unsigned)grid * (unsigned)count);
  const int ps = in->prec, pg = g->prec;
  if (ps == QCD_HIP_F64 && pg == QCD_HIP_F64)
    return eo_launch_pp<QCD_HIP_F64, QCD_HIP_F64>(epi, red, e, gr, out, ro, in, ri, g, b, rb, partials, st, grid_out);
  if (ps == QCD_HIP_F32 && pg == QCD_HIP_F32)
    return eo_launch_pp<QCD_HIP_F32, QCD_HIP_F32>(epi, red, e, gr, out, ro, in, ri, g, b, rb, partials, st, grid_out);
  if (ps == QCD_HIP_F16 && pg == QCD_HIP_F16)
    return eo_launch_pp<QCD_HIP_F16, QCD_HIP_F16>(epi, red, e, gr, out, ro, in, ri, g, b, rb, partials, st, grid_out);
  set_error("eo: unsupported precision pair (spinor %d, gauge %d)", ps, pg);
  return 1;
}

int eo_split(qcd_hip_field full, int f0, qcd_hip_field ev, qcd_hip_field od, const int *ids, int count,
             hipStream_t st) {
  const EoArgs e = make_eo_args(full->lat, 0.0, 0.0, 0);
  k_eo_split<<<dim3(grid_for(12 * round_up(e.d.V, kTile)), count), kBlock, 0, st>>>(e, SpinorSet<QCD_HIP_F64>::of(full), f0,
                                                              SpinorSet<QCD_HIP_F64>::of(ev),
                                                              SpinorSet<QCD_HIP_F64>::of(od), ids);
  QH_CHECK(hipGetLastError());
  return 0;
}

int eo_accumulate(qcd_hip_field x, int f0, qcd_hip_field ev, qcd_hip_field od, const int *ids, int count,
                  hipStream_t st) {
  const EoArgs e = make_eo_args(x->lat, 0.0, 0.0, 0);
  k_eo_accumulate<<<dim3(grid_for(12 * round_up(e.d.V, kTile)), count), kBlock, 0, st>>>(e, SpinorSet<QCD_HIP_F64>::of(x), f0,
                                                                   SpinorSet<QCD_HIP_F64>::of(ev),
                                                                   SpinorSet<QCD_HIP_F64>::of(od), ids);
  QH_CHECK(hipGetLastError());
  return 0;
}

int eo_minv(qcd_hip_lattice full, qcd_hip_field in, qcd_hip_field out, double m, double tm, const int *ids, int count,
            hipStream_t st) {
  const EoArgs e = make_eo_args(full, m, tm, 0);
  k_eo_minv<<<dim3(grid_for(e.Vh), count), kBlock, 0, st>>>(e, SpinorSet<QCD_HIP_F64>::of(in),
                                                             SpinorSet<QCD_HIP_F64>::of(out), ids);
  QH_CHECK(hipGetLastError());
  return 0;
}

int eo_scale64(qcd_hip_lattice full, qcd_hip_field in, qcd_hip_field out, const double *scale, size_t scale_stride,
               const int *ids, int count, hipStream_t st) {
  const EoArgs e = make_eo_args(full, 0.0, 0.0, 0);
  const dim3 gr(grid_for(e.Vh), count);
  switch (in->prec) {
    case QCD_HIP_F64:
      k_eo_scale64<QCD_HIP_F64><<<gr, kBlock, 0, st>>>(e, SpinorSet<QCD_HIP_F64>::of(in),
                                                       SpinorSet<QCD_HIP_F64>::of(out), scale, scale_stride, ids);
      break;
    case QCD_HIP_F32:
      k_eo_scale64<QCD_HIP_F32><<<gr, kBlock, 0, st>>>(e, SpinorSet<QCD_HIP_F32>::of(in),
                                                       SpinorSet<QCD_HIP_F64>::of(out), scale, scale_stride, ids);
      break;
    default:
      k_eo_scale64<QCD_HIP_F16><<<gr, kBlock, 0, st>>>(e, SpinorSet<QCD_HIP_F16>::of(in),
                                                       SpinorSet<QCD_HIP_F64>::of(out), scale, scale_stride, ids);
  }
  QH_CHECK(hipGetLastError());
  return 0;
}

template <int P>
static int eo_mixed_launch(int what, qcd_hip_lattice full, qcd_hip_field a, int f0, qcd_hip_field ev, qcd_hip_field od,
                           qcd_hip_field tev, double m, double tm, const int *ids, int count, hipStream_t st) {
  const EoArgs e = make_eo_args(full, m, tm, 0);
  if (what == 0) {
    k_eo_split_minv<P><<<dim3(grid_for(e.d.V), count), kBlock, 0, st>>>(
        e, SpinorSet<QCD_HIP_F64>::of(a), f0, SpinorSet<P>::of(ev), SpinorSet<P>::of(od), SpinorSet<P>::of(tev), ids);
  } else {
    k_eo_accumulate_p<P><<<dim3(grid_for(e.d.V), count), kBlock, 0, st>>>(e, SpinorSet<QCD_HIP_F64>::of(a), f0,
                                                                         SpinorSet<P>::of(ev), SpinorSet<P>::of(od),
                                                                         ids);
  }
  QH_CHECK(hipGetLastError());
  return 0;
}

static int eo_mixed(int what, qcd_hip_lattice full, qcd_hip_field a, int f0, qcd_hip_field ev, qcd_hip_field od,
                    qcd_hip_field tev, double m, double tm, const int *ids, int count, hipStream_t st) {
  switch (ev->prec) {
    case QCD_HIP_F32:
      return eo_mixed_launch<QCD_HIP_F32>(what, full, a, f0, ev, od, tev, m, tm, ids, count, st);
    case QCD_HIP_F16:
      return eo_mixed_launch<QCD_HIP_F16>(what, full, a, f0, ev, od, tev, m, tm, ids, count, st);
    default:
      set_error("eo: the inner-precision outer step needs FP32 or FP16 half vectors");
      return 1;
  }
}

int eo_split_minv(qcd_hip_field r, int f0, qcd_hip_field ev, qcd_hip_field od, qcd_hip_field tev, double m, double tm,
                  const int *ids, int count, hipStream_t st) {
  return eo_mixed(0, r->lat, r, f0, ev, od, tev, m, tm, ids, count, st);
}

int eo_accumulate_p(qcd_hip_field x, int f0, qcd_hip_field ev, qcd_hip_field od, const int *ids, int count,
                    hipStream_t st) {
  return eo_mixed(2, x->lat, x, f0, ev, od, nullptr, 0.0, 0.0, ids, count, st);
}

}  // namespace qcdhip

using namespace qcdhip;

extern "C" {

int qcd_hip_eo_set_zlen(int zlen) {
  if (zlen < 0) {
    set_error("eo_set_zlen: zlen must be >= 0");
    return 1;
  }
  eo_zlen_setting().store(zlen);
  return 0;
}

int qcd_hip_dslash_schur(qcd_hip_field out, qcd_hip_field in, qcd_hip_field gauge, double m, double tm) {
  QH_TRY(ensure_init());
  if (!out || !in || !gauge || out->kind != QCD_HIP_SPINOR || in->kind != QCD_HIP_SPINOR ||
      gauge->kind != QCD_HIP_GAUGE || out == in) {
    set_error("dslash_schur: expected (spinor out, spinor in, gauge), out not aliasing in");
    return 1;
  }
  const qcd_hip_lattice full = gauge->lat;
  bool ok = eo_supported(full) && in->lat->V * 2 == full->V && out->lat->V == in->lat->V &&
            in->lat->L[0] * 2 == full->L[0] && out->prec == in->prec && gauge->prec == in->prec &&
            out->nrhs == in->nrhs;
  for (int d = 1; d < 4; d++) ok = ok && in->lat->L[d] == full->L[d];
  if (!ok) {
    set_error("dslash_schur: in/out must be fields of the half lattice {T/2, X, Y, Z} of an all-even gauge "
              "lattice, with equal precisions and rhs counts");
    return 1;
  }
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  hipStream_t st = ctx()->stream;
  qcd_hip_field te = nullptr;
  QH_TRY(qcd_hip_spinor_create(&te, in->lat, in->prec, in->nrhs));
  int rc = eo_launch(EO_MINV, false, te, 0, 1, in, 0, 1, gauge, nullptr, 0, 0, nullptr, in->nrhs, m, tm, 0, nullptr,
                     st, nullptr);
  if (!rc)
    rc = eo_launch(EO_SCHUR, false, out, 0, 1, te, 0, 1, gauge, in, 0, 1, nullptr, in->nrhs, m, tm, 1, nullptr, st,
                   nullptr);
  if (!rc && hipStreamSynchronize(st) != hipSuccess) {
    set_error("dslash_schur: %s", hipGetErrorString(hipGetLastError()));
    rc = 2;
  }
  qcd_hip_field_destroy(te);
  return rc;
}

}  // extern "C"
