// mma.hip -- multi-right-hand-side Wilson hops on the matrix cores (MFMA).
//
// For FP16-storage spinors (per-site norm) and FP16 links, batches of
// right-hand sides apply each link to a panel of colour vectors with
// v_mfma_f32_4x4x4_16b_f16: 16 independent 4x4x4 products per instruction,
// one per site.  In the instruction's layout lane 4 b + n holds site b and
// rhs n.  For one hop of site b:
//   A (4 x 4 per block, lane n holds row n): row n of the phased link
//     W = e^{+-i theta/L} U (or U^+), as (re, -im) pairs for the real part of
//     the product and (im, re) pairs for the imaginary part; row 3 is never
//     read back;
//   B (4 x 4 per block, lane n holds column n): the spin-projected and
//     rebuilt neighbour spinor chi_s = ((1 -+ g_mu) psi)_s of rhs n for one
//     spin s, colours 0, 1 in the first k step, colour 2 (and a zero) in the
//     second;
//   D: lane n receives rows 0..2 (colours) of column n: Re or Im of
//     (W chi_s)_c for its own site and rhs.
// So one hop is 4 spins x {Re, Im} x 2 k steps = 16 MFMAs accumulating into
// eight 4-float accumulators that hold the site's whole hopping sum: the
// colour products, the rebuild of the lower spins and the sum over the eight
// hops all run on the matrix cores.
//
// The loads do not use that layout (four rhs, four separate fields, in every
// four adjacent lanes): lanes load 16 consecutive sites of one rhs, project
// and scale there, and hand the six projected half2 of a hop to the MFMA
// layout by ds_bpermute; the hopping sum comes back the same way for a
// coalesced epilogue and store.  A wavefront walks 12 z planes of its 16
// columns with the two previous input planes in registers (mma_march).
//
// Per-site norms: the MFMA inputs are FP16, so each neighbour spinor enters
// scaled by n_nbr / n_ref, with n_ref the largest of the site's eight
// neighbour norms for that rhs (all factors <= 1, no overflow), and the sum is
// multiplied by n_ref at the end.
//
// The same core serves the full-lattice Dslash (D, Q) and the even-odd hops
// (M^-1 H, M b - H) of the batched inner solve.
#include "eo_common.h"

namespace qcdhip {

typedef _Float16 h4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ h2v h2_zero() { return h2v{(_Float16)0.0f, (_Float16)0.0f}; }
__device__ __forceinline__ h4v cat(h2v a, h2v b) { return h4v{a.x, a.y, b.x, b.y}; }

// Per-lane view of one right-hand side of an FP16 spinor field: the lane's
// rhs differs inside a wavefront, so base pointers are per lane.
struct RhsF16 {
  const char *base;   // component group 0, tile 0
  const float *norm;
  // the stored values of site s (its norm is read separately)
  __device__ __forceinline__ void load(unsigned s, __half2 r[12]) const {
    const unsigned off = SpinorIn<QCD_HIP_F16>::offset(s);
#pragma unroll
    for (int g = 0; g < 3; g++) {
      const u32x4 q = *(const u32x4 *)(base + g * 4 * kTile * 4 + off);
#pragma unroll
      for (int i = 0; i < 4; i++) r[4 * g + i] = as_half2(q[i]);
    }
  }
  __device__ __forceinline__ float load_norm(unsigned s) const { return norm[s]; }
};

// ------------------------------------------------------------ hop links
// The A operands depend only on the gauge field and the boundary phases, so
// they are built once per field (and rebuilt when a library call rewrites the
// field or the phases change): for direction mu and hop kind k of site x
// (0: forward, W = e^{+i theta_mu/L_mu} U_mu(x); 1: backward,
// W = e^{-i theta_mu/L_mu} U_mu(x - mu)^+), row i of W as three
// (Re W_ic, -Im W_ic) FP16 pairs, 12 bytes, at byte (hop_record(x) * 3 + i) * 12
// of array 2 mu + k.  Both kinds sit at the site that applies them, so a
// site's eight link rows need only its own index.  The Im operand
// (Im W_ic, Re W_ic) is one packed multiply away.  288 bytes per site.
//
// Records go parity-major when the time extent is even: the sites of parity p
// at records p V/2 + (x >> 1), the even-odd half-lattice order.  An even-odd
// hop (all output sites of one parity) then reads one contiguous run of
// records per block of half columns; in site order it would read every other
// record and fetch the lines of both parities.  The full-lattice hop reads
// two runs, one per parity.
struct HopLinks {
  const char *p[8];   // array 2 mu + k
};
constexpr size_t kHopLinkBytes = 8 * 36;   // per site

// record of site s = t + L0 (x + L1 (y + L2 z)) of parity par
__device__ __forceinline__ unsigned hop_record(unsigned s, unsigned par, const DslashArgs &a) {
  return (a.L[0] & 1u) ? s : par * (a.V / 2) + (s >> 1);
}

template <int PG>
__global__ void __launch_bounds__(kBlock) k_hop_links(DslashArgs a, GaugeIn<PG> U, HopLinks out) {
  const unsigned s = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (s >= a.V) return;
  const int mu = blockIdx.y;
  const Site st = site_coords(s, a);
  const unsigned sb = mu == 0 ? neighbour<0, -1>(st, a)
                      : mu == 1 ? neighbour<1, -1>(st, a)
                      : mu == 2 ? neighbour<2, -1>(st, a)
                                : neighbour<3, -1>(st, a);
  const float pr = (float)a.ph_re[mu], pi = (float)a.ph_im[mu];
#pragma unroll
  for (int kind = 0; kind < 2; kind++) {
    // W = ph U_mu(x) (forward) or conj(ph) U_mu(x - mu)^+ (backward), both
    // stored at x, the site whose hop applies them
    Cx<float> u[9];
    {
      typename Prec<PG>::T r[9];
      U.load_raw(mu, kind ? sb : s, r);
#pragma unroll
      for (int k = 0; k < 9; k++) u[k] = to_cx<float>(r[k]);
    }
    const float qi = kind ? -pi : pi;
    unsigned w[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const Cx<float> x = kind ? Cx<float>{u[3 * c + i].re, -u[3 * c + i].im} : u[3 * i + c];
        const float wr = fmaf(x.re, pr, -(x.im * qi)), wi = fmaf(x.re, qi, x.im * pr);
        w[3 * i + c] = as_u32(__floats2half2_rn(wr, -wi));
      }
    U32x3 *o = (U32x3 *)(const_cast<char *>(out.p[2 * mu + kind]) +
                         (size_t)hop_record(s, (st.t + st.x + st.y + st.z) & 1u, a) * 36);
#pragma unroll
    for (int i = 0; i < 3; i++) o[i] = U32x3{w[3 * i], w[3 * i + 1], w[3 * i + 2]};
  }
}

// The hop links of FP16 gauge field g with the phases of a, built on stream st
// if g changed since they were last built or the phases differ.  Building
// allocates, so it must not happen under graph capture: callers that capture
// (the solver) prepare the links first.
int hop_links_for(qcd_hip_field g, const DslashArgs &a, hipStream_t st, HopLinks *out) {
  std::lock_guard<std::recursive_mutex> lock(state_mutex());
  bool same = g->hop_links && g->hop_version == g->version;
  for (int mu = 0; mu < 4 && same; mu++) same = g->hop_ph[mu][0] == a.ph_re[mu] && g->hop_ph[mu][1] == a.ph_im[mu];
  const size_t per = (size_t)36 * a.V;
  if (!same) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    QH_CHECK(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) {
      set_error("matrix-core hops: the hop links of this gauge field are not built (build them before capture)");
      return 1;
    }
    const size_t bytes = 8 * per;
    if (g->hop_bytes != bytes) {
      dev_free(g->hop_links, g->hop_bytes);
      g->hop_links = nullptr;
      g->hop_bytes = 0;
      QH_TRY(dev_alloc(&g->hop_links, bytes));
      g->hop_bytes = bytes;
    }
    HopLinks hl;
    for (int k = 0; k < 8; k++) hl.p[k] = (const char *)g->hop_links + k * per;
    if (g->prec == QCD_HIP_F16)
      k_hop_links<QCD_HIP_F16><<<dim3(grid_for(a.V), 4), kBlock, 0, st>>>(a, GaugeIn<QCD_HIP_F16>::of(g), hl);
    else {
      set_error("matrix-core hops: FP16 links expected");
      return 1;
    }
    QH_CHECK(hipGetLastError());
    g->hop_version = g->version;
    for (int mu = 0; mu < 4; mu++) {
      g->hop_ph[mu][0] = a.ph_re[mu];
      g->hop_ph[mu][1] = a.ph_im[mu];
    }
  }
  for (int k = 0; k < 8; k++) out->p[k] = (const char *)g->hop_links + k * per;
  return 0;
}

// The two A operands (Re and Im parts) of one hop from the loaded link row q
// (row min(i, 2) of W for link-row lane i: rows of A only reach the same rows
// of D, and row 3 of D is never read, so the padding row may hold anything)
__device__ __forceinline__ void link_operands(h4v are[2], h4v aim[2], const U32x3 &q) {
  const unsigned w[3] = {q.x, q.y, q.z};
  h2v re[3], im[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    re[c] = __builtin_bit_cast(h2v, w[c]);               // (Re W, -Im W)
    im[c] = h2v{-re[c].y, re[c].x};                       // (Im W, Re W)
  }
  are[0] = cat(re[0], re[1]);
  are[1] = cat(re[2], h2_zero());
  aim[0] = cat(im[0], im[1]);
  aim[1] = cat(im[2], h2_zero());
}

// D[s][0] += Re(W chi_s), D[s][1] += Im(W chi_s) for s = 0..3
__device__ __forceinline__ void hop_mma(f4v D[4][2], const h4v are[2], const h4v aim[2], const h2v chi[4][3]) {
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const h4v b0 = cat(chi[s][0], chi[s][1]), b1 = cat(chi[s][2], h2_zero());
    D[s][0] = __builtin_amdgcn_mfma_f32_4x4x4f16(are[0], b0, D[s][0], 0, 0, 0);
    D[s][1] = __builtin_amdgcn_mfma_f32_4x4x4f16(aim[0], b0, D[s][1], 0, 0, 0);
    D[s][0] = __builtin_amdgcn_mfma_f32_4x4x4f16(are[1], b1, D[s][0], 0, 0, 0);
    D[s][1] = __builtin_amdgcn_mfma_f32_4x4x4f16(aim[1], b1, D[s][1], 0, 0, 0);
  }
}

// Upper half spinor of (1 + S g_MU) psi scaled by f (packed FP16): h0, h1 are
// spins 0 and 1; spins 2 and 3 are i^k rotations of them (rebuild4)
template <int MU, int S>
__device__ __forceinline__ void project2(h2v h0[3], h2v h1[3], const __half2 r[12], h2v f) {
  using PJ = Proj<MU>;
  constexpr int sp = S > 0 ? 0 : 2;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    h0[c] = (as_h2v(r[c]) + ipow_h<PJ::A + sp>(as_h2v(r[3 * PJ::P + c]))) * f;
    h1[c] = (as_h2v(r[3 + c]) + ipow_h<PJ::B + sp>(as_h2v(r[3 * PJ::Q + c]))) * f;
  }
}

// chi_s, the four spins of (1 + S g_MU) psi, from its upper half
template <int MU, int S>
__device__ __forceinline__ void rebuild4(h2v chi[4][3], const h2v h0[3], const h2v h1[3]) {
  using PJ = Proj<MU>;
  constexpr int sp = S > 0 ? 0 : 2;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    chi[0][c] = h0[c];
    chi[1][c] = h1[c];
    chi[2][c] = ipow_h<PJ::C + sp>(PJ::R == 0 ? h0[c] : h1[c]);
    chi[3][c] = ipow_h<PJ::D + sp>(PJ::W == 0 ? h0[c] : h1[c]);
  }
}

// ------------------------------------------------------------ lane roles
// A wavefront owns one block of 16 sites x one group of 4 batch rows, and
// every lane plays two roles.  As a load lane, lane 16 n + b reads site b of
// row n: 16 consecutive sites of one row, 256 contiguous bytes per component
// group, so one load instruction covers four runs of whole cache lines.  The
// load lane also reads the neighbour norms and projects and scales the
// neighbour spinors of its (site, row).  As an MFMA lane, lane 4 b + n holds
// site b, row n, the layout the matrix instruction fixes: it reads its site's
// link rows and receives the six projected half2 of each hop from its load
// lane through ds_bpermute.  Loading in the MFMA layout instead would put four
// rows -- four separate fields -- in every four adjacent lanes: four L1 tag
// lookups per lane group and far more address translations.
__device__ __forceinline__ unsigned load_lane_of(unsigned lane) { return 16u * (lane & 3u) + (lane >> 2); }

// value v of the lane whose byte address (4 x lane) is addr
template <typename T>
__device__ __forceinline__ T pull(T v, int addr) {
  static_assert(sizeof(T) == 4, "one dword");
  return __builtin_bit_cast(T, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, v)));
}

// D += W chi for hop (MU, K): forward e^{+i th/L} U_mu(x) (1 - g_mu) psi(x + mu),
// backward e^{-i th/L} U_mu(x - mu)^+ (1 + g_mu) psi(x - mu).  r: the load
// lane's neighbour spinor, f: its norm times 1 / n_ref; w: the MFMA lane's
// link row; addr: byte address of the MFMA lane's load lane
template <int MU, int K>
__device__ __forceinline__ void hop_apply(f4v D[4][2], const __half2 r[12], const U32x3 &w, float f, int addr) {
  constexpr int S = K ? +1 : -1;
  h2v h0[3], h1[3];
  const _Float16 fh = (_Float16)f;
  project2<MU, S>(h0, h1, r, h2v{fh, fh});
#pragma unroll
  for (int c = 0; c < 3; c++) {
    h0[c] = pull(h0[c], addr);
    h1[c] = pull(h1[c], addr);
  }
  h4v are[2], aim[2];
  link_operands(are, aim, w);
  h2v chi[4][3];
  rebuild4<MU, S>(chi, h0, h1);
  hop_mma(D, are, aim, chi);
}

// ------------------------------------------------------------ z-march
// A wavefront owns 16 consecutive (t, x, y) columns -- even-odd: half
// columns, the pairs (2 j, 2 j + 1) of t with x, y fixed -- of one group of 4
// batch rows and steps through zlen consecutive planes z.  Its load lanes keep
// the stored input spinors of their column at planes z - 1 and z (loaded as
// forward z neighbours one and two steps earlier) and load plane z + 1, so
// every input plane of a column is read once per z chunk plus two planes of
// halo.  Plane z - 1 is the backward z neighbour and plane z the centre of the
// full-lattice operator.  Even-odd: the output site of half column j at plane
// z has its z neighbours at half column j of the input field, planes z -+ 1,
// so the same ring serves the hops.  The epilogue runs in the load lanes
// (the hopping sum comes back from the MFMA layout by ds_bpermute), so the
// output stores are as coalesced as the loads.
struct Plane {
  __half2 r[12];
  float n;
};

// Wave -> (group of 4 rhs, column block, z chunk): workgroups remapped per
// XCD.  The rhs groups go in super-groups of sg: within one, the group is the
// fastest index, so the groups of one column block run together on one XCD
// and read the block's links through one L2; then the column blocks of a z
// chunk follow each other, then the z chunks.  sg = all groups shares the
// links most; a smaller sg keeps more columns of fewer rhs in flight, so the
// y neighbours of a plane are read while still in L2.
struct MmaTile {
  unsigned blk, zc, group;   // column block, z chunk, group of 4 batch rows
  bool valid;                // the wave has work
};
__device__ __forceinline__ MmaTile mma_tile(unsigned ncb, unsigned nzc, unsigned count, unsigned sg) {
  const unsigned groups = (count + 3) / 4, nblk = ncb * nzc;
  const unsigned w = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x / 64) + threadIdx.x / 64;
  MmaTile t;
  t.valid = w < nblk * groups;
  t.blk = t.zc = t.group = 0;
  if (!t.valid) return t;
  const unsigned R = sg < 1 || sg > groups ? groups : sg;
  const unsigned full = groups / R, span = nblk * R;
  unsigned rest;
  if (w < full * span) {
    const unsigned q = w / span, r = w - q * span;
    t.group = q * R + r % R;
    rest = r / R;
  } else {
    const unsigned rem = groups - full * R, r = w - full * span;
    t.group = full * R + r % rem;
    rest = r / rem;
  }
  t.zc = rest / ncb;
  t.blk = rest - t.zc * ncb;
  return t;
}

// The z-march body shared by the full-lattice and even-odd kernels.  PP:
// columns per plane (EO: half columns); par: parity of the output sites (EO).
// epi(acc, centre, s, h, view, store) finishes one site in the load lane:
// acc = H psi of its (site, row), centre = input plane z of its column, s =
// full-lattice output site, h = its half-lattice entry (EO), view = the row's
// fields, store = the site and row exist.
template <bool EO, typename Epi>
__device__ __forceinline__ void mma_march(const DslashArgs &a, unsigned par, const SpinorIn<QCD_HIP_F16> &in_all,
                                          const HopLinks &hl, const SpinorOut<QCD_HIP_F16> &out_all,
                                          const SpinorIn<QCD_HIP_F16> &b_all, Epi epi) {
  const unsigned P3 = a.L[0] * a.L[1] * a.L[2], Lz = a.L[3];
  const unsigned PP = EO ? P3 / 2 : P3;
  const MmaTile t = mma_tile(a.nblk, (Lz + a.zlen - 1) / a.zlen, a.count, a.rgroup);
  if (!t.valid) return;   // whole wavefronts only
  const unsigned lane = threadIdx.x % 64;
  // load role: column lane % 16 of the block, row lane / 16 of the group
  // (columns and rows past the end read the last ones and store nothing)
  const unsigned cl0 = 16 * t.blk + (lane & 15u), rl = 4 * t.group + (lane >> 4);
  const bool store = cl0 < PP && rl < a.count;
  const unsigned cl = cl0 < PP ? cl0 : PP - 1;
  const RhsView<QCD_HIP_F16> lv = rhs_view(a, in_all, out_all, b_all, rl < a.count ? rl : a.count - 1);
  const RhsF16 in{(const char *)lv.in.p[0], lv.in.norm};
  Site st;
  {
    const unsigned l0 = EO ? 2 * cl : cl;
    const unsigned r = l0 / a.L[0];
    st.t = l0 - r * a.L[0];
    st.y = r / a.L[1];
    st.x = r - st.y * a.L[1];
  }
  const unsigned t0 = st.t;
  // MFMA role: column lane / 4, row lane % 4
  const unsigned cm0 = 16 * t.blk + (lane >> 2);
  const unsigned cm = cm0 < PP ? cm0 : PP - 1;
  const int addr = (int)(4u * load_lane_of(lane));                      // MFMA lane <- load lane
  const int back = (int)(4u * (4u * (lane & 15u) + (lane >> 4)));     // load lane <- MFMA lane
  const unsigned ii = (lane & 3u) < 3u ? (lane & 3u) : 2u;
  // hop-link record of the MFMA lane's site at plane z: even-odd, the output
  // parity's run (par V/2 + half-lattice index); full lattice, the parity of
  // (t, x, y) of column cm flips with z
  unsigned pcm = 0;
  if constexpr (!EO) {
    const unsigned r = cm / a.L[0], y = r / a.L[1];
    pcm = (cm - r * a.L[0] + (r - y * a.L[1]) + y) & 1u;
  }
  const unsigned z0 = t.zc * a.zlen, z1 = min(z0 + a.zlen, Lz);
  Plane pm, pc;   // input planes z - 1 and z of the load lane's column
  {
    const unsigned zm = z0 == 0 ? Lz - 1 : z0 - 1;
    in.load(cl + PP * zm, pm.r);
    pm.n = in.load_norm(cl + PP * zm);
    in.load(cl + PP * z0, pc.r);
    pc.n = in.load_norm(cl + PP * z0);
  }
  const unsigned sh = EO ? 1 : 0;
  for (unsigned z = z0; z < z1; z++) {
    st.z = z;
    if constexpr (EO) {
      const unsigned odd = (par + st.x + st.y + z) & 1u;
      st.t = t0 + odd;
      st.s = 2 * cl + P3 * z + odd;
    } else {
      st.s = cl + P3 * z;
    }
    const unsigned rec = EO ? par * (a.V / 2) + cm + PP * z : hop_record(cm + P3 * z, (pcm + z) & 1u, a);
    const unsigned ip = cl + PP * (z + 1 == Lz ? 0 : z + 1);
    const unsigned nbr[6] = {neighbour<0, +1>(st, a) >> sh, neighbour<0, -1>(st, a) >> sh,
                             neighbour<1, +1>(st, a) >> sh, neighbour<1, -1>(st, a) >> sh,
                             neighbour<2, +1>(st, a) >> sh, neighbour<2, -1>(st, a) >> sh};
    // every load of the step is issued before the first product (splitting
    // the step into two halves whose loads fly one half ahead needs about 300
    // registers: one wavefront per SIMD, measured slower)
    float n[7];
#pragma unroll
    for (int k = 0; k < 6; k++) n[k] = in.load_norm(nbr[k]);
    n[6] = in.load_norm(ip);
    __half2 r[7][12];
    U32x3 w[8];
#pragma unroll
    for (int k = 0; k < 6; k++) in.load(nbr[k], r[k]);
    in.load(ip, r[6]);
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = *(const U32x3 *)(hl.p[k] + (rec * 36u + ii * 12u));   // below 2^32 for V <= 2^24
    __builtin_amdgcn_sched_barrier(0);   // keep the scheduler from sinking the loads
    float nref = pm.n;
#pragma unroll
    for (int k = 0; k < 7; k++) nref = fmaxf(nref, n[k]);
    // the scale factors cancel against nref below: an approximate reciprocal
    // only moves each factor by an ulp of FP32
    const float inv_ref = nref > 0.f ? __builtin_amdgcn_rcpf(nref) : 0.f;
    f4v D[4][2];
#pragma unroll
    for (int q = 0; q < 4; q++) D[q][0] = D[q][1] = f4v{0.f, 0.f, 0.f, 0.f};
    hop_apply<0, 0>(D, r[0], w[0], n[0] * inv_ref, addr);
    hop_apply<0, 1>(D, r[1], w[1], n[1] * inv_ref, addr);
    hop_apply<1, 0>(D, r[2], w[2], n[2] * inv_ref, addr);
    hop_apply<1, 1>(D, r[3], w[3], n[3] * inv_ref, addr);
    hop_apply<2, 0>(D, r[4], w[4], n[4] * inv_ref, addr);
    hop_apply<2, 1>(D, r[5], w[5], n[5] * inv_ref, addr);
    hop_apply<3, 0>(D, r[6], w[6], n[6] * inv_ref, addr);
    hop_apply<3, 1>(D, pm.r, w[7], pm.n * inv_ref, addr);
    Cx<float> acc[4][3];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int c = 0; c < 3; c++) acc[q][c] = Cx<float>{pull(D[q][0][c], back) * nref, pull(D[q][1][c], back) * nref};
    epi(acc, pc, st.s, cl + PP * z, lv, store);
    pm = pc;
#pragma unroll
    for (int k = 0; k < 12; k++) pc.r[k] = r[6][k];
    pc.n = n[6];
  }
}

// Full-lattice D or Q over a batch (FP16 spinors and links).
template <int EPI>
__global__ void __launch_bounds__(kBlock) k_dslash_mma(DslashArgs a, SpinorIn<QCD_HIP_F16> in_all, HopLinks hl,
                                                       SpinorOut<QCD_HIP_F16> out_all, SpinorIn<QCD_HIP_F16> b_all) {
  const float mp4 = (float)a.mp4, tm = (float)a.tm;
  mma_march<false>(a, 0u, in_all, hl, out_all, b_all,
                   [&](const Cx<float> acc[4][3], const Plane &centre, unsigned s, unsigned, const RhsView<QCD_HIP_F16> &v,
                       bool store) {
                     // out = (m + 4 + i tm g5) psi - 1/2 H psi
                     Cx<float> psi[12];
                     SpinorIn<QCD_HIP_F16>::convert(centre.r, centre.n, psi);
                     Cx<float> o[12];
#pragma unroll
                     for (int sp = 0; sp < 4; sp++) {
                       const float g = sp < 2 ? tm : -tm;
#pragma unroll
                       for (int c = 0; c < 3; c++) {
                         const Cx<float> x = psi[3 * sp + c];
                         o[3 * sp + c] = Cx<float>{fmaf(-0.5f, acc[sp][c].re, fmaf(mp4, x.re, -(g * x.im))),
                                                   fmaf(-0.5f, acc[sp][c].im, fmaf(mp4, x.im, g * x.re))};
                       }
                     }
                     apply_epilogue<EPI>(o, v.b, s);
                     if (store) v.out.template store<true>(s, o);
                   });
}

// Even-odd hop of an FP16 batch: output sites of parity e.par, neighbours
// read from the half field of the other parity; epilogue M^-1 H in (EO_MINV)
// or M b - H in (EO_SCHUR).
template <int EPI>
__global__ void __launch_bounds__(kBlock) k_eo_hop_mma(EoArgs e, SpinorIn<QCD_HIP_F16> in_all, HopLinks hl,
                                                       SpinorOut<QCD_HIP_F16> out_all, SpinorIn<QCD_HIP_F16> b_all) {
  mma_march<true>(e.d, e.par, in_all, hl, out_all, b_all,
                  [&](const Cx<float> acc[4][3], const Plane &, unsigned, unsigned h, const RhsView<QCD_HIP_F16> &v,
                      bool store) {
                    Cx<float> o[12];
                    eo_epilogue<EPI>(o, acc, e, v.b, h);
                    if (store) v.out.template store<true>(h, o);
                  });
}

// build the hop links of FP16 gauge field g for the phases of lattice lat
// (outside graph capture); no-op when they are current
int mma_prepare(qcd_hip_field g, qcd_hip_lattice lat, hipStream_t st) {
  if (!g || g->prec != QCD_HIP_F16) return 0;
  const DslashArgs a = make_args(lat, 0.0, 0.0);
  HopLinks hl;
  return hop_links_for(g, a, st, &hl);
}

// rhs groups per super-group (mma_tile); QCD_HIP_MMA_RGROUP overrides the
// default (0 = all groups)
static unsigned mma_rgroup() {
  static const unsigned r = [] {
    const char *e = getenv("QCD_HIP_MMA_RGROUP");
    return e ? (unsigned)std::atoi(e) : 3u;
  }();
  return r;
}

// z steps per wavefront (mma_march); QCD_HIP_MMA_ZLEN overrides the default
static unsigned mma_zlen(unsigned Lz) {
  static const unsigned r = [] {
    const char *e = getenv("QCD_HIP_MMA_ZLEN");
    return e ? (unsigned)std::atoi(e) : 12u;
  }();
  return r < 1 || r > Lz ? Lz : r;
}

// the z-march fields of a for PP columns per plane; returns the grid:
// workgroups of kBlock / 64 wavefronts over column blocks x z chunks x rhs
// groups
static unsigned mma_setup(DslashArgs &a, unsigned PP) {
  a.nblk = (PP + 15) / 16;
  a.zlen = mma_zlen(a.L[3]);
  a.rgroup = mma_rgroup();
  const unsigned nzc = (a.L[3] + a.zlen - 1) / a.zlen;
  const unsigned waves = a.nblk * nzc * ((a.count + 3) / 4), per = kBlock / 64;
  return (waves + per - 1) / per;
}

int dslash_mma_launch(int epi, DslashArgs a, qcd_hip_field out, int ro, qcd_hip_field in, int ri, qcd_hip_field g,
                      qcd_hip_field b, int rb, hipStream_t st) {
  if (in->prec != QCD_HIP_F16 || g->prec != QCD_HIP_F16 || (epi != EPI_D && epi != EPI_Q)) {
    set_error("dslash: the matrix-core traversal needs FP16 spinors and links and the D or Q epilogue");
    return 1;
  }
  const unsigned grid = mma_setup(a, a.L[0] * a.L[1] * a.L[2]);
  const SpinorIn<QCD_HIP_F16> si = SpinorIn<QCD_HIP_F16>::of(in, ri);
  const SpinorIn<QCD_HIP_F16> sb = b ? SpinorIn<QCD_HIP_F16>::of(b, rb) : si;
  const SpinorOut<QCD_HIP_F16> so = SpinorOut<QCD_HIP_F16>::of(out, ro);
  HopLinks hl;
  QH_TRY(hop_links_for(g, a, st, &hl));
  if (epi == EPI_D) k_dslash_mma<EPI_D><<<grid, kBlock, 0, st>>>(a, si, hl, so, sb);
  else k_dslash_mma<EPI_Q><<<grid, kBlock, 0, st>>>(a, si, hl, so, sb);
  QH_CHECK(hipGetLastError());
  return 0;
}

int eo_mma_launch(int epi, const EoArgs &e0, qcd_hip_field out, int ro, qcd_hip_field in, int ri, qcd_hip_field g,
                  qcd_hip_field b, int rb, hipStream_t st) {
  if (in->prec != QCD_HIP_F16 || g->prec != QCD_HIP_F16 || (epi != EO_MINV && epi != EO_SCHUR)) {
    set_error("eo: the matrix-core hop needs FP16 spinors and links and the M^-1 H or M b - H epilogue");
    return 1;
  }
  EoArgs e = e0;
  if (e.d.L[0] % 2) {
    set_error("eo: the matrix-core hop needs an even time extent");
    return 1;
  }
  const unsigned grid = mma_setup(e.d, e.d.L[0] * e.d.L[1] * e.d.L[2] / 2);
  const SpinorIn<QCD_HIP_F16> si = SpinorIn<QCD_HIP_F16>::of(in, ri);
  const SpinorIn<QCD_HIP_F16> sb = b ? SpinorIn<QCD_HIP_F16>::of(b, rb) : si;
  const SpinorOut<QCD_HIP_F16> so = SpinorOut<QCD_HIP_F16>::of(out, ro);
  HopLinks hl;
  QH_TRY(hop_links_for(g, e.d, st, &hl));
  if (epi == EO_MINV) k_eo_hop_mma<EO_MINV><<<grid, kBlock, 0, st>>>(e, si, hl, so, sb);
  else k_eo_hop_mma<EO_SCHUR><<<grid, kBlock, 0, st>>>(e, si, hl, so, sb);
  QH_CHECK(hipGetLastError());
  return 0;
}

}  // namespace qcdhip
