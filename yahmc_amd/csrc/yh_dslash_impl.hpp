// Hopping term with fused epilogues (library/dirac.cpp:227-704, 706-781).
//
//   out(x) = scale * sum_mu [ (1 - g_mu) U(x,mu) psi(x+mu)
//                           + (1 + g_mu) U(x-mu,mu)^+ psi(x-mu) ]
//
// One thread per output site of one parity; a wave covers 64 consecutive
// EO sites so every neighbour / link / clover component load is a
// contiguous 512 B - 1 KB wave access.  The spin projection keeps two of
// the four spin rows (2 * D complex) per direction, the link multiplies
// the projected half spinor, and the other two rows are rebuilt with a
// multiplication by one of {+-1, +-i}.  Projection codes per direction
// (codes: 0 = +1, 1 = -1, 2 = +i, 3 = -i) follow from the chiral gamma
// basis of library/spinor.cpp:191-264:
//   h_k = psi_{2+k} + PC_k * psi_{PI_k};   r = U h (or U^+ h)
//   acc_{2+k} += r_k;  acc_k += RC_k * r_{RI_k}
// Boundary conditions are applied in-kernel: antiperiodic flips the sign
// of U_0 on links leaving t = 0 (library/bcs.cpp:264-272), open / SF zero
// the output on the boundary slices (library/bcs.cpp:139-164).
//
// Epilogues:
//   PLAIN    out = scale * sum                           (Dphi_eo / Dphi_oe)
//   CLOVINV  h = scale * sum; out = D_oo^-1 h; aux = h   (first half of Dphi)
//   CLOV     out = scale * sum + D_ee self [; gamma5] [; |out|^2]
//                                                        (second half of Dphi)
#pragma once
#include "yh_internal.hpp"
#include "yh_reduce.hpp"

namespace yh {

// threads per workgroup of the hop (one output site each)
#ifndef YH_HOP_BLOCK
#define YH_HOP_BLOCK 256
#endif

template <int MU, bool FWD>
struct Proj;
template <> struct Proj<0, true> { static constexpr int PI0 = 0, PI1 = 1, PC0 = 0, PC1 = 0, RI0 = 0, RI1 = 1, RC0 = 0, RC1 = 0; };
template <> struct Proj<0, false> { static constexpr int PI0 = 0, PI1 = 1, PC0 = 1, PC1 = 1, RI0 = 0, RI1 = 1, RC0 = 1, RC1 = 1; };
template <> struct Proj<1, true> { static constexpr int PI0 = 1, PI1 = 0, PC0 = 3, PC1 = 3, RI0 = 1, RI1 = 0, RC0 = 2, RC1 = 2; };
template <> struct Proj<1, false> { static constexpr int PI0 = 1, PI1 = 0, PC0 = 2, PC1 = 2, RI0 = 1, RI1 = 0, RC0 = 3, RC1 = 3; };
template <> struct Proj<2, true> { static constexpr int PI0 = 1, PI1 = 0, PC0 = 1, PC1 = 0, RI0 = 1, RI1 = 0, RC0 = 0, RC1 = 1; };
template <> struct Proj<2, false> { static constexpr int PI0 = 1, PI1 = 0, PC0 = 0, PC1 = 1, RI0 = 1, RI1 = 0, RC0 = 1, RC1 = 0; };
template <> struct Proj<3, true> { static constexpr int PI0 = 0, PI1 = 1, PC0 = 3, PC1 = 2, RI0 = 0, RI1 = 1, RC0 = 2, RC1 = 3; };
template <> struct Proj<3, false> { static constexpr int PI0 = 0, PI1 = 1, PC0 = 2, PC1 = 3, RI0 = 0, RI1 = 1, RC0 = 3, RC1 = 2; };

struct HopArgs {
  const void* in;
  const float* in_nrm;
  void* out;
  float* out_nrm;
  void* aux;
  float* aux_nrm;
  const void* self;
  const float* self_nrm;
  const void* lf;   // links of the output parity, [mu][vec][site]
  const void* lb;   // links of the input parity
  const void* clov; // packed clover (CLOV: term of the output parity; CLOVINV: inverse)
  const float* clov_nrm;  // 16-bit clover: block scales [block][site] of the same sites
  double scale, diag;
  int parity, gamma5, zero_self_bc, reduce;
  double* partials;
  unsigned* ticket;
  double* red_out;
  const int* skip;
  Geom g;
};

// one direction of the hop: the neighbour spinor is element `nidx` of `in`
// (component stride `stride`: the input field, or its copy staged in LDS),
// the link is site `lidx` of the link field
template <int MU, bool FWD, int D, int PREC, int RECON, typename R>
__device__ __forceinline__ void hop_dir_src(Cx<R> (&acc)[4][D], const HopArgs& a, const void* in, const float* in_nrm,
                                            int stride, int nidx, int lidx, bool flip) {
  typedef Proj<MU, FWD> P;
  constexpr int LNV = LinkStore<PREC, RECON, D>::NV;
  typedef typename LinkStore<PREC, RECON, D>::vec LV;
  Cx<R> s[4][D];
  R sc = (R)1;
  if constexpr (PREC == PREC_F16) sc = load_spinor_q<D, R>(s, in, in_nrm, nidx, stride);
  else load_spinor<PREC, D, R>(s, in, in_nrm, nidx, stride);
  Cx<R> h[2][D];
#pragma unroll
  for (int i = 0; i < D; i++) {
    h[0][i] = s[2][i] + unit_mul<P::PC0>(s[P::PI0][i]);
    h[1][i] = s[3][i] + unit_mul<P::PC1>(s[P::PI1][i]);
  }
  const LV* lbase = (const LV*)(FWD ? a.lf : a.lb) + (size_t)MU * LNV * a.g.VHpad;
  Cx<R> r[2][D];
  if constexpr (PREC == PREC_F64 && RECON == RECON_18 && D >= 6) {
    // a D_R x D_R FP64 link is 2 D_R^2 doubles (144 registers at D_R = 6,
    // 256 at 8): it is read in NP parts of RP rows (forward: U h) or columns
    // (backward: U^+ h), three at D_R = 6 and two at 8, each part giving RP
    // complete output elements that go straight into the accumulator; the
    // next part's loads wait (opaque zero) for that, so one part is live at
    // a time.  Every output element sums over j in the order of the
    // one-piece product: bit-identical.
    constexpr int NP = D >= 8 ? 4 : 2, RP = D / NP;
    const R sgn = flip ? -(R)1 : (R)1, nsgn = -sgn;
    // the first part waits for the projected half spinor, so the full
    // neighbour spinor is not live beside the link
#pragma unroll
    for (int i = 0; i < D; i++) {
      __asm__ volatile("" : "+v"(h[0][i].re), "+v"(h[0][i].im));
      __asm__ volatile("" : "+v"(h[1][i].re), "+v"(h[1][i].im));
    }
    int dep;
    __asm__ volatile("v_mov_b32 %0, 0" : "=v"(dep));
#pragma unroll
    for (int part = 0; part < NP; part++) {
      Cx<R> u[RP][D];
#pragma unroll
      for (int k = 0; k < RP * D; k++) {
        const int o = part * RP + k / D, j = k % D;  // output element o, summed index j
        const double2 v = ld_stream(&lbase[(size_t)(FWD ? o * D + j : j * D + o) * a.g.VHpad + lidx + dep]);
        u[k / D][j] = Cx<R>((R)v.x, (R)v.y);
      }
#pragma unroll
      for (int l = 0; l < RP; l++) {
        const int o = part * RP + l;
        Cx<R> r0((R)0, (R)0), r1((R)0, (R)0);
#pragma unroll
        for (int j = 0; j < D; j++) {
          if constexpr (FWD) {
            cmac(r0, u[l][j], h[0][j]);
            cmac(r1, u[l][j], h[1][j]);
          } else {
            cmac_conj(r0, u[l][j], h[0][j]);
            cmac_conj(r1, u[l][j], h[1][j]);
          }
        }
        r[0][o] = r0;
        r[1][o] = r1;
        unit_fma<0>(acc[2][o], sgn, nsgn, r0);
        unit_fma<0>(acc[3][o], sgn, nsgn, r1);
        unit_fma<P::RC0>(acc[0][o], sgn, nsgn, r[P::RI0][o]);
        unit_fma<P::RC1>(acc[1][o], sgn, nsgn, r[P::RI1][o]);
      }
      if (part + 1 < NP) {
#pragma unroll
        for (int l = 0; l < RP; l++)
#pragma unroll
          for (int sp = 0; sp < 4; sp++) __asm__ volatile("" : "+v"(acc[sp][part * RP + l].re), "+v"(acc[sp][part * RP + l].im));
        __asm__ volatile("v_mov_b32 %0, 0" : "=v"(dep));
      }
    }
    return;
  } else {
    Cx<R> u[D][D];
    // FP64 hops stream the links (each is read once per hop) with
    // non-temporal loads and write their output with non-temporal stores:
    // measured faster for FP64 (326 -> 301 us, 18 reals, 32^4), neutral for
    // FP32 and slower for FP16, which keep ordinary accesses
    load_link<PREC, RECON, D, R, PREC == PREC_F64, PREC == PREC_F16>(u, lbase, lidx, a.g.VHpad);
    if constexpr (PREC == PREC_F16) sc *= (R)Q_UNIT;  // the links were loaded as 32768 U
#pragma unroll
    for (int i = 0; i < D; i++) {
      Cx<R> r0((R)0, (R)0), r1((R)0, (R)0);
#pragma unroll
      for (int j = 0; j < D; j++) {
        if constexpr (FWD) {
          cmac(r0, u[i][j], h[0][j]);
          cmac(r1, u[i][j], h[1][j]);
        } else {
          cmac_conj(r0, u[j][i], h[0][j]);
          cmac_conj(r1, u[j][i], h[1][j]);
        }
      }
      r[0][i] = r0;
      r[1][i] = r1;
    }
  }
  // the antiperiodic sign and the 16-bit scales multiply the accumulation
  // (a branch here splits the kernel into blocks and serialises the
  // neighbour loads of the t direction); folded into its FMAs they cost no
  // separate multiplies
  const R sgn = flip ? -sc : sc;
  const R nsgn = -sgn;
#pragma unroll
  for (int i = 0; i < D; i++) {
    unit_fma<0>(acc[2][i], sgn, nsgn, r[0][i]);
    unit_fma<0>(acc[3][i], sgn, nsgn, r[1][i]);
    unit_fma<P::RC0>(acc[0][i], sgn, nsgn, r[P::RI0][i]);
    unit_fma<P::RC1>(acc[1][i], sgn, nsgn, r[P::RI1][i]);
  }
}

template <int MU, bool FWD, int D, int PREC, int RECON, typename R>
__device__ __forceinline__ void hop_dir(Cx<R> (&acc)[4][D], const HopArgs& a, int nidx, int lidx, bool flip) {
  hop_dir_src<MU, FWD, D, PREC, RECON, R>(acc, a, a.in, a.in_nrm, a.g.VHpad, nidx, lidx, flip);
}

// LDS staging of the neighbour spinors a workgroup shares.  The 256 output
// sites of a workgroup are consecutive in the EO order (z fastest) and, when
// a (t, x) plane holds a multiple of 256 sites, lie in one plane.  The input
// sites of their y and z neighbours are then (a) within H = L_Z / 2 + 1 of
// the workgroup's own index range, or (b) across the y wrap: row L_Y - 1 of
// the plane for the row y = 0 and row 0 for the row y = L_Y - 1.  The
// workgroup copies the window [k0 - H, k0 + 256 + H) and the two wrap rows of
// the input field into LDS once (global -> LDS, no registers), and its y and
// z neighbours (four reads per staged site) come from there; t and x
// neighbours, which no other site of the workgroup uses, come from the
// field.  Layout per component plane (stride NT): the window, row L_Y - 1,
// row 0.
template <int PREC, int D>
struct Stage {
  typedef typename SpinorStore<PREC, D>::vec V;
  static constexpr int NV = SpinorStore<PREC, D>::NV;
  static __host__ __device__ int halo(const Geom& g) { return g.L[3] / 2 + 1; }
  static __host__ __device__ int window(const Geom& g) { return YH_HOP_BLOCK + 2 * halo(g); }
  static __host__ __device__ int stride(const Geom& g) { return window(g) + g.L[3]; }
  static __host__ __device__ size_t bytes(const Geom& g) {
    return (size_t)stride(g) * (NV * sizeof(V) + (PREC == PREC_F16 ? sizeof(float) : 0));
  }
  // workgroups never straddle a (t, x) plane, every thread has a site, and
  // a row fits one 64-site chunk
  static __host__ bool usable(const Geom& g) {
    const long plane = (long)g.L[2] * g.L[3] / 2;
    return g.VH % YH_HOP_BLOCK == 0 && plane % YH_HOP_BLOCK == 0 && g.L[3] <= 128 && bytes(g) <= 64 * 1024;
  }
};

// Issue the copy: one 64-site chunk of one component plane per wave
// instruction, each lane moving one 16-byte element to LDS base + lane * 16.
// Lanes of sites outside the field load nothing (never read).
template <int PREC, int D>
__device__ __forceinline__ void stage_issue(unsigned char* lds_bytes, const void* in, const float* in_nrm, int k0,
                                            const Geom& g) {
  typedef Stage<PREC, D> St;
  typedef typename St::V V;
  static_assert(sizeof(V) == 16, "spinor component vectors are 16 bytes");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int NW = St::window(g), NT = St::stride(g), hz = g.L[3] / 2;
  const int s0 = k0 - St::halo(g);
  const int plane = g.L[2] * hz;
  const int p0 = k0 / plane * plane;
  const int nchw = (NW + 63) / 64;  // window chunks (the last may be partial)
  const int nch = nchw + 2;          // + the two wrap rows
  // source site of LDS slot (chunk c, lane); -1: none
  auto src_site = [&](int c) -> int {
    if (c < nchw) {
      const int j = c * 64 + lane;
      const int site = s0 + j;
      return (j < NW && site >= 0 && site < g.VH) ? site : -1;
    }
    if (lane >= hz) return -1;
    return c == nchw ? p0 + (g.L[2] - 1) * hz + lane : p0 + lane;
  };
  auto slot = [&](int c) -> int { return c < nchw ? c * 64 : NW + (c - nchw) * hz; };
  for (int cidx = wave; cidx < nch * St::NV; cidx += YH_HOP_BLOCK / 64) {
    const int v = cidx / nch, c = cidx - v * nch;
    const int site = src_site(c);
    __attribute__((address_space(3))) void* dst =
        (__attribute__((address_space(3))) void*)(lds_bytes + ((size_t)v * NT + slot(c)) * sizeof(V));
    if (site >= 0)
      __builtin_amdgcn_global_load_lds((const void*)((const V*)in + (size_t)v * g.VHpad + site), dst, 16, 0, 0);
  }
  if constexpr (PREC == PREC_F16) {
    unsigned char* nb = lds_bytes + (size_t)NT * St::NV * sizeof(V);
    for (int c = wave; c < nch; c += YH_HOP_BLOCK / 64) {
      const int site = src_site(c);
      __attribute__((address_space(3))) void* dst =
          (__attribute__((address_space(3))) void*)(nb + (size_t)slot(c) * sizeof(float));
      if (site >= 0) __builtin_amdgcn_global_load_lds((const void*)(in_nrm + site), dst, 4, 0, 0);
    }
  }
}

// minimum waves per SIMD requested from the compiler per precision (caps
// the VGPR budget: 512 / waves).  A sweep on MI355X (32^4) found every
// tighter bound slower than none: the extra waves do not pay for the spills.
#ifndef YH_MINW_F64
#define YH_MINW_F64 1
#endif
#ifndef YH_MINW_F32
#define YH_MINW_F32 1
#endif
#ifndef YH_MINW_F16
#define YH_MINW_F16 1
#endif
template <int PREC>
constexpr int hop_min_waves() {
  return PREC == PREC_F64 ? YH_MINW_F64 : PREC == PREC_F32 ? YH_MINW_F32 : YH_MINW_F16;
}

// scheduling groups: bit m - 1 of the mask ends a group after direction m
// (t = 1, x = 2, y = 3; 0 = one group, 7 = one group per direction).  The
// neighbour indices of a group are made to depend, via an opaque zero, on
// the accumulated result of the previous group, so the compiler cannot
// hoist the loads of all eight neighbours at once (they are loads of
// read-only kernel data and would otherwise move freely) and the registers
// of a finished group are free again.  Defaults from a sweep of all masks on
// MI355X (SU(3), 32^4, GB/s by algorithmic bytes): FP64-18 t|xy|z 5763
// (one group: 2670, spilling), FP64-12 t|x|y|z 5374 (4771), FP32-12 t|xyz
// 5174 (5111), FP16-12 txy|z 4302 (3998).  The 16-bit kernels were swept
// again by the time of the 16-bit mixed CG at 32^4, which runs the
// clover-epilogue variants: t x|y|z 0.2203-0.2210 s against 0.2221-0.2226 s
// for txy|z (one group 0.2394, t|x y|z 0.2251, one per direction 0.2246).
#ifndef YH_SCHED_F64_18
#define YH_SCHED_F64_18 5
#endif
#ifndef YH_SCHED_F64_12
#define YH_SCHED_F64_12 7
#endif
#ifndef YH_SCHED_F32
#define YH_SCHED_F32 1
#endif
#ifndef YH_SCHED_F16
#define YH_SCHED_F16 6
#endif
// representations with D >= 4: bit 3 also splits each direction between its
// forward and backward hop (one neighbour spinor and link in flight)
#ifndef YH_SCHED_WIDE
#define YH_SCHED_WIDE 63
#endif
// FP32 hops with LDS-staged y / z neighbours: t x | y | z (their own sweep:
// 57.0 against 59.2 us unstaged at 32^4, FP32-12)
#ifndef YH_SCHED_F32_STG
#define YH_SCHED_F32_STG 6
#endif
template <int PREC, int RECON, int D, bool STG = false>
constexpr int hop_sched() {
  return D >= 4 ? YH_SCHED_WIDE
                : PREC == PREC_F64 ? (RECON == RECON_18 ? YH_SCHED_F64_18 : YH_SCHED_F64_12)
                                   : PREC == PREC_F32 ? (STG ? YH_SCHED_F32_STG : YH_SCHED_F32) : YH_SCHED_F16;
}
// an opaque zero after every accumulator is final, if bit B of the mask is set
template <int PREC, int RECON, int B, int D, typename R, bool STG = false>
__device__ __forceinline__ int hop_group_dep(Cx<R> (&acc)[4][D]) {
  constexpr int s = hop_sched<PREC, RECON, D, STG>();
  if constexpr ((s >> B) & 1) {
    // every accumulator passes through an (empty) volatile asm, so the whole
    // group's arithmetic completes here and its loaded operands die
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int i = 0; i < D; i++) {
        __asm__ volatile("" : "+v"(acc[sp][i].re));
        __asm__ volatile("" : "+v"(acc[sp][i].im));
      }
    int z;
    __asm__ volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
  } else {
    return 0;
  }
}

template <int D, int PREC, int RECON, int EPI, bool STG = false>
__global__ void __launch_bounds__(YH_HOP_BLOCK, hop_min_waves<PREC>()) k_hop(HopArgs a) {
  typedef typename SpinorStore<PREC, D>::real R;
  if (a.skip && *a.skip) return;
  const int nblk = gridDim.x;
  const int blk = xcd_block(nblk, a.g.xq);
  const int k = blk * YH_HOP_BLOCK + threadIdx.x;
  const Geom& g = a.g;
  double nrm2 = 0.0;
  // staged input (STG): per component plane NT slots, the window from s0
  // then the wrap rows L_Y - 1 and 0 of the workgroup's plane
  extern __shared__ __align__(16) unsigned char hop_lds[];
  typedef Stage<PREC, D> St;
  const int NT = STG ? St::stride(g) : 0;
  const int s0 = blk * YH_HOP_BLOCK - (STG ? St::halo(g) : 0);
  const typename St::V* lds = (const typename St::V*)hop_lds;
  const float* lds_nrm = (const float*)(hop_lds + (size_t)NT * St::NV * sizeof(typename St::V));
  // the copy is issued now and waited for after the t and x directions
  if constexpr (STG) stage_issue<PREC, D>(hop_lds, a.in, a.in_nrm, blk * YH_HOP_BLOCK, g);
  if (k < g.VH) {
    int c[4];
    site_coords(k, a.parity, g, c);
    const int t = c[0], x = c[1], y = c[2], z = c[3];
    const int T = g.L[0];
    Cx<R> acc[4][D];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < D; i++) acc[s][i] = Cx<R>((R)0, (R)0);
    const bool ap = g.bc == BC_ANTIPERIODIC;
    // t
    {
      int tf = (t + 1 == T) ? 0 : t + 1;
      int tb = (t == 0) ? T - 1 : t - 1;
      int nf = half_index(tf, x, y, z, g), nb = half_index(tb, x, y, z, g);
      hop_dir<0, true, D, PREC, RECON, R>(acc, a, nf, k, ap && t == 0);
      const int h = hop_group_dep<PREC, RECON, 3, D, R, STG>(acc);
      hop_dir<0, false, D, PREC, RECON, R>(acc, a, nb + h, nb + h, ap && tb == 0);
    }
    const int dep1 = hop_group_dep<PREC, RECON, 0, D, R, STG>(acc);
    {
      int xf = (x + 1 == g.L[1]) ? 0 : x + 1;
      int xb = (x == 0) ? g.L[1] - 1 : x - 1;
      int nf = half_index(t, xf, y, z, g) + dep1, nb = half_index(t, xb, y, z, g) + dep1;
      hop_dir<1, true, D, PREC, RECON, R>(acc, a, nf, k + dep1, false);
      const int h = hop_group_dep<PREC, RECON, 3, D, R, STG>(acc);
      hop_dir<1, false, D, PREC, RECON, R>(acc, a, nb + h, nb + h, false);
    }
    const int dep2 = hop_group_dep<PREC, RECON, 1, D, R, STG>(acc);
    if constexpr (STG) {
      // every lane's copy has landed; then every wave's (the staged
      // kernel runs only when every thread has a site, VH a multiple of
      // the workgroup, so the whole workgroup reaches this barrier)
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
    {
      int yf = (y + 1 == g.L[2]) ? 0 : y + 1;
      int yb = (y == 0) ? g.L[2] - 1 : y - 1;
      int nf = half_index(t, x, yf, z, g) + dep2, nb = half_index(t, x, yb, z, g) + dep2;
      if constexpr (STG) {
        // slots: the window, or the wrap rows (same z offset in the row)
        const int hz = g.L[3] / 2, NW = St::window(g), zo = k % hz;
        const int jf = (yf == 0) ? NW + hz + zo + dep2 : nf - s0;
        const int jb = (yb == g.L[2] - 1 && y == 0) ? NW + zo : nb - s0;
        hop_dir_src<2, true, D, PREC, RECON, R>(acc, a, lds, lds_nrm, NT, jf, k + dep2, false);
        const int h = hop_group_dep<PREC, RECON, 3, D, R, STG>(acc);
        hop_dir_src<2, false, D, PREC, RECON, R>(acc, a, lds, lds_nrm, NT, jb + h, nb + h, false);
      } else {
        hop_dir<2, true, D, PREC, RECON, R>(acc, a, nf, k + dep2, false);
        const int h = hop_group_dep<PREC, RECON, 3, D, R, STG>(acc);
        hop_dir<2, false, D, PREC, RECON, R>(acc, a, nb + h, nb + h, false);
      }
    }
    const int dep3 = hop_group_dep<PREC, RECON, 2, D, R, STG>(acc);
    {
      int zf = (z + 1 == g.L[3]) ? 0 : z + 1;
      int zb = (z == 0) ? g.L[3] - 1 : z - 1;
      int nf = half_index(t, x, y, zf, g) + dep3, nb = half_index(t, x, y, zb, g) + dep3;
      if constexpr (STG) {
        hop_dir_src<3, true, D, PREC, RECON, R>(acc, a, lds, lds_nrm, NT, nf - s0, k + dep3, false);
        const int h = hop_group_dep<PREC, RECON, 3, D, R, STG>(acc);
        hop_dir_src<3, false, D, PREC, RECON, R>(acc, a, lds, lds_nrm, NT, nb - s0 + h, nb + h, false);
      } else {
        hop_dir<3, true, D, PREC, RECON, R>(acc, a, nf, k + dep3, false);
        const int h = hop_group_dep<PREC, RECON, 3, D, R, STG>(acc);
        hop_dir<3, false, D, PREC, RECON, R>(acc, a, nb + h, nb + h, false);
      }
    }
    // epilogue loads (clover block, own spinor) after the hop is final
    const int ke = k + hop_group_dep<PREC, RECON, 4, D, R, STG>(acc);
    const bool bz = bc_zero_slice(t, g);
    const R sc = bz ? (R)0 : (R)a.scale;
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < D; i++) acc[s][i] = sc * acc[s][i];

    if constexpr (EPI == EPI_PLAIN) {
      store_spinor<PREC, D, R, PREC == PREC_F64>(a.out, a.out_nrm, k, g.VHpad, acc);
    } else if constexpr (EPI == EPI_CLOVINV) {
      if (a.aux) store_spinor<PREC, D, R>(a.aux, a.aux_nrm, k, g.VHpad, acc);
      constexpr int N = 2 * D;
      Cx<R> o[4][D];
      int kb = ke;
#pragma unroll
      for (int blk2 = 0; blk2 < 2; blk2++) {
        Cx<R> xin[N], yout[N];
#pragma unroll
        for (int i = 0; i < D; i++) { xin[i] = acc[2 * blk2][i]; xin[D + i] = acc[2 * blk2 + 1][i]; }
        if constexpr (PREC == PREC_F16) {
          uint2 cq[N * N / 4];
          const R cs = (R)load_clover_block_q<D>(cq, a.clov, blk2, kb, g.VHpad, a.clov_nrm);
          herm_mv_q<N, R>(yout, cq, cs, xin);
        } else if constexpr (PREC == PREC_F64 && D >= 6) {
          herm_mv_f64_stream<N>(yout, a.clov, blk2, kb, g.VHpad, xin);
        } else {
          R cl[N * N];
          load_clover_block<PREC, D, R>(cl, a.clov, blk2, kb, g.VHpad, a.clov_nrm);
          herm_mv<N, R>(yout, cl, xin);
        }
#pragma unroll
        for (int i = 0; i < D; i++) { o[2 * blk2][i] = yout[i]; o[2 * blk2 + 1][i] = yout[D + i]; }
        kb = ke + hop_group_dep<PREC, RECON, 5, D, R, STG>(o);
      }
      store_spinor<PREC, D, R, PREC == PREC_F64>(a.out, a.out_nrm, k, g.VHpad, o);
    } else {
      constexpr int N = 2 * D;
      // FP64 at D_R >= 6: each block's two spins of the own spinor are read
      // with the block (the whole spinor is 8 D_R doubles)
      constexpr bool SPLIT = PREC == PREC_F64 && D >= 6;
      Cx<R> xs[4][D];
      if constexpr (!SPLIT) load_spinor<PREC, D, R>(xs, a.self, a.self_nrm, ke, g.VHpad);
      const R dg = (R)(a.diag + bc_cf_shift(t, g));
      const bool zs = a.zero_self_bc && bz;
      int kb = ke;
#pragma unroll
      for (int blk2 = 0; blk2 < 2; blk2++) {
        Cx<R> xin[N], yout[N];
        if constexpr (SPLIT) {
          const double2* sp = (const double2*)a.self;
#pragma unroll
          for (int c = 0; c < 2 * D; c++) {
            const double2 v = sp[(size_t)(2 * blk2 * D + c) * g.VHpad + kb];
            xs[2 * blk2 + c / D][c % D] = Cx<R>((R)v.x, (R)v.y);
          }
        }
#pragma unroll
        for (int i = 0; i < D; i++) {
          xin[i] = zs ? Cx<R>((R)0, (R)0) : xs[2 * blk2][i];
          xin[D + i] = zs ? Cx<R>((R)0, (R)0) : xs[2 * blk2 + 1][i];
        }
        if constexpr (PREC == PREC_F16) {
          uint2 cq[N * N / 4];
          const R cs = (R)load_clover_block_q<D>(cq, a.clov, blk2, kb, g.VHpad, a.clov_nrm);
          herm_mv_q<N, R>(yout, cq, cs, xin);
        } else if constexpr (PREC == PREC_F64 && D >= 6) {
          herm_mv_f64_stream<N>(yout, a.clov, blk2, kb, g.VHpad, xin);
        } else {
          R cl[N * N];
          load_clover_block<PREC, D, R>(cl, a.clov, blk2, kb, g.VHpad, a.clov_nrm);
          herm_mv<N, R>(yout, cl, xin);
        }
#pragma unroll
        for (int i = 0; i < D; i++) {
          acc[2 * blk2][i] = acc[2 * blk2][i] + yout[i] + dg * xin[i];
          acc[2 * blk2 + 1][i] = acc[2 * blk2 + 1][i] + yout[D + i] + dg * xin[D + i];
        }
        kb = ke + hop_group_dep<PREC, RECON, 5, D, R, STG>(acc);
      }
      if (a.gamma5) {
#pragma unroll
        for (int i = 0; i < D; i++) {
          acc[2][i] = Cx<R>(-acc[2][i].re, -acc[2][i].im);
          acc[3][i] = Cx<R>(-acc[3][i].re, -acc[3][i].im);
        }
      }
      store_spinor<PREC, D, R, PREC == PREC_F64>(a.out, a.out_nrm, k, g.VHpad, acc);
      if (a.reduce) {
        if constexpr (PREC == PREC_F16) {
          // reduce what was stored (the rounded field), as a consumer sees it
          Cx<R> back[4][D];
          load_spinor<PREC, D, R>(back, a.out, a.out_nrm, k, g.VHpad);
#pragma unroll
          for (int s = 0; s < 4; s++)
#pragma unroll
            for (int i = 0; i < D; i++) nrm2 += (double)back[s][i].re * back[s][i].re + (double)back[s][i].im * back[s][i].im;
        } else {
#pragma unroll
          for (int s = 0; s < 4; s++)
#pragma unroll
            for (int i = 0; i < D; i++) nrm2 += (double)acc[s][i].re * acc[s][i].re + (double)acc[s][i].im * acc[s][i].im;
        }
      }
    }
  }
  if constexpr (EPI == EPI_CLOV) {
    if (a.reduce) {
      double v[1] = {nrm2};
      block_reduce_finish<1, YH_HOP_BLOCK>(v, a.partials, nblk, blockIdx.x, a.ticket, a.red_out);
    }
  }
}

// standalone clover apply (Dphi_ee / Dphi_oo) and inverse (Dphi_oo_inv)
// (library/dirac.cpp:25-225)
template <int D, int PREC, bool INV>
__global__ void __launch_bounds__(256) k_clover(void* out, float* out_nrm, const void* in, const float* in_nrm,
                                                const void* clov, const float* clov_nrm, double diag, int parity,
                                                Geom g) {
  typedef typename SpinorStore<PREC, D>::real R;
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= g.VH) return;
  constexpr int N = 2 * D;
  int c[4];
  site_coords(k, parity, g, c);
  Cx<R> xs[4][D], o[4][D];
  load_spinor<PREC, D, R>(xs, in, in_nrm, k, g.VHpad);
  const R dg = (R)(diag + bc_cf_shift(c[0], g));
#pragma unroll
  for (int blk2 = 0; blk2 < 2; blk2++) {
    R cl[N * N];
    load_clover_block<PREC, D, R>(cl, clov, blk2, k, g.VHpad, clov_nrm);
    Cx<R> xin[N], yout[N];
#pragma unroll
    for (int i = 0; i < D; i++) { xin[i] = xs[2 * blk2][i]; xin[D + i] = xs[2 * blk2 + 1][i]; }
    herm_mv<N, R>(yout, cl, xin);
#pragma unroll
    for (int i = 0; i < D; i++) {
      if (INV) {
        o[2 * blk2][i] = yout[i];
        o[2 * blk2 + 1][i] = yout[D + i];
      } else {
        o[2 * blk2][i] = yout[i] + dg * xin[i];
        o[2 * blk2 + 1][i] = yout[D + i] + dg * xin[D + i];
      }
    }
  }
  store_spinor<PREC, D, R>(out, out_nrm, k, g.VHpad, o);
}

// ---------------------------------------------------------------------------
// Several right-hand sides in one launch: NR adjacent lanes take the same
// output site, lane r working on right-hand side r with the single-RHS code.
// Their link and clover loads are the same addresses within one wave
// instruction, so each link and clover block is fetched from memory once
// for the whole batch (at FP64 18-real links 1,152 of the 1,536 bytes per
// site and right-hand side), while each lane keeps the registers of the
// single-RHS hop.  Right-hand sides whose skip flag is set do nothing; the
// norm reduction of the CLOV epilogue is kept per right-hand side.
enum { MR_MAX = 8 };
struct HopArgsM {
  HopArgs a;                  // links, clover, scalars, geometry (field pointers unused)
  const void* in[MR_MAX];
  void* out[MR_MAX];
  const void* self[MR_MAX];
  double* red_out[MR_MAX];
  const int* skip[MR_MAX];
};

template <int D, int PREC, int RECON, int EPI, int NR>
__global__ void __launch_bounds__(YH_HOP_BLOCK, hop_min_waves<PREC>()) k_hop_mr(HopArgsM m) {
  static_assert(PREC != PREC_F16, "16-bit fields carry per-site scales");
  typedef typename SpinorStore<PREC, D>::real R;
  const int nblk = gridDim.x;
  const int blk = xcd_block(nblk, m.a.g.xq);
  const int t = blk * YH_HOP_BLOCK + threadIdx.x;
  const int k = t / NR, rr = t - (t / NR) * NR;
  HopArgs a = m.a;
  a.in = m.in[rr];
  a.out = m.out[rr];
  a.self = m.self[rr];
  const Geom& g = a.g;
  const bool live = k < g.VH && !(m.skip[rr] && *m.skip[rr]);
  double nrm2 = 0.0;
  if (live) {
    int c[4];
    site_coords(k, a.parity, g, c);
    const int tt = c[0], x = c[1], y = c[2], z = c[3];
    const int T = g.L[0];
    Cx<R> acc[4][D];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < D; i++) acc[s][i] = Cx<R>((R)0, (R)0);
    const bool ap = g.bc == BC_ANTIPERIODIC;
    {
      int tf = (tt + 1 == T) ? 0 : tt + 1;
      int tb = (tt == 0) ? T - 1 : tt - 1;
      int nf = half_index(tf, x, y, z, g), nb = half_index(tb, x, y, z, g);
      hop_dir<0, true, D, PREC, RECON, R>(acc, a, nf, k, ap && tt == 0);
      const int h = hop_group_dep<PREC, RECON, 3>(acc);
      hop_dir<0, false, D, PREC, RECON, R>(acc, a, nb + h, nb + h, ap && tb == 0);
    }
    const int dep1 = hop_group_dep<PREC, RECON, 0>(acc);
    {
      int xf = (x + 1 == g.L[1]) ? 0 : x + 1;
      int xb = (x == 0) ? g.L[1] - 1 : x - 1;
      int nf = half_index(tt, xf, y, z, g) + dep1, nb = half_index(tt, xb, y, z, g) + dep1;
      hop_dir<1, true, D, PREC, RECON, R>(acc, a, nf, k + dep1, false);
      const int h = hop_group_dep<PREC, RECON, 3>(acc);
      hop_dir<1, false, D, PREC, RECON, R>(acc, a, nb + h, nb + h, false);
    }
    const int dep2 = hop_group_dep<PREC, RECON, 1>(acc);
    {
      int yf = (y + 1 == g.L[2]) ? 0 : y + 1;
      int yb = (y == 0) ? g.L[2] - 1 : y - 1;
      int nf = half_index(tt, x, yf, z, g) + dep2, nb = half_index(tt, x, yb, z, g) + dep2;
      hop_dir<2, true, D, PREC, RECON, R>(acc, a, nf, k + dep2, false);
      const int h = hop_group_dep<PREC, RECON, 3>(acc);
      hop_dir<2, false, D, PREC, RECON, R>(acc, a, nb + h, nb + h, false);
    }
    const int dep3 = hop_group_dep<PREC, RECON, 2>(acc);
    {
      int zf = (z + 1 == g.L[3]) ? 0 : z + 1;
      int zb = (z == 0) ? g.L[3] - 1 : z - 1;
      int nf = half_index(tt, x, y, zf, g) + dep3, nb = half_index(tt, x, y, zb, g) + dep3;
      hop_dir<3, true, D, PREC, RECON, R>(acc, a, nf, k + dep3, false);
      const int h = hop_group_dep<PREC, RECON, 3>(acc);
      hop_dir<3, false, D, PREC, RECON, R>(acc, a, nb + h, nb + h, false);
    }
    const int ke = k + hop_group_dep<PREC, RECON, 4>(acc);
    const bool bz = bc_zero_slice(tt, g);
    const R sc = bz ? (R)0 : (R)a.scale;
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < D; i++) acc[s][i] = sc * acc[s][i];
    constexpr int N = 2 * D;
    if constexpr (EPI == EPI_CLOVINV) {
      Cx<R> o[4][D];
      int kb = ke;
#pragma unroll
      for (int blk2 = 0; blk2 < 2; blk2++) {
        R cl[N * N];
        load_clover_block<PREC, D, R>(cl, a.clov, blk2, kb, g.VHpad, a.clov_nrm);
        Cx<R> xin[N], yout[N];
#pragma unroll
        for (int i = 0; i < D; i++) { xin[i] = acc[2 * blk2][i]; xin[D + i] = acc[2 * blk2 + 1][i]; }
        herm_mv<N, R>(yout, cl, xin);
#pragma unroll
        for (int i = 0; i < D; i++) { o[2 * blk2][i] = yout[i]; o[2 * blk2 + 1][i] = yout[D + i]; }
        kb = ke + hop_group_dep<PREC, RECON, 5>(o);
      }
      store_spinor<PREC, D, R, PREC == PREC_F64>(a.out, nullptr, k, g.VHpad, o);
    } else {
      Cx<R> xs[4][D];
      load_spinor<PREC, D, R>(xs, a.self, nullptr, ke, g.VHpad);
      const R dg = (R)(a.diag + bc_cf_shift(tt, g));
      const bool zs = a.zero_self_bc && bz;
      int kb = ke;
#pragma unroll
      for (int blk2 = 0; blk2 < 2; blk2++) {
        R cl[N * N];
        load_clover_block<PREC, D, R>(cl, a.clov, blk2, kb, g.VHpad, a.clov_nrm);
        Cx<R> xin[N], yout[N];
#pragma unroll
        for (int i = 0; i < D; i++) {
          xin[i] = zs ? Cx<R>((R)0, (R)0) : xs[2 * blk2][i];
          xin[D + i] = zs ? Cx<R>((R)0, (R)0) : xs[2 * blk2 + 1][i];
        }
        herm_mv<N, R>(yout, cl, xin);
#pragma unroll
        for (int i = 0; i < D; i++) {
          acc[2 * blk2][i] = acc[2 * blk2][i] + yout[i] + dg * xin[i];
          acc[2 * blk2 + 1][i] = acc[2 * blk2 + 1][i] + yout[D + i] + dg * xin[D + i];
        }
        kb = ke + hop_group_dep<PREC, RECON, 5>(acc);
      }
      if (a.gamma5) {
#pragma unroll
        for (int i = 0; i < D; i++) {
          acc[2][i] = Cx<R>(-acc[2][i].re, -acc[2][i].im);
          acc[3][i] = Cx<R>(-acc[3][i].re, -acc[3][i].im);
        }
      }
      store_spinor<PREC, D, R, PREC == PREC_F64>(a.out, nullptr, k, g.VHpad, acc);
      if (a.reduce) {
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
          for (int i = 0; i < D; i++) nrm2 += (double)acc[s][i].re * acc[s][i].re + (double)acc[s][i].im * acc[s][i].im;
      }
    }
  }
  if constexpr (EPI == EPI_CLOV) {
    if (a.reduce) {
      double v[NR];
#pragma unroll
      for (int q = 0; q < NR; q++) v[q] = q == rr ? nrm2 : 0.0;
      double res[NR];
      if (block_reduce_finish<NR, YH_HOP_BLOCK>(v, a.partials, nblk, blockIdx.x, a.ticket, res) && threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NR; q++)
          if (m.red_out[q] && !(m.skip[q] && *m.skip[q])) *m.red_out[q] = res[q];
      }
    }
  }
}

template <int D, int PREC, int NR>
static int hop_mr_epi(Ctx* c, const HopArgsM& m, int epi) {
  dim3 grid((unsigned)(((long)c->g.VH * NR + YH_HOP_BLOCK - 1) / YH_HOP_BLOCK)), block(YH_HOP_BLOCK);
  if (epi == EPI_CLOVINV) hipLaunchKernelGGL((k_hop_mr<D, PREC, RECON_18, EPI_CLOVINV, NR>), grid, block, 0, c->stream, m);
  else hipLaunchKernelGGL((k_hop_mr<D, PREC, RECON_18, EPI_CLOV, NR>), grid, block, 0, c->stream, m);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

// NR right-hand sides (2 or 4), FP64 or FP32 fields, 18-real links,
// CLOVINV or CLOV epilogue
template <int D>
int hop_mr_dispatch_d(Ctx* c, const HopArgsM& m, int prec, int nr, int epi) {
  if (prec == PREC_F64) {
    if (nr == 2) return hop_mr_epi<D, PREC_F64, 2>(c, m, epi);
    if (nr == 4) return hop_mr_epi<D, PREC_F64, 4>(c, m, epi);
  } else if (prec == PREC_F32) {
    if (nr == 2) return hop_mr_epi<D, PREC_F32, 2>(c, m, epi);
    if (nr == 4) return hop_mr_epi<D, PREC_F32, 4>(c, m, epi);
  }
  set_error("multi-RHS hop: unsupported precision / batch");
  return YH_ERR_UNSUPPORTED;
}

// ---------------------------------------------------------------------------
template <int D, int PREC, int RECON>
static int hop_dispatch_epi(Ctx* c, const HopArgs& a, int epi, hipStream_t st) {
  dim3 grid((c->g.VH + YH_HOP_BLOCK - 1) / YH_HOP_BLOCK), block(YH_HOP_BLOCK);
  if constexpr (D <= 3) {
    // LDS-staged y / z neighbours (yh_tune YH_TUNE_HOP_STAGE) where the
    // geometry allows it (Stage::usable; the kernel has a barrier)
    const size_t lb = Stage<PREC, D>::bytes(c->g);
    // FP64 clover epilogues run one wave per SIMD (256 VGPRs) and measured
    // slower staged (H^2 at 32^4: 880 against 859 us); only the plain FP64
    // hop is staged (134 against 141 us)
    const bool stage_epi = PREC != PREC_F64 || epi == EPI_PLAIN;
    if (((c->hop_stage >> PREC) & 1) && stage_epi && Stage<PREC, D>::usable(c->g)) {
      switch (epi) {
        case EPI_PLAIN: hipLaunchKernelGGL((k_hop<D, PREC, RECON, EPI_PLAIN, true>), grid, block, lb, st, a); break;
        case EPI_CLOVINV: hipLaunchKernelGGL((k_hop<D, PREC, RECON, EPI_CLOVINV, true>), grid, block, lb, st, a); break;
        default: hipLaunchKernelGGL((k_hop<D, PREC, RECON, EPI_CLOV, true>), grid, block, lb, st, a); break;
      }
      YH_HIP(hipGetLastError());
      return YH_OK;
    }
  }
  switch (epi) {
    case EPI_PLAIN: hipLaunchKernelGGL((k_hop<D, PREC, RECON, EPI_PLAIN>), grid, block, 0, st, a); break;
    case EPI_CLOVINV: hipLaunchKernelGGL((k_hop<D, PREC, RECON, EPI_CLOVINV>), grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL((k_hop<D, PREC, RECON, EPI_CLOV>), grid, block, 0, st, a); break;
  }
  YH_HIP(hipGetLastError());
  return YH_OK;
}

template <int D>
int hop_dispatch_d(Ctx* c, const HopArgs& a, int prec, int recon, int epi) {
  hipStream_t st = c->stream;
  if (recon == RECON_12H) {
    if constexpr (D == 3) {
      if (prec == PREC_F32) return hop_dispatch_epi<3, PREC_F32, RECON_12H>(c, a, epi, st);
    }
    set_error("FP16 12-real links under FP32 spinors need SU(3) fundamental");
    return YH_ERR_UNSUPPORTED;
  }
  if (recon == RECON_12) {
    if constexpr (D == 3) {
      if (prec == PREC_F64) return hop_dispatch_epi<3, PREC_F64, RECON_12>(c, a, epi, st);
      if (prec == PREC_F32) return hop_dispatch_epi<3, PREC_F32, RECON_12>(c, a, epi, st);
      return hop_dispatch_epi<3, PREC_F16, RECON_12>(c, a, epi, st);
    }
    set_error("12-real links need SU(3) fundamental");
    return YH_ERR_UNSUPPORTED;
  }
  if (prec == PREC_F64) return hop_dispatch_epi<D, PREC_F64, RECON_18>(c, a, epi, st);
  if (prec == PREC_F32) return hop_dispatch_epi<D, PREC_F32, RECON_18>(c, a, epi, st);
  return hop_dispatch_epi<D, PREC_F16, RECON_18>(c, a, epi, st);
}

template <int D>
int clover_apply_d(Ctx* c, int prec, int parity, double diag, Field* out, const Field* in, int inverse) {
  int N = 2 * c->dr;
  const int cprec = prec;
  size_t cvb = cprec == PREC_F16 ? sizeof(uint2) : 16;
  size_t cnv = cprec == PREC_F64 ? (size_t)N * N : (size_t)N * N / 2;
  const void* cl;
  const float* cn = nullptr;
  if (inverse) {
    cl = c->clinv[cprec];
    cn = c->clinv_nrm;
  } else {
    cl = c->clov[cprec] ? (const char*)c->clov[cprec] + (size_t)parity * cnv * c->g.VHpad * cvb : nullptr;
    cn = c->clov_nrm ? c->clov_nrm + (size_t)parity * 2 * c->g.VHpad : nullptr;
  }
  if (!cl || (cprec == PREC_F16 && !cn)) { set_error("clover data missing"); return YH_ERR_STATE; }
  dim3 grid((c->g.VH + 255) / 256);
  void* o = out->half[parity];
  float* on = out->nrm[parity];
  const void* i = in->half[parity];
  const float* inn = in->nrm[parity];
#define YH_CL(P)                                                                                               \
  if (inverse) hipLaunchKernelGGL((k_clover<D, P, true>), grid, dim3(256), 0, c->stream, o, on, i, inn, cl, cn, diag, parity, c->g); \
  else hipLaunchKernelGGL((k_clover<D, P, false>), grid, dim3(256), 0, c->stream, o, on, i, inn, cl, cn, diag, parity, c->g);
  if (prec == PREC_F64) { YH_CL(PREC_F64) }
  else if (prec == PREC_F32) { YH_CL(PREC_F32) }
  else { YH_CL(PREC_F16) }
#undef YH_CL
  YH_HIP(hipGetLastError());
  return YH_OK;
}

}  // namespace yh
