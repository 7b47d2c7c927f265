// Learned preconditioner: a gauge-covariant parallel-transport convolution
// network (no counterpart in the reference).
//
// Layer:   (L h)_o(x) = sum_q sum_i W_q[o][i] (T_q h_i)(x)
//   q = 0 identity, 1+mu: U(x,mu) h(x+mu), 5+mu: U(x-mu,mu)^+ h(x-mu), with
//   the FP32 represented links and the antiperiodic sign of the Dirac
//   operator; W_q[o][i] is a complex 4x4 spin matrix, so colour is only
//   touched by the transport and the layer is gauge covariant.
// Adjoint: (L^+ g)_i(x) = sum_q sum_o W_qbar[o][i]^H (T_q g_o)(x): the same
//   kernel with a re-indexed weight matrix.
// Network on the even sublattice: A = P_e L_n ... L_1 E, preconditioner
//   M = A A^+ (Hermitian positive semi-definite, usable inside CG).
//
// Kernels.  The channel mixing is a GEMM over rows = (site, colour) with
// K = (path, input channel, spin, re/im) and N = (output channel, spin,
// re/im); complex weights are expanded to real 2x2 blocks and held as BF16
// in LDS in MFMA fragment order.  Each wave takes 16 sites; a lane computes
// the transported spinor of one (site, path, channel) in FP32 registers and
// feeds it, rounded to BF16, straight into v_mfma_f32_16x16x32_bf16 as the A
// operand (one M-block per colour), accumulating in FP32.
// The weight gradient sum_{x,a} G_o(x)[a][s] conj((T_q h_i)(x)[a][s']) is
// the transposed product: cotangent and transported input are staged
// through LDS in [colour][column][site] order so that the reduction over
// 32 sites is the MFMA's K dimension; per-workgroup partials are summed in
// FP64 in a fixed order.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include "yh_internal.hpp"
#include "yh_reduce.hpp"

namespace yh {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f4;
union Frag {
  uint4 u;
  bf16x8 v;
};

constexpr int ND = 3;  // colours: SU(3) fundamental

__device__ __forceinline__ unsigned bf16_bits(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return u >> 16;
}
__device__ __forceinline__ unsigned pack2(float a, float b) { return bf16_bits(a) | (bf16_bits(b) << 16); }

// t[a][s] = (T_q h)(x) for one channel (h: [V][3][4] complex FP32)
__device__ __forceinline__ void net_transport(float2 (&t)[ND][4], const float2* __restrict__ h,
                                              const float2* __restrict__ links, int x, int q, const Geom& g) {
  if (q == 0) {
    const float4* p = (const float4*)(h + (size_t)x * 12);
#pragma unroll
    for (int k = 0; k < 6; k++) {
      float4 v = p[k];
      t[(2 * k) / 4][(2 * k) % 4] = make_float2(v.x, v.y);
      t[(2 * k + 1) / 4][(2 * k + 1) % 4] = make_float2(v.z, v.w);
    }
    return;
  }
  int c[4];
  full_coords(x, g, c);
  const int mu = (q - 1) & 3;
  const bool fw = q <= 4;
  int n[4] = {c[0], c[1], c[2], c[3]};
  n[mu] = fw ? (c[mu] + 1 == g.L[mu] ? 0 : c[mu] + 1) : (c[mu] == 0 ? g.L[mu] - 1 : c[mu] - 1);
  const int y = full_index(n[0], n[1], n[2], n[3], g);
  const int ls = fw ? x : y;
  const int lp = ls >= g.VH;
  const float2* lb = links + (size_t)lp * 4 * 9 * g.VHpad + (size_t)mu * 9 * g.VHpad + (ls - lp * g.VH);
  float2 u[3][3];
#pragma unroll
  for (int e = 0; e < 9; e++) u[e / 3][e % 3] = lb[(size_t)e * g.VHpad];
  const bool flip = g.bc == BC_ANTIPERIODIC && mu == 0 && (fw ? c[0] : n[0]) == 0;
  float2 hv[ND][4];
  const float4* p = (const float4*)(h + (size_t)y * 12);
#pragma unroll
  for (int k = 0; k < 6; k++) {
    float4 v = p[k];
    hv[(2 * k) / 4][(2 * k) % 4] = make_float2(v.x, v.y);
    hv[(2 * k + 1) / 4][(2 * k + 1) % 4] = make_float2(v.z, v.w);
  }
  const float sg = flip ? -1.f : 1.f;
#pragma unroll
  for (int a = 0; a < ND; a++)
#pragma unroll
    for (int s = 0; s < 4; s++) {
      float re = 0.f, im = 0.f;
#pragma unroll
      for (int b = 0; b < ND; b++) {
        float2 w = fw ? u[a][b] : make_float2(u[b][a].x, -u[b][a].y);
        re += w.x * hv[b][s].x - w.y * hv[b][s].y;
        im += w.x * hv[b][s].y + w.y * hv[b][s].x;
      }
      t[a][s] = make_float2(sg * re, sg * im);
    }
}

template <int CIN, int COUT>
struct PtcShape {
  static constexpr int NP = 9 * CIN;              // (path, channel) pairs
  static constexpr int NPP = (NP + 7) / 8 * 8;    // padded: 4 per K-step, 2 per wgrad column block, 4 waves
  static constexpr int KS = NPP / 4;              // forward K-steps of 32
  static constexpr int NB = (COUT * 8 + 15) / 16;  // forward N-blocks
  static constexpr int KCOLS = NPP * 8;            // real K columns
  static constexpr int NROWS = NB * 16;            // real N (padded)
  static constexpr int BFRAG = KS * NB * 64;       // uint4 per B matrix
};

// ---------------------------------------------------------------------------
// forward layer (also the adjoint, with the adjoint weight matrix)
// in: [CIN][V][3][4], out: [COUT][V][3][4]; tiles of 16 sites
// ---------------------------------------------------------------------------
typedef __attribute__((ext_vector_type(8))) float f8;

// A-operand fragments (one per colour) of a transported spinor t[a][s]:
// k order within the 8 = (spin, re/im); hardware RNE conversion to BF16
__device__ __forceinline__ void make_frags(Frag (&A)[ND], const float2 (&t)[ND][4]) {
#pragma unroll
  for (int a = 0; a < ND; a++) {
    f8 v = {t[a][0].x, t[a][0].y, t[a][1].x, t[a][1].y, t[a][2].x, t[a][2].y, t[a][3].x, t[a][3].y};
    A[a].v = __builtin_convertvector(v, bf16x8);
  }
}

// neighbour and link of path q (1..8) for a site with coordinates c
struct PathRef {
  int y;           // full index of the neighbour
  const float2* l; // link base (element 0)
  bool fw, flip;
};

__device__ __forceinline__ PathRef path_ref(const int (&c)[4], int x, int q, const float2* links, const Geom& g) {
  PathRef P;
  const int mu = (q - 1) & 3;
  P.fw = q <= 4;
  int n[4] = {c[0], c[1], c[2], c[3]};
  n[mu] = P.fw ? (c[mu] + 1 == g.L[mu] ? 0 : c[mu] + 1) : (c[mu] == 0 ? g.L[mu] - 1 : c[mu] - 1);
  P.y = full_index(n[0], n[1], n[2], n[3], g);
  const int ls = P.fw ? x : P.y;
  const int lp = ls >= g.VH;
  P.l = links + (size_t)(lp * 4 + mu) * 9 * g.VHpad + (ls - lp * g.VH);
  P.flip = g.bc == BC_ANTIPERIODIC && mu == 0 && (P.fw ? c[0] : n[0]) == 0;
  return P;
}

__device__ __forceinline__ void load_h(float4 (&h)[6], const float2* base, int y) {
  const float4* p = (const float4*)(base + (size_t)y * 12);
#pragma unroll
  for (int k = 0; k < 6; k++) h[k] = p[k];
}

__device__ __forceinline__ void load_u(float2 (&u)[9], const float2* l, int VHpad) {
#pragma unroll
  for (int e = 0; e < 9; e++) u[e] = l[(size_t)e * VHpad];
}

// t = sign * U h (fw) or sign * U^+ h; h as 6 float4 in [colour][spin] order
__device__ __forceinline__ void apply_link(float2 (&t)[ND][4], const float2 (&u)[9], const float4 (&h)[6], bool fw,
                                           bool flip) {
  float2 hv[ND][4];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    hv[(2 * k) / 4][(2 * k) % 4] = make_float2(h[k].x, h[k].y);
    hv[(2 * k + 1) / 4][(2 * k + 1) % 4] = make_float2(h[k].z, h[k].w);
  }
  const float sg = flip ? -1.f : 1.f;
#pragma unroll
  for (int a = 0; a < ND; a++)
#pragma unroll
    for (int s = 0; s < 4; s++) {
      float re = 0.f, im = 0.f;
#pragma unroll
      for (int b = 0; b < ND; b++) {
        float2 w = fw ? u[a * 3 + b] : make_float2(u[b * 3 + a].x, -u[b * 3 + a].y);
        re += w.x * hv[b][s].x - w.y * hv[b][s].y;
        im += w.x * hv[b][s].y + w.y * hv[b][s].x;
      }
      t[a][s] = make_float2(sg * re, sg * im);
    }
}

__device__ __forceinline__ void unpack_h(float2 (&t)[ND][4], const float4 (&h)[6]) {
#pragma unroll
  for (int k = 0; k < 6; k++) {
    t[(2 * k) / 4][(2 * k) % 4] = make_float2(h[k].x, h[k].y);
    t[(2 * k + 1) / 4][(2 * k + 1) % 4] = make_float2(h[k].z, h[k].w);
  }
}

template <int CIN, int COUT, bool PF>
__global__ void __launch_bounds__(256) k_ptc(const float2* __restrict__ in, float2* __restrict__ out,
                                             const uint4* __restrict__ Bg, const float2* __restrict__ links, Geom g,
                                             int ntiles) {
  typedef PtcShape<CIN, COUT> S;
  extern __shared__ uint4 Bs[];
  // output staging: per wave, one output channel x 16 sites x 24 floats
  // (6 KiB per workgroup, so that two workgroups with the 72 KiB weight
  // matrix of an 8 -> 8 layer still share a CU)
  __shared__ __align__(16) float stage[4][16][24];
  for (int i = threadIdx.x; i < S::BFRAG; i += 256) Bs[i] = Bg[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, grp = lane >> 4;
  for (int tile = blockIdx.x * 4 + wave; tile < ntiles; tile += gridDim.x * 4) {
    const int xb = tile * 16;
    const int x = xb + r;
    int c[4];
    full_coords(x, g, c);
    f4 acc[ND][S::NB];
#pragma unroll
    for (int a = 0; a < ND; a++)
#pragma unroll
      for (int nb = 0; nb < S::NB; nb++) acc[a][nb] = f4{0.f, 0.f, 0.f, 0.f};
    auto mfma_step = [&](int ks, const Frag (&A)[ND]) {
#pragma unroll
      for (int nb = 0; nb < S::NB; nb++) {
        Frag B;
        B.u = Bs[(ks * S::NB + nb) * 64 + lane];
#pragma unroll
        for (int a = 0; a < ND; a++) acc[a][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[a].v, B.v, acc[a][nb], 0, 0, 0);
      }
    };
    if constexpr (CIN % 4 == 0 && PF) {
      // every lane group takes the same path in a K-step: one link per path,
      // CIN/4 channels per lane; the next path's loads are issued before
      // this path's arithmetic
      constexpr int CPL = CIN / 4;
      float4 h[CPL][6];
      float2 u[9];
#pragma unroll
      for (int cc = 0; cc < CPL; cc++) load_h(h[cc], in + (size_t)(4 * cc + grp) * g.V * 12, x);
      PathRef P{};
#pragma unroll
      for (int q = 0; q < 9; q++) {
        float4 hn[CPL][6];
        float2 un[9];
        PathRef Pn{};
        if (q < 8) {
          Pn = path_ref(c, x, q + 1, links, g);
          load_u(un, Pn.l, g.VHpad);
#pragma unroll
          for (int cc = 0; cc < CPL; cc++) load_h(hn[cc], in + (size_t)(4 * cc + grp) * g.V * 12, Pn.y);
        }
#pragma unroll
        for (int cc = 0; cc < CPL; cc++) {
          float2 t[ND][4];
          if (q == 0) unpack_h(t, h[cc]);
          else apply_link(t, u, h[cc], P.fw, P.flip);
          Frag A[ND];
          make_frags(A, t);
          mfma_step(q * CPL + cc, A);
        }
        if (q < 8) {
#pragma unroll
          for (int cc = 0; cc < CPL; cc++)
#pragma unroll
            for (int k = 0; k < 6; k++) h[cc][k] = hn[cc][k];
#pragma unroll
          for (int e = 0; e < 9; e++) u[e] = un[e];
          P = Pn;
        }
      }
    } else if constexpr (CIN % 4 == 0) {
      // same order without the prefetch (fewer registers)
      constexpr int CPL = CIN / 4;
      for (int q = 0; q < 9; q++) {
        PathRef P{};
        float2 u[9];
        if (q > 0) {
          P = path_ref(c, x, q, links, g);
          load_u(u, P.l, g.VHpad);
        }
        const int y = q > 0 ? P.y : x;
#pragma unroll
        for (int cc = 0; cc < CPL; cc++) {
          float4 h[6];
          load_h(h, in + (size_t)(4 * cc + grp) * g.V * 12, y);
          float2 t[ND][4];
          if (q == 0) unpack_h(t, h);
          else apply_link(t, u, h, P.fw, P.flip);
          Frag A[ND];
          make_frags(A, t);
          mfma_step(q * CPL + cc, A);
        }
      }
    } else {
      // CIN < 4: the lane groups of a K-step take different (path, channel) pairs
      for (int ks = 0; ks < S::KS; ks++) {
        const int j = ks * 4 + grp;
        float2 t[ND][4];
        if (j < S::NP) {
          const int q = j / CIN, i = j % CIN;
          float4 h[6];
          if (q == 0) {
            load_h(h, in + (size_t)i * g.V * 12, x);
            unpack_h(t, h);
          } else {
            PathRef P = path_ref(c, x, q, links, g);
            float2 u[9];
            load_u(u, P.l, g.VHpad);
            load_h(h, in + (size_t)i * g.V * 12, P.y);
            apply_link(t, u, h, P.fw, P.flip);
          }
        } else {
#pragma unroll
          for (int a = 0; a < ND; a++)
#pragma unroll
            for (int s = 0; s < 4; s++) t[a][s] = make_float2(0.f, 0.f);
        }
        Frag A[ND];
        make_frags(A, t);
        mfma_step(ks, A);
      }
    }
    // epilogue: lane holds column r of each N-block for rows (sites) 4 grp + i;
    // stage one output channel at a time in LDS, then write its 16 sites x
    // 96 B with coalesced float4 stores
#pragma unroll
    for (int nb = 0; nb < S::NB; nb++)
#pragma unroll
      for (int ch = 0; ch < 2; ch++) {
        const int oc = 2 * nb + ch;
        if (oc >= COUT) break;
        const int n = nb * 16 + r;
        if ((n >> 3) == oc) {
#pragma unroll
          for (int i = 0; i < 4; i++)
#pragma unroll
            for (int a = 0; a < ND; a++) stage[wave][grp * 4 + i][a * 8 + (n & 7)] = acc[a][nb][i];
        }
        // the staging area is private to this wave: order its LDS writes
        // before the reads (LDS executes one wave's operations in order)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (int idx = lane; idx < 96; idx += 64) {
          const int site = idx / 6, part = idx % 6;
          const float4 v = *(const float4*)&stage[wave][site][part * 4];
          *(float4*)((float*)out + ((size_t)oc * g.V + xb + site) * 24 + part * 4) = v;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
  }
}

// ---------------------------------------------------------------------------
// weight gradient: D[m][n] = sum_{x,a} Gcot[x,a][m] * A[x,a][n]
//   m = (o, s, re/im) (NROWS), n = (q, i, s', re/im) (KCOLS)
// one partial matrix per workgroup; tiles of 32 sites
// ---------------------------------------------------------------------------
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_ptc_wgrad(const float2* __restrict__ gcot, const float2* __restrict__ h,
                                                   const float2* __restrict__ links, Geom g, int ntiles,
                                                   float* __restrict__ partial) {
  typedef PtcShape<CIN, COUT> S;
  constexpr int PW = S::NPP / 4;   // pairs per wave
  constexpr int NBW = PW / 2;      // 16-column blocks per wave
  constexpr int MB = S::NB;        // 16-row blocks over m
  __shared__ __align__(16) unsigned short Gt[ND][S::NROWS][32];
  __shared__ __align__(16) unsigned short At[4][ND][16][32];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < ND * S::NROWS * 32; i += 256) (&Gt[0][0][0])[i] = 0;
  f4 acc[MB][NBW];
#pragma unroll
  for (int mb = 0; mb < MB; mb++)
#pragma unroll
    for (int nb = 0; nb < NBW; nb++) acc[mb][nb] = f4{0.f, 0.f, 0.f, 0.f};
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    __syncthreads();
    {
      const int site = threadIdx.x & 31;
      const int x = tile * 32 + site;
      for (int oc = threadIdx.x >> 5; oc < COUT; oc += 8) {
        const float2* p = gcot + ((size_t)oc * g.V + x) * 12;
#pragma unroll
        for (int a = 0; a < ND; a++)
#pragma unroll
          for (int s = 0; s < 4; s++) {
            float2 v = p[a * 4 + s];
            Gt[a][(oc * 4 + s) * 2][site] = (unsigned short)bf16_bits(v.x);
            Gt[a][(oc * 4 + s) * 2 + 1][site] = (unsigned short)bf16_bits(v.y);
          }
      }
    }
    __syncthreads();
    for (int nb = 0; nb < NBW; nb++) {
      const int half = lane >> 5, site = lane & 31;
      const int j = wave * PW + nb * 2 + half;
      float2 t[ND][4];
      if (j < S::NP) {
        net_transport(t, h + (size_t)(j % CIN) * g.V * 12, links, tile * 32 + site, j / CIN, g);
      } else {
#pragma unroll
        for (int a = 0; a < ND; a++)
#pragma unroll
          for (int s = 0; s < 4; s++) t[a][s] = make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int a = 0; a < ND; a++)
#pragma unroll
        for (int s = 0; s < 4; s++) {
          At[wave][a][half * 8 + s * 2][site] = (unsigned short)bf16_bits(t[a][s].x);
          At[wave][a][half * 8 + s * 2 + 1][site] = (unsigned short)bf16_bits(t[a][s].y);
        }
      __syncthreads();
#pragma unroll
      for (int a = 0; a < ND; a++) {
        Frag B;
        B.u = *(const uint4*)&At[wave][a][lane & 15][(lane >> 4) * 8];
#pragma unroll
        for (int mb = 0; mb < MB; mb++) {
          Frag A;
          A.u = *(const uint4*)&Gt[a][mb * 16 + (lane & 15)][(lane >> 4) * 8];
          acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.v, B.v, acc[mb][nb], 0, 0, 0);
        }
      }
      __syncthreads();
    }
  }
  // D rows m = mb*16 + 4*(lane>>4) + i, column n = (wave*PW + nb*2)*8 + (lane & 15)
  float* P = partial + (size_t)blockIdx.x * S::NROWS * S::KCOLS;
#pragma unroll
  for (int mb = 0; mb < MB; mb++)
#pragma unroll
    for (int nb = 0; nb < NBW; nb++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int m = mb * 16 + 4 * (lane >> 4) + i;
        const int n = (wave * PW + nb * 2) * 8 + (lane & 15);
        P[(size_t)m * S::KCOLS + n] = acc[mb][nb][i];
      }
}

// sum the per-workgroup partials (FP64, fixed order) and fold the real 2x2
// blocks into the complex gradient d/dRe W + i d/dIm W, added to grad
__global__ void k_wgrad_fold(const float* __restrict__ partial, int nparts, int nrows, int kcols, int cin, int cout,
                             double coeff, double2* __restrict__ grad) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;  // complex weight [q][o][i][s][s']
  const int nw = 9 * cout * cin * 16;
  if (w >= nw) return;
  const int sp = w & 3, s = (w >> 2) & 3, rest = w >> 4;
  const int i = rest % cin, o = (rest / cin) % cout, q = rest / (cin * cout);
  const int kre = ((q * cin + i) * 4 + sp) * 2, nre = (o * 4 + s) * 2;
  double rr = 0, ii = 0, ri = 0, ir = 0;
  const size_t stride = (size_t)nrows * kcols;
  for (int b = 0; b < nparts; b++) {
    const float* P = partial + b * stride;
    rr += P[(size_t)nre * kcols + kre];
    ii += P[(size_t)(nre + 1) * kcols + kre + 1];
    ri += P[(size_t)nre * kcols + kre + 1];
    ir += P[(size_t)(nre + 1) * kcols + kre];
  }
  grad[w].x += coeff * (rr + ii);
  grad[w].y += coeff * (ir - ri);
}

// real-expanded BF16 weight matrix in MFMA B-fragment order.
// forward: kernel input channels ci = cin, output co = cout, w = W_q[o][i][s][s'];
// adjoint: input channels = cout, output = cin, w = conj(W_qbar[i][o][s'][s])
__global__ void k_net_bmat(const float2* __restrict__ W, uint4* __restrict__ B, int cin, int cout, int adjoint, int ks_n,
                           int nb_n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ks_n * nb_n * 64) return;
  const int lane = t & 63, nb = (t >> 6) % nb_n, ks = (t >> 6) / nb_n;
  const int kin = adjoint ? cout : cin, kout = adjoint ? cin : cout;
  const int n = nb * 16 + (lane & 15);
  const int oc = n >> 3, s = (n & 7) >> 1, rio = n & 1;
  unsigned v[8];
  for (int e = 0; e < 8; e++) {
    const int k = ks * 32 + (lane >> 4) * 8 + e;
    const int j = k >> 3, sp = (k & 7) >> 1, rii = k & 1;
    float val = 0.f;
    if (j < 9 * kin && oc < kout) {
      const int q = j / kin, i = j % kin;
      float2 w;
      if (!adjoint) {
        w = W[((size_t)(q * cout + oc) * cin + i) * 16 + s * 4 + sp];
      } else {
        const int qb = q == 0 ? 0 : (q <= 4 ? q + 4 : q - 4);
        float2 f = W[((size_t)(qb * cout + i) * cin + oc) * 16 + sp * 4 + s];
        w = make_float2(f.x, -f.y);
      }
      // [[re, im], [-im, re]] over (input re/im) x (output re/im)
      val = (rii == 0) ? (rio == 0 ? w.x : w.y) : (rio == 0 ? -w.y : w.x);
    }
    v[e] = bf16_bits(val);
  }
  B[t] = make_uint4(v[0] | (v[1] << 16), v[2] | (v[3] << 16), v[4] | (v[5] << 16), v[6] | (v[7] << 16));
}

// even half of a device field -> channel 0 of an activation (odd sites zero)
template <int PREC>
__global__ void k_field_to_act(const void* __restrict__ f, const float* __restrict__ nrm, float2* __restrict__ act,
                               Geom g) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= g.V) return;
  float4* o = (float4*)(act + (size_t)x * 12);
  if (x >= g.VH) {
#pragma unroll
    for (int k = 0; k < 6; k++) o[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  Cx<float> s[4][ND];
  load_spinor<PREC, ND, float>(s, f, nrm, x, g.VHpad);
  float2 v[12];
#pragma unroll
  for (int a = 0; a < ND; a++)
#pragma unroll
    for (int sp = 0; sp < 4; sp++) v[a * 4 + sp] = make_float2(s[sp][a].re, s[sp][a].im);
#pragma unroll
  for (int k = 0; k < 6; k++) o[k] = make_float4(v[2 * k].x, v[2 * k].y, v[2 * k + 1].x, v[2 * k + 1].y);
}

template <int PREC>
__global__ void k_act_to_field(const float2* __restrict__ act, void* __restrict__ f, float* __restrict__ nrm, Geom g) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= g.VH) return;
  const float4* p = (const float4*)(act + (size_t)x * 12);
  float2 v[12];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    float4 q = p[k];
    v[2 * k] = make_float2(q.x, q.y);
    v[2 * k + 1] = make_float2(q.z, q.w);
  }
  Cx<float> s[4][ND];
#pragma unroll
  for (int a = 0; a < ND; a++)
#pragma unroll
    for (int sp = 0; sp < 4; sp++) s[sp][a] = Cx<float>(v[a * 4 + sp].x, v[a * 4 + sp].y);
  store_spinor<PREC, ND, float>(f, nrm, x, g.VHpad, s);
}

// training residual on even sites: d = f - x; red = (|d|^2, |x|^2)
__global__ void __launch_bounds__(256) k_net_diff(const float2* __restrict__ f, const float2* __restrict__ xs,
                                                  float2* __restrict__ d, int n, double* partials, unsigned* ticket,
                                                  double* out) {
  const int i0 = blockIdx.x * 256 + threadIdx.x;
  double v[2] = {0.0, 0.0};
  for (int i = i0; i < n; i += gridDim.x * 256) {
    float2 a = f[i], b = xs[i];
    float2 e = make_float2(a.x - b.x, a.y - b.y);
    d[i] = e;
    v[0] += (double)e.x * e.x + (double)e.y * e.y;
    v[1] += (double)b.x * b.x + (double)b.y * b.y;
  }
  block_reduce_finish<2, 256>(v, partials, gridDim.x, blockIdx.x, ticket, out);
}

__global__ void k_net_scale(float2* __restrict__ a, int n, const double* __restrict__ red, double c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float f = (float)(c / red[1]);
  a[i] = make_float2(a[i].x * f, a[i].y * f);
}

// Adam on the FP32 master weights (re and im are independent parameters)
__global__ void k_net_adam(float2* __restrict__ W, float2* __restrict__ m1, float2* __restrict__ m2,
                           const double2* __restrict__ grad, int nw, float lr, float b1, float b2, float eps, float c1,
                           float c2, float gscale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  float2 gr = make_float2((float)(grad[i].x * gscale), (float)(grad[i].y * gscale));
  float2 a = m1[i], b = m2[i], w = W[i];
  a.x = b1 * a.x + (1.f - b1) * gr.x;
  a.y = b1 * a.y + (1.f - b1) * gr.y;
  b.x = b2 * b.x + (1.f - b2) * gr.x * gr.x;
  b.y = b2 * b.y + (1.f - b2) * gr.y * gr.y;
  w.x -= lr * (a.x / c1) / (sqrtf(b.x / c2) + eps);
  w.y -= lr * (a.y / c1) / (sqrtf(b.y / c2) + eps);
  m1[i] = a;
  m2[i] = b;
  W[i] = w;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct Net {
  Ctx* c = nullptr;
  int nl = 0;
  int ch[5] = {0, 0, 0, 0, 0};
  size_t woff[5] = {0, 0, 0, 0, 0};  // complex weight offsets
  size_t nw = 0;
  float2* W = nullptr;
  float2* m1 = nullptr;
  float2* m2 = nullptr;
  double2* grad = nullptr;
  int adam_t = 0;
  uint4* Bf[4] = {nullptr, nullptr, nullptr, nullptr};
  uint4* Ba[4] = {nullptr, nullptr, nullptr, nullptr};
  bool bdirty = true;
  // activations: fwd[l] = input of layer l (fwd[nl] = output), same for the
  // adjoint chain (adj[l] = cotangent at the output of layer l, adj[-1] in
  // adj[nl]); plus training scratch
  float2* fwd[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  float2* adj[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  float2* keepU[4] = {nullptr, nullptr, nullptr, nullptr};
  float2* keepB[4] = {nullptr, nullptr, nullptr, nullptr};
  float2* ab = nullptr;  // source b as activation
  float2* ax = nullptr;  // solution x as activation
  float* partial = nullptr;
  size_t partial_floats = 0;
  int wg_blocks = 0;
};

static size_t act_bytes(const Ctx* c, int ch) { return (size_t)ch * c->g.V * 12 * sizeof(float2); }

#define YH_NET_SHAPES(CI, CO, ...)                                             \
  if (CI == 1 && CO == 4) { constexpr int CI_ = 1, CO_ = 4; __VA_ARGS__; }       \
  else if (CI == 4 && CO == 4) { constexpr int CI_ = 4, CO_ = 4; __VA_ARGS__; }  \
  else if (CI == 4 && CO == 1) { constexpr int CI_ = 4, CO_ = 1; __VA_ARGS__; }  \
  else if (CI == 1 && CO == 8) { constexpr int CI_ = 1, CO_ = 8; __VA_ARGS__; }  \
  else if (CI == 8 && CO == 8) { constexpr int CI_ = 8, CO_ = 8; __VA_ARGS__; }  \
  else if (CI == 8 && CO == 1) { constexpr int CI_ = 8, CO_ = 1; __VA_ARGS__; }  \
  else if (CI == 1 && CO == 1) { constexpr int CI_ = 1, CO_ = 1; __VA_ARGS__; }  \
  else { set_error("network: unsupported channel pair"); return YH_ERR_UNSUPPORTED; }

static int bfrag_count(int ci, int co) {
  YH_NET_SHAPES(ci, co, return PtcShape<CI_, CO_>::BFRAG);
  return 0;
}

static int net_bmats(Net* n) {
  if (!n->bdirty) return YH_OK;
  Ctx* c = n->c;
  for (int l = 0; l < n->nl; l++) {
    int ci = n->ch[l], co = n->ch[l + 1];
    for (int adjoint = 0; adjoint < 2; adjoint++) {
      int ki = adjoint ? co : ci, ko = adjoint ? ci : co;
      int ks = 0, nb = 0;
      YH_NET_SHAPES(ki, ko, ks = PtcShape<CI_, CO_>::KS; nb = PtcShape<CI_, CO_>::NB);
      int tot = ks * nb * 64;
      hipLaunchKernelGGL(k_net_bmat, dim3((tot + 255) / 256), dim3(256), 0, c->stream, n->W + n->woff[l],
                         adjoint ? n->Ba[l] : n->Bf[l], ci, co, adjoint, ks, nb);
    }
  }
  YH_HIP(hipGetLastError());
  n->bdirty = false;
  return YH_OK;
}

// one layer application: in has ki channels, out ko channels
static int net_layer(Net* n, int l, int adjoint, const float2* in, float2* out, bool even_out) {
  Ctx* c = n->c;
  int ci = n->ch[l], co = n->ch[l + 1];
  int ki = adjoint ? co : ci, ko = adjoint ? ci : co;
  const float2* links = (const float2*)c->links[PREC_F32][RECON_18];
  int ntiles = (even_out ? c->g.VH : c->g.V) / 16;
  int grid = std::min((ntiles + 3) / 4, 2 * 256);
  YH_NET_SHAPES(ki, ko, {
    size_t lds = sizeof(uint4) * PtcShape<CI_, CO_>::BFRAG;
    static bool lds_set = false;  // the weight matrix can exceed 64 KiB of LDS
    if (!lds_set) {
      YH_HIP(hipFuncSetAttribute((const void*)k_ptc<CI_, CO_, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      YH_HIP(hipFuncSetAttribute((const void*)k_ptc<CI_, CO_, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      lds_set = true;
    }
    if (c->net_prefetch)
      hipLaunchKernelGGL((k_ptc<CI_, CO_, true>), dim3(grid), dim3(256), lds, c->stream, in, out,
                         adjoint ? n->Ba[l] : n->Bf[l], links, c->g, ntiles);
    else
      hipLaunchKernelGGL((k_ptc<CI_, CO_, false>), dim3(grid), dim3(256), lds, c->stream, in, out,
                         adjoint ? n->Ba[l] : n->Bf[l], links, c->g, ntiles);
  });
  YH_HIP(hipGetLastError());
  return YH_OK;
}

// A (adjoint = 0) or A^+ on the activation in channel-0 buffer `src`
// (odd sites zero); result (even sites valid) in the returned buffer.
// With keep, the per-layer inputs (forward) / output cotangents (adjoint)
// are copied to `keep`.
static int net_chain(Net* n, int adjoint, float2* src, float2** result, float2** keep) {
  Ctx* c = n->c;
  int rc;
  if ((rc = net_bmats(n))) return rc;
  float2* cur = src;
  for (int k = 0; k < n->nl; k++) {
    int l = adjoint ? n->nl - 1 - k : k;
    float2* dst = adjoint ? n->adj[l] : n->fwd[l + 1];
    if (keep) keep[l] = cur;
    bool last = k == n->nl - 1;
    if ((rc = net_layer(n, l, adjoint, cur, dst, last))) return rc;
    if (last) {
      // odd half of the single output channel is not written: clear it
      YH_HIP(hipMemsetAsync(dst + (size_t)c->g.VH * 12, 0, (size_t)c->g.VH * 12 * sizeof(float2), c->stream));
    }
    cur = dst;
  }
  *result = cur;
  return YH_OK;
}

static int field_to_act(Ctx* c, const Field* f, float2* act) {
  dim3 grid((c->g.V + 255) / 256);
  if (f->prec == PREC_F64)
    hipLaunchKernelGGL(k_field_to_act<PREC_F64>, grid, dim3(256), 0, c->stream, f->half[0], f->nrm[0], act, c->g);
  else if (f->prec == PREC_F32)
    hipLaunchKernelGGL(k_field_to_act<PREC_F32>, grid, dim3(256), 0, c->stream, f->half[0], f->nrm[0], act, c->g);
  else
    hipLaunchKernelGGL(k_field_to_act<PREC_F16>, grid, dim3(256), 0, c->stream, f->half[0], f->nrm[0], act, c->g);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

static int act_to_field(Ctx* c, const float2* act, Field* f) {
  dim3 grid((c->g.VH + 255) / 256);
  if (f->prec == PREC_F64)
    hipLaunchKernelGGL(k_act_to_field<PREC_F64>, grid, dim3(256), 0, c->stream, act, f->half[0], f->nrm[0], c->g);
  else if (f->prec == PREC_F32)
    hipLaunchKernelGGL(k_act_to_field<PREC_F32>, grid, dim3(256), 0, c->stream, act, f->half[0], f->nrm[0], c->g);
  else
    hipLaunchKernelGGL(k_act_to_field<PREC_F16>, grid, dim3(256), 0, c->stream, act, f->half[0], f->nrm[0], c->g);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

// gradient of Re<y, A w> for every layer, added (times coeff) to n->grad.
// fwd_in[l] = layer l input on w; cot[l] = cotangent at layer l output.
static int net_wgrad_all(Net* n, float2* const* fwd_in, float2* const* cot, double coeff) {
  Ctx* c = n->c;
  const float2* links = (const float2*)c->links[PREC_F32][RECON_18];
  int ntiles = c->g.V / 32;
  int grid = std::min(ntiles, n->wg_blocks);
  for (int l = 0; l < n->nl; l++) {
    int ci = n->ch[l], co = n->ch[l + 1];
    int nrows = 0, kcols = 0;
    YH_NET_SHAPES(ci, co, {
      nrows = PtcShape<CI_, CO_>::NROWS;
      kcols = PtcShape<CI_, CO_>::KCOLS;
      hipLaunchKernelGGL((k_ptc_wgrad<CI_, CO_>), dim3(grid), dim3(256), 0, c->stream, cot[l], fwd_in[l], links, c->g,
                         ntiles, n->partial);
    });
    int nwl = 9 * ci * co * 16;
    hipLaunchKernelGGL(k_wgrad_fold, dim3((nwl + 127) / 128), dim3(128), 0, c->stream, n->partial, grid, nrows, kcols,
                       ci, co, coeff, n->grad + n->woff[l]);
  }
  YH_HIP(hipGetLastError());
  return YH_OK;
}

}  // namespace yh

// ---------------------------------------------------------------------------
using namespace yh;

struct yh_net : public yh::Net {};

namespace {

void to_device_layout(const double* in, std::vector<float2>& out, int V, int ch) {
  // [V][ch][4][3] complex double -> [ch][V][3][4] complex float
  out.resize((size_t)ch * V * 12);
  for (int x = 0; x < V; x++)
    for (int k = 0; k < ch; k++)
      for (int s = 0; s < 4; s++)
        for (int a = 0; a < 3; a++) {
          const double* p = in + 2 * ((((size_t)x * ch + k) * 4 + s) * 3 + a);
          out[((size_t)k * V + x) * 12 + a * 4 + s] = make_float2((float)p[0], (float)p[1]);
        }
}

void from_device_layout(const std::vector<float2>& in, double* out, int V, int ch) {
  for (int x = 0; x < V; x++)
    for (int k = 0; k < ch; k++)
      for (int s = 0; s < 4; s++)
        for (int a = 0; a < 3; a++) {
          double* p = out + 2 * ((((size_t)x * ch + k) * 4 + s) * 3 + a);
          float2 v = in[((size_t)k * V + x) * 12 + a * 4 + s];
          p[0] = v.x;
          p[1] = v.y;
        }
}

int net_check(yh_net* n) {
  if (!n) { set_error("null network"); return YH_ERR_ARG; }
  if (!n->c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  YH_HIP(hipSetDevice(n->c->device));
  return YH_OK;
}

}  // namespace

extern "C" {

int yh_net_create(yh_ctx* c, int nlayers, const int* channels, yh_net** out) {
  if (!c || !channels || !out || nlayers < 1 || nlayers > 4) { set_error("bad argument"); return YH_ERR_ARG; }
  if (c->nc != 3 || c->repr != YH_REPR_FUNDAMENTAL) {
    set_error("the network is built for SU(3) fundamental fields");
    return YH_ERR_UNSUPPORTED;
  }
  if (c->g.V % 32 != 0 || c->g.VH % 16 != 0) { set_error("volume must be a multiple of 32"); return YH_ERR_UNSUPPORTED; }
  if (channels[0] != 1 || channels[nlayers] != 1) { set_error("first and last channel counts must be 1"); return YH_ERR_ARG; }
  for (int l = 0; l < nlayers; l++)
    if (bfrag_count(channels[l], channels[l + 1]) == 0 || bfrag_count(channels[l + 1], channels[l]) == 0) {
      set_error("unsupported channel counts (hidden layers: 4 or 8 channels, all equal)");
      return YH_ERR_UNSUPPORTED;
    }
  YH_HIP(hipSetDevice(c->device));
  yh_net* n = new yh_net();
  n->c = c;
  n->nl = nlayers;
  size_t off = 0;
  int maxch = 1;
  for (int l = 0; l <= nlayers; l++) {
    n->ch[l] = channels[l];
    maxch = std::max(maxch, channels[l]);
  }
  for (int l = 0; l < nlayers; l++) {
    n->woff[l] = off;
    off += (size_t)9 * channels[l] * channels[l + 1] * 16;
  }
  n->woff[nlayers] = off;
  n->nw = off;
  n->wg_blocks = 512;
  size_t maxpart = 0;
  for (int l = 0; l < nlayers; l++) {
    int ci = channels[l], co = channels[l + 1];
    size_t p = 0;
    YH_NET_SHAPES(ci, co, p = (size_t)PtcShape<CI_, CO_>::NROWS * PtcShape<CI_, CO_>::KCOLS);
    maxpart = std::max(maxpart, p);
  }
  n->partial_floats = maxpart * n->wg_blocks;
  hipError_t e = hipSuccess;
  auto A = [&](void** p, size_t b) { if (e == hipSuccess) e = hipMalloc(p, b); };
  A((void**)&n->W, n->nw * sizeof(float2));
  A((void**)&n->m1, n->nw * sizeof(float2));
  A((void**)&n->m2, n->nw * sizeof(float2));
  A((void**)&n->grad, n->nw * sizeof(double2));
  for (int l = 0; l < nlayers; l++) {
    A((void**)&n->Bf[l], sizeof(uint4) * bfrag_count(channels[l], channels[l + 1]));
    A((void**)&n->Ba[l], sizeof(uint4) * bfrag_count(channels[l + 1], channels[l]));
  }
  for (int l = 0; l <= nlayers; l++) {
    A((void**)&n->fwd[l], act_bytes(c, channels[l]));
    A((void**)&n->adj[l], act_bytes(c, channels[l]));
  }
  for (int l = 0; l < nlayers; l++) {
    A((void**)&n->keepU[l], act_bytes(c, channels[l]));
    A((void**)&n->keepB[l], act_bytes(c, channels[l + 1]));
  }
  A((void**)&n->ab, act_bytes(c, 1));
  A((void**)&n->ax, act_bytes(c, 1));
  A((void**)&n->partial, n->partial_floats * sizeof(float));
  if (e == hipSuccess) e = hipMemsetAsync(n->W, 0, n->nw * sizeof(float2), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(n->m1, 0, n->nw * sizeof(float2), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(n->m2, 0, n->nw * sizeof(float2), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(n->grad, 0, n->nw * sizeof(double2), c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    set_error(std::string("network allocation: ") + hipGetErrorString(e));
    yh_net_destroy(n);
    return YH_ERR_HIP;
  }
  (void)maxch;
  *out = n;
  return YH_OK;
}

int yh_net_destroy(yh_net* n) {
  if (!n) return YH_OK;
  (void)hipSetDevice(n->c->device);
  void* ptrs[] = {n->W, n->m1, n->m2, n->grad, n->ab, n->ax, n->partial};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int l = 0; l < 5; l++) {
    if (n->fwd[l]) (void)hipFree(n->fwd[l]);
    if (n->adj[l]) (void)hipFree(n->adj[l]);
  }
  for (int l = 0; l < 4; l++) {
    if (n->Bf[l]) (void)hipFree(n->Bf[l]);
    if (n->Ba[l]) (void)hipFree(n->Ba[l]);
    if (n->keepU[l]) (void)hipFree(n->keepU[l]);
    if (n->keepB[l]) (void)hipFree(n->keepB[l]);
  }
  delete n;
  return YH_OK;
}

long yh_net_num_weights(const yh_net* n) { return n ? (long)n->nw : -1; }

int yh_net_set_weights(yh_net* n, const double* w) {
  if (!n || !w) { set_error("null argument"); return YH_ERR_ARG; }
  YH_HIP(hipSetDevice(n->c->device));
  std::vector<float2> h(n->nw);
  for (size_t i = 0; i < n->nw; i++) h[i] = make_float2((float)w[2 * i], (float)w[2 * i + 1]);
  YH_HIP(hipMemcpyAsync(n->W, h.data(), n->nw * sizeof(float2), hipMemcpyHostToDevice, n->c->stream));
  YH_HIP(hipStreamSynchronize(n->c->stream));
  n->bdirty = true;
  n->adam_t = 0;
  YH_HIP(hipMemsetAsync(n->m1, 0, n->nw * sizeof(float2), n->c->stream));
  YH_HIP(hipMemsetAsync(n->m2, 0, n->nw * sizeof(float2), n->c->stream));
  YH_HIP(hipStreamSynchronize(n->c->stream));
  return YH_OK;
}

int yh_net_get_weights(yh_net* n, double* w) {
  if (!n || !w) { set_error("null argument"); return YH_ERR_ARG; }
  YH_HIP(hipSetDevice(n->c->device));
  std::vector<float2> h(n->nw);
  YH_HIP(hipMemcpyAsync(h.data(), n->W, n->nw * sizeof(float2), hipMemcpyDeviceToHost, n->c->stream));
  YH_HIP(hipStreamSynchronize(n->c->stream));
  for (size_t i = 0; i < n->nw; i++) {
    w[2 * i] = h[i].x;
    w[2 * i + 1] = h[i].y;
  }
  return YH_OK;
}

// one layer on host arrays in the layout [V][C][4][3] complex double
int yh_net_layer_host(yh_net* n, int layer, int adjoint, const double* in, double* out) {
  int rc = net_check(n);
  if (rc) return rc;
  if (layer < 0 || layer >= n->nl || !in || !out) { set_error("bad argument"); return YH_ERR_ARG; }
  if ((rc = net_bmats(n))) return rc;
  Ctx* c = n->c;
  int ki = adjoint ? n->ch[layer + 1] : n->ch[layer], ko = adjoint ? n->ch[layer] : n->ch[layer + 1];
  std::vector<float2> h;
  to_device_layout(in, h, c->g.V, ki);
  float2 *din = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc(&din, h.size() * sizeof(float2));
  if (e == hipSuccess) e = hipMalloc(&dout, act_bytes(c, ko));
  if (e == hipSuccess) e = hipMemcpyAsync(din, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) rc = net_layer(n, layer, adjoint, din, dout, false);
  std::vector<float2> o((size_t)ko * c->g.V * 12);
  if (e == hipSuccess && !rc)
    e = hipMemcpyAsync(o.data(), dout, o.size() * sizeof(float2), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (din) (void)hipFree(din);
  if (dout) (void)hipFree(dout);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return YH_ERR_HIP; }
  if (rc) return rc;
  from_device_layout(o, out, c->g.V, ko);
  return YH_OK;
}

// weight gradient of one layer on host arrays: g [V][cout][4][3], h [V][cin][4][3];
// grad [9][cout][cin][4][4] complex double (overwritten)
int yh_net_wgrad_host(yh_net* n, int layer, const double* g, const double* h, double* grad) {
  int rc = net_check(n);
  if (rc) return rc;
  if (layer < 0 || layer >= n->nl || !g || !h || !grad) { set_error("bad argument"); return YH_ERR_ARG; }
  Ctx* c = n->c;
  int ci = n->ch[layer], co = n->ch[layer + 1];
  std::vector<float2> hg, hh;
  to_device_layout(g, hg, c->g.V, co);
  to_device_layout(h, hh, c->g.V, ci);
  float2 *dg = nullptr, *dh = nullptr;
  hipError_t e = hipMalloc(&dg, hg.size() * sizeof(float2));
  if (e == hipSuccess) e = hipMalloc(&dh, hh.size() * sizeof(float2));
  if (e == hipSuccess) e = hipMemcpyAsync(dg, hg.data(), hg.size() * sizeof(float2), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dh, hh.data(), hh.size() * sizeof(float2), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(n->grad, 0, n->nw * sizeof(double2), c->stream);
  if (e == hipSuccess) {
    float2* fw[4] = {nullptr, nullptr, nullptr, nullptr};
    float2* ct[4] = {nullptr, nullptr, nullptr, nullptr};
    fw[layer] = dh;
    ct[layer] = dg;
    // run the single layer's gradient through the shared path
    const float2* links = (const float2*)c->links[PREC_F32][RECON_18];
    int ntiles = c->g.V / 32;
    int grid = std::min(ntiles, n->wg_blocks);
    int nrows = 0, kcols = 0;
    YH_NET_SHAPES(ci, co, {
      nrows = PtcShape<CI_, CO_>::NROWS;
      kcols = PtcShape<CI_, CO_>::KCOLS;
      hipLaunchKernelGGL((k_ptc_wgrad<CI_, CO_>), dim3(grid), dim3(256), 0, c->stream, ct[layer], fw[layer], links,
                         c->g, ntiles, n->partial);
    });
    int nwl = 9 * ci * co * 16;
    hipLaunchKernelGGL(k_wgrad_fold, dim3((nwl + 127) / 128), dim3(128), 0, c->stream, n->partial, grid, nrows, kcols,
                       ci, co, 1.0, n->grad + n->woff[layer]);
    e = hipGetLastError();
  }
  std::vector<double2> out((size_t)9 * ci * co * 16);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out.data(), n->grad + n->woff[layer], out.size() * sizeof(double2), hipMemcpyDeviceToHost,
                       c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (dg) (void)hipFree(dg);
  if (dh) (void)hipFree(dh);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return YH_ERR_HIP; }
  memcpy(grad, out.data(), out.size() * sizeof(double2));
  return YH_OK;
}

// dst = A src (adjoint = 0) or A^+ src on the even halves
int yh_net_apply(yh_net* n, int adjoint, yh_field* dst, const yh_field* src) {
  int rc = net_check(n);
  if (rc) return rc;
  if (!dst || !src) { set_error("null field"); return YH_ERR_ARG; }
  Ctx* c = n->c;
  float2* in = adjoint ? n->adj[n->nl] : n->fwd[0];
  if ((rc = field_to_act(c, src, in))) return rc;
  float2* res;
  if ((rc = net_chain(n, adjoint, in, &res, nullptr))) return rc;
  if ((rc = act_to_field(c, res, dst))) return rc;
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

}  // extern "C"

namespace yh {

// dst = A A^+ src on even halves, asynchronous on the context stream
int net_precondition(Net* n, Field* dst, const Field* src) {
  Ctx* c = n->c;
  int rc;
  if ((rc = field_to_act(c, src, n->adj[n->nl]))) return rc;
  float2* u;
  if ((rc = net_chain(n, 1, n->adj[n->nl], &u, nullptr))) return rc;
  // u is adj[0] (1 channel, odd cleared): feed it to the forward chain
  float2* f;
  if ((rc = net_chain(n, 0, u, &f, nullptr))) return rc;
  return act_to_field(c, f, dst);
}

// loss |A A^+ b - x|^2 / |x|^2 and its gradient, accumulated into n->grad
int net_loss_grad(Net* n, const Field* b, const Field* x, double* loss) {
  Ctx* c = n->c;
  int rc;
  const int nl = n->nl;
  if ((rc = field_to_act(c, b, n->ab))) return rc;
  if ((rc = field_to_act(c, x, n->ax))) return rc;
  // u = A^+ b, keeping the cotangents of b (keepB[l] = layer l output cotangent)
  float2* cotb[4];
  float2* u;
  if ((rc = net_chain(n, 1, n->ab, &u, cotb))) return rc;
  for (int l = 0; l < nl; l++)
    YH_HIP(hipMemcpyAsync(n->keepB[l], cotb[l], act_bytes(c, n->ch[l + 1]), hipMemcpyDeviceToDevice, c->stream));
  // f = A u, keeping the layer inputs on u
  float2* inu[4];
  float2* f;
  if ((rc = net_chain(n, 0, u, &f, inu))) return rc;
  for (int l = 0; l < nl; l++)
    YH_HIP(hipMemcpyAsync(n->keepU[l], inu[l], act_bytes(c, n->ch[l]), hipMemcpyDeviceToDevice, c->stream));
  // e = 2 (f - x) / |x|^2 on even sites, in place of f's buffer copy adj[nl]
  const int ne = c->g.VH * 12;
  float2* e = n->adj[nl];
  double* red = &c->sc->red[0];
  {
    int grid = std::min((ne + 255) / 256, c->nblk);
    hipLaunchKernelGGL(k_net_diff, dim3(grid), dim3(256), 0, c->stream, f, n->ax, e, ne, c->partials, c->ticket, red);
    hipLaunchKernelGGL(k_net_scale, dim3((ne + 255) / 256), dim3(256), 0, c->stream, e, ne, red, 2.0);
    YH_HIP(hipMemsetAsync(e + ne, 0, (size_t)ne * sizeof(float2), c->stream));
  }
  double r2[2];
  YH_HIP(hipMemcpyAsync(r2, red, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  // gradient part 1: WG(u, e): layer inputs on u (keepU), cotangents of e
  float2* cote[4];
  float2* v;
  if ((rc = net_chain(n, 1, e, &v, cote))) return rc;
  if ((rc = net_wgrad_all(n, n->keepU, cote, 1.0))) return rc;
  // gradient part 2: WG(v, b): layer inputs on v, cotangents of b (keepB)
  float2* inv[4];
  float2* dummy;
  YH_HIP(hipMemcpyAsync(n->ax, v, act_bytes(c, 1), hipMemcpyDeviceToDevice, c->stream));
  if ((rc = net_chain(n, 0, n->ax, &dummy, inv))) return rc;
  if ((rc = net_wgrad_all(n, inv, n->keepB, 1.0))) return rc;
  YH_HIP(hipStreamSynchronize(c->stream));
  if (loss) *loss = r2[0] / r2[1];
  return YH_OK;
}

}  // namespace yh

extern "C" {

int yh_net_precondition(yh_net* n, yh_field* dst, const yh_field* src) {
  int rc = net_check(n);
  if (rc) return rc;
  if (!dst || !src) { set_error("null field"); return YH_ERR_ARG; }
  if ((rc = net_precondition(n, dst, src))) return rc;
  YH_HIP(hipStreamSynchronize(n->c->stream));
  return YH_OK;
}

// loss and gradient summed over pairs; grad (optional) receives the summed
// complex gradient [nw] (re, im doubles)
int yh_net_loss_grad(yh_net* n, int npairs, const yh_field* const* b, const yh_field* const* x, double* loss,
                     double* grad) {
  int rc = net_check(n);
  if (rc) return rc;
  if (npairs < 1 || !b || !x) { set_error("bad argument"); return YH_ERR_ARG; }
  Ctx* c = n->c;
  YH_HIP(hipMemsetAsync(n->grad, 0, n->nw * sizeof(double2), c->stream));
  double tot = 0;
  for (int k = 0; k < npairs; k++) {
    double l;
    if ((rc = net_loss_grad(n, b[k], x[k], &l))) return rc;
    tot += l;
  }
  if (loss) *loss = tot;
  if (grad) {
    YH_HIP(hipMemcpyAsync(grad, n->grad, n->nw * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    YH_HIP(hipStreamSynchronize(c->stream));
  }
  return YH_OK;
}

// one Adam step on the mean gradient of the last yh_net_loss_grad call
int yh_net_adam_step(yh_net* n, double lr, double beta1, double beta2, double eps, int npairs) {
  int rc = net_check(n);
  if (rc) return rc;
  if (npairs < 1) { set_error("bad argument"); return YH_ERR_ARG; }
  n->adam_t++;
  float c1 = (float)(1.0 - pow(beta1, n->adam_t)), c2 = (float)(1.0 - pow(beta2, n->adam_t));
  int nw = (int)n->nw;
  hipLaunchKernelGGL(k_net_adam, dim3((nw + 255) / 256), dim3(256), 0, n->c->stream, n->W, n->m1, n->m2, n->grad, nw,
                     (float)lr, (float)beta1, (float)beta2, (float)eps, c1, c2, (float)(1.0 / npairs));
  YH_HIP(hipGetLastError());
  n->bdirty = true;
  YH_HIP(hipStreamSynchronize(n->c->stream));
  return YH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Flexible CG (Notay 2000; beta = <z', r' - r> / <z, r>) on H^2 in FP32 with
// the network preconditioner M = A A^+ (or none), inside FP64 reliable
// updates: the correction is folded into x and the residual recomputed
// with the FP64 operator whenever |r| has dropped by `delta` or
// `reliable_every` iterations have passed.  Stops on the FP64 true residual.
// ---------------------------------------------------------------------------
namespace {

double rd_sq(Ctx* c, const Field* a, int* rc) {
  double v[2] = {0, 0};
  *rc = reduce_read(c, 0, a, nullptr, 0, v);
  return v[0];
}

double rd_re(Ctx* c, const Field* a, const Field* b, int* rc) {
  double v[2] = {0, 0};
  *rc = reduce_read(c, 2, a, b, 0, v);
  return v[0];
}

}  // namespace

extern "C" int yh_cg_net(yh_ctx* cc, const yh_inv_par* par, yh_net* net, yh_field* dptr, const yh_field* sptr,
                         int* iters, int* updates) {
  Ctx* c = cc;
  if (!c || !par || !dptr || !sptr || !iters) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (net && net->c != c) { set_error("network belongs to another context"); return YH_ERR_ARG; }
  if (dptr->prec != PREC_F64 || sptr->prec != PREC_F64) { set_error("solver fields must be FP64"); return YH_ERR_ARG; }
  if (par->mvm != YH_MVM_HPHI_SQ) { set_error("the preconditioned CG needs Hphi_sq"); return YH_ERR_ARG; }
  if (!(par->prec > 0.0)) { set_error("prec must be positive"); return YH_ERR_ARG; }
  YH_HIP(hipSetDevice(c->device));
  int rc;
  Field *r64 = tmp_field(c, PREC_F64, 2, &rc), *t64 = nullptr;
  if (rc) return rc;
  t64 = tmp_field(c, PREC_F64, 3, &rc);
  if (rc) return rc;
  Field* f[6];
  for (int i = 0; i < 6; i++) {
    f[i] = tmp_field(c, PREC_F32, 2 + i, &rc);
    if (rc) return rc;
  }
  Field *xs = f[0], *r = f[1], *rp = f[2], *z = f[3], *p = f[4], *q = f[5];
  const int recon64 = default_recon(c, PREC_F64), recon32 = default_recon(c, PREC_F32);
  const int K = par->reliable_every > 0 ? par->reliable_every : 50;
  const double delta = (par->delta > 0.0 && par->delta < 1.0) ? par->delta : 0.1;
  Field* x = dptr;
  const double norm_b = rd_sq(c, sptr, &rc);
  if (rc) return rc;
  const double target = par->prec * norm_b;
  auto true_residual = [&](double* rr) -> int {
    int e;
    if ((e = apply_op(c, PREC_F64, recon64, YH_MVM_HPHI_SQ, par->mass, t64, x, nullptr, -1, nullptr))) return e;
    if ((e = blas_copy(c, r64, sptr, 0))) return e;
    if ((e = blas_axpby(c, r64, -1.0, t64, 1.0, 0))) return e;
    *rr = rd_sq(c, r64, &e);
    return e;
  };
  auto precond = [&]() -> int {
    if (net) return net_precondition(net, z, r);
    return blas_copy(c, z, r, 0);
  };
  double rr;
  if ((rc = true_residual(&rr))) return rc;
  *iters = 0;
  int nupd = 0;
  if (rr < target) {
    if (updates) *updates = 0;
    return YH_OK;
  }
  if ((rc = field_convert(c, r, 0, r64, 0)) || (rc = field_zero(c, xs, YH_EVEN)) || (rc = precond()) ||
      (rc = blas_copy(c, p, z, 0)))
    return rc;
  double rz = rd_re(c, z, r, &rc);
  if (rc) return rc;
  double maxrr = rr;
  int k = 0, at_update = 0;
  while (true) {
    if ((rc = apply_op(c, PREC_F32, recon32, YH_MVM_HPHI_SQ, par->mass, q, p, nullptr, -1, nullptr))) return rc;
    double pq = rd_re(c, p, q, &rc);
    if (rc) return rc;
    double alpha = rz / pq;
    if ((rc = blas_axpby(c, xs, alpha, p, 1.0, 0)) || (rc = blas_copy(c, rp, r, 0)) ||
        (rc = blas_axpby(c, r, -alpha, q, 1.0, 0)))
      return rc;
    double rrs = rd_sq(c, r, &rc);
    if (rc) return rc;
    k++;
    if (!(rrs == rrs) || !(alpha == alpha)) { set_error("cg_net: non-finite residual"); return YH_ERR_NONFINITE; }
    if (par->log)
      fprintf(stderr, "[INVERTER][DEBUG] cg_net: iter=%d rr=%.6e alpha=%.6e\n", k, rrs / norm_b, alpha);
    if (rrs < delta * delta * maxrr || k - at_update >= K || rrs < target) {
      // reliable update
      if ((rc = field_convert(c, t64, 0, xs, 0)) || (rc = blas_axpby(c, x, 1.0, t64, 1.0, 0))) return rc;
      if ((rc = true_residual(&rr))) return rc;
      nupd++;
      if (!(rr == rr)) { set_error("cg_net: non-finite true residual"); return YH_ERR_NONFINITE; }
      if (rr < target) break;
      if ((rc = field_convert(c, r, 0, r64, 0)) || (rc = field_zero(c, xs, YH_EVEN))) return rc;
      maxrr = rr;
      at_update = k;
    }
    if (par->max_iter > 0 && k >= par->max_iter) {
      *iters = k;
      if (updates) *updates = nupd;
      set_error("cg_net: max_iter reached");
      return YH_ERR_NOCONV;
    }
    if ((rc = precond())) return rc;
    double rzn = rd_re(c, z, r, &rc);
    if (rc) return rc;
    double zrp = rd_re(c, z, rp, &rc);
    if (rc) return rc;
    double beta = (rzn - zrp) / rz;
    rz = rzn;
    if ((rc = blas_axpby(c, p, 1.0, z, beta, 0))) return rc;
  }
  *iters = k;
  if (updates) *updates = nupd;
  if (par->log)
    fprintf(stderr, "[INVERTER][NOTICE] cg_net: iterations = %d, updates = %d, precision = %1.4e\n", k, nupd,
            rr / norm_b);
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

// ---------------------------------------------------------------------------
// Flexible GCR (Eisenstat, Elman & Schultz 1983; flexible form of Saad's
// FGMRES family: each direction is the preconditioned residual, so M may
// change from step to step) on H^2 in FP32 with the network preconditioner
// M = A A^+ (or none), inside the same FP64 reliable updates as the
// flexible CG.  Per step:
//   z = M r;  w = H^2 z;  for the kept j: c = (w_j^H w) / |w_j|^2,
//   w -= c w_j, z -= c z_j;  a = (w^H r) / |w|^2;  x += a z;  r -= a w
// which minimises |r| over the kept directions (the w_j are orthogonal).
// At most `nkrylov` directions are kept; the oldest is dropped first
// (truncated GCR), and a reliable update restarts the history.
// ---------------------------------------------------------------------------
extern "C" int yh_gcr_net(yh_ctx* cc, const yh_inv_par* par, yh_net* net, int nkrylov, yh_field* dptr,
                          const yh_field* sptr, int* iters, int* updates) {
  Ctx* c = cc;
  if (!c || !par || !dptr || !sptr || !iters) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (net && net->c != c) { set_error("network belongs to another context"); return YH_ERR_ARG; }
  if (dptr->prec != PREC_F64 || sptr->prec != PREC_F64) { set_error("solver fields must be FP64"); return YH_ERR_ARG; }
  if (par->mvm != YH_MVM_HPHI_SQ) { set_error("the preconditioned GCR needs Hphi_sq"); return YH_ERR_ARG; }
  if (!(par->prec > 0.0)) { set_error("prec must be positive"); return YH_ERR_ARG; }
  if (nkrylov < 1 || nkrylov > 32) { set_error("nkrylov must be 1..32"); return YH_ERR_ARG; }
  YH_HIP(hipSetDevice(c->device));
  int rc;
  Field *r64 = tmp_field(c, PREC_F64, 2, &rc), *t64 = nullptr;
  if (rc) return rc;
  t64 = tmp_field(c, PREC_F64, 3, &rc);
  if (rc) return rc;
  Field *xs = tmp_field(c, PREC_F32, 2, &rc), *r = nullptr;
  if (rc) return rc;
  r = tmp_field(c, PREC_F32, 3, &rc);
  if (rc) return rc;
  // the kept directions z_j and w_j = H^2 z_j, plus one slot for the new pair
  std::vector<Field*> Z(nkrylov + 1, nullptr), W(nkrylov + 1, nullptr);
  std::vector<double> wn(nkrylov + 1, 0.0);
  auto release = [&]() {
    for (auto f : Z) field_free(f);
    for (auto f : W) field_free(f);
  };
  for (int j = 0; j <= nkrylov; j++)
    if ((rc = field_alloc(c, PREC_F32, &Z[j])) || (rc = field_alloc(c, PREC_F32, &W[j]))) {
      release();
      return rc;
    }
  const int recon64 = default_recon(c, PREC_F64), recon32 = default_recon(c, PREC_F32);
  const int K = par->reliable_every > 0 ? par->reliable_every : 50;
  const double delta = (par->delta > 0.0 && par->delta < 1.0) ? par->delta : 0.1;
  Field* x = dptr;
  int k = 0, nupd = 0;
  auto run = [&]() -> int {
    int e;
    const double norm_b = rd_sq(c, sptr, &e);
    if (e) return e;
    const double target = par->prec * norm_b;
    auto true_residual = [&](double* rr) -> int {
      int q;
      if ((q = apply_op(c, PREC_F64, recon64, YH_MVM_HPHI_SQ, par->mass, t64, x, nullptr, -1, nullptr))) return q;
      if ((q = blas_copy(c, r64, sptr, 0))) return q;
      if ((q = blas_axpby(c, r64, -1.0, t64, 1.0, 0))) return q;
      *rr = rd_sq(c, r64, &q);
      return q;
    };
    double rr;
    if ((e = true_residual(&rr))) return e;
    if (rr < target) return YH_OK;
    if ((e = field_convert(c, r, 0, r64, 0)) || (e = field_zero(c, xs, YH_EVEN))) return e;
    double maxrr = rr;
    int kept = 0, head = 0, at_update = 0;  // ring of kept directions, `head` = next free slot
    while (true) {
      Field* z = Z[head];
      Field* w = W[head];
      if (net) {
        if ((e = net_precondition(net, z, r))) return e;
      } else if ((e = blas_copy(c, z, r, 0))) {
        return e;
      }
      if ((e = apply_op(c, PREC_F32, recon32, YH_MVM_HPHI_SQ, par->mass, w, z, nullptr, -1, nullptr))) return e;
      for (int i = 0; i < kept; i++) {
        const int j = (head - 1 - i + 2 * (nkrylov + 1)) % (nkrylov + 1);
        double d[2];
        if ((e = reduce_read(c, 1, w, W[j], 0, d))) return e;  // sum w conj(w_j) = w_j^H w
        const double cr = d[0] / wn[j], ci = d[1] / wn[j];
        if ((e = blas_caxpy(c, w, -cr, -ci, W[j], 0)) || (e = blas_caxpy(c, z, -cr, -ci, Z[j], 0))) return e;
      }
      double ww = rd_sq(c, w, &e);
      if (e) return e;
      double d[2];
      if ((e = reduce_read(c, 1, r, w, 0, d))) return e;  // w^H r
      const double ar = d[0] / ww, ai = d[1] / ww;
      if (!(ar == ar) || !(ai == ai)) { set_error("gcr_net: non-finite coefficient"); return YH_ERR_NONFINITE; }
      if ((e = blas_caxpy(c, xs, ar, ai, z, 0)) || (e = blas_caxpy(c, r, -ar, -ai, w, 0))) return e;
      wn[head] = ww;
      head = (head + 1) % (nkrylov + 1);
      kept = kept < nkrylov ? kept + 1 : nkrylov;
      k++;
      const double rrs = rd_sq(c, r, &e);
      if (e) return e;
      if (!(rrs == rrs)) { set_error("gcr_net: non-finite residual"); return YH_ERR_NONFINITE; }
      if (par->log) fprintf(stderr, "[INVERTER][DEBUG] gcr_net: iter=%d rr=%.6e\n", k, rrs / norm_b);
      if (rrs < delta * delta * maxrr || k - at_update >= K || rrs < target) {
        // reliable update: fold the FP32 correction into x, recompute r in
        // FP64 and restart the direction history
        if ((e = field_convert(c, t64, 0, xs, 0)) || (e = blas_axpby(c, x, 1.0, t64, 1.0, 0))) return e;
        if ((e = true_residual(&rr))) return e;
        nupd++;
        if (!(rr == rr)) { set_error("gcr_net: non-finite true residual"); return YH_ERR_NONFINITE; }
        if (rr < target) return YH_OK;
        if ((e = field_convert(c, r, 0, r64, 0)) || (e = field_zero(c, xs, YH_EVEN))) return e;
        maxrr = rr;
        at_update = k;
        kept = 0;
      }
      if (par->max_iter > 0 && k >= par->max_iter) {
        set_error("gcr_net: max_iter reached");
        return YH_ERR_NOCONV;
      }
    }
  };
  rc = run();
  (void)hipStreamSynchronize(c->stream);
  release();
  *iters = k;
  if (updates) *updates = nupd;
  if (par->log && !rc) fprintf(stderr, "[INVERTER][NOTICE] gcr_net: iterations = %d, updates = %d\n", k, nupd);
  return rc;
}
