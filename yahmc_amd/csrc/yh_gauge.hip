// Gauge-side kernels (FP64): representation of the links, clover term,
// odd-site clover inverse and log det, and the derived low-precision copies.
//
// represent  : library/repr.cpp:107-280 (written as the action of U on the
//              representation basis, repr.cpp:25-61)
// clover     : library/wilsonloop.cpp:5-31 + library/clover.cpp:93-168
// inverse    : library/clover.cpp:25-91 + library/solver.cpp:97-142 (the
//              same unpivoted LDL^+ factorisation, then the explicit
//              Hermitian inverse is stored instead of the factors, so the
//              apply is one Hermitian mat-vec)
// logdet     : library/clover.cpp:281-305
#include <math.h>
#include <string.h>
#include <vector>
#include "yh_internal.hpp"

namespace yh {

// representation basis b_n (DR <= 15, NC <= 4): the context's table
// tab_basis[n * 16 + row * NC + col]

static const double kSqrtHalf = 0.70710678118654752440;

// anti-Hermitian generator matrix for algebra vector e_n
// (library/suNg.cpp:15-58)
static void alg_unit(int nc, int n, std::vector<double2>& m) {
  m.assign(nc * nc, make_double2(0, 0));
  int k = 0;
  for (int i = 0; i < nc; i++)
    for (int j = 0; j < nc; j++) {
      if (i > j) {
        if (k == n) { m[i * nc + j].y += 1; m[j * nc + i].y += 1; }
        k++;
      } else if (i < j) {
        if (k == n) { m[i * nc + j].x += 1; m[j * nc + i].x -= 1; }
        k++;
      } else if (i > 0) {
        if (k == n) {
          double v = sqrt(2.0 / (i * (i + 1)));
          for (int q = 0; q < i; q++) m[q * nc + q].y += v;
          m[i * nc + i].y -= i * v;
        }
        k++;
      }
    }
}

int gauge_init_basis(Ctx* c) {
  int nc = c->nc, dr = c->dr;
  std::vector<double2> b(15 * 16, make_double2(0, 0));
  if (c->repr == YH_REPR_SYMMETRIC) {
    for (int i = 0, n = 0; i < nc; i++)
      for (int j = 0; j <= i; j++, n++) {
        if (i == j) b[n * 16 + i * nc + i].x = 1;
        else { b[n * 16 + i * nc + j].x = kSqrtHalf; b[n * 16 + j * nc + i].x = kSqrtHalf; }
      }
  } else if (c->repr == YH_REPR_ANTISYMMETRIC) {
    for (int i = 0, n = 0; i < nc; i++)
      for (int j = 0; j < i; j++, n++) {
        b[n * 16 + j * nc + i].x = kSqrtHalf;
        b[n * 16 + i * nc + j].x = -kSqrtHalf;
      }
  } else if (c->repr == YH_REPR_ADJOINT) {
    std::vector<double2> g;
    for (int n = 0; n < dr; n++) {
      alg_unit(nc, n, g);
      // (-i / sqrt 2) * g
      for (int k = 0; k < nc * nc; k++) b[n * 16 + k] = make_double2(kSqrtHalf * g[k].y, -kSqrtHalf * g[k].x);
    }
  }
  return upload_table(c, c->tab_basis, b.data(), b.size());
}

typedef Cx<double> cd;

template <int NC>
__device__ __forceinline__ void load_fund(cd (&u)[NC][NC], const double* fund, int id, int mu) {
  const double* p = fund + (size_t)(4 * id + mu) * 2 * NC * NC;
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) u[i][j] = cd(p[i * NC + j], p[NC * NC + i * NC + j]);
}

template <int NC>
__device__ __forceinline__ void mm(cd (&r)[NC][NC], const cd (&a)[NC][NC], const cd (&b)[NC][NC]) {
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) {
      cd s(0, 0);
#pragma unroll
      for (int k = 0; k < NC; k++) cmac(s, a[i][k], b[k][j]);
      r[i][j] = s;
    }
}
// r = a * b^+
template <int NC>
__device__ __forceinline__ void mmd(cd (&r)[NC][NC], const cd (&a)[NC][NC], const cd (&b)[NC][NC]) {
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) {
      cd s(0, 0);
#pragma unroll
      for (int k = 0; k < NC; k++) cmac(s, a[i][k], conj(b[j][k]));
      r[i][j] = s;
    }
}
// r = a^+ * b
template <int NC>
__device__ __forceinline__ void mdm(cd (&r)[NC][NC], const cd (&a)[NC][NC], const cd (&b)[NC][NC]) {
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) {
      cd s(0, 0);
#pragma unroll
      for (int k = 0; k < NC; k++) cmac_conj(s, a[k][i], b[k][j]);
      r[i][j] = s;
    }
}

// R(U) for the fermion representation
template <int NC, int REPR, int D>
__device__ __forceinline__ void represent(cd (&R)[D][D], const cd (&U)[NC][NC], const double2* __restrict__ basis) {
  if constexpr (REPR == YH_REPR_FUNDAMENTAL) {
#pragma unroll
    for (int i = 0; i < NC; i++)
#pragma unroll
      for (int j = 0; j < NC; j++) R[i][j] = U[i][j];
  } else {
    for (int m = 0; m < D; m++) {
      // W = U b_m ; X = W U^T (sym/asym) or W U^+ (adjoint)
      cd W[NC][NC], X[NC][NC];
#pragma unroll
      for (int a = 0; a < NC; a++)
#pragma unroll
        for (int d = 0; d < NC; d++) {
          cd s(0, 0);
#pragma unroll
          for (int q = 0; q < NC; q++) {
            double2 bb = basis[m * 16 + q * NC + d];
            cmac(s, U[a][q], cd(bb.x, bb.y));
          }
          W[a][d] = s;
        }
#pragma unroll
      for (int a = 0; a < NC; a++)
#pragma unroll
        for (int b = 0; b < NC; b++) {
          cd s(0, 0);
#pragma unroll
          for (int d = 0; d < NC; d++) {
            if constexpr (REPR == YH_REPR_ADJOINT) cmac(s, W[a][d], conj(U[b][d]));
            else cmac(s, W[a][d], U[b][d]);
          }
          X[a][b] = s;
        }
      for (int n = 0; n < D; n++) {
        cd s(0, 0);
#pragma unroll
        for (int a = 0; a < NC; a++)
#pragma unroll
          for (int b = 0; b < NC; b++) {
            if constexpr (REPR == YH_REPR_ADJOINT) {
              double2 bb = basis[n * 16 + b * NC + a];  // Tr(b_n X) = sum (b_n)_{ba} X_{ab}
              cmac(s, cd(bb.x, bb.y), X[a][b]);
            } else {
              double2 bb = basis[n * 16 + a * NC + b];
              cmac_conj(s, cd(bb.x, bb.y), X[a][b]);
            }
          }
        R[n][m] = s;
      }
    }
  }
}

// fund links (reference AoS) -> F64 recon-18 SoA links (no BC sign)
template <int NC, int REPR, int D>
__global__ void k_represent(const double* __restrict__ fund, double2* __restrict__ out, Geom g,
                            const double2* __restrict__ basis) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * g.V) return;
  int id = t >> 2, mu = t & 3;
  cd U[NC][NC], R[D][D];
  load_fund<NC>(U, fund, id, mu);
  represent<NC, REPR, D>(R, U, basis);
  int p = id >= g.VH;
  int k = id - p * g.VH;
#pragma unroll
  for (int e = 0; e < D * D; e++)
    out[((size_t)(p * 4 + mu) * D * D + e) * g.VHpad + k] = make_double2(R[e / D][e % D].re, R[e / D][e % D].im);
}

__device__ __forceinline__ void coords_of(int id, const Geom& g, int c[4]) { full_coords(id, g, c); }

__device__ __forceinline__ int shift(const int c[4], int mu, int d, const Geom& g) {
  int n[4] = {c[0], c[1], c[2], c[3]};
  n[mu] = (n[mu] + d + g.L[mu]) % g.L[mu];
  return full_index(n[0], n[1], n[2], n[3], g);
}

// packed Hermitian store: n reals then strictly-lower triangle
template <int N>
__device__ __forceinline__ void store_packed(double2* base, int stride, int site, int vec0, const cd (&M)[N][N]) {
  double v[N * N];
  int o = 0;
#pragma unroll
  for (int i = 0; i < N; i++) v[o++] = M[i][i].re;
#pragma unroll
  for (int i = 1; i < N; i++)
#pragma unroll
    for (int j = 0; j < i; j++) { v[o++] = M[i][j].re; v[o++] = M[i][j].im; }
#pragma unroll
  for (int k = 0; k < N * N / 2; k++) base[(size_t)(vec0 + k) * stride + site] = make_double2(v[2 * k], v[2 * k + 1]);
}

template <int N>
__device__ __forceinline__ void load_packed(cd (&M)[N][N], const double2* base, int stride, int site, int vec0) {
  double v[N * N];
#pragma unroll
  for (int k = 0; k < N * N / 2; k++) {
    double2 w = base[(size_t)(vec0 + k) * stride + site];
    v[2 * k] = w.x;
    v[2 * k + 1] = w.y;
  }
  int o = 0;
#pragma unroll
  for (int i = 0; i < N; i++) M[i][i] = cd(v[o++], 0);
#pragma unroll
  for (int i = 1; i < N; i++)
#pragma unroll
    for (int j = 0; j < i; j++) {
      M[i][j] = cd(v[o], v[o + 1]);
      M[j][i] = cd(v[o], -v[o + 1]);
      o += 2;
    }
}

// ---------------------------------------------------------------------------
// Clover term per site -> packed blocks [parity][2 blocks * n^2/2 vec][site]
// (library/clover.cpp:120-168; the dense form is -(c_sw/16) sum_{mu<nu}
// (1/2)[g_mu, g_nu] (x) (Q_mu_nu - Q_mu_nu^+), Q_mu_nu the represented sum
// of the four plaquette leaves, wilsonloop.cpp:5-31).  A group of D_R lanes
// per site, lane i computing row i of every D_R x D_R matrix, so that no
// lane holds more than a few rows.  Per plane, lanes 0..3 of the group each
// form one leaf in the fundamental as a product of four links,
// L = F1 F2 F3 F4, and put it in LDS; every lane then adds row i of R(L) for
// the four leaves:
//   fundamental: R = L
//   two-index:   R[i][m] = sum_cd (b_m)_cd Y_cd,  Y = L^T conj(b_i) L
//   adjoint:     R[i][m] = sum_cd Y_cd (b_m)_dc,  Y = L^+ b_i L
// (both follow from represent()).  The at/bt/ct/dt combinations need column
// i of the plane sum too, which the group exchanges through LDS.  The block
// is one wave; LDS traffic needs only wavefront-scope fences (a wave's LDS
// operations execute in order).
// LDS exchange between lanes of one wave: a wavefront-scope fence (the
// wave's LDS operations execute in order, so it emits no wait; it keeps the
// compiler from moving LDS accesses across it)
__device__ __forceinline__ void cb_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

template <int NC, int REPR, int D>
__global__ void __launch_bounds__(64) k_clover_build_rows(const double* __restrict__ fund, double2* __restrict__ out,
                                                         Geom g, double csw, const double2* __restrict__ basis) {
  constexpr int G = 64 / D, N = 2 * D, NN = NC * NC;
  __shared__ double2 lv[G][4][NN];  // the plane's four leaves
  __shared__ double2 qx[G][D * D];  // row exchange
  __shared__ double2 bl[D * 16];    // the representation basis
  const int lane = threadIdx.x;
  for (int e = lane; e < D * 16; e += 64) bl[e] = basis[e];
  cb_fence();
  const int gi = min(lane / D, G - 1);
  const int i = lane % D;
  const int grp = blockIdx.x * G + gi;
  const bool on = lane < G * D && grp < g.V;
  const int id = grp < g.V ? grp : g.V - 1;  // tail lanes compute on a valid site and write nothing
  int c[4];
  full_coords(id, g, c);
  bool zero = csw == 0.0;
  if (g.bc == BC_SF) zero = zero || c[0] == 0 || c[0] == g.L[0] - 1 || c[0] == g.L[0] - 2;
  if (g.bc == BC_OPEN) zero = zero || c[0] == 0 || c[0] == g.L[0] - 1;
  cd r1[D], r2[D];  // at, bt rows, then ct, dt rows
  const double cc = -csw / 16.0;
  const int p = id >= g.VH, k = id - p * g.VH;
  double2* base = out + (size_t)p * (N * N) * g.VHpad;
  // output addresses are formed where they are stored (from an opaque base:
  // hoisted out of the plane loop they would hold dozens of registers)
  double2* ob = base;
  auto put_re = [&](int vec0, int o, double v) {
    if (on) ((double*)(ob + (size_t)(vec0 + o / 2) * g.VHpad + k))[o & 1] = v;
  };
  auto put_cx = [&](int vec0, int o, cd v) {  // o even
    if (on) ob[(size_t)(vec0 + o / 2) * g.VHpad + k] = make_double2(v.re, v.im);
  };
  auto lo = [](int r, int s) { return N + 2 * (r * (r - 1) / 2 + s); };  // packed index of (r, s), r > s
  // planes 03, 12, 01, 02, 13, 23 (clover.cpp:135-156)
#pragma unroll 1
  for (int pl = 0; pl < 6; pl++) {
    const int mu = pl == 0 ? 0 : pl == 1 ? 1 : pl == 2 ? 0 : pl == 3 ? 0 : pl == 4 ? 1 : 2;
    const int nu = pl == 0 ? 3 : pl == 1 ? 2 : pl == 2 ? 1 : pl == 3 ? 2 : pl == 4 ? 3 : 3;
    // leaves: lane q < 4 of the group forms leaf q
    for (int q = i; q < 4; q += D) {
      int xm = shift(c, mu, 1, g), xn = shift(c, nu, 1, g);
      int bm = shift(c, mu, -1, g), bn = shift(c, nu, -1, g);
      int cb[4], cn[4];
      coords_of(bm, g, cb);
      coords_of(bn, g, cn);
      const int bm_pn = shift(cb, nu, 1, g), bm_mn = shift(cb, nu, -1, g), bn_pm = shift(cn, mu, 1, g);
      // the four leaves as products F1 F2 F3 F4 (leaf 0: U(x,mu) U(x+mu,nu)
      // U(x+nu,mu)^+ U(x,nu)^+, and the others around x)
      int s[4], d[4];
      bool h[4];
      if (q == 0) {
        s[0] = id; d[0] = mu; h[0] = false; s[1] = xm; d[1] = nu; h[1] = false;
        s[2] = xn; d[2] = mu; h[2] = true; s[3] = id; d[3] = nu; h[3] = true;
      } else if (q == 1) {
        s[0] = id; d[0] = nu; h[0] = false; s[1] = bm_pn; d[1] = mu; h[1] = true;
        s[2] = bm; d[2] = nu; h[2] = true; s[3] = bm; d[3] = mu; h[3] = false;
      } else if (q == 2) {
        s[0] = bm; d[0] = mu; h[0] = true; s[1] = bm_mn; d[1] = nu; h[1] = true;
        s[2] = bm_mn; d[2] = mu; h[2] = false; s[3] = bn; d[3] = nu; h[3] = false;
      } else {
        s[0] = bn; d[0] = nu; h[0] = true; s[1] = bn; d[1] = mu; h[1] = false;
        s[2] = bn_pm; d[2] = nu; h[2] = false; s[3] = id; d[3] = mu; h[3] = true;
      }
      // L = F1 F2 F3 F4, each factor a link or its dagger, multiplied in
      // place row by row
      cd L[NC][NC], F[NC][NC];
      auto ldf = [&](cd (&u)[NC][NC], int st, int dr, bool dg) {
        const double* pf = fund + (size_t)(4 * st + dr) * 2 * NC * NC;
#pragma unroll
        for (int a = 0; a < NC; a++)
#pragma unroll
          for (int b = 0; b < NC; b++) {
            const int e = dg ? b * NC + a : a * NC + b;
            u[a][b] = cd(pf[e], dg ? -pf[NC * NC + e] : pf[NC * NC + e]);
          }
      };
      ldf(L, s[0], d[0], h[0]);
#pragma unroll
      for (int f = 1; f < 4; f++) {
        ldf(F, s[f], d[f], h[f]);
#pragma unroll
        for (int a = 0; a < NC; a++) {
          cd row[NC];
#pragma unroll
          for (int b = 0; b < NC; b++) {
            cd t(0, 0);
#pragma unroll
            for (int kk = 0; kk < NC; kk++) cmac(t, L[a][kk], F[kk][b]);
            row[b] = t;
          }
#pragma unroll
          for (int b = 0; b < NC; b++) L[a][b] = row[b];
        }
      }
#pragma unroll
      for (int a = 0; a < NC; a++)
#pragma unroll
        for (int b = 0; b < NC; b++) lv[gi][q][a * NC + b] = make_double2(L[a][b].re, L[a][b].im);
    }
    cb_fence();
    // row i of the plane sum Q = sum_leaves R(L)
    cd qr[D];
#pragma unroll
    for (int m = 0; m < D; m++) qr[m] = cd(0, 0);
#pragma unroll 1
    for (int q = 0; q < 4 && REPR == YH_REPR_FUNDAMENTAL; q++) {
      // R(L) = L: row i straight from LDS
#pragma unroll
      for (int m = 0; m < D; m++) {
        const double2 v = lv[gi][q][i * NC + m];
        qr[m] = qr[m] + cd(v.x, v.y);
      }
    }
#pragma unroll 1
    for (int q = 0; q < 4 && REPR != YH_REPR_FUNDAMENTAL; q++) {
      cd L[NC][NC];
#pragma unroll
      for (int a = 0; a < NC; a++)
#pragma unroll
        for (int b = 0; b < NC; b++) {
          const double2 v = lv[gi][q][a * NC + b];
          L[a][b] = cd(v.x, v.y);
        }
      // column dd of T = b_i' L and of Y = L^T T (two-index, b_i' = conj b_i)
      // or L^+ T (adjoint), and its terms of row i of R(L)
#pragma unroll
      for (int dd = 0; dd < NC; dd++) {
        // basis from LDS through an opaque zero per column, so that the
        // compiler neither hoists the reads out of the loops nor issues them
        // all at once
        int z = 0;
        __asm__ volatile("" : "+v"(z));
        const double2* bs = bl + z;
        cd tc[NC], yc[NC];
#pragma unroll
        for (int a = 0; a < NC; a++) {
          cd t(0, 0);
#pragma unroll
          for (int b = 0; b < NC; b++) {
            const double2 v = bs[i * 16 + a * NC + b];
            cmac(t, REPR == YH_REPR_ADJOINT ? cd(v.x, v.y) : cd(v.x, -v.y), L[b][dd]);
          }
          tc[a] = t;
        }
#pragma unroll
        for (int cc2 = 0; cc2 < NC; cc2++) {
          cd y(0, 0);
#pragma unroll
          for (int a = 0; a < NC; a++) {
            if constexpr (REPR == YH_REPR_ADJOINT) cmac_conj(y, L[a][cc2], tc[a]);
            else cmac(y, L[a][cc2], tc[a]);
          }
          yc[cc2] = y;
        }
#pragma unroll
        for (int m = 0; m < D; m++) {
#pragma unroll
          for (int cc2 = 0; cc2 < NC; cc2++) {
            const double2 bb = REPR == YH_REPR_ADJOINT ? bs[m * 16 + dd * NC + cc2] : bs[m * 16 + cc2 * NC + dd];
            cmac(qr[m], yc[cc2], cd(bb.x, bb.y));
          }
        }
      }
    }
    // column i through LDS
    cb_fence();
#pragma unroll
    for (int m = 0; m < D; m++) qx[gi][i * D + m] = make_double2(qr[m].re, qr[m].im);
    cb_fence();
    cd qc[D];
#pragma unroll
    for (int m = 0; m < D; m++) {
      const double2 v = qx[gi][m * D + i];
      qc[m] = cd(v.x, v.y);
    }
    cb_fence();
    // Q[i][j] = qr[j], Q[j][i] = qc[j] (clover.cpp:149-156)
#pragma unroll
    for (int j = 0; j < D; j++) {
      const cd a = qr[j], b = qc[j];
      if (pl == 0) r1[j] = cd(a.im + b.im, b.re - a.re);                    // at
      else if (pl == 1) r2[j] = cd(a.im + b.im, a.re - b.re);               // bt
      else if (pl == 2) r1[j] = cd(a.im + b.im, b.re - a.re);               // ct
      else if (pl == 3) r1[j] = r1[j] + cd(-a.re + b.re, -a.im - b.im);     // ct
      else if (pl == 4) r2[j] = cd(a.re - b.re, a.im + b.im);               // dt
      else r2[j] = r2[j] + cd(a.im + b.im, -a.re + b.re);                   // dt
    }
    if (pl == 1 || pl == 5) {
      int z = 0;
      __asm__ volatile("" : "+v"(z));
      ob = base + z;
    }
    if (pl == 1) {
      // A = cc (at - bt) (block 0, diagonal blocks), C (block 1)
#pragma unroll
      for (int j = 0; j < D; j++) {
        const cd at = r1[j], bt = r2[j];
        cd A(cc * (at.re - bt.re), cc * (at.im + bt.im));
        cd C(-cc * (at.re + bt.re), cc * (bt.im - at.im));
        if (zero) A = C = cd(0, 0);
        if (j == i) {
          put_re(0, i, A.re);
          put_re(0, D + i, -A.re);
          put_re(N * N / 2, i, C.re);
          put_re(N * N / 2, D + i, -C.re);
        } else if (j < i) {
          put_cx(0, lo(i, j), A);
          put_cx(0, lo(D + i, D + j), cd(-A.re, -A.im));
          put_cx(N * N / 2, lo(i, j), C);
          put_cx(N * N / 2, lo(D + i, D + j), cd(-C.re, -C.im));
        }
      }
    } else if (pl == 5) {
      // B (block 0) and D (block 1) off-diagonal blocks, stored as their
      // conjugates in the lower left
#pragma unroll
      for (int j = 0; j < D; j++) {
        const cd ct = r1[j], dt = r2[j];
        cd B(cc * (ct.re - dt.re), cc * (ct.im - dt.im));
        cd Dm(-cc * (ct.re + dt.re), -cc * (ct.im + dt.im));
        if (zero) B = Dm = cd(0, 0);
        put_cx(0, lo(D + j, i), conj(B));
        put_cx(N * N / 2, lo(D + j, i), conj(Dm));
      }
    }
  }
}

// odd-site inverse of sigma + c_f shift + clover, per chiral block, and its
// log det (the reference factors with an unpivoted LDL^+, solver.cpp:97-117).
// A group of N = 2 D_R lanes per odd site, lane r holding row r of the
// block: in-place Gauss-Jordan without pivoting, whose pivots are the d_i of
// the unpivoted LDL^+ (both are the ratios of the leading principal minors),
// so log det = sum log |d_i|.  Each
// step's normalised pivot row reaches the group through LDS (the block is
// one wave: a wave's LDS operations execute in order, wavefront-scope
// fences suffice).  The two chiral blocks of a site run one after the other in
// the same group, which then sums the 2N log |d_i| in fixed order.
template <int D>
__global__ void __launch_bounds__(64) k_clover_invert_rows(const double2* __restrict__ clov, double2* __restrict__ inv,
                                                          double* __restrict__ logdet, Geom g, double sigma) {
  constexpr int N = 2 * D, G = 64 / N;
  __shared__ double2 pr[G][N];   // pivot row
  __shared__ double lds_ld[G][N];
  const int lane = threadIdx.x;
  const int gi = min(lane / N, G - 1);
  const int r = lane % N;
  const int grp = blockIdx.x * G + gi;
  const bool on = lane < G * N && grp < g.VH;
  const int k = grp < g.VH ? grp : g.VH - 1;  // tail lanes compute on a valid site and write nothing
  int c[4];
  site_coords(k, 1, g, c);
  const double s = sigma + bc_cf_shift(c[0], g);
  const double2* src = clov + (size_t)1 * (N * N) * g.VHpad;  // odd parity
  auto lo = [](int a, int b) { return N + 2 * (a * (a - 1) / 2 + b); };  // packed real index of (a, b), a > b
  auto get = [&](const double2* base, int vec0, int o) -> double {
    return ((const double*)(base + (size_t)(vec0 + o / 2) * g.VHpad + k))[o & 1];
  };
  double ld = 0.0;
#pragma unroll 1
  for (int blk = 0; blk < 2; blk++) {
    const int vec0 = blk * N * N / 2;
    // row r of sigma + clover
    cd a[N];
#pragma unroll
    for (int j = 0; j < N; j++) {
      if (j == r) a[j] = cd(get(src, vec0, r) + s, 0.0);
      else if (j < r) a[j] = cd(get(src, vec0, lo(r, j)), get(src, vec0, lo(r, j) + 1));
      else a[j] = cd(get(src, vec0, lo(j, r)), -get(src, vec0, lo(j, r) + 1));
    }
#pragma unroll
    for (int kk = 0; kk < N; kk++) {
      if (r == kk) {
        const double d = a[kk].re;  // real for a Hermitian block, as the LDL^+ d
        const double ip = 1.0 / d;
        ld += log(fabs(d));
#pragma unroll
        for (int j = 0; j < N; j++) a[j] = j == kk ? cd(ip, 0.0) : ip * a[j];
#pragma unroll
        for (int j = 0; j < N; j++) pr[gi][j] = make_double2(a[j].re, a[j].im);
      }
      cb_fence();
      if (r != kk) {
        const cd f = a[kk];
#pragma unroll
        for (int j = 0; j < N; j++) {
          const double2 v = pr[gi][j];
          const cd pj(v.x, v.y);
          a[j] = j == kk ? cd(0.0, 0.0) - f * pj : a[j] - f * pj;
        }
      }
      cb_fence();
    }
    // row r of the inverse: diagonal and strictly lower part, packed
    if (on) {
      double2* base = inv;
      double dr = 0.0;  // a[r] without indexing the register array by a lane value
#pragma unroll
      for (int j = 0; j < N; j++) dr = j == r ? a[j].re : dr;
      ((double*)(base + (size_t)(vec0 + r / 2) * g.VHpad + k))[r & 1] = dr;
#pragma unroll
      for (int j = 0; j < N; j++)
        if (j < r) base[(size_t)(vec0 + lo(r, j) / 2) * g.VHpad + k] = make_double2(a[j].re, a[j].im);
    }
  }
  // log det of both blocks: lane r holds the log |d| of pivot r of each
  lds_ld[gi][r] = ld;
  cb_fence();
  if (r == 0) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < N; j++) t += lds_ld[gi][j];
    if (on) logdet[k] = t;
  }
}

// recon-18 F64 -> other link precisions
template <int D, int PREC, int RECON>
__global__ void k_derive_links(const double2* __restrict__ src, void* __restrict__ dst, Geom g) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.VH) return;
  int pm = blockIdx.y;  // parity * 4 + mu
  const double2* s = src + (size_t)pm * D * D * g.VHpad;
  typedef typename LinkStore<PREC, RECON, D>::vec VT;
  constexpr int NV = LinkStore<PREC, RECON, D>::NV;
  VT* d = (VT*)dst + (size_t)pm * NV * g.VHpad;
  if constexpr (PREC == PREC_F64) {
    for (int v = 0; v < NV; v++) d[(size_t)v * g.VHpad + k] = s[(size_t)v * g.VHpad + k];
  } else if constexpr (PREC == PREC_F32 && RECON == RECON_18) {
    for (int v = 0; v < NV; v++) {
      double2 a = s[(size_t)v * g.VHpad + k];
      d[(size_t)v * g.VHpad + k] = make_float2((float)a.x, (float)a.y);
    }
  } else if constexpr (PREC == PREC_F32 && RECON == RECON_12) {
    for (int v = 0; v < 3; v++) {
      double2 a = s[(size_t)(2 * v) * g.VHpad + k], b = s[(size_t)(2 * v + 1) * g.VHpad + k];
      d[(size_t)v * g.VHpad + k] = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
    }
  } else if constexpr (PREC == PREC_F16 && RECON == RECON_18) {
    for (int v = 0; v < NV; v++) {
      double2 a = s[(size_t)v * g.VHpad + k];
      d[(size_t)v * g.VHpad + k] = f2q((float)a.x) | (f2q((float)a.y) << 16);
    }
  } else {
    for (int v = 0; v < 3; v++) {
      double2 a = s[(size_t)(2 * v) * g.VHpad + k], b = s[(size_t)(2 * v + 1) * g.VHpad + k];
      d[(size_t)v * g.VHpad + k] = make_uint2(f2q((float)a.x) | (f2q((float)a.y) << 16),
                                              f2q((float)b.x) | (f2q((float)b.y) << 16));
    }
  }
}

// packed clover F64 double2[vec] -> F32 float4[vec/2] (same real order)
__global__ void k_clover_to_f32(const double2* __restrict__ src, float4* __restrict__ dst, int nvec4, int VHpad, int VH) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= VH) return;
  int v = blockIdx.y;
  if (v >= nvec4) return;
  double2 a = src[(size_t)(2 * v) * VHpad + k], b = src[(size_t)(2 * v + 1) * VHpad + k];
  dst[(size_t)v * VHpad + k] = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
}

// packed clover F64 -> 16-bit fixed point: per site and block the largest
// |real| becomes the block scale, 4 reals per uint2 (same real order)
__global__ void k_clover_to_q16(const double2* __restrict__ src, uint2* __restrict__ dst, float* __restrict__ nrm,
                                int n2, int VHpad, int VH) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= VH) return;
  for (int blk = 0; blk < 2; blk++) {
    const double2* s = src + (size_t)blk * (n2 / 2) * VHpad + k;
    double mx = 0.0;
    for (int v = 0; v < n2 / 2; v++) {
      double2 a = s[(size_t)v * VHpad];
      mx = fmax(mx, fmax(fabs(a.x), fabs(a.y)));
    }
    const float sc = (float)mx;
    const double inv = mx > 0.0 ? 1.0 / (double)sc : 0.0;
    nrm[(size_t)blk * VHpad + k] = sc;
    uint2* d = dst + (size_t)blk * (n2 / 4) * VHpad + k;
    for (int v = 0; v < n2 / 4; v++) {
      double2 a = s[(size_t)(2 * v) * VHpad], b = s[(size_t)(2 * v + 1) * VHpad];
      d[(size_t)v * VHpad] = make_uint2(pack_q2((float)(a.x * inv), (float)(a.y * inv)),
                                        pack_q2((float)(b.x * inv), (float)(b.y * inv)));
    }
  }
}

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------
#define YH_DISPATCH_MODEL(NCV, REPRV, DV, ...) \
  if (c->nc == NCV && c->repr == REPRV) { constexpr int NC_ = NCV, REPR_ = REPRV, D_ = DV; __VA_ARGS__; return YH_OK; }

#define YH_ALL_MODELS(...)                                     \
  YH_DISPATCH_MODEL(3, YH_REPR_FUNDAMENTAL, 3, __VA_ARGS__)    \
  YH_DISPATCH_MODEL(2, YH_REPR_FUNDAMENTAL, 2, __VA_ARGS__)    \
  YH_DISPATCH_MODEL(4, YH_REPR_FUNDAMENTAL, 4, __VA_ARGS__)    \
  YH_DISPATCH_MODEL(2, YH_REPR_ADJOINT, 3, __VA_ARGS__)        \
  YH_DISPATCH_MODEL(3, YH_REPR_SYMMETRIC, 6, __VA_ARGS__)      \
  YH_DISPATCH_MODEL(2, YH_REPR_SYMMETRIC, 3, __VA_ARGS__)      \
  YH_DISPATCH_MODEL(3, YH_REPR_ANTISYMMETRIC, 3, __VA_ARGS__)  \
  YH_DISPATCH_MODEL(4, YH_REPR_ANTISYMMETRIC, 6, __VA_ARGS__)  \
  YH_DISPATCH_MODEL(3, YH_REPR_ADJOINT, 8, __VA_ARGS__)

static int grid1(long n) { return (int)((n + 255) / 256); }
static int grid_bs(long n, int bs) { return (int)((n + bs - 1) / bs); }

int gauge_represent(Ctx* c) {
  if (!c->links[PREC_F64][RECON_18]) {
    size_t bytes = (size_t)8 * c->dr * c->dr * c->g.VHpad * sizeof(double2);
    YH_HIP(hipMalloc(&c->links[PREC_F64][RECON_18], bytes));
  }
  double2* out = (double2*)c->links[PREC_F64][RECON_18];
  YH_ALL_MODELS(hipLaunchKernelGGL((k_represent<NC_, REPR_, D_>), dim3(grid1(4L * c->g.V)), dim3(256), 0, c->stream,
                                   c->fund, out, c->g, c->tab_basis);
                YH_HIP(hipGetLastError()));
  set_error("unsupported (nc, repr) model");
  return YH_ERR_UNSUPPORTED;
}

template <int D>
static int derive_links_d(Ctx* c) {
  const double2* src = (const double2*)c->links[PREC_F64][RECON_18];
  dim3 grid(grid1(c->g.VH), 8);
  auto alloc = [&](int prec, int recon, size_t vbytes, int nv) -> int {
    if (!c->links[prec][recon]) {
      YH_HIP(hipMalloc(&c->links[prec][recon], (size_t)8 * nv * c->g.VHpad * vbytes));
    }
    return YH_OK;
  };
  int rc;
  if ((rc = alloc(PREC_F32, RECON_18, sizeof(float2), D * D))) return rc;
  hipLaunchKernelGGL((k_derive_links<D, PREC_F32, RECON_18>), grid, dim3(256), 0, c->stream, src,
                     c->links[PREC_F32][RECON_18], c->g);
  if ((rc = alloc(PREC_F16, RECON_18, sizeof(unsigned), D * D))) return rc;
  hipLaunchKernelGGL((k_derive_links<D, PREC_F16, RECON_18>), grid, dim3(256), 0, c->stream, src,
                     c->links[PREC_F16][RECON_18], c->g);
  if constexpr (D == 3) {
    if (c->repr == YH_REPR_FUNDAMENTAL && c->nc == 3) {
      if ((rc = alloc(PREC_F64, RECON_12, sizeof(double2), 6))) return rc;
      hipLaunchKernelGGL((k_derive_links<D, PREC_F64, RECON_12>), grid, dim3(256), 0, c->stream, src,
                         c->links[PREC_F64][RECON_12], c->g);
      if ((rc = alloc(PREC_F32, RECON_12, sizeof(float4), 3))) return rc;
      hipLaunchKernelGGL((k_derive_links<D, PREC_F32, RECON_12>), grid, dim3(256), 0, c->stream, src,
                         c->links[PREC_F32][RECON_12], c->g);
      if ((rc = alloc(PREC_F16, RECON_12, sizeof(uint2), 3))) return rc;
      hipLaunchKernelGGL((k_derive_links<D, PREC_F16, RECON_12>), grid, dim3(256), 0, c->stream, src,
                         c->links[PREC_F16][RECON_12], c->g);
    }
  }
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int gauge_derive_links(Ctx* c) {
  switch (c->dr) {
    case 2: return derive_links_d<2>(c);
    case 3: return derive_links_d<3>(c);
    case 4: return derive_links_d<4>(c);
    case 6: return derive_links_d<6>(c);
    case 8: return derive_links_d<8>(c);
  }
  set_error("unsupported representation dimension");
  return YH_ERR_UNSUPPORTED;
}

int clover_build(Ctx* c) {
  // per site and parity: two N x N Hermitian blocks = 2 N^2 reals
  //   FP64: N^2 double2 vectors, FP32: N^2 / 2 float4 vectors
  int N = 2 * c->dr;
  size_t bytes = (size_t)2 * N * N * c->g.VHpad * sizeof(double2);
  if (!c->clov[PREC_F64]) YH_HIP(hipMalloc(&c->clov[PREC_F64], bytes));
  double2* out = (double2*)c->clov[PREC_F64];
  c->cl_sigma = 0xF00F;  // invalidate the inverse (clover.cpp:337)
  int rc = YH_OK;
  auto launch = [&]() -> int {
    // a group of D_R lanes per site
    YH_ALL_MODELS(hipLaunchKernelGGL((k_clover_build_rows<NC_, REPR_, D_>), dim3(grid_bs(c->g.V, 64 / D_)), dim3(64), 0,
                                     c->stream, c->fund, out, c->g, c->csw, c->tab_basis);
                  YH_HIP(hipGetLastError()));
    set_error("unsupported (nc, repr) model");
    return YH_ERR_UNSUPPORTED;
  };
  if ((rc = launch())) return rc;
  // FP32 copy (both parities)
  size_t b32 = (size_t)2 * (N * N / 2) * c->g.VHpad * sizeof(float4);
  if (!c->clov[PREC_F32]) YH_HIP(hipMalloc(&c->clov[PREC_F32], b32));
  for (int p = 0; p < 2; p++) {
    dim3 grid(grid1(c->g.VH), N * N / 2);
    hipLaunchKernelGGL(k_clover_to_f32, grid, dim3(256), 0, c->stream, out + (size_t)p * (N * N) * c->g.VHpad,
                       (float4*)c->clov[PREC_F32] + (size_t)p * (N * N / 2) * c->g.VHpad, N * N / 2, c->g.VHpad,
                       c->g.VH);
  }
  YH_HIP(hipGetLastError());
  // 16-bit copy for the 16-bit operators
  {
    size_t b16 = (size_t)2 * (N * N / 2) * c->g.VHpad * sizeof(uint2);
    if (!c->clov[PREC_F16]) YH_HIP(hipMalloc(&c->clov[PREC_F16], b16));
    if (!c->clov_nrm) YH_HIP(hipMalloc(&c->clov_nrm, sizeof(float) * 4 * c->g.VHpad));
    for (int p = 0; p < 2; p++)
      hipLaunchKernelGGL(k_clover_to_q16, dim3(grid1(c->g.VH)), dim3(256), 0, c->stream,
                         out + (size_t)p * (N * N) * c->g.VHpad,
                         (uint2*)c->clov[PREC_F16] + (size_t)p * (N * N / 2) * c->g.VHpad,
                         c->clov_nrm + (size_t)p * 2 * c->g.VHpad, N * N, c->g.VHpad, c->g.VH);
    YH_HIP(hipGetLastError());
  }
  return YH_OK;
}

// the inverse buffers of every precision, allocated once per context (at
// yh_set_gauge), so that solves never allocate
int clover_reserve_inverse(Ctx* c) {
  int N = 2 * c->dr;
  size_t bytes = (size_t)N * N * c->g.VHpad * sizeof(double2);  // odd sites only
  if (!c->clinv[PREC_F64]) YH_HIP(hipMalloc(&c->clinv[PREC_F64], bytes));
  if (!c->logdet_site) YH_HIP(hipMalloc(&c->logdet_site, sizeof(double) * c->g.VHpad));
  size_t b32 = (size_t)(N * N / 2) * c->g.VHpad * sizeof(float4);
  if (!c->clinv[PREC_F32]) YH_HIP(hipMalloc(&c->clinv[PREC_F32], b32));
  size_t b16 = (size_t)(N * N / 2) * c->g.VHpad * sizeof(uint2);
  if (!c->clinv[PREC_F16]) YH_HIP(hipMalloc(&c->clinv[PREC_F16], b16));
  if (!c->clinv_nrm) YH_HIP(hipMalloc(&c->clinv_nrm, sizeof(float) * 2 * c->g.VHpad));
  return YH_OK;
}

int clover_invert(Ctx* c, double sigma) {
  if (sigma == c->cl_sigma) return YH_OK;  // clover.cpp:314-321
  int N = 2 * c->dr;
  int rc = clover_reserve_inverse(c);
  if (rc) return rc;
  const double2* src = (const double2*)c->clov[PREC_F64];
  double2* inv = (double2*)c->clinv[PREC_F64];
  switch (c->dr) {
    case 2: hipLaunchKernelGGL(k_clover_invert_rows<2>, dim3(grid_bs(c->g.VH, 64 / (2 * 2))), dim3(64), 0, c->stream, src, inv, c->logdet_site, c->g, sigma); break;
    case 3: hipLaunchKernelGGL(k_clover_invert_rows<3>, dim3(grid_bs(c->g.VH, 64 / (2 * 3))), dim3(64), 0, c->stream, src, inv, c->logdet_site, c->g, sigma); break;
    case 4: hipLaunchKernelGGL(k_clover_invert_rows<4>, dim3(grid_bs(c->g.VH, 64 / (2 * 4))), dim3(64), 0, c->stream, src, inv, c->logdet_site, c->g, sigma); break;
    case 6: hipLaunchKernelGGL(k_clover_invert_rows<6>, dim3(grid_bs(c->g.VH, 64 / (2 * 6))), dim3(64), 0, c->stream, src, inv, c->logdet_site, c->g, sigma); break;
    case 8: hipLaunchKernelGGL(k_clover_invert_rows<8>, dim3(grid_bs(c->g.VH, 64 / (2 * 8))), dim3(64), 0, c->stream, src, inv, c->logdet_site, c->g, sigma); break;
    default: set_error("unsupported representation dimension"); return YH_ERR_UNSUPPORTED;
  }
  YH_HIP(hipGetLastError());
  dim3 grid(grid1(c->g.VH), N * N / 2);
  hipLaunchKernelGGL(k_clover_to_f32, grid, dim3(256), 0, c->stream, inv, (float4*)c->clinv[PREC_F32], N * N / 2,
                     c->g.VHpad, c->g.VH);
  YH_HIP(hipGetLastError());
  {
    hipLaunchKernelGGL(k_clover_to_q16, dim3(grid1(c->g.VH)), dim3(256), 0, c->stream, inv,
                       (uint2*)c->clinv[PREC_F16], c->clinv_nrm, N * N, c->g.VHpad, c->g.VH);
    YH_HIP(hipGetLastError());
  }
  c->cl_sigma = sigma;
  return YH_OK;
}

// reference AoS layouts for read-back
__global__ void k_repr_to_aos(const double2* __restrict__ src, double* __restrict__ dst, Geom g, int D) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * g.V) return;
  int id = t >> 2, mu = t & 3;
  int p = id >= g.VH, k = id - p * g.VH;
  int c[4];
  site_coords(k, p, g, c);
  double sgn = (g.bc == BC_ANTIPERIODIC && mu == 0 && c[0] == 0) ? -1.0 : 1.0;
  double* o = dst + (size_t)t * 2 * D * D;
  for (int e = 0; e < D * D; e++) {
    double2 v = src[((size_t)(p * 4 + mu) * D * D + e) * g.VHpad + k];
    o[e] = sgn * v.x;
    o[D * D + e] = sgn * v.y;
  }
}

__global__ void k_clover_to_aos(const double2* __restrict__ src, double* __restrict__ dst, Geom g, int D) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= g.V) return;
  int p = id >= g.VH, k = id - p * g.VH;
  int N = 2 * D;
  const double2* base = src + (size_t)p * (N * N) * g.VHpad;
  for (int blk = 0; blk < 2; blk++) {
    // unpack into A (or C) = M[i][j], B (or D) = M[i][D + j]
    double v[16 * 16];
    for (int q = 0; q < N * N / 2; q++) {
      double2 w = base[(size_t)(blk * N * N / 2 + q) * g.VHpad + k];
      v[2 * q] = w.x;
      v[2 * q + 1] = w.y;
    }
    auto get = [&](int i, int j, double& re, double& im) {
      if (i == j) { re = v[i]; im = 0; return; }
      bool lower = i > j;
      int a = lower ? i : j, b = lower ? j : i;
      int o = N + 2 * (a * (a - 1) / 2 + b);
      re = v[o];
      im = lower ? v[o + 1] : -v[o + 1];
    };
    double* A = dst + (size_t)(4 * id + 2 * blk) * 2 * D * D;
    double* B = dst + (size_t)(4 * id + 2 * blk + 1) * 2 * D * D;
    for (int i = 0; i < D; i++)
      for (int j = 0; j < D; j++) {
        get(i, j, A[i * D + j], A[D * D + i * D + j]);
        get(i, D + j, B[i * D + j], B[D * D + i * D + j]);
      }
  }
}

int gauge_download_repr(Ctx* c, double* host) {
  size_t n = (size_t)4 * c->g.V * 2 * c->dr * c->dr;
  double* d;
  YH_HIP(hipMalloc(&d, n * sizeof(double)));
  hipLaunchKernelGGL(k_repr_to_aos, dim3(grid1(4L * c->g.V)), dim3(256), 0, c->stream,
                     (const double2*)c->links[PREC_F64][RECON_18], d, c->g, c->dr);
  hipError_t e = hipMemcpyAsync(host, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(d);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return YH_ERR_HIP; }
  return YH_OK;
}

int clover_download(Ctx* c, double* host) {
  size_t n = (size_t)4 * c->g.V * 2 * c->dr * c->dr;
  double* d;
  YH_HIP(hipMalloc(&d, n * sizeof(double)));
  hipLaunchKernelGGL(k_clover_to_aos, dim3(grid_bs(c->g.V, 64)), dim3(64), 0, c->stream, (const double2*)c->clov[PREC_F64],
                     d, c->g, c->dr);
  hipError_t e = hipMemcpyAsync(host, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(d);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return YH_ERR_HIP; }
  return YH_OK;
}

}  // namespace yh
