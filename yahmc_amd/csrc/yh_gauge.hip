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

// sum of the four plaquette leaves at x in the (mu, nu) plane, represented
// (library/wilsonloop.cpp:5-31)
template <int NC, int REPR, int D>
__device__ void leaf_sum(cd (&Q)[D][D], const double* fund, int id, const int c[4], int mu, int nu, const Geom& g,
                         const double2* __restrict__ basis) {
  cd A[NC][NC], B[NC][NC], T1[NC][NC], T2[NC][NC], L[NC][NC], R[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) Q[i][j] = cd(0, 0);
  int xm = shift(c, mu, 1, g), xn = shift(c, nu, 1, g);
  int bm = shift(c, mu, -1, g), bn = shift(c, nu, -1, g);
  int cb[4];
  coords_of(bm, g, cb);
  int bm_pn = shift(cb, nu, 1, g);   // x - mu + nu
  int bm_mn = shift(cb, nu, -1, g);  // x - mu - nu
  int cn[4];
  coords_of(bn, g, cn);
  int bn_pm = shift(cn, mu, 1, g);   // x - nu + mu
  // leaf 0: U(x,mu) U(x+mu,nu) [U(x,nu) U(x+nu,mu)]^+
  load_fund<NC>(A, fund, id, mu);
  load_fund<NC>(B, fund, xm, nu);
  mm<NC>(T1, A, B);
  load_fund<NC>(A, fund, id, nu);
  load_fund<NC>(B, fund, xn, mu);
  mm<NC>(T2, A, B);
  mmd<NC>(L, T1, T2);
  represent<NC, REPR, D>(R, L, basis);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) Q[i][j] = Q[i][j] + R[i][j];
  // leaf 1: U(x,nu) [U(x-mu,nu) U(x-mu+nu,mu)]^+ U(x-mu,mu)
  load_fund<NC>(A, fund, bm, nu);
  load_fund<NC>(B, fund, bm_pn, mu);
  mm<NC>(T1, A, B);
  load_fund<NC>(A, fund, id, nu);
  mmd<NC>(T2, A, T1);
  load_fund<NC>(A, fund, bm, mu);
  mm<NC>(L, T2, A);
  represent<NC, REPR, D>(R, L, basis);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) Q[i][j] = Q[i][j] + R[i][j];
  // leaf 2: [U(x-mu-nu,nu) U(x-mu,mu)]^+ U(x-mu-nu,mu) U(x-nu,nu)
  load_fund<NC>(A, fund, bm_mn, nu);
  load_fund<NC>(B, fund, bm, mu);
  mm<NC>(T1, A, B);
  load_fund<NC>(A, fund, bm_mn, mu);
  mdm<NC>(T2, T1, A);
  load_fund<NC>(A, fund, bn, nu);
  mm<NC>(L, T2, A);
  represent<NC, REPR, D>(R, L, basis);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) Q[i][j] = Q[i][j] + R[i][j];
  // leaf 3: U(x-nu,nu)^+ U(x-nu,mu) U(x-nu+mu,nu) U(x,mu)^+
  load_fund<NC>(A, fund, bn, nu);
  load_fund<NC>(B, fund, bn, mu);
  mdm<NC>(T1, A, B);
  load_fund<NC>(A, fund, bn_pm, nu);
  mm<NC>(T2, T1, A);
  load_fund<NC>(A, fund, id, mu);
  mmd<NC>(L, T2, A);
  represent<NC, REPR, D>(R, L, basis);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) Q[i][j] = Q[i][j] + R[i][j];
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

// clover term per site -> packed blocks [parity][2 blocks * n^2/2 vec][site]
// (library/clover.cpp:120-168; the dense form is -(c_sw/16) sum_{mu<nu}
//  (1/2)[g_mu, g_nu] (x) (Q_mu_nu - Q_mu_nu^+))
template <int NC, int REPR, int D>
__global__ void k_clover_build(const double* __restrict__ fund, double2* __restrict__ out, Geom g, double csw,
                               const double2* __restrict__ basis) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= g.V) return;
  constexpr int N = 2 * D;
  int c[4];
  coords_of(id, g, c);
  int p = id >= g.VH;
  int k = id - p * g.VH;
  double2* base = out + (size_t)p * (N * N) * g.VHpad;
  cd at[D][D], bt[D][D], ct[D][D], dt[D][D], Q[D][D];
  bool zero = false;
  if (g.bc == BC_SF) zero = (c[0] == 0 || c[0] == g.L[0] - 1 || c[0] == g.L[0] - 2);
  if (g.bc == BC_OPEN) zero = (c[0] == 0 || c[0] == g.L[0] - 1);
  if (csw == 0.0) zero = true;
  if (zero) {
    for (int v = 0; v < N * N; v++) base[(size_t)v * g.VHpad + k] = make_double2(0, 0);
    return;
  }
  // plane order of clover.cpp:135-140: 01, 02, 03, 12, 13, 23
  // at: from Q03; bt: Q12; ct: Q01, Q02; dt: Q13, Q23 (clover.cpp:149-156)
  leaf_sum<NC, REPR, D>(Q, fund, id, c, 0, 3, g, basis);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) at[i][j] = cd(Q[i][j].im + Q[j][i].im, Q[j][i].re - Q[i][j].re);
  leaf_sum<NC, REPR, D>(Q, fund, id, c, 1, 2, g, basis);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) bt[i][j] = cd(Q[i][j].im + Q[j][i].im, Q[i][j].re - Q[j][i].re);
  leaf_sum<NC, REPR, D>(Q, fund, id, c, 0, 1, g, basis);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) ct[i][j] = cd(Q[i][j].im + Q[j][i].im, Q[j][i].re - Q[i][j].re);
  leaf_sum<NC, REPR, D>(Q, fund, id, c, 0, 2, g, basis);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) ct[i][j] = ct[i][j] + cd(-Q[i][j].re + Q[j][i].re, -Q[i][j].im - Q[j][i].im);
  leaf_sum<NC, REPR, D>(Q, fund, id, c, 1, 3, g, basis);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) dt[i][j] = cd(Q[i][j].re - Q[j][i].re, Q[i][j].im + Q[j][i].im);
  leaf_sum<NC, REPR, D>(Q, fund, id, c, 2, 3, g, basis);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) dt[i][j] = dt[i][j] + cd(Q[i][j].im + Q[j][i].im, -Q[i][j].re + Q[j][i].re);
  double cc = -csw / 16.0;
  cd M[N][N];
  // block 0 = [[A, B], [B^+, -A]], block 1 = [[C, D], [D^+, -C]]  (dirac.cpp:63-89)
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      cd A(cc * (at[i][j].re - bt[i][j].re), cc * (at[i][j].im + bt[i][j].im));
      cd B(cc * (ct[i][j].re - dt[i][j].re), cc * (ct[i][j].im - dt[i][j].im));
      M[i][j] = A;
      M[D + i][D + j] = cd(-A.re, -A.im);
      M[i][D + j] = B;
      M[D + j][i] = conj(B);
    }
  store_packed<N>(base, g.VHpad, k, 0, M);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      cd C(-cc * (at[i][j].re + bt[i][j].re), cc * (bt[i][j].im - at[i][j].im));
      cd Dm(-cc * (ct[i][j].re + dt[i][j].re), -cc * (ct[i][j].im + dt[i][j].im));
      M[i][j] = C;
      M[D + i][D + j] = cd(-C.re, -C.im);
      M[i][D + j] = Dm;
      M[D + j][i] = conj(Dm);
    }
  store_packed<N>(base, g.VHpad, k, N * N / 2, M);
}

// odd-site inverse of sigma + c_f shift + clover, per chiral block, through
// the reference's unpivoted LDL^+ (solver.cpp:97-117); also log det
template <int D>
__global__ void k_clover_invert(const double2* __restrict__ clov, double2* __restrict__ inv, double* __restrict__ logdet,
                                Geom g, double sigma) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.VH) return;
  constexpr int N = 2 * D;
  int c[4];
  site_coords(k, 1, g, c);
  double s = sigma + bc_cf_shift(c[0], g);
  const double2* src = clov + (size_t)1 * (N * N) * g.VHpad;  // odd parity
  double ld = 0.0;
  for (int blk = 0; blk < 2; blk++) {
    cd A[N][N];
    load_packed<N>(A, src, g.VHpad, k, blk * N * N / 2);
    for (int i = 0; i < N; i++) A[i][i].re += s;
    // LDL^+: A = L D L^+, L unit lower, D real
    for (int i = 0; i < N; i++) {
      double d = A[i][i].re;
      for (int q = 0; q < i; q++) d -= (A[i][q].re * A[i][q].re + A[i][q].im * A[i][q].im) * A[q][q].re;
      A[i][i] = cd(d, 0);
      for (int j = i + 1; j < N; j++) {
        cd v = A[j][i];
        for (int q = 0; q < i; q++) {
          cd t = A[j][q] * conj(A[i][q]);
          v = v - A[q][q].re * t;
        }
        A[j][i] = (1.0 / d) * v;
      }
    }
    double prod = 1.0;
    for (int i = 0; i < N; i++) prod *= A[i][i].re;
    ld += log(fabs(prod));
    // explicit inverse column by column
    cd X[N][N];
    for (int col = 0; col < N; col++) {
      cd x[N];
      for (int i = 0; i < N; i++) x[i] = cd(i == col ? 1.0 : 0.0, 0);
      for (int i = 1; i < N; i++)
        for (int q = 0; q < i; q++) x[i] = x[i] - A[i][q] * x[q];
      for (int i = N - 1; i >= 0; i--) {
        x[i] = (1.0 / A[i][i].re) * x[i];
        for (int q = i + 1; q < N; q++) x[i] = x[i] - conj(A[q][i]) * x[q];
      }
      for (int i = 0; i < N; i++) X[i][col] = x[i];
    }
    store_packed<N>(inv, g.VHpad, k, blk * N * N / 2, X);
  }
  logdet[k] = ld;
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
    YH_ALL_MODELS(hipLaunchKernelGGL((k_clover_build<NC_, REPR_, D_>), dim3(grid_bs(c->g.V, 128)), dim3(128), 0, c->stream,
                                     c->fund, out, c->g, c->csw, c->tab_basis);
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
    case 2: hipLaunchKernelGGL(k_clover_invert<2>, dim3(grid_bs(c->g.VH, 64)), dim3(64), 0, c->stream, src, inv, c->logdet_site, c->g, sigma); break;
    case 3: hipLaunchKernelGGL(k_clover_invert<3>, dim3(grid_bs(c->g.VH, 64)), dim3(64), 0, c->stream, src, inv, c->logdet_site, c->g, sigma); break;
    case 4: hipLaunchKernelGGL(k_clover_invert<4>, dim3(grid_bs(c->g.VH, 64)), dim3(64), 0, c->stream, src, inv, c->logdet_site, c->g, sigma); break;
    case 6: hipLaunchKernelGGL(k_clover_invert<6>, dim3(grid_bs(c->g.VH, 64)), dim3(64), 0, c->stream, src, inv, c->logdet_site, c->g, sigma); break;
    case 8: hipLaunchKernelGGL(k_clover_invert<8>, dim3(grid_bs(c->g.VH, 64)), dim3(64), 0, c->stream, src, inv, c->logdet_site, c->g, sigma); break;
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
