// Gauge molecular dynamics on the device (FP64): Wilson gauge force, link
// update, reunitarisation, plaquette and kinetic energy.
//
// wilson_gauge_force : library/forces.cpp:262-283 + staples.cpp:82-107
// update_links       : library/update.cpp:228-247 + suNa.cpp:113-134 (exp)
// reunitarize        : library/suNg.cpp:171-359
// plaquette          : library/wilsonloop.cpp:58-133 (avr_wilson(1,1))
// improved_gauge_force: library/forces.cpp:285-310 + staples.cpp:7-80
// avr_wilson(h, w)   : library/wilsonloop.cpp:58-133 (rectangle weights of
//                      bcs.cpp:30-31, 72-120)
// hamiltonian (P^2)  : library/update.cpp:201-216
// Fundamental links stay in the reference layout (4*id+mu, re plane then im
// plane); momenta are SoA mom[(mu * NG + a) * V + id].  Periodic, open and
// SF gauge boundaries (bcs.cpp: boundary plaquette weights, fixed links).
#include <algorithm>
#include <vector>
#include "yh_internal.hpp"
#include "yh_reduce.hpp"

namespace yh {

typedef Cx<double> cd;

// diagonal phases of the SF boundary links (bcs.cpp:208-232)
struct SfPhases {
  double c1[4], c2[4];
};

template <int NC>
struct M {
  cd m[NC][NC];
};

template <int NC>
__device__ __forceinline__ M<NC> ld(const double* fund, int id, int mu) {
  M<NC> u;
  const double* p = fund + (size_t)(4 * id + mu) * 2 * NC * NC;
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) u.m[i][j] = cd(p[i * NC + j], p[NC * NC + i * NC + j]);
  return u;
}

template <int NC>
__device__ __forceinline__ void st(double* fund, int id, int mu, const M<NC>& u) {
  double* p = fund + (size_t)(4 * id + mu) * 2 * NC * NC;
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) {
      p[i * NC + j] = u.m[i][j].re;
      p[NC * NC + i * NC + j] = u.m[i][j].im;
    }
}

template <int NC>
__device__ __forceinline__ M<NC> mul(const M<NC>& a, const M<NC>& b) {
  M<NC> r;
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) {
      cd s(0, 0);
#pragma unroll
      for (int k = 0; k < NC; k++) cmac(s, a.m[i][k], b.m[k][j]);
      r.m[i][j] = s;
    }
  return r;
}

// a * b^+
template <int NC>
__device__ __forceinline__ M<NC> muld(const M<NC>& a, const M<NC>& b) {
  M<NC> r;
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) {
      cd s(0, 0);
#pragma unroll
      for (int k = 0; k < NC; k++) cmac(s, a.m[i][k], conj(b.m[j][k]));
      r.m[i][j] = s;
    }
  return r;
}

// a^+ * b
template <int NC>
__device__ __forceinline__ M<NC> dmul(const M<NC>& a, const M<NC>& b) {
  M<NC> r;
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) {
      cd s(0, 0);
#pragma unroll
      for (int k = 0; k < NC; k++) cmac_conj(s, a.m[k][i], b.m[k][j]);
      r.m[i][j] = s;
    }
  return r;
}

__device__ __forceinline__ int mshift(const int c[4], int mu, int d, const Geom& g) {
  int n[4] = {c[0], c[1], c[2], c[3]};
  n[mu] = (n[mu] + d + g.L[mu]) % g.L[mu];
  return full_index(n[0], n[1], n[2], n[3], g);
}

// anti-Hermitian matrix of an algebra vector (library/suNg.cpp:15-58)
template <int NC>
__device__ __forceinline__ M<NC> alg(const double* mom, int V, int mu, int id) {
  constexpr int NG = NC * NC - 1;
  M<NC> r;
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) r.m[i][j] = cd(0, 0);
  int n = 0;
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) {
      if (i > j) {
        double v = mom[(size_t)(mu * NG + n) * V + id];
        r.m[i][j].im += v;
        r.m[j][i].im += v;
        n++;
      } else if (i < j) {
        double v = mom[(size_t)(mu * NG + n) * V + id];
        r.m[i][j].re += v;
        r.m[j][i].re -= v;
        n++;
      } else if (i > 0) {
        double v = mom[(size_t)(mu * NG + n) * V + id];
        double w = sqrt(2.0 / (i * (i + 1)));
        for (int k = 0; k < i; k++) r.m[k][k].im += w * v;
        r.m[i][i].im -= i * w * v;
        n++;
      }
    }
  return r;
}

// determinant by Gaussian elimination with partial pivoting (the row swaps
// are selects, so every index stays static); the reference uses an
// unpivoted LU (suNg.cpp:296-330) for the same value
template <int NC>
__device__ __forceinline__ cd det_gauss(M<NC> m) {
  cd det(1, 0);
#pragma unroll
  for (int c = 0; c < NC; c++) {
#pragma unroll
    for (int i = c + 1; i < NC; i++) {
      const double ai = m.m[i][c].re * m.m[i][c].re + m.m[i][c].im * m.m[i][c].im;
      const double ac = m.m[c][c].re * m.m[c][c].re + m.m[c][c].im * m.m[c][c].im;
      const bool sw = ai > ac;
#pragma unroll
      for (int j = 0; j < NC; j++) {
        const cd a = m.m[c][j], b = m.m[i][j];
        m.m[c][j] = sw ? b : a;
        m.m[i][j] = sw ? a : b;
      }
      det = sw ? cd(-det.re, -det.im) : det;
    }
    const cd p = m.m[c][c];
    det = det * p;
    const double n = 1.0 / (p.re * p.re + p.im * p.im);
    const cd inv(p.re * n, -p.im * n);
#pragma unroll
    for (int i = c + 1; i < NC; i++) {
      const cd f = m.m[i][c] * inv;
#pragma unroll
      for (int j = c; j < NC; j++) m.m[i][j] = m.m[i][j] - f * m.m[c][j];
    }
  }
  return det;
}

// suNg::reunitarize (library/suNg.cpp:171-359); a zero row returns the
// input unchanged (the reference returns early there, with the rows before
// it already processed in the general branch)
//   NC = 2: normalise row 0, row 1 = (-conj r01, conj r00)
//   NC = 3: normalise row 0, orthogonalise and normalise row 1,
//           row 2 = conj(r0 x r1)
//   other:  modified Gram-Schmidt over all rows, then the last row times
//           exp(-i arg det) so that det = 1
template <int NC>
__device__ __forceinline__ void reunit(M<NC>& u) {
  if constexpr (NC == 2 || NC == 3) {
    double n0 = 0;
#pragma unroll
    for (int j = 0; j < NC; j++) n0 += u.m[0][j].re * u.m[0][j].re + u.m[0][j].im * u.m[0][j].im;
    if (n0 == 0.0) return;
    n0 = 1.0 / sqrt(n0);
#pragma unroll
    for (int j = 0; j < NC; j++) u.m[0][j] = n0 * u.m[0][j];
    if constexpr (NC == 2) {
      u.m[1][0] = cd(-u.m[0][1].re, u.m[0][1].im);
      u.m[1][1] = conj(u.m[0][0]);
    } else {
      cd p(0, 0);
#pragma unroll
      for (int j = 0; j < NC; j++) cmac_conj(p, u.m[0][j], u.m[1][j]);
#pragma unroll
      for (int j = 0; j < NC; j++) u.m[1][j] = u.m[1][j] - p * u.m[0][j];
      double n1 = 0;
#pragma unroll
      for (int j = 0; j < NC; j++) n1 += u.m[1][j].re * u.m[1][j].re + u.m[1][j].im * u.m[1][j].im;
      n1 = 1.0 / sqrt(n1);
#pragma unroll
      for (int j = 0; j < NC; j++) u.m[1][j] = n1 * u.m[1][j];
      u.m[2][0] = conj(u.m[0][1] * u.m[1][2] - u.m[0][2] * u.m[1][1]);
      u.m[2][1] = conj(u.m[0][2] * u.m[1][0] - u.m[0][0] * u.m[1][2]);
      u.m[2][2] = conj(u.m[0][0] * u.m[1][1] - u.m[0][1] * u.m[1][0]);
    }
  } else {
    const M<NC> orig = u;
    bool zero = false;
#pragma unroll
    for (int i = 0; i < NC; i++) {
      double n = 0;
#pragma unroll
      for (int j = 0; j < NC; j++) n += u.m[i][j].re * u.m[i][j].re + u.m[i][j].im * u.m[i][j].im;
      zero = zero || n == 0.0;
      n = n == 0.0 ? 1.0 : 1.0 / sqrt(n);
#pragma unroll
      for (int j = 0; j < NC; j++) u.m[i][j] = n * u.m[i][j];
#pragma unroll
      for (int k = i + 1; k < NC; k++) {
        cd p(0, 0);
#pragma unroll
        for (int j = 0; j < NC; j++) cmac_conj(p, u.m[i][j], u.m[k][j]);
#pragma unroll
        for (int j = 0; j < NC; j++) u.m[k][j] = u.m[k][j] - p * u.m[i][j];
      }
    }
    if (zero) {
      u = orig;
      return;
    }
    const cd d = det_gauss<NC>(u);
    const double inv = 1.0 / sqrt(d.re * d.re + d.im * d.im);
    const cd ph(d.re * inv, -d.im * inv);
#pragma unroll
    for (int j = 0; j < NC; j++) u.m[NC - 1][j] = u.m[NC - 1][j] * ph;
  }
}

// fundamental generators iT_a, NC x NC: the context's table
// tab_itf[a * 16 + i * NC + k], passed to the kernels

// Open and SF boundaries of the gauge field (library/bcs.cpp:28-130, 166-262):
// spatial plaquettes based on the boundary slices (open: t = 0 and T-1; SF:
// t = 0 and T-2) count half in the action and its force, and the links the
// boundary fixes (open: U_0 at t = T-1; SF: the spatial links at t = 0, every
// link at t = T-2 and T-1) keep their values: their momenta are zeroed.
__device__ __forceinline__ double plaq_w(int t, int mu, int nu, const Geom& g) {
  if (mu == 0 || nu == 0) return 1.0;
  if (g.bc == BC_OPEN && (t == 0 || t == g.L[0] - 1)) return 0.5;
  if (g.bc == BC_SF && (t == 0 || t == g.L[0] - 2)) return 0.5;
  return 1.0;
}

__device__ __forceinline__ bool link_fixed(int t, int mu, const Geom& g) {
  if (g.bc == BC_OPEN) return t == g.L[0] - 1 && mu == 0;
  if (g.bc == BC_SF) return (t == 0 && mu > 0) || t == g.L[0] - 2 || t == g.L[0] - 1;
  return false;
}

// rect_weight(id, a, b) of the 1 x 2 rectangle based on time slice t with
// its short side along a and long side along b: open boundaries use the
// plaquette weights (bcs.cpp:31); SF weights 0 for the spatial a < b
// rectangles on t = 0 and T-2 and 1.5 for a = 0 on t = 0 and T-3
// (bcs.cpp:89-116)
__device__ __forceinline__ double rect_w(int t, int a, int b, const Geom& g) {
  if (g.bc == BC_OPEN) return plaq_w(t, a, b, g);
  if (g.bc != BC_SF) return 1.0;
  if ((t == 0 || t == g.L[0] - 2) && a >= 1 && a < b) return 0.0;
  if ((t == 0 || t == g.L[0] - 3) && a == 0 && b >= 1) return 1.5;
  return 1.0;
}

// one step of a path on coordinates c: forward (d, +1) multiplies P by
// U(c, d) and moves to c + d; backward (d, -1) moves to c - d and multiplies
// by U(c - d, d)^+
template <int NC>
__device__ __forceinline__ void path_step(M<NC>& P, const double* fund, int (&c)[4], int d, bool fw, const Geom& g) {
  if (fw) {
    P = mul<NC>(P, ld<NC>(fund, full_index(c[0], c[1], c[2], c[3], g), d));
    c[d] = c[d] + 1 == g.L[d] ? 0 : c[d] + 1;
  } else {
    c[d] = c[d] == 0 ? g.L[d] - 1 : c[d] - 1;
    P = muld<NC>(P, ld<NC>(fund, full_index(c[0], c[1], c[2], c[3], g), d));
  }
}

template <int NC>
__device__ __forceinline__ M<NC> unit_matrix() {
  M<NC> r;
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) r.m[i][j] = cd(i == j ? 1.0 : 0.0, 0.0);
  return r;
}

// The six 1 x 2 rectangles through the link (x, mu) for one nu, as paths
// from x + mu back to x (bits 2j, 2j+1 of the code: step j is 0 +mu, 1 -mu,
// 2 +nu, 3 -nu), with the offset of the rectangle's base corner from x
// (bits 10-11: -mu steps, 12-13: -nu steps) and whether its long side is
// along nu (bit 14), which orders the arguments of its weight
#define YH_R(s0, s1, s2, s3, s4, bm, bn, lnu) \
  ((s0) | (s1) << 2 | (s2) << 4 | (s3) << 6 | (s4) << 8 | (bm) << 10 | (bn) << 12 | (lnu) << 14)
__constant__ int c_rect_paths[6] = {
    YH_R(2, 2, 1, 3, 3, 0, 0, 1),  // long side nu, +nu side
    YH_R(3, 3, 1, 2, 2, 0, 2, 1),  // long side nu, -nu side
    YH_R(0, 2, 1, 1, 3, 0, 0, 0),  // long side mu, the link first, +nu side
    YH_R(2, 1, 1, 3, 0, 1, 0, 0),  // long side mu, the link second, +nu side
    YH_R(0, 3, 1, 1, 2, 0, 1, 0),  // long side mu, the link first, -nu side
    YH_R(3, 1, 1, 2, 0, 1, 1, 0),  // long side mu, the link second, -nu side
};
#undef YH_R

// apply_bcs_on_momentum_field (bcs.cpp:166-199)
__global__ void __launch_bounds__(256) k_mom_bcs(double* __restrict__ mom, int NG, Geom g) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 4 * g.V) return;
  int id = t >> 2, mu = t & 3;
  int c[4];
  full_coords(id, g, c);
  if (!link_fixed(c[0], mu, g)) return;
  for (int a = 0; a < NG; a++) mom[(size_t)(mu * NG + a) * g.V + id] = 0.0;
}

// apply_bcs_on_gauge_field (bcs.cpp:201-262): the fixed links take their
// boundary values: zero, or the SF matrices diag(exp(i c1_k)), diag(exp(i c2_k))
template <int NC>
__global__ void __launch_bounds__(256) k_gauge_bcs(double* __restrict__ fund, Geom g, SfPhases ph) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 4 * g.V) return;
  int id = t >> 2, mu = t & 3;
  int c[4];
  full_coords(id, g, c);
  if (!link_fixed(c[0], mu, g)) return;
  M<NC> u;
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) u.m[i][j] = cd(0, 0);
  if (g.bc == BC_SF && mu > 0 && (c[0] == 0 || c[0] == g.L[0] - 2)) {
    const double* v = c[0] == 0 ? ph.c1 : ph.c2;
#pragma unroll
    for (int i = 0; i < NC; i++) u.m[i][i] = cd(cos(v[i]), sin(v[i]));
  }
  st<NC>(fund, id, mu, u);
}

// Wilson gauge force: mom -= algebra_project(U staple, dt beta / NC), the
// staples weighted by their plaquettes' weights (staples.cpp:82-107).  With
// RECT the improved force: staple c0 S_plaq + c1 S_rect, S_rect the six
// weighted rectangles through the link for each nu (staples.cpp:7-80)
template <int NC, bool RECT>
__global__ void __launch_bounds__(128) k_gauge_force(const double* __restrict__ fund, double* __restrict__ mom,
                                                     double coeff, double c0, double c1, Geom g,
                                                     const double2* __restrict__ itf) {
  int t = blockIdx.x * 128 + threadIdx.x;
  if (t >= 4 * g.V) return;
  int id = t >> 2, mu = t & 3;
  constexpr int NG = NC * NC - 1;
  int c[4];
  full_coords(id, g, c);
  M<NC> S;
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) S.m[i][j] = cd(0, 0);
  int o1m = mshift(c, mu, 1, g);
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    int o2 = mshift(c, nu, 1, g);
    M<NC> a = mul<NC>(ld<NC>(fund, id, nu), ld<NC>(fund, o2, mu));
    M<NC> s1 = muld<NC>(ld<NC>(fund, o1m, nu), a);
    int b = mshift(c, nu, -1, g);
    int cb[4];
    full_coords(b, g, cb);
    int b2 = mshift(cb, mu, 1, g);
    M<NC> q = mul<NC>(ld<NC>(fund, b, mu), ld<NC>(fund, b2, nu));
    M<NC> s2 = dmul<NC>(q, ld<NC>(fund, b, nu));
    double w1 = plaq_w(c[0], mu, nu, g), w2 = plaq_w(cb[0], nu, mu, g);
    if constexpr (RECT) {
      w1 *= c0;
      w2 *= c0;
    }
#pragma unroll
    for (int i = 0; i < NC; i++)
#pragma unroll
      for (int j = 0; j < NC; j++) S.m[i][j] = S.m[i][j] + w1 * s1.m[i][j] + w2 * s2.m[i][j];
    if constexpr (RECT) {
      for (int r = 0; r < 6; r++) {
        const int code = c_rect_paths[r];
        int p[4] = {c[0], c[1], c[2], c[3]};
        p[mu] = p[mu] + 1 == g.L[mu] ? 0 : p[mu] + 1;
        M<NC> P = unit_matrix<NC>();
        for (int j = 0; j < 5; j++) {
          const int sc = (code >> (2 * j)) & 3;
          path_step<NC>(P, fund, p, sc >> 1 ? nu : mu, (sc & 1) == 0, g);
        }
        // time slice of the base corner: x - bm mu - bn nu
        int tb = c[0] - (mu == 0 ? (code >> 10) & 3 : 0) - (nu == 0 ? (code >> 12) & 3 : 0);
        tb = tb < 0 ? tb + g.L[0] : tb;
        const double w = c1 * ((code >> 14) & 1 ? rect_w(tb, mu, nu, g) : rect_w(tb, nu, mu, g));
#pragma unroll
        for (int i = 0; i < NC; i++)
#pragma unroll
          for (int j = 0; j < NC; j++) S.m[i][j] = S.m[i][j] + w * P.m[i][j];
      }
    }
  }
  M<NC> F = mul<NC>(ld<NC>(fund, id, mu), S);
#pragma unroll
  for (int a = 0; a < NG; a++) {
    double re = 0;
#pragma unroll
    for (int i = 0; i < NC; i++)
#pragma unroll
      for (int k = 0; k < NC; k++) {
        double2 tg = itf[a * 16 + i * NC + k];
        re += tg.x * F.m[k][i].re - tg.y * F.m[k][i].im;
      }
    mom[(size_t)(mu * NG + a) * g.V + id] -= -(coeff * 0.5) * re;
  }
}

// U <- exp(eps P) U, Taylor series to |X^k/k!|^2 < 1e-24 then reunitarised
template <int NC>
__global__ void __launch_bounds__(128) k_update_links(double* __restrict__ fund, const double* __restrict__ mom,
                                                      double eps, Geom g) {
  int t = blockIdx.x * 128 + threadIdx.x;
  if (t >= 4 * g.V) return;
  int id = t >> 2, mu = t & 3;
  M<NC> X = alg<NC>(mom, g.V, mu, id);
  M<NC> R, Xk;
#pragma unroll
  for (int i = 0; i < NC; i++)
#pragma unroll
    for (int j = 0; j < NC; j++) {
      R.m[i][j] = cd(i == j ? 1.0 : 0.0, 0);
      Xk.m[i][j] = R.m[i][j];
    }
  double err = 1.0;
  for (int k = 1; k < 64 && err > 1e-24; k++) {
    Xk = mul<NC>(Xk, X);
    double f = eps / k;
    err = 0;
#pragma unroll
    for (int i = 0; i < NC; i++)
#pragma unroll
      for (int j = 0; j < NC; j++) {
        Xk.m[i][j] = f * Xk.m[i][j];
        R.m[i][j] = R.m[i][j] + Xk.m[i][j];
        err += Xk.m[i][j].re * Xk.m[i][j].re + Xk.m[i][j].im * Xk.m[i][j].im;
      }
  }
  reunit<NC>(R);
  st<NC>(fund, id, mu, mul<NC>(R, ld<NC>(fund, id, mu)));
}

template <int NC>
__global__ void __launch_bounds__(128) k_reunitarize(double* __restrict__ fund, Geom g) {
  int t = blockIdx.x * 128 + threadIdx.x;
  if (t >= 4 * g.V) return;
  M<NC> u = ld<NC>(fund, t >> 2, t & 3);
  reunit<NC>(u);
  st<NC>(fund, t >> 2, t & 3, u);
}

// sum over planes of Re Tr of the plaquette
template <int NC>
__global__ void __launch_bounds__(256) k_plaquette(const double* __restrict__ fund, double* partials, unsigned* ticket,
                                                   double* out, Geom g) {
  int id = blockIdx.x * 256 + threadIdx.x;
  double v[1] = {0.0};
  if (id < g.V) {
    int c[4];
    full_coords(id, g, c);
    for (int mu = 1; mu < 4; mu++)
      for (int nu = 0; nu < mu; nu++) {
        int a = mshift(c, mu, 1, g), b = mshift(c, nu, 1, g);
        M<NC> one = mul<NC>(ld<NC>(fund, id, mu), ld<NC>(fund, a, nu));
        M<NC> two = mul<NC>(ld<NC>(fund, id, nu), ld<NC>(fund, b, mu));
        M<NC> P = muld<NC>(one, two);
        double tr = 0.0;
#pragma unroll
        for (int i = 0; i < NC; i++) tr += P.m[i][i].re;
        v[0] += plaq_w(c[0], mu, nu, g) * tr;
      }
  }
  block_reduce_finish<1, 256>(v, partials, gridDim.x, blockIdx.x, ticket, out);
}

// sum over planes mu > nu of the weighted Re Tr of the h x w loop based at
// each site: h links along mu, w along nu, back along mu, back along nu
// (wilson_loop, wilsonloop.cpp:58-113)
template <int NC>
__global__ void __launch_bounds__(256) k_wilson_loop(const double* __restrict__ fund, int h, int w, double* partials,
                                                     unsigned* ticket, double* out, Geom g) {
  int id = blockIdx.x * 256 + threadIdx.x;
  double v[1] = {0.0};
  if (id < g.V) {
    int c[4];
    full_coords(id, g, c);
    for (int mu = 1; mu < 4; mu++)
      for (int nu = 0; nu < mu; nu++) {
        int p[4] = {c[0], c[1], c[2], c[3]};
        M<NC> P = unit_matrix<NC>();
        for (int k = 0; k < h; k++) path_step<NC>(P, fund, p, mu, true, g);
        for (int k = 0; k < w; k++) path_step<NC>(P, fund, p, nu, true, g);
        for (int k = 0; k < h; k++) path_step<NC>(P, fund, p, mu, false, g);
        for (int k = 0; k < w; k++) path_step<NC>(P, fund, p, nu, false, g);
        double tr = 0.0;
#pragma unroll
        for (int i = 0; i < NC; i++) tr += P.m[i][i].re;
        if (h == 1 && w == 1) tr *= plaq_w(c[0], mu, nu, g);
        else if (h == 1 && w == 2) tr *= rect_w(c[0], mu, nu, g);
        else if (h == 2 && w == 1) tr *= rect_w(c[0], nu, mu, g);
        v[0] += tr;
      }
  }
  block_reduce_finish<1, 256>(v, partials, gridDim.x, blockIdx.x, ticket, out);
}

__global__ void __launch_bounds__(256) k_sumsq(const double* __restrict__ a, long n, double* partials, unsigned* ticket,
                                               double* out) {
  long i0 = (long)blockIdx.x * 256 + threadIdx.x;
  double v[1] = {0.0};
  for (long i = i0; i < n; i += (long)gridDim.x * 256) v[0] += a[i] * a[i];
  block_reduce_finish<1, 256>(v, partials, gridDim.x, blockIdx.x, ticket, out);
}

// ---------------------------------------------------------------------------
int md_init_constants(Ctx* c) {
  int nc = c->nc, ng = nc * nc - 1;
  std::vector<double2> h(15 * 16, make_double2(0, 0));
  for (int a = 0; a < ng; a++) {
    int k = 0;
    for (int i = 0; i < nc; i++)
      for (int j = 0; j < nc; j++) {
        if (i > j) {
          if (k == a) { h[a * 16 + i * nc + j].y += 1; h[a * 16 + j * nc + i].y += 1; }
          k++;
        } else if (i < j) {
          if (k == a) { h[a * 16 + i * nc + j].x += 1; h[a * 16 + j * nc + i].x -= 1; }
          k++;
        } else if (i > 0) {
          if (k == a) {
            double w = sqrt(2.0 / (i * (i + 1)));
            for (int q = 0; q < i; q++) h[a * 16 + q * nc + q].y += w;
            h[a * 16 + i * nc + i].y -= i * w;
          }
          k++;
        }
      }
  }
  return upload_table(c, c->tab_itf, h.data(), h.size());
}

#define YH_NC_SWITCH(...)                                     \
  switch (c->nc) {                                            \
    case 2: { constexpr int NC_ = 2; __VA_ARGS__; } break;    \
    case 3: { constexpr int NC_ = 3; __VA_ARGS__; } break;    \
    case 4: { constexpr int NC_ = 4; __VA_ARGS__; } break;    \
    default: set_error("unsupported NC"); return YH_ERR_UNSUPPORTED; \
  }

int md_gauge_force(Ctx* c, double dt, double beta) {
  int rc;
  if ((rc = force_alloc(c))) return rc;
  double coeff = dt * beta / c->nc;
  dim3 grid((4 * c->g.V + 127) / 128);
  YH_NC_SWITCH(hipLaunchKernelGGL((k_gauge_force<NC_, false>), grid, dim3(128), 0, c->stream, c->fund, c->mom, coeff,
                                  1.0, 0.0, c->g, c->tab_itf));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int md_improved_gauge_force(Ctx* c, double dt, double beta, double c0, double c1) {
  int rc;
  if ((rc = force_alloc(c))) return rc;
  double coeff = dt * beta / c->nc;
  dim3 grid((4 * c->g.V + 127) / 128);
  YH_NC_SWITCH(hipLaunchKernelGGL((k_gauge_force<NC_, true>), grid, dim3(128), 0, c->stream, c->fund, c->mom, coeff,
                                  c0, c1, c->g, c->tab_itf));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

static bool gauge_bcs(const Ctx* c) { return c->g.bc == BC_OPEN || c->g.bc == BC_SF; }

static int mom_bcs(Ctx* c) {
  if (!gauge_bcs(c)) return YH_OK;
  hipLaunchKernelGGL(k_mom_bcs, dim3((4 * c->g.V + 255) / 256), dim3(256), 0, c->stream, c->mom, c->nc * c->nc - 1, c->g);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

// the `volume` the plaquette is normalised by (bcs.cpp:32, 76; geometry.cpp:1150)
static double gauge_volume(const Ctx* c) {
  const double v3 = (double)c->g.V / c->g.L[0];
  if (c->g.bc == BC_OPEN) return (c->g.L[0] - 1) * v3;
  if (c->g.bc == BC_SF) return (c->g.L[0] - 2) * v3;
  return (double)c->g.V;
}

int md_apply_gauge_bcs(Ctx* c) {
  if (!gauge_bcs(c)) return YH_OK;
  SfPhases ph = {};
  if (c->g.bc == BC_SF) {
    // bcs.cpp:208-232: phases v1 / L, v2 / L with L = 2 NC L_X
    const double pi = 3.14159265358979323846;
    const double L = 2.0 * c->nc * c->g.L[1];
    if (c->nc == 2) {
      ph.c1[0] = -pi / L; ph.c1[1] = pi / L;
      ph.c2[0] = -3 * pi / L; ph.c2[1] = 3 * pi / L;
    } else if (c->nc == 3) {
      ph.c1[0] = -pi / L; ph.c1[1] = 0; ph.c1[2] = pi / L;
      ph.c2[0] = -5 * pi / L; ph.c2[1] = 2 * pi / L; ph.c2[2] = 3 * pi / L;
    } else {
      set_error("SF boundary values are defined for NC = 2 and 3");
      return YH_ERR_UNSUPPORTED;
    }
  }
  dim3 grid((4 * c->g.V + 255) / 256);
  YH_NC_SWITCH(hipLaunchKernelGGL((k_gauge_bcs<NC_>), grid, dim3(256), 0, c->stream, c->fund, c->g, ph));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int md_update_links(Ctx* c, double eps) {
  int rc;
  if ((rc = force_alloc(c))) return rc;
  // the fixed links' momenta are zeroed, so exp(0) U leaves them as they are
  // (update.cpp:231-243 zeroes them, updates, and sets the fixed links again)
  if ((rc = mom_bcs(c))) return rc;
  dim3 grid((4 * c->g.V + 127) / 128);
  YH_NC_SWITCH(hipLaunchKernelGGL((k_update_links<NC_>), grid, dim3(128), 0, c->stream, c->fund, c->mom, eps, c->g));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int md_reunitarize(Ctx* c) {
  dim3 grid((4 * c->g.V + 127) / 128);
  YH_NC_SWITCH(hipLaunchKernelGGL((k_reunitarize<NC_>), grid, dim3(128), 0, c->stream, c->fund, c->g));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int md_plaquette(Ctx* c, double* out) {
  dim3 grid((c->g.V + 255) / 256);
  double* dst = &c->sc->red[0];
  YH_NC_SWITCH(hipLaunchKernelGGL((k_plaquette<NC_>), grid, dim3(256), 0, c->stream, c->fund, c->partials, c->ticket,
                                  dst, c->g));
  YH_HIP(hipGetLastError());
  double v;
  YH_HIP(hipMemcpyAsync(&v, dst, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  YH_HIP(hipStreamSynchronize(c->stream));
  *out = v / (c->nc * 6.0 * gauge_volume(c));
  return YH_OK;
}

int md_avr_wilson(Ctx* c, int h, int w, double* out) {
  dim3 grid((c->g.V + 255) / 256);
  double* dst = &c->sc->red[0];
  YH_NC_SWITCH(hipLaunchKernelGGL((k_wilson_loop<NC_>), grid, dim3(256), 0, c->stream, c->fund, h, w, c->partials,
                                  c->ticket, dst, c->g));
  YH_HIP(hipGetLastError());
  double v;
  YH_HIP(hipMemcpyAsync(&v, dst, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  YH_HIP(hipStreamSynchronize(c->stream));
  *out = v / (c->nc * 6.0 * gauge_volume(c));
  return YH_OK;
}

int md_kinetic(Ctx* c, double* out) {
  int rc;
  if ((rc = force_alloc(c))) return rc;
  // hamiltonian() applies the momentum BCs first (update.cpp:201-206)
  if ((rc = mom_bcs(c))) return rc;
  long n = 4L * c->g.V * (c->nc * c->nc - 1);
  int grid = (int)std::min<long>((n + 255) / 256, (long)c->nblk);
  double* dst = &c->sc->red[0];
  hipLaunchKernelGGL(k_sumsq, dim3(grid), dim3(256), 0, c->stream, c->mom, n, c->partials, c->ticket, dst);
  YH_HIP(hipGetLastError());
  double v;
  YH_HIP(hipMemcpyAsync(&v, dst, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  YH_HIP(hipStreamSynchronize(c->stream));
  *out = v;  // TF / 2 = 1
  return YH_OK;
}

}  // namespace yh

// ---------------------------------------------------------------------------
using namespace yh;

static int md_check(Ctx* c) {
  if (!c) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (c->g.bc == YH_BC_SF && c->nc > 3) {
    set_error("SF boundary values are defined for NC = 2 and 3");
    return YH_ERR_UNSUPPORTED;
  }
  YH_HIP(hipSetDevice(c->device));
  return YH_OK;
}

static int md_sync(Ctx* c, int rc) {
  if (rc) return rc;
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

extern "C" {

int yh_get_gauge(yh_ctx* c, double* links) {
  if (!c || !links) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  YH_HIP(hipSetDevice(c->device));
  size_t n = (size_t)4 * c->g.V * 2 * c->nc * c->nc;
  YH_HIP(hipMemcpyAsync(links, c->fund, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

int yh_represent_gauge_field(yh_ctx* c) {
  if (!c) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  YH_HIP(hipSetDevice(c->device));
  int rc;
  if ((rc = gauge_represent(c))) return rc;
  if ((rc = gauge_derive_links(c))) return rc;
  if ((rc = clover_build(c))) return rc;
  return md_sync(c, YH_OK);
}

int yh_wilson_gauge_force(yh_ctx* c, double dt, double beta) {
  int rc = md_check(c);
  if (rc) return rc;
  return md_sync(c, md_gauge_force(c, dt, beta));
}

int yh_improved_gauge_force(yh_ctx* c, double dt, double beta, double c0, double c1) {
  int rc = md_check(c);
  if (rc) return rc;
  return md_sync(c, md_improved_gauge_force(c, dt, beta, c0, c1));
}

int yh_avr_wilson(yh_ctx* c, int height, int width, double* out) {
  if (!out) { set_error("null argument"); return YH_ERR_ARG; }
  if (height < 1 || width < 1 || height > 16 || width > 16) { set_error("avr_wilson: loop sides must be 1..16"); return YH_ERR_ARG; }
  int rc = md_check(c);
  if (rc) return rc;
  return md_avr_wilson(c, height, width, out);
}

int yh_update_links(yh_ctx* c, double eps) {
  int rc = md_check(c);
  if (rc) return rc;
  return md_sync(c, md_update_links(c, eps));
}

int yh_reunitarize(yh_ctx* c) {
  int rc = md_check(c);
  if (rc) return rc;
  return md_sync(c, md_reunitarize(c));
}

int yh_plaquette(yh_ctx* c, double* out) {
  if (!out) { set_error("null argument"); return YH_ERR_ARG; }
  int rc = md_check(c);
  if (rc) return rc;
  return md_plaquette(c, out);
}

int yh_apply_bcs_on_gauge_field(yh_ctx* c) {
  int rc = md_check(c);
  if (rc) return rc;
  return md_sync(c, md_apply_gauge_bcs(c));
}

int yh_momenta_energy(yh_ctx* c, double* out) {
  if (!c || !out) { set_error("null argument"); return YH_ERR_ARG; }
  YH_HIP(hipSetDevice(c->device));
  return md_kinetic(c, out);
}

}  // extern "C"
