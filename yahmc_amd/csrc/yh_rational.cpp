// Rational approximations for RHMC and the partial-fraction application.
//
// The reference looks approximations up in a precomputed table
// (library/rational.cpp:56-132, include/approx.h) that stops at orders and
// spectral ranges too small for a 12-pole x^(-1/2) on [1e-4, 8].  Here the
// approximation of x^(-1/2) is generated directly: Zolotarev's optimal
// rational approximation (Zolotarev 1877; in the lattice form of Chiu,
// Hsieh, Huang & Lee, Phys. Rev. D 66 (2002) 114502):
//
//   on y in [1, b]:   1/sqrt(y) ~ d0 * prod_{l=1..n} (y + c_{2l}) / (y + c_{2l-1})
//   c_l = sn^2(l K'/(2n+1); k') / (1 - sn^2(l K'/(2n+1); k')),
//   k' = sqrt(1 - 1/b), K' = K(k'),
//
// with d0 fixed so the relative error equioscillates.  The Jacobi function
// sn and the complete elliptic integral K come from the arithmetic-geometric
// mean (Abramowitz & Stegun 16.4, 17.6), parametrised by the complementary
// modulus sqrt(1/b) to avoid cancellation when b is large.  Positive powers
// swap roots and poles and invert a0 as the reference does
// (rational.cpp:111-122); residues follow rational.cpp:10-36.
#include <math.h>
#include <cmath>
#include <string.h>
#include <algorithm>
#include <vector>
#include "yh_internal.hpp"

using namespace yh;

namespace {

typedef long double LD;

// AGM sequence for modulus k given its complement kc = sqrt(1 - k^2); long
// double, so that high-order nodes keep their accuracy after rounding to
// double
struct Agm {
  std::vector<LD> a, c;
  int n;
};

Agm agm(LD kc) {
  Agm s;
  LD a = 1.0L, b = kc, c = sqrtl(std::max(0.0L, (1.0L - kc) * (1.0L + kc)));
  s.a.push_back(a);
  s.c.push_back(c);
  for (int it = 0; it < 64 && fabsl(c) > 1e-20L * a; it++) {
    LD an = 0.5L * (a + b), bn = sqrtl(a * b), cn = 0.5L * (a - b);
    a = an;
    b = bn;
    c = cn;
    s.a.push_back(a);
    s.c.push_back(c);
  }
  s.n = (int)s.a.size() - 1;
  return s;
}

LD ellip_k(LD kc) {
  Agm s = agm(kc);
  return 3.14159265358979323846264338327950288L / (2.0L * s.a[s.n]);
}

// sn, cn, dn(u; k) with k given through its complement kc
// (A&S 16.4.3-16.4.4: phi_{j-1} = (phi_j + asin(c_j / a_j sin phi_j)) / 2,
//  sn = sin phi_0, cn = cos phi_0, dn = cos phi_0 / cos(phi_1 - phi_0))
void jacobi(LD u, LD kc, LD* sn, LD* cn, LD* dn) {
  Agm s = agm(kc);
  LD phi = ldexpl(1.0L, s.n) * s.a[s.n] * u;
  LD phi1 = phi;
  for (int j = s.n; j >= 1; j--) {
    phi1 = phi;
    phi = 0.5L * (phi + asinl(s.c[j] / s.a[j] * sinl(phi)));
  }
  *sn = sinl(phi);
  *cn = cosl(phi);
  *dn = s.n >= 1 ? cosl(phi) / cosl(phi1 - phi) : 1.0L;
}

LD jacobi_sn(LD u, LD kc) {
  LD sn, cn, dn;
  jacobi(u, kc, &sn, &cn, &dn);
  return sn;
}

// sn^2 / cn^2 at u, accurate up to u -> K: for u > K/2 use
// sn(K - v) = cn(v)/dn(v), cn(K - v) = kc sn(v)/dn(v)
LD sc2(LD u, LD K, LD kc) {
  LD sn, cn, dn;
  if (u <= 0.5L * K) {
    jacobi(u, kc, &sn, &cn, &dn);
    return (sn / cn) * (sn / cn);
  }
  jacobi(K - u, kc, &sn, &cn, &dn);
  LD t = cn / (kc * sn);
  return t * t;
}

double product_form(const yh_rational* r, double x) {
  double v = r->a0;
  for (int i = 0; i < r->order; i++) v *= (x + r->root[i]) / (x + r->pole[i]);
  return v;
}

// residue of the partial fraction at pole n (library/rational.cpp:10-36)
double residue_of(const yh_rational* r, int n) {
  double x = r->pole[n];
  double q = 1.0, p = 0.0;
  for (int i = 0; i < r->order; i++) q *= (r->root[i] - x);
  for (int i = 0; i < r->order; i++) {
    double t = 1.0;
    for (int j = 0; j < r->order; j++)
      if (i != j) t *= (r->pole[j] - x);
    p += t;
  }
  return (q / p) * r->a0;
}

// Zolotarev's roots and poles of y^(-1/2) on [1, b], y = x / min
void zolotarev_nodes(int order, double b, LD* root, LD* pole) {
  const LD kc = sqrtl(1.0L / b);  // complementary modulus of k' = sqrt(1 - 1/b)
  const LD K = ellip_k(kc);
  const int m = 2 * order + 1;
  for (int l = 1; l <= order; l++) {
    root[l - 1] = sc2(2 * l * K / m, K, kc);
    pole[l - 1] = sc2((2 * l - 1) * K / m, K, kc);
  }
}

// ---------------------------------------------------------------------------
// Remez exchange for the best relative approximation of y^s, -1 < s < 0, on
// [1, b] by R(y) = A prod_i (y + r_i) / (y + p_i).
//
// Unknowns u = (log r_1..n, log p_1..n, log A, E).  On a reference set of
// 2n + 2 points t_j = log y_j the equioscillation conditions
//   g(t_j) = log R(y_j) - s t_j = log(1 + (-1)^j E)
// are solved by Newton's method; the log parametrisation keeps roots and
// poles positive and the system well scaled.  The extrema of the error
// e(t) = exp(g(t)) - 1 then become the next reference set, until the
// extremal values are level (Remez's second algorithm; Cheney, Introduction
// to Approximation Theory, ch. 3).  The optimal approximation of y^(-1/2)
// is known in closed form (Zolotarev), and the optimum depends smoothly on
// s, so the iteration starts there and follows s to its target in steps
// (continuation), each step starting from the previous optimum.  Long double
// keeps the Newton residuals well below the 1e-15 errors of high orders.
// ---------------------------------------------------------------------------
struct Remez {
  int n = 0;
  LD s = -0.5L, lb = 0;            // exponent, log b
  std::vector<LD> lr, lp;          // log roots, log poles
  LD la = 0, E = 0;                // log A, levelled error
  std::vector<LD> ref;             // reference set (t values)

  LD g(LD t) const {
    const LD y = expl(t);
    LD v = la - s * t;
    for (int i = 0; i < n; i++) v += logl(y + expl(lr[i])) - logl(y + expl(lp[i]));
    return v;
  }
  LD err(LD t) const { return expm1l(g(t)); }

  // Newton's method on the reference set; false if it does not converge
  bool newton() {
    const int N = 2 * n + 2;
    std::vector<LD> J((size_t)N * N), F(N), rr(n), pp(n);
    for (int it = 0; it < 60; it++) {
      for (int i = 0; i < n; i++) { rr[i] = expl(lr[i]); pp[i] = expl(lp[i]); }
      for (int j = 0; j < N; j++) {
        const LD t = ref[j], y = expl(t), sg = (j & 1) ? -1.0L : 1.0L;
        if (!(1 + sg * E > 0)) return false;
        LD v = la - s * t;
        for (int i = 0; i < n; i++) {
          v += logl(y + rr[i]) - logl(y + pp[i]);
          J[(size_t)j * N + i] = rr[i] / (y + rr[i]);
          J[(size_t)j * N + n + i] = -pp[i] / (y + pp[i]);
        }
        J[(size_t)j * N + 2 * n] = 1;
        J[(size_t)j * N + 2 * n + 1] = -sg / (1 + sg * E);
        F[j] = -(v - log1pl(sg * E));
      }
      // converged when the conditions hold to far below E (the individual
      // roots and poles are determined less sharply than the values)
      LD res = 0;
      for (int j = 0; j < N; j++) res = std::max(res, fabsl(F[j]));
      if (res < std::max(1e-16L, 1e-10L * fabsl(E))) return true;
      // Gaussian elimination with partial pivoting, J du = F
      for (int k = 0; k < N; k++) {
        int piv = k;
        for (int r = k + 1; r < N; r++)
          if (fabsl(J[(size_t)r * N + k]) > fabsl(J[(size_t)piv * N + k])) piv = r;
        if (J[(size_t)piv * N + k] == 0) return false;
        if (piv != k) {
          for (int c = 0; c < N; c++) std::swap(J[(size_t)k * N + c], J[(size_t)piv * N + c]);
          std::swap(F[k], F[piv]);
        }
        for (int r = k + 1; r < N; r++) {
          const LD f = J[(size_t)r * N + k] / J[(size_t)k * N + k];
          if (f == 0) continue;
          for (int c = k; c < N; c++) J[(size_t)r * N + c] -= f * J[(size_t)k * N + c];
          F[r] -= f * F[k];
        }
      }
      for (int k = N - 1; k >= 0; k--) {
        LD v = F[k];
        for (int c = k + 1; c < N; c++) v -= J[(size_t)k * N + c] * F[c];
        F[k] = v / J[(size_t)k * N + k];
      }
      // damp steps of the log parameters to at most 0.5
      LD big = 0;
      for (int k = 0; k < 2 * n + 1; k++) big = std::max(big, fabsl(F[k]));
      const LD damp = big > 0.5L ? 0.5L / big : 1.0L;
      for (int i = 0; i < n; i++) { lr[i] += damp * F[i]; lp[i] += damp * F[n + i]; }
      la += damp * F[2 * n];
      E += damp * F[2 * n + 1];
      if (!std::isfinite((double)la) || !std::isfinite((double)E)) return false;
    }
    return false;
  }

  // local extrema of e with alternating signs, one per sign run of e on a
  // grid of log y, refined by golden-section search; returns their count
  int extrema(std::vector<LD>& pts, std::vector<LD>& vals) const {
    const int G = std::max(2000, 120 * (2 * n + 2));
    const LD h = lb / G;
    std::vector<LD> e(G + 1);
    for (int k = 0; k <= G; k++) e[k] = err(k * h);
    pts.clear();
    vals.clear();
    int k = 0;
    while (k <= G) {
      const bool pos = e[k] >= 0;
      int best = k;
      while (k <= G && ((e[k] >= 0) == pos || e[k] == 0)) {
        if (fabsl(e[k]) > fabsl(e[best])) best = k;
        k++;
      }
      LD lo = best > 0 ? (best - 1) * h : 0, hi = best < G ? (best + 1) * h : lb;
      const LD sg = pos ? 1 : -1, gr = 0.61803398874989484820L;
      LD m1 = hi - gr * (hi - lo), m2 = lo + gr * (hi - lo);
      LD f1 = sg * err(m1), f2 = sg * err(m2);
      for (int it = 0; it < 50; it++) {
        if (f1 > f2) { hi = m2; m2 = m1; f2 = f1; m1 = hi - gr * (hi - lo); f1 = sg * err(m1); }
        else { lo = m1; m1 = m2; f1 = f2; m2 = lo + gr * (hi - lo); f2 = sg * err(m2); }
      }
      LD t = 0.5L * (lo + hi), v = err(t);
      if (fabsl(v) < fabsl(e[best])) { t = best * h; v = e[best]; }
      pts.push_back(t);
      vals.push_back(v);
    }
    return (int)pts.size();
  }

  // one exchange: new reference set from the extrema; returns the relative
  // spread of the extremal values, or a negative number on failure
  LD exchange(LD* emax) {
    std::vector<LD> pts, vals;
    int m = extrema(pts, vals);
    const int N = 2 * n + 2;
    if (m < N) return -1;
    // the maximum error is over all extrema, not only the reference set
    LD all = 0;
    for (int j = 0; j < m; j++) all = std::max(all, fabsl(vals[j]));
    *emax = all;
    // too many alternations (spurious sign runs near zeros of e, or early
    // iterates): drop the smallest extremum; inside the set its two
    // neighbours then share a sign and the smaller of them goes too
    while ((int)pts.size() > N) {
      size_t k = 0;
      for (size_t j = 1; j < vals.size(); j++)
        if (fabsl(vals[j]) < fabsl(vals[k])) k = j;
      if (k == 0 || k + 1 == vals.size()) {
        pts.erase(pts.begin() + k);
        vals.erase(vals.begin() + k);
        continue;
      }
      size_t drop = fabsl(vals[k - 1]) < fabsl(vals[k + 1]) ? k - 1 : k + 1;
      size_t first = std::min(k, drop);
      pts.erase(pts.begin() + first, pts.begin() + first + 2);
      vals.erase(vals.begin() + first, vals.begin() + first + 2);
      if ((int)pts.size() < N) return -1;
    }
    LD lo = 1e300L;
    for (int j = 0; j < N; j++) {
      ref[j] = pts[j];
      lo = std::min(lo, fabsl(vals[j]));
    }
    // keep E's sign convention: reference point 0 carries +E
    if ((vals[0] >= 0) != (E >= 0)) E = -E;
    return (all - lo) / all;
  }

  // Remez iterations at the current s until the extrema are level
  bool solve(LD* emax) {
    for (int it = 0; it < 40; it++) {
      if (!newton()) return false;
      LD spread = exchange(emax);
      if (spread < 0) return false;
      // level to 1e-9, or to what long double resolves for tiny errors
      if (spread < std::max(1e-9L, 1e-16L / *emax)) return true;
    }
    return false;
  }
};

// best approximation of y^s on [1, b] of the given order; fills normalised
// roots and poles, log A and the maximum relative error
int remez(int order, double s_target, double b, double* root, double* pole, double* A, double* prec) {
  Remez R;
  R.n = order;
  R.lb = logl((LD)b);
  R.lr.resize(order);
  R.lp.resize(order);
  R.ref.resize(2 * order + 2);
  {
    std::vector<LD> r0(order), p0(order);
    zolotarev_nodes(order, b, r0.data(), p0.data());
    for (int i = 0; i < order; i++) { R.lr[i] = logl(r0[i]); R.lp[i] = logl(p0[i]); }
  }
  // centre the closed-form error on zero, then take its extrema
  R.la = 0;
  {
    LD gmin = 1e300L, gmax = -1e300L;
    const int G = 4000;
    for (int k = 0; k <= G; k++) {
      LD v = R.g(R.lb * k / G);
      gmin = std::min(gmin, v);
      gmax = std::max(gmax, v);
    }
    R.la = -0.5L * (gmin + gmax);
    R.E = 0.5L * (gmax - gmin);
    LD emax;
    if (R.exchange(&emax) < 0) return YH_ERR_NOCONV;
  }
  LD emax = 0;
  if (!R.solve(&emax)) return YH_ERR_NOCONV;
  // continuation in s from -1/2 to the target
  const LD target = s_target;
  LD ds = (target - R.s) / 4;
  Remez good = R;
  LD good_emax = emax;
  while (fabsl(good.s - target) > 0) {
    if (fabsl(ds) < 1e-4L) return YH_ERR_NOCONV;
    R = good;
    R.s = fabsl(target - good.s) <= fabsl(ds) ? target : good.s + ds;
    if (R.solve(&emax)) {
      good = R;
      good_emax = emax;
      ds *= 1.5L;
    } else {
      ds *= 0.5L;
    }
  }
  for (int i = 0; i < order; i++) { root[i] = (double)expl(good.lr[i]); pole[i] = (double)expl(good.lp[i]); }
  *A = (double)good.la;
  *prec = (double)good_emax;
  return YH_OK;
}

}  // namespace

extern "C" {

double yh_jacobi_sn(double u, double kc) { return (double)jacobi_sn(u, kc); }
double yh_ellip_k(double kc) { return (double)ellip_k(kc); }

int yh_rational_zolotarev(int order, int n, int d, double min, double max, yh_rational* out) {
  if (!out || order < 1 || order > 48 || !(min > 0.0) || !(max > min)) { set_error("bad argument"); return YH_ERR_ARG; }
  if (abs(n) != 1 || abs(d) != 2) { set_error("only x^(+-1/2) is generated (Zolotarev)"); return YH_ERR_UNSUPPORTED; }
  const double b = max / min;
  memset(out, 0, sizeof(*out));
  out->n = n;
  out->d = d;
  out->order = order;
  out->min = min;
  out->max = max;
  // x^(-1/2) on [min, max]: 1/sqrt(x) = min^(-1/2) * 1/sqrt(x / min)
  {
    std::vector<LD> r0(order), p0(order);
    zolotarev_nodes(order, b, r0.data(), p0.data());
    for (int l = 0; l < order; l++) {
      out->root[l] = (double)(min * r0[l]);
      out->pole[l] = (double)(min * p0[l]);
    }
  }
  out->a0 = 1.0 / sqrt(min);
  // equioscillation: scale so that max and min of sqrt(x) r(x) straddle 1.
  // Extremes are located on a log grid and refined by golden-section search
  // (the error curve is smooth between grid points).
  const int grid = 4000;
  // (long double, so that the fit resolves errors down to 1e-15)
  auto err = [&](LD lx) {
    const LD x = min * expl(lx);
    LD v = out->a0 * sqrtl(x);
    for (int i = 0; i < order; i++) v *= (x + out->root[i]) / (x + out->pole[i]);
    return v;
  };
  const LD lb = logl((LD)b), h = lb / grid;
  std::vector<LD> e(grid + 1);
  for (int i = 0; i <= grid; i++) e[i] = err(i * h);
  LD emax = std::max(e[0], e[grid]), emin = std::min(e[0], e[grid]);
  for (int i = 1; i < grid; i++) {
    bool is_max = e[i] >= e[i - 1] && e[i] >= e[i + 1];
    bool is_min = e[i] <= e[i - 1] && e[i] <= e[i + 1];
    if (!is_max && !is_min) continue;
    LD lo = (i - 1) * h, hi = (i + 1) * h;
    const LD gr = 0.61803398874989484820L;
    for (int it = 0; it < 60; it++) {
      LD m1 = hi - gr * (hi - lo), m2 = lo + gr * (hi - lo);
      LD f1 = err(m1), f2 = err(m2);
      if (is_max ? (f1 > f2) : (f1 < f2)) hi = m2;
      else lo = m1;
    }
    LD v = err(0.5L * (lo + hi));
    if (is_max) emax = std::max(emax, std::max(v, e[i]));
    else emin = std::min(emin, std::min(v, e[i]));
  }
  out->a0 = (double)(out->a0 * 2.0L / (emax + emin));
  out->prec = (double)((emax - emin) / (emax + emin));
  if (n * d > 0) {
    // x^(+1/2) = x * x^(-1/2): the reference swaps roots and poles of the
    // x^(-1/2) approximation and inverts a0 (rational.cpp:111-122)
    for (int i = 0; i < order; i++) std::swap(out->root[i], out->pole[i]);
    out->a0 = 1.0 / out->a0;
  }
  for (int i = 0; i < order; i++) out->residue[i] = residue_of(out, i);
  return YH_OK;
}

int yh_rational_remez(int order, int n, int d, double min, double max, yh_rational* out) {
  if (!out || order < 1 || order > 48 || !(min > 0.0) || !(max > min)) { set_error("bad argument"); return YH_ERR_ARG; }
  if (n == 0 || d == 0 || abs(n) >= abs(d)) { set_error("powers 0 < |n/d| < 1 only"); return YH_ERR_UNSUPPORTED; }
  const double s = -(double)abs(n) / abs(d);
  memset(out, 0, sizeof(*out));
  double la, prec;
  int rc = remez(order, s, max / min, out->root, out->pole, &la, &prec);
  if (rc) {
    set_error("Remez iteration did not converge");
    return rc;
  }
  out->n = n;
  out->d = d;
  out->order = order;
  out->min = min;
  out->max = max;
  // y = x / min: R(x) = A min^s prod (x + min r) / (x + min p)
  for (int i = 0; i < order; i++) {
    out->root[i] *= min;
    out->pole[i] *= min;
  }
  out->a0 = exp(la) * pow(min, s);
  out->prec = prec;
  if (n * d > 0) {
    // positive power: the reciprocal, as the table lookup does
    // (rational.cpp:111-122); its relative error is E / (1 - E)
    for (int i = 0; i < order; i++) std::swap(out->root[i], out->pole[i]);
    out->a0 = 1.0 / out->a0;
    out->prec = prec / (1.0 - prec);
  }
  for (int i = 0; i < order; i++) out->residue[i] = residue_of(out, i);
  return YH_OK;
}

int yh_rational_find(int n, int d, double prec, double min, double max, yh_rational* out) {
  // lowest order whose maximum relative error is below prec
  // (library/rational.cpp:56-132 picks the lowest-order table entry likewise):
  // x^(+-1/2) from the closed form, other powers by the Remez exchange with a
  // bisection over the order (the optimal error falls monotonically with it)
  if (!out || !(prec > 0.0)) { set_error("bad argument"); return YH_ERR_ARG; }
  const bool half = abs(n) * 2 == abs(d);
  auto make = [&](int order) {
    return half ? yh_rational_zolotarev(order, n, d, min, max, out) : yh_rational_remez(order, n, d, min, max, out);
  };
  if (half) {
    for (int order = 1; order <= 48; order++) {
      int rc = make(order);
      if (rc) return rc;
      if (out->prec < prec) return YH_OK;
    }
  } else {
    // upper bound: the order x^(-1/2) needs on the interval (powers nearer
    // zero need fewer poles), stepped up if it is not enough
    int hi = 48;
    yh_rational z;
    for (int order = 1; order <= 48; order++) {
      if (yh_rational_zolotarev(order, -1, 2, min, max, &z)) return YH_ERR_ARG;
      if (z.prec < prec) { hi = order; break; }
    }
    int rc = make(hi);
    while (!rc && out->prec >= prec && hi < 48) rc = make(++hi);
    if (rc) return rc;
    if (out->prec < prec) {
      int lo = 0;  // prec fails at lo, holds at hi
      while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        rc = make(mid);
        if (rc) return rc;
        if (out->prec < prec) hi = mid;
        else lo = mid;
      }
      return make(hi);
    }
  }
  set_error("no rational approximation of order <= 48 reaches the requested precision");
  return YH_ERR_NOCONV;
}

int yh_rational_calc(yh_ctx* cc, const yh_rational* r, const yh_inv_par* par, yh_field* dst, const yh_field* src,
                     int* iters) {
  Ctx* c = cc;
  if (!c || !r || !par || !dst || !src || r->order < 1 || r->order > 48) { set_error("bad argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (dst->prec != PREC_F64 || src->prec != PREC_F64) { set_error("fields must be FP64"); return YH_ERR_ARG; }
  std::vector<Field*> xs(r->order, nullptr);
  int rc = YH_OK;
  for (int i = 0; i < r->order && !rc; i++) rc = field_alloc(c, PREC_F64, &xs[i]);
  int it = 0;
  if (!rc) {
    yh_inv_par p = *par;
    p.mvm = YH_MVM_HPHI_SQ;
    const Field* b = src;
    rc = mshift_device(c, &p, 1, r->order, r->pole, xs.data(), &b, &it);
  }
  // dst = a0 src + sum_i residue_i x_i  (rational.cpp:158-162)
  if (!rc) rc = blas_axpby(c, dst, r->a0, src, 0.0, 0);
  for (int i = 0; i < r->order && !rc; i++) rc = blas_axpby(c, dst, r->residue[i], xs[i], 1.0, 0);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); rc = YH_ERR_HIP; }
  }
  for (Field* f : xs) field_free(f);
  if (iters) *iters = it;
  return rc;
}

double yh_rational_eval(const yh_rational* r, double x) {
  double v = r->a0;
  for (int i = 0; i < r->order; i++) v += r->residue[i] / (x + r->pole[i]);
  return v;
}

}  // extern "C"
