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
#include <string.h>
#include <algorithm>
#include <vector>
#include "yh_internal.hpp"

using namespace yh;

namespace {

// AGM sequence for modulus k given its complement kc = sqrt(1 - k^2)
struct Agm {
  std::vector<double> a, c;
  int n;
};

Agm agm(double kc) {
  Agm s;
  double a = 1.0, b = kc, c = sqrt(std::max(0.0, (1.0 - kc) * (1.0 + kc)));
  s.a.push_back(a);
  s.c.push_back(c);
  for (int it = 0; it < 64 && fabs(c) > 1e-17 * a; it++) {
    double an = 0.5 * (a + b), bn = sqrt(a * b), cn = 0.5 * (a - b);
    a = an;
    b = bn;
    c = cn;
    s.a.push_back(a);
    s.c.push_back(c);
  }
  s.n = (int)s.a.size() - 1;
  return s;
}

double ellip_k(double kc) {
  Agm s = agm(kc);
  return M_PI / (2.0 * s.a[s.n]);
}

// sn, cn, dn(u; k) with k given through its complement kc
// (A&S 16.4.3-16.4.4: phi_{j-1} = (phi_j + asin(c_j / a_j sin phi_j)) / 2,
//  sn = sin phi_0, cn = cos phi_0, dn = cos phi_0 / cos(phi_1 - phi_0))
void jacobi(double u, double kc, double* sn, double* cn, double* dn) {
  Agm s = agm(kc);
  double phi = ldexp(1.0, s.n) * s.a[s.n] * u;
  double phi1 = phi;
  for (int j = s.n; j >= 1; j--) {
    phi1 = phi;
    phi = 0.5 * (phi + asin(s.c[j] / s.a[j] * sin(phi)));
  }
  *sn = sin(phi);
  *cn = cos(phi);
  *dn = s.n >= 1 ? cos(phi) / cos(phi1 - phi) : 1.0;
}

double jacobi_sn(double u, double kc) {
  double sn, cn, dn;
  jacobi(u, kc, &sn, &cn, &dn);
  return sn;
}

// sn^2 / cn^2 at u, accurate up to u -> K: for u > K/2 use
// sn(K - v) = cn(v)/dn(v), cn(K - v) = kc sn(v)/dn(v)
double sc2(double u, double K, double kc) {
  double sn, cn, dn;
  if (u <= 0.5 * K) {
    jacobi(u, kc, &sn, &cn, &dn);
    return (sn / cn) * (sn / cn);
  }
  jacobi(K - u, kc, &sn, &cn, &dn);
  double t = cn / (kc * sn);
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

}  // namespace

extern "C" {

double yh_jacobi_sn(double u, double kc) { return jacobi_sn(u, kc); }
double yh_ellip_k(double kc) { return ellip_k(kc); }

int yh_rational_zolotarev(int order, int n, int d, double min, double max, yh_rational* out) {
  if (!out || order < 1 || order > 48 || !(min > 0.0) || !(max > min)) { set_error("bad argument"); return YH_ERR_ARG; }
  if (abs(n) != 1 || abs(d) != 2) { set_error("only x^(+-1/2) is generated (Zolotarev)"); return YH_ERR_UNSUPPORTED; }
  const double b = max / min;
  const double kc = sqrt(1.0 / b);  // complementary modulus of k' = sqrt(1 - 1/b)
  const double K = ellip_k(kc);
  const int m = 2 * order + 1;
  std::vector<double> c(2 * order + 1);
  for (int l = 1; l <= 2 * order; l++) c[l] = sc2(l * K / m, K, kc);
  memset(out, 0, sizeof(*out));
  out->n = n;
  out->d = d;
  out->order = order;
  out->min = min;
  out->max = max;
  // x^(-1/2) on [min, max]: 1/sqrt(x) = min^(-1/2) * 1/sqrt(x / min)
  for (int l = 1; l <= order; l++) {
    out->root[l - 1] = min * c[2 * l];
    out->pole[l - 1] = min * c[2 * l - 1];
  }
  out->a0 = 1.0 / sqrt(min);
  // equioscillation: scale so that max and min of sqrt(x) r(x) straddle 1.
  // Extremes are located on a log grid and refined by golden-section search
  // (the error curve is smooth between grid points).
  const int grid = 4000;
  auto err = [&](double lx) { double x = min * exp(lx); return sqrt(x) * product_form(out, x); };
  const double lb = log(b), h = lb / grid;
  std::vector<double> e(grid + 1);
  for (int i = 0; i <= grid; i++) e[i] = err(i * h);
  double emax = std::max(e[0], e[grid]), emin = std::min(e[0], e[grid]);
  for (int i = 1; i < grid; i++) {
    bool is_max = e[i] >= e[i - 1] && e[i] >= e[i + 1];
    bool is_min = e[i] <= e[i - 1] && e[i] <= e[i + 1];
    if (!is_max && !is_min) continue;
    double lo = (i - 1) * h, hi = (i + 1) * h;
    const double gr = 0.6180339887498949;
    for (int it = 0; it < 60; it++) {
      double m1 = hi - gr * (hi - lo), m2 = lo + gr * (hi - lo);
      double f1 = err(m1), f2 = err(m2);
      if (is_max ? (f1 > f2) : (f1 < f2)) hi = m2;
      else lo = m1;
    }
    double v = err(0.5 * (lo + hi));
    if (is_max) emax = std::max(emax, std::max(v, e[i]));
    else emin = std::min(emin, std::min(v, e[i]));
  }
  out->a0 *= 2.0 / (emax + emin);
  out->prec = (emax - emin) / (emax + emin);
  if (n * d > 0) {
    // x^(+1/2) = x * x^(-1/2): the reference swaps roots and poles of the
    // x^(-1/2) approximation and inverts a0 (rational.cpp:111-122)
    for (int i = 0; i < order; i++) std::swap(out->root[i], out->pole[i]);
    out->a0 = 1.0 / out->a0;
  }
  for (int i = 0; i < order; i++) out->residue[i] = residue_of(out, i);
  return YH_OK;
}

int yh_rational_find(int n, int d, double prec, double min, double max, yh_rational* out) {
  // lowest order whose maximum relative error is below prec
  // (library/rational.cpp:56-132 picks the lowest-order table entry likewise)
  for (int order = 1; order <= 48; order++) {
    int rc = yh_rational_zolotarev(order, n, d, min, max, out);
    if (rc) return rc;
    if (out->prec < prec) return YH_OK;
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
