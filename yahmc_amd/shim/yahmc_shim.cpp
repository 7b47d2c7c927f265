// Drop-in C++ shim: the reference's own call surface (same C++ signatures)
// implemented over the C ABI of libyahmc.
//
// A YaHMC build links this file and libyahmc.so in place of eight library
// translation units, and changes nothing else:
//   library/dirac.cpp     dirac.h     mvm_reset .. Hphi_sq
//   library/clover.cpp    clover.h    clover_init, get_csw, clover_logdet, ...
//   library/cg.cpp        inverter.h  cg_inverter, cg_mshift
//   library/bicgstab.cpp  inverter.h  bicgstab_inverter
//   library/minres.cpp    inverter.h  minres_inverter
//   library/inverter.cpp  inverter.h  auto_inverter
//   library/forces.cpp    forces.h    fermion and gauge forces
//   library/rational.cpp  rational.h  rational_find, rational_calc
// Build: the reference's flags (EO update required), plus
//   -I<this repo>/include, linking -L<this repo>/yahmc_amd/lib -lyahmc.
//
// Fields stay where the reference keeps them (host SpinorField, gfield_fund,
// afield_momenta); every call moves what it reads to the device and what it
// writes back, so host code that touches the arrays directly stays correct.
// Solvers move one source and one solution per solve; the iterations run on
// the device.  Two things are cached on the device:
//   - the gauge field: represent_gauge_field() (repr.cpp:265-280) ends every
//     change of the links with compute_clover_term(), which here marks the
//     device copy stale; the next operator call uploads gfield_fund;
//   - the momenta between begin_fermion_force() and end_fermion_force()
//     (monomials/list.cpp:11-30 brackets every fermion level with them).
//
// Operators are recognised by function-pointer identity, as auto_inverter
// does (inverter.cpp:27-54): Dphi, Hphi and Hphi_sq run on the device; any
// other mvm_operator is called as a black box by host solvers written over
// linalg.h (each of its calls may itself be device-backed).  auto_inverter
// keeps the reference's CRITICAL for an unknown operator.
//
// Errors of the C ABI become lprintf(CRITICAL), the reference's own error
// path (logger.cpp:53-56).
//
// Environment: YAHMC_DEVICE (HIP device ordinal, default 0); YAHMC_SLOPPY
// = f32 | f16 makes cg_inverter on Hphi_sq a mixed-precision solve with FP64
// reliable updates (same FP64 true residual; default: plain FP64 CG).
#include <global.h>
#include <dirac.h>
#include <inverter.h>
#include <forces.h>
#include <linalg.h>
#include <rational.h>
#include <clover.h>
#include <bcs.h>
#include <repr.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "yahmc.h"

#ifndef UPDATE_EO
#error "the device build implements the even/odd-preconditioned operators (UPDATE_EO)"
#endif

static_assert(sizeof(Spinor) == sizeof(double) * 8 * REPR_DIM, "Spinor must be re[DR][4], im[DR][4]");
static_assert(sizeof(suNg) == sizeof(double) * 2 * NC * NC, "suNg must be re[NC*NC], im[NC*NC]");
static_assert(sizeof(suNa) == sizeof(double) * NG, "suNa must be av[NG]");
static_assert(sizeof(rational_app) == sizeof(yh_rational), "rational_app and yh_rational must match");

// the reference's clover globals (clover.h:35-37): the clover term, its LDL
// factors and the clover force live on the device in this build
suNf* cl_force = nullptr;
suNf* cl_term = nullptr;
ldl_t* cl_ldl = nullptr;

namespace {

struct Device {
  yh_ctx* ctx = nullptr;
  bool gauge_stale = true;
  bool mom_on_device = false;  // inside begin_fermion_force .. end_fermion_force
  double csw = 0.0;
  std::vector<yh_field*> pool;  // FP64 scratch fields, reused across calls
};

Device& dev() {
  static Device d;
  return d;
}

void check(int rc, const char* where) {
  if (rc != YH_OK) lprintf("YAHMC", CRITICAL, "%s: %s (status %d)", where, yh_last_error(), rc);
}

int bc_id() {
#if defined(BC_SF)
  return YH_BC_SF;
#elif defined(BC_OPEN)
  return YH_BC_OPEN;
#elif defined(BC_ANTIPERIODIC)
  return YH_BC_ANTIPERIODIC;
#else
  return YH_BC_PERIODIC;
#endif
}

yh_ctx* context() {
  Device& d = dev();
  if (d.ctx) return d.ctx;
  if (np_t * np_x * np_y * np_z != 1)
    lprintf("YAHMC", CRITICAL, "the device build runs the whole lattice on one GPU (one process)");
  yh_lattice lat;
  for (int m = 0; m < 4; m++) lat.dims[m] = global.dim[m];
  lat.nc = NC;
  lat.repr = REPR_ID;
  lat.bc = bc_id();
  lat.csw = d.csw;
  lat.cf = bcs.cf;
  const char* env = getenv("YAHMC_DEVICE");
  lat.device = env ? atoi(env) : 0;
  check(yh_create(&lat, &d.ctx), "yh_create");
  lprintf("YAHMC", INFO, "device context: %dx%dx%dx%d, NC = %d, %s, csw = %1.4f", lat.dims[0], lat.dims[1],
          lat.dims[2], lat.dims[3], NC, REPR_NAME, lat.csw);
  return d.ctx;
}

// the context with the current gauge field
yh_ctx* ready() {
  yh_ctx* c = context();
  Device& d = dev();
  if (d.gauge_stale) {
    check(yh_set_gauge(c, &gfield_fund[0].re[0]), "yh_set_gauge");
    d.gauge_stale = false;
  }
  return c;
}

yh_field* scratch(int i) {
  Device& d = dev();
  while ((int)d.pool.size() <= i) {
    yh_field* f = nullptr;
    check(yh_field_create(context(), YH_PREC_F64, &f), "yh_field_create");
    d.pool.push_back(f);
  }
  return d.pool[i];
}

double* host(SpinorField& s) { return &s.ptr[0].re[0][0]; }

void up(yh_field* f, SpinorField& s, int mask) { check(yh_field_upload(context(), f, host(s), mask), "upload"); }
void down(yh_field* f, SpinorField& s, int mask) { check(yh_field_download(context(), f, host(s), mask), "download"); }

bool is_device_op(mvm_operator m) { return m == Dphi || m == Hphi || m == Hphi_sq; }

int mvm_id(mvm_operator m) { return m == Dphi ? YH_MVM_DPHI : m == Hphi ? YH_MVM_HPHI : YH_MVM_HPHI_SQ; }

yh_inv_par device_par(const inv_par& par) {
  yh_inv_par p;
  memset(&p, 0, sizeof(p));
  p.mass = par.mass;
  p.prec = par.prec;
  p.max_iter = 0;  // the reference has no iteration cap
  p.mvm = mvm_id(par.mvm);
  p.log = par.log;
  return p;
}

// momenta on the device for one force call (or for the whole bracket)
void momenta_in(yh_ctx* c) {
  if (!dev().mom_on_device) check(yh_set_momenta(c, &afield_momenta[0].av[0]), "yh_set_momenta");
}
void momenta_out(yh_ctx* c) {
  if (!dev().mom_on_device) check(yh_get_momenta(c, &afield_momenta[0].av[0]), "yh_get_momenta");
}

// ---- host solvers for operators given as black boxes ----------------------
// (the recurrences of cg.cpp:11-202, bicgstab.cpp:9-97 and a textbook
// MINRES, over linalg.h so that every operation runs where the caller's
// operator expects its fields)
struct HostFields {
  std::vector<SpinorField> f;
  explicit HostFields(int n) : f(n) {
    for (auto& s : f) spinor_allocate(s);
  }
};

int host_cg(inv_par& par, SpinorField& x, SpinorField& b) {
  HostFields w(3);
  SpinorField &r = w.f[0], &p = w.f[1], &q = w.f[2];
  spinor_copy(p, x);
  par.mvm(par.mass, q, p);
  spinor_copy(r, b);
  spinor_sub_assign(r, q);
  spinor_copy(p, r);
  const double target = spinor_sqnorm(b) * par.prec;
  double rr = spinor_sqnorm(r);
  int it = 0;
  while (rr >= target) {
    it++;
    par.mvm(par.mass, q, p);
    const double alpha = rr / spinor_product_re(p, q);
    spinor_mulr_add_assign(x, alpha, p);
    spinor_mulr_sub_assign(r, alpha, q);
    const double rr_new = spinor_sqnorm(r);
    if (rr_new < target) {
      rr = rr_new;
      break;
    }
    spinor_mulr_assign(p, rr_new / rr);
    spinor_add_assign(p, r);
    rr = rr_new;
  }
  if (par.log) lprintf("INVERTER", NOTICE, "cg (host): iterations = %d, precision = %1.4e", it, rr / (target / par.prec));
  return it;
}

int host_bicgstab(inv_par& par, SpinorField& x, SpinorField& b) {
  HostFields w(5);
  SpinorField &r = w.f[0], &r0 = w.f[1], &p = w.f[2], &v = w.f[3], &t = w.f[4];
  par.mvm(par.mass, p, x);
  spinor_copy(r, b);
  spinor_sub_assign(r, p);
  spinor_copy(r0, r);
  spinor_copy(p, r);
  const double norm = spinor_sqnorm(b);
  complex rho = spinor_product(r, r0);
  double res = rho.re;
  int it = 0;
  while (res >= norm * par.prec) {
    it++;
    par.mvm(par.mass, v, p);
    const complex alpha = rho / spinor_product(v, r0);
    spinor_mulc_sub_assign(r, alpha, v);
    par.mvm(par.mass, t, r);
    const complex omega = spinor_product(r, t) / complex(spinor_sqnorm(t), 0);
    spinor_mulc_add_assign(x, alpha, p);
    spinor_mulc_add_assign(x, omega, r);
    spinor_mulc_sub_assign(r, omega, t);
    const complex rho_new = spinor_product(r, r0);
    const complex beta = (rho_new / rho) * (alpha / omega);
    spinor_mulc_assign(p, beta);
    spinor_add_assign(p, r);
    spinor_mulc_sub_assign(p, beta * omega, v);
    res = spinor_sqnorm(r);
    rho = rho_new;
  }
  if (par.log) lprintf("INVERTER", NOTICE, "bicgstab (host): iterations = %d, precision = %1.4e", 2 * it, res / norm);
  return 2 * it;
}

int host_mshift(inv_par& par, SpinorField* x, SpinorField& b) {
  const int num = par.num;
  HostFields w(3 + num);
  SpinorField &r = w.f[0], &p = w.f[1], &q = w.f[2];
  std::vector<double> z_old(num, 1.0), z(num, 1.0);
  std::vector<char> done(num, 0);
  spinor_copy(r, b);
  spinor_copy(p, b);
  for (int n = 0; n < num; n++) {
    spinor_allocate(x[n]);
    spinor_zero(x[n]);
    spinor_copy(w.f[3 + n], b);
  }
  const double target = spinor_sqnorm(b) * par.prec;
  double rr = spinor_sqnorm(r), beta = 1.0, alpha = 0.0;
  int it = 0;
  for (;;) {
    it++;
    par.mvm(par.mass, q, p);
    const double beta_old = beta;
    beta = -rr / spinor_product_re(p, q);
    std::vector<double> z_new(num, 0.0);
    for (int n = 0; n < num; n++) {
      if (done[n]) continue;
      z_new[n] = z[n] * z_old[n] * beta_old /
                 (beta * alpha * (z_old[n] - z[n]) + z_old[n] * beta_old * (1.0 - par.shift[n] * beta));
      spinor_mulr_sub_assign(x[n], beta * z_new[n] / z[n], w.f[3 + n]);
    }
    spinor_mulr_add_assign(r, beta, q);
    const double rr_new = spinor_sqnorm(r);
    alpha = rr_new / rr;
    rr = rr_new;
    int open = 0;
    for (int n = 0; n < num; n++) {
      if (done[n]) continue;
      // k_n = z_new r + alpha (z_new / z)^2 k_n
      spinor_mulr_assign(w.f[3 + n], alpha * z_new[n] * z_new[n] / (z[n] * z[n]));
      spinor_mulr_add_assign(w.f[3 + n], z_new[n], r);
      if (rr * z_new[n] * z_new[n] < target) done[n] = 1;
      else open++;
      z_old[n] = z[n];
      z[n] = z_new[n];
    }
    if (!open) break;
    spinor_mulr_assign(p, alpha);
    spinor_add_assign(p, r);
  }
  if (par.log) lprintf("INVERTER", NOTICE, "cg_mshift (host): iterations = %d", it);
  return it;
}

// MINRES for a Hermitian operator (Lanczos with Givens rotations)
int host_minres(inv_par& par, SpinorField& x, SpinorField& b) {
  HostFields w(6);
  SpinorField &v_old = w.f[0], &v = w.f[1], &v_new = w.f[2], &d_old = w.f[3], &d = w.f[4], &d_new = w.f[5];
  par.mvm(par.mass, v_new, x);
  spinor_copy(v, b);
  spinor_sub_assign(v, v_new);  // r0 = b - A x
  const double norm = spinor_sqnorm(b);
  double beta = std::sqrt(spinor_sqnorm(v));
  double eta = beta, res = beta * beta;
  int it = 0;
  if (res < norm * par.prec) return 0;
  spinor_mulr_assign(v, 1.0 / beta);
  spinor_zero(v_old);
  spinor_zero(d_old);
  spinor_zero(d);
  double c_old = 1, c = 1, s_old = 0, s = 0, beta_prev = 0;
  while (res >= norm * par.prec) {
    it++;
    par.mvm(par.mass, v_new, v);
    const double alpha = spinor_product_re(v, v_new);
    spinor_mulr_sub_assign(v_new, alpha, v);
    spinor_mulr_sub_assign(v_new, beta, v_old);
    beta_prev = beta;
    beta = std::sqrt(spinor_sqnorm(v_new));
    // apply the previous two rotations to the new column (beta_prev, alpha, beta)
    const double e = s_old * beta_prev;
    const double f = c * c_old * beta_prev + s * alpha;
    double g = -s * c_old * beta_prev + c * alpha;
    const double h = std::sqrt(g * g + beta * beta);
    c_old = c;
    s_old = s;
    c = g / h;
    s = beta / h;
    // d_new = (v - f d - e d_old) / h
    spinor_copy(d_new, v);
    spinor_mulr_sub_assign(d_new, f, d);
    spinor_mulr_sub_assign(d_new, e, d_old);
    spinor_mulr_assign(d_new, 1.0 / h);
    spinor_mulr_add_assign(x, c * eta, d_new);
    eta = -s * eta;
    res = eta * eta;
    spinor_copy(d_old, d);
    spinor_copy(d, d_new);
    spinor_copy(v_old, v);
    spinor_copy(v, v_new);
    if (beta == 0) break;
    spinor_mulr_assign(v, 1.0 / beta);
  }
  if (par.log) lprintf("INVERTER", NOTICE, "minres (host): iterations = %d, precision = %1.4e", it, res / norm);
  return it;
}

}  // namespace

// ---- dirac.h ----------------------------------------------------------------
void mvm_reset() { yh_mvm_reset(context()); }
int mvm_get() { return yh_mvm_get(context()); }

void Dphi_oe(SpinorField& dptr, SpinorField& sptr) {
  yh_ctx* c = ready();
  yh_field *s = scratch(0), *d = scratch(1);
  up(s, sptr, YH_EVEN);
  check(yh_Dphi_oe(c, d, s), "Dphi_oe");
  down(d, dptr, YH_ODD);
  apply_bcs_on_spinor_field(dptr);
}

void Dphi_eo(SpinorField& dptr, SpinorField& sptr) {
  yh_ctx* c = ready();
  yh_field *s = scratch(0), *d = scratch(1);
  up(s, sptr, YH_ODD);
  check(yh_Dphi_eo(c, d, s), "Dphi_eo");
  down(d, dptr, YH_EVEN);
  apply_bcs_on_spinor_field(dptr);
}

void Dphi_ee(double mass, SpinorField& dptr, SpinorField& sptr) {
  yh_ctx* c = ready();
  yh_field *s = scratch(0), *d = scratch(1);
  up(s, sptr, YH_EVEN);
  check(yh_Dphi_ee(c, mass, d, s), "Dphi_ee");
  down(d, dptr, YH_EVEN);
}

void Dphi_oo(double mass, SpinorField& dptr, SpinorField& sptr) {
  yh_ctx* c = ready();
  yh_field *s = scratch(0), *d = scratch(1);
  up(s, sptr, YH_ODD);
  check(yh_Dphi_oo(c, mass, d, s), "Dphi_oo");
  down(d, dptr, YH_ODD);
}

void Dphi_oo_inv(double mass, SpinorField& dptr, SpinorField& sptr) {
  yh_ctx* c = ready();
  yh_field *s = scratch(0), *d = scratch(1);
  up(s, sptr, YH_ODD);
  check(yh_Dphi_oo_inv(c, mass, d, s), "Dphi_oo_inv");
  down(d, dptr, YH_ODD);
}

// Dphi, Hphi and Hphi_sq write -D_oe sptr_e into the odd half of sptr and,
// for open / SF, zero the boundary slices of both fields (dirac.cpp:738-781,
// bcs.cpp:139-164)
static void schur(int which, double mass, SpinorField& dptr, SpinorField& sptr) {
  yh_ctx* c = ready();
  yh_field *s = scratch(0), *d = scratch(1);
  up(s, sptr, YH_EVEN);
  int rc = which == YH_MVM_DPHI ? yh_Dphi(c, mass, d, s) : which == YH_MVM_HPHI ? yh_Hphi(c, mass, d, s)
                                                                                : yh_Hphi_sq(c, mass, d, s);
  check(rc, "Dphi / Hphi / Hphi_sq");
  down(d, dptr, YH_EVEN);
  down(s, sptr, YH_ODD);
  apply_bcs_on_spinor_field(sptr);
  apply_bcs_on_spinor_field(dptr);
}

void Dphi(double mass, SpinorField& dptr, SpinorField& sptr) { schur(YH_MVM_DPHI, mass, dptr, sptr); }
void Hphi(double mass, SpinorField& dptr, SpinorField& sptr) { schur(YH_MVM_HPHI, mass, dptr, sptr); }
void Hphi_sq(double mass, SpinorField& dptr, SpinorField& sptr) { schur(YH_MVM_HPHI_SQ, mass, dptr, sptr); }

// ---- clover.h ---------------------------------------------------------------
void clover_init(double csw) {
#ifdef CLOVER_TERM
  dev().csw = csw;
  lprintf("CLOVER", INFO, "Coefficient: csw = %1.4f", csw);
#else
  (void)csw;
#endif
}

double get_csw() { return dev().csw; }

// called at the end of represent_gauge_field (repr.cpp:276) after every
// change of the links: the device copy of the gauge field is now stale
void compute_clover_term() { dev().gauge_stale = true; }

// the device factorises D_oo for each mass when it is needed
void compute_ldl_decomp(double) {}

double clover_logdet(double mass) {
  double v = 0.0;
  check(yh_clover_logdet(ready(), mass, &v), "clover_logdet");
  return v;
}

void compute_clover_force(double mass, double coeff) {
  yh_ctx* c = ready();
  check(yh_clover_logdet_force(c, mass, coeff), "compute_clover_force");
}

// ---- inverter.h ---------------------------------------------------------------
int cg_inverter(inv_par& par, SpinorField& dptr, SpinorField& sptr) {
  if (par.mvm == 0) lprintf("INVERTER", CRITICAL, "cg: no MVM operator specified");
  if (!is_device_op(par.mvm)) return host_cg(par, dptr, sptr);
  yh_ctx* c = ready();
  yh_field *x = scratch(0), *b = scratch(1);
  up(x, dptr, YH_EVEN);
  up(b, sptr, YH_EVEN);
  yh_inv_par p = device_par(par);
  int it = 0, upd = 0;
  const char* sloppy = getenv("YAHMC_SLOPPY");
  if (sloppy && par.mvm == Hphi_sq && (!strcmp(sloppy, "f32") || !strcmp(sloppy, "f16"))) {
    p.sloppy_prec = !strcmp(sloppy, "f16") ? YH_PREC_F16 : YH_PREC_F32;
    check(yh_cg_mixed(c, &p, x, b, &it, &upd), "cg_inverter (mixed)");
  } else {
    check(yh_cg(c, &p, x, b, &it), "cg_inverter");
  }
  down(x, dptr, YH_EVEN);
  return it;
}

int bicgstab_inverter(inv_par& par, SpinorField& dptr, SpinorField& sptr) {
  if (par.mvm == 0) lprintf("INVERTER", CRITICAL, "bicgstab: no MVM operator specified");
  if (!is_device_op(par.mvm)) return host_bicgstab(par, dptr, sptr);
  yh_ctx* c = ready();
  yh_field *x = scratch(0), *b = scratch(1);
  up(x, dptr, YH_EVEN);
  up(b, sptr, YH_EVEN);
  yh_inv_par p = device_par(par);
  int it = 0;
  check(yh_bicgstab(c, &p, x, b, &it), "bicgstab_inverter");
  down(x, dptr, YH_EVEN);
  return it;
}

int cg_mshift(inv_par& par, SpinorField* dptr, SpinorField& sptr) {
  if (par.mvm == 0) lprintf("INVERTER", CRITICAL, "cg_mshift: no MVM operator specified");
  if (par.num < 1 || par.num > 48) lprintf("INVERTER", CRITICAL, "cg_mshift: %d shifts", par.num);
  if (par.mvm != Hphi_sq) return host_mshift(par, dptr, sptr);
  yh_ctx* c = ready();
  yh_field* b = scratch(0);
  up(b, sptr, YH_EVEN);
  std::vector<yh_field*> x(par.num);
  for (int n = 0; n < par.num; n++) x[n] = scratch(1 + n);
  yh_inv_par p = device_par(par);
  int it = 0;
  check(yh_cg_mshift(c, &p, par.num, par.shift, x.data(), b, &it), "cg_mshift");
  for (int n = 0; n < par.num; n++) {
    spinor_allocate(dptr[n]);  // callers allocate them (rational.cpp:148-152): a no-op then
    down(x[n], dptr[n], YH_BOTH);
  }
  return it;
}

int minres_inverter(inv_par& par, SpinorField& dptr, SpinorField& sptr) {
  if (par.mvm == 0) lprintf("INVERTER", CRITICAL, "minres: no MVM operator specified");
  if (par.mvm != Hphi && par.mvm != Hphi_sq) return host_minres(par, dptr, sptr);
  yh_ctx* c = ready();
  yh_field *x = scratch(0), *b = scratch(1);
  up(x, dptr, YH_EVEN);
  up(b, sptr, YH_EVEN);
  yh_inv_par p = device_par(par);
  int it = 0;
  check(yh_minres(c, &p, x, b, &it), "minres_inverter");
  down(x, dptr, YH_EVEN);
  return it;
}

int auto_inverter(inv_par& par, SpinorField& dptr, SpinorField& sptr) {
  if (par.num > 0) lprintf("INVERTER", CRITICAL, "auto: no support for multishift inversion");
  if (!is_device_op(par.mvm)) lprintf("INVERTER", CRITICAL, "auto: MVM operator not supported");
  yh_ctx* c = ready();
  yh_field *x = scratch(0), *b = scratch(1);
  up(x, dptr, YH_EVEN);
  up(b, sptr, YH_EVEN);
  yh_inv_par p = device_par(par);
  int it = 0;
  check(yh_auto_inverter(c, &p, x, b, &it), "auto_inverter");
  down(x, dptr, YH_EVEN);
  return it;
}

// ---- forces.h -----------------------------------------------------------------
void begin_fermion_force() {
  yh_ctx* c = ready();
  check(yh_set_momenta(c, &afield_momenta[0].av[0]), "begin_fermion_force");
  check(yh_begin_fermion_force(c), "begin_fermion_force");
  dev().mom_on_device = true;
}

void end_fermion_force(double dt) {
  yh_ctx* c = ready();
  check(yh_end_fermion_force(c, dt), "end_fermion_force");
  check(yh_get_momenta(c, &afield_momenta[0].av[0]), "end_fermion_force");
  dev().mom_on_device = false;
}

void clover_fermion_force(double residue, SpinorField& lhs, SpinorField& rhs) {
  yh_ctx* c = ready();
  yh_field *x = scratch(0), *y = scratch(1);
  up(x, lhs, YH_BOTH);
  up(y, rhs, YH_BOTH);
  check(yh_clover_fermion_force(c, residue, x, y), "clover_fermion_force");
}

void wilson_fermion_force(double dt, double mass, double residue, SpinorField& lhs, SpinorField& rhs) {
  yh_ctx* c = ready();
  yh_field *x = scratch(0), *y = scratch(1);
  up(x, lhs, YH_EVEN);
  up(y, rhs, YH_EVEN);
  momenta_in(c);
  check(yh_wilson_fermion_force(c, dt, mass, residue, x, y), "wilson_fermion_force");
  momenta_out(c);
  // the odd halves now hold D_oo^-1 D_oe X_e (forces.cpp:183-187); the
  // reference's in-place Dphi_oe also zeroed the boundary slices of both
  // fields for open / SF (dirac.cpp:700)
  down(x, lhs, YH_ODD);
  down(y, rhs, YH_ODD);
  apply_bcs_on_spinor_field(lhs);
  apply_bcs_on_spinor_field(rhs);
}

void wilson_fermion_force_q(double dt, double mass, double residue, SpinorField& rhs) {
  yh_ctx* c = ready();
  yh_field* y = scratch(0);
  up(y, rhs, YH_EVEN);
  momenta_in(c);
  check(yh_wilson_fermion_force_q(c, dt, mass, residue, y), "wilson_fermion_force_q");
  momenta_out(c);
  down(y, rhs, YH_ODD);
  apply_bcs_on_spinor_field(rhs);  // as the reference's in-place Hphi and Dphi_oe leave it
}

void clover_logdet_force(double mass, double residue) {
  yh_ctx* c = ready();
  check(yh_clover_logdet_force(c, mass, residue), "clover_logdet_force");
}

void wilson_gauge_force(double dt, double beta) {
  yh_ctx* c = ready();
  momenta_in(c);
  check(yh_wilson_gauge_force(c, dt, beta), "wilson_gauge_force");
  momenta_out(c);
}

void improved_gauge_force(double dt, double beta, double c0, double c1) {
  yh_ctx* c = ready();
  momenta_in(c);
  check(yh_improved_gauge_force(c, dt, beta, c0, c1), "improved_gauge_force");
  momenta_out(c);
}

// ---- rational.h -----------------------------------------------------------
// generated instead of looked up in approx.h: Zolotarev's optimal
// approximation for x^(+-1/2), a Remez exchange otherwise, at the lowest
// order that reaches `prec` on [min, max]
void rational_find(rational_app& r, int n, int d, double prec, double min, double max) {
  yh_rational y;
  check(yh_rational_find(n, d, prec, min, max, &y), "rational_find");
  memcpy(&r, &y, sizeof(r));
}

void rational_calc(rational_app& data, mvm_operator mvm, double mass, double prec, SpinorField& dptr,
                   SpinorField& sptr) {
  if (mvm != Hphi_sq) {
    // black box: multishift on the host over mvm, then the FP64 sum
    inv_par par;
    par.num = data.order;
    par.prec = prec;
    par.shift = data.pole;
    par.mass = mass;
    par.log = 1;
    par.mvm = mvm;
    std::vector<SpinorField> out(data.order);
    cg_mshift(par, out.data(), sptr);
    spinor_mulr(dptr, data.a0, sptr);
    for (int i = 0; i < data.order; i++) spinor_mulr_add_assign(dptr, data.residue[i], out[i]);
    return;
  }
  yh_ctx* c = ready();
  yh_field *x = scratch(0), *b = scratch(1);
  up(b, sptr, YH_EVEN);
  yh_rational y;
  memcpy(&y, &data, sizeof(y));
  yh_inv_par p;
  memset(&p, 0, sizeof(p));
  p.mass = mass;
  p.prec = prec;
  p.mvm = YH_MVM_HPHI_SQ;
  p.log = 1;
  int it = 0;
  check(yh_rational_calc(c, &y, &p, x, b, &it), "rational_calc");
  down(x, dptr, YH_EVEN);
}
