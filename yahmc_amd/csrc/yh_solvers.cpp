// BiCGstab, MINRES, auto_inverter, MRE, multishift CG and the
// mixed-precision reliable-update CG.
//
// bicgstab_inverter : library/bicgstab.cpp:9-97
// minres_inverter   : library/minres.cpp:12-118
// auto_inverter     : library/inverter.cpp:9-55
// mre_*             : library/mre.cpp:15-117 (+ the pivoted LU of solver.cpp:4-95)
// cg_mshift         : library/cg.cpp:11-130
// cg (mixed)        : project-new; FP64 true residual b - A x recomputed with
//                     the FP64 operator at every reliable update, sloppy
//                     FP32 / FP16 operator and vectors in between
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <algorithm>
#include <complex>
#include <deque>
#include <vector>
#include "yh_internal.hpp"

// iteration cap of the learned attempt when the caller gives none (the
// fallback then runs with the caller's cap, 0 = none as in the reference)
static const int LEARNED_MAX_ITER = 20000;

using namespace yh;
typedef std::complex<double> zd;

namespace {

int need(Ctx* c, int prec, int slot, Field** f) {
  int rc;
  *f = tmp_field(c, prec, slot, &rc);
  return rc;
}

int sync(Ctx* c) {
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

double sqnorm(Ctx* c, const Field* a, int* rc) {
  double v[2] = {0, 0};
  *rc = reduce_read(c, 0, a, nullptr, 0, v);
  return v[0];
}

zd dot(Ctx* c, const Field* a, const Field* b, int* rc) {
  double v[2] = {0, 0};
  *rc = reduce_read(c, 1, a, b, 0, v);
  return zd(v[0], v[1]);
}

// the first n reduction slots, device to host, in one synchronisation
int read_red(Ctx* c, int n, const double** out) {
  YH_HIP(hipMemcpyAsync(c->sc_host->red, c->sc->red, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  YH_HIP(hipStreamSynchronize(c->stream));
  *out = c->sc_host->red;
  return YH_OK;
}

double dot_re(Ctx* c, const Field* a, const Field* b, int* rc) {
  double v[2] = {0, 0};
  *rc = reduce_read(c, 2, a, b, 0, v);
  return v[0];
}

}  // namespace

extern "C" {

int yh_bicgstab(yh_ctx* cc, const yh_inv_par* par, yh_field* dptr, const yh_field* sptr, int* iters) {
  Ctx* c = cc;
  if (!c || !par || !dptr || !sptr || !iters) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (dptr->prec != PREC_F64 || sptr->prec != PREC_F64) { set_error("solver fields must be FP64"); return YH_ERR_ARG; }
  if (!(par->prec > 0.0)) { set_error("prec must be positive"); return YH_ERR_ARG; }
  if (par->mvm < YH_MVM_DPHI || par->mvm > YH_MVM_HPHI_SQ) { set_error("bad mvm"); return YH_ERR_ARG; }
  return bicgstab_device(c, par, PREC_F64, dptr, sptr, iters, nullptr);
}

int yh_bicgstab_mixed(yh_ctx* cc, const yh_inv_par* par, yh_field* dptr, const yh_field* sptr, int* iters,
                      int* updates) {
  Ctx* c = cc;
  if (!c || !par || !dptr || !sptr || !iters) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (dptr->prec != PREC_F64 || sptr->prec != PREC_F64) { set_error("solver fields must be FP64"); return YH_ERR_ARG; }
  if (!(par->prec > 0.0)) { set_error("prec must be positive"); return YH_ERR_ARG; }
  if (par->mvm < YH_MVM_DPHI || par->mvm > YH_MVM_HPHI_SQ) { set_error("bad mvm"); return YH_ERR_ARG; }
  return bicgstab_device(c, par, PREC_F32, dptr, sptr, iters, updates);
}

int yh_cg_mshift_batched(yh_ctx* cc, const yh_inv_par* par, int nrhs, int num, const double* shift, yh_field** dptr,
                         const yh_field** sptr, int* iters) {
  Ctx* c = cc;
  if (!c || !par || !shift || !dptr || !sptr || !iters || num < 1 || num > 48 || nrhs < 1 || nrhs > 16) {
    set_error("bad argument");
    return YH_ERR_ARG;
  }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (!(par->prec > 0.0)) { set_error("prec must be positive"); return YH_ERR_ARG; }
  if (par->mvm != YH_MVM_HPHI_SQ) { set_error("batched multishift needs Hphi_sq"); return YH_ERR_ARG; }
  std::vector<Field*> x(nrhs * num);
  std::vector<const Field*> b(nrhs);
  for (int j = 0; j < nrhs; j++) {
    if (!sptr[j] || sptr[j]->prec != PREC_F64) { set_error("fields must be FP64"); return YH_ERR_ARG; }
    b[j] = sptr[j];
  }
  for (int i = 0; i < nrhs * num; i++) {
    if (!dptr[i] || dptr[i]->prec != PREC_F64) { set_error("fields must be FP64"); return YH_ERR_ARG; }
    x[i] = dptr[i];
  }
  return mshift_device(c, par, nrhs, num, shift, x.data(), b.data(), iters);
}

int yh_cg_mshift(yh_ctx* cc, const yh_inv_par* par, int num, const double* shift, yh_field** dptr,
                 const yh_field* sptr, int* iters) {
  Ctx* c = cc;
  if (!c || !par || !shift || !dptr || !sptr || !iters || num < 1 || num > 64) { set_error("bad argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (!(par->prec > 0.0)) { set_error("prec must be positive"); return YH_ERR_ARG; }
  for (int n = 0; n < num; n++)
    if (!dptr[n] || dptr[n]->prec != PREC_F64) { set_error("solution fields must be FP64"); return YH_ERR_ARG; }
  if (par->mvm == YH_MVM_HPHI_SQ && num <= 48) {
    // device-resident recurrences (the usual case: H^2 in RHMC)
    std::vector<Field*> x(dptr, dptr + num);
    const Field* b = sptr;
    return mshift_device(c, par, 1, num, shift, x.data(), &b, iters);
  }
  int rc;
  Field *r, *p, *q;
  if ((rc = need(c, PREC_F64, 2, &r)) || (rc = need(c, PREC_F64, 3, &p)) || (rc = need(c, PREC_F64, 4, &q))) return rc;
  std::vector<Field*> k(num);
  for (int n = 0; n < num; n++)
    if ((rc = field_alloc(c, PREC_F64, &k[n]))) {
      for (int m = 0; m < n; m++) field_free(k[m]);
      return rc;
    }
  auto cleanup = [&]() { for (auto f : k) field_free(f); };
  const int recon = default_recon(c, PREC_F64);
  std::vector<double> c1(num, 1.0), c2(num, 1.0), c3(num, 0.0);
  std::vector<int> cflag(num, 0);
  if ((rc = blas_copy(c, r, sptr, 0)) || (rc = blas_copy(c, p, sptr, 0))) { cleanup(); return rc; }
  for (int n = 0; n < num; n++) {
    if ((rc = field_zero(c, dptr[n], YH_BOTH)) || (rc = blas_copy(c, k[n], sptr, 0))) { cleanup(); return rc; }
  }
  double beta = 1.0, alpha = 0.0, beta_old, atmp, btmp = 0.0;
  double norm = sqnorm(c, sptr, &rc);
  if (rc) { cleanup(); return rc; }
  int it = 0;
  while (++it) {
    if ((rc = apply_op(c, PREC_F64, recon, par->mvm, par->mass, q, p, nullptr, -1, nullptr))) { cleanup(); return rc; }
    atmp = sqnorm(c, r, &rc);
    if (!rc) btmp = dot_re(c, p, q, &rc);
    if (rc) { cleanup(); return rc; }
    beta_old = beta;
    beta = -(atmp / btmp);
    for (int n = 0; n < num; n++) {
      if (cflag[n]) continue;
      c3[n] = c2[n] * c1[n] * beta_old;
      c3[n] /= (beta * alpha * (c1[n] - c2[n])) + (c1[n] * beta_old * (1 - shift[n] * beta));
      double bt = beta * (c3[n] / c2[n]);
      if ((rc = blas_axpby(c, dptr[n], -bt, k[n], 1.0, 0))) { cleanup(); return rc; }
    }
    if ((rc = blas_axpby(c, r, beta, q, 1.0, 0))) { cleanup(); return rc; }
    btmp = sqnorm(c, r, &rc);
    if (rc) { cleanup(); return rc; }
    if (!(btmp == btmp)) { cleanup(); set_error("cg_mshift: non-finite residual"); return YH_ERR_NONFINITE; }
    alpha = btmp / atmp;
    int notconverged = 0;
    for (int n = 0; n < num; n++) {
      if (cflag[n]) continue;
      double at = alpha * (c3[n] * c3[n]) / (c2[n] * c2[n]);
      if ((rc = blas_axpby(c, k[n], c3[n], r, at, 0))) { cleanup(); return rc; }
      if (btmp * c3[n] * c3[n] < norm * par->prec) cflag[n] = 1;
      else notconverged++;
      c1[n] = c2[n];
      c2[n] = c3[n];
    }
    if (notconverged == 0) break;
    if (par->max_iter > 0 && it >= par->max_iter) {
      cleanup();
      *iters = it;
      set_error("cg_mshift: max_iter reached");
      return YH_ERR_NOCONV;
    }
    if ((rc = blas_axpby(c, p, 1.0, r, alpha, 0))) { cleanup(); return rc; }
  }
  rc = sync(c);
  cleanup();
  *iters = it;
  if (par->log) fprintf(stderr, "[INVERTER][NOTICE] cg_mshift: iterations = %d, precision = %1.4e\n", it, btmp / norm);
  return rc;
}

int yh_cg_mixed(yh_ctx* cc, const yh_inv_par* par, yh_field* dptr, const yh_field* sptr, int* iters, int* updates) {
  Ctx* c = cc;
  if (!c || !par || !dptr || !sptr || !iters) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (dptr->prec != PREC_F64 || sptr->prec != PREC_F64) { set_error("solver fields must be FP64"); return YH_ERR_ARG; }
  if (par->mvm != YH_MVM_HPHI_SQ) { set_error("mixed CG needs the Hermitian operator Hphi_sq"); return YH_ERR_ARG; }
  if (par->sloppy_prec != PREC_F32 && par->sloppy_prec != PREC_F16) { set_error("sloppy precision must be F32 or F16"); return YH_ERR_ARG; }
  if (!(par->prec > 0.0)) { set_error("prec must be positive"); return YH_ERR_ARG; }
  const int S = par->sloppy_prec;
  // default cap: 200 iterations (at 32^4, 0.229 against 0.232 s for 16-bit and
  // 0.358 against 0.363 s for FP32 sloppy operators, against a cap of 50); 50
  // for FP32 spinors with 16-bit links, which need the frequent updates (688
  // iterations against 787 at 200)
  const int K = par->reliable_every > 0 ? par->reliable_every : (S == PREC_F32 && c->half_links ? 50 : 200);
  int rc;
  Field *r64, *t64, *xs, *rs, *p, *q;
  if ((rc = need(c, PREC_F64, 2, &r64)) || (rc = need(c, PREC_F64, 3, &t64))) return rc;
  if ((rc = need(c, PREC_F32, 2, &xs)) || (rc = need(c, PREC_F32, 3, &rs))) return rc;
  if (S == PREC_F32) {
    if ((rc = need(c, PREC_F32, 4, &p)) || (rc = need(c, PREC_F32, 5, &q))) return rc;
  } else {
    if ((rc = need(c, PREC_F16, 2, &p)) || (rc = need(c, PREC_F16, 3, &q))) return rc;
  }
  const int recon64 = default_recon(c, PREC_F64);
  const int recons = default_recon(c, S);
  Field* x = dptr;
  double norm_b = sqnorm(c, sptr, &rc);
  if (rc) return rc;
  const double target = norm_b * par->prec;
  auto true_residual = [&](double* rr) -> int {
    int e;
    if ((e = apply_op(c, PREC_F64, recon64, YH_MVM_HPHI_SQ, par->mass, t64, x, nullptr, -1, nullptr))) return e;
    if ((e = blas_copy(c, r64, sptr, 0))) return e;
    if ((e = blas_axpby(c, r64, -1.0, t64, 1.0, 0))) return e;
    *rr = sqnorm(c, r64, &e);
    return e;
  };
  double rr;
  if ((rc = true_residual(&rr))) return rc;
  *iters = 0;
  if (rr < target) {
    if (updates) *updates = 0;
    return YH_OK;
  }
  if ((rc = field_convert(c, rs, 0, r64, 0)) || (rc = field_convert(c, p, 0, r64, 0)) || (rc = field_zero(c, xs, YH_EVEN)))
    return rc;
  // Reliable updates (Sleijpen & van der Vorst; Clark et al. for lattice
  // QCD): the sloppy recurrence runs on the device until |r| has dropped by
  // `delta` since the last update (flag raised by the CG update kernel) or
  // `reliable_every` iterations have passed; then x += xs and the residual
  // is recomputed in FP64.  The search direction p is kept.
  //
  // The update itself runs on the device (ru_begin, the FP64 operator,
  // ru_end, whose last workgroup decides whether the solve is finished), so
  // the host never waits inside an update.  The host queues a fixed plan
  // (segments of K sloppy iterations in chunks of at most 16, each followed
  // by an update) up to two items ahead of the one whose scalar snapshot it
  // reads; items queued after the solve finished, or sloppy items queued
  // after the recurrence stopped on its own flag, return at once.  An early
  // stop moves the next update forward.  Once a segment has ended on the
  // delta flag, later segments are planned to that length: the chunk that
  // reaches it ends there, and after it the host queues two iterations at a
  // time, so that a stop leaves at most a few queued iterations idle instead
  // of a whole chunk (a light system that reaches delta in ~16 iterations
  // otherwise queued twice the iterations it ran).
  const double delta = (par->delta > 0.0 && par->delta < 1.0) ? par->delta : 0.1;
  SolverScalars* h = c->sc_host;
  memset(h, 0, sizeof(SolverScalars));
  h->rr = rr;
  h->maxrr = rr;
  h->best_rr = rr;
  h->delta2 = delta * delta;
  h->target = target;
  h->norm_b = norm_b;
  if ((rc = write_scalars(c))) return rc;
  const int chunk = 16, depth = 2;
  struct Item {
    bool ru;  // reliable update, else n sloppy iterations
    int n;
    int slot;
  };
  std::deque<Item> inflight;
  int slot = 0;
  int seg = 0;          // sloppy iterations planned in the current segment
  long planned = 0;     // sloppy iterations planned in total
  bool ru_next = false; // the recurrence stopped early: update next
  bool plan_done = false;
  int est = 0;          // length of the last segment that ended on the delta flag (0: none yet)
  int seg_start = 0;    // iterations done when the current segment started
  SolverScalars v;
  memset(&v, 0, sizeof(v));
  auto enqueue = [&](Item it) -> int {
    int e;
    if (it.ru) {
      if ((e = ru_begin(c, x, xs, p, rs))) return e;
      if ((e = apply_op(c, PREC_F64, recon64, YH_MVM_HPHI_SQ, par->mass, t64, x, nullptr, -1, &c->sc->finished))) return e;
      if ((e = ru_end(c, sptr, t64, rs))) return e;
    } else {
      for (int i = 0; i < it.n; i++) {
        if ((e = apply_op(c, S, recons, YH_MVM_HPHI_SQ, par->mass, q, p, nullptr, 0, &c->sc->converged))) return e;
        if ((e = cg_update_xr(c, xs, rs, p, q, 0, PREC_F32))) return e;
        if ((e = cg_update_p(c, p, rs, 0, PREC_F32))) return e;
      }
    }
    it.slot = slot;
    YH_HIP(hipMemcpyAsync(&c->sc_ring[slot], c->sc, sizeof(SolverScalars), hipMemcpyDeviceToHost, c->stream));
    YH_HIP(hipEventRecord(c->ev_ring[slot], c->stream));
    inflight.push_back(it);
    slot = (slot + 1) % Ctx::SC_RING;
    return YH_OK;
  };
  int err = YH_OK;
  while (true) {
    while (!plan_done && err == YH_OK && (int)inflight.size() < depth) {
      Item it = {false, 0, 0};
      const bool capped = par->max_iter > 0 && planned >= par->max_iter;
      if (ru_next || seg >= K || capped) {
        it.ru = true;
        seg = 0;
        ru_next = false;
        if (capped) plan_done = true;  // the last update at the iteration cap
      } else {
        it.n = std::min(chunk, K - seg);
        if (est > 0) it.n = std::min(it.n, std::max(2, est - seg));
        if (par->max_iter > 0) it.n = (int)std::min<long>(it.n, par->max_iter - planned);
        seg += it.n;
        planned += it.n;
      }
      if ((rc = enqueue(it))) return rc;
    }
    if (inflight.empty()) break;
    const Item it = inflight.front();
    inflight.pop_front();
    YH_HIP(hipEventSynchronize(c->ev_ring[it.slot]));
    v = c->sc_ring[it.slot];
    if (par->log)
      fprintf(stderr, "[INVERTER][DEBUG] cg_mixed: %s iters=%d rr=%.6e true_rr=%.6e updates=%d stop=%d reason=%d finished=%d\n",
              it.ru ? "update" : "sloppy", v.iters, v.rr, v.rr_true, v.nupd, v.converged, v.reason, v.finished);
    if (err != YH_OK || v.finished) {
      plan_done = true;  // what is still queued returns at once; drain it
      continue;
    }
    if (!it.ru && !(v.rr == v.rr)) {
      set_error("cg_mixed: non-finite residual");
      err = YH_ERR_NONFINITE;
      plan_done = true;
      continue;
    }
    if (it.ru) seg_start = v.iters;
    if (!it.ru && v.converged) {
      // the recurrence stopped inside this item: the sloppy items queued
      // after it did nothing, and the update comes next unless queued
      planned = v.iters;
      if (v.reason == 2 && v.iters > seg_start) est = v.iters - seg_start;
      if (inflight.empty() || !inflight.front().ru) ru_next = true;
    }
  }
  YH_HIP(hipStreamSynchronize(c->stream));
  *iters = v.iters;
  if (updates) *updates = v.nupd;
  // the operator applications of the sloppy iterations and of the updates
  // that ran (their launches carry stop flags and are not counted at launch)
  c->mvm_counter += (long)(v.iters + v.nupd) * hops_per_apply(YH_MVM_HPHI_SQ);
  if (err != YH_OK) return err;
  if (v.finished == 2) { set_error("cg_mixed: non-finite true residual"); return YH_ERR_NONFINITE; }
  if (v.finished == 3) { set_error("cg_mixed: true residual stalled"); return YH_ERR_NOCONV; }
  if (v.finished != 1) { set_error("cg_mixed: max_iter reached"); return YH_ERR_NOCONV; }
  if (par->log)
    fprintf(stderr, "[INVERTER][NOTICE] cg_mixed: iterations = %d, updates = %d, precision = %1.4e\n", v.iters, v.nupd,
            v.rr_true / norm_b);
  return YH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
extern "C" {

int yh_minres(yh_ctx* cc, const yh_inv_par* par, yh_field* dptr, const yh_field* sptr, int* iters) {
  Ctx* c = cc;
  if (!c || !par || !dptr || !sptr || !iters) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (dptr->prec != PREC_F64 || sptr->prec != PREC_F64) { set_error("solver fields must be FP64"); return YH_ERR_ARG; }
  if (!(par->prec > 0.0)) { set_error("prec must be positive"); return YH_ERR_ARG; }
  if (par->mvm == YH_MVM_DPHI) { set_error("minres needs a Hermitian operator (Hphi or Hphi_sq)"); return YH_ERR_ARG; }
  int rc;
  Field *v0, *v1, *w0, *w1, *t;
  if ((rc = need(c, PREC_F64, 2, &v0)) || (rc = need(c, PREC_F64, 3, &v1)) || (rc = need(c, PREC_F64, 4, &w0)) ||
      (rc = need(c, PREC_F64, 5, &w1)) || (rc = need(c, PREC_F64, 6, &t)))
    return rc;
  const int recon = default_recon(c, PREC_F64);
  Field* x = dptr;
  // v1 = b - A x ; v0 = w0 = w1 = 0
  if ((rc = apply_op(c, PREC_F64, recon, par->mvm, par->mass, t, x, nullptr, -1, nullptr))) return rc;
  if ((rc = blas_copy(c, v1, sptr, 0)) || (rc = blas_axpby(c, v1, -1.0, t, 1.0, 0))) return rc;
  if ((rc = field_zero(c, v0, YH_EVEN)) || (rc = field_zero(c, w0, YH_EVEN)) || (rc = field_zero(c, w1, YH_EVEN)))
    return rc;
  const double norm = sqnorm(c, sptr, &rc);
  if (rc) return rc;
  double residue = sqnorm(c, v1, &rc);
  if (rc) return rc;
  double beta0 = sqrt(residue), eta = beta0;
  double g0 = 1, g1 = 1, s0 = 0, s1 = 0;
  int it = 0;
  if (residue < norm * par->prec) it = -1;
  while (++it) {
    // Lanczos step: v1 <- v1 / beta0, t = A v1 - alpha v1 - beta0 v0
    if ((rc = blas_axpby(c, v1, 0.0, v1, 1.0 / beta0, 0))) return rc;
    if ((rc = apply_op(c, PREC_F64, recon, par->mvm, par->mass, t, v1, nullptr, -1, nullptr))) return rc;
    double alpha = dot_re(c, v1, t, &rc);
    if (rc) return rc;
    if ((rc = blas_axpby(c, t, -alpha, v1, 1.0, 0)) || (rc = blas_axpby(c, t, -beta0, v0, 1.0, 0))) return rc;
    std::swap(v0, v1);  // (v0, v1, t) <- (v1, t, v0)
    std::swap(v1, t);
    double beta1 = sqrt(sqnorm(c, v1, &rc));
    if (rc) return rc;
    // Givens rotation
    double delta = g1 * alpha - g0 * s1 * beta0;
    double rho1 = sqrt(delta * delta + beta1 * beta1);
    double rho2 = s1 * alpha + g0 * g1 * beta0;
    double rho3 = s0 * beta0;
    g0 = g1;
    g1 = delta / rho1;
    s0 = s1;
    s1 = beta1 / rho1;
    // t = (v0 - rho3 w0 - rho2 w1) / rho1 ; (w0, w1, t) <- (w1, t, w0)
    if ((rc = blas_axpby(c, t, 1.0 / rho1, v0, 0.0, 0)) || (rc = blas_axpby(c, t, -rho3 / rho1, w0, 1.0, 0)) ||
        (rc = blas_axpby(c, t, -rho2 / rho1, w1, 1.0, 0)))
      return rc;
    std::swap(w0, w1);
    std::swap(w1, t);
    if ((rc = blas_axpby(c, x, g1 * eta, w1, 1.0, 0))) return rc;
    eta = -s1 * eta;
    beta0 = beta1;
    residue = eta * eta;
    if (!(residue == residue)) { set_error("minres: non-finite residual"); return YH_ERR_NONFINITE; }
    if (residue < norm * par->prec) break;
    if (par->max_iter > 0 && it >= par->max_iter) {
      *iters = it;
      set_error("minres: max_iter reached");
      return YH_ERR_NOCONV;
    }
  }
  *iters = it;
  if (par->log) fprintf(stderr, "[INVERTER][NOTICE] minres: iterations = %d, precision = %1.4e\n", it, residue / norm);
  return sync(c);
}

// auto_inverter: Hphi_sq -> two BiCGstab solves on Dphi with gamma5,
// Hphi -> one, Dphi -> one.  The intermediate field persists between calls
// and is the first solve's initial guess, as the reference's static
// scratch field is.
int yh_auto_inverter(yh_ctx* cc, const yh_inv_par* par, yh_field* dptr, yh_field* sptr, int* iters) {
  Ctx* c = cc;
  if (!c || !par || !dptr || !sptr || !iters) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  int rc;
  if (!c->tmp[4]) {
    Field* f;
    if ((rc = field_alloc(c, PREC_F64, &f))) return rc;
    c->tmp[4] = f;
  }
  Field* tmp = c->tmp[4];
  yh_inv_par p = *par;
  p.mvm = YH_MVM_DPHI;
  int it = 0, n = 0;
  if (par->mvm == YH_MVM_HPHI_SQ) {
    if ((rc = blas_g5(c, sptr, 0))) return rc;
    rc = yh_bicgstab(cc, &p, static_cast<yh_field*>(tmp), sptr, &n);
    it += n;
    if (!rc) rc = blas_g5(c, tmp, 0);
    if (!rc) rc = yh_bicgstab(cc, &p, dptr, static_cast<yh_field*>(tmp), &n);
    it += n;
    int rc2 = blas_g5(c, sptr, 0);
    if (rc) return rc;
    if (rc2) return rc2;
  } else if (par->mvm == YH_MVM_HPHI) {
    if ((rc = blas_g5(c, sptr, 0))) return rc;
    rc = yh_bicgstab(cc, &p, dptr, sptr, &n);
    it += n;
    int rc2 = blas_g5(c, sptr, 0);
    if (rc) return rc;
    if (rc2) return rc2;
  } else if (par->mvm == YH_MVM_DPHI) {
    if ((rc = yh_bicgstab(cc, &p, dptr, sptr, &n))) return rc;
    it += n;
  } else {
    set_error("auto_inverter: operator not supported");
    return YH_ERR_ARG;
  }
  *iters = it;
  return sync(c);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// MRE: minimal residual extrapolation of the initial guess from the last
// `max` solutions (Brower, Ivanenko, Levkova, Orginos, hep-lat/9509012):
// x0 = sum_k x_k s_k with the Galerkin condition s_j^+ (b - M x0) = 0 on
// the span of the stored solutions.
// ---------------------------------------------------------------------------
struct yh_mre {
  Ctx* c = nullptr;
  int max = 0, num = 0, mvm = YH_MVM_DPHI;
  double mass = 0;
  std::vector<Field*> s;  // s[0] is the most recent solution
  Field* out = nullptr;
};

namespace {

// Gaussian elimination with partial pivoting on a small dense complex system
int small_solve(int n, std::vector<zd> A, std::vector<zd> b, std::vector<zd>& x) {
  for (int k = 0; k < n; k++) {
    int piv = k;
    for (int i = k + 1; i < n; i++)
      if (std::abs(A[i * n + k]) > std::abs(A[piv * n + k])) piv = i;
    if (std::abs(A[piv * n + k]) < 1e-14) { set_error("mre: singular Galerkin matrix"); return YH_ERR_NOCONV; }
    if (piv != k) {
      for (int j = 0; j < n; j++) std::swap(A[k * n + j], A[piv * n + j]);
      std::swap(b[k], b[piv]);
    }
    for (int i = k + 1; i < n; i++) {
      zd f = A[i * n + k] / A[k * n + k];
      for (int j = k; j < n; j++) A[i * n + j] -= f * A[k * n + j];
      b[i] -= f * b[k];
    }
  }
  x.assign(n, zd(0, 0));
  for (int i = n - 1; i >= 0; i--) {
    zd v = b[i];
    for (int j = i + 1; j < n; j++) v -= A[i * n + j] * x[j];
    x[i] = v / A[i * n + i];
  }
  return YH_OK;
}

}  // namespace

extern "C" {

int yh_mre_create(yh_ctx* cc, int max, int mvm, double mass, yh_mre** out) {
  Ctx* c = cc;
  if (!c || !out || max < 0 || mvm < YH_MVM_DPHI || mvm > YH_MVM_HPHI_SQ) { set_error("bad argument"); return YH_ERR_ARG; }
  yh_mre* m = new yh_mre();
  m->c = c;
  m->max = std::min(max, 15);  // MRE_MAX (include/mre.h:8)
  m->mvm = mvm;
  m->mass = mass;
  int rc = YH_OK;
  for (int i = 0; i < m->max && !rc; i++) {
    Field* f = nullptr;
    rc = field_alloc(c, PREC_F64, &f);
    if (!rc) m->s.push_back(f);
  }
  if (!rc && m->max > 0) rc = field_alloc(c, PREC_F64, &m->out);
  if (rc) {
    yh_mre_destroy(m);
    return rc;
  }
  *out = m;
  return YH_OK;
}

int yh_mre_destroy(yh_mre* m) {
  if (!m) return YH_OK;
  for (Field* f : m->s) field_free(f);
  field_free(m->out);
  delete m;
  return YH_OK;
}

int yh_mre_reset(yh_mre* m) {
  if (!m) { set_error("null argument"); return YH_ERR_ARG; }
  m->num = 0;
  return YH_OK;
}

// mre_store: the oldest slot becomes the newest and receives x (even half)
int yh_mre_store(yh_mre* m, const yh_field* x) {
  if (!m || !x) { set_error("null argument"); return YH_ERR_ARG; }
  if (m->max == 0) return YH_OK;
  std::rotate(m->s.begin(), m->s.end() - 1, m->s.end());
  int rc = blas_copy(m->c, m->s[0], x, 0);
  if (rc) return rc;
  m->num++;
  return sync(m->c);
}

// mre_guess: dst = extrapolated guess for M dst = b (dst = 0 without history)
int yh_mre_guess(yh_mre* m, yh_field* dst, const yh_field* b) {
  if (!m || !dst || !b) { set_error("null argument"); return YH_ERR_ARG; }
  Ctx* c = m->c;
  if (m->max == 0) return YH_OK;
  int rc;
  const int n = std::min(m->num, m->max);
  if ((rc = field_zero(c, dst, YH_EVEN))) return rc;
  if (n == 0) return sync(c);
  // Gram-Schmidt on the stored solutions (mre.cpp:15-33); the span, and so
  // the Galerkin solution, is what matters.  Each step queues |s_i|^2 and
  // the overlaps s_i^+ s_j (j > i) into the reduction slots and reads them
  // back with one synchronisation.
  const double* red;
  for (int i = 0; i < n; i++) {
    if ((rc = blas_sqnorm(c, m->s[i], 0, 0))) return rc;
    for (int j = i + 1; j < n; j++)
      if ((rc = blas_dot(c, m->s[i], m->s[j], 0, 2 * j))) return rc;  // sum s_i conj(s_j)
    if ((rc = read_red(c, 2 * n, &red))) return rc;
    const double inv = 1.0 / sqrt(red[0]);
    std::vector<zd> r(n);
    for (int j = i + 1; j < n; j++) r[j] = zd(red[2 * j], red[2 * j + 1]) * inv;
    if ((rc = blas_cscale(c, m->s[i], inv, 0.0, 0))) return rc;
    // subtract the projection of s_j on the unit s_i: (s_i^+ s_j) s_i = conj(r) s_i
    for (int j = i + 1; j < n; j++)
      if ((rc = blas_caxpy(c, m->s[j], -r[j].real(), r[j].imag(), m->s[i], 0))) return rc;
  }
  // Galerkin matrix A_ji = s_j^+ M s_i and right-hand side s_i^+ b, one
  // synchronisation per column; for the hermitian H and H^2 the entries
  // above the diagonal are the conjugates of those below
  const bool herm = m->mvm == YH_MVM_HPHI || m->mvm == YH_MVM_HPHI_SQ;
  const int recon = default_recon(c, PREC_F64);
  std::vector<zd> A(n * n), rhs(n), xs;
  for (int i = 0; i < n; i++) {
    if ((rc = apply_op(c, PREC_F64, recon, m->mvm, m->mass, m->out, m->s[i], nullptr, -1, nullptr))) return rc;
    const int j0 = herm ? i : 0;
    for (int j = j0; j < n; j++)
      if ((rc = blas_dot(c, m->out, m->s[j], 0, 2 * j))) return rc;  // sum (M s_i) conj(s_j)
    if ((rc = blas_dot(c, b, m->s[i], 0, 2 * n))) return rc;         // s_i^+ b
    if ((rc = read_red(c, 2 * n + 2, &red))) return rc;
    for (int j = j0; j < n; j++) {
      A[j * n + i] = zd(red[2 * j], red[2 * j + 1]);
      if (herm && j > i) A[i * n + j] = std::conj(A[j * n + i]);
    }
    rhs[i] = zd(red[2 * n], red[2 * n + 1]);
  }
  if ((rc = small_solve(n, A, rhs, xs))) return rc;
  for (int k = 0; k < n; k++)
    if ((rc = blas_caxpy(c, dst, xs[k].real(), xs[k].imag(), m->s[k], 0))) return rc;
  return sync(c);
}


// Learned-preconditioner solve with its fallback: flexible CG (or, with
// YH_TUNE_LEARNED_GCR, flexible GCR keeping 8 directions) preconditioned by
// the network, started from the MRE guess; if it stops without converging
// (iteration cap or a non-finite scalar), the plain mixed-precision CG
// restarts from the MRE guess.  The converged solution joins the MRE history.
int yh_cg_learned(yh_ctx* cc, const yh_inv_par* par, yh_net* net, yh_mre* mre, yh_field* dptr, const yh_field* sptr,
                  int* iters, int* updates, int* fellback) {
  Ctx* c = cc;
  if (!c || !par || !dptr || !sptr || !iters) { set_error("null argument"); return YH_ERR_ARG; }
  if (mre && mre->c != c) { set_error("MRE history belongs to another context"); return YH_ERR_ARG; }
  int rc, it = 0, upd = 0;
  if (fellback) *fellback = 0;
  if (mre && (rc = yh_mre_guess(mre, dptr, sptr))) return rc;
  rc = YH_ERR_NOCONV;
  if (net) {
    yh_inv_par p = *par;
    if (p.max_iter <= 0) p.max_iter = LEARNED_MAX_ITER;
    rc = c->learned_gcr ? yh_gcr_net(cc, &p, net, 8, dptr, sptr, &it, &upd)
                        : yh_cg_net(cc, &p, net, dptr, sptr, &it, &upd);
    if (rc != YH_OK && rc != YH_ERR_NOCONV && rc != YH_ERR_NONFINITE) return rc;
  }
  if (rc != YH_OK) {
    if (fellback) *fellback = 1;
    const int it0 = it, upd0 = upd;
    // restart from the guess: the abandoned iterate may be non-finite
    if (mre && mre->max > 0) {
      if ((rc = yh_mre_guess(mre, dptr, sptr))) return rc;
    } else if ((rc = field_zero(c, dptr, YH_EVEN))) {
      return rc;
    }
    yh_inv_par p = *par;
    if (p.sloppy_prec != PREC_F32 && p.sloppy_prec != PREC_F16) p.sloppy_prec = PREC_F32;
    rc = yh_cg_mixed(cc, &p, dptr, sptr, &it, &upd);
    it += it0;
    upd += upd0;
  }
  *iters = it;
  if (updates) *updates = upd;
  if (rc) return rc;
  if (mre) rc = yh_mre_store(mre, dptr);
  return rc;
}

}  // extern "C"
