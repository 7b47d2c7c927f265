// BiCGstab with every scalar on the device (library/bicgstab.cpp:9-97), and
// its mixed-precision form with FP64 reliable updates (project-new).
//
// One iteration of the reference recurrence
//   v = A p;  alpha = rho0 / <v, r0>;  s = r - alpha v;  t = A s;
//   omega = <s, t> / |t|^2;  x += alpha p + omega s;  r = s - omega t;
//   rho1 = <r, r0>;  beta = (rho1 / rho0) (alpha / omega);
//   p = r + beta (p - omega v);  stop when |r|^2 < prec |b|^2
// with <a, b> = sum a conj(b) (spinor_product, library/spinor.cpp:553-570)
// runs as two operator applications and five kernels; each scalar is
// computed by the last workgroup of the kernel that reduces its inputs, so
// the host only reads the scalars once per chunk of iterations.  Kernels
// of iterations queued past convergence return at once.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <algorithm>
#include "yh_internal.hpp"
#include "yh_reduce.hpp"

namespace yh {

// complex scalars kept in SolverScalars::red
enum { BI_RHO = 8, BI_ALPHA = 10, BI_OMEGA = 12, BI_BETA = 14 };

template <int P, int D>
using BR = typename SpinorStore<P, D>::real;

__device__ __forceinline__ Cx<double> cdiv(Cx<double> a, Cx<double> b) {
  const double n = b.re * b.re + b.im * b.im;
  return Cx<double>((a.re * b.re + a.im * b.im) / n, (a.im * b.re - a.re * b.im) / n);
}

// alpha = rho0 / <v, r0>
template <int P, int D>
__global__ void __launch_bounds__(256) k_bi_alpha(const void* v, const void* r0, SolverScalars* sc, double* partials,
                                                  unsigned* ticket, Geom g) {
  typedef BR<P, D> R;
  if (sc->converged) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  double w[2] = {0.0, 0.0};
  if (k < g.VH) {
    Cx<R> A[4][D], B[4][D];
    load_spinor<P, D, R>(A, v, nullptr, k, g.VHpad);
    load_spinor<P, D, R>(B, r0, nullptr, k, g.VHpad);
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < D; i++) {
        w[0] += (double)A[s][i].re * B[s][i].re + (double)A[s][i].im * B[s][i].im;
        w[1] += (double)A[s][i].im * B[s][i].re - (double)A[s][i].re * B[s][i].im;
      }
  }
  double res[2];
  if (block_reduce_finish<2, 256>(w, partials, gridDim.x, blockIdx.x, ticket, res) && threadIdx.x == 0) {
    const Cx<double> a = cdiv(Cx<double>(sc->red[BI_RHO], sc->red[BI_RHO + 1]), Cx<double>(res[0], res[1]));
    sc->red[BI_ALPHA] = a.re;
    sc->red[BI_ALPHA + 1] = a.im;
  }
}

// s = r - alpha v (in place in r)
template <int P, int D>
__global__ void __launch_bounds__(256) k_bi_s(void* r, const void* v, const SolverScalars* sc, Geom g) {
  typedef BR<P, D> R;
  if (sc->converged) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= g.VH) return;
  const Cx<R> a((R)sc->red[BI_ALPHA], (R)sc->red[BI_ALPHA + 1]);
  Cx<R> Rv[4][D], V[4][D];
  load_spinor<P, D, R>(Rv, r, nullptr, k, g.VHpad);
  load_spinor<P, D, R>(V, v, nullptr, k, g.VHpad);
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int i = 0; i < D; i++) Rv[s][i] = Rv[s][i] - a * V[s][i];
  store_spinor<P, D, R>(r, nullptr, k, g.VHpad, Rv);
}

// omega = <s, t> / |t|^2
template <int P, int D>
__global__ void __launch_bounds__(256) k_bi_omega(const void* s_, const void* t_, SolverScalars* sc, double* partials,
                                                  unsigned* ticket, Geom g) {
  typedef BR<P, D> R;
  if (sc->converged) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  double w[3] = {0.0, 0.0, 0.0};
  if (k < g.VH) {
    Cx<R> S[4][D], T[4][D];
    load_spinor<P, D, R>(S, s_, nullptr, k, g.VHpad);
    load_spinor<P, D, R>(T, t_, nullptr, k, g.VHpad);
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < D; i++) {
        w[0] += (double)S[s][i].re * T[s][i].re + (double)S[s][i].im * T[s][i].im;
        w[1] += (double)S[s][i].im * T[s][i].re - (double)S[s][i].re * T[s][i].im;
        w[2] += (double)T[s][i].re * T[s][i].re + (double)T[s][i].im * T[s][i].im;
      }
  }
  double res[3];
  if (block_reduce_finish<3, 256>(w, partials, gridDim.x, blockIdx.x, ticket, res) && threadIdx.x == 0) {
    sc->red[BI_OMEGA] = res[0] / res[2];
    sc->red[BI_OMEGA + 1] = res[1] / res[2];
  }
}

// x += alpha p + omega s;  r = s - omega t;  rho1 = <r, r0>, |r|^2;
// beta = (rho1 / rho0)(alpha / omega); convergence and reliable-update flags
template <int P, int D>
__global__ void __launch_bounds__(256) k_bi_xr(void* x, void* r, const void* p, const void* t, const void* r0,
                                               SolverScalars* sc, double* partials, unsigned* ticket, Geom g) {
  typedef BR<P, D> R;
  if (sc->converged) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  const Cx<R> a((R)sc->red[BI_ALPHA], (R)sc->red[BI_ALPHA + 1]);
  const Cx<R> om((R)sc->red[BI_OMEGA], (R)sc->red[BI_OMEGA + 1]);
  double w[3] = {0.0, 0.0, 0.0};
  if (k < g.VH) {
    Cx<R> X[4][D], S[4][D], Pv[4][D], T[4][D], Z[4][D];
    load_spinor<P, D, R>(Pv, p, nullptr, k, g.VHpad);
    load_spinor<P, D, R>(S, r, nullptr, k, g.VHpad);
    load_spinor<P, D, R>(T, t, nullptr, k, g.VHpad);
    load_spinor<P, D, R>(Z, r0, nullptr, k, g.VHpad);
    load_spinor<P, D, R>(X, x, nullptr, k, g.VHpad);
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < D; i++) {
        X[s][i] = X[s][i] + a * Pv[s][i] + om * S[s][i];
        S[s][i] = S[s][i] - om * T[s][i];
        w[0] += (double)S[s][i].re * Z[s][i].re + (double)S[s][i].im * Z[s][i].im;
        w[1] += (double)S[s][i].im * Z[s][i].re - (double)S[s][i].re * Z[s][i].im;
        w[2] += (double)S[s][i].re * S[s][i].re + (double)S[s][i].im * S[s][i].im;
      }
    store_spinor<P, D, R>(x, nullptr, k, g.VHpad, X);
    store_spinor<P, D, R>(r, nullptr, k, g.VHpad, S);
  }
  double res[3];
  if (block_reduce_finish<3, 256>(w, partials, gridDim.x, blockIdx.x, ticket, res) && threadIdx.x == 0) {
    const Cx<double> rho0(sc->red[BI_RHO], sc->red[BI_RHO + 1]);
    const Cx<double> rho1(res[0], res[1]);
    const Cx<double> al(sc->red[BI_ALPHA], sc->red[BI_ALPHA + 1]);
    const Cx<double> omd(sc->red[BI_OMEGA], sc->red[BI_OMEGA + 1]);
    const Cx<double> beta = cdiv(rho1, rho0) * cdiv(al, omd);
    sc->red[BI_BETA] = beta.re;
    sc->red[BI_BETA + 1] = beta.im;
    sc->red[BI_RHO] = rho1.re;
    sc->red[BI_RHO + 1] = rho1.im;
    const double rr = res[2];
    sc->rr = rr;
    sc->iters += 1;
    if (rr < sc->target || !(rr == rr) || !(beta.re == beta.re)) {
      sc->converged = 1;
      sc->reason = 1;
    } else if (sc->delta2 > 0.0 && rr < sc->delta2 * sc->maxrr) {
      sc->converged = 1;
      sc->reason = 2;
    }
    if (rr > sc->maxrr) sc->maxrr = rr;
  }
}

// p = r + beta (p - omega v)
template <int P, int D>
__global__ void __launch_bounds__(256) k_bi_p(void* p, const void* r, const void* v, const SolverScalars* sc, Geom g) {
  typedef BR<P, D> R;
  if (sc->converged) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= g.VH) return;
  const Cx<R> be((R)sc->red[BI_BETA], (R)sc->red[BI_BETA + 1]);
  const Cx<R> om((R)sc->red[BI_OMEGA], (R)sc->red[BI_OMEGA + 1]);
  Cx<R> Pv[4][D], Rv[4][D], V[4][D];
  load_spinor<P, D, R>(Pv, p, nullptr, k, g.VHpad);
  load_spinor<P, D, R>(Rv, r, nullptr, k, g.VHpad);
  load_spinor<P, D, R>(V, v, nullptr, k, g.VHpad);
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int i = 0; i < D; i++) Pv[s][i] = Rv[s][i] + be * (Pv[s][i] - om * V[s][i]);
  store_spinor<P, D, R>(p, nullptr, k, g.VHpad, Pv);
}

#define YH_BI_D(CALL)                                       \
  switch (c->dr) {                                          \
    case 2: { constexpr int D_ = 2; CALL; } break;          \
    case 3: { constexpr int D_ = 3; CALL; } break;          \
    case 4: { constexpr int D_ = 4; CALL; } break;          \
    case 6: { constexpr int D_ = 6; CALL; } break;          \
    case 8: { constexpr int D_ = 8; CALL; } break;          \
    default: set_error("unsupported representation dimension"); return YH_ERR_UNSUPPORTED; \
  }
#define YH_BI_P(PREC, CALL)                                 \
  if ((PREC) == PREC_F64) { constexpr int P_ = PREC_F64; CALL; } \
  else { constexpr int P_ = PREC_F32; CALL; }

// one BiCGstab iteration on fields of precision prec (FP64 or FP32)
static int bi_iteration(Ctx* c, int prec, int recon, const yh_inv_par* par, Field* x, Field* r, Field* r0, Field* p,
                        Field* v, Field* t) {
  int rc;
  const dim3 grid((c->g.VH + 255) / 256), blk(256);
  SolverScalars* sc = c->sc;
  const int* stop = &c->sc->converged;
  if ((rc = apply_op(c, prec, recon, par->mvm, par->mass, v, p, nullptr, -1, stop))) return rc;
  YH_BI_D(YH_BI_P(prec, hipLaunchKernelGGL((k_bi_alpha<P_, D_>), grid, blk, 0, c->stream, v->half[0], r0->half[0], sc,
                                           c->partials, c->ticket, c->g)));
  YH_BI_D(YH_BI_P(prec, hipLaunchKernelGGL((k_bi_s<P_, D_>), grid, blk, 0, c->stream, r->half[0], v->half[0], sc, c->g)));
  if ((rc = apply_op(c, prec, recon, par->mvm, par->mass, t, r, nullptr, -1, stop))) return rc;
  YH_BI_D(YH_BI_P(prec, hipLaunchKernelGGL((k_bi_omega<P_, D_>), grid, blk, 0, c->stream, r->half[0], t->half[0], sc,
                                           c->partials, c->ticket, c->g)));
  YH_BI_D(YH_BI_P(prec, hipLaunchKernelGGL((k_bi_xr<P_, D_>), grid, blk, 0, c->stream, x->half[0], r->half[0],
                                           p->half[0], t->half[0], r0->half[0], sc, c->partials, c->ticket, c->g)));
  YH_BI_D(YH_BI_P(prec, hipLaunchKernelGGL((k_bi_p<P_, D_>), grid, blk, 0, c->stream, p->half[0], r->half[0],
                                           v->half[0], sc, c->g)));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

static int need_field(Ctx* c, int prec, int slot, Field** f) {
  int rc;
  *f = tmp_field(c, prec, slot, &rc);
  return rc;
}

// r = b - A x in FP64; returns |r|^2 and rho = <r, r0> (r0 may be r)
static int true_residual(Ctx* c, const yh_inv_par* par, Field* x, const Field* b, Field* r, Field* tmp) {
  int rc;
  if ((rc = apply_op(c, PREC_F64, default_recon(c, PREC_F64), par->mvm, par->mass, tmp, x, nullptr, -1, nullptr)))
    return rc;
  if ((rc = blas_copy(c, r, b, 0))) return rc;
  return blas_axpby(c, r, -1.0, tmp, 1.0, 0);
}

// FP64 BiCGstab (the reference's recurrence) or, with sloppy = PREC_F32,
// the same recurrence on FP32 fields inside FP64 reliable updates
int bicgstab_device(Ctx* c, const yh_inv_par* par, int sloppy, Field* x, const Field* b, int* iters, int* updates) {
  int rc;
  const bool mixed = sloppy == PREC_F32;
  const int S = mixed ? PREC_F32 : PREC_F64;
  Field *r64, *t64, *xs, *r, *r0, *p, *v, *t;
  if ((rc = need_field(c, PREC_F64, 2, &r64)) || (rc = need_field(c, PREC_F64, 3, &t64))) return rc;
  if (mixed) {
    if ((rc = need_field(c, PREC_F32, 2, &xs)) || (rc = need_field(c, PREC_F32, 3, &r)) ||
        (rc = need_field(c, PREC_F32, 4, &r0)) || (rc = need_field(c, PREC_F32, 5, &p)) ||
        (rc = need_field(c, PREC_F32, 6, &v)) || (rc = need_field(c, PREC_F32, 7, &t)))
      return rc;
  } else {
    xs = x;
    r = r64;
    if ((rc = need_field(c, PREC_F64, 4, &r0)) || (rc = need_field(c, PREC_F64, 5, &p)) ||
        (rc = need_field(c, PREC_F64, 6, &v)) || (rc = need_field(c, PREC_F64, 7, &t)))
      return rc;
  }
  const int recon = default_recon(c, S);
  double nb[2], rho[2];
  if ((rc = reduce_read(c, 0, b, nullptr, 0, nb))) return rc;
  const double norm = nb[0], target = norm * par->prec;
  // r = b - A x ; r0 = r ; p = r (bicgstab.cpp:36-44)
  if ((rc = true_residual(c, par, x, b, r64, t64))) return rc;
  if ((rc = reduce_read(c, 1, r64, r64, 0, rho))) return rc;
  double rr = rho[0];
  *iters = 0;
  if (updates) *updates = 0;
  if (rr < target) return YH_OK;  // bicgstab.cpp:50-55
  if ((rc = field_convert(c, r, 0, r64, 0)) || (rc = field_convert(c, r0, 0, r64, 0)) ||
      (rc = field_convert(c, p, 0, r64, 0)))
    return rc;
  if (mixed && (rc = field_zero(c, xs, YH_EVEN))) return rc;
  // reliable updates every K iterations at the latest, or when |r| dropped by delta
  const int K = par->reliable_every > 0 ? par->reliable_every : 100;
  const double delta = (par->delta > 0.0 && par->delta < 1.0) ? par->delta : 0.1;
  SolverScalars* h = c->sc_host;
  memset(h, 0, sizeof(SolverScalars));
  h->target = target;
  h->norm_b = norm;
  h->rr = rr;
  h->maxrr = rr;
  h->delta2 = mixed ? delta * delta : 0.0;
  h->red[BI_RHO] = rho[0];
  h->red[BI_RHO + 1] = rho[1];
  if ((rc = write_scalars(c))) return rc;
  int total = 0, at_update = 0, nupd = 0, stalls = 0;
  double best = rr;
  const int chunk = 16;
  while (true) {
    int n = mixed ? std::min(chunk, std::max(1, K - (total - at_update))) : chunk;
    if (par->max_iter > 0) n = std::min(n, std::max(1, par->max_iter - total));
    for (int i = 0; i < n; i++)
      if ((rc = bi_iteration(c, S, recon, par, xs, r, r0, p, v, t))) return rc;
    if ((rc = read_scalars(c))) return rc;
    c->mvm_counter += (long)(h->iters - total) * 2 * hops_per_apply(par->mvm);
    total = h->iters;
    if (!(h->rr == h->rr)) { set_error("bicgstab: non-finite residual"); return YH_ERR_NONFINITE; }
    const bool capped = par->max_iter > 0 && total >= par->max_iter;
    if (!mixed) {
      if (h->converged && !(h->rr < target)) {
        // the flag came from a non-finite scalar (rho0 = 0 or omega = 0)
        *iters = 2 * total;
        set_error("bicgstab: breakdown (non-finite scalar)");
        return YH_ERR_NONFINITE;
      }
      if (h->converged) break;
      if (capped) {
        *iters = 2 * total;
        set_error("bicgstab: max_iter reached");
        return YH_ERR_NOCONV;
      }
      continue;
    }
    if (!h->converged && total - at_update < K && !capped) continue;
    if (h->converged) {
      // the iteration that raised the stop flag skipped its p update;
      // complete it with the iterated residual before r is replaced
      h->converged = 0;
      if ((rc = write_scalars(c))) return rc;
      const dim3 grid((c->g.VH + 255) / 256), blk(256);
      YH_BI_D(hipLaunchKernelGGL((k_bi_p<PREC_F32, D_>), grid, blk, 0, c->stream, p->half[0], r->half[0], v->half[0],
                                 c->sc, c->g));
      YH_HIP(hipGetLastError());
    }
    // reliable update: x += xs ; r = b - A x in FP64 ; restart the sloppy
    // residual and rho from it, keeping p and the shadow residual
    if ((rc = field_convert(c, t64, 0, xs, 0)) || (rc = blas_axpby(c, x, 1.0, t64, 1.0, 0))) return rc;
    if ((rc = true_residual(c, par, x, b, r64, t64))) return rc;
    nupd++;
    double r2[2];
    if ((rc = reduce_read(c, 0, r64, nullptr, 0, r2))) return rc;
    rr = r2[0];
    if (!(rr == rr)) { set_error("bicgstab: non-finite true residual"); return YH_ERR_NONFINITE; }
    if (rr < target) break;
    if (rr < best) {
      best = rr;
      stalls = 0;
    } else {
      stalls++;
    }
    if (capped || stalls >= 10) {
      *iters = 2 * total;
      if (updates) *updates = nupd;
      set_error(capped ? "bicgstab: max_iter reached" : "bicgstab: true residual stalled");
      return YH_ERR_NOCONV;
    }
    if ((rc = field_convert(c, r, 0, r64, 0)) || (rc = field_zero(c, xs, YH_EVEN))) return rc;
    if ((rc = reduce_read(c, 1, r, r0, 0, rho))) return rc;
    if ((rc = read_scalars(c))) return rc;
    h->red[BI_RHO] = rho[0];
    h->red[BI_RHO + 1] = rho[1];
    h->rr = rr;
    h->maxrr = rr;
    h->converged = 0;
    h->reason = 0;
    h->iters = total;
    at_update = total;
    if ((rc = write_scalars(c))) return rc;
  }
  *iters = 2 * total;  // two operator applications per iteration (bicgstab.cpp:89)
  if (updates) *updates = nupd;
  if (par->log)
    fprintf(stderr, "[INVERTER][NOTICE] bicgstab: iterations = %d, precision = %1.4e\n", 2 * total,
            (mixed ? rr : h->rr) / norm);
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

}  // namespace yh
