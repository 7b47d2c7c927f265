// BLAS-1 on one parity half (library/linalg.cpp:29-271) and the fused CG
// update kernels.  Every kernel reads each operand once and writes once;
// reductions are FP64 and deterministic (yh_reduce.hpp) whatever the
// storage precision.
#include "yh_internal.hpp"
#include "yh_reduce.hpp"

namespace yh {

#define YH_SITE_LOOP_BEGIN                      \
  const int k = blockIdx.x * 256 + threadIdx.x; \
  const bool act = k < g.VH;

template <int P, int D>
using RT = typename SpinorStore<P, D>::real;

template <int P, int D>
__global__ void __launch_bounds__(256) k_axpby(void* y, float* yn, double a, const void* x, const float* xn, double b,
                                               Geom g) {
  typedef RT<P, D> R;
  YH_SITE_LOOP_BEGIN
  if (!act) return;
  Cx<R> X[4][D], Y[4][D];
  load_spinor<P, D, R>(X, x, xn, k, g.VHpad);
  load_spinor<P, D, R>(Y, y, yn, k, g.VHpad);
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int i = 0; i < D; i++) Y[s][i] = Cx<R>((R)a * X[s][i].re + (R)b * Y[s][i].re, (R)a * X[s][i].im + (R)b * Y[s][i].im);
  store_spinor<P, D, R>(y, yn, k, g.VHpad, Y);
}

template <int P, int D>
__global__ void __launch_bounds__(256) k_caxpy(void* y, float* yn, double ar, double ai, const void* x, const float* xn,
                                               Geom g) {
  typedef RT<P, D> R;
  YH_SITE_LOOP_BEGIN
  if (!act) return;
  Cx<R> X[4][D], Y[4][D];
  load_spinor<P, D, R>(X, x, xn, k, g.VHpad);
  load_spinor<P, D, R>(Y, y, yn, k, g.VHpad);
  Cx<R> c((R)ar, (R)ai);
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int i = 0; i < D; i++) cmac(Y[s][i], c, X[s][i]);
  store_spinor<P, D, R>(y, yn, k, g.VHpad, Y);
}

template <int P, int D>
__global__ void __launch_bounds__(256) k_cscale(void* y, float* yn, double ar, double ai, Geom g) {
  typedef RT<P, D> R;
  YH_SITE_LOOP_BEGIN
  if (!act) return;
  Cx<R> Y[4][D];
  load_spinor<P, D, R>(Y, y, yn, k, g.VHpad);
  Cx<R> c((R)ar, (R)ai);
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int i = 0; i < D; i++) Y[s][i] = c * Y[s][i];
  store_spinor<P, D, R>(y, yn, k, g.VHpad, Y);
}

template <int P, int D>
__global__ void __launch_bounds__(256) k_g5(void* y, float* yn, Geom g) {
  typedef RT<P, D> R;
  YH_SITE_LOOP_BEGIN
  if (!act) return;
  Cx<R> Y[4][D];
  load_spinor<P, D, R>(Y, y, yn, k, g.VHpad);
#pragma unroll
  for (int s = 2; s < 4; s++)
#pragma unroll
    for (int i = 0; i < D; i++) Y[s][i] = Cx<R>(-Y[s][i].re, -Y[s][i].im);
  store_spinor<P, D, R>(y, yn, k, g.VHpad, Y);
}

template <int PD, int PS, int D>
__global__ void __launch_bounds__(256) k_convert(void* y, float* yn, const void* x, const float* xn, Geom g) {
  YH_SITE_LOOP_BEGIN
  if (!act) return;
  Cx<double> X[4][D];
  load_spinor<PS, D, double>(X, x, xn, k, g.VHpad);
  store_spinor<PD, D, double>(y, yn, k, g.VHpad, X);
}

// OP 0: |a|^2 ; 1: <a, b> = sum a conj(b) (2 values); 2: Re <a, b>
template <int P, int D, int OP>
__global__ void __launch_bounds__(256) k_reduce(const void* a, const float* an, const void* b, const float* bn,
                                                double* partials, unsigned* ticket, double* out, Geom g) {
  typedef RT<P, D> R;
  YH_SITE_LOOP_BEGIN
  double v[2] = {0.0, 0.0};
  if (act) {
    Cx<R> A[4][D];
    load_spinor<P, D, R>(A, a, an, k, g.VHpad);
    if constexpr (OP == 0) {
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int i = 0; i < D; i++) v[0] += (double)A[s][i].re * A[s][i].re + (double)A[s][i].im * A[s][i].im;
    } else {
      Cx<R> B[4][D];
      load_spinor<P, D, R>(B, b, bn, k, g.VHpad);
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int i = 0; i < D; i++) {
          v[0] += (double)A[s][i].re * B[s][i].re + (double)A[s][i].im * B[s][i].im;
          v[1] += (double)A[s][i].im * B[s][i].re - (double)A[s][i].re * B[s][i].im;
        }
    }
  }
  block_reduce_finish<2, 256>(v, partials, gridDim.x, blockIdx.x, ticket, out);
}

// CG: alpha = rr / pap; x += alpha p; r -= alpha q; rr_new = |r|^2.
// The last block updates the scalars (library/cg.cpp:174-189).
template <int P, int PP, int D>
__global__ void __launch_bounds__(256) k_cg_xr(void* x, float* xn, void* r, float* rn, const void* p, const float* pn,
                                               const void* q, const float* qn, SolverScalars* sc, double* partials,
                                               unsigned* ticket, Geom g) {
  typedef RT<P, D> R;
  if (sc->converged) return;
  const double alpha = sc->rr / sc->red[0];
  YH_SITE_LOOP_BEGIN
  double v[1] = {0.0};
  if (act) {
    // all four operands are loaded before the first store: the fields are
    // distinct, but the compiler cannot know that, and a load after a store
    // would wait for it
    Cx<R> X[4][D], Pv[4][D], Q[4][D], Rv[4][D];
    load_spinor<PP, D, R>(Pv, p, pn, k, g.VHpad);
    load_spinor<P, D, R>(X, x, xn, k, g.VHpad);
    load_spinor<PP, D, R>(Q, q, qn, k, g.VHpad);
    load_spinor<P, D, R>(Rv, r, rn, k, g.VHpad);
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < D; i++) {
        X[s][i] = Cx<R>(X[s][i].re + (R)alpha * Pv[s][i].re, X[s][i].im + (R)alpha * Pv[s][i].im);
        Rv[s][i] = Cx<R>(Rv[s][i].re - (R)alpha * Q[s][i].re, Rv[s][i].im - (R)alpha * Q[s][i].im);
      }
    store_spinor<P, D, R>(x, xn, k, g.VHpad, X);
    store_spinor<P, D, R>(r, rn, k, g.VHpad, Rv);
    // a 16-bit residual is reduced as stored (quantised)
    if constexpr (P == PREC_F16) load_spinor<P, D, R>(Rv, r, rn, k, g.VHpad);
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < D; i++) v[0] += (double)Rv[s][i].re * Rv[s][i].re + (double)Rv[s][i].im * Rv[s][i].im;
  }
  double res[1];
  if (block_reduce_finish<1, 256>(v, partials, gridDim.x, blockIdx.x, ticket, res)) {
    if (threadIdx.x == 0) {
      double rr_new = res[0];
      sc->alpha = alpha;
      sc->beta = rr_new / sc->rr;
      sc->rr_new = rr_new;
      sc->rr = rr_new;
      sc->iters += 1;
      if (rr_new < sc->target || !(rr_new == rr_new)) {
        sc->converged = 1;
        sc->reason = 1;
      } else if (sc->delta2 > 0.0 && rr_new < sc->delta2 * sc->maxrr) {
        sc->converged = 1;
        sc->reason = 2;
      }
      if (rr_new > sc->maxrr) sc->maxrr = rr_new;
    }
  }
}

// CG: p = r + beta p   (p in precision PP, r in P)
template <int P, int PP, int D>
__global__ void __launch_bounds__(256) k_cg_p(void* p, float* pn, const void* r, const float* rn, const SolverScalars* sc,
                                              Geom g) {
  typedef RT<P, D> R;
  if (sc->converged) return;
  const R beta = (R)sc->beta;
  YH_SITE_LOOP_BEGIN
  if (!act) return;
  Cx<R> Pv[4][D], Rv[4][D];
  load_spinor<PP, D, R>(Pv, p, pn, k, g.VHpad);
  load_spinor<P, D, R>(Rv, r, rn, k, g.VHpad);
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int i = 0; i < D; i++) Pv[s][i] = Cx<R>(Rv[s][i].re + beta * Pv[s][i].re, Rv[s][i].im + beta * Pv[s][i].im);
  store_spinor<PP, D, R>(p, pn, k, g.VHpad, Pv);
}

// Reliable update of the mixed CG, first half: x (FP64) += xs (FP32 sloppy
// solution), xs = 0, and, if the sloppy recurrence stopped on its own flag
// (the iteration that raised it skipped p = r + beta p), that p step.
template <int PP, int D>
__global__ void __launch_bounds__(256) k_ru_begin(void* x, void* xs, void* p, float* pn, const void* rs,
                                                  const SolverScalars* sc, Geom g) {
  if (sc->finished) return;
  const bool complete_p = sc->converged != 0;
  const float beta = (float)sc->beta;
  YH_SITE_LOOP_BEGIN
  if (!act) return;
  Cx<float> Xs[4][D];
  Cx<double> X[4][D];
  load_spinor<PREC_F32, D, float>(Xs, xs, nullptr, k, g.VHpad);
  load_spinor<PREC_F64, D, double>(X, x, nullptr, k, g.VHpad);
  if (complete_p) {
    Cx<float> Pv[4][D], Rv[4][D];
    load_spinor<PP, D, float>(Pv, p, pn, k, g.VHpad);
    load_spinor<PREC_F32, D, float>(Rv, rs, nullptr, k, g.VHpad);
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < D; i++) Pv[s][i] = Cx<float>(Rv[s][i].re + beta * Pv[s][i].re, Rv[s][i].im + beta * Pv[s][i].im);
    store_spinor<PP, D, float>(p, pn, k, g.VHpad, Pv);
  }
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int i = 0; i < D; i++) {
      X[s][i] = Cx<double>(X[s][i].re + (double)Xs[s][i].re, X[s][i].im + (double)Xs[s][i].im);
      Xs[s][i] = Cx<float>(0.0f, 0.0f);
    }
  store_spinor<PREC_F64, D, double>(x, nullptr, k, g.VHpad, X);
  store_spinor<PREC_F32, D, float>(xs, nullptr, k, g.VHpad, Xs);
}

// Second half: r = b - t with t = A x in FP64; rs = r in FP32; |r|^2 in
// FP64.  The last workgroup ends the solve (target reached, non-finite, or
// 10 updates in a row without a lower true residual) or restarts the
// sloppy recurrence from the true residual.
template <int D>
__global__ void __launch_bounds__(256) k_ru_end(const void* b, const void* t, void* rs, SolverScalars* sc,
                                                double* partials, unsigned* ticket, Geom g) {
  if (sc->finished) return;
  YH_SITE_LOOP_BEGIN
  double v[1] = {0.0};
  if (act) {
    Cx<double> B[4][D], T[4][D];
    load_spinor<PREC_F64, D, double>(B, b, nullptr, k, g.VHpad);
    load_spinor<PREC_F64, D, double>(T, t, nullptr, k, g.VHpad);
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < D; i++) {
        B[s][i] = Cx<double>(B[s][i].re - T[s][i].re, B[s][i].im - T[s][i].im);
        v[0] += B[s][i].re * B[s][i].re + B[s][i].im * B[s][i].im;
      }
    store_spinor<PREC_F32, D, double>(rs, nullptr, k, g.VHpad, B);
  }
  double res[1];
  if (block_reduce_finish<1, 256>(v, partials, gridDim.x, blockIdx.x, ticket, res) && threadIdx.x == 0) {
    const double rr = res[0];
    sc->rr_true = rr;
    sc->nupd += 1;
    if (!(rr == rr)) {
      sc->finished = 2;
      sc->converged = 1;
    } else if (rr < sc->target) {
      sc->finished = 1;
      sc->converged = 1;
      sc->reason = 1;
    } else {
      if (rr < sc->best_rr) {
        sc->best_rr = rr;
        sc->stalls = 0;
      } else {
        sc->stalls += 1;
      }
      if (sc->stalls >= 10) {
        sc->finished = 3;
        sc->converged = 1;
      } else {
        sc->rr = rr;
        sc->maxrr = rr;
        sc->converged = 0;
        sc->reason = 0;
      }
    }
  }
}

// ---------------------------------------------------------------------------
#define YH_D_SWITCH(CALL)                                   \
  switch (c->dr) {                                          \
    case 2: { constexpr int D_ = 2; CALL; } break;          \
    case 3: { constexpr int D_ = 3; CALL; } break;          \
    case 4: { constexpr int D_ = 4; CALL; } break;          \
    case 6: { constexpr int D_ = 6; CALL; } break;          \
    case 8: { constexpr int D_ = 8; CALL; } break;          \
    default: set_error("unsupported representation dimension"); return YH_ERR_UNSUPPORTED; \
  }

#define YH_P_SWITCH(PREC, CALL)                             \
  switch (PREC) {                                           \
    case PREC_F64: { constexpr int P_ = PREC_F64; CALL; } break; \
    case PREC_F32: { constexpr int P_ = PREC_F32; CALL; } break; \
    default: { constexpr int P_ = PREC_F16; CALL; } break;  \
  }

static dim3 gsz(Ctx* c) { return dim3((c->g.VH + 255) / 256); }

int blas_axpby(Ctx* c, Field* y, double a, const Field* x, double b, int par) {
  if (x->prec != y->prec) { set_error("precision mismatch"); return YH_ERR_ARG; }
  YH_D_SWITCH(YH_P_SWITCH(y->prec, hipLaunchKernelGGL((k_axpby<P_, D_>), gsz(c), dim3(256), 0, c->stream, y->half[par],
                                                      y->nrm[par], a, x->half[par], x->nrm[par], b, c->g)));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int blas_caxpy(Ctx* c, Field* y, double ar, double ai, const Field* x, int par) {
  if (x->prec != y->prec) { set_error("precision mismatch"); return YH_ERR_ARG; }
  YH_D_SWITCH(YH_P_SWITCH(y->prec, hipLaunchKernelGGL((k_caxpy<P_, D_>), gsz(c), dim3(256), 0, c->stream, y->half[par],
                                                      y->nrm[par], ar, ai, x->half[par], x->nrm[par], c->g)));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int blas_cscale(Ctx* c, Field* y, double ar, double ai, int par) {
  YH_D_SWITCH(YH_P_SWITCH(y->prec, hipLaunchKernelGGL((k_cscale<P_, D_>), gsz(c), dim3(256), 0, c->stream, y->half[par],
                                                      y->nrm[par], ar, ai, c->g)));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int blas_g5(Ctx* c, Field* y, int par) {
  YH_D_SWITCH(YH_P_SWITCH(y->prec, hipLaunchKernelGGL((k_g5<P_, D_>), gsz(c), dim3(256), 0, c->stream, y->half[par],
                                                      y->nrm[par], c->g)));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int field_convert(Ctx* c, Field* dst, int dpar, const Field* src, int spar) {
#define YH_CV(PD, PS) \
  YH_D_SWITCH(hipLaunchKernelGGL((k_convert<PD, PS, D_>), gsz(c), dim3(256), 0, c->stream, dst->half[dpar], dst->nrm[dpar], src->half[spar], src->nrm[spar], c->g))
  int key = dst->prec * 3 + src->prec;
  switch (key) {
    case 0: YH_CV(PREC_F64, PREC_F64); break;
    case 1: YH_CV(PREC_F64, PREC_F32); break;
    case 2: YH_CV(PREC_F64, PREC_F16); break;
    case 3: YH_CV(PREC_F32, PREC_F64); break;
    case 4: YH_CV(PREC_F32, PREC_F32); break;
    case 5: YH_CV(PREC_F32, PREC_F16); break;
    case 6: YH_CV(PREC_F16, PREC_F64); break;
    case 7: YH_CV(PREC_F16, PREC_F32); break;
    default: YH_CV(PREC_F16, PREC_F16); break;
  }
#undef YH_CV
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int blas_copy(Ctx* c, Field* dst, const Field* src, int par) { return field_convert(c, dst, par, src, par); }

static int reduce_launch(Ctx* c, int op, const Field* a, const Field* b, int par, int slot) {
  if (b && a->prec != b->prec) { set_error("precision mismatch"); return YH_ERR_ARG; }
  double* out = &c->sc->red[slot];
  const void* bp = b ? b->half[par] : nullptr;
  const float* bn = b ? b->nrm[par] : nullptr;
#define YH_RL(OPV) \
  YH_D_SWITCH(YH_P_SWITCH(a->prec, hipLaunchKernelGGL((k_reduce<P_, D_, OPV>), gsz(c), dim3(256), 0, c->stream, a->half[par], a->nrm[par], bp, bn, c->partials, c->ticket, out, c->g)))
  if (op == 0) { YH_RL(0); }
  else if (op == 1) { YH_RL(1); }
  else { YH_RL(2); }
#undef YH_RL
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int blas_sqnorm(Ctx* c, const Field* a, int par, int slot) { return reduce_launch(c, 0, a, nullptr, par, slot); }
int blas_dot(Ctx* c, const Field* a, const Field* b, int par, int slot) { return reduce_launch(c, 1, a, b, par, slot); }
int blas_dot_re(Ctx* c, const Field* a, const Field* b, int par, int slot) { return reduce_launch(c, 2, a, b, par, slot); }

// supported (x/r precision, p/q precision) pairs: (64,64), (32,32), (32,16)
int cg_update_xr(Ctx* c, Field* x, Field* r, const Field* p, const Field* q, int par, int prec) {
  (void)prec;
  if (x->prec != r->prec || p->prec != q->prec) { set_error("precision mismatch"); return YH_ERR_ARG; }
#define YH_XR(PX, PPV) \
  YH_D_SWITCH(hipLaunchKernelGGL((k_cg_xr<PX, PPV, D_>), gsz(c), dim3(256), 0, c->stream, x->half[par], x->nrm[par], r->half[par], r->nrm[par], p->half[par], p->nrm[par], q->half[par], q->nrm[par], c->sc, c->partials, c->ticket, c->g))
  if (x->prec == PREC_F64 && p->prec == PREC_F64) { YH_XR(PREC_F64, PREC_F64); }
  else if (x->prec == PREC_F32 && p->prec == PREC_F32) { YH_XR(PREC_F32, PREC_F32); }
  else if (x->prec == PREC_F32 && p->prec == PREC_F16) { YH_XR(PREC_F32, PREC_F16); }
  else { set_error("unsupported CG precision pair"); return YH_ERR_UNSUPPORTED; }
#undef YH_XR
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int cg_update_p(Ctx* c, Field* p, const Field* r, int par, int prec) {
  (void)prec;
#define YH_PU(PX, PPV) \
  YH_D_SWITCH(hipLaunchKernelGGL((k_cg_p<PX, PPV, D_>), gsz(c), dim3(256), 0, c->stream, p->half[par], p->nrm[par], r->half[par], r->nrm[par], c->sc, c->g))
  if (r->prec == PREC_F64 && p->prec == PREC_F64) { YH_PU(PREC_F64, PREC_F64); }
  else if (r->prec == PREC_F32 && p->prec == PREC_F32) { YH_PU(PREC_F32, PREC_F32); }
  else if (r->prec == PREC_F32 && p->prec == PREC_F16) { YH_PU(PREC_F32, PREC_F16); }
  else { set_error("unsupported CG precision pair"); return YH_ERR_UNSUPPORTED; }
#undef YH_PU
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int ru_begin(Ctx* c, Field* x, Field* xs, Field* p, const Field* rs) {
  if (x->prec != PREC_F64 || xs->prec != PREC_F32 || rs->prec != PREC_F32 || p->prec == PREC_F64) {
    set_error("reliable update: unsupported precisions");
    return YH_ERR_ARG;
  }
  if (p->prec == PREC_F32) {
    YH_D_SWITCH(hipLaunchKernelGGL((k_ru_begin<PREC_F32, D_>), gsz(c), dim3(256), 0, c->stream, x->half[0], xs->half[0],
                                   p->half[0], p->nrm[0], rs->half[0], c->sc, c->g));
  } else {
    YH_D_SWITCH(hipLaunchKernelGGL((k_ru_begin<PREC_F16, D_>), gsz(c), dim3(256), 0, c->stream, x->half[0], xs->half[0],
                                   p->half[0], p->nrm[0], rs->half[0], c->sc, c->g));
  }
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int ru_end(Ctx* c, const Field* b, const Field* t, Field* rs) {
  if (b->prec != PREC_F64 || t->prec != PREC_F64 || rs->prec != PREC_F32) {
    set_error("reliable update: unsupported precisions");
    return YH_ERR_ARG;
  }
  YH_D_SWITCH(hipLaunchKernelGGL((k_ru_end<D_>), gsz(c), dim3(256), 0, c->stream, b->half[0], t->half[0], rs->half[0],
                                 c->sc, c->partials, c->ticket, c->g));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

}  // namespace yh
