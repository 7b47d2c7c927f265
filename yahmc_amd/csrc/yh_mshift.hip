// Multishift CG with device-resident recurrences, for one or several
// right-hand sides (library/cg.cpp:11-130).
//
// Per iteration and right-hand side: H^2 p (four fused hop kernels, the last
// one also reducing <p, H^2 p> = |H p|^2), then two fused update kernels:
//   k_ms_x : r += beta q ; |r|^2 -> alpha; the last workgroup computes the
//            zeta coefficients c3_s, the solution step bt_s = beta c3_s / c2_s
//            and the direction factor of every active shift and the
//            convergence flags
//   k_ms_k : for the shifts that were active in this iteration, in one pass
//            over x_s and k_s:  x_s -= bt_s k_s ;  k_s = at_s k_s + c3_s r
//            (the solution step uses k_s before its update, as the
//            reference's order x_s -= bt k_s, then k_s = ... has it) ;
//            p = alpha p + r
// The host only queues chunks of iterations and reads one flag per chunk.
// Right-hand sides are independent recurrences sharing the gauge field; they
// are interleaved iteration by iteration so consecutive hop launches reuse
// the links from the MALL.
#include <string.h>
#include <vector>
#include "yh_internal.hpp"
#include "yh_reduce.hpp"

namespace yh {

enum { MS_MAX = 48 };

struct MsState {
  double rr, pap, beta, beta_old, alpha, norm, target;
  int done, iters, nshift, notconv;
  double shift[MS_MAX], c1[MS_MAX], c2[MS_MAX], c3[MS_MAX], at[MS_MAX], bt[MS_MAX];
  int act[MS_MAX], upd[MS_MAX];
};

struct MsFields {
  void* x[MS_MAX];   // even halves of the solutions (FP64)
  void* k[MS_MAX];   // even halves of the shifted search vectors
};

template <int D>
__global__ void __launch_bounds__(256) k_ms_x(MsState* st, const MsFields* __restrict__ mf, void* r, const void* q,
                                              double* partials, unsigned* ticket, Geom g) {
  (void)mf;
  if (st->done) return;
  const double atmp = st->rr, btmp = st->pap;
  const double beta_old = st->beta;
  const double beta = -(atmp / btmp);
  const double alpha = st->alpha;
  const int ns = st->nshift;
  const int k = blockIdx.x * 256 + threadIdx.x;
  double v[1] = {0.0};
  if (k < g.VH) {
    Cx<double> R[4][D], Q[4][D];
    load_spinor<PREC_F64, D, double>(R, r, nullptr, k, g.VHpad);
    load_spinor<PREC_F64, D, double>(Q, q, nullptr, k, g.VHpad);
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int i = 0; i < D; i++) {
        R[a][i] = Cx<double>(R[a][i].re + beta * Q[a][i].re, R[a][i].im + beta * Q[a][i].im);
        v[0] += R[a][i].re * R[a][i].re + R[a][i].im * R[a][i].im;
      }
    store_spinor<PREC_F64, D, double>(r, nullptr, k, g.VHpad, R);
  }
  double res[1];
  if (block_reduce_finish<1, 256>(v, partials, gridDim.x, blockIdx.x, ticket, res) && threadIdx.x == 0) {
    const double rr_new = res[0];
    const double alpha_new = rr_new / atmp;
    int notconv = 0;
    for (int s = 0; s < ns; s++) {
      st->upd[s] = st->act[s];
      if (!st->act[s]) continue;
      double c1 = st->c1[s], c2 = st->c2[s];
      double c3 = c2 * c1 * beta_old / (beta * alpha * (c1 - c2) + c1 * beta_old * (1 - st->shift[s] * beta));
      st->bt[s] = beta * (c3 / c2);
      st->c3[s] = c3;
      st->at[s] = alpha_new * (c3 * c3) / (c2 * c2);
      if (rr_new * c3 * c3 < st->target) st->act[s] = 0;
      else notconv++;
      st->c1[s] = c2;
      st->c2[s] = c3;
    }
    st->beta_old = beta_old;
    st->beta = beta;
    st->alpha = alpha_new;
    st->rr = rr_new;
    st->iters += 1;
    st->notconv = notconv;
    if (notconv == 0 || !(rr_new == rr_new)) st->done = 1;
  }
}

template <int D>
__global__ void __launch_bounds__(256) k_ms_k(MsState* st, const MsFields* __restrict__ mf, const void* r, void* p,
                                              unsigned* ticket, Geom g) {
  // a chunk can run past the final iteration: then no shift is marked for
  // update (cleared below) and p is left alone
  const int ns = st->nshift;
  const bool move_p = !st->done;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < g.VH) {
    Cx<double> R[4][D];
    load_spinor<PREC_F64, D, double>(R, r, nullptr, k, g.VHpad);
    for (int s = 0; s < ns; s++) {
      if (!st->upd[s]) continue;
      const double at = st->at[s], c3 = st->c3[s], bt = st->bt[s];
      Cx<double> K[4][D], X[4][D];
      load_spinor<PREC_F64, D, double>(K, mf->k[s], nullptr, k, g.VHpad);
      load_spinor<PREC_F64, D, double>(X, mf->x[s], nullptr, k, g.VHpad);
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int i = 0; i < D; i++) {
          X[a][i] = Cx<double>(X[a][i].re - bt * K[a][i].re, X[a][i].im - bt * K[a][i].im);
          K[a][i] = Cx<double>(at * K[a][i].re + c3 * R[a][i].re, at * K[a][i].im + c3 * R[a][i].im);
        }
      store_spinor<PREC_F64, D, double>(mf->x[s], nullptr, k, g.VHpad, X);
      store_spinor<PREC_F64, D, double>(mf->k[s], nullptr, k, g.VHpad, K);
    }
    if (move_p) {
      const double alpha = st->alpha;
      Cx<double> P[4][D];
      load_spinor<PREC_F64, D, double>(P, p, nullptr, k, g.VHpad);
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int i = 0; i < D; i++) P[a][i] = Cx<double>(alpha * P[a][i].re + R[a][i].re, alpha * P[a][i].im + R[a][i].im);
      store_spinor<PREC_F64, D, double>(p, nullptr, k, g.VHpad, P);
    }
  }
  // last workgroup: clear the update marks once every workgroup has used them
  // (every wave of this workgroup has read the marks before the barrier;
  // the ticket is a relaxed agent-scope atomic, as in block_reduce_finish)
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) {
    stores_complete();
    unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    for (int s = 0; s < ns; s++) st->upd[s] = 0;
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ---------------------------------------------------------------------------
template <int D>
static int launch_x(Ctx* c, MsState* st, const MsFields* mf, Field* r, const Field* q) {
  dim3 grid((c->g.VH + 255) / 256);
  hipLaunchKernelGGL(k_ms_x<D>, grid, dim3(256), 0, c->stream, st, mf, r->half[0], q->half[0], c->partials, c->ticket,
                     c->g);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

template <int D>
static int launch_k(Ctx* c, MsState* st, const MsFields* mf, const Field* r, Field* p) {
  dim3 grid((c->g.VH + 255) / 256);
  hipLaunchKernelGGL(k_ms_k<D>, grid, dim3(256), 0, c->stream, st, mf, r->half[0], p->half[0], c->ticket, c->g);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

static int ms_x(Ctx* c, MsState* st, const MsFields* mf, Field* r, const Field* q) {
  switch (c->dr) {
    case 2: return launch_x<2>(c, st, mf, r, q);
    case 3: return launch_x<3>(c, st, mf, r, q);
    case 4: return launch_x<4>(c, st, mf, r, q);
    case 6: return launch_x<6>(c, st, mf, r, q);
    case 8: return launch_x<8>(c, st, mf, r, q);
  }
  set_error("unsupported representation dimension");
  return YH_ERR_UNSUPPORTED;
}

static int ms_k(Ctx* c, MsState* st, const MsFields* mf, const Field* r, Field* p) {
  switch (c->dr) {
    case 2: return launch_k<2>(c, st, mf, r, p);
    case 3: return launch_k<3>(c, st, mf, r, p);
    case 4: return launch_k<4>(c, st, mf, r, p);
    case 6: return launch_k<6>(c, st, mf, r, p);
    case 8: return launch_k<8>(c, st, mf, r, p);
  }
  set_error("unsupported representation dimension");
  return YH_ERR_UNSUPPORTED;
}

// Solves (H^2 + shift_s) x[j][s] = b[j] for nrhs right-hand sides at once.
// x is nrhs * num fields (row j = right-hand side j).
int mshift_device(Ctx* c, const yh_inv_par* par, int nrhs, int num, const double* shift, Field** x, const Field** b,
                  int* iters) {
  if (nrhs < 1 || nrhs > 16 || num < 1 || num > MS_MAX) { set_error("bad shift / rhs count"); return YH_ERR_ARG; }
  int rc = YH_OK;
  const int recon = default_recon(c, PREC_F64);
  struct Rhs {
    Field *r, *p, *q;
    Field *t1 = nullptr, *t2 = nullptr;  // operator temporaries of the multi-RHS path
    std::vector<Field*> k;
  };
  // H^2 of up to four right-hand sides per hop launch (FP64, 18-real links,
  // clover term): batches of 4, then 2, the rest one by one
  // (not instantiated for D_R = 8, whose single-RHS hop already fills the registers)
  const bool batched = c->multi_rhs && nrhs > 1 && c->csw != 0.0 && recon == RECON_18 && c->dr <= 6;
  std::vector<Rhs> rh(nrhs);
  std::vector<Field*> owned;
  auto alloc = [&](Field** f) -> int {
    int e = field_alloc(c, PREC_F64, f);
    if (!e) owned.push_back(*f);
    return e;
  };
  auto cleanup = [&]() { for (Field* f : owned) field_free(f); };
  MsState* st = nullptr;
  MsFields* mf = nullptr;
  hipError_t he = hipMalloc(&st, sizeof(MsState) * nrhs);
  if (he == hipSuccess) he = hipMalloc(&mf, sizeof(MsFields) * nrhs);
  if (he != hipSuccess) {
    if (st) (void)hipFree(st);
    set_error(std::string("mshift allocation: ") + hipGetErrorString(he));
    return YH_ERR_HIP;
  }
  std::vector<MsState> hs(nrhs);
  std::vector<MsFields> hf(nrhs);
  for (int j = 0; j < nrhs && !rc; j++) {
    Rhs& R = rh[j];
    if ((rc = alloc(&R.r)) || (rc = alloc(&R.p)) || (rc = alloc(&R.q))) break;
    if (batched && ((rc = alloc(&R.t1)) || (rc = alloc(&R.t2)))) break;
    R.k.resize(num);
    for (int s = 0; s < num && !rc; s++) rc = alloc(&R.k[s]);
    if (rc) break;
    // r = b ; p = b ; x_s = 0 ; k_s = b  (cg.cpp:35-43)
    if ((rc = blas_copy(c, R.r, b[j], 0)) || (rc = blas_copy(c, R.p, b[j], 0))) break;
    for (int s = 0; s < num && !rc; s++) {
      if ((rc = field_zero(c, x[j * num + s], YH_BOTH))) break;
      rc = blas_copy(c, R.k[s], b[j], 0);
    }
    if (rc) break;
    double nb[2];
    if ((rc = reduce_read(c, 0, b[j], nullptr, 0, nb))) break;
    MsState& S = hs[j];
    memset(&S, 0, sizeof(S));
    S.rr = nb[0];  // |r|^2 of the starting residual r = b
    S.beta = 1.0;
    S.alpha = 0.0;
    S.norm = nb[0];
    S.target = nb[0] * par->prec;
    S.nshift = num;
    for (int s = 0; s < num; s++) {
      S.shift[s] = shift[s];
      S.c1[s] = 1.0;
      S.c2[s] = 1.0;
      S.act[s] = 1;
      hf[j].x[s] = x[j * num + s]->half[0];
      hf[j].k[s] = R.k[s]->half[0];
    }
  }
  if (!rc) {
    he = hipMemcpyAsync(st, hs.data(), sizeof(MsState) * nrhs, hipMemcpyHostToDevice, c->stream);
    if (he == hipSuccess) he = hipMemcpyAsync(mf, hf.data(), sizeof(MsFields) * nrhs, hipMemcpyHostToDevice, c->stream);
    if (he != hipSuccess) { set_error(hipGetErrorString(he)); rc = YH_ERR_HIP; }
  }
  const int chunk = 16;
  int done_iters = 0;
  // one chunk of iterations for every right-hand side; all its launches take
  // their scalars from device memory, so the same sequence can be replayed
  auto enqueue_chunk = [&]() -> int {
    int e = YH_OK;
    for (int i = 0; i < chunk && !e; i++) {
      int j0 = 0;
      while (batched && !e && nrhs - j0 >= 2) {
        // q = H^2 p and pap = |H p|^2 for a batch of right-hand sides
        const int nb = nrhs - j0 >= 4 ? 4 : 2;
        Field *q[4], *p[4], *t1[4], *t2[4];
        double* rd[4];
        const int* sk[4];
        for (int b = 0; b < nb; b++) {
          Rhs& R = rh[j0 + b];
          q[b] = R.q; p[b] = R.p; t1[b] = R.t1; t2[b] = R.t2;
          rd[b] = &st[j0 + b].pap;
          sk[b] = &st[j0 + b].done;
        }
        e = apply_hsq_mr(c, PREC_F64, par->mass, nb, q, p, t1, t2, rd, sk);
        j0 += nb;
      }
      for (int j = j0; j < nrhs && !e; j++) {
        Rhs& R = rh[j];
        // q = H^2 p, and pap = |H p|^2 into st[j].pap via the hop reduction
        e = apply_op(c, PREC_F64, recon, YH_MVM_HPHI_SQ, par->mass, R.q, R.p, nullptr, 0, &st[j].done, &st[j].pap);
      }
      for (int j = 0; j < nrhs && !e; j++) {
        Rhs& R = rh[j];
        if ((e = ms_x(c, st + j, mf + j, R.r, R.q))) break;
        e = ms_k(c, st + j, mf + j, R.r, R.p);
      }
    }
    return e;
  };
  // The first chunk runs eagerly (it allocates the operator's scratch fields
  // and the clover inverse); later chunks replay a HIP graph of one chunk.
  // The chunk records no memset node.  If capture fails the solver stays eager.
  hipGraphExec_t exec = nullptr;
  bool graphs = c->use_graphs != 0;
  // iterations of each right-hand side already counted in mvm_counter
  std::vector<int> counted(nrhs, 0);
  while (!rc) {
    if (done_iters > 0 && graphs && !exec) {
      hipGraph_t gr = nullptr;
      hipError_t ge = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
      int crc = ge == hipSuccess ? enqueue_chunk() : YH_ERR_HIP;
      if (ge == hipSuccess) ge = hipStreamEndCapture(c->stream, &gr);
      if (ge == hipSuccess && crc == YH_OK) ge = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
      if (gr) (void)hipGraphDestroy(gr);
      if (ge != hipSuccess || crc != YH_OK || !exec) {
        (void)hipGetLastError();
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        graphs = false;
      }
    }
    if (exec) {
      he = hipGraphLaunch(exec, c->stream);
      if (he != hipSuccess) { set_error(hipGetErrorString(he)); rc = YH_ERR_HIP; break; }
    } else if ((rc = enqueue_chunk())) {
      break;
    }
    done_iters += chunk;
    he = hipMemcpyAsync(hs.data(), st, sizeof(MsState) * nrhs, hipMemcpyDeviceToHost, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) { set_error(hipGetErrorString(he)); rc = YH_ERR_HIP; break; }
    bool all = true;
    for (int j = 0; j < nrhs; j++) {
      if (!(hs[j].rr == hs[j].rr)) { set_error("cg_mshift: non-finite residual"); rc = YH_ERR_NONFINITE; }
      all = all && hs[j].done;
      // the operator ran once per iteration of this right-hand side (the
      // launches queued after it converged return at once)
      c->mvm_counter += (long)(hs[j].iters - counted[j]) * hops_per_apply(YH_MVM_HPHI_SQ);
      counted[j] = hs[j].iters;
    }
    if (rc || all) break;
    if (par->max_iter > 0 && done_iters >= par->max_iter) { set_error("cg_mshift: max_iter reached"); rc = YH_ERR_NOCONV; }
  }
  int itmax = 0;
  for (int j = 0; j < nrhs; j++) {
    if (iters) iters[j] = hs[j].iters;
    itmax = hs[j].iters > itmax ? hs[j].iters : itmax;
  }
  if (par->log && !rc)
    fprintf(stderr, "[INVERTER][NOTICE] cg_mshift (device, %d rhs, %d shifts): iterations = %d\n", nrhs, num, itmax);
  (void)hipStreamSynchronize(c->stream);
  if (exec) (void)hipGraphExecDestroy(exec);
  cleanup();
  (void)hipFree(st);
  (void)hipFree(mf);
  return rc;
}

}  // namespace yh
