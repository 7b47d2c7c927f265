// C ABI entry points: context, fields, gauge, operators, linear algebra and
// solvers.  Every call validates its arguments on the host before any
// kernel is launched (shapes, precisions, state), and returns a status.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>
#include "yh_internal.hpp"

namespace yh {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

int field_layout_kernels_upload(Ctx* c, Field* f, const double* dev_aos, int parity);
int field_layout_kernels_download(Ctx* c, const Field* f, double* dev_aos, int parity);
int copy_kernel_bench(Ctx* c, long bytes, int reps, double* ms);

static int repr_dim(int nc, int repr) {
  switch (repr) {
    case YH_REPR_FUNDAMENTAL: return nc;
    case YH_REPR_ADJOINT: return nc * nc - 1;
    case YH_REPR_SYMMETRIC: return nc * (nc + 1) / 2;
    case YH_REPR_ANTISYMMETRIC: return nc * (nc - 1) / 2;
  }
  return -1;
}

size_t half_bytes(const Ctx* c, int prec) {
  size_t per = prec == PREC_F64 ? 64 * c->dr : prec == PREC_F32 ? 32 * c->dr : 16 * c->dr;
  return per * c->g.VHpad;
}

int field_alloc(Ctx* c, int prec, Field** out) {
  if (prec < PREC_F64 || prec > PREC_F16) { set_error("bad precision"); return YH_ERR_ARG; }
  yh_field* f = new yh_field();
  f->prec = prec;
  size_t hb = half_bytes(c, prec);
  size_t nb = prec == PREC_F16 ? sizeof(float) * c->g.VHpad : 0;
  f->bytes = 2 * (hb + nb);
  hipError_t e = hipMalloc(&f->mem, f->bytes);
  if (e != hipSuccess) {
    delete f;
    set_error(std::string("field allocation: ") + hipGetErrorString(e));
    return YH_ERR_HIP;
  }
  char* m = (char*)f->mem;
  f->half[0] = m;
  f->half[1] = m + hb;
  f->nrm[0] = nb ? (float*)(m + 2 * hb) : nullptr;
  f->nrm[1] = nb ? (float*)(m + 2 * hb + nb) : nullptr;
  e = hipMemsetAsync(f->mem, 0, f->bytes, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    (void)hipFree(f->mem);
    delete f;
    set_error(hipGetErrorString(e));
    return YH_ERR_HIP;
  }
  *out = f;
  return YH_OK;
}

void field_free(Field* f) {
  if (!f) return;
  if (f->mem) (void)hipFree(f->mem);
  delete static_cast<yh_field*>(f);
}

int field_zero(Ctx* c, Field* f, int mask) {
  size_t hb = half_bytes(c, f->prec);
  for (int p = 0; p < 2; p++)
    if (mask & (1 << p)) {
      YH_HIP(hipMemsetAsync(f->half[p], 0, hb, c->stream));
      if (f->nrm[p]) YH_HIP(hipMemsetAsync(f->nrm[p], 0, sizeof(float) * c->g.VHpad, c->stream));
    }
  return YH_OK;
}

// staging buffer for host transfers (grown on demand)
static int staging(Ctx* c, double** out) {
  static std::mutex mtx;
  std::lock_guard<std::mutex> lk(mtx);
  size_t need = (size_t)c->g.V * 8 * c->dr * sizeof(double);
  Field* t = c->tmp[5];
  if (!t) {
    yh_field* f = new yh_field();
    memset((Field*)f, 0, sizeof(Field));
    YH_HIP(hipMalloc(&f->mem, need));
    f->bytes = need;
    c->tmp[5] = f;
    t = f;
  }
  *out = (double*)t->mem;
  return YH_OK;
}

int field_upload(Ctx* c, Field* f, const double* host, int mask) {
  double* st;
  int rc = staging(c, &st);
  if (rc) return rc;
  YH_HIP(hipMemcpyAsync(st, host, (size_t)c->g.V * 8 * c->dr * sizeof(double), hipMemcpyHostToDevice, c->stream));
  for (int p = 0; p < 2; p++)
    if (mask & (1 << p))
      if ((rc = field_layout_kernels_upload(c, f, st, p))) return rc;
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

int field_download(Ctx* c, const Field* f, double* host, int mask) {
  double* st;
  int rc = staging(c, &st);
  if (rc) return rc;
  size_t per_site = (size_t)8 * c->dr;
  for (int p = 0; p < 2; p++)
    if (mask & (1 << p))
      if ((rc = field_layout_kernels_download(c, f, st, p))) return rc;
  for (int p = 0; p < 2; p++)
    if (mask & (1 << p))
      YH_HIP(hipMemcpyAsync(host + (size_t)p * c->g.VH * per_site, st + (size_t)p * c->g.VH * per_site,
                            (size_t)c->g.VH * per_site * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

// ---------------------------------------------------------------------------
// operators
// ---------------------------------------------------------------------------
// per-context, per-precision scratch fields (slots 0, 1: operator
// temporaries; 2..7: solver vectors)
enum { NSCRATCH = 8 };
struct Scratch {
  Field* f[3][NSCRATCH] = {};
};
static std::vector<std::pair<Ctx*, Scratch*>> g_scratch;
static std::mutex g_scratch_mtx;

Field* tmp_field(Ctx* c, int prec, int i, int* rc) {
  std::lock_guard<std::mutex> lk(g_scratch_mtx);
  Scratch* s = nullptr;
  for (auto& e : g_scratch)
    if (e.first == c) s = e.second;
  if (!s) {
    s = new Scratch();
    g_scratch.push_back({c, s});
  }
  if (!s->f[prec][i]) {
    Field* f;
    *rc = field_alloc(c, prec, &f);
    if (*rc) return nullptr;
    s->f[prec][i] = f;
  }
  *rc = YH_OK;
  return s->f[prec][i];
}

static void free_scratch(Ctx* c) {
  std::lock_guard<std::mutex> lk(g_scratch_mtx);
  for (auto it = g_scratch.begin(); it != g_scratch.end(); ++it)
    if (it->first == c) {
      for (int p = 0; p < 3; p++)
        for (int i = 0; i < NSCRATCH; i++) field_free(it->second->f[p][i]);
      delete it->second;
      g_scratch.erase(it);
      return;
    }
}

int default_recon(Ctx* c, int prec) {
  // SU(3) fundamental sloppy operators use 12-real links; FP64 keeps the
  // full 18 reals unless YH_TUNE_RECON12_F64 asks for 12.
  if (c->nc == 3 && c->repr == YH_REPR_FUNDAMENTAL) {
    if (prec != PREC_F64) return prec == PREC_F32 && c->half_links ? RECON_12H : RECON_12;
    if (c->recon12_f64) return RECON_12;
  }
  return RECON_18;
}

// out_e = Op(in_e) for Op in {Dphi, Hphi, Hphi_sq} at precision `prec`.
// side: if non-null receives -D_oe in_e in its odd half (dirac.cpp:743).
// reduce_slot >= 0 (Hphi_sq only): |Hphi in|^2 into sc->red[reduce_slot]
int apply_op(Ctx* c, int prec, int recon, int mvm, double mass, Field* out, Field* in, Field* side, int reduce_slot,
             const int* skip, double* reduce_ptr) {
  int rc;
  const bool clover = c->csw != 0.0;
  if (clover && (rc = clover_invert(c, 4.0 + mass))) return rc;
  const bool own = c->op_tmp[0] && c->op_tmp[0]->prec == prec;
  Field* t1 = own ? c->op_tmp[0] : tmp_field(c, prec, 0, &rc);
  if (!own && rc) return rc;
  Field* t2 = nullptr;
  if (mvm == YH_MVM_HPHI_SQ) {
    t2 = own ? c->op_tmp[1] : tmp_field(c, prec, 1, &rc);
    if (!own && rc) return rc;
  }
  auto half_op = [&](Field* dst, Field* src, int g5, int red, Field* sidef) -> int {
    HopCall h1 = {};
    h1.prec = prec; h1.recon = recon; h1.out_parity = 1; h1.in = src; h1.out = t1;
    h1.aux = sidef; h1.scale = 0.5; h1.epi = clover ? EPI_CLOVINV : EPI_PLAIN; h1.skip_flag = skip;
    if (!clover && sidef) { h1.out = sidef; }
    int r = hop_launch(c, h1);
    if (r) return r;
    HopCall h2 = {};
    h2.prec = prec; h2.recon = recon; h2.out_parity = 0; h2.in = (!clover && sidef) ? sidef : t1; h2.out = dst;
    h2.self = src; h2.scale = -0.5; h2.epi = EPI_CLOV; h2.gamma5 = g5; h2.skip_flag = skip;
    h2.diag = clover ? 4.0 + mass : (4.0 + mass) * (4.0 + mass);
    h2.zero_self_bc = 1;
    h2.reduce_norm = red >= 0; h2.reduce_slot = red >= 0 ? red : 0; h2.reduce_ptr = reduce_ptr;
    return hop_launch(c, h2);
  };
  if (mvm == YH_MVM_DPHI) return half_op(out, in, 0, -1, side);
  if (mvm == YH_MVM_HPHI) return half_op(out, in, 1, -1, side);
  if ((rc = half_op(t2, in, 1, reduce_slot, side))) return rc;
  return half_op(out, t2, 1, -1, nullptr);
}

int apply_hsq_mr(Ctx* c, int prec, double mass, int nr, Field* const* out, Field* const* in, Field* const* t1,
                 Field* const* t2, double* const* red, const int* const* skip) {
  int rc;
  if (c->csw == 0.0) { set_error("multi-RHS operator needs the clover term"); return YH_ERR_UNSUPPORTED; }
  if ((rc = clover_invert(c, 4.0 + mass))) return rc;
  // the two halves of Hphi, each as in apply_op (dirac.cpp:738-781)
  auto half_op = [&](Field* const* dst, Field* const* src, double* const* rd) -> int {
    HopCall h1 = {};
    h1.prec = prec; h1.recon = RECON_18; h1.out_parity = 1; h1.scale = 0.5; h1.epi = EPI_CLOVINV;
    int r = hop_launch_mr(c, h1, nr, src, t1, nullptr, nullptr, skip);
    if (r) return r;
    HopCall h2 = {};
    h2.prec = prec; h2.recon = RECON_18; h2.out_parity = 0; h2.scale = -0.5; h2.epi = EPI_CLOV; h2.gamma5 = 1;
    h2.diag = 4.0 + mass; h2.zero_self_bc = 1; h2.reduce_norm = rd != nullptr;
    return hop_launch_mr(c, h2, nr, t1, dst, src, rd, skip);
  };
  if ((rc = half_op(t2, in, red))) return rc;
  return half_op(out, t2, nullptr);
}

// ---------------------------------------------------------------------------
// scalars
// ---------------------------------------------------------------------------
int upload_table(Ctx* c, double2* dst, const double2* host, size_t n) {
  YH_HIP(hipMemcpyAsync(dst, host, sizeof(double2) * n, hipMemcpyHostToDevice, c->stream));
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

int read_scalars(Ctx* c) {
  YH_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(SolverScalars), hipMemcpyDeviceToHost, c->stream));
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}
int write_scalars(Ctx* c) {
  YH_HIP(hipMemcpyAsync(c->sc, c->sc_host, sizeof(SolverScalars), hipMemcpyHostToDevice, c->stream));
  return YH_OK;
}
int reduce_read(Ctx* c, int which, const Field* a, const Field* b, int par, double* out2) {
  int rc;
  if (which == 0) rc = blas_sqnorm(c, a, par, 0);
  else if (which == 1) rc = blas_dot(c, a, b, par, 0);
  else rc = blas_dot_re(c, a, b, par, 0);
  if (rc) return rc;
  YH_HIP(hipMemcpyAsync(c->red_host, &c->sc->red[0], 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  YH_HIP(hipStreamSynchronize(c->stream));
  out2[0] = c->red_host[0];
  out2[1] = c->red_host[1];
  return YH_OK;
}

// ---------------------------------------------------------------------------
// solvers
// ---------------------------------------------------------------------------
// CG in one precision with device-resident scalars (library/cg.cpp:132-202).
// x holds the initial guess.  Chunks of `chunk` iterations are queued
// between host checks; iterations after convergence are no-ops.
int cg_device(Ctx* c, int prec, int recon, const yh_inv_par* par, Field* x, const Field* b, int* iters,
                     double* final_rr, double* norm_b_out) {
  int rc = YH_OK;
  const bool own = c->cg_tmp[0] && c->cg_tmp[0]->prec == prec;
  Field* r = own ? c->cg_tmp[0] : tmp_field(c, prec, 2, &rc);
  if (rc) return rc;
  Field* p = own ? c->cg_tmp[1] : tmp_field(c, prec, 3, &rc);
  if (rc) return rc;
  Field* qq = own ? c->cg_tmp[2] : tmp_field(c, prec, 4, &rc);
  if (rc) return rc;
  double norm_b[2];
  if ((rc = reduce_read(c, 0, b, nullptr, 0, norm_b))) return rc;
  if (norm_b_out) *norm_b_out = norm_b[0];
  // r = b - A x ; p = r
  if ((rc = blas_copy(c, p, x, 0))) return rc;
  if ((rc = apply_op(c, prec, recon, par->mvm, par->mass, qq, p, nullptr, -1, nullptr))) return rc;
  if ((rc = blas_copy(c, r, b, 0))) return rc;
  if ((rc = blas_axpby(c, r, -1.0, qq, 1.0, 0))) return rc;
  if ((rc = blas_copy(c, p, r, 0))) return rc;
  double rr[2];
  if ((rc = reduce_read(c, 0, r, nullptr, 0, rr))) return rc;
  SolverScalars* h = c->sc_host;
  memset(h, 0, sizeof(SolverScalars));
  h->rr = rr[0];
  h->target = norm_b[0] * par->prec;
  h->norm_b = norm_b[0];
  h->converged = rr[0] < h->target ? 1 : 0;
  if (h->converged) {
    *iters = 0;
    if (final_rr) *final_rr = rr[0];
    return YH_OK;
  }
  if ((rc = write_scalars(c))) return rc;
  const int chunk = 16;
  int done = 0;
  const bool sq = par->mvm == YH_MVM_HPHI_SQ;
  while (true) {
    // the last chunk stops at the iteration cap
    const int n = par->max_iter > 0 ? std::min(chunk, par->max_iter - done) : chunk;
    for (int i = 0; i < n; i++) {
      if ((rc = apply_op(c, prec, recon, par->mvm, par->mass, qq, p, nullptr, sq ? 0 : -1, &c->sc->converged))) return rc;
      if (!sq) {
        if ((rc = blas_dot_re(c, p, qq, 0, 0))) return rc;  // pap into red[0]
      }
      if ((rc = cg_update_xr(c, x, r, p, qq, 0, prec))) return rc;
      if ((rc = cg_update_p(c, p, r, 0, prec))) return rc;
    }
    if ((rc = read_scalars(c))) return rc;
    // every iteration that ran applied the operator once
    c->mvm_counter += (long)(h->iters - done) * hops_per_apply(par->mvm);
    done = h->iters;
    if (!(h->rr == h->rr)) { set_error("cg: non-finite residual"); return YH_ERR_NONFINITE; }
    if (h->converged) break;
    if (par->max_iter > 0 && done >= par->max_iter) {
      *iters = h->iters;
      if (final_rr) *final_rr = h->rr;
      set_error("cg: max_iter reached");
      return YH_ERR_NOCONV;
    }
  }
  *iters = h->iters;
  if (final_rr) *final_rr = h->rr;
  if (par->log) fprintf(stderr, "[INVERTER][NOTICE] cg: iterations = %d, precision = %1.4e\n", h->iters, h->rr / norm_b[0]);
  return YH_OK;
}

}  // namespace yh

using namespace yh;

extern "C" {

const char* yh_last_error(void) { return g_err.c_str(); }
int yh_version(void) { return 1; }

int yh_create(const yh_lattice* lat, yh_ctx** out) {
  if (!lat || !out) { set_error("null argument"); return YH_ERR_ARG; }
  for (int m = 0; m < 4; m++)
    if (lat->dims[m] < 2 || (lat->dims[m] & 1)) { set_error("every lattice extent must be even and >= 2"); return YH_ERR_ARG; }
  int dr = repr_dim(lat->nc, lat->repr);
  if (dr < 2 || !(dr == 2 || dr == 3 || dr == 4 || dr == 6 || dr == 8)) {
    set_error("unsupported (nc, representation)");
    return YH_ERR_UNSUPPORTED;
  }
  if (lat->bc < YH_BC_PERIODIC || lat->bc > YH_BC_SF) { set_error("bad boundary condition"); return YH_ERR_ARG; }
  if (lat->bc == YH_BC_SF && lat->dims[0] < 4) { set_error("SF needs T >= 4"); return YH_ERR_ARG; }
  long V = (long)lat->dims[0] * lat->dims[1] * lat->dims[2] * lat->dims[3];
  if (V > (1L << 30)) { set_error("volume too large for 32-bit site indexing"); return YH_ERR_ARG; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || lat->device < 0 || lat->device >= ndev) {
    set_error("no such HIP device");
    return YH_ERR_HIP;
  }
  YH_HIP(hipSetDevice(lat->device));
  yh_ctx* c = new yh_ctx();
  c->nc = lat->nc;
  c->repr = lat->repr;
  c->dr = dr;
  c->csw = lat->csw;
  c->device = lat->device;
  for (int m = 0; m < 4; m++) c->g.L[m] = lat->dims[m];
  c->g.V = (int)V;
  c->g.VH = (int)(V / 2);
  // pad the half-field stride off a power of two so the component streams
  // of one field do not land on the same HBM channels
  c->g.VHpad = ((c->g.VH + 63) / 64) * 64 + 64;
  c->g.bc = lat->bc;
  c->g.cf = lat->cf;
  c->g.Ta = c->g.Tb = -1;
  if (lat->bc == YH_BC_OPEN) { c->g.Ta = 1; c->g.Tb = lat->dims[0] - 2; }
  if (lat->bc == YH_BC_SF) { c->g.Ta = 1; c->g.Tb = lat->dims[0] - 3; }
  c->nblk = (c->g.VH + 255) / 256;
  // XCD chunk: each XCD takes 1/8 of a t-slice in turn, so the t +- 1
  // neighbours of its sites stay in its own L2 (best of a sweep at 32^4:
  // +10% over one slab per XCD, +15% over the hardware round robin)
  c->g.xq = std::max(1, (int)lround((double)c->g.VH / c->g.L[0] / (256.0 * 8.0)));
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  // partial sums: up to 8 values per workgroup of a one-site-per-thread
  // kernel, and up to 4 values per workgroup of the multi-RHS hop, whose grid
  // is 4 times larger
  c->partials_cap = (size_t)(c->nblk + 64) * 16;
  if (e == hipSuccess) e = hipMalloc(&c->partials, sizeof(double) * c->partials_cap);
  if (e == hipSuccess) e = hipMalloc(&c->ticket, sizeof(unsigned) * 64);
  if (e == hipSuccess) e = hipMemset(c->ticket, 0, sizeof(unsigned) * 64);
  if (e == hipSuccess) e = hipMalloc(&c->sc, sizeof(SolverScalars));
  if (e == hipSuccess) e = hipMemset(c->sc, 0, sizeof(SolverScalars));
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->sc_host, sizeof(SolverScalars), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->red_host, 2 * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&c->sc_ring, Ctx::SC_RING * sizeof(SolverScalars), hipHostMallocDefault);
  for (int i = 0; i < Ctx::SC_RING && e == hipSuccess; i++)
    e = hipEventCreateWithFlags(&c->ev_ring[i], hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&c->fund, sizeof(double) * 4 * (size_t)V * 2 * lat->nc * lat->nc);
  if (e == hipSuccess) e = hipMalloc(&c->tab, sizeof(double2) * (TAB_BASIS + TAB_ITR + TAB_ITF));
  if (e != hipSuccess) {
    set_error(std::string("context allocation: ") + hipGetErrorString(e));
    yh_destroy(c);
    return YH_ERR_HIP;
  }
  c->tab_basis = c->tab;
  c->tab_itr = c->tab + TAB_BASIS;
  c->tab_itf = c->tab + TAB_BASIS + TAB_ITR;
  int rc = noalloc_init(c->device);
  if (!rc) rc = gauge_init_basis(c);
  if (!rc) rc = force_init_constants(c);
  if (!rc) rc = md_init_constants(c);
  if (!rc) rc = meson_init_tables();
  if (rc) { yh_destroy(c); return rc; }
  *out = c;
  return YH_OK;
}

int yh_destroy(yh_ctx* c) {
  if (!c) return YH_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_scratch(c);
  for (int p = 0; p < 3; p++) {
    for (int r = 0; r < 2; r++)
      if (c->links[p][r]) (void)hipFree(c->links[p][r]);
    if (c->clov[p]) (void)hipFree(c->clov[p]);
    if (c->clinv[p]) (void)hipFree(c->clinv[p]);
  }
  {
    if (c->clov_nrm) (void)hipFree(c->clov_nrm);
    if (c->clinv_nrm) (void)hipFree(c->clinv_nrm);
  }
  for (int i = 4; i < 6; i++)
    if (c->tmp[i]) { (void)hipFree(c->tmp[i]->mem); delete static_cast<yh_field*>(c->tmp[i]); }
  if (c->fund) (void)hipFree(c->fund);
  if (c->tab) (void)hipFree(c->tab);
  if (c->logdet_site) (void)hipFree(c->logdet_site);
  if (c->mom) (void)hipFree(c->mom);
  if (c->clf) (void)hipFree(c->clf);
  if (c->partials) (void)hipFree(c->partials);
  if (c->ticket) (void)hipFree(c->ticket);
  if (c->sc) (void)hipFree(c->sc);
  if (c->sc_host) (void)hipHostFree(c->sc_host);
  if (c->red_host) (void)hipHostFree(c->red_host);
  if (c->sc_ring) (void)hipHostFree(c->sc_ring);
  for (int i = 0; i < Ctx::SC_RING; i++)
    if (c->ev_ring[i]) (void)hipEventDestroy(c->ev_ring[i]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return YH_OK;
}

int yh_repr_dim(const yh_ctx* c) { return c ? c->dr : -1; }
long yh_volume(const yh_ctx* c) { return c ? c->g.V : -1; }

int yh_set_gauge(yh_ctx* c, const double* links) {
  if (!c || !links) { set_error("null argument"); return YH_ERR_ARG; }
  YH_HIP(hipSetDevice(c->device));
  size_t n = (size_t)4 * c->g.V * 2 * c->nc * c->nc;
  YH_HIP(hipMemcpyAsync(c->fund, links, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int rc;
  if ((rc = gauge_represent(c))) return rc;
  if ((rc = gauge_derive_links(c))) return rc;
  if ((rc = clover_build(c))) return rc;
  if (c->csw != 0.0 && (rc = clover_reserve_inverse(c))) return rc;
  YH_HIP(hipStreamSynchronize(c->stream));
  c->gauge_set = true;
  return YH_OK;
}

int yh_get_gauge_repr(yh_ctx* c, double* out) {
  if (!c || !out) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  return gauge_download_repr(c, out);
}

int yh_get_clover_term(yh_ctx* c, double* out) {
  if (!c || !out) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  return clover_download(c, out);
}

int yh_clover_logdet(yh_ctx* c, double mass, double* out) {
  if (!c || !out) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (c->csw == 0.0) { *out = 0.0; return YH_OK; }  // clover.cpp:283-286
  int rc = clover_invert(c, 4.0 + mass);
  if (rc) return rc;
  return clover_logdet_sum(c, out);
}

int yh_field_create(yh_ctx* c, int prec, yh_field** out) {
  if (!c || !out) { set_error("null argument"); return YH_ERR_ARG; }
  Field* f;
  int rc = field_alloc(c, prec, &f);
  if (rc) return rc;
  *out = static_cast<yh_field*>(f);
  return YH_OK;
}

int yh_field_destroy(yh_field* f) {
  field_free(f);
  return YH_OK;
}

static int mask_ok(int m) { return m >= 1 && m <= 3; }

int yh_field_upload(yh_ctx* c, yh_field* f, const double* host, int mask) {
  if (!c || !f || !host || !mask_ok(mask)) { set_error("bad argument"); return YH_ERR_ARG; }
  return field_upload(c, f, host, mask);
}

int yh_field_download(yh_ctx* c, const yh_field* f, double* host, int mask) {
  if (!c || !f || !host || !mask_ok(mask)) { set_error("bad argument"); return YH_ERR_ARG; }
  return field_download(c, f, host, mask);
}

int yh_field_zero(yh_ctx* c, yh_field* f, int mask) {
  if (!c || !f || !mask_ok(mask)) { set_error("bad argument"); return YH_ERR_ARG; }
  int rc = field_zero(c, f, mask);
  if (rc) return rc;
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

int yh_field_copy(yh_ctx* c, yh_field* dst, const yh_field* src, int mask) {
  if (!c || !dst || !src || !mask_ok(mask)) { set_error("bad argument"); return YH_ERR_ARG; }
  for (int p = 0; p < 2; p++)
    if (mask & (1 << p)) {
      int rc = field_convert(c, dst, p, src, p);
      if (rc) return rc;
    }
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

void yh_mvm_reset(yh_ctx* c) {
  if (c) c->mvm_counter = 0;
}
int yh_mvm_get(yh_ctx* c) { return c ? (int)(c->mvm_counter >> 1) : -1; }

static int op_check(yh_ctx* c, const yh_field* d, const yh_field* s) {
  if (!c || !d || !s) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (d->prec != s->prec) { set_error("field precision mismatch"); return YH_ERR_ARG; }
  return YH_OK;
}

static int hop_plain(yh_ctx* c, yh_field* d, const yh_field* s, int out_parity) {
  int rc = op_check(c, d, s);
  if (rc) return rc;
  HopCall h = {};
  h.prec = d->prec;
  h.recon = default_recon(c, d->prec);
  h.out_parity = out_parity;
  h.in = s;
  h.out = d;
  h.scale = -0.5;  // dirac_kernel with sign +1 (dirac.cpp:234, 265-275)
  h.epi = EPI_PLAIN;
  if ((rc = hop_launch(c, h))) return rc;
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

int yh_Dphi_oe(yh_ctx* c, yh_field* d, const yh_field* s) { return hop_plain(c, d, s, 1); }
int yh_Dphi_eo(yh_ctx* c, yh_field* d, const yh_field* s) { return hop_plain(c, d, s, 0); }

static int clov_plain(yh_ctx* c, double mass, yh_field* d, const yh_field* s, int parity, int inverse) {
  int rc = op_check(c, d, s);
  if (rc) return rc;
  if (inverse && (rc = clover_invert(c, 4.0 + mass))) return rc;
  if ((rc = clover_apply(c, d->prec, parity, 4.0 + mass, d, s, inverse))) return rc;
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

int yh_Dphi_ee(yh_ctx* c, double m, yh_field* d, const yh_field* s) { return clov_plain(c, m, d, s, 0, 0); }
int yh_Dphi_oo(yh_ctx* c, double m, yh_field* d, const yh_field* s) { return clov_plain(c, m, d, s, 1, 0); }
int yh_Dphi_oo_inv(yh_ctx* c, double m, yh_field* d, const yh_field* s) {
  if (c && c->csw == 0.0) { set_error("Dphi_oo_inv needs the clover term"); return YH_ERR_STATE; }
  return clov_plain(c, m, d, s, 1, 1);
}

static int op_public(yh_ctx* c, int mvm, double mass, yh_field* d, yh_field* s) {
  int rc = op_check(c, d, s);
  if (rc) return rc;
  if (d == s) { set_error("dst and src must differ"); return YH_ERR_ARG; }
  // the odd half of src receives -D_oe src_e (dirac.cpp:743)
  if ((rc = apply_op(c, d->prec, default_recon(c, d->prec), mvm, mass, d, s, s, -1, nullptr))) return rc;
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

int yh_Dphi(yh_ctx* c, double m, yh_field* d, yh_field* s) { return op_public(c, YH_MVM_DPHI, m, d, s); }
int yh_Hphi(yh_ctx* c, double m, yh_field* d, yh_field* s) { return op_public(c, YH_MVM_HPHI, m, d, s); }
int yh_Hphi_sq(yh_ctx* c, double m, yh_field* d, yh_field* s) { return op_public(c, YH_MVM_HPHI_SQ, m, d, s); }

int yh_spinor_sqnorm(yh_ctx* c, const yh_field* a, double* out) {
  if (!c || !a || !out) { set_error("null argument"); return YH_ERR_ARG; }
  double v[2];
  int rc = reduce_read(c, 0, a, nullptr, 0, v);
  *out = v[0];
  return rc;
}

int yh_spinor_product(yh_ctx* c, const yh_field* a, const yh_field* b, double* out) {
  if (!c || !a || !b || !out) { set_error("null argument"); return YH_ERR_ARG; }
  return reduce_read(c, 1, a, b, 0, out);
}

int yh_spinor_product_re(yh_ctx* c, const yh_field* a, const yh_field* b, double* out) {
  if (!c || !a || !b || !out) { set_error("null argument"); return YH_ERR_ARG; }
  double v[2];
  int rc = reduce_read(c, 2, a, b, 0, v);
  *out = v[0];
  return rc;
}

int yh_spinor_g5(yh_ctx* c, yh_field* a) {
  if (!c || !a) { set_error("null argument"); return YH_ERR_ARG; }
  int rc = blas_g5(c, a, 0);
  if (rc) return rc;
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

int yh_spinor_mulr_add_assign(yh_ctx* c, yh_field* a, double r, const yh_field* b) {
  if (!c || !a || !b) { set_error("null argument"); return YH_ERR_ARG; }
  int rc = blas_axpby(c, a, r, b, 1.0, 0);
  if (rc) return rc;
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

int yh_spinor_mulc_add_assign(yh_ctx* c, yh_field* a, double re, double im, const yh_field* b) {
  if (!c || !a || !b) { set_error("null argument"); return YH_ERR_ARG; }
  int rc = blas_caxpy(c, a, re, im, b, 0);
  if (rc) return rc;
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

static int par_check(yh_ctx* c, const yh_inv_par* par, const yh_field* d, const yh_field* s, int* iters) {
  if (!c || !par || !d || !s || !iters) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (par->mvm < YH_MVM_DPHI || par->mvm > YH_MVM_HPHI_SQ) { set_error("bad mvm"); return YH_ERR_ARG; }
  if (!(par->prec > 0.0)) { set_error("prec must be positive"); return YH_ERR_ARG; }
  if (d->prec != PREC_F64 || s->prec != PREC_F64) { set_error("solver fields must be FP64"); return YH_ERR_ARG; }
  return YH_OK;
}

int yh_cg(yh_ctx* c, const yh_inv_par* par, yh_field* dst, const yh_field* src, int* iters) {
  int rc = par_check(c, par, dst, src, iters);
  if (rc) return rc;
  return cg_device(c, PREC_F64, default_recon(c, PREC_F64), par, dst, src, iters, nullptr, nullptr);
}

int yh_tune(yh_ctx* c, int key, int value) {
  if (!c) { set_error("null context"); return YH_ERR_ARG; }
  if (key == YH_TUNE_XCD_CHUNK) {
    if (value < 0) { set_error("chunk must be >= 0"); return YH_ERR_ARG; }
    c->g.xq = value;
    return YH_OK;
  }
  if (key == YH_TUNE_NET_PREFETCH) {
    c->net_prefetch = value != 0;
    return YH_OK;
  }
  if (key == YH_TUNE_GRAPHS) {
    c->use_graphs = value != 0;
    return YH_OK;
  }
  if (key == YH_TUNE_HALF_LINKS) {
    c->half_links = value != 0;
    return YH_OK;
  }
  if (key == YH_TUNE_LEARNED_GCR) {
    c->learned_gcr = value != 0;
    return YH_OK;
  }
  if (key == YH_TUNE_HOP_STAGE) {
    if (value < 0 || value > 7) { set_error("hop staging mask: bits 0 (FP64), 1 (FP32), 2 (16-bit)"); return YH_ERR_ARG; }
    c->hop_stage = value;
    return YH_OK;
  }
  if (key == YH_TUNE_FORCE_ORDER) {
    if (value < -1) { set_error("force order: -1, or a run length >= 0"); return YH_ERR_ARG; }
    c->force_order = value;
    return YH_OK;
  }
  if (key == YH_TUNE_MULTI_RHS) {
    c->multi_rhs = value != 0;
    return YH_OK;
  }
  if (key == YH_TUNE_RECON12_F64) {
    if (value && !(c->nc == 3 && c->repr == YH_REPR_FUNDAMENTAL)) {
      set_error("12-real links need SU(3) fundamental");
      return YH_ERR_UNSUPPORTED;
    }
    c->recon12_f64 = value != 0;
    return YH_OK;
  }
  set_error("unknown tuning key");
  return YH_ERR_ARG;
}

int yh_synchronize(yh_ctx* c) {
  if (!c) return YH_ERR_ARG;
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

int yh_bench_hop(yh_ctx* c, int prec, int recon, int reps, yh_field* d, const yh_field* s, double* ms) {
  int rc = op_check(c, d, s);
  if (rc) return rc;
  if (!ms || reps < 1 || d->prec != prec) { set_error("bad argument"); return YH_ERR_ARG; }
  HopCall h = {};
  h.prec = prec; h.recon = recon; h.out_parity = 0; h.in = s; h.out = d; h.scale = -0.5; h.epi = EPI_PLAIN;
  if ((rc = hop_launch(c, h))) return rc;  // warm
  hipEvent_t e0, e1;
  YH_HIP(hipEventCreate(&e0));
  YH_HIP(hipEventCreate(&e1));
  YH_HIP(hipEventRecord(e0, c->stream));
  for (int i = 0; i < reps; i++)
    if ((rc = hop_launch(c, h))) return rc;
  YH_HIP(hipEventRecord(e1, c->stream));
  YH_HIP(hipEventSynchronize(e1));
  float t;
  YH_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t / reps;
  return YH_OK;
}

// Online autotuner: for every hop variant this context can run, time the
// plain hop over the even sites for a set of XCD chunk sizes and keep the
// fastest (the block -> XCD tiling is the shape that matters for this
// HBM-bound stencil: it sets which t-slices share an XCD's L2).
int yh_autotune(yh_ctx* c, int reps, int* chosen) {
  if (!c || reps < 1) { set_error("bad argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  YH_HIP(hipSetDevice(c->device));
  const int nblocks = (c->g.VH + 255) / 256;
  std::vector<int> cand = {0};
  for (int q = 1; 8 * q <= nblocks; q *= 2) cand.push_back(q);
  for (int prec = 0; prec < 3; prec++)
    for (int recon = 0; recon < 2; recon++) {
      if (chosen) chosen[prec * 2 + recon] = -1;
      if (!c->links[prec][recon]) continue;
      int rc;
      Field *s = nullptr, *d = nullptr;
      if ((rc = field_alloc(c, prec, &s)) || (rc = field_alloc(c, prec, &d))) {
        field_free(s);
        return rc;
      }
      double best = 1e30;
      int bq = -1;
      for (int q : cand) {
        c->xq_tab[prec][recon] = q;
        double ms;
        rc = yh_bench_hop(static_cast<yh_ctx*>(c), prec, recon, reps, static_cast<yh_field*>(d),
                          static_cast<const yh_field*>(s), &ms);
        if (rc) break;
        if (ms < best) { best = ms; bq = q; }
      }
      field_free(s);
      field_free(d);
      if (rc) { c->xq_tab[prec][recon] = -1; return rc; }
      c->xq_tab[prec][recon] = bq;
      if (chosen) chosen[prec * 2 + recon] = bq;
    }
  return YH_OK;
}

int yh_bench_copy(yh_ctx* c, long bytes, int reps, double* ms) {
  if (!c || !ms || bytes <= 0 || reps < 1) { set_error("bad argument"); return YH_ERR_ARG; }
  return copy_kernel_bench(c, bytes, reps, ms);
}

}  // extern "C"
