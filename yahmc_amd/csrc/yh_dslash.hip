// Dispatch of the hopping-term and clover kernels; the kernels themselves
// are instantiated per representation dimension in yh_dslash_d*.hip so
// that the large unrolled instantiations compile in parallel.
#include <string.h>
#include "yh_dslash_impl.hpp"

namespace yh {

template <int D>
int hop_dispatch_d(Ctx* c, const HopArgs& a, int prec, int recon, int epi);
template <int D>
int clover_apply_d(Ctx* c, int prec, int parity, double diag, Field* out, const Field* in, int inverse);
extern template int hop_dispatch_d<2>(Ctx*, const HopArgs&, int, int, int);
extern template int hop_dispatch_d<3>(Ctx*, const HopArgs&, int, int, int);
extern template int hop_dispatch_d<4>(Ctx*, const HopArgs&, int, int, int);
extern template int hop_dispatch_d<6>(Ctx*, const HopArgs&, int, int, int);
extern template int clover_apply_d<2>(Ctx*, int, int, double, Field*, const Field*, int);
extern template int clover_apply_d<3>(Ctx*, int, int, double, Field*, const Field*, int);
extern template int clover_apply_d<4>(Ctx*, int, int, double, Field*, const Field*, int);
extern template int clover_apply_d<6>(Ctx*, int, int, double, Field*, const Field*, int);
extern template int hop_dispatch_d<8>(Ctx*, const HopArgs&, int, int, int);
extern template int clover_apply_d<8>(Ctx*, int, int, double, Field*, const Field*, int);
template <int D>
int hop_mr_dispatch_d(Ctx* c, const HopArgsM& m, int prec, int nr, int epi);
extern template int hop_mr_dispatch_d<2>(Ctx*, const HopArgsM&, int, int, int);
extern template int hop_mr_dispatch_d<3>(Ctx*, const HopArgsM&, int, int, int);
extern template int hop_mr_dispatch_d<4>(Ctx*, const HopArgsM&, int, int, int);
extern template int hop_mr_dispatch_d<6>(Ctx*, const HopArgsM&, int, int, int);

int hop_launch(Ctx* c, const HopCall& h) {
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  int p = h.out_parity;
  // RECON_12H reads the FP16 12-real link array under FP32 spinors
  void* links = h.recon == RECON_12H ? (h.prec == PREC_F32 ? c->links[PREC_F16][RECON_12] : nullptr)
                                     : c->links[h.prec][h.recon];
  if (!links) { set_error("links for this precision / reconstruction are not available"); return YH_ERR_UNSUPPORTED; }
  if (h.in->prec != h.prec || h.out->prec != h.prec) { set_error("field precision mismatch"); return YH_ERR_ARG; }
  HopArgs a;
  a.in = h.in->half[1 - p];
  a.in_nrm = h.in->nrm[1 - p];
  a.out = h.out->half[p];
  a.out_nrm = h.out->nrm[p];
  a.aux = h.aux ? h.aux->half[p] : nullptr;
  a.aux_nrm = h.aux ? h.aux->nrm[p] : nullptr;
  a.self = h.self ? h.self->half[p] : nullptr;
  a.self_nrm = h.self ? h.self->nrm[p] : nullptr;
  size_t lvbytes;
  int lnv;
  switch (h.recon == RECON_12H ? 5 : h.prec * 2 + h.recon) {
    case 0: lvbytes = 16; lnv = c->dr * c->dr; break;
    case 1: lvbytes = 16; lnv = 6; break;
    case 2: lvbytes = 8; lnv = c->dr * c->dr; break;
    case 3: lvbytes = 16; lnv = 3; break;
    case 4: lvbytes = 4; lnv = c->dr * c->dr; break;
    default: lvbytes = 8; lnv = 3; break;
  }
  size_t par_bytes = (size_t)4 * lnv * c->g.VHpad * lvbytes;
  a.lf = (const char*)links + (size_t)p * par_bytes;
  a.lb = (const char*)links + (size_t)(1 - p) * par_bytes;
  int N = 2 * c->dr;
  // the clover of an operator is stored at its spinor precision (16-bit:
  // fixed point with block scales)
  const int cprec = h.prec;
  size_t cvb = cprec == PREC_F16 ? sizeof(uint2) : 16;
  size_t cnv = cprec == PREC_F64 ? (size_t)N * N : (size_t)N * N / 2;  // vec per site (2 blocks)
  a.clov = nullptr;
  a.clov_nrm = nullptr;
  if (h.epi == EPI_CLOV) {
    if (!c->clov[cprec] || (cprec == PREC_F16 && !c->clov_nrm)) { set_error("clover term missing"); return YH_ERR_STATE; }
    a.clov = (const char*)c->clov[cprec] + (size_t)p * cnv * c->g.VHpad * cvb;
    if (cprec == PREC_F16) a.clov_nrm = c->clov_nrm + (size_t)p * 2 * c->g.VHpad;
  } else if (h.epi == EPI_CLOVINV) {
    if (!c->clinv[cprec] || p != 1 || (cprec == PREC_F16 && !c->clinv_nrm)) {
      set_error("clover inverse missing");
      return YH_ERR_STATE;
    }
    a.clov = c->clinv[cprec];
    a.clov_nrm = c->clinv_nrm;
  }
  a.scale = h.scale;
  a.diag = h.diag;
  a.parity = p;
  a.gamma5 = h.gamma5;
  a.zero_self_bc = h.zero_self_bc;
  a.reduce = h.reduce_norm;
  a.partials = c->partials;
  a.ticket = c->ticket;
  a.red_out = h.reduce_norm ? (h.reduce_ptr ? h.reduce_ptr : &c->sc->red[0] + h.reduce_slot) : nullptr;
  a.skip = h.skip_flag;
  a.g = c->g;
  if (c->xq_tab[h.prec][h.recon] >= 0) a.g.xq = c->xq_tab[h.prec][h.recon];
  // the reference counts every dirac_kernel call (dirac.cpp:233); a launch
  // with a skip flag may return at once (solver chunks queued past
  // convergence), so its solver counts the hops that ran from the device
  // iteration counter instead
  if (!h.skip_flag) c->mvm_counter++;
  switch (c->dr) {
    case 2: return hop_dispatch_d<2>(c, a, h.prec, h.recon, h.epi);
    case 3: return hop_dispatch_d<3>(c, a, h.prec, h.recon, h.epi);
    case 4: return hop_dispatch_d<4>(c, a, h.prec, h.recon, h.epi);
    case 6: return hop_dispatch_d<6>(c, a, h.prec, h.recon, h.epi);
    case 8: return hop_dispatch_d<8>(c, a, h.prec, h.recon, h.epi);
  }
  set_error("unsupported representation dimension");
  return YH_ERR_UNSUPPORTED;
}

// nr right-hand sides through one launch: h describes the launch (precision,
// parity, epilogue, scalars; its field pointers are ignored), in/out/self
// give the fields per right-hand side, red (CLOV with reduce_norm) the
// per-RHS destinations of |out|^2, skip the per-RHS stop flags (optional).
// FP64 / FP32 fields with 18-real links, CLOVINV or CLOV epilogues.  Not
// counted in mvm_counter: its callers are solvers that count the
// iterations that ran.
int hop_launch_mr(Ctx* c, const HopCall& h, int nr, const Field* const* in, Field* const* out, const Field* const* self,
                  double* const* red, const int* const* skip) {
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (nr != 2 && nr != 4) { set_error("multi-RHS hop: batch of 2 or 4"); return YH_ERR_ARG; }
  // the reduction keeps nr partial sums per workgroup of an nr-times larger grid
  const size_t nblk = ((size_t)c->g.VH * nr + YH_HOP_BLOCK - 1) / YH_HOP_BLOCK;
  if (h.reduce_norm && (size_t)nr * nblk > c->partials_cap) { set_error("multi-RHS hop: partials too small"); return YH_ERR_STATE; }
  if (h.recon != RECON_18 || (h.prec != PREC_F64 && h.prec != PREC_F32) || h.epi == EPI_PLAIN) {
    set_error("multi-RHS hop: FP64 / FP32, 18-real links, clover epilogues");
    return YH_ERR_UNSUPPORTED;
  }
  const int p = h.out_parity;
  void* links = c->links[h.prec][h.recon];
  if (!links) { set_error("links for this precision are not available"); return YH_ERR_UNSUPPORTED; }
  HopArgsM m;
  memset(&m, 0, sizeof(m));
  HopArgs& a = m.a;
  const size_t lvbytes = h.prec == PREC_F64 ? 16 : 8;
  const size_t par_bytes = (size_t)4 * c->dr * c->dr * c->g.VHpad * lvbytes;
  a.lf = (const char*)links + (size_t)p * par_bytes;
  a.lb = (const char*)links + (size_t)(1 - p) * par_bytes;
  const int N = 2 * c->dr;
  const size_t cnv = h.prec == PREC_F64 ? (size_t)N * N : (size_t)N * N / 2;
  if (h.epi == EPI_CLOV) {
    if (!c->clov[h.prec]) { set_error("clover term missing"); return YH_ERR_STATE; }
    a.clov = (const char*)c->clov[h.prec] + (size_t)p * cnv * c->g.VHpad * 16;
  } else {
    if (!c->clinv[h.prec] || p != 1) { set_error("clover inverse missing"); return YH_ERR_STATE; }
    a.clov = c->clinv[h.prec];
  }
  a.scale = h.scale;
  a.diag = h.diag;
  a.parity = p;
  a.gamma5 = h.gamma5;
  a.zero_self_bc = h.zero_self_bc;
  a.reduce = h.reduce_norm;
  a.partials = c->partials;
  a.ticket = c->ticket;
  a.g = c->g;
  if (c->xq_tab[h.prec][h.recon] >= 0) a.g.xq = c->xq_tab[h.prec][h.recon];
  for (int r = 0; r < nr; r++) {
    if (in[r]->prec != h.prec || out[r]->prec != h.prec) { set_error("field precision mismatch"); return YH_ERR_ARG; }
    m.in[r] = in[r]->half[1 - p];
    m.out[r] = out[r]->half[p];
    m.self[r] = self ? self[r]->half[p] : nullptr;
    m.red_out[r] = red ? red[r] : nullptr;
    m.skip[r] = skip ? skip[r] : nullptr;
  }
  switch (c->dr) {
    case 2: return hop_mr_dispatch_d<2>(c, m, h.prec, nr, h.epi);
    case 3: return hop_mr_dispatch_d<3>(c, m, h.prec, nr, h.epi);
    case 4: return hop_mr_dispatch_d<4>(c, m, h.prec, nr, h.epi);
    case 6: return hop_mr_dispatch_d<6>(c, m, h.prec, nr, h.epi);
  }
  set_error("unsupported representation dimension");
  return YH_ERR_UNSUPPORTED;
}

int clover_apply(Ctx* c, int prec, int parity, double diag, Field* out, const Field* in, int inverse) {
  if (in->prec != prec || out->prec != prec) { set_error("field precision mismatch"); return YH_ERR_ARG; }
  switch (c->dr) {
    case 2: return clover_apply_d<2>(c, prec, parity, diag, out, in, inverse);
    case 3: return clover_apply_d<3>(c, prec, parity, diag, out, in, inverse);
    case 4: return clover_apply_d<4>(c, prec, parity, diag, out, in, inverse);
    case 6: return clover_apply_d<6>(c, prec, parity, diag, out, in, inverse);
    case 8: return clover_apply_d<8>(c, prec, parity, diag, out, in, inverse);
  }
  set_error("unsupported representation dimension");
  return YH_ERR_UNSUPPORTED;
}

}  // namespace yh
