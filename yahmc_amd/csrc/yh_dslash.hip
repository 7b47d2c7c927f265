// Dispatch of the hopping-term and clover kernels; the kernels themselves
// are instantiated per representation dimension in yh_dslash_d*.hip so
// that the large unrolled instantiations compile in parallel.
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
  }
  set_error("unsupported representation dimension");
  return YH_ERR_UNSUPPORTED;
}

}  // namespace yh
