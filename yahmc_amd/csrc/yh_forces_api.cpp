// C ABI of the fermion forces (include/forces.h:6-15): argument checks and
// the reference's sequence of operator applications around the force
// kernels of yh_force.hip.
#include <hip/hip_runtime.h>
#include "yh_internal.hpp"

using namespace yh;

namespace {

int ready(Ctx* c) {
  if (!c) { set_error("null context"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  return YH_OK;
}

int sync(Ctx* c) {
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

// X_o <- D_oo^-1 D_oe X_e (forces.cpp:182-188), in place on the odd half
int prepare_odd(Ctx* c, double mass, Field* f) {
  HopCall h = {};
  h.prec = PREC_F64;
  h.recon = RECON_18;
  h.out_parity = 1;
  h.in = f;
  h.out = f;
  h.scale = -0.5;  // Dphi_oe: dirac_kernel sign +1
  h.epi = EPI_PLAIN;
  int rc = hop_launch(c, h);
  if (rc) return rc;
  if (c->csw != 0.0) {
    if ((rc = clover_invert(c, 4.0 + mass))) return rc;
    if ((rc = clover_apply(c, PREC_F64, 1, 4.0 + mass, f, f, 1))) return rc;
  }
  return YH_OK;
}

}  // namespace

extern "C" {

int yh_set_momenta(yh_ctx* c, const double* host) {
  int rc = ready(c);
  if (rc) return rc;
  if (!host) { set_error("null argument"); return YH_ERR_ARG; }
  return momenta_upload(c, host);
}

int yh_get_momenta(yh_ctx* c, double* host) {
  int rc = ready(c);
  if (rc) return rc;
  if (!host) { set_error("null argument"); return YH_ERR_ARG; }
  return momenta_download(c, host);
}

int yh_get_clover_force(yh_ctx* c, double* host) {
  int rc = ready(c);
  if (rc) return rc;
  if (!host) { set_error("null argument"); return YH_ERR_ARG; }
  return clover_force_download(c, host);
}

int yh_begin_fermion_force(yh_ctx* c) {
  int rc = ready(c);
  if (rc) return rc;
  if (c->csw == 0.0) return YH_OK;
  if ((rc = clover_force_zero(c))) return rc;
  return sync(c);
}

int yh_end_fermion_force(yh_ctx* c, double dt) {
  int rc = ready(c);
  if (rc) return rc;
  if (c->csw == 0.0) return YH_OK;
  if ((rc = force_clover_core(c, dt))) return rc;
  return sync(c);
}

int yh_clover_fermion_force(yh_ctx* c, double residue, const yh_field* lhs, const yh_field* rhs) {
  int rc = ready(c);
  if (rc) return rc;
  if (!lhs || !rhs || lhs->prec != PREC_F64 || rhs->prec != PREC_F64) { set_error("force fields must be FP64"); return YH_ERR_ARG; }
  if (c->csw == 0.0) return YH_OK;
  if ((rc = force_clover_site(c, residue, lhs, rhs))) return rc;
  return sync(c);
}

int yh_wilson_fermion_force(yh_ctx* c, double dt, double mass, double residue, yh_field* lhs, yh_field* rhs) {
  int rc = ready(c);
  if (rc) return rc;
  if (!lhs || !rhs || lhs->prec != PREC_F64 || rhs->prec != PREC_F64) { set_error("force fields must be FP64"); return YH_ERR_ARG; }
  double coeff = dt * residue;
  if ((rc = prepare_odd(c, mass, rhs))) return rc;
  if ((rc = prepare_odd(c, mass, lhs))) return rc;
  coeff = -coeff;
  if (c->csw != 0.0 && (rc = force_clover_site(c, residue, lhs, rhs))) return rc;
  if ((rc = force_hop(c, coeff, lhs, rhs))) return rc;
  return sync(c);
}

int yh_wilson_fermion_force_q(yh_ctx* c, double dt, double mass, double residue, yh_field* rhs) {
  int rc = ready(c);
  if (rc) return rc;
  if (!rhs || rhs->prec != PREC_F64) { set_error("force fields must be FP64"); return YH_ERR_ARG; }
  Field* tmp = tmp_field(c, PREC_F64, 7, &rc);
  if (rc) return rc;
  if ((rc = apply_op(c, PREC_F64, RECON_18, YH_MVM_HPHI, mass, tmp, rhs, rhs, -1, nullptr))) return rc;
  // the reference's Hphi(tmp, rhs) leaves the boundary slices of rhs zeroed
  // (dirac.cpp:700 on the in-place odd half), and the force's Dphi_oe then
  // reads that zeroed even half
  if ((rc = zero_bc_half(c, rhs, 0))) return rc;
  return yh_wilson_fermion_force(c, dt, mass, residue, static_cast<yh_field*>(tmp), rhs);
}

int yh_clover_logdet_force(yh_ctx* c, double mass, double residue) {
  int rc = ready(c);
  if (rc) return rc;
  if (c->csw == 0.0) return YH_OK;
  if ((rc = force_logdet(c, mass, residue))) return rc;
  return sync(c);
}

}  // extern "C"
