// Host-side context, field objects and the internal launch wrappers shared
// by the translation units of libyahmc.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string>
#include "yh_common.hpp"
#include "../../include/yahmc.h"

namespace yh {

#define YH_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));          \
      return YH_ERR_HIP;                                                     \
    }                                                                        \
  } while (0)

void set_error(const std::string& msg);

// a device spinor field: both parity halves, one precision
struct Field {
  int prec;
  void* half[2];     // even, odd
  float* nrm[2];     // FP16 per-site scale (null otherwise)
  void* mem;         // single allocation
  size_t bytes;
};

enum { RED_SLOTS = 64 };

// scalars of the device-resident solvers (FP64, one cache line each group)
struct SolverScalars {
  double rr;        // |r|^2 of the current iterate
  double rr_new;    // |r|^2 after the update
  double pap;       // <p, A p>
  double alpha, beta;
  double target;    // prec * |b|^2
  double norm_b;
  double maxrr;     // reliable-update reference norm: max |r|^2 since the last update
  double delta2;    // stop for a reliable update once |r|^2 < delta2 * maxrr (0: off)
  double best_rr;   // mixed CG: lowest true |r|^2 so far
  double rr_true;   // mixed CG: true |r|^2 of the last reliable update
  double red[RED_SLOTS];  // general reduction results
  int converged;    // stop flag: every solver kernel returns at once while set
  int iters;
  int reason;       // why `converged` was set: 1 target reached / non-finite, 2 reliable update due
  int finished;     // mixed CG: 1 converged, 2 non-finite, 3 stalled (the reliable-update kernels stop)
  int nupd;         // mixed CG: reliable updates that ran
  int stalls;       // mixed CG: reliable updates in a row that did not lower the true residual
  int pad[10];
};

struct Ctx {
  Geom g;
  int nc, dr, repr;
  double csw;
  int device;
  hipStream_t stream;
  int nblk;            // blocks of 256 threads over one half
  // gauge field: fundamental links in the reference layout (4V suNg)
  double* fund = nullptr;
  // represented links, SoA, per (prec, recon); null if not built
  void* links[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  // packed clover term (no mass) for both parities, per precision
  void* clov[3] = {nullptr, nullptr, nullptr};
  // packed inverse of the odd-site diagonal at sigma = cl_sigma
  void* clinv[3] = {nullptr, nullptr, nullptr};
  float* clov_nrm = nullptr;   // 16-bit clover block scales [parity][block][site]
  float* clinv_nrm = nullptr;  // 16-bit clover inverse block scales [block][site]
  double cl_sigma = 0xF00F;
  double* logdet_site = nullptr;  // per odd site log det (for the sum)
  bool gauge_set = false;
  // reductions
  double* partials = nullptr;      // per-workgroup partial sums of the grid reductions
  size_t partials_cap = 0;         // doubles in `partials`
  unsigned* ticket = nullptr;      // last-block counter
  SolverScalars* sc = nullptr;     // device scalars
  SolverScalars* sc_host = nullptr;  // pinned mirror
  // pinned snapshots of the scalars read back while later work is queued
  // (mixed CG), and the events that mark each copy complete
  enum { SC_RING = 4 };
  SolverScalars* sc_ring = nullptr;
  hipEvent_t ev_ring[SC_RING] = {};
  long mvm_counter = 0;
  // operator scratch (odd halves / even halves)
  Field* tmp[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // molecular dynamics: momenta (SoA, FP64) and clover force matrices
  double* mom = nullptr;
  void* clf = nullptr;
  // learned preconditioner: layer kernels prefetch the next path (tunable)
  int net_prefetch = 0;
  // per (precision, reconstruction) XCD chunk chosen by the autotuner (-1: g.xq)
  int xq_tab[3][3] = {{-1, -1, -1}, {-1, -1, -1}, {-1, -1, -1}};
  int half_links = 0;  // FP32 operators of SU(3) fundamental read FP16 12-real links
  int recon12_f64 = 0;  // FP64 operators of SU(3) fundamental read 12-real links
  int multi_rhs = 1;    // batched multishift: one hop launch per batch of right-hand sides
  int learned_gcr = 0;  // the learned solve uses flexible GCR instead of flexible CG
  // hop kernels (D_R <= 3) of the precisions in this mask (bit = precision)
  // read y / z neighbours from an LDS-staged window
  int hop_stage = (1 << PREC_F64) | (1 << PREC_F32);
  int force_order = 0;  // YH_TUNE_FORCE_ORDER
  // solver chunks are captured once into a HIP graph and replayed
  int use_graphs = 1;
  // model tables of this context's (nc, representation), in device memory
  // and passed to the kernels as arguments (two live contexts of different
  // models each keep their own): representation basis b_n, represented
  // generators iT^R_a and fundamental generators iT_a
  double2* tab = nullptr;
  double2* tab_basis = nullptr;  // TAB_BASIS entries
  double2* tab_itr = nullptr;    // TAB_ITR entries
  double2* tab_itf = nullptr;    // TAB_ITF entries
  // pinned host copy of the last reduction read back (2 doubles)
  double* red_host = nullptr;
  // caller-provided temporaries (the no-allocation entry point): when set,
  // apply_op uses op_tmp[0] (odd half) and op_tmp[1] (even half) and
  // cg_device uses cg_tmp[0..2] (r, p, q) instead of the context's own
  Field* op_tmp[2] = {nullptr, nullptr};
  Field* cg_tmp[3] = {nullptr, nullptr, nullptr};
};
enum { TAB_BASIS = 15 * 16, TAB_ITR = 15 * 64, TAB_ITF = 15 * 16 };
// copies `n` table entries to device memory on the context stream and waits
int upload_table(Ctx* c, double2* dst, const double2* host, size_t n);

// ---- fields ---------------------------------------------------------------
int field_alloc(Ctx* c, int prec, Field** out);
void field_free(Field* f);
size_t half_bytes(const Ctx* c, int prec);
int field_zero(Ctx* c, Field* f, int parity_mask);
int field_upload(Ctx* c, Field* f, const double* host, int parity_mask);
int field_download(Ctx* c, const Field* f, double* host, int parity_mask);
int field_convert(Ctx* c, Field* dst, int dpar, const Field* src, int spar);

// ---- gauge ------------------------------------------------------------------
int gauge_init_basis(Ctx* c);
int gauge_represent(Ctx* c);                 // fund -> F64 R18 links
int gauge_derive_links(Ctx* c);              // F64 R18 -> other precisions
int clover_build(Ctx* c);                    // packed FP64 clover
int clover_invert(Ctx* c, double sigma);     // odd-site inverses (+ logdet)
int clover_reserve_inverse(Ctx* c);          // allocate the inverse buffers once
int clover_logdet_sum(Ctx* c, double* out);
int gauge_download_repr(Ctx* c, double* host);    // reference gfield_repr layout, BC sign applied
int clover_download(Ctx* c, double* host);         // reference cl_term layout

// ---- dslash -----------------------------------------------------------------
enum { EPI_PLAIN = 0, EPI_CLOVINV = 1, EPI_CLOV = 2 };
struct HopCall {
  int prec, recon;
  int out_parity;
  const Field* in;        // reads half[1 - out_parity]
  Field* out;             // writes half[out_parity]
  Field* aux;             // EPI_CLOVINV: optional copy of the plain hop result (same half)
  const Field* self;      // EPI_CLOV: same-parity input for the clover term
  double scale;           // multiplies the hop sum
  double diag;            // 4 + m  (EPI_CLOV)
  int epi;
  int gamma5;             // EPI_CLOV: apply gamma5 to the result
  int zero_self_bc;       // EPI_CLOV: the reference zeroes boundary slices of the input
  int reduce_norm;        // EPI_CLOV: |out|^2 into sc->red[reduce_slot]
  int reduce_slot;
  double* reduce_ptr;     // if non-null, the reduction goes here instead of sc->red[reduce_slot]
  const int* skip_flag;   // if non-null and *skip_flag, every thread returns
};
int hop_launch(Ctx* c, const HopCall& h);
int hop_launch_mr(Ctx* c, const HopCall& h, int nr, const Field* const* in, Field* const* out, const Field* const* self,
                  double* const* red, const int* const* skip);
int clover_apply(Ctx* c, int prec, int parity, double diag, Field* out, const Field* in, int inverse);

// ---- blas -------------------------------------------------------------------
int blas_launch_reduce_finish(Ctx* c);
int blas_sqnorm(Ctx* c, const Field* a, int parity, int slot);
int blas_dot(Ctx* c, const Field* a, const Field* b, int parity, int slot);       // complex -> slot, slot+1
int blas_dot_re(Ctx* c, const Field* a, const Field* b, int parity, int slot);
int blas_axpby(Ctx* c, Field* y, double a, const Field* x, double b, int parity);  // y = a x + b y
int blas_caxpy(Ctx* c, Field* y, double ar, double ai, const Field* x, int parity); // y += (ar + i ai) x
int blas_cscale(Ctx* c, Field* y, double ar, double ai, int parity);
int blas_g5(Ctx* c, Field* f, int parity);
int blas_copy(Ctx* c, Field* dst, const Field* src, int parity);

// ---- forces -----------------------------------------------------------------
int force_init_constants(Ctx* c);
int force_alloc(Ctx* c);
int force_hop(Ctx* c, double coeff, const Field* lhs, const Field* rhs);
int force_clover_site(Ctx* c, double residue, const Field* lhs, const Field* rhs);
int force_logdet(Ctx* c, double mass, double coeff);
int force_clover_core(Ctx* c, double dt);
int clover_force_zero(Ctx* c);
int zero_bc_half(Ctx* c, Field* f, int parity);
int momenta_upload(Ctx* c, const double* host);
int momenta_download(Ctx* c, double* host);
int clover_force_download(Ctx* c, double* host);

// ---- gauge molecular dynamics ----------------------------------------------
int md_init_constants(Ctx* c);
int md_gauge_force(Ctx* c, double dt, double beta);
int md_improved_gauge_force(Ctx* c, double dt, double beta, double c0, double c1);
int md_avr_wilson(Ctx* c, int h, int w, double* out);
int md_update_links(Ctx* c, double eps);
int md_reunitarize(Ctx* c);
int md_plaquette(Ctx* c, double* out);
int md_kinetic(Ctx* c, double* out);

// ---- spectroscopy -----------------------------------------------------------
int meson_init_tables();

// ---- learned preconditioner ---------------------------------------------------
struct Net;
int net_precondition(Net* n, Field* dst, const Field* src);

// ---- host helpers shared by the solver translation units -------------------
Field* tmp_field(Ctx* c, int prec, int slot, int* rc);
int default_recon(Ctx* c, int prec);
int apply_op(Ctx* c, int prec, int recon, int mvm, double mass, Field* out, Field* in, Field* side, int reduce_slot,
             const int* skip, double* reduce_ptr = nullptr);
// Hphi_sq on nr right-hand sides (2 or 4) through multi-RHS hop launches:
// out[r] = H^2 in[r] with t1[r] (odd halves) and t2[r] (even halves) as
// temporaries; red[r] receives |H in[r]|^2; skip[r] stops right-hand side r
int apply_hsq_mr(Ctx* c, int prec, double mass, int nr, Field* const* out, Field* const* in, Field* const* t1,
                 Field* const* t2, double* const* red, const int* const* skip);
// hop launches of one apply_op (Hphi_sq: 4, Dphi / Hphi: 2)
inline int hops_per_apply(int mvm) { return mvm == YH_MVM_HPHI_SQ ? 4 : 2; }
int mshift_device(Ctx* c, const yh_inv_par* par, int nrhs, int num, const double* shift, Field** x, const Field** b,
                  int* iters);
int read_scalars(Ctx* c);
int write_scalars(Ctx* c);
int reduce_read(Ctx* c, int which, const Field* a, const Field* b, int par, double* out2);
// once per device (yh_create): device address of the no-allocation scratch array
int noalloc_init(int device);
int cg_device(Ctx* c, int prec, int recon, const yh_inv_par* par, Field* x, const Field* b, int* iters,
              double* final_rr, double* norm_b_out);
// BiCGstab with device-resident scalars: FP64 (sloppy = PREC_F64) or FP32
// inside FP64 reliable updates (sloppy = PREC_F32); iters = 2 x iterations
int bicgstab_device(Ctx* c, const yh_inv_par* par, int sloppy, Field* x, const Field* b, int* iters, int* updates);

// CG step kernels operating on device scalars (no host round trip)
int cg_update_xr(Ctx* c, Field* x, Field* r, const Field* p, const Field* q, int parity, int prec);
int cg_update_p(Ctx* c, Field* p, const Field* r, int parity, int prec);
// reliable update of the mixed CG on the device, both halves no-ops once
// sc->finished is set.  begin: complete the p = r + beta p step an iteration
// that raised the stop flag skipped (if sc->converged), x += xs, xs = 0.
// end: rs = b - t (t = A x in FP64), |b - t|^2; the last workgroup decides:
// finished (target reached, non-finite, or 10 updates without progress) or
// restart the sloppy recurrence (rr = maxrr = true |r|^2, stop flag cleared)
int ru_begin(Ctx* c, Field* x, Field* xs, Field* p, const Field* rs);
int ru_end(Ctx* c, const Field* b, const Field* t, Field* rs);

}  // namespace yh

// the opaque C handles are the internal objects
struct yh_ctx : public yh::Ctx {};
struct yh_field : public yh::Field {};
