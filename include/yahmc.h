/*
 * yahmc.h — C ABI of the MI355X-native YaHMC fermion hot path.
 *
 * Plain pointers, sizes and integer status codes; no C++ or torch types.
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference repository root).
 *
 * Host-side field formats are the reference's own, so a caller converts
 * nothing:
 *   spinor field   V sites x Spinor{re[DR][4], im[DR][4]} doubles, in the
 *                  reference's EO site order, both parities
 *                  (include/spinor.h:7-9, include/spinorfield.h:6-39,
 *                  library/geometry.cpp:756-878)
 *   gauge field    4V suNg{re[NC*NC], im[NC*NC]} doubles, index 4*id+mu
 *                  (include/gaugefield.h:6-18)
 * Device fields live in HBM in the library's own layout (structure of
 * arrays, site-minor) and are referred to by opaque handles.
 *
 * All calls are synchronous with respect to the host unless stated, run on
 * the context's HIP stream, and return YH_OK (0) or a negative YH_ERR_*.
 * A context is not reentrant: one host thread per context.
 */
#ifndef YAHMC_H
#define YAHMC_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define YH_OK 0
#define YH_ERR_ARG -1        /* invalid argument */
#define YH_ERR_HIP -2        /* HIP runtime error (see yh_last_error) */
#define YH_ERR_STATE -3      /* call out of order (e.g. no gauge field) */
#define YH_ERR_NOCONV -4     /* solver hit max_iter without converging */
#define YH_ERR_NONFINITE -5  /* NaN or Inf in a solver scalar */
#define YH_ERR_UNSUPPORTED -6

/* representations: the reference's REPR_ID values (include/repr.h:9-39) */
#define YH_REPR_FUNDAMENTAL 0
#define YH_REPR_ADJOINT 1
#define YH_REPR_SYMMETRIC 2
#define YH_REPR_ANTISYMMETRIC 3

/* fermion boundary conditions in time (include/bcs.h:1-31) */
#define YH_BC_PERIODIC 0
#define YH_BC_ANTIPERIODIC 1
#define YH_BC_OPEN 2
#define YH_BC_SF 3

/* storage precisions of device spinor fields */
#define YH_PREC_F64 0
#define YH_PREC_F32 1
/* 16-bit storage: signed fixed point (value / 32768) of the components divided
 * by a per-site float scale for spinors and clover blocks; plain fixed point
 * for links (unitary, elements of modulus <= 1); arithmetic in FP32 */
#define YH_PREC_F16 2

/* parity masks (include/global.h:10-13) */
#define YH_EVEN 1
#define YH_ODD 2
#define YH_BOTH 3

/* operator selectors (the reference passes function pointers of type
 * mvm_operator, include/inverter.h:6) */
#define YH_MVM_DPHI 0
#define YH_MVM_HPHI 1
#define YH_MVM_HPHI_SQ 2

typedef struct yh_ctx yh_ctx;
typedef struct yh_field yh_field;

typedef struct {
  int dims[4];     /* T, X, Y, Z; all even (EO order needs Z even) */
  int nc;          /* colours: 2, 3 or 4 */
  int repr;        /* YH_REPR_* */
  int bc;          /* YH_BC_* */
  double csw;      /* clover coefficient; 0 disables the clover term */
  double cf;       /* boundary improvement c_f for open / SF */
  int device;      /* HIP device ordinal */
} yh_lattice;

/* solver parameters: mirrors inv_par (include/inverter.h:8-16) plus the
 * controls the reference lacks (iteration cap, precision ladder) */
typedef struct {
  double mass;
  double prec;        /* squared relative residual |r|^2 < prec |b|^2 */
  int max_iter;       /* 0 = no cap (the reference has none) */
  int mvm;            /* YH_MVM_* */
  int sloppy_prec;    /* mixed solvers: YH_PREC_F32 or YH_PREC_F16 */
  int reliable_every; /* mixed / flexible CG: FP64 reliable update at the latest
                        every N iterations (0 = 200 for the mixed CG, 50 for
                        it with YH_TUNE_HALF_LINKS and for the flexible CG) */
  double delta;       /* mixed CG: reliable update when |r| drops by delta */
  int log;
} yh_inv_par;

/* --- context / lattice ---------------------------------------------------- */
/* geometry_init + repr_init + clover_init + bcs_init
 * (library/geometry.cpp:1032-1191, library/repr.cpp:17-105,
 *  library/clover.cpp:349-361, library/bcs.cpp:11-137) */
int yh_create(const yh_lattice* lat, yh_ctx** out);
int yh_destroy(yh_ctx* ctx);
const char* yh_last_error(void);
int yh_version(void);
int yh_repr_dim(const yh_ctx* ctx);
long yh_volume(const yh_ctx* ctx);

/* --- gauge field ----------------------------------------------------------- */
/* upload gfield_fund then represent_gauge_field: represent the links,
 * build the clover term, apply the represented-field BCs
 * (library/repr.cpp:265-298, library/clover.cpp:330-347,
 *  library/bcs.cpp:264-272).  `links` is 4V suNg in the reference layout,
 * with the gauge-field BCs (bcs.cpp:201-262) already applied. */
int yh_set_gauge(yh_ctx* ctx, const double* links);
/* read back gfield_repr (with the antiperiodic sign) and cl_term in the
 * reference layouts (include/gaugefield.h:9-10, include/clover.h:12-22) */
int yh_get_gauge_repr(yh_ctx* ctx, double* out);
int yh_get_clover_term(yh_ctx* ctx, double* out);
/* clover_logdet (library/clover.cpp:281-305) */
int yh_clover_logdet(yh_ctx* ctx, double mass, double* out);

/* --- spinor fields (library/linalg.cpp:10-27, include/spinorfield.h) ------- */
int yh_field_create(yh_ctx* ctx, int prec, yh_field** out);
int yh_field_destroy(yh_field* f);
/* host <-> device in the reference Spinor layout; only the parities in
 * `parity_mask` are transferred (host array always holds all V sites) */
int yh_field_upload(yh_ctx* ctx, yh_field* f, const double* host, int parity_mask);
int yh_field_download(yh_ctx* ctx, const yh_field* f, double* host, int parity_mask);
/* spinor_zero (library/linalg.cpp:29-40) */
int yh_field_zero(yh_ctx* ctx, yh_field* f, int parity_mask);
/* spinor_copy with precision conversion (library/linalg.cpp:71-76) */
int yh_field_copy(yh_ctx* ctx, yh_field* dst, const yh_field* src, int parity_mask);

/* --- operators (include/dirac.h:6-16, library/dirac.cpp) ------------------- */
void yh_mvm_reset(yh_ctx* ctx);
int yh_mvm_get(yh_ctx* ctx);
int yh_Dphi_oe(yh_ctx* ctx, yh_field* dst, const yh_field* src);
int yh_Dphi_eo(yh_ctx* ctx, yh_field* dst, const yh_field* src);
int yh_Dphi_ee(yh_ctx* ctx, double mass, yh_field* dst, const yh_field* src);
int yh_Dphi_oo(yh_ctx* ctx, double mass, yh_field* dst, const yh_field* src);
int yh_Dphi_oo_inv(yh_ctx* ctx, double mass, yh_field* dst, const yh_field* src);
/* Dphi writes -D_oe src_e into the odd half of src, as the reference does
 * (library/dirac.cpp:738-751); Hphi = gamma5 Dphi; Hphi_sq = Hphi Hphi */
int yh_Dphi(yh_ctx* ctx, double mass, yh_field* dst, yh_field* src);
int yh_Hphi(yh_ctx* ctx, double mass, yh_field* dst, yh_field* src);
int yh_Hphi_sq(yh_ctx* ctx, double mass, yh_field* dst, yh_field* src);

/* --- linear algebra on the even half (include/linalg.h:7-34) --------------- */
int yh_spinor_sqnorm(yh_ctx* ctx, const yh_field* a, double* out);
int yh_spinor_product(yh_ctx* ctx, const yh_field* a, const yh_field* b, double* out_re_im);
int yh_spinor_product_re(yh_ctx* ctx, const yh_field* a, const yh_field* b, double* out);
int yh_spinor_g5(yh_ctx* ctx, yh_field* a);
int yh_spinor_mulr_add_assign(yh_ctx* ctx, yh_field* a, double r, const yh_field* b);
int yh_spinor_mulc_add_assign(yh_ctx* ctx, yh_field* a, double re, double im, const yh_field* b);

/* --- solvers (include/inverter.h:18-22) ------------------------------------ */
/* cg_inverter (library/cg.cpp:132-202): dst holds the initial guess */
int yh_cg(yh_ctx* ctx, const yh_inv_par* par, yh_field* dst, const yh_field* src, int* iters);
/* No counterpart in the reference: yh_cg without any allocation, for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The five temporary half
 * fields (5 x 64 * D_R * VHpad bytes, VHpad = 64 * ceil(V / 128) + 64, each
 * rounded up to 256 bytes) come from a library-owned device array of
 * YH_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a larger
 * lattice returns YH_ERR_ARG.  The gauge field must be set.  Calls that
 * overlap on one GPU (any contexts) share that array and race. */
#define YH_NOALLOC_SCRATCH_BYTES (32u << 20)
int yh_cg_noalloc(yh_ctx* ctx, const yh_inv_par* par, yh_field* dst, const yh_field* src, int* iters);
/* bicgstab_inverter (library/bicgstab.cpp:9-97); returns 2 x iterations.
 * dst holds the initial guess.  All scalars stay on the device. */
int yh_bicgstab(yh_ctx* ctx, const yh_inv_par* par, yh_field* dst, const yh_field* src, int* iters);
/* project-new: the same BiCGstab recurrence on FP32 fields and operator
 * inside FP64 reliable updates (x += x_sloppy, r = b - A x in FP64 when |r|
 * has dropped by par->delta or every par->reliable_every iterations, 100 if
 * 0); stops on the FP64 true residual |r|^2 < prec |b|^2.  iters = 2 x
 * sloppy iterations, updates = reliable updates */
int yh_bicgstab_mixed(yh_ctx* ctx, const yh_inv_par* par, yh_field* dst, const yh_field* src, int* iters, int* updates);
/* cg_mshift (library/cg.cpp:11-130): (mvm + shift[n]) dst[n] = src */
int yh_cg_mshift(yh_ctx* ctx, const yh_inv_par* par, int num, const double* shift, yh_field** dst,
                 const yh_field* src, int* iters);
/* minres_inverter (library/minres.cpp:12-118): Hphi or Hphi_sq; dst holds
 * the initial guess */
int yh_minres(yh_ctx* ctx, const yh_inv_par* par, yh_field* dst, const yh_field* src, int* iters);
/* auto_inverter (library/inverter.cpp:9-55): by par->mvm, BiCGstab on Dphi
 * (two solves with gamma5 for Hphi_sq, one for Hphi and Dphi); src is
 * restored on return */
int yh_auto_inverter(yh_ctx* ctx, const yh_inv_par* par, yh_field* dst, yh_field* src, int* iters);
/* MRE initial guess from past solutions (library/mre.cpp:15-117,
 * include/mre.h:10-21): mre_init / mre_guess / mre_store / mre_reset.
 * max = 0 disables it (guess and store do nothing), max is capped at 15. */
typedef struct yh_mre yh_mre;
int yh_mre_create(yh_ctx* ctx, int max, int mvm, double mass, yh_mre** out);
int yh_mre_destroy(yh_mre* mre);
int yh_mre_reset(yh_mre* mre);
int yh_mre_store(yh_mre* mre, const yh_field* x);
int yh_mre_guess(yh_mre* mre, yh_field* dst, const yh_field* src);
/* project-new: FP64 reliable-update CG with an FP32 / 16-bit inner operator;
 * reaches the same FP64 true residual as yh_cg.  iters = sloppy iterations,
 * updates = FP64 reliable updates */
int yh_cg_mixed(yh_ctx* ctx, const yh_inv_par* par, yh_field* dst, const yh_field* src, int* iters, int* updates);

/* batched multishift CG (project-new, generalises cg_mshift, cg.cpp:11-130):
 * nrhs independent right-hand sides src[j], num shifts each; solutions
 * dst[j * num + s] solve (H^2 + shift[s]) dst = src[j] (mvm must be
 * YH_MVM_HPHI_SQ); all recurrences run on the device; iters[j] per rhs */
int yh_cg_mshift_batched(yh_ctx* ctx, const yh_inv_par* par, int nrhs, int num, const double* shift, yh_field** dst,
                         const yh_field** src, int* iters);

/* --- rational approximations (include/rational.h:7-21) --------------------- */
/* same fields as rational_app: r(x) = a0 prod (x + root) / (x + pole)
 * = a0 + sum residue / (x + pole) approximates x^(n/d) on [min, max] */
typedef struct {
  int n;
  int d;
  int order;
  double prec;   /* maximum relative error on [min, max] */
  double min;
  double max;
  double a0;
  double pole[48];
  double root[48];
  double residue[48];
} yh_rational;
/* Zolotarev's optimal approximation of x^(+-1/2) of the given order
 * (replaces the table lookup of rational_find, rational.cpp:56-132) */
int yh_rational_zolotarev(int order, int n, int d, double min, double max, yh_rational* out);
/* best (minimax relative error) approximation of x^(n/d), 0 < |n/d| < 1,
 * of the given order by the Remez exchange, e.g. the x^(+-1/4) entries of
 * the reference's table (include/approx.h); positive powers are the
 * reciprocal of the negative one (rational.cpp:111-122) */
int yh_rational_remez(int order, int n, int d, double min, double max, yh_rational* out);
/* lowest order with maximum relative error below prec (rational_find,
 * rational.cpp:56-132): Zolotarev for x^(+-1/2), Remez otherwise */
int yh_rational_find(int n, int d, double prec, double min, double max, yh_rational* out);
double yh_rational_eval(const yh_rational* r, double x);
/* rational_calc (rational.cpp:134-166): dst = a0 src + sum residue_i
 * (H^2 + pole_i)^-1 src, through one multishift solve; FP64 sum */
int yh_rational_calc(yh_ctx* ctx, const yh_rational* r, const yh_inv_par* par, yh_field* dst, const yh_field* src,
                     int* iters);
/* Jacobi sn(u; k) and K(k), with k given through its complement
 * kc = sqrt(1 - k^2) (helpers of the Zolotarev construction) */
double yh_jacobi_sn(double u, double kc);
double yh_ellip_k(double kc);

/* --- forces (include/forces.h:6-15, library/forces.cpp) -------------------- */
/* momenta afield_momenta: 4V suNa{av[NC*NC-1]} doubles, index 4*id+mu
 * (include/gaugefield.h:12-18) */
int yh_set_momenta(yh_ctx* ctx, const double* host);
int yh_get_momenta(yh_ctx* ctx, double* host);
/* cl_force: 6V suNf, index 6*id+k, planes 01,02,12,03,13,23
 * (include/clover.h:15-16) */
int yh_get_clover_force(yh_ctx* ctx, double* host);
/* begin_fermion_force / end_fermion_force (forces.cpp:239-259) */
int yh_begin_fermion_force(yh_ctx* ctx);
int yh_end_fermion_force(yh_ctx* ctx, double dt);
/* clover_fermion_force (forces.cpp:122-172) */
int yh_clover_fermion_force(yh_ctx* ctx, double residue, const yh_field* lhs, const yh_field* rhs);
/* wilson_fermion_force (forces.cpp:175-227): overwrites the odd halves of
 * lhs and rhs with D_oo^-1 D_oe X_e as the reference does */
int yh_wilson_fermion_force(yh_ctx* ctx, double dt, double mass, double residue, yh_field* lhs, yh_field* rhs);
/* wilson_fermion_force_q (forces.cpp:230-236): overwrites the odd half of
 * rhs with D_oo^-1 D_oe X_e; for open and SF boundaries it also zeroes the
 * boundary time slices of rhs's even half, as the reference's in-place
 * Hphi(tmp, rhs) does (dirac.cpp:700) before the force reads it */
int yh_wilson_fermion_force_q(yh_ctx* ctx, double dt, double mass, double residue, yh_field* rhs);
/* clover_logdet_force (forces.cpp:109-119) */
int yh_clover_logdet_force(yh_ctx* ctx, double mass, double residue);

/* --- gauge molecular dynamics (FP64; open and SF boundaries as in
 * library/bcs.cpp: spatial boundary plaquettes weighted 1/2, the fixed
 * links' momenta zeroed before every link update and energy; SF for
 * NC = 2, 3 only) ------------------------------------------------------- */
/* the fundamental links (gfield_fund layout) as they are on the device */
int yh_get_gauge(yh_ctx* ctx, double* links);
/* represent_gauge_field (repr.cpp:265-280): rebuild the represented links
 * and the clover term from the device fundamental links */
int yh_represent_gauge_field(yh_ctx* ctx);
/* wilson_gauge_force (forces.cpp:262-283): mom -= proj(U staple, dt beta/NC) */
int yh_wilson_gauge_force(yh_ctx* ctx, double dt, double beta);
/* improved_gauge_force (forces.cpp:285-310): mom -= proj(c0 U S_plaq +
 * c1 U S_rect, dt beta/NC), S_rect the six weighted 1 x 2 rectangles through
 * the link for each nu != mu (staples_symanzik, staples.cpp:7-80); the
 * rectangle weights of the open and SF boundaries are bcs.cpp:30-31, 72-120 */
int yh_improved_gauge_force(yh_ctx* ctx, double dt, double beta, double c0, double c1);
/* update_links (update.cpp:228-247): U <- exp(eps P) U; the represented
 * links go stale until yh_represent_gauge_field, as set_gauge_field_state(true)
 * marks them in the reference */
int yh_update_links(yh_ctx* ctx, double eps);
/* reunitarize (update.cpp:309-320) */
int yh_reunitarize(yh_ctx* ctx);
/* avr_wilson(1,1) (wilsonloop.cpp:116-133): average plaquette, boundary
 * plaquettes weighted, normalised by NC 6 volume (volume = (T-1) V3 for open,
 * (T-2) V3 for SF boundaries, V otherwise) */
int yh_plaquette(yh_ctx* ctx, double* out);
/* avr_wilson(height, width) (wilsonloop.cpp:58-133): the average over sites
 * and planes mu > nu of Re Tr of the loop with height links along mu and
 * width along nu, plaquettes and 1 x 2 rectangles weighted by their boundary
 * weights, normalised like yh_plaquette; 1 <= height, width <= 16 */
int yh_avr_wilson(yh_ctx* ctx, int height, int width, double* out);
/* momentum part of hamiltonian (update.cpp:201-216): (TF/2) sum |P|^2,
 * after apply_bcs_on_momentum_field (bcs.cpp:166-199) */
int yh_momenta_energy(yh_ctx* ctx, double* out);
/* apply_bcs_on_gauge_field (bcs.cpp:201-262) on the device fundamental
 * links: open zeroes U_0 at t = T-1; SF sets the spatial links at t = 0 and
 * T-2 to their boundary matrices and zeroes the other links of t = T-2 and
 * T-1; periodic / antiperiodic leave the links alone */
int yh_apply_bcs_on_gauge_field(yh_ctx* ctx);

/* --- spectroscopy (observables/propagator.cpp, observables/mesons.cpp) ------ */
/* calc_propagator (propagator.cpp:75-166): full-lattice D dst = src through
 * the EO Schur complement; the even solve is the mixed-precision CG on H^2
 * to par->prec (par->mvm is ignored); both parities of src are read */
int yh_calc_propagator(yh_ctx* ctx, const yh_inv_par* par, yh_field* dst, yh_field* src, int* iters);
/* contract_all_correlators (mesons.cpp:55-147) for one source slice tau:
 * corr[16][T] (overwritten), channels id, g1, g2, g3, g5, g5g1, g5g2, g5g3,
 * g0g1, g0g2, g0g3, g0g5, g0g5g1, g0g5g2, g0g5g3, g5_g0g5; the time index is
 * (t - tau) mod T; no sign or normalisation (print_correlator applies them) */
#define YH_MESON_WALL 1  /* SEMWALL: 4 propagators, one per source spin */
#define YH_MESON_POINT 2 /* POINT: 4 D_R propagators, index spin * D_R + colour */
int yh_meson_contract(yh_ctx* ctx, int method, int nprop, const yh_field* const* props, int tau, double* corr);

/* --- gauge configuration files (host only; no device involved) ------------
 * YH_CFG_NATIVE: gauge_field_save / gauge_field_load (library/gaugefield.cpp:
 *   44-179; header_t of include/gaugefield.h:20-36, host byte order);
 * YH_CFG_HIREP: gauge_field_save_hirep / gauge_field_load_hirep
 *   (library/hirep.cpp:56-212, big-endian).
 * links: 4V suNg in the gfield_fund layout.  yh_config_read with links ==
 * NULL reads the header only (to size the array).  HiRep files carry no
 * representation (repr = -1) and take num_cnfg from the last "n%d" of the
 * file name, as the reference does. */
#define YH_CFG_NATIVE 0
#define YH_CFG_HIREP 1
typedef struct {
  int ident, version;
  int dims[4]; /* T, X, Y, Z */
  int nc, repr, num_cnfg, num_accept;
  double plaquette;
} yh_config_info;
int yh_config_write(const char* path, int format, const yh_config_info* info, const double* links);
int yh_config_read(const char* path, int format, yh_config_info* info, double* links);

/* --- learned preconditioner (no counterpart in the reference) --------------
 * A gauge-covariant parallel-transport convolution network on the even
 * sublattice, A = P_e L_n ... L_1 E with
 *   (L h)_o(x) = sum_q sum_i W_q[o][i] (T_q h_i)(x),
 * q = 0 identity, 1+mu: U(x,mu) h(x+mu), 5+mu: U(x-mu,mu)^+ h(x-mu) (the
 * Dirac operator's links), W_q[o][i] complex 4x4 spin matrices.  Weights of
 * layer l: [9][channels[l+1]][channels[l]][4][4] complex (re, im doubles),
 * layers concatenated.  The preconditioner is M = A A^+.  SU(3)
 * fundamental only; channels[0] = channels[n] = 1, hidden 4 or 8. */
typedef struct yh_net yh_net;
int yh_net_create(yh_ctx* ctx, int nlayers, const int* channels, yh_net** out);
int yh_net_destroy(yh_net* net);
long yh_net_num_weights(const yh_net* net);
int yh_net_set_weights(yh_net* net, const double* w);
int yh_net_get_weights(yh_net* net, double* w);
/* dst = A src (adjoint 0) or A^+ src on the even halves */
int yh_net_apply(yh_net* net, int adjoint, yh_field* dst, const yh_field* src);
/* dst = A A^+ src */
int yh_net_precondition(yh_net* net, yh_field* dst, const yh_field* src);
/* sum over pairs of |A A^+ b_k - x_k|^2 / |x_k|^2 and its gradient
 * d/dRe W + i d/dIm W (grad optional: nw complex) */
int yh_net_loss_grad(yh_net* net, int npairs, const yh_field* const* b, const yh_field* const* x, double* loss,
                     double* grad);
/* one Adam step with the mean gradient of the last yh_net_loss_grad */
int yh_net_adam_step(yh_net* net, double lr, double beta1, double beta2, double eps, int npairs);
/* single-layer access on host arrays [V][C][4][3] complex (testing) */
int yh_net_layer_host(yh_net* net, int layer, int adjoint, const double* in, double* out);
int yh_net_wgrad_host(yh_net* net, int layer, const double* g, const double* h, double* grad);
/* flexible CG on Hphi_sq in FP32, preconditioned by A A^+ (net may be NULL:
 * plain CG through the same path), inside FP64 reliable updates; stops on
 * the FP64 true residual |b - H^2 x|^2 < prec |b|^2 */
int yh_cg_net(yh_ctx* ctx, const yh_inv_par* par, yh_net* net, yh_field* dst, const yh_field* src, int* iters,
              int* updates);
/* flexible GCR on Hphi_sq in FP32 preconditioned by A A^+ (net may be NULL),
 * keeping the last `nkrylov` (1..32) directions, inside FP64 reliable
 * updates (which restart the direction history); stops on the FP64 true
 * residual |b - H^2 x|^2 < prec |b|^2 */
int yh_gcr_net(yh_ctx* ctx, const yh_inv_par* par, yh_net* net, int nkrylov, yh_field* dst, const yh_field* src,
               int* iters, int* updates);
/* The learned solve with its fallback: flexible CG preconditioned by `net`,
 * started from the MRE guess of `mre` (either may be NULL); if it stops
 * without converging (cap par->max_iter, 20000 when 0, or a non-finite
 * scalar) the plain mixed-precision CG (par->sloppy_prec, FP32 if unset)
 * restarts from the MRE guess with the caller's cap.  *fellback = 1 when the
 * fallback ran; iterations count both attempts.  The converged solution is
 * stored in the MRE history.  No counterpart in the reference (its callers
 * use auto_inverter + mre_guess, inverter.cpp:9-55, mre.cpp:54-85). */
int yh_cg_learned(yh_ctx* ctx, const yh_inv_par* par, yh_net* net, yh_mre* mre, yh_field* dst, const yh_field* src,
                  int* iters, int* updates, int* fellback);

/* --- measurement support ---------------------------------------------------- */
/* time `reps` back-to-back hop launches (Dphi_eo shape) of the given
 * precision / recon on the context stream with HIP events; returns the
 * average milliseconds per launch */
int yh_bench_hop(yh_ctx* ctx, int prec, int recon, int reps, yh_field* dst, const yh_field* src, double* ms);
/* time `reps` plain copies of `bytes` bytes with a vector copy kernel */
int yh_bench_copy(yh_ctx* ctx, long bytes, int reps, double* ms);
int yh_synchronize(yh_ctx* ctx);
/* tuning knobs of the kernels (no counterpart in the reference):
 * YH_TUNE_XCD_CHUNK = consecutive workgroups each XCD takes (0 = one slab
 * per XCD; default 1/8 of a t-slice) */
#define YH_TUNE_XCD_CHUNK 1
/* YH_TUNE_NET_PREFETCH = network layer kernels issue the next path's loads
 * before the current path's arithmetic (1) or not (0, default: the extra
 * registers cost more occupancy than the prefetch gains) */
#define YH_TUNE_NET_PREFETCH 2
/* YH_TUNE_GRAPHS = the multishift solver captures its 16-iteration chunk
 * in a HIP graph once and replays it (1, default) or launches eagerly (0) */
#define YH_TUNE_GRAPHS 3
/* 1: FP32 operators of SU(3) fundamental fields read the 16-bit 12-real links
 * (spinors and clover stay FP32): the sloppy operator of the mixed CG moves
 * 384 instead of 576 bytes per hop site */
#define YH_TUNE_HALF_LINKS 4
/* 1: FP64 operators of SU(3) fundamental fields read 12-real links (the third
 * row rebuilt as conj(r0 x r1)): 1,152 instead of 1,536 bytes per hop site;
 * 0 (default): all 18 reals */
#define YH_TUNE_RECON12_F64 5
/* 1 (default): the batched multishift applies H^2 to up to four right-hand
 * sides per hop launch, loading each link and clover block once for all of
 * them; 0: one hop launch per right-hand side */
#define YH_TUNE_MULTI_RHS 6
/* 1: yh_cg_learned runs flexible GCR (8 kept directions) instead of
 * flexible CG; 0 (default): flexible CG */
#define YH_TUNE_LEARNED_GCR 7
/* mask of storage precisions (bit YH_PREC_F64 / _F32 / _F16) whose hop
 * kernels (D_R <= 3) copy the window of input spinors their y / z
 * neighbours come from into LDS once per workgroup and read those
 * neighbours there (when a (t, x) plane holds a multiple of 256 sites);
 * default 3: FP64 and FP32 */
#define YH_TUNE_HOP_STAGE 8
/* order in which the clover-force kernel's workgroups take their (site
 * block, mu) work: -1 = all sites of one mu after another; q >= 0 = the
 * four mu of a site block together, in runs of q workgroups per XCD
 * (0, default = one contiguous slab of the lattice per XCD) */
#define YH_TUNE_FORCE_ORDER 9
int yh_tune(yh_ctx* ctx, int key, int value);
/* autotuner (no counterpart in the reference): times the hop of every
 * precision / reconstruction this context holds for each XCD chunk size
 * (`reps` launches each) and keeps the fastest per variant; chosen[6]
 * (optional) receives them, index prec * 2 + recon (-1: variant absent) */
int yh_autotune(yh_ctx* ctx, int reps, int* chosen);

#ifdef __cplusplus
}
#endif

#endif
