// Propagators and meson contractions (observables/propagator.cpp:75-166,
// observables/mesons.cpp:55-147).
//
// calc_propagator_eo: full-lattice solve D x = b through the EO Schur
//   complement: s_e = b_e - D_eo D_oo^-1 b_o, x_e = D^-1 s_e (here with the
//   mixed-precision CG on H^2: x_e = (H^2)^-1 H g5 s_e), then
//   x_o = D_oo^-1 (b_o - D_oe x_e).
// Contractions: per site, for each of the 16 channels (Gamma1, Gamma2) given
//   as g5-multiplied permutation-phase tables (mesons.cpp:10-26), the point
//   formula sum_{a,c,x,y} g1[a] g2[c] S_{c,x}(g1col[a], y) conj(S_{g2col[c],x}(a, y))
//   or the wall formula sum_{a,A,x} g1[a] g2[A] S_A(g1col[a], x) conj(S_{g2col[A]}(a, x)),
//   real part; then summed per time slice in a fixed order (FP64).
#include <vector>
#include <mutex>
#include "yh_internal.hpp"

namespace yh {

enum { MES_CH = 16 };

struct MesTables {
  int col[MES_CH][2][4];
  double2 val[MES_CH][2][4];
};
__constant__ MesTables c_mes;

struct PropPtrs {
  const double2* p[48][2];  // [propagator][parity] FP64 halves
};

template <int D>
__device__ __forceinline__ double2 sget(const PropPtrs& P, int n, int par, int k, int VHpad, int spin, int colour) {
  return P.p[n][par][(size_t)(spin * D + colour) * VHpad + k];
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

template <int D, bool POINT>
__global__ void __launch_bounds__(256) k_meson_site(PropPtrs P, double* __restrict__ out, Geom g) {
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= g.V) return;
  const int par = id >= g.VH;
  const int k = id - par * g.VH;
  for (int ch = 0; ch < MES_CH; ch++) {
    double2 sum = make_double2(0.0, 0.0);
    for (int a = 0; a < 4; a++) {
      const int b = c_mes.col[ch][0][a];
      const double2 v1 = c_mes.val[ch][0][a];
      for (int c = 0; c < 4; c++) {
        const int d = c_mes.col[ch][1][c];
        const double2 gv = cmul(v1, c_mes.val[ch][1][c]);
        double2 acc = make_double2(0.0, 0.0);
        if constexpr (POINT) {
          for (int x = 0; x < D; x++)
#pragma unroll
            for (int y = 0; y < D; y++) {
              const double2 s1 = sget<D>(P, c * D + x, par, k, g.VHpad, b, y);
              const double2 s2 = sget<D>(P, d * D + x, par, k, g.VHpad, a, y);
              const double2 t = cmulc(s1, s2);
              acc.x += t.x;
              acc.y += t.y;
            }
        } else {
#pragma unroll
          for (int x = 0; x < D; x++) {
            const double2 s1 = sget<D>(P, c, par, k, g.VHpad, b, x);
            const double2 s2 = sget<D>(P, d, par, k, g.VHpad, a, x);
            const double2 t = cmulc(s1, s2);
            acc.x += t.x;
            acc.y += t.y;
          }
        }
        const double2 t = cmul(gv, acc);
        sum.x += t.x;
        sum.y += t.y;
      }
    }
    out[(size_t)ch * g.V + id] = sum.x;
  }
}

// one block per (channel, time slice): the slice's sites of both parities
// are two contiguous ranges of half indices
__global__ void __launch_bounds__(256) k_meson_tsum(const double* __restrict__ site, double* __restrict__ corr, int tau,
                                                    Geom g) {
  const int ch = blockIdx.y, t = blockIdx.x;
  const int T = g.L[0];
  const int hs = g.L[1] * g.L[2] * g.L[3] / 2;  // half sites per slice
  double v = 0.0;
  for (int i = threadIdx.x; i < 2 * hs; i += 256) {
    const int par = i >= hs;
    const int id = par * g.VH + t * hs + (i - par * hs);
    v += site[(size_t)ch * g.V + id];
  }
  __shared__ double sm[256];
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) corr[ch * T + (t - tau + T) % T] = sm[0];
}

int meson_init_tables() {
  // g5 * Gamma as (column, phase) per row (mesons.cpp:10-26); channel order
  // id, g1, g2, g3, g5, g5g1, g5g2, g5g3, g0g1, g0g2, g0g3, g0g5, g0g5g1,
  // g0g5g2, g0g5g3, (g0g5, g5)
  struct T4 { int col[4]; double re[4], im[4]; };
  const T4 ID = {{0, 1, 2, 3}, {1, 1, -1, -1}, {0, 0, 0, 0}};
  const T4 G1 = {{3, 2, 1, 0}, {0, 0, 0, 0}, {-1, -1, -1, -1}};
  const T4 G2 = {{3, 2, 1, 0}, {-1, 1, -1, 1}, {0, 0, 0, 0}};
  const T4 G3 = {{2, 3, 0, 1}, {0, 0, 0, 0}, {-1, 1, -1, 1}};
  const T4 G5 = {{0, 1, 2, 3}, {1, 1, 1, 1}, {0, 0, 0, 0}};
  const T4 G5G1 = {{3, 2, 1, 0}, {0, 0, 0, 0}, {-1, -1, 1, 1}};
  const T4 G5G2 = {{3, 2, 1, 0}, {-1, 1, 1, -1}, {0, 0, 0, 0}};
  const T4 G5G3 = {{2, 3, 0, 1}, {0, 0, 0, 0}, {-1, 1, 1, -1}};
  const T4 G0G1 = {{1, 0, 3, 2}, {0, 0, 0, 0}, {-1, -1, -1, -1}};
  const T4 G0G2 = {{1, 0, 3, 2}, {-1, 1, -1, 1}, {0, 0, 0, 0}};
  const T4 G0G3 = {{0, 1, 2, 3}, {0, 0, 0, 0}, {-1, 1, -1, 1}};
  const T4 G0G5 = {{2, 3, 0, 1}, {1, 1, 1, 1}, {0, 0, 0, 0}};
  const T4 G0G5G1 = {{1, 0, 3, 2}, {0, 0, 0, 0}, {1, 1, -1, -1}};
  const T4 G0G5G2 = {{1, 0, 3, 2}, {1, -1, -1, 1}, {0, 0, 0, 0}};
  const T4 G0G5G3 = {{0, 1, 2, 3}, {0, 0, 0, 0}, {1, -1, -1, 1}};
  const T4* pairs[MES_CH][2] = {{&ID, &ID},         {&G1, &G1},         {&G2, &G2},         {&G3, &G3},
                                {&G5, &G5},         {&G5G1, &G5G1},     {&G5G2, &G5G2},     {&G5G3, &G5G3},
                                {&G0G1, &G0G1},     {&G0G2, &G0G2},     {&G0G3, &G0G3},     {&G0G5, &G0G5},
                                {&G0G5G1, &G0G5G1}, {&G0G5G2, &G0G5G2}, {&G0G5G3, &G0G5G3}, {&G0G5, &G5}};
  MesTables h;
  for (int c = 0; c < MES_CH; c++)
    for (int w = 0; w < 2; w++)
      for (int a = 0; a < 4; a++) {
        h.col[c][w][a] = pairs[c][w]->col[a];
        h.val[c][w][a] = make_double2(pairs[c][w]->re[a], pairs[c][w]->im[a]);
      }
  // the tables do not depend on the lattice: uploaded once per device (the
  // current one), so a new context never rewrites memory that another
  // context's kernels may be reading
  static std::mutex mtx;
  static unsigned done_mask = 0;
  std::lock_guard<std::mutex> lk(mtx);
  int dev = 0;
  YH_HIP(hipGetDevice(&dev));
  if (dev < 32 && (done_mask >> dev) & 1u) return YH_OK;
  YH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_mes), &h, sizeof(h)));
  YH_HIP(hipDeviceSynchronize());
  if (dev < 32) done_mask |= 1u << dev;
  return YH_OK;
}

}  // namespace yh

using namespace yh;

extern "C" {

int yh_calc_propagator(yh_ctx* cc, const yh_inv_par* par, yh_field* dst, yh_field* src, int* iters) {
  Ctx* c = cc;
  if (!c || !par || !dst || !src) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (dst->prec != PREC_F64 || src->prec != PREC_F64) { set_error("fields must be FP64"); return YH_ERR_ARG; }
  int rc;
  Field *s = nullptr, *t = nullptr, *u = nullptr;
  if ((rc = field_alloc(c, PREC_F64, &s)) || (rc = field_alloc(c, PREC_F64, &t)) || (rc = field_alloc(c, PREC_F64, &u))) {
    field_free(s);
    field_free(t);
    return rc;
  }
  const double m = par->mass;
  int it = 0;
  auto body = [&]() -> int {
    int e;
    if (c->csw != 0.0) {
      // s_e = b_e - D_eo D_oo^-1 b_o  (propagator.cpp:92-97)
      if ((e = yh_field_copy(cc, static_cast<yh_field*>(t), src, YH_ODD))) return e;
      if ((e = yh_Dphi_oo_inv(cc, m, static_cast<yh_field*>(u), static_cast<yh_field*>(t)))) return e;
      if ((e = yh_Dphi_eo(cc, static_cast<yh_field*>(t), static_cast<yh_field*>(u)))) return e;
      if ((e = yh_field_copy(cc, static_cast<yh_field*>(s), src, YH_EVEN))) return e;
      if ((e = blas_axpby(c, s, -1.0, t, 1.0, 0))) return e;
    } else {
      // s_e = (4 + m) b_e - D_eo b_o, the source of the unnormalised
      // (4 + m)^2 - D_eo D_oe (propagator.cpp:125-129)
      if ((e = yh_field_copy(cc, static_cast<yh_field*>(t), src, YH_ODD))) return e;
      if ((e = yh_Dphi_eo(cc, static_cast<yh_field*>(u), static_cast<yh_field*>(t)))) return e;
      if ((e = yh_field_copy(cc, static_cast<yh_field*>(s), src, YH_EVEN))) return e;
      if ((e = blas_axpby(c, s, -1.0, u, 4.0 + m, 0))) return e;
    }
    // x_e = (H^2)^-1 H g5 s_e
    if ((e = blas_g5(c, s, 0))) return e;
    if ((e = yh_Hphi(cc, m, static_cast<yh_field*>(u), static_cast<yh_field*>(s)))) return e;
    if ((e = yh_field_zero(cc, dst, YH_BOTH))) return e;
    yh_inv_par p = *par;
    p.mvm = YH_MVM_HPHI_SQ;
    int upd;
    if ((e = yh_cg_mixed(cc, &p, dst, static_cast<yh_field*>(u), &it, &upd))) return e;
    // x_o = D_oo^-1 (b_o - D_oe x_e)
    if ((e = yh_Dphi_oe(cc, static_cast<yh_field*>(t), dst))) return e;
    if ((e = blas_axpby(c, t, 1.0, src, -1.0, 1))) return e;  // t_o = b_o - t_o
    if (c->csw != 0.0) {
      if ((e = yh_Dphi_oo_inv(cc, m, dst, static_cast<yh_field*>(t)))) return e;
    } else if ((e = blas_axpby(c, dst, 1.0 / (4.0 + m), t, 0.0, 1))) {  // D_oo = 4 + m
      return e;
    }
    return YH_OK;
  };
  rc = body();
  field_free(s);
  field_free(t);
  field_free(u);
  if (iters) *iters = it;
  return rc;
}

// Adds nothing: writes corr[16][T] for one source time slice tau.
// method 1 = wall (4 propagators, one per source spin), 2 = point
// (4 * D_R propagators, index spin * D_R + colour); all FP64, full lattice.
int yh_meson_contract(yh_ctx* cc, int method, int nprop, const yh_field* const* props, int tau, double* corr) {
  Ctx* c = cc;
  if (!c || !props || !corr) { set_error("null argument"); return YH_ERR_ARG; }
  const int D = c->dr;
  const int need = method == YH_MESON_POINT ? 4 * D : 4;
  if ((method != YH_MESON_POINT && method != YH_MESON_WALL) || nprop != need) {
    set_error("method must be YH_MESON_WALL (4 propagators) or YH_MESON_POINT (4 D_R propagators)");
    return YH_ERR_ARG;
  }
  const int T = c->g.L[0];
  if (tau < 0 || tau >= T) { set_error("tau out of range"); return YH_ERR_ARG; }
  PropPtrs P = {};
  for (int n = 0; n < nprop; n++) {
    if (!props[n] || props[n]->prec != PREC_F64) { set_error("propagators must be FP64 fields"); return YH_ERR_ARG; }
    P.p[n][0] = (const double2*)props[n]->half[0];
    P.p[n][1] = (const double2*)props[n]->half[1];
  }
  YH_HIP(hipSetDevice(c->device));
  double *site = nullptr, *dc = nullptr;
  YH_HIP(hipMalloc(&site, sizeof(double) * MES_CH * (size_t)c->g.V));
  hipError_t e = hipMalloc(&dc, sizeof(double) * MES_CH * T);
  if (e != hipSuccess) {
    (void)hipFree(site);
    set_error(hipGetErrorString(e));
    return YH_ERR_HIP;
  }
  dim3 grid((c->g.V + 255) / 256);
  int rc = YH_OK;
  const bool point = method == YH_MESON_POINT;
#define YH_MES(DD)                                                                                         \
  if (point) hipLaunchKernelGGL((k_meson_site<DD, true>), grid, dim3(256), 0, c->stream, P, site, c->g);   \
  else hipLaunchKernelGGL((k_meson_site<DD, false>), grid, dim3(256), 0, c->stream, P, site, c->g);
  switch (D) {
    case 2: YH_MES(2); break;
    case 3: YH_MES(3); break;
    case 4: YH_MES(4); break;
    case 6: YH_MES(6); break;
    case 8: YH_MES(8); break;
    default: set_error("unsupported representation dimension"); rc = YH_ERR_UNSUPPORTED;
  }
#undef YH_MES
  if (!rc) {
    hipLaunchKernelGGL(k_meson_tsum, dim3(T, MES_CH), dim3(256), 0, c->stream, site, dc, tau, c->g);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(corr, dc, sizeof(double) * MES_CH * T, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); rc = YH_ERR_HIP; }
  }
  (void)hipFree(site);
  (void)hipFree(dc);
  return rc;
}

}  // extern "C"
