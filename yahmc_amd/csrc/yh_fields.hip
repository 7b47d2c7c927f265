// Host-format conversion of spinor fields (reference AoS Spinor layout <->
// device SoA), the clover log-det sum and the copy kernel used to
// calibrate HBM bandwidth.
#include "yh_internal.hpp"
#include "yh_reduce.hpp"

namespace yh {

// reference layout per site: re[D][4] then im[D][4] (include/spinor.h:7-9)
template <int P, int D>
__global__ void __launch_bounds__(256) k_from_aos(void* out, float* out_nrm, const double* __restrict__ aos, int parity,
                                                  Geom g) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= g.VH) return;
  const double* s = aos + ((size_t)parity * g.VH + k) * 8 * D;
  Cx<double> v[4][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int sp = 0; sp < 4; sp++) v[sp][i] = Cx<double>(s[i * 4 + sp], s[4 * D + i * 4 + sp]);
  store_spinor<P, D, double>(out, out_nrm, k, g.VHpad, v);
}

template <int P, int D>
__global__ void __launch_bounds__(256) k_to_aos(const void* in, const float* in_nrm, double* __restrict__ aos, int parity,
                                                Geom g) {
  int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= g.VH) return;
  double* s = aos + ((size_t)parity * g.VH + k) * 8 * D;
  Cx<double> v[4][D];
  load_spinor<P, D, double>(v, in, in_nrm, k, g.VHpad);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int sp = 0; sp < 4; sp++) {
      s[i * 4 + sp] = v[sp][i].re;
      s[4 * D + i * 4 + sp] = v[sp][i].im;
    }
}

#define YH_DP(CALL)                                                     \
  switch (c->dr) {                                                      \
    case 2: { constexpr int D_ = 2; CALL; } break;                      \
    case 3: { constexpr int D_ = 3; CALL; } break;                      \
    case 4: { constexpr int D_ = 4; CALL; } break;                      \
    case 6: { constexpr int D_ = 6; CALL; } break;                      \
    case 8: { constexpr int D_ = 8; CALL; } break;                      \
    default: set_error("unsupported representation dimension"); return YH_ERR_UNSUPPORTED; \
  }

int field_layout_kernels_upload(Ctx* c, Field* f, const double* dev_aos, int p) {
  dim3 grid((c->g.VH + 255) / 256);
  if (f->prec == PREC_F64) { YH_DP(hipLaunchKernelGGL((k_from_aos<PREC_F64, D_>), grid, dim3(256), 0, c->stream, f->half[p], f->nrm[p], dev_aos, p, c->g)); }
  else if (f->prec == PREC_F32) { YH_DP(hipLaunchKernelGGL((k_from_aos<PREC_F32, D_>), grid, dim3(256), 0, c->stream, f->half[p], f->nrm[p], dev_aos, p, c->g)); }
  else { YH_DP(hipLaunchKernelGGL((k_from_aos<PREC_F16, D_>), grid, dim3(256), 0, c->stream, f->half[p], f->nrm[p], dev_aos, p, c->g)); }
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int field_layout_kernels_download(Ctx* c, const Field* f, double* dev_aos, int p) {
  dim3 grid((c->g.VH + 255) / 256);
  if (f->prec == PREC_F64) { YH_DP(hipLaunchKernelGGL((k_to_aos<PREC_F64, D_>), grid, dim3(256), 0, c->stream, f->half[p], f->nrm[p], dev_aos, p, c->g)); }
  else if (f->prec == PREC_F32) { YH_DP(hipLaunchKernelGGL((k_to_aos<PREC_F32, D_>), grid, dim3(256), 0, c->stream, f->half[p], f->nrm[p], dev_aos, p, c->g)); }
  else { YH_DP(hipLaunchKernelGGL((k_to_aos<PREC_F16, D_>), grid, dim3(256), 0, c->stream, f->half[p], f->nrm[p], dev_aos, p, c->g)); }
  YH_HIP(hipGetLastError());
  return YH_OK;
}

__global__ void __launch_bounds__(256) k_sum(const double* __restrict__ v, int n, double* partials, unsigned* ticket,
                                             double* out) {
  int k = blockIdx.x * 256 + threadIdx.x;
  double s[1] = {k < n ? v[k] : 0.0};
  block_reduce_finish<1, 256>(s, partials, gridDim.x, blockIdx.x, ticket, out);
}

int clover_logdet_sum(Ctx* c, double* out) {
  hipLaunchKernelGGL(k_sum, dim3((c->g.VH + 255) / 256), dim3(256), 0, c->stream, c->logdet_site, c->g.VH, c->partials,
                     c->ticket, &c->sc->red[0]);
  YH_HIP(hipGetLastError());
  YH_HIP(hipMemcpyAsync(out, &c->sc->red[0], sizeof(double), hipMemcpyDeviceToHost, c->stream));
  YH_HIP(hipStreamSynchronize(c->stream));
  return YH_OK;
}

// plain 16-byte vector copy: the calibration kernel for HBM bandwidth and
// for the FETCH_SIZE / WRITE_SIZE counters.  One 16-byte element per lane
// with non-temporal loads and stores was the fastest form measured on
// MI355X (6.5 TB/s for 1 GiB; four elements per lane with ordinary
// accesses reached 5.6 TB/s).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ a, uint4* __restrict__ b, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load((const u32x4*)a + i), (u32x4*)b + i);
}

int copy_kernel_bench(Ctx* c, long bytes, int reps, double* ms) {
  long n = bytes / 16;
  if (n < 1) { set_error("copy size too small"); return YH_ERR_ARG; }
  uint4 *a, *b;
  YH_HIP(hipMalloc(&a, n * 16));
  YH_HIP(hipMalloc(&b, n * 16));
  YH_HIP(hipMemsetAsync(a, 1, n * 16, c->stream));
  long gl = (n + 255) / 256;
  if (gl > (1L << 30)) { set_error("copy size too large"); return YH_ERR_ARG; }
  int grid = (int)gl;
  hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, c->stream, a, b, n);
  hipEvent_t e0, e1;
  YH_HIP(hipEventCreate(&e0));
  YH_HIP(hipEventCreate(&e1));
  YH_HIP(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, c->stream, a, b, n);
  YH_HIP(hipEventRecord(e1, c->stream));
  YH_HIP(hipEventSynchronize(e1));
  float t;
  YH_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  *ms = t / reps;
  return YH_OK;
}

}  // namespace yh
