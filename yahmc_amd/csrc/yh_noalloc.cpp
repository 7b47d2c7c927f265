// CG with no allocation: yh_cg_noalloc.
//
// The same solver as yh_cg (cg_device: the reference recurrence of
// library/cg.cpp:132-202 with device-resident scalars), for callers that
// cannot call hipMalloc (code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation).  Its five temporary half fields
// (operator: odd half of D_oo^-1 D_oe p and the even half Hphi p; solver:
// r, p, q) are carved from `scratch`, a file-scope device array of
// YH_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until
// the process ends.  The reductions use the context's partial-sum buffer
// and scalars, which yh_create allocated; the clover inverse buffers were
// reserved by yh_set_gauge.  Nothing here allocates device or host memory.
//
// The array's name is a host-side stand-in, not a device address: the
// address is looked up once per device (noalloc_init, called by yh_create
// with that device current) and only that pointer is handed to kernels.
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; the solver writes every temporary before reading it, so nothing
// needs clearing.  Calls that overlap on one GPU (two host threads, or two
// contexts' streams) race on the array.
#include <hip/hip_runtime.h>
#include <mutex>
#include <string>
#include "yh_internal.hpp"

namespace yh {

static __device__ __align__(256) unsigned char scratch[YH_NOALLOC_SCRATCH_BYTES];

enum { NOALLOC_MAX_DEVICES = 64 };
static void* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_scratch_dev_mtx;

int noalloc_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) { set_error("device index out of range"); return YH_ERR_ARG; }
  std::lock_guard<std::mutex> lk(g_scratch_dev_mtx);
  if (g_scratch_dev[device]) return YH_OK;
  // the caller (yh_create) has made `device` current
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_error(std::string("no-allocation scratch lookup: ") + hipGetErrorString(e));
    return YH_ERR_HIP;
  }
  g_scratch_dev[device] = ptr;
  return YH_OK;
}

static size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace yh

using namespace yh;

extern "C" int yh_cg_noalloc(yh_ctx* c, const yh_inv_par* par, yh_field* dst, const yh_field* src, int* iters) {
  // the checks of yh_cg
  if (!c || !par || !dst || !src || !iters) { set_error("null argument"); return YH_ERR_ARG; }
  if (!c->gauge_set) { set_error("no gauge field"); return YH_ERR_STATE; }
  if (par->mvm < YH_MVM_DPHI || par->mvm > YH_MVM_HPHI_SQ) { set_error("bad mvm"); return YH_ERR_ARG; }
  if (!(par->prec > 0.0)) { set_error("prec must be positive"); return YH_ERR_ARG; }
  if (dst->prec != PREC_F64 || src->prec != PREC_F64) { set_error("solver fields must be FP64"); return YH_ERR_ARG; }
  int rc;
  if (c->op_tmp[0] || c->cg_tmp[0]) { set_error("cg_noalloc: temporaries already in use on this context"); return YH_ERR_STATE; }
  const size_t hb = round256(half_bytes(c, PREC_F64));
  const size_t need = 5 * hb;
  if (need > (size_t)YH_NOALLOC_SCRATCH_BYTES) {
    set_error("cg_noalloc: the lattice needs " + std::to_string(need) + " bytes of scratch, more than the " +
              std::to_string((size_t)YH_NOALLOC_SCRATCH_BYTES) + " available");
    return YH_ERR_ARG;
  }
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_dev_mtx);
    if (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES) base = (unsigned char*)g_scratch_dev[c->device];
  }
  if (!base) { set_error("cg_noalloc: scratch not set up for this device"); return YH_ERR_STATE; }
  // host-side field descriptors over the scratch array; each temporary
  // has only the half the operator and solver use
  Field f[5];
  for (int i = 0; i < 5; i++) {
    f[i].prec = PREC_F64;
    f[i].half[0] = f[i].half[1] = nullptr;
    f[i].nrm[0] = f[i].nrm[1] = nullptr;
    f[i].mem = nullptr;
    f[i].bytes = 0;
  }
  f[0].half[1] = base;           // operator: D_oo^-1 D_oe p (odd half)
  f[1].half[0] = base + hb;      // operator: Hphi p inside Hphi_sq (even half)
  f[2].half[0] = base + 2 * hb;  // r
  f[3].half[0] = base + 3 * hb;  // p
  f[4].half[0] = base + 4 * hb;  // q
  c->op_tmp[0] = &f[0];
  c->op_tmp[1] = &f[1];
  c->cg_tmp[0] = &f[2];
  c->cg_tmp[1] = &f[3];
  c->cg_tmp[2] = &f[4];
  // the descriptors are host-only: each launch takes the device pointers
  // by value, so they may go out of scope with work still queued
  rc = cg_device(c, PREC_F64, default_recon(c, PREC_F64), par, dst, src, iters, nullptr, nullptr);
  c->op_tmp[0] = c->op_tmp[1] = nullptr;
  c->cg_tmp[0] = c->cg_tmp[1] = c->cg_tmp[2] = nullptr;
  return rc;
}
