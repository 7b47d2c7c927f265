// Shared device/host definitions for the MI355X Wilson-clover hot path.
//
// HBM layouts (all per parity half, "site-minor" structure-of-arrays so that
// a wave of 64 consecutive output sites issues fully coalesced 8/16-byte
// loads for every component):
//
//   spinor half   vec[v * VHpad + site], v < SPV
//                 complex component k = spin * D + colour
//                 FP64: double2 (1 complex / vec), SPV = 4D
//                 FP32: float4  (2 complex / vec), SPV = 2D
//                 FP16: uint4   (4 complex / vec, IEEE half), SPV = D,
//                       plus one float scale per site (max |component|)
//   links         for output parity p and direction mu, the forward link
//                 U(x, mu) of every parity-p site x:
//                 vec[((p * 4 + mu) * LV + v) * VHpad + site]
//                 stored WITHOUT the antiperiodic sign (applied in-kernel)
//   clover        per site two Hermitian 2D x 2D chiral blocks (spins 0,1
//                 and spins 2,3; row a = spin_in_pair * D + colour), packed
//                 as n real diagonal entries then the strictly lower
//                 triangle row-major as complex: n^2 reals per block
//
// Site order inside a half is the reference's EO order restricted to one
// parity: half index = lex / 2 with lex = ((t*X + x)*Y + y)*Z + z
// (library/geometry.cpp:756-878, single process), so host fields convert
// site by site.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace yh {

enum { PREC_F64 = 0, PREC_F32 = 1, PREC_F16 = 2 };
enum { BC_PERIODIC = 0, BC_ANTIPERIODIC = 1, BC_OPEN = 2, BC_SF = 3 };
// RECON_12H: 12-real SU(3) links stored as FP16 under FP32 spinors (the
// FP16 path's link array; the spinors and the clover stay FP32)
enum { RECON_18 = 0, RECON_12 = 1, RECON_12H = 2 };

struct Geom {
  int L[4];       // T, X, Y, Z
  int V, VH;      // volume, half volume
  int VHpad;      // padded half-field stride (elements)
  int bc;         // boundary condition
  int Ta, Tb;     // time slices carrying the c_f improvement (open / SF)
  double cf;      // c_f (open / SF); the diagonal gets c_f - 1 extra there
  int xq;         // XCD chunk: consecutive workgroups one XCD takes in a row (0: nblocks / 8)
};

template <typename R>
struct Cx {
  R re, im;
  __host__ __device__ Cx() {}
  __host__ __device__ Cx(R a, R b) : re(a), im(b) {}
};

template <typename R>
__host__ __device__ inline Cx<R> operator+(Cx<R> a, Cx<R> b) { return Cx<R>(a.re + b.re, a.im + b.im); }
template <typename R>
__host__ __device__ inline Cx<R> operator-(Cx<R> a, Cx<R> b) { return Cx<R>(a.re - b.re, a.im - b.im); }
template <typename R>
__host__ __device__ inline Cx<R> operator*(Cx<R> a, Cx<R> b) {
  return Cx<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
__host__ __device__ inline Cx<R> operator*(R s, Cx<R> a) { return Cx<R>(s * a.re, s * a.im); }
template <typename R>
__host__ __device__ inline Cx<R> conj(Cx<R> a) { return Cx<R>(a.re, -a.im); }
// a += b * c, as four fused multiply-adds (written as a + (x - y) the
// compiler needs a multiply, an FMA and an add per component)
template <typename R>
__host__ __device__ inline void cmac(Cx<R>& a, Cx<R> b, Cx<R> c) {
  a.re = fma(b.re, c.re, a.re);
  a.re = fma(-b.im, c.im, a.re);
  a.im = fma(b.re, c.im, a.im);
  a.im = fma(b.im, c.re, a.im);
}
// a += conj(b) * c
template <typename R>
__host__ __device__ inline void cmac_conj(Cx<R>& a, Cx<R> b, Cx<R> c) {
  a.re = fma(b.re, c.re, a.re);
  a.re = fma(b.im, c.im, a.re);
  a.im = fma(b.re, c.im, a.im);
  a.im = fma(-b.im, c.re, a.im);
}

// multiplication by one of {+1, -1, +i, -i} (codes 0..3)
template <int C, typename R>
__device__ __forceinline__ Cx<R> unit_mul(Cx<R> z) {
  if constexpr (C == 0) return z;
  else if constexpr (C == 1) return Cx<R>(-z.re, -z.im);
  else if constexpr (C == 2) return Cx<R>(-z.im, z.re);
  else return Cx<R>(z.im, -z.re);
}

// a += s * unit_mul<C>(z) as two FMAs (ns = -s)
template <int C, typename R>
__device__ __forceinline__ void unit_fma(Cx<R>& a, R s, R ns, Cx<R> z) {
  if constexpr (C == 0) {
    a.re = fma(s, z.re, a.re);
    a.im = fma(s, z.im, a.im);
  } else if constexpr (C == 1) {
    a.re = fma(ns, z.re, a.re);
    a.im = fma(ns, z.im, a.im);
  } else if constexpr (C == 2) {
    a.re = fma(ns, z.im, a.re);
    a.im = fma(s, z.re, a.im);
  } else {
    a.re = fma(s, z.im, a.re);
    a.im = fma(ns, z.re, a.im);
  }
}

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
__device__ __forceinline__ void site_coords(int k, int parity, const Geom& g, int c[4]) {
  int lex = 2 * k;
  int z = lex % g.L[3];
  int r = lex / g.L[3];
  int y = r % g.L[2];
  r /= g.L[2];
  int x = r % g.L[1];
  int t = r / g.L[1];
  if (((t + x + y + z) & 1) != parity) z += 1;
  c[0] = t; c[1] = x; c[2] = y; c[3] = z;
}

__device__ __forceinline__ int half_index(int t, int x, int y, int z, const Geom& g) {
  return (((t * g.L[1] + x) * g.L[2] + y) * g.L[3] + z) >> 1;
}

// full-lattice EO index (parity * VH + half index) and its inverse
__device__ __forceinline__ int full_index(int t, int x, int y, int z, const Geom& g) {
  return ((t + x + y + z) & 1) * g.VH + half_index(t, x, y, z, g);
}

__device__ __forceinline__ void full_coords(int id, const Geom& g, int c[4]) {
  int p = id >= g.VH;
  site_coords(id - p * g.VH, p, g, c);
}

// true where the boundary conditions force spinors to zero
// (library/bcs.cpp:139-164: open t in {0, T-1}; SF t in {0, T-2, T-1})
__device__ __forceinline__ bool bc_zero_slice(int t, const Geom& g) {
  if (g.bc == BC_OPEN) return t == 0 || t == g.L[0] - 1;
  if (g.bc == BC_SF) return t == 0 || t == g.L[0] - 2 || t == g.L[0] - 1;
  return false;
}

__device__ __forceinline__ double bc_cf_shift(int t, const Geom& g) {
  if ((g.bc == BC_OPEN || g.bc == BC_SF) && (t == g.Ta || t == g.Tb)) return g.cf - 1.0;
  return 0.0;
}

// XCD-aware block order.  The hardware deals consecutive workgroups round
// robin to the 8 XCDs (physical block b runs on XCD b % 8 as that XCD's
// (b / 8)-th block).  Remap so that XCD x takes `q` consecutive logical
// blocks of every group of 8q: with q = nblocks / 8 each XCD sweeps one
// contiguous slab; smaller q interleaves the XCDs at a finer grain so the
// t +- 1 reuse window of each XCD stays small.  The map is a bijection on
// [0, G*8q) and the identity on the tail.
__device__ __forceinline__ int xcd_block(int nblocks, int q) {
  const int b = blockIdx.x;
  if (q <= 0) q = nblocks / 8;
  if (q <= 0) return b;
  const int group = 8 * q;
  const int G = nblocks / group;
  if (b >= G * group) return b;
  const int x = b & 7, s = b >> 3;
  return (s / q) * group + x * q + (s % q);
}

// ---------------------------------------------------------------------------
// non-temporal (streaming) accesses for data a kernel touches once: the
// links of a hop and its output field
// ---------------------------------------------------------------------------
template <typename V>
__device__ __forceinline__ V ld_stream(const V* p) {
  if constexpr (sizeof(V) == 16 && alignof(V) >= 16) {
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    u4 r = __builtin_nontemporal_load((const u4*)p);
    V v;
    __builtin_memcpy(&v, &r, 16);
    return v;
  } else if constexpr (sizeof(V) == 8) {
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    u2 r = __builtin_nontemporal_load((const u2*)p);
    V v;
    __builtin_memcpy(&v, &r, 8);
    return v;
  } else {
    return *p;
  }
}

template <typename V>
__device__ __forceinline__ void st_stream(V* p, const V& v) {
  if constexpr (sizeof(V) == 16 && alignof(V) >= 16) {
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    u4 r;
    __builtin_memcpy(&r, &v, 16);
    __builtin_nontemporal_store(r, (u4*)p);
  } else {
    *p = v;
  }
}

// ---------------------------------------------------------------------------
// spinor storage traits
// ---------------------------------------------------------------------------
template <int PREC, int D>
struct SpinorStore;

template <int D>
struct SpinorStore<PREC_F64, D> {
  typedef double2 vec;
  typedef double real;
  static constexpr int NV = 4 * D;
  static constexpr int BYTES = 64 * D;  // per site
};
template <int D>
struct SpinorStore<PREC_F32, D> {
  typedef float4 vec;
  typedef float real;
  static constexpr int NV = 2 * D;
  static constexpr int BYTES = 32 * D;
};
template <int D>
struct SpinorStore<PREC_F16, D> {
  typedef uint4 vec;
  typedef float real;
  static constexpr int NV = D;
  static constexpr int BYTES = 16 * D + 4;
};

__device__ __forceinline__ float h2f(unsigned short h) { return __half2float(__ushort_as_half(h)); }

// 16-bit links and spinors are fixed point: element = q / 32768 with q a
// signed (two's complement) 16-bit integer.  Every element of a unitary
// matrix in any representation has modulus <= 1 (a spinor is divided by its
// largest component first), and fixed point resolves 1.5e-5 over the whole
// range where FP16 rounds elements near 1 to 2.4e-4.  That error is the
// sloppy operator's, and the mixed CG's iteration count follows it: 983
// iterations with FP16 storage, 764 with fixed point, 681 with FP32 at 32^4.
// Decoding is one sign-extending convert per word (q2r_lo / q2r_hi compile
// to v_cvt_f32_i32 with a WORD_0 / WORD_1 source select); the 2^-15 is a
// separate multiply or is folded into a later scale.
// encoding: 1.5 * 2^23 + 32768 f, which the FMA rounds to an integer
// (round to nearest even), holds 2^22 + q in its mantissa, so its low 16
// bits are q in two's complement
__device__ __forceinline__ unsigned q_enc_bits(float f) {
  return __float_as_uint(fmaf(fminf(fmaxf(f, -1.0f), 32767.0f / 32768.0f), 32768.0f, 12582912.0f));
}
__device__ __forceinline__ unsigned f2q(float f) { return q_enc_bits(f) & 0xffffu; }
__device__ __forceinline__ unsigned pack_q2(float lo, float hi) {
  return (q_enc_bits(lo) & 0xffffu) | (q_enc_bits(hi) << 16);
}
// q itself, as a float
__device__ __forceinline__ float q2r_lo(unsigned v) { return (float)(int)(short)(v & 0xffffu); }
__device__ __forceinline__ float q2r_hi(unsigned v) { return (float)((int)v >> 16); }
constexpr float Q_UNIT = 1.0f / 32768.0f;
__device__ __forceinline__ float q2f_lo(unsigned v) { return q2r_lo(v) * Q_UNIT; }
__device__ __forceinline__ float q2f_hi(unsigned v) { return q2r_hi(v) * Q_UNIT; }
__device__ __forceinline__ unsigned short f2h(float f) { return __half_as_ushort(__float2half_rn(f)); }

// load the full spinor of one site into registers s[spin][colour]
template <int PREC, int D, typename R>
__device__ __forceinline__ void load_spinor(Cx<R> (&s)[4][D], const void* base, const float* nrm, int site,
                                            int stride) {
  if constexpr (PREC == PREC_F64) {
    const double2* p = (const double2*)base;
#pragma unroll
    for (int k = 0; k < 4 * D; k++) {
      double2 v = p[(size_t)k * stride + site];
      s[k / D][k % D] = Cx<R>((R)v.x, (R)v.y);
    }
  } else if constexpr (PREC == PREC_F32) {
    const float4* p = (const float4*)base;
#pragma unroll
    for (int k = 0; k < 2 * D; k++) {
      float4 v = p[(size_t)k * stride + site];
      s[(2 * k) / D][(2 * k) % D] = Cx<R>((R)v.x, (R)v.y);
      s[(2 * k + 1) / D][(2 * k + 1) % D] = Cx<R>((R)v.z, (R)v.w);
    }
  } else {
    const uint4* p = (const uint4*)base;
    const float sc = nrm[site] * Q_UNIT;
#pragma unroll
    for (int k = 0; k < D; k++) {
      uint4 v = p[(size_t)k * stride + site];
      unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int c = 4 * k + q;
        s[c / D][c % D] = Cx<R>((R)(q2r_lo(w[q]) * sc), (R)(q2r_hi(w[q]) * sc));
      }
    }
  }
}

// a 16-bit spinor as the integers q, without 2^-15 and its site scale (their
// product is returned): the hop applies the scale once to the projected,
// transported half spinor
template <int D, typename R>
__device__ __forceinline__ R load_spinor_q(Cx<R> (&s)[4][D], const void* base, const float* nrm, int site,
                                           int stride) {
  const uint4* p = (const uint4*)base;
#pragma unroll
  for (int k = 0; k < D; k++) {
    uint4 v = p[(size_t)k * stride + site];
    unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int c = 4 * k + q;
      s[c / D][c % D] = Cx<R>((R)q2r_lo(w[q]), (R)q2r_hi(w[q]));
    }
  }
  return (R)(nrm[site] * Q_UNIT);
}

// NT: write with non-temporal stores (a field no later access of the same
// kernel reads)
template <int PREC, int D, typename R, bool NT = false>
__device__ __forceinline__ void store_spinor(void* base, float* nrm, int site, int stride, const Cx<R> (&s)[4][D]) {
  if constexpr (PREC == PREC_F64) {
    double2* p = (double2*)base;
#pragma unroll
    for (int k = 0; k < 4 * D; k++) {
      const double2 v = make_double2(s[k / D][k % D].re, s[k / D][k % D].im);
      if constexpr (NT) st_stream(&p[(size_t)k * stride + site], v);
      else p[(size_t)k * stride + site] = v;
    }
  } else if constexpr (PREC == PREC_F32) {
    float4* p = (float4*)base;
#pragma unroll
    for (int k = 0; k < 2 * D; k++) {
      Cx<R> a = s[(2 * k) / D][(2 * k) % D], b = s[(2 * k + 1) / D][(2 * k + 1) % D];
      const float4 v = make_float4((float)a.re, (float)a.im, (float)b.re, (float)b.im);
      if constexpr (NT) st_stream(&p[(size_t)k * stride + site], v);
      else p[(size_t)k * stride + site] = v;
    }
  } else {
    uint4* p = (uint4*)base;
    float mx = 0.f;
#pragma unroll
    for (int c = 0; c < 4 * D; c++) {
      mx = fmaxf(mx, fabsf((float)s[c / D][c % D].re));
      mx = fmaxf(mx, fabsf((float)s[c / D][c % D].im));
    }
    float inv = mx > 0.f ? 1.0f / mx : 0.f;
    nrm[site] = mx;
#pragma unroll
    for (int k = 0; k < D; k++) {
      unsigned w[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int c = 4 * k + q;
        w[q] = pack_q2((float)s[c / D][c % D].re * inv, (float)s[c / D][c % D].im * inv);
      }
      const uint4 v = make_uint4(w[0], w[1], w[2], w[3]);
      if constexpr (NT) st_stream(&p[(size_t)k * stride + site], v);
      else p[(size_t)k * stride + site] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// link storage traits
// ---------------------------------------------------------------------------
// number of storage vectors per link
template <int PREC, int RECON, int D>
struct LinkStore;
template <int D>
struct LinkStore<PREC_F64, RECON_18, D> { typedef double2 vec; static constexpr int NV = D * D; static constexpr int BYTES = 16 * D * D; };
template <int D>
struct LinkStore<PREC_F64, RECON_12, D> { typedef double2 vec; static constexpr int NV = 6; static constexpr int BYTES = 96; };
template <int D>
struct LinkStore<PREC_F32, RECON_18, D> { typedef float2 vec; static constexpr int NV = D * D; static constexpr int BYTES = 8 * D * D; };
template <int D>
struct LinkStore<PREC_F32, RECON_12, D> { typedef float4 vec; static constexpr int NV = 3; static constexpr int BYTES = 48; };
template <int D>
struct LinkStore<PREC_F16, RECON_18, D> { typedef unsigned vec; static constexpr int NV = D * D; static constexpr int BYTES = 4 * D * D; };
template <int D>
struct LinkStore<PREC_F16, RECON_12, D> { typedef uint2 vec; static constexpr int NV = 3; static constexpr int BYTES = 24; };
template <int D>
struct LinkStore<PREC_F32, RECON_12H, D> { typedef uint2 vec; static constexpr int NV = 3; static constexpr int BYTES = 24; };

// third row of an SU(3) matrix from the first two: conj(r0 x r1)
template <typename R>
__device__ __forceinline__ Cx<R> conj_cross(Cx<R> a, Cx<R> b, Cx<R> c, Cx<R> d) {
  // conj(a b - c d) as two four-term FMA chains
  R re = a.re * b.re;
  re = fma(-a.im, b.im, re);
  re = fma(-c.re, d.re, re);
  re = fma(c.im, d.im, re);
  R im = c.re * d.im;
  im = fma(c.im, d.re, im);
  im = fma(-a.re, b.im, im);
  im = fma(-a.im, b.re, im);
  return Cx<R>(re, im);
}
template <typename R>
__device__ __forceinline__ void su3_third_row(Cx<R> (&u)[3][3]) {
  u[2][0] = conj_cross(u[0][1], u[1][2], u[0][2], u[1][1]);
  u[2][1] = conj_cross(u[0][2], u[1][0], u[0][0], u[1][2]);
  u[2][2] = conj_cross(u[0][0], u[1][1], u[0][1], u[1][0]);
}

// NT: non-temporal loads (each link is read once per hop)
// RAW (16-bit links): the matrix times 32768 (the integers q; a reconstructed
// third row is brought to the same scale), for callers that fold the 2^-15
// into a later scale
template <int PREC, int RECON, int D, typename R, bool NT = false, bool RAW = false>
__device__ __forceinline__ void load_link(Cx<R> (&u)[D][D], const void* base, int site, int stride) {
  typedef typename LinkStore<PREC, RECON, D>::vec V;
  const V* p = (const V*)base;
  auto ld = [&](size_t i) -> V {
    if constexpr (NT) return ld_stream(&p[i]);
    else return p[i];
  };
  if constexpr (RECON == RECON_18) {
#pragma unroll
    for (int k = 0; k < D * D; k++) {
      V v = ld((size_t)k * stride + site);
      if constexpr (PREC == PREC_F64 || PREC == PREC_F32) u[k / D][k % D] = Cx<R>((R)v.x, (R)v.y);
      else if constexpr (RAW) u[k / D][k % D] = Cx<R>((R)q2r_lo(v), (R)q2r_hi(v));
      else u[k / D][k % D] = Cx<R>((R)q2f_lo(v), (R)q2f_hi(v));
    }
  } else {
    static_assert(D == 3, "12-real reconstruction is SU(3) only");
    Cx<R>(&w)[3][3] = *reinterpret_cast<Cx<R>(*)[3][3]>(&u);
    if constexpr (PREC == PREC_F64) {
#pragma unroll
      for (int k = 0; k < 6; k++) {
        double2 v = ld((size_t)k * stride + site);
        w[k / 3][k % 3] = Cx<R>((R)v.x, (R)v.y);
      }
    } else if constexpr (PREC == PREC_F32 && RECON == RECON_12) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        float4 v = ld((size_t)k * stride + site);
        w[(2 * k) / 3][(2 * k) % 3] = Cx<R>((R)v.x, (R)v.y);
        w[(2 * k + 1) / 3][(2 * k + 1) % 3] = Cx<R>((R)v.z, (R)v.w);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        uint2 v = ld((size_t)k * stride + site);
        if constexpr (RAW) {
          w[(2 * k) / 3][(2 * k) % 3] = Cx<R>((R)q2r_lo(v.x), (R)q2r_hi(v.x));
          w[(2 * k + 1) / 3][(2 * k + 1) % 3] = Cx<R>((R)q2r_lo(v.y), (R)q2r_hi(v.y));
        } else {
          w[(2 * k) / 3][(2 * k) % 3] = Cx<R>((R)q2f_lo(v.x), (R)q2f_hi(v.x));
          w[(2 * k + 1) / 3][(2 * k + 1) % 3] = Cx<R>((R)q2f_lo(v.y), (R)q2f_hi(v.y));
        }
      }
    }
    su3_third_row(w);
    if constexpr (RAW) {
#pragma unroll
      for (int j = 0; j < 3; j++) w[2][j] = Cx<R>(w[2][j].re * (R)Q_UNIT, w[2][j].im * (R)Q_UNIT);
    }
  }
}

// ---------------------------------------------------------------------------
// clover storage traits (n = 2D, n*n reals per block, two blocks per site)
// ---------------------------------------------------------------------------
template <int PREC, int D>
struct CloverStore;
template <int D>
struct CloverStore<PREC_F64, D> { typedef double2 vec; static constexpr int NV = (2 * D) * (2 * D); static constexpr int BYTES = 8 * 2 * (2 * D) * (2 * D); };
template <int D>
struct CloverStore<PREC_F32, D> { typedef float4 vec; static constexpr int NV = (2 * D) * (2 * D) / 2; static constexpr int BYTES = 4 * 2 * (2 * D) * (2 * D); };
// 16-bit: fixed point (as the 16-bit spinors) with one float scale per block
// and site, the block's largest |real|: 4 reals per uint2
template <int D>
struct CloverStore<PREC_F16, D> { typedef uint2 vec; static constexpr int NV = (2 * D) * (2 * D) / 2; static constexpr int BYTES = 2 * 2 * (2 * D) * (2 * D) + 8; };

// load block b's packed n*n reals into a register array
// (16-bit: cnrm holds the scales, [block][site])
template <int PREC, int D, typename R>
__device__ __forceinline__ void load_clover_block(R (&c)[(2 * D) * (2 * D)], const void* base, int blk, int site,
                                                  int stride, const float* cnrm = nullptr) {
  constexpr int n2 = (2 * D) * (2 * D);
  if constexpr (PREC == PREC_F16) {
    const uint2* p = (const uint2*)base + (size_t)blk * (n2 / 4) * stride;
    const R sc = (R)(cnrm[(size_t)blk * stride + site] * Q_UNIT);
#pragma unroll
    for (int k = 0; k < n2 / 4; k++) {
      uint2 v = p[(size_t)k * stride + site];
      c[4 * k] = sc * (R)q2r_lo(v.x);
      c[4 * k + 1] = sc * (R)q2r_hi(v.x);
      c[4 * k + 2] = sc * (R)q2r_lo(v.y);
      c[4 * k + 3] = sc * (R)q2r_hi(v.y);
    }
  } else if constexpr (PREC == PREC_F64) {
    const double2* p = (const double2*)base + (size_t)blk * (n2 / 2) * stride;
#pragma unroll
    for (int k = 0; k < n2 / 2; k++) {
      double2 v = p[(size_t)k * stride + site];
      c[2 * k] = (R)v.x;
      c[2 * k + 1] = (R)v.y;
    }
  } else {
    const float4* p = (const float4*)base + (size_t)blk * (n2 / 4) * stride;
#pragma unroll
    for (int k = 0; k < n2 / 4; k++) {
      float4 v = p[(size_t)k * stride + site];
      c[4 * k] = (R)v.x;
      c[4 * k + 1] = (R)v.y;
      c[4 * k + 2] = (R)v.z;
      c[4 * k + 3] = (R)v.w;
    }
  }
}

// 16-bit block as loaded: the packed integers (4 per uint2) and the block's
// scale; herm_mv_q decodes each entry where it is used, so the registers
// hold the 16-bit words rather than n*n floats
template <int D>
__device__ __forceinline__ float load_clover_block_q(uint2 (&cq)[(2 * D) * (2 * D) / 4], const void* base, int blk,
                                                     int site, int stride, const float* cnrm) {
  constexpr int n2 = (2 * D) * (2 * D);
  const uint2* p = (const uint2*)base + (size_t)blk * (n2 / 4) * stride;
#pragma unroll
  for (int k = 0; k < n2 / 4; k++) cq[k] = p[(size_t)k * stride + site];
  return cnrm[(size_t)blk * stride + site] * Q_UNIT;
}

template <int N>
__device__ __forceinline__ float clq(const uint2 (&cq)[N * N / 4], int k) {
  const unsigned w = (k & 2) ? cq[k >> 2].y : cq[k >> 2].x;
  return (k & 1) ? q2r_hi(w) : q2r_lo(w);
}

// y = sc H x for the packed Hermitian block held as 16-bit integers
template <int N, typename R>
__device__ __forceinline__ void herm_mv_q(Cx<R> (&y)[N], const uint2 (&cq)[N * N / 4], R sc, const Cx<R> (&x)[N]) {
#pragma unroll
  for (int i = 0; i < N; i++) y[i] = clq<N>(cq, i) * x[i];
  int o = N;
#pragma unroll
  for (int i = 1; i < N; i++) {
#pragma unroll
    for (int j = 0; j < i; j++) {
      Cx<R> h(clq<N>(cq, o), clq<N>(cq, o + 1));
      o += 2;
      cmac(y[i], h, x[j]);
      cmac_conj(y[j], h, x[i]);
    }
  }
#pragma unroll
  for (int i = 0; i < N; i++) y[i] = sc * y[i];
}

// y = H x for a packed FP64 block of D_R >= 6 (2 D_R^2 doubles per block)
// read from memory as it is used: after every ~24 complex entries the next
// loads wait (opaque zero) for the products so far, so at most that many
// entries are live.  The products are those of herm_mv, in its order.
template <int N>
__device__ __forceinline__ void herm_mv_f64_stream(Cx<double> (&y)[N], const void* base, int blk, int site,
                                                   int stride, const Cx<double> (&x)[N]) {
  const double2* p = (const double2*)base + (size_t)blk * (N * N / 2) * stride;
  int dep = 0;
#pragma unroll
  for (int k = 0; k < N / 2; k++) {
    const double2 v = p[(size_t)k * stride + site];
    y[2 * k] = v.x * x[2 * k];
    y[2 * k + 1] = v.y * x[2 * k + 1];
  }
  int cnt = 0;
#pragma unroll
  for (int i = 1; i < N; i++) {
    if (cnt >= 24) {
#pragma unroll
      for (int q = 0; q < N; q++) __asm__ volatile("" : "+v"(y[q].re), "+v"(y[q].im));
      __asm__ volatile("v_mov_b32 %0, 0" : "=v"(dep));
      cnt = 0;
    }
#pragma unroll
    for (int j = 0; j < i; j++) {
      const double2 v = p[(size_t)(N / 2 + i * (i - 1) / 2 + j) * stride + site + dep];
      const Cx<double> h(v.x, v.y);
      cmac(y[i], h, x[j]);
      cmac_conj(y[j], h, x[i]);
    }
    cnt += i;
  }
}

// y = H x for the packed Hermitian block
template <int N, typename R>
__device__ __forceinline__ void herm_mv(Cx<R> (&y)[N], const R (&c)[N * N], const Cx<R> (&x)[N]) {
#pragma unroll
  for (int i = 0; i < N; i++) y[i] = c[i] * x[i];
  int o = N;
#pragma unroll
  for (int i = 1; i < N; i++) {
#pragma unroll
    for (int j = 0; j < i; j++) {
      Cx<R> h(c[o], c[o + 1]);
      o += 2;
      cmac(y[i], h, x[j]);
      cmac_conj(y[j], h, x[i]);
    }
  }
}

}  // namespace yh
