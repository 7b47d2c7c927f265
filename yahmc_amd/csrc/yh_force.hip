// Fermion forces (FP64): hopping force, clover force, clover log-det force
// and the clover-leaf derivative that turns the clover force matrices into
// momentum updates.
//
// wilson_fermion_force : library/forces.cpp:175-227
// clover_fermion_force : library/forces.cpp:122-172
// clover_logdet_force  : library/forces.cpp:109-119 + clover.cpp:170-279
// end_fermion_force    : library/forces.cpp:33-106, 254-259
// algebra_project      : library/suNa.cpp:102-111 (TF = 2)
//
// Device layouts: momenta mom[(mu * NG + a) * V + id] (FP64, id = full EO
// index), clover force matrices clf[(plane * D * D + e) * V + id] (double2),
// planes in the reference order 01, 02, 12, 03, 13, 23.
#include <vector>
#include <mutex>
#include "yh_internal.hpp"

namespace yh {

// represented generators iT^R_a, a < NG, D x D row-major: the context's
// table tab_itr[(a * D + i) * D + k], passed to the kernels
__constant__ double2 c_g5omg[4 * 16];    // gamma5 (1 - gamma_mu)
__constant__ double2 c_g5sig[6 * 16];    // gamma5 (1/2)[gamma_mu, gamma_nu], reference plane order

typedef Cx<double> cd;

// ---------------------------------------------------------------------------
// host: generators of the fermion representation and the spin matrices
// ---------------------------------------------------------------------------
static void gen_fund(int nc, int a, std::vector<cd>& g) {
  g.assign(nc * nc, cd(0, 0));
  int k = 0;
  for (int i = 0; i < nc; i++)
    for (int j = 0; j < nc; j++) {
      if (i > j) {
        if (k == a) { g[i * nc + j].im += 1; g[j * nc + i].im += 1; }
        k++;
      } else if (i < j) {
        if (k == a) { g[i * nc + j].re += 1; g[j * nc + i].re -= 1; }
        k++;
      } else if (i > 0) {
        if (k == a) {
          double v = sqrt(2.0 / (i * (i + 1)));
          for (int q = 0; q < i; q++) g[q * nc + q].im += v;
          g[i * nc + i].im -= i * v;
        }
        k++;
      }
    }
}

// basis of the representation space (matches the device basis used by the
// representation kernel): sym/asym real tensors, adjoint (-i/sqrt2) iT_a
static void rep_basis(int nc, int repr, int dr, std::vector<std::vector<cd>>& b) {
  const double s2 = 0.70710678118654752440;
  b.assign(dr, std::vector<cd>(nc * nc, cd(0, 0)));
  if (repr == YH_REPR_SYMMETRIC) {
    for (int i = 0, n = 0; i < nc; i++)
      for (int j = 0; j <= i; j++, n++) {
        if (i == j) b[n][i * nc + i].re = 1;
        else { b[n][i * nc + j].re = s2; b[n][j * nc + i].re = s2; }
      }
  } else if (repr == YH_REPR_ANTISYMMETRIC) {
    for (int i = 0, n = 0; i < nc; i++)
      for (int j = 0; j < i; j++, n++) {
        b[n][j * nc + i].re = s2;
        b[n][i * nc + j].re = -s2;
      }
  } else if (repr == YH_REPR_ADJOINT) {
    std::vector<cd> g;
    for (int n = 0; n < dr; n++) {
      gen_fund(nc, n, g);
      for (int k = 0; k < nc * nc; k++) b[n][k] = cd(s2 * g[k].im, -s2 * g[k].re);
    }
  }
}

// iT^R_a = d/dt R(exp(t iT_a)) at t = 0 (library/repr.cpp:71-103):
//   fund: iT_a;  sym/asym: sum conj(b_n)_{xy} (G b_m + b_m G^T)_{xy};
//   adjoint: Tr(b_n [G, b_m])
static void rep_generators(int nc, int repr, int dr, std::vector<cd>& out) {
  int ng = nc * nc - 1;
  out.assign(ng * dr * dr, cd(0, 0));
  std::vector<std::vector<cd>> b;
  rep_basis(nc, repr, dr, b);
  std::vector<cd> G;
  for (int a = 0; a < ng; a++) {
    gen_fund(nc, a, G);
    for (int n = 0; n < dr; n++)
      for (int m = 0; m < dr; m++) {
        cd s(0, 0);
        if (repr == YH_REPR_FUNDAMENTAL) {
          s = G[n * nc + m];
        } else if (repr == YH_REPR_ADJOINT) {
          for (int x = 0; x < nc; x++)
            for (int y = 0; y < nc; y++)
              for (int z = 0; z < nc; z++)
                s = s + b[n][x * nc + y] * (G[y * nc + z] * b[m][z * nc + x] - b[m][y * nc + z] * G[z * nc + x]);
        } else {
          for (int x = 0; x < nc; x++)
            for (int y = 0; y < nc; y++) {
              cd v(0, 0);
              for (int z = 0; z < nc; z++) v = v + G[x * nc + z] * b[m][z * nc + y] + b[m][x * nc + z] * G[y * nc + z];
              s = s + conj(b[n][x * nc + y]) * v;
            }
        }
        out[(a * dr + n) * dr + m] = s;
      }
  }
}

// chiral gamma basis as applied by library/spinor.cpp:191-264
static void gamma_mats(cd g[4][4][4]) {
  for (int m = 0; m < 4; m++)
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) g[m][a][b] = cd(0, 0);
  g[0][0][2] = g[0][1][3] = g[0][2][0] = g[0][3][1] = cd(-1, 0);
  g[1][0][3] = g[1][1][2] = cd(0, -1);
  g[1][2][1] = g[1][3][0] = cd(0, 1);
  g[2][0][3] = cd(-1, 0); g[2][1][2] = cd(1, 0); g[2][2][1] = cd(1, 0); g[2][3][0] = cd(-1, 0);
  g[3][0][2] = cd(0, -1); g[3][1][3] = cd(0, 1); g[3][2][0] = cd(0, 1); g[3][3][1] = cd(0, -1);
}

int force_init_constants(Ctx* c) {
  std::vector<cd> itr;
  rep_generators(c->nc, c->repr, c->dr, itr);
  std::vector<double2> h(TAB_ITR, make_double2(0, 0));
  if (itr.size() > h.size()) { set_error("generator table too small"); return YH_ERR_UNSUPPORTED; }
  for (size_t i = 0; i < itr.size(); i++) h[i] = make_double2(itr[i].re, itr[i].im);
  int rc = upload_table(c, c->tab_itr, h.data(), h.size());
  if (rc) return rc;
  // gamma5 (1 - gamma_mu) and gamma5 sigma_mu_nu do not depend on the model:
  // uploaded once per device
  static std::mutex mtx;
  static unsigned done_mask = 0;
  std::lock_guard<std::mutex> lk(mtx);
  if (c->device < 32 && (done_mask >> c->device) & 1u) return YH_OK;
  cd g[4][4][4];
  gamma_mats(g);
  const double g5[4] = {1, 1, -1, -1};
  std::vector<double2> om(4 * 16), sg(6 * 16);
  for (int m = 0; m < 4; m++)
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        cd v = cd(a == b ? 1.0 : 0.0, 0) - g[m][a][b];
        om[m * 16 + a * 4 + b] = make_double2(g5[a] * v.re, g5[a] * v.im);
      }
  const int planes[6][2] = {{0, 1}, {0, 2}, {1, 2}, {0, 3}, {1, 3}, {2, 3}};
  for (int k = 0; k < 6; k++) {
    int mu = planes[k][0], nu = planes[k][1];
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        cd s(0, 0);
        for (int q = 0; q < 4; q++) s = s + g[mu][a][q] * g[nu][q][b] - g[nu][a][q] * g[mu][q][b];
        sg[k * 16 + a * 4 + b] = make_double2(g5[a] * 0.5 * s.re, g5[a] * 0.5 * s.im);
      }
  }
  YH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_g5omg), om.data(), sizeof(double2) * om.size()));
  YH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_g5sig), sg.data(), sizeof(double2) * sg.size()));
  YH_HIP(hipDeviceSynchronize());
  if (c->device < 32) done_mask |= 1u << c->device;
  return YH_OK;
}

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
struct FF {
  const void* x[2];   // lhs halves (FP64)
  const void* y[2];   // rhs halves (FP64)
};

__device__ __forceinline__ int fidx(int t, int x, int y, int z, const Geom& g) {
  return ((t + x + y + z) & 1) * g.VH + half_index(t, x, y, z, g);
}

__device__ __forceinline__ void fcoords(int id, const Geom& g, int c[4]) {
  int p = id >= g.VH;
  site_coords(id - p * g.VH, p, g, c);
}

__device__ __forceinline__ int fshift(const int c[4], int mu, int d, const Geom& g) {
  int n[4] = {c[0], c[1], c[2], c[3]};
  n[mu] = (n[mu] + d + g.L[mu]) % g.L[mu];
  return fidx(n[0], n[1], n[2], n[3], g);
}

template <int D>
__device__ __forceinline__ void load_full_spinor(cd (&s)[4][D], const void* const (&h)[2], int id, const Geom& g) {
  int p = id >= g.VH;
  load_spinor<PREC_F64, D, double>(s, h[p], nullptr, id - p * g.VH, g.VHpad);
}

// represented link U_R(id, mu) with the antiperiodic sign of the reference's
// gfield_repr (library/bcs.cpp:264-272)
template <int D>
__device__ __forceinline__ void load_link_bc(cd (&u)[D][D], const double2* links, int id, int mu, const Geom& g) {
  int p = id >= g.VH, k = id - p * g.VH;
  const double2* base = links + (size_t)(p * 4 + mu) * D * D * g.VHpad;
#pragma unroll
  for (int e = 0; e < D * D; e++) {
    double2 v = base[(size_t)e * g.VHpad + k];
    u[e / D][e % D] = cd(v.x, v.y);
  }
  if (g.bc == BC_ANTIPERIODIC && mu == 0) {
    int c[4];
    site_coords(k, p, g, c);
    if (c[0] == 0) {
#pragma unroll
      for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) u[i][j] = cd(-u[i][j].re, -u[i][j].im);
    }
  }
}

template <int D>
__device__ __forceinline__ void spin_mat(cd (&o)[4][D], const double2* M, const cd (&s)[4][D]) {
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < D; c++) {
      cd v(0, 0);
#pragma unroll
      for (int b = 0; b < 4; b++) {
        double2 m = M[a * 4 + b];
        if (m.x != 0.0 || m.y != 0.0) cmac(v, cd(m.x, m.y), s[b][c]);
      }
      o[a][c] = v;
    }
}

// F_ij += sum_spin a_i conj(b_j)
template <int D>
__device__ __forceinline__ void outer_acc(cd (&F)[D][D], const cd (&a)[4][D], const cd (&b)[4][D]) {
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++)
#pragma unroll
      for (int s = 0; s < 4; s++) {
        F[i][j].re += a[s][i].re * b[s][j].re + a[s][i].im * b[s][j].im;
        F[i][j].im += a[s][i].im * b[s][j].re - a[s][i].re * b[s][j].im;
      }
}

template <int D>
__device__ __forceinline__ void mmul(cd (&R)[D][D], const cd (&A)[D][D], const cd (&B)[D][D]) {
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      cd s(0, 0);
#pragma unroll
      for (int k = 0; k < D; k++) cmac(s, A[i][k], B[k][j]);
      R[i][j] = s;
    }
}

// mom_a += -(coeff / TF) Re Tr(iT^R_a F)
template <int NG, int D>
__device__ __forceinline__ void project_add(double* mom, int mu, int id, int V, const cd (&F)[D][D], double coeff,
                                            const double2* __restrict__ itr) {
#pragma unroll
  for (int a = 0; a < NG; a++) {
    double re = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int k = 0; k < D; k++) {
        double2 t = itr[(a * D + i) * D + k];
        re += t.x * F[k][i].re - t.y * F[k][i].im;
      }
    mom[(size_t)(mu * NG + a) * V + id] += -(coeff * 0.5) * re;
  }
}

// A D x D matrix left in memory: element (i, j) is read where it is used, so
// a product with one operand in memory keeps only the other operand and the
// result in registers.  `dag` reads the conjugate transpose, `sgn` carries
// the antiperiodic sign of a represented link.
struct MSrc {
  const double2* p;
  size_t stride;
  double sgn;
  bool dag;
};

template <int D>
__device__ __forceinline__ cd mget(const MSrc& m, int i, int j) {
  const double2 v = m.p[(size_t)(m.dag ? j * D + i : i * D + j) * m.stride];
  return m.dag ? cd(m.sgn * v.x, -m.sgn * v.y) : cd(m.sgn * v.x, m.sgn * v.y);
}

// represented link U_R(id, mu) (or its dagger) with the antiperiodic sign
template <int D>
__device__ __forceinline__ MSrc link_src(const double2* links, int id, int mu, bool dag, const Geom& g) {
  int p = id >= g.VH, k = id - p * g.VH;
  double sgn = 1.0;
  if (g.bc == BC_ANTIPERIODIC && mu == 0) {
    int c[4];
    site_coords(k, p, g, c);
    if (c[0] == 0) sgn = -1.0;
  }
  return MSrc{links + (size_t)(p * 4 + mu) * D * D * g.VHpad + k, (size_t)g.VHpad, sgn, dag};
}

// clover-force matrix of plane `num` at site id
template <int D>
__device__ __forceinline__ MSrc clf_src(const double2* clf, int num, int id, int V) {
  return MSrc{clf + (size_t)num * D * D * V + id, (size_t)V, 1.0, false};
}

// the same source, but readable only after `dep` is final: every element of
// dep passes through an empty volatile asm and the pointer through one more,
// so the compiler can neither hoist the reads above the computation of dep
// nor merge them with an earlier read of the same matrix (the reads are of
// read-only kernel data and would otherwise all be issued at the start)
template <int D>
__device__ __forceinline__ MSrc after(MSrc m, cd (&dep)[D][D]) {
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      __asm__ volatile("" : "+v"(dep[i][j].re));
      __asm__ volatile("" : "+v"(dep[i][j].im));
    }
  __asm__ volatile("" : "+v"(m.p));
  return m;
}

template <int D>
__device__ __forceinline__ void mload(cd (&A)[D][D], const MSrc& m) {
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) A[i][j] = mget<D>(m, i, j);
}

// C (+)= s * M B, M in memory
template <int D, bool ACC>
__device__ __forceinline__ void mul_mb(cd (&C)[D][D], const MSrc& M, const cd (&B)[D][D], double s) {
#pragma unroll
  for (int i = 0; i < D; i++) {
    cd row[D];
#pragma unroll
    for (int j = 0; j < D; j++) row[j] = cd(0, 0);
#pragma unroll
    for (int k = 0; k < D; k++) {
      const cd m = mget<D>(M, i, k);
#pragma unroll
      for (int j = 0; j < D; j++) cmac(row[j], m, B[k][j]);
    }
#pragma unroll
    for (int j = 0; j < D; j++) C[i][j] = ACC ? C[i][j] + s * row[j] : s * row[j];
  }
}

// C (+)= s * A M, M in memory
template <int D, bool ACC>
__device__ __forceinline__ void mul_am(cd (&C)[D][D], const cd (&A)[D][D], const MSrc& M, double s) {
#pragma unroll
  for (int j = 0; j < D; j++) {
    cd col[D];
#pragma unroll
    for (int i = 0; i < D; i++) col[i] = cd(0, 0);
#pragma unroll
    for (int k = 0; k < D; k++) {
      const cd m = mget<D>(M, k, j);
#pragma unroll
      for (int i = 0; i < D; i++) cmac(col[i], A[i][k], m);
    }
#pragma unroll
    for (int i = 0; i < D; i++) C[i][j] = ACC ? C[i][j] + s * col[i] : s * col[i];
  }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// hopping force, one thread per (site, mu)  (forces.cpp:204-223)
template <int NG, int D>
__global__ void __launch_bounds__(256) k_force_hop(FF f, const double2* __restrict__ links, double* __restrict__ mom,
                                                   double coeff, Geom g, const double2* __restrict__ itr) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 4 * g.V) return;
  int id = t >> 2, mu = t & 3;
  int c[4];
  fcoords(id, g, c);
  int fw = fshift(c, mu, 1, g);
  cd F[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) F[i][j] = cd(0, 0);
  {
    cd s[4][D], ps[4][D], o[4][D];
    load_full_spinor<D>(s, f.x, fw, g);          // lhs(x + mu)
    spin_mat<D>(ps, c_g5omg + mu * 16, s);
    load_full_spinor<D>(o, f.y, id, g);          // rhs(x)
    outer_acc<D>(F, ps, o);
    // the second pair of spinors is read after F is updated (fewer live registers)
    const MSrc nf = after<D>(MSrc{nullptr, 0, 1.0, false}, F);
    const int z = (int)(size_t)nf.p;
    load_full_spinor<D>(s, f.y, fw + z, g);      // rhs(x + mu)
    spin_mat<D>(ps, c_g5omg + mu * 16, s);
    load_full_spinor<D>(o, f.x, id, g);          // lhs(x)
    outer_acc<D>(F, ps, o);
  }
  cd R[D][D];
  mul_mb<D, false>(R, after<D>(link_src<D>(links, id, mu, false, g), F), F, 1.0);
  project_add<NG, D>(mom, mu, id, g.V, R, coeff, itr);
}

// clover force matrices, one thread per site  (forces.cpp:130-168)
template <int D>
__global__ void __launch_bounds__(256) k_force_clover_site(FF f, double2* __restrict__ clf, double residue, Geom g) {
  int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= g.V) return;
  cd X[4][D], Y[4][D];
  load_full_spinor<D>(X, f.x, id, g);
  load_full_spinor<D>(Y, f.y, id, g);
  for (int k = 0; k < 6; k++) {
    cd sx[4][D], sy[4][D], F[D][D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) F[i][j] = cd(0, 0);
    spin_mat<D>(sx, c_g5sig + k * 16, X);
    spin_mat<D>(sy, c_g5sig + k * 16, Y);
    outer_acc<D>(F, sx, Y);  // (g5 sigma lhs) # rhs
    outer_acc<D>(F, sy, X);  // (g5 sigma rhs) # lhs
#pragma unroll
    for (int e = 0; e < D * D; e++) {
      double2* p = clf + (size_t)(k * D * D + e) * g.V + id;
      double2 v = *p;
      *p = make_double2(v.x + residue * F[e / D][e % D].re, v.y + residue * F[e / D][e % D].im);
    }
  }
}

// log-det force: coeff * (sigma + clover)^-1 per odd site -> six plane
// combinations (clover.cpp:184-262); the inverse is the stored explicit one
template <int D>
__global__ void __launch_bounds__(128) k_force_logdet(const double2* __restrict__ inv, double2* __restrict__ clf,
                                                      double coeff, Geom g) {
  int k = blockIdx.x * 128 + threadIdx.x;
  if (k >= g.VH) return;
  constexpr int N = 2 * D;
  int id = g.VH + k;
  // element (i, j) of block b from the packed Hermitian store
  auto get = [&](int b, int i, int j) -> cd {
    int base = b * N * N / 2;
    double re, im;
    if (i == j) {
      double2 w = inv[(size_t)(base + i / 2) * g.VHpad + k];
      return cd(coeff * ((i & 1) ? w.y : w.x), 0.0);
    }
    bool lower = i > j;
    int a = lower ? i : j, c = lower ? j : i;
    int o = N + 2 * (a * (a - 1) / 2 + c);  // real index of the lower entry (a, c)
    double2 w0 = inv[(size_t)(base + o / 2) * g.VHpad + k];
    double2 w1 = inv[(size_t)(base + (o + 1) / 2) * g.VHpad + k];
    re = (o & 1) ? w0.y : w0.x;
    im = ((o + 1) & 1) ? w1.y : w1.x;
    return lower ? cd(coeff * re, coeff * im) : cd(coeff * re, -coeff * im);
  };
  for (int i = 0; i < D; i++)
    for (int j = 0; j < D; j++) {
      cd a21 = get(0, i + D, j), a12 = conj(get(0, j + D, i));
      cd a43 = get(1, i + D, j), a34 = conj(get(1, j + D, i));
      cd a11, a22, a33, a44;
      if (i < j) {
        a11 = conj(get(0, j, i)); a22 = conj(get(0, j + D, i + D));
        a33 = conj(get(1, j, i)); a44 = conj(get(1, j + D, i + D));
      } else {
        a11 = get(0, i, j); a22 = get(0, i + D, j + D);
        a33 = get(1, i, j); a44 = get(1, i + D, j + D);
      }
      cd xs[6];
      xs[0] = cd(a12.im + a21.im - a34.im - a43.im, -a12.re - a21.re + a34.re + a43.re);  // X_01
      xs[1] = cd(a12.re - a21.re + a43.re - a34.re, a12.im - a21.im + a43.im - a34.im);   // X_02
      xs[2] = cd(a22.im - a11.im + a44.im - a33.im, a11.re - a22.re + a33.re - a44.re);   // X_12
      xs[3] = cd(a11.im - a22.im + a44.im - a33.im, a22.re - a11.re + a33.re - a44.re);   // X_03
      xs[4] = cd(a12.re - a21.re + a34.re - a43.re, a12.im - a21.im + a34.im - a43.im);   // X_13
      xs[5] = cd(-a12.im - a21.im - a34.im - a43.im, a12.re + a21.re + a34.re + a43.re);  // X_23
      for (int q = 0; q < 6; q++) {
        double2* p = clf + (size_t)(q * D * D + i * D + j) * g.V + id;
        double2 v = *p;
        *p = make_double2(v.x + xs[q].re, v.y + xs[q].im);
      }
    }
}

template <int D>
__device__ __forceinline__ void load_clf(cd (&Z)[D][D], const double2* clf, int num, int id, int V) {
#pragma unroll
  for (int e = 0; e < D * D; e++) {
    double2 v = clf[(size_t)(num * D * D + e) * V + id];
    Z[e / D][e % D] = cd(v.x, v.y);
  }
}

template <int D>
__device__ __forceinline__ void dag(cd (&A)[D][D]) {
  cd T[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) T[i][j] = conj(A[j][i]);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) A[i][j] = T[i][j];
}

template <int D>
__device__ __forceinline__ void madd(cd (&A)[D][D], const cd (&B)[D][D], double s) {
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) A[i][j] = A[i][j] + s * B[i][j];
}

// clover-leaf derivative (forces.cpp:33-106).  With the links
// W0 = U(x-nu, mu)^+, W1 = U(x-nu, nu), W2 = U(x+mu, nu), W3 = U(x+nu, mu)^+,
// W4 = U(x, nu)^+, W5 = U(x+mu-nu, nu)^+, W6 = W0 W1, W7 = W2 W3,
// W8 = W7 W4 - W5 W6 and Z0..Z5 the clover-force matrices of the plane at
// x, x+mu, x-nu, x+mu-nu, x+mu+nu, x+nu:
//   F = U(x, mu) [W8 Z0 + Z1 W8 - W5 (W0 Z2 W1 + Z3 W6) + (W2 Z4 W3 + W7 Z5) W4]
// and P(x, mu) += proj(F, sign c_sw dt / 8).
//
// Work split: one launch per nu (the three nu != mu in ascending order, so
// every momentum receives its three terms in the reference's order and no
// two lanes write the same momentum); within a launch a group of D lanes
// handles one (x, mu), lane i computing row i of F.  Row i is a chain of
// row-vector by matrix products starting from u = row i of U(x, mu):
//   a = u W2, b = a W3, c = b W4, d = u W5, e = d W0, f = e W1, h = u Z1
//   F_i = (a Z4 W3 + b Z5) W4 + (c - f) Z0 - (e Z2 + d Z3 W0) W1
//         + h W2 W3 W4 - h W5 W0 W1
// (22 products, no D x D intermediate).  For each product the group stages
// the right-hand matrix in LDS (lane i fetches row i, or column i for a
// dagger, one step ahead), and every lane of the group reads the whole
// matrix from there.  The projection is summed over the group's lanes in
// fixed order.  A lane holds about seven row vectors (D complex each).
template <int D>
struct Op {
  const double2* p;  // element (r, c) of the stored matrix at p[(r * D + c) * stride]
  double sgn;
  bool dag;          // the operand is the conjugate transpose of the stored matrix
};

template <int D>
__device__ __forceinline__ Op<D> link_op(const double2* links, int id, int mu, bool dag, const Geom& g) {
  const MSrc m = link_src<D>(links, id, mu, dag, g);
  return Op<D>{m.p, m.sgn, dag};
}

template <int D>
__device__ __forceinline__ Op<D> clf_op(const double2* clf, int num, int id, int V) {
  return Op<D>{clf + (size_t)num * D * D * V + id, 1.0, false};
}

// row i of the operand (sign and conjugation applied); `stride` is VHpad for
// links and V for clover-force matrices.  The index passes through an empty
// volatile asm first: the loads are of read-only kernel data, and without it
// the compiler issues the loads of every step at the start of the kernel
template <int D>
__device__ __forceinline__ void fetch_row(double2 (&r)[D], const Op<D>& o, int i, size_t stride) {
  int z = 0;
  __asm__ volatile("" : "+v"(z));
  const double2* p = o.p + z;
#pragma unroll
  for (int j = 0; j < D; j++) {
    const double2 v = p[(size_t)(o.dag ? j * D + i : i * D + j) * stride];
    r[j] = o.dag ? make_double2(o.sgn * v.x, -o.sgn * v.y) : make_double2(o.sgn * v.x, o.sgn * v.y);
  }
}

template <int D>
__device__ __forceinline__ void put_row(double2* buf, const double2 (&r)[D], int i, bool on) {
  if (on) {
#pragma unroll
    for (int j = 0; j < D; j++) buf[i * D + j] = r[j];
  }
}

// an opaque zero that the compiler can only compute after every component
// of v is final
template <int D>
__device__ __forceinline__ int after_row(cd (&v)[D]) {
#pragma unroll
  for (int k = 0; k < D; k++) {
    __asm__ volatile("" : "+v"(v[k].re));
    __asm__ volatile("" : "+v"(v[k].im));
  }
  int z;
  __asm__ volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

// out (+)= s * (r M), M row-major in LDS.  M is read two rows at a time, and
// each pair's reads take an index that depends (through an empty volatile
// asm) on the previous pair's products: the workgroup is one wave, so the
// compiler drops its barriers and would otherwise issue the LDS reads of
// every step at once and spill them
template <int D, bool ACC>
__device__ __forceinline__ void vm(cd (&out)[D], cd (&r)[D], const double2* M, double s) {
  cd t[D];
#pragma unroll
  for (int j = 0; j < D; j++) t[j] = cd(0, 0);
  int z = after_row<D>(r);
#pragma unroll
  for (int k = 0; k < D; k++) {
    if (k > 0 && (k & 1) == 0) z = after_row<D>(t);
#pragma unroll
    for (int j = 0; j < D; j++) {
      const double2 m = M[z + k * D + j];
      cmac(t[j], r[k], cd(m.x, m.y));
    }
  }
#pragma unroll
  for (int j = 0; j < D; j++) out[j] = ACC ? out[j] + s * t[j] : s * t[j];
}

template <int NG, int D>
__global__ void __launch_bounds__(64) k_force_clover_rows(const double2* __restrict__ clf,
                                                          const double2* __restrict__ links, double* __restrict__ mom,
                                                          double coeff, Geom g, const double2* __restrict__ itr,
                                                          int nu_index) {
  constexpr int G = 64 / D;  // groups per wave
  __shared__ double2 sbuf[G][2][D * D];
  __shared__ double spart[64][NG];
  const int lane = threadIdx.x;
  const int gi = min(lane / D, G - 1);
  const int i = lane % D;
  const int group = blockIdx.x * G + gi;
  const bool on = lane < G * D && group < 4 * g.V;
  // tail lanes compute on a valid site and write nothing
  const int gq = group < 4 * g.V ? group : 4 * g.V - 1;
  const int mu = gq / g.V, id = gq - mu * g.V;
  const int nu = nu_index + (nu_index >= mu ? 1 : 0);
  int c[4];
  fcoords(id, g, c);
  const int o1 = fshift(c, mu, 1, g);  // x + mu
  const int o2 = fshift(c, nu, 1, g);  // x + nu
  const int o3 = fshift(c, nu, -1, g); // x - nu
  int c3[4], c2[4];
  fcoords(o3, g, c3);
  fcoords(o2, g, c2);
  const int o4 = fshift(c3, mu, 1, g); // x + mu - nu
  const int o5 = fshift(c2, mu, 1, g); // x + mu + nu
  int num, sign;
  if (mu < nu) { num = nu * (nu - 1) / 2 + mu; sign = 1; }
  else { num = mu * (mu - 1) / 2 + nu; sign = -1; }
  const Op<D> W0 = link_op<D>(links, o3, mu, true, g), W1 = link_op<D>(links, o3, nu, false, g),
              W2 = link_op<D>(links, o1, nu, false, g), W3 = link_op<D>(links, o2, mu, true, g),
              W4 = link_op<D>(links, id, nu, true, g), W5 = link_op<D>(links, o4, nu, true, g);
  const Op<D> Z0 = clf_op<D>(clf, num, id, g.V), Z1 = clf_op<D>(clf, num, o1, g.V),
              Z2 = clf_op<D>(clf, num, o3, g.V), Z3 = clf_op<D>(clf, num, o4, g.V),
              Z4 = clf_op<D>(clf, num, o5, g.V), Z5 = clf_op<D>(clf, num, o2, g.V);
  const size_t SL = g.VHpad, SZ = g.V;  // strides of links and of clover-force matrices
  // u = row i of U(x, mu), loaded again where it is used rather than kept live
  const Op<D> U = link_op<D>(links, id, mu, false, g);
  cd u[D];
  auto load_u = [&]() {
    double2 r[D];
    fetch_row<D>(r, U, i, SL);
#pragma unroll
    for (int j = 0; j < D; j++) u[j] = cd(r[j].x, r[j].y);
  };
  double2* B0 = sbuf[gi][0];
  double2* B1 = sbuf[gi][1];
  double2 pre[D];
  cd a[D], b[D], cc[D], d[D], e[D], s[D], t[D], acc[D];
  // each step: stage the prefetched operand, fetch the next one, multiply
#define YH_STEP(BUF, NEXT, STRIDE)  \
  put_row<D>(BUF, pre, i, on);      \
  __syncthreads();                  \
  fetch_row<D>(pre, NEXT, i, STRIDE);
  load_u();
  fetch_row<D>(pre, W2, i, SL);
  YH_STEP(B0, W3, SL) vm<D, false>(a, u, B0, 1.0);   // a = u W2
  YH_STEP(B1, Z4, SZ) vm<D, false>(b, a, B1, 1.0);   // b = a W3
  YH_STEP(B0, W3, SL) vm<D, false>(s, a, B0, 1.0);   // s = a Z4
  YH_STEP(B1, Z5, SZ) vm<D, false>(t, s, B1, 1.0);   // t = a Z4 W3
  YH_STEP(B0, W4, SL) vm<D, true>(t, b, B0, 1.0);    // t += b Z5
  YH_STEP(B1, W5, SL)
  vm<D, false>(cc, b, B1, 1.0);                      // c = b W4
  vm<D, false>(acc, t, B1, 1.0);                     // acc = (a Z4 W3 + b Z5) W4
  load_u();
  YH_STEP(B0, W0, SL) vm<D, false>(d, u, B0, 1.0);   // d = u W5
  YH_STEP(B1, W1, SL) vm<D, false>(e, d, B1, 1.0);   // e = d W0
  YH_STEP(B0, Z0, SZ)
  vm<D, false>(s, e, B0, 1.0);                       // s = e W1 (= u W5 W6)
#pragma unroll
  for (int j = 0; j < D; j++) cc[j] = cc[j] - s[j];  // c - f = u W8
  YH_STEP(B1, Z2, SZ) vm<D, true>(acc, cc, B1, 1.0); // acc += u W8 Z0
  YH_STEP(B0, Z3, SZ) vm<D, false>(s, e, B0, 1.0);   // s = e Z2
  YH_STEP(B1, W0, SL) vm<D, false>(t, d, B1, 1.0);   // t = d Z3
  YH_STEP(B0, W1, SL) vm<D, true>(s, t, B0, 1.0);    // s += d Z3 W0
  YH_STEP(B1, Z1, SZ) vm<D, true>(acc, s, B1, -1.0); // acc -= (e Z2 + d Z3 W0) W1
  load_u();
  YH_STEP(B0, W2, SL) vm<D, false>(d, u, B0, 1.0);   // h = u Z1 (in d)
  YH_STEP(B1, W3, SL) vm<D, false>(s, d, B1, 1.0);   // s = h W2
  YH_STEP(B0, W4, SL) vm<D, false>(t, s, B0, 1.0);   // t = h W2 W3
  YH_STEP(B1, W5, SL) vm<D, true>(acc, t, B1, 1.0);  // acc += h W2 W3 W4
  YH_STEP(B0, W0, SL) vm<D, false>(s, d, B0, 1.0);   // s = h W5
  YH_STEP(B1, W1, SL) vm<D, false>(t, s, B1, 1.0);   // t = h W5 W0
  put_row<D>(B0, pre, i, on);
  __syncthreads();
  vm<D, true>(acc, t, B0, -1.0);                     // acc -= h W5 W0 W1
#undef YH_STEP
  // row i of U(x, mu) F, projected: sum_k Re (iT_a)_{k i} M_{i k}
#pragma unroll
  for (int q = 0; q < NG; q++) {
    double re = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) {
      const double2 tg = itr[(q * D + k) * D + i];
      re += tg.x * acc[k].re - tg.y * acc[k].im;
    }
    spart[lane][q] = re;
  }
  __syncthreads();
  if (on && i == 0) {
    const double f = -(sign * coeff * 0.5);
#pragma unroll
    for (int q = 0; q < NG; q++) {
      double sum = 0.0;
#pragma unroll
      for (int r = 0; r < D; r++) sum += spart[lane + r][q];
      mom[(size_t)(mu * NG + q) * g.V + id] += f * sum;
    }
  }
}

// momenta: reference layout (4*id + mu)[NG] <-> device SoA
__global__ void k_mom_from_ref(const double* __restrict__ src, double* __restrict__ dst, int V, int NG) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 4 * V) return;
  int id = t >> 2, mu = t & 3;
  for (int a = 0; a < NG; a++) dst[(size_t)(mu * NG + a) * V + id] = src[(size_t)t * NG + a];
}
__global__ void k_mom_to_ref(const double* __restrict__ src, double* __restrict__ dst, int V, int NG) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 4 * V) return;
  int id = t >> 2, mu = t & 3;
  for (int a = 0; a < NG; a++) dst[(size_t)t * NG + a] = src[(size_t)(mu * NG + a) * V + id];
}
// clover force matrices -> reference cl_force layout [6*id + k] suNf
__global__ void k_clf_to_ref(const double2* __restrict__ src, double* __restrict__ dst, int V, int D) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 6 * V) return;
  int id = t / 6, k = t % 6;
  double* o = dst + (size_t)t * 2 * D * D;
  for (int e = 0; e < D * D; e++) {
    double2 v = src[(size_t)(k * D * D + e) * V + id];
    o[e] = v.x;
    o[D * D + e] = v.y;
  }
}

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------
static int ng_of(Ctx* c) { return c->nc * c->nc - 1; }

int force_alloc(Ctx* c) {
  if (!c->mom) {
    YH_HIP(hipMalloc(&c->mom, sizeof(double) * 4 * (size_t)ng_of(c) * c->g.V));
    YH_HIP(hipMemsetAsync(c->mom, 0, sizeof(double) * 4 * (size_t)ng_of(c) * c->g.V, c->stream));
  }
  if (!c->clf) {
    YH_HIP(hipMalloc(&c->clf, sizeof(double2) * 6 * (size_t)c->dr * c->dr * c->g.V));
    YH_HIP(hipMemsetAsync(c->clf, 0, sizeof(double2) * 6 * (size_t)c->dr * c->dr * c->g.V, c->stream));
  }
  return YH_OK;
}

#define YH_FORCE_MODELS(...)                                                             \
  if (c->nc == 3 && c->dr == 3) { constexpr int NG_ = 8, D_ = 3; __VA_ARGS__; }          \
  else if (c->nc == 2 && c->dr == 2) { constexpr int NG_ = 3, D_ = 2; __VA_ARGS__; }     \
  else if (c->nc == 2 && c->dr == 3) { constexpr int NG_ = 3, D_ = 3; __VA_ARGS__; }     \
  else if (c->nc == 3 && c->dr == 6) { constexpr int NG_ = 8, D_ = 6; __VA_ARGS__; }     \
  else if (c->nc == 3 && c->dr == 8) { constexpr int NG_ = 8, D_ = 8; __VA_ARGS__; }     \
  else if (c->nc == 4 && c->dr == 4) { constexpr int NG_ = 15, D_ = 4; __VA_ARGS__; }    \
  else if (c->nc == 4 && c->dr == 6) { constexpr int NG_ = 15, D_ = 6; __VA_ARGS__; }    \
  else { set_error("forces: unsupported (nc, representation)"); return YH_ERR_UNSUPPORTED; }

int force_hop(Ctx* c, double coeff, const Field* lhs, const Field* rhs) {
  int rc;
  if ((rc = force_alloc(c))) return rc;
  FF f;
  f.x[0] = lhs->half[0]; f.x[1] = lhs->half[1];
  f.y[0] = rhs->half[0]; f.y[1] = rhs->half[1];
  const double2* links = (const double2*)c->links[PREC_F64][RECON_18];
  dim3 grid((4 * c->g.V + 255) / 256);
  YH_FORCE_MODELS(hipLaunchKernelGGL((k_force_hop<NG_, D_>), grid, dim3(256), 0, c->stream, f, links, c->mom, coeff, c->g, c->tab_itr));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int force_clover_site(Ctx* c, double residue, const Field* lhs, const Field* rhs) {
  int rc;
  if ((rc = force_alloc(c))) return rc;
  FF f;
  f.x[0] = lhs->half[0]; f.x[1] = lhs->half[1];
  f.y[0] = rhs->half[0]; f.y[1] = rhs->half[1];
  dim3 grid((c->g.V + 255) / 256);
  switch (c->dr) {
    case 2: hipLaunchKernelGGL(k_force_clover_site<2>, grid, dim3(256), 0, c->stream, f, (double2*)c->clf, residue, c->g); break;
    case 3: hipLaunchKernelGGL(k_force_clover_site<3>, grid, dim3(256), 0, c->stream, f, (double2*)c->clf, residue, c->g); break;
    case 4: hipLaunchKernelGGL(k_force_clover_site<4>, grid, dim3(256), 0, c->stream, f, (double2*)c->clf, residue, c->g); break;
    case 6: hipLaunchKernelGGL(k_force_clover_site<6>, grid, dim3(256), 0, c->stream, f, (double2*)c->clf, residue, c->g); break;
    case 8: hipLaunchKernelGGL(k_force_clover_site<8>, grid, dim3(256), 0, c->stream, f, (double2*)c->clf, residue, c->g); break;
    default: set_error("unsupported representation dimension"); return YH_ERR_UNSUPPORTED;
  }
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int force_logdet(Ctx* c, double mass, double coeff) {
  int rc;
  if ((rc = force_alloc(c))) return rc;
  if ((rc = clover_invert(c, 4.0 + mass))) return rc;
  dim3 grid((c->g.VH + 127) / 128);
  const double2* inv = (const double2*)c->clinv[PREC_F64];
  switch (c->dr) {
    case 2: hipLaunchKernelGGL(k_force_logdet<2>, grid, dim3(128), 0, c->stream, inv, (double2*)c->clf, coeff, c->g); break;
    case 3: hipLaunchKernelGGL(k_force_logdet<3>, grid, dim3(128), 0, c->stream, inv, (double2*)c->clf, coeff, c->g); break;
    case 4: hipLaunchKernelGGL(k_force_logdet<4>, grid, dim3(128), 0, c->stream, inv, (double2*)c->clf, coeff, c->g); break;
    case 6: hipLaunchKernelGGL(k_force_logdet<6>, grid, dim3(128), 0, c->stream, inv, (double2*)c->clf, coeff, c->g); break;
    case 8: hipLaunchKernelGGL(k_force_logdet<8>, grid, dim3(128), 0, c->stream, inv, (double2*)c->clf, coeff, c->g); break;
    default: set_error("unsupported representation dimension"); return YH_ERR_UNSUPPORTED;
  }
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int force_clover_core(Ctx* c, double dt) {
  int rc;
  if ((rc = force_alloc(c))) return rc;
  double coeff = c->csw * dt / 8.0;
  const double2* links = (const double2*)c->links[PREC_F64][RECON_18];
  // one launch per nu (ascending), so each momentum gets its three terms in
  // the reference's order
  for (int j = 0; j < 3; j++) {
    YH_FORCE_MODELS(const int groups = 64 / D_;
                    dim3 grid((unsigned)((4L * c->g.V + groups - 1) / groups));
                    hipLaunchKernelGGL((k_force_clover_rows<NG_, D_>), grid, dim3(64), 0, c->stream,
                                       (const double2*)c->clf, links, c->mom, coeff, c->g, c->tab_itr, j));
    YH_HIP(hipGetLastError());
  }
  return YH_OK;
}

int clover_force_zero(Ctx* c) {
  int rc;
  if ((rc = force_alloc(c))) return rc;
  YH_HIP(hipMemsetAsync(c->clf, 0, sizeof(double2) * 6 * (size_t)c->dr * c->dr * c->g.V, c->stream));
  return YH_OK;
}

int momenta_upload(Ctx* c, const double* host) {
  int rc;
  if ((rc = force_alloc(c))) return rc;
  size_t n = (size_t)4 * c->g.V * ng_of(c);
  double* d;
  YH_HIP(hipMalloc(&d, n * sizeof(double)));
  hipError_t e = hipMemcpyAsync(d, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    hipLaunchKernelGGL(k_mom_from_ref, dim3((4 * c->g.V + 255) / 256), dim3(256), 0, c->stream, d, c->mom, c->g.V, ng_of(c));
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return YH_ERR_HIP; }
  return YH_OK;
}

int momenta_download(Ctx* c, double* host) {
  int rc;
  if ((rc = force_alloc(c))) return rc;
  size_t n = (size_t)4 * c->g.V * ng_of(c);
  double* d;
  YH_HIP(hipMalloc(&d, n * sizeof(double)));
  hipLaunchKernelGGL(k_mom_to_ref, dim3((4 * c->g.V + 255) / 256), dim3(256), 0, c->stream, c->mom, d, c->g.V, ng_of(c));
  hipError_t e = hipMemcpyAsync(host, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return YH_ERR_HIP; }
  return YH_OK;
}

int clover_force_download(Ctx* c, double* host) {
  int rc;
  if ((rc = force_alloc(c))) return rc;
  size_t n = (size_t)6 * c->g.V * 2 * c->dr * c->dr;
  double* d;
  YH_HIP(hipMalloc(&d, n * sizeof(double)));
  hipLaunchKernelGGL(k_clf_to_ref, dim3((6 * c->g.V + 255) / 256), dim3(256), 0, c->stream, (const double2*)c->clf, d,
                     c->g.V, c->dr);
  hipError_t e = hipMemcpyAsync(host, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) { set_error(hipGetErrorString(e)); return YH_ERR_HIP; }
  return YH_OK;
}

}  // namespace yh
