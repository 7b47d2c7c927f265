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
#include <utility>
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

// spinor of site id; zero on the time slices the open and SF boundaries
// clear (bcs.cpp:139-164): the reference's Dphi_oe(X, X) at the start of the
// force zeroes those slices of the whole field, both halves
template <int D>
__device__ __forceinline__ void load_full_spinor(cd (&s)[4][D], const void* const (&h)[2], int id, const Geom& g) {
  int p = id >= g.VH;
  load_spinor<PREC_F64, D, double>(s, h[p], nullptr, id - p * g.VH, g.VHpad);
  if (g.bc == BC_OPEN || g.bc == BC_SF) {
    int c[4];
    fcoords(id, g, c);
    if (bc_zero_slice(c[0], g)) {
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int i = 0; i < D; i++) s[a][i] = cd(0, 0);
    }
  }
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
// k_force_clover_x: the three nu of one (x, mu) in one pass (the sign folded
// into F, one projection per (x, mu)), with the reference's association of
// the products (forces.cpp:80-96), row by row:
//   W6 = W0 L(W1)          T1 = W0 L(Z2)          S = T1 L(W1) + Z3 L(W6)
//   W8 = W7 L(W4) - W5 L(W6)   with W7 = W2 L(W3)
//   Q  = (W2 L(Z4)) L(W3) + W7 L(Z5)
//   F += sign [Q L(W4) + W8 L(Z0) + Z1 L(W8) - W5 L(S)]
// and after the three nu one product M = U(x, mu) L(F), projected.  L(X)
// is a matrix the group reads whole from LDS; the others are row vectors in
// registers (row i in lane i of the group).  14 products per nu and one at
// the end: 43 per (x, mu), against 48 for the reference's per-nu U fmat.
//
// The right-hand link and clover-force matrices reach LDS by direct DMA
// (global_load_lds, 16 bytes a lane): for column j lane (g, i) fetches the
// stored element (i, j) into slot word j * 64 + lane, so the stored (r, c)
// of group g sits at c * 64 + g D + r, and a dagger operand is read
// transposed and conjugated.  Computed matrices are written the same way.
// Three slots (three LDS objects) rotate.  A step issues the DMAs of later
// steps (lgkmcnt 0 first, so a slot's earlier reads and writes are done
// before it is refilled), waits with vmcnt for the DMA of its own slot
// (vector memory operations complete in issue order; the count is the DMA
// instructions issued after it, D per matrix, and row loads only add to the
// operations after it), and multiplies.  The slot reads are inline asm that
// completes its own reads, since the compiler would otherwise wait for every
// outstanding DMA before each one.  The block is one wave: LDS traffic needs
// no barrier, only the waits and compiler fences.  Work items run mu
// fastest, each XCD taking a contiguous slab of the lattice (xcd_block), so
// the four mu of a site share its links and clover-force matrices in L2.
template <int D>
struct FX {
  static constexpr int G = 64 / D;          // groups (x, mu) per wave
  static constexpr int SLOT = D * 64;       // double2 words per slot
};

// step boundary: memory operations stay on their side of it (program order
// of the issues); the waits are the compiler's own
__device__ __forceinline__ void fx_wait() { __asm__ volatile("" ::: "memory"); }

// an integer the compiler cannot see through: addresses built from it are
// computed where they are used instead of being hoisted and kept live
__device__ __forceinline__ int fx_opq(int v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}

// element 0 of the stored matrix: represented link (site, dir) or clover
// force matrix (plane num, site), and its element stride
template <int D>
__device__ __forceinline__ const double2* fx_link(const double2* links, int site, int dir, const Geom& g) {
  const int p = site >= g.VH;
  return links + (size_t)(p * 4 + dir) * D * D * g.VHpad + (site - p * g.VH);
}
template <int D>
__device__ __forceinline__ const double2* fx_clf(const double2* clf, int num, int site, const Geom& g) {
  return clf + (size_t)num * D * D * g.V + site;
}

// DMA the stored matrix into a slot: lane (g, i) fetches its row i.  The
// slot's earlier LDS reads and writes are complete first (lgkmcnt 0), so the
// asynchronous DMA write cannot overtake them
template <int D>
__device__ __forceinline__ void fx_dma(double2* slot, const double2* p, int i, size_t stride) {
  __asm__ volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0) only
  __asm__ volatile("" ::: "memory");
  const double2* q = p + (size_t)(i * D) * stride;
#pragma unroll
  for (int j = 0; j < D; j++) {
    __attribute__((address_space(3))) void* dst = (__attribute__((address_space(3))) void*)(slot + j * 64);
    __builtin_amdgcn_global_load_lds((const void*)(q + (size_t)j * stride), dst, 16, 0, 0);
  }
}

// row i of the stored matrix (DAG: of its conjugate transpose) into registers
template <int D, bool DAG>
__device__ __forceinline__ void fx_row(cd (&r)[D], const double2* p, int i, size_t stride) {
  const double2* q = p + (size_t)(DAG ? i : i * D) * stride;
#pragma unroll
  for (int j = 0; j < D; j++) {
    const double2 v = q[(size_t)(DAG ? j * D : j) * stride];
    r[j] = DAG ? cd(v.x, -v.y) : cd(v.x, v.y);
  }
}

// row i of a register matrix into a slot (stored layout)
template <int D>
__device__ __forceinline__ void fx_put(double2* slot, const cd (&r)[D], int lane) {
#pragma unroll
  for (int j = 0; j < D; j++) slot[j * 64 + lane] = make_double2(r[j].re, r[j].im);
}

// LDS reads of the clover-force kernel's slots, written as inline asm so
// that the compiler does not add its own wait for every outstanding DMA in
// front of them (it cannot tell the slots apart across the loop): N 16-byte
// reads at immediate offsets from one address, completed inside the block
// (lgkmcnt 0), so no output register is left in flight.  The caller waits
// for the DMA that filled the slot (vmcnt) before the block.
typedef double fx_d2 __attribute__((ext_vector_type(2)));
template <int O0, int O1, int O2>
__device__ __forceinline__ void fx_rd(fx_d2 (&v)[3], unsigned a) {
  __asm__ volatile(
      "ds_read_b128 %0, %3 offset:%4\n\t"
      "ds_read_b128 %1, %3 offset:%5\n\t"
      "ds_read_b128 %2, %3 offset:%6\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=v"(v[0]), "=v"(v[1]), "=v"(v[2])
      : "v"(a), "i"(O0), "i"(O1), "i"(O2));
}
template <int O0, int O1, int O2, int O3>
__device__ __forceinline__ void fx_rd(fx_d2 (&v)[4], unsigned a) {
  __asm__ volatile(
      "ds_read_b128 %0, %4 offset:%5\n\t"
      "ds_read_b128 %1, %4 offset:%6\n\t"
      "ds_read_b128 %2, %4 offset:%7\n\t"
      "ds_read_b128 %3, %4 offset:%8\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=v"(v[0]), "=v"(v[1]), "=v"(v[2]), "=v"(v[3])
      : "v"(a), "i"(O0), "i"(O1), "i"(O2), "i"(O3));
}
template <int O0, int O1, int O2, int O3, int O4, int O5>
__device__ __forceinline__ void fx_rd(fx_d2 (&v)[6], unsigned a) {
  __asm__ volatile(
      "ds_read_b128 %0, %6 offset:%7\n\t"
      "ds_read_b128 %1, %6 offset:%8\n\t"
      "ds_read_b128 %2, %6 offset:%9\n\t"
      "ds_read_b128 %3, %6 offset:%10\n\t"
      "ds_read_b128 %4, %6 offset:%11\n\t"
      "ds_read_b128 %5, %6 offset:%12\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=v"(v[0]), "=v"(v[1]), "=v"(v[2]), "=v"(v[3]), "=v"(v[4]), "=v"(v[5])
      : "v"(a), "i"(O0), "i"(O1), "i"(O2), "i"(O3), "i"(O4), "i"(O5));
}
template <int O0, int O1, int O2, int O3, int O4, int O5, int O6, int O7>
__device__ __forceinline__ void fx_rd(fx_d2 (&v)[8], unsigned a) {
  __asm__ volatile(
      "ds_read_b128 %0, %8 offset:%9\n\t"
      "ds_read_b128 %1, %8 offset:%10\n\t"
      "ds_read_b128 %2, %8 offset:%11\n\t"
      "ds_read_b128 %3, %8 offset:%12\n\t"
      "ds_read_b128 %4, %8 offset:%13\n\t"
      "ds_read_b128 %5, %8 offset:%14\n\t"
      "ds_read_b128 %6, %8 offset:%15\n\t"
      "ds_read_b128 %7, %8 offset:%16\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=v"(v[0]), "=v"(v[1]), "=v"(v[2]), "=v"(v[3]), "=v"(v[4]), "=v"(v[5]), "=v"(v[6]), "=v"(v[7])
      : "v"(a), "i"(O0), "i"(O1), "i"(O2), "i"(O3), "i"(O4), "i"(O5), "i"(O6), "i"(O7));
}
template <int O0, int O1, int O2, int O3, int O4, int O5, int O6, int O7, int O8, int O9, int O10, int O11>
__device__ __forceinline__ void fx_rd(fx_d2 (&v)[12], unsigned a) {
  __asm__ volatile(
      "ds_read_b128 %0, %12 offset:%13\n\t"
      "ds_read_b128 %1, %12 offset:%14\n\t"
      "ds_read_b128 %2, %12 offset:%15\n\t"
      "ds_read_b128 %3, %12 offset:%16\n\t"
      "ds_read_b128 %4, %12 offset:%17\n\t"
      "ds_read_b128 %5, %12 offset:%18\n\t"
      "ds_read_b128 %6, %12 offset:%19\n\t"
      "ds_read_b128 %7, %12 offset:%20\n\t"
      "ds_read_b128 %8, %12 offset:%21\n\t"
      "ds_read_b128 %9, %12 offset:%22\n\t"
      "ds_read_b128 %10, %12 offset:%23\n\t"
      "ds_read_b128 %11, %12 offset:%24\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=v"(v[0]), "=v"(v[1]), "=v"(v[2]), "=v"(v[3]), "=v"(v[4]), "=v"(v[5]), "=v"(v[6]), "=v"(v[7]), "=v"(v[8]), "=v"(v[9]), "=v"(v[10]), "=v"(v[11])
      : "v"(a), "i"(O0), "i"(O1), "i"(O2), "i"(O3), "i"(O4), "i"(O5), "i"(O6), "i"(O7), "i"(O8), "i"(O9), "i"(O10), "i"(O11));
}
template <int O0, int O1, int O2, int O3, int O4, int O5, int O6, int O7, int O8, int O9, int O10, int O11, int O12, int O13, int O14, int O15>
__device__ __forceinline__ void fx_rd(fx_d2 (&v)[16], unsigned a) {
  __asm__ volatile(
      "ds_read_b128 %0, %16 offset:%17\n\t"
      "ds_read_b128 %1, %16 offset:%18\n\t"
      "ds_read_b128 %2, %16 offset:%19\n\t"
      "ds_read_b128 %3, %16 offset:%20\n\t"
      "ds_read_b128 %4, %16 offset:%21\n\t"
      "ds_read_b128 %5, %16 offset:%22\n\t"
      "ds_read_b128 %6, %16 offset:%23\n\t"
      "ds_read_b128 %7, %16 offset:%24\n\t"
      "ds_read_b128 %8, %16 offset:%25\n\t"
      "ds_read_b128 %9, %16 offset:%26\n\t"
      "ds_read_b128 %10, %16 offset:%27\n\t"
      "ds_read_b128 %11, %16 offset:%28\n\t"
      "ds_read_b128 %12, %16 offset:%29\n\t"
      "ds_read_b128 %13, %16 offset:%30\n\t"
      "ds_read_b128 %14, %16 offset:%31\n\t"
      "ds_read_b128 %15, %16 offset:%32\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=v"(v[0]), "=v"(v[1]), "=v"(v[2]), "=v"(v[3]), "=v"(v[4]), "=v"(v[5]), "=v"(v[6]), "=v"(v[7]), "=v"(v[8]), "=v"(v[9]), "=v"(v[10]), "=v"(v[11]), "=v"(v[12]), "=v"(v[13]), "=v"(v[14]), "=v"(v[15])
      : "v"(a), "i"(O0), "i"(O1), "i"(O2), "i"(O3), "i"(O4), "i"(O5), "i"(O6), "i"(O7), "i"(O8), "i"(O9), "i"(O10), "i"(O11), "i"(O12), "i"(O13), "i"(O14), "i"(O15));
}

// byte offset of element (k, j) of the group's matrix B in a slot, from the
// group's base: stored (k, j) at word j * 64 + k, or for B = stored^+ the
// stored (j, k) at word k * 64 + j
template <bool DAG>
__host__ __device__ constexpr int fx_off(int k, int j) { return DAG ? k * 1024 + j * 16 : j * 1024 + k * 16; }

template <int D, bool DAG, int K0, int NK, size_t... I>
__device__ __forceinline__ void fx_rd_rows(fx_d2 (&v)[NK * D], unsigned a, std::index_sequence<I...>) {
  fx_rd<fx_off<DAG>(K0 + (int)I / D, (int)I % D)...>(v, a);
}

// out (+)= s (r B), B the group's matrix in a slot (DAG: B = stored^+),
// `a` the LDS byte address of the slot plus 16 g D; out and r are different
// arrays.  B is read two rows at a time.
template <int D, bool DAG, bool ACC, int K0 = 0>
__device__ __forceinline__ void fx_vm(cd (&out)[D], cd (&r)[D], unsigned a, double s) {
  if constexpr (K0 < D) {
    constexpr int NK = (D - K0 >= 2 && D != 8) ? 2 : 1;
    fx_d2 v[NK * D];
    fx_rd_rows<D, DAG, K0, NK>(v, a, std::make_index_sequence<NK * D>{});
#pragma unroll
    for (int dk = 0; dk < NK; dk++) {
      const int k = K0 + dk;
      const cd rk = s * r[k];
#pragma unroll
      for (int j = 0; j < D; j++) {
        const fx_d2 m = v[dk * D + j];
        const double mi = DAG ? -m.y : m.y;
        // out += rk m as four fused multiply-adds
        double re = (!ACC && k == 0) ? 0.0 : out[j].re, im = (!ACC && k == 0) ? 0.0 : out[j].im;
        re = __builtin_fma(rk.re, m.x, re);
        re = __builtin_fma(-rk.im, mi, re);
        im = __builtin_fma(rk.re, mi, im);
        im = __builtin_fma(rk.im, m.x, im);
        out[j] = cd(re, im);
      }
    }
    fx_vm<D, DAG, ACC, K0 + NK>(out, r, a, s);
  }
}

// wait until at most N vector memory operations are outstanding (the DMA
// that filled the slot about to be read has landed: DMAs and loads complete
// in issue order)
template <int N>
__device__ __forceinline__ void fx_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __asm__ volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
  __asm__ volatile("" ::: "memory");
}

__device__ __forceinline__ unsigned fx_lds(const double2* p) {
  return (unsigned)(size_t)(__attribute__((address_space(3))) const double2*)p;
}

// antiperiodic sign of the represented link (site with coordinates c, dir)
// (bcs.cpp:264-272)
__device__ __forceinline__ double fx_sgn(const int (&c)[4], int dir, const Geom& g) {
  return (g.bc == BC_ANTIPERIODIC && dir == 0 && c[0] == 0) ? -1.0 : 1.0;
}

// coordinates of c + d e_dir (d = +-1, periodic wrap) and their EO index
__device__ __forceinline__ int fx_nb(const int (&c)[4], int dir, int d, const Geom& g, int (&o)[4]) {
#pragma unroll
  for (int a = 0; a < 4; a++) o[a] = c[a];
#pragma unroll
  for (int a = 0; a < 4; a++)
    if (a == dir) {
      int v = o[a] + d;
      v = v < 0 ? v + g.L[a] : (v >= g.L[a] ? v - g.L[a] : v);
      o[a] = v;
    }
  return full_index(o[0], o[1], o[2], o[3], g);
}

template <int NG, int D>
__global__ void __launch_bounds__(64) k_force_clover_x(const double2* __restrict__ clf,
                                                       const double2* __restrict__ links, double* __restrict__ mom,
                                                       double coeff, Geom g, const double2* __restrict__ itr,
                                                       int order) {
  typedef FX<D> F;
  __shared__ __align__(16) double2 sl0[F::SLOT], sl1[F::SLOT], sl2[F::SLOT];
  double2* const sl[3] = {sl0, sl1, sl2};
  const int lane = threadIdx.x;
  const int gi = min(lane / D, F::G - 1);
  const int i = lane % D;
  const int gD = gi * D;
  const unsigned a0 = fx_lds(sl0) + 16 * gD, a1 = fx_lds(sl1) + 16 * gD, a2 = fx_lds(sl2) + 16 * gD;
  // work item: (site block, mu), from a 2D grid (order < 0) or a 1D grid
  // with mu fastest, remapped so that each XCD takes runs of `order` items
  int sb, mu;
  if (order < 0) { sb = blockIdx.x; mu = blockIdx.y; }
  else { const int b = xcd_block(gridDim.x, order); sb = b >> 2; mu = b & 3; }
  const int grp = sb * F::G + gi;
  const bool on = lane < F::G * D && grp < g.V;
  const int id = grp < g.V ? grp : g.V - 1;  // tail lanes compute on a valid site and write nothing
  const size_t SL = g.VHpad, SZ = g.V;
  int c[4], c1[4];
  fcoords(id, g, c);
  const int o1 = fx_nb(c, mu, 1, g, c1);  // x + mu
  cd Ft[D], S[D], W8[D], W7[D], Q[D], T[D], A[D];
#pragma unroll
  for (int j = 0; j < D; j++) Ft[j] = cd(0, 0);
  // Each step issues DMAs for later steps (in the order they are needed),
  // waits for the DMA of its own slot by counting the DMA instructions
  // issued after it (D per matrix), and multiplies.  The row operands are
  // ordinary loads the compiler waits for.
  // prologue: W3 = U(x + nu, mu)^+ and the row of W2 = U(x + mu, nu) for the first nu
  {
    const int nu = mu == 0 ? 1 : 0;
    int cn[4];
    fx_dma<D>(sl[0], fx_link<D>(links, fx_nb(c, nu, 1, g, cn), mu, g), i, SL);
    fx_row<D, false>(A, fx_link<D>(links, o1, nu, g), i, SL);
  }
#pragma unroll 1
  for (int jn = 0; jn < 3; jn++) {
    const int nu = jn + (jn >= mu ? 1 : 0);
    int c2[4], c3[4], c4[4], c5[4];
    const int o2 = fx_nb(c, nu, 1, g, c2);    // x + nu
    const int o3 = fx_nb(c, nu, -1, g, c3);   // x - nu
    const int o4 = fx_nb(c3, mu, 1, g, c4);   // x + mu - nu
    const int o5 = fx_nb(c2, mu, 1, g, c5);   // x + mu + nu
    int num;
    double sign;
    if (mu < nu) { num = nu * (nu - 1) / 2 + mu; sign = 1.0; }
    else { num = mu * (mu - 1) / 2 + nu; sign = -1.0; }
    // antiperiodic signs of W0..W5
    const double s0 = fx_sgn(c3, mu, g), s1 = fx_sgn(c3, nu, g), s2 = fx_sgn(c1, nu, g),
                 s3 = fx_sgn(c2, mu, g), s4 = fx_sgn(c, nu, g), s5 = fx_sgn(c4, nu, g);
    // W7 = W2 W3                                (A: row of W2; W3 in s0)
    fx_dma<D>(sl[1], fx_clf<D>(clf, num, fx_opq(o5), g), fx_opq(i), SZ);  // Z4
    fx_dma<D>(sl[2], fx_clf<D>(clf, num, fx_opq(o2), g), fx_opq(i), SZ);  // Z5
    fx_vmcnt<2 * D>();                                                      // W3
    fx_vm<D, true, false>(W7, A, a0, s2 * s3);
    fx_wait();
    // T = W2 Z4                                                  Z4 s1
    fx_vmcnt<D>();
    fx_vm<D, false, false>(T, A, a1, s2);
    fx_wait();
    // Q = T W3                                                   W3 s0
    fx_dma<D>(sl[1], fx_link<D>(links, fx_opq(id), nu, g), fx_opq(i), SL);  // W4
    fx_vm<D, true, false>(Q, T, a0, s3);
    fx_wait();
    // Q += W7 Z5                                                 Z5 s2
    fx_dma<D>(sl[0], fx_link<D>(links, fx_opq(o3), nu, g), fx_opq(i), SL);  // W1
    fx_row<D, true>(A, fx_link<D>(links, fx_opq(o3), mu, g), fx_opq(i), SL);  // W0
    fx_vmcnt<2 * D>();
    fx_vm<D, false, true>(Q, W7, a2, 1.0);
    fx_wait();
    // F += sign Q W4                                             W4 s1
    fx_dma<D>(sl[2], fx_clf<D>(clf, num, fx_opq(o3), g), fx_opq(i), SZ);  // Z2
    fx_vmcnt<2 * D>();
    fx_vm<D, true, true>(Ft, Q, a1, sign * s4);
    fx_wait();
    // W8 = W7 W4                                                 W4 s1
    fx_vm<D, true, false>(W8, W7, a1, s4);
    fx_wait();
    // W6 = W0 W1 (in T)                                          W1 s0
    fx_vmcnt<D>();
    fx_vm<D, false, false>(T, A, a0, s0 * s1);
    fx_wait();
    // T1 = W0 Z2 (in T), W6 -> s1                                Z2 s2
    fx_put<D>(sl[1], T, lane);
    fx_vmcnt<0>();
    fx_vm<D, false, false>(T, A, a2, s0);
    fx_wait();
    // S = T1 W1                                                  W1 s0
    fx_dma<D>(sl[2], fx_clf<D>(clf, num, fx_opq(id), g), fx_opq(i), SZ);  // Z0
    fx_row<D, true>(A, fx_link<D>(links, fx_opq(o4), nu, g), fx_opq(i), SL);  // W5
    fx_vm<D, false, false>(S, T, a0, s1);
    fx_wait();
    // W8 -= W5 W6                                                W6 s1
    fx_vm<D, false, true>(W8, A, a1, -s5);
    fx_row<D, false>(A, fx_clf<D>(clf, num, fx_opq(o4), g), fx_opq(i), SZ);  // Z3
    fx_wait();
    // S += Z3 W6                                                 W6 s1
    fx_vm<D, false, true>(S, A, a1, 1.0);
    fx_row<D, false>(A, fx_clf<D>(clf, num, fx_opq(o1), g), fx_opq(i), SZ);  // Z1
    fx_wait();
    // F += sign W8 Z0; W8 -> s0, S -> s1                         Z0 s2
    fx_put<D>(sl[0], W8, lane);
    fx_put<D>(sl[1], S, lane);
    fx_vmcnt<0>();
    fx_vm<D, false, true>(Ft, W8, a2, sign);
    fx_wait();
    // F += sign Z1 W8                                            W8 s0
    fx_vm<D, false, true>(Ft, A, a0, sign);
    fx_row<D, true>(A, fx_link<D>(links, fx_opq(o4), nu, g), fx_opq(i), SL);  // W5
    fx_wait();
    // F -= sign W5 S                                             S s1
    // with the next nu's W3 and W2 row, or after the last nu the row of U
    if (jn < 2) {
      const int nx = jn + 1 + (jn + 1 >= mu ? 1 : 0);
      int cn[4];
      fx_dma<D>(sl[0], fx_link<D>(links, fx_opq(fx_nb(c, nx, 1, g, cn)), mu, g), fx_opq(i), SL);
      fx_vm<D, false, true>(Ft, A, a1, -sign * s5);
      fx_row<D, false>(A, fx_link<D>(links, fx_opq(o1), nx, g), fx_opq(i), SL);
    } else {
      fx_vm<D, false, true>(Ft, A, a1, -sign * s5);
      fx_row<D, false>(A, fx_link<D>(links, fx_opq(id), mu, g), fx_opq(i), SL);
    }
    fx_wait();
  }
  // M = U(x, mu) F, then the projection sum_{k,r} Re (iT_q)_{k r} M_{r k}
  // with lane i of the group taking q = i, i + D, ...
  fx_put<D>(sl[2], Ft, lane);
  fx_wait();
  fx_vm<D, false, false>(T, A, a2, fx_sgn(c, mu, g));
  fx_put<D>(sl[0], T, lane);
  fx_wait();
  const double f = -coeff * 0.5;
#pragma unroll
  for (int q0 = 0; q0 < NG; q0 += D) {
    const int q = q0 + i;
    if (q < NG) {
      double re = 0.0;
#pragma unroll
      for (int k = 0; k < D; k++)
#pragma unroll
        for (int r = 0; r < D; r++) {
          const double2 tg = itr[(q * D + k) * D + r];
          const double2 m = sl0[k * 64 + gD + r];
          re += tg.x * m.x - tg.y * m.y;
        }
      if (on) mom[(size_t)(mu * NG + q) * g.V + id] += f * re;
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
  YH_FORCE_MODELS(const int groups = 64 / D_;
                  const unsigned nsb = (unsigned)((c->g.V + groups - 1) / groups);
                  dim3 grid = c->force_order < 0 ? dim3(nsb, 4) : dim3(4 * nsb);
                  hipLaunchKernelGGL((k_force_clover_x<NG_, D_>), grid, dim3(64), 0, c->stream,
                                     (const double2*)c->clf, links, c->mom, coeff, c->g, c->tab_itr,
                                     c->force_order));
  YH_HIP(hipGetLastError());
  return YH_OK;
}

// zero the time slices the open / SF boundaries clear (bcs.cpp:139-164) in
// one half of an FP64 field
__global__ void __launch_bounds__(256) k_zero_bc_half(double2* __restrict__ h, int parity, int ncomp, Geom g) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= g.VH) return;
  int c[4];
  site_coords(k, parity, g, c);
  if (!bc_zero_slice(c[0], g)) return;
  for (int v = 0; v < ncomp; v++) h[(size_t)v * g.VHpad + k] = make_double2(0.0, 0.0);
}

int zero_bc_half(Ctx* c, Field* f, int parity) {
  if (c->g.bc != BC_OPEN && c->g.bc != BC_SF) return YH_OK;
  if (f->prec != PREC_F64) { set_error("zero_bc_half: FP64 fields only"); return YH_ERR_ARG; }
  hipLaunchKernelGGL(k_zero_bc_half, dim3((c->g.VH + 255) / 256), dim3(256), 0, c->stream, (double2*)f->half[parity],
                     parity, 4 * c->dr, c->g);
  YH_HIP(hipGetLastError());
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
