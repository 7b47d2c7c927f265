"""CPU oracle — test infrastructure only.

ctypes bindings for the plain-C FP64 restatement of the reference's fermion
hot path (oracle/oracle.c).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this package, and only as the checker;
the product package (yahmc_amd) never does.

All arrays use the reference's host layouts (see oracle.c's header):
spinor fields are float64 arrays of shape (V, 8*DR), fundamental links
(V*4, 2*NC*NC), represented links (V*4, 2*DR*DR).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

BC_PERIODIC, BC_ANTIPERIODIC, BC_OPEN, BC_SF = 0, 1, 2, 3
MVM_DPHI, MVM_HPHI, MVM_HPHI_SQ = 0, 1, 2

MODELS = {
    "su3_fund": (3, 0),
    "su2_fund": (2, 0),
    "su2_adj": (2, 1),
    "su3_adj": (3, 1),
    "su3_sym": (3, 2),
    "su3_asym": (3, 3),
    "su4_asym": (4, 3),
}

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build(quiet=True):
    """Compile every oracle model with gcc (oracle/Makefile)."""
    out = subprocess.run(["make", "-C", HERE, "-j8"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


class Oracle:
    """One (N, representation) model of the oracle.  The C side keeps global
    state (like the reference), so one lattice is live per model at a time."""

    _cache = {}

    def __init__(self, model="su3_fund"):
        path = os.path.join(LIBDIR, f"liboracle_{model}.so")
        if not os.path.exists(path):
            build()
        if path not in Oracle._cache:
            Oracle._cache[path] = ctypes.CDLL(path)
        self.lib = lib = Oracle._cache[path]
        self.model = model
        self.nc = lib.or_nc()
        self.dr = lib.or_dr()
        self.repr = lib.or_repr()
        lib.or_init.argtypes = [ctypes.c_int] * 5 + [ctypes.c_double, ctypes.c_double]
        for name in ("or_Dphi_oe", "or_Dphi_eo"):
            getattr(lib, name).argtypes = [_D, _D]
        for name in ("or_Dphi_ee", "or_Dphi_oo", "or_Dphi_oo_inv", "or_Dphi", "or_Hphi", "or_Hphi_sq"):
            getattr(lib, name).argtypes = [ctypes.c_double, _D, _D]
        lib.or_mvm.argtypes = [ctypes.c_int, ctypes.c_double, _D, _D]
        lib.or_sqnorm.argtypes = [_D]
        lib.or_sqnorm.restype = ctypes.c_double
        lib.or_product.argtypes = [_D, _D, _D]
        lib.or_product_re.argtypes = [_D, _D]
        lib.or_product_re.restype = ctypes.c_double
        lib.or_cg.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, _D, _D]
        lib.or_bicgstab.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, _D, _D]
        lib.or_auto_inverter.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, _D]
        lib.or_cg_mshift.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _D, _D, _D]
        lib.or_clover_logdet.argtypes = [ctypes.c_double]
        lib.or_clover_logdet.restype = ctypes.c_double
        lib.or_compute_ldl_decomp.argtypes = [ctypes.c_double]
        lib.or_represent_link.argtypes = [_D, _D]
        lib.or_get_generators.argtypes = [_D, _D]
        for name in ("or_set_gauge", "or_get_gauge", "or_get_repr", "or_set_repr", "or_get_clover",
                     "or_get_ldl", "or_get_clover_force", "or_get_momenta", "or_set_momenta",
                     "or_apply_bcs_on_spinor", "or_g5"):
            getattr(lib, name).argtypes = [_D]
        lib.or_coords.argtypes = [ctypes.c_int, _I]
        lib.or_clover_fermion_force.argtypes = [ctypes.c_double, _D, _D]
        lib.or_wilson_fermion_force.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D]
        lib.or_wilson_fermion_force_q.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _D]
        lib.or_clover_logdet_force.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.or_end_fermion_force.argtypes = [ctypes.c_double]
        lib.or_hmc_action.argtypes = [ctypes.c_double, ctypes.c_double, _D]
        lib.or_hmc_action.restype = ctypes.c_double
        lib.or_spin_matrix.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]
        lib.or_wilson_gauge_force.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.or_plaquette.restype = ctypes.c_double
        lib.or_gauge_action.argtypes = [ctypes.c_double]
        lib.or_avr_wilson.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.or_avr_wilson.restype = ctypes.c_double
        lib.or_improved_gauge_force.argtypes = [ctypes.c_double] * 4
        lib.or_gauge_volume.restype = ctypes.c_double
        lib.or_gauge_action.restype = ctypes.c_double
        lib.or_kinetic.restype = ctypes.c_double
        lib.or_update_links.argtypes = [ctypes.c_double]
        lib.or_minres.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, _D, _D]
        lib.or_minres.restype = ctypes.c_int
        lib.or_ptc_layer.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D, ctypes.c_int]
        lib.or_ptc_wgrad.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D]

    # ---- lattice -------------------------------------------------------
    def init(self, dims, bc=BC_ANTIPERIODIC, csw=1.0, cf=1.0):
        self.dims = tuple(int(d) for d in dims)
        self.bc, self.csw, self.cf = bc, csw, cf
        rc = self.lib.or_init(*self.dims, bc, csw, cf)
        if rc != 0:
            raise ValueError(f"unsupported lattice {dims}")
        self.V = int(np.prod(self.dims))
        self.VH = self.V // 2
        return self

    def index(self, t, x, y, z):
        return self.lib.or_index(t, x, y, z)

    def coords(self):
        out = np.zeros((self.V, 4), dtype=np.int32)
        c = (ctypes.c_int * 4)()
        for i in range(self.V):
            self.lib.or_coords(i, c)
            out[i] = list(c)
        return out

    def neighbour(self, i, mu, d):
        return self.lib.or_neighbour(i, mu, d)

    # ---- fields --------------------------------------------------------
    @property
    def nsp(self):
        return 8 * self.dr

    def spinor(self):
        return np.zeros((self.V, self.nsp))

    def set_gauge(self, links):
        links = np.ascontiguousarray(links, dtype=np.float64)
        assert links.size == self.V * 4 * 2 * self.nc ** 2
        self.lib.or_set_gauge(_p(links))

    def get_gauge(self):
        out = np.zeros((self.V * 4, 2 * self.nc ** 2))
        self.lib.or_get_gauge(_p(out))
        return out

    def apply_bcs_on_gauge(self):
        self.lib.or_apply_bcs_on_gauge_field()

    def represent(self):
        self.lib.or_represent()

    def get_repr(self):
        out = np.zeros((self.V * 4, 2 * self.dr ** 2))
        self.lib.or_get_repr(_p(out))
        return out

    def set_repr(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.lib.or_set_repr(_p(a))

    def get_clover(self):
        out = np.zeros((self.V * 4, 2 * self.dr ** 2))
        self.lib.or_get_clover(_p(out))
        return out

    def get_ldl(self):
        n = self.dr * (2 * self.dr + 1)
        out = np.zeros((self.V, 4 * n))
        self.lib.or_get_ldl(_p(out))
        flat = out.reshape(self.V, 2, n, 2)
        return flat[..., 0] + 1j * flat[..., 1]  # (V, {upper,lower}, n)

    def compute_ldl(self, sigma):
        self.lib.or_compute_ldl_decomp(sigma)

    def represent_link(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(2 * self.dr ** 2)
        self.lib.or_represent_link(_p(out), _p(u))
        return out

    def generators(self):
        ng = self.nc ** 2 - 1
        f = np.zeros((ng, 2 * self.nc ** 2))
        r = np.zeros((ng, 2 * self.dr ** 2))
        self.lib.or_get_generators(_p(f), _p(r))
        return f, r

    # ---- operators (dirac.h) ------------------------------------------
    def Dphi_oe(self, d, s):
        self.lib.or_Dphi_oe(_p(d), _p(s))

    def Dphi_eo(self, d, s):
        self.lib.or_Dphi_eo(_p(d), _p(s))

    def Dphi_ee(self, m, d, s):
        self.lib.or_Dphi_ee(m, _p(d), _p(s))

    def Dphi_oo(self, m, d, s):
        self.lib.or_Dphi_oo(m, _p(d), _p(s))

    def Dphi_oo_inv(self, m, d, s):
        self.lib.or_Dphi_oo_inv(m, _p(d), _p(s))

    def Dphi(self, m, d, s):
        self.lib.or_Dphi(m, _p(d), _p(s))

    def Hphi(self, m, d, s):
        self.lib.or_Hphi(m, _p(d), _p(s))

    def Hphi_sq(self, m, d, s):
        self.lib.or_Hphi_sq(m, _p(d), _p(s))

    def mvm(self, which, m, d, s):
        self.lib.or_mvm(which, m, _p(d), _p(s))

    def mvm_reset(self):
        self.lib.or_mvm_reset()

    def mvm_get(self):
        return self.lib.or_mvm_get()

    # ---- linalg / solvers ---------------------------------------------
    def sqnorm(self, a):
        return self.lib.or_sqnorm(_p(a))

    def product(self, a, b):
        out = np.zeros(2)
        self.lib.or_product(_p(a), _p(b), _p(out))
        return complex(out[0], out[1])

    def product_re(self, a, b):
        return self.lib.or_product_re(_p(a), _p(b))

    def g5(self, s):
        self.lib.or_g5(_p(s))

    def cg(self, mvm, mass, prec, d, s):
        r = np.zeros(1)
        it = self.lib.or_cg(mvm, mass, prec, _p(d), _p(s), _p(r))
        return it, r[0]

    def bicgstab(self, mvm, mass, prec, d, s):
        r = np.zeros(1)
        it = self.lib.or_bicgstab(mvm, mass, prec, _p(d), _p(s), _p(r))
        return it, r[0]

    def minres(self, mvm, mass, prec, d, s):
        """minres_inverter (library/minres.cpp:12-118); d holds the guess."""
        res = np.zeros(1)
        it = self.lib.or_minres(mvm, mass, prec, _p(d), _p(s), _p(res))
        return it, float(res[0])

    def auto_inverter(self, mvm, mass, prec, d, s):
        return self.lib.or_auto_inverter(mvm, mass, prec, _p(d), _p(s))

    def cg_mshift(self, mvm, mass, prec, shifts, s):
        shifts = np.ascontiguousarray(shifts, dtype=np.float64)
        out = np.zeros((len(shifts),) + s.shape)
        it = self.lib.or_cg_mshift(mvm, mass, prec, len(shifts), _p(shifts), _p(out), _p(s))
        return it, out

    def clover_logdet(self, mass):
        return self.lib.or_clover_logdet(mass)

    def apply_bcs_on_spinor(self, s):
        self.lib.or_apply_bcs_on_spinor(_p(s))

    def set_threads(self, n):
        self.lib.or_set_threads(n)

    # ---- forces (forces.h) ---------------------------------------------
    def get_momenta(self):
        out = np.zeros((4 * self.V, self.nc ** 2 - 1))
        self.lib.or_get_momenta(_p(out))
        return out

    def set_momenta(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        self.lib.or_set_momenta(_p(m))

    def get_clover_force(self):
        out = np.zeros((6 * self.V, 2 * self.dr ** 2))
        self.lib.or_get_clover_force(_p(out))
        return out

    def begin_fermion_force(self):
        self.lib.or_begin_fermion_force()

    def end_fermion_force(self, dt):
        self.lib.or_end_fermion_force(dt)

    def clover_fermion_force(self, residue, lhs, rhs):
        self.lib.or_clover_fermion_force(residue, _p(lhs), _p(rhs))

    def wilson_fermion_force(self, dt, mass, residue, lhs, rhs):
        self.lib.or_wilson_fermion_force(dt, mass, residue, _p(lhs), _p(rhs))

    def wilson_fermion_force_q(self, dt, mass, residue, rhs):
        self.lib.or_wilson_fermion_force_q(dt, mass, residue, _p(rhs))

    def clover_logdet_force(self, mass, coeff):
        self.lib.or_clover_logdet_force(mass, coeff)

    def hmc_action(self, mass, prec, pf):
        return self.lib.or_hmc_action(mass, prec, _p(np.ascontiguousarray(pf)))

    # ---- gauge MD (forces.cpp:262-283, update.cpp:201-247) -------------
    def wilson_gauge_force(self, dt, beta):
        self.lib.or_wilson_gauge_force(dt, beta)

    def plaquette(self):
        return self.lib.or_plaquette()

    def gauge_action(self, beta):
        return self.lib.or_gauge_action(beta)

    def avr_wilson(self, h, w):
        """avr_wilson(h, w) (wilsonloop.cpp:116-133), h, w <= 16"""
        assert 1 <= h <= 16 and 1 <= w <= 16
        return self.lib.or_avr_wilson(h, w)

    def improved_gauge_force(self, dt, beta, c0, c1):
        self.lib.or_improved_gauge_force(dt, beta, c0, c1)

    def improved_gauge_action(self, beta, c0):
        """MonomialGauge::action (monomials/gauge.cpp:20-38) with
        c1 = (1 - c0) / 8"""
        c1 = (1.0 - c0) / 8.0
        v1 = self.avr_wilson(2, 1) + self.avr_wilson(1, 2)
        return -beta * (c0 * self.plaquette() + c1 * v1) * 6.0 * self.gauge_volume()

    def gauge_volume(self):
        """`volume` of the gauge action (bcs.cpp:32, 76)"""
        return self.lib.or_gauge_volume()

    def kinetic(self):
        return self.lib.or_kinetic()

    def update_links(self, eps):
        self.lib.or_update_links(eps)

    def reunitarize(self):
        self.lib.or_reunitarize()

    # ---- parallel-transport convolution layers (learned preconditioner) --
    def ptc_layer(self, W, h, adjoint=False):
        """W: complex (9, cout, cin, 4, 4); h: complex (V, nin, 4, DR) with
        nin = cin (forward) or cout (adjoint)."""
        W = np.ascontiguousarray(W, dtype=np.complex128)
        _, cout, cin = W.shape[:3]
        assert max(cin, cout) <= 16
        nout = cin if adjoint else cout
        h = np.ascontiguousarray(h, dtype=np.complex128)
        out = np.zeros((self.V, nout, 4, self.dr), dtype=np.complex128)
        self.lib.or_ptc_layer(cin, cout, _p(W.view(np.float64)), _p(h.view(np.float64)),
                              _p(out.view(np.float64)), int(adjoint))
        return out

    def ptc_wgrad(self, g, h):
        """sum_{x,a} g_co(x)[s][a] conj((T_p h_ci)(x)[s'][a]) as (9, cout, cin, 4, 4)."""
        g = np.ascontiguousarray(g, dtype=np.complex128)
        h = np.ascontiguousarray(h, dtype=np.complex128)
        cout, cin = g.shape[1], h.shape[1]
        out = np.zeros((9, cout, cin, 4, 4), dtype=np.complex128)
        self.lib.or_ptc_wgrad(cin, cout, _p(g.view(np.float64)), _p(h.view(np.float64)),
                              _p(out.view(np.float64)))
        return out

    def spin_matrix(self, kind, mu, nu=0):
        out = np.zeros(32)
        self.lib.or_spin_matrix(kind, mu, nu, _p(out))
        return (out[0::2] + 1j * out[1::2]).reshape(4, 4)


# ---------------------------------------------------------------------------
# The learned preconditioner's network on the even sublattice, restated in
# FP64 on top of the layer functions above:
#   A b = P_e L_n ... L_1 E b      (E: even -> full lattice, P_e: keep even)
#   A^+ g = P_e L_1^+ ... L_n^+ E g
#   preconditioner  M = A A^+      (Hermitian, positive semi-definite)
# Training loss over (source b, solution x) pairs: |A A^+ b - x|^2 / |x|^2.
# Its weight gradient (d/dRe W + i d/dIm W) is WG(u, e) + WG(v, b) with
# u = A^+ b, e = 2 (A u - x) / |x|^2, v = A^+ e, where WG(w, y) is the
# gradient of Re<y, A w>: the per-layer sum of ptc_wgrad(cotangent of the
# layer output, layer input).
# ---------------------------------------------------------------------------
def spinor_to_act(o, s):
    """Spinor layout (V, 8*DR) doubles -> activation (V, 1, 4, DR) complex."""
    s = np.asarray(s, dtype=np.float64).reshape(o.V, 2, o.dr, 4)
    c = s[:, 0] + 1j * s[:, 1]                      # (V, DR, 4)
    return np.ascontiguousarray(c.transpose(0, 2, 1))[:, None]


def act_to_spinor(o, a):
    c = np.asarray(a)[:, 0].transpose(0, 2, 1)      # (V, DR, 4)
    return np.ascontiguousarray(np.stack([c.real, c.imag], axis=1).reshape(o.V, 8 * o.dr))


class OracleNet:
    def __init__(self, o, weights):
        self.o = o
        self.W = [np.asarray(w, dtype=np.complex128) for w in weights]

    def _even(self, a):
        a = a.copy()
        a[self.o.VH:] = 0
        return a

    def forward(self, h, keep=False):
        """h: (V, 1, 4, DR) activation; returns P_e A E h (and layer inputs)."""
        acts = [self._even(h)]
        for W in self.W:
            acts.append(self.o.ptc_layer(W, acts[-1]))
        out = self._even(acts[-1])
        return (out, acts[:-1]) if keep else out

    def adjoint(self, g, keep=False):
        """P_e A^+ E g; with keep, the cotangents entering each layer output."""
        cot = [self._even(g)]
        for W in reversed(self.W):
            cot.append(self.o.ptc_layer(W, cot[-1], adjoint=True))
        out = self._even(cot[-1])
        # cot[k] enters layer n-1-k; reorder so cots[l] is layer l's output cotangent
        return (out, cot[:-1][::-1]) if keep else out

    def precondition(self, r):
        return self.forward(self.adjoint(r))

    def wgrad(self, w, y):
        """Gradient of Re<y, A w> for every layer."""
        _, acts = self.forward(w, keep=True)
        _, cots = self.adjoint(y, keep=True)
        return [self.o.ptc_wgrad(c, a) for c, a in zip(cots, acts)]

    def loss_and_grad(self, b, x):
        """b, x: activations (V, 1, 4, DR), even sites."""
        u = self.adjoint(b)
        f = self.forward(u)
        nx = np.vdot(x, x).real
        d = f - x
        loss = np.vdot(d, d).real / nx
        e = 2.0 * d / nx
        v = self.adjoint(e)
        g1 = self.wgrad(u, e)
        g2 = self.wgrad(v, b)
        return loss, [a + c for a, c in zip(g1, g2)]


class OracleMRE:
    """mre_init / mre_store / mre_guess (library/mre.cpp:15-117) restated
    over the oracle's operators: Gram-Schmidt of the stored solutions with
    the reference's product sum a conj(b) and update s_j -= r_ij s_i
    (mre.cpp:15-33), Galerkin matrix A[j][i] = <M s_i, s_j>, b_i = <b, s_i>
    (mre.cpp:66-77), dense solve, x0 = sum x_k s_k."""

    def __init__(self, o, max_past, mvm, mass):
        self.o, self.max, self.mvm, self.mass = o, min(max_past, 15), mvm, mass
        self.s = []

    def store(self, x):
        self.s.insert(0, np.array(x, dtype=np.float64, copy=True))
        self.s = self.s[: self.max]

    def _c(self, a):
        return spinor_to_act(self.o, a)[: self.o.VH].reshape(-1)

    def guess(self, b):
        o = self.o
        n = len(self.s)
        out = np.zeros((o.V, 8 * o.dr))
        if n == 0:
            return out
        S = [self._c(v) for v in self.s]
        for i in range(n):
            S[i] = S[i] / np.sqrt(np.vdot(S[i], S[i]).real)
            for j in range(i + 1, n):
                rij = np.sum(S[i] * np.conj(S[j]))
                S[j] = S[j] - rij * S[i]
        A = np.zeros((n, n), dtype=complex)
        rhs = np.zeros(n, dtype=complex)
        bc = self._c(b)
        for i in range(n):
            si = np.zeros((o.V, 1, 4, o.dr), dtype=complex)
            si[: o.VH] = S[i].reshape(o.VH, 1, 4, o.dr)
            d = np.zeros((o.V, 8 * o.dr))
            o.mvm(self.mvm, self.mass, d, act_to_spinor(o, si))
            mi = self._c(d)
            for j in range(n):
                A[j, i] = np.sum(mi * np.conj(S[j]))
            rhs[i] = np.sum(bc * np.conj(S[i]))
        x = np.linalg.solve(A, rhs)
        g = sum(xk * sk for xk, sk in zip(x, S))
        ga = np.zeros((o.V, 1, 4, o.dr), dtype=complex)
        ga[: o.VH] = g.reshape(o.VH, 1, 4, o.dr)
        return act_to_spinor(o, ga)
