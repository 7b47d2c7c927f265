"""Pins the oracle's gauge molecular dynamics (library/forces.cpp:262-283,
staples.cpp:82-136, wilsonloop.cpp:58-133, update.cpp:201-247,
integrator.cpp:42-63, suNa.cpp:113-134):
  - plaquette of the unit field is 1, and agrees with a numpy loop;
  - the Wilson gauge force is -(1/TF) times the gradient of
    S_g = -beta * plaq * 6V (finite differences along random directions);
  - the link update equals scipy's matrix exponential;
  - a 2nd-order minimal-norm trajectory conserves H up to O(dt^2) and is
    reversible.
The Symanzik-improved action (forces.cpp:285-310, staples.cpp:7-80,
monomials/gauge.cpp:20-49, the rectangle weights of bcs.cpp:30-31, 72-120)
is pinned the same way: every h x w Wilson loop average against a numpy
loop over explicit link products, the improved force against the gradient of
the improved action, c0 = 1 against the Wilson force, and an improved o2mn
trajectory for dt^2 scaling and reversibility.
Every check runs for N = 2, 3 and 4 (the models su2_fund, su3_fund and
su4_asym; the gauge MD uses only the fundamental links), and with the open
and SF gauge boundaries of library/bcs.cpp (weighted boundary plaquettes,
fixed links) where the weighted plaquette is checked against a numpy loop
with the weights written out, the force against the gradient of the weighted
action, and the o2mn trajectory for dt^2 scaling and reversibility."""
import numpy as np
import pytest
from scipy.linalg import expm

from fields import complex_to_mat, eo_geometry, mat_to_complex, random_gauge, unit_gauge
from oracle import BC_OPEN, BC_PERIODIC, BC_SF, Oracle
from test_oracle_forces import alg_to_mat

TF = 2.0


MODELS = pytest.mark.parametrize("model,nc", [("su2_fund", 2), ("su3_fund", 3), ("su4_asym", 4)])


BCS = pytest.mark.parametrize("model,nc,bc", [("su2_fund", 2, BC_PERIODIC), ("su3_fund", 3, BC_PERIODIC),
                                               ("su4_asym", 4, BC_PERIODIC), ("su3_fund", 3, BC_OPEN),
                                               ("su2_fund", 2, BC_OPEN), ("su3_fund", 3, BC_SF),
                                               ("su2_fund", 2, BC_SF)])


def make(model, nc, seed=4, dims=(4, 4, 4, 4), bc=BC_PERIODIC):
    if bc != BC_PERIODIC:
        dims = (6,) + tuple(dims[1:])
    o = Oracle(model).init(dims, bc=bc, csw=0.0)
    assert o.nc == nc
    rng = np.random.default_rng(seed)
    o.set_gauge(random_gauge(rng, o.V, nc))
    o.apply_bcs_on_gauge()
    return o, rng


def fixed_links(o, bc):
    """Mask (4V,) of the links the boundary fixes (bcs.cpp:166-262)."""
    T = o.dims[0]
    t = np.repeat(np.array([c[0] for c in eo_geometry(o.dims)[0]]), 4)
    mu = np.tile(np.arange(4), o.V)
    if bc == BC_OPEN:
        return (t == T - 1) & (mu == 0)
    if bc == BC_SF:
        return ((t == 0) & (mu > 0)) | (t == T - 2) | (t == T - 1)
    return np.zeros(4 * o.V, bool)


def weight(o, bc, x_t, mu, nu):
    """plaq_weight (bcs.cpp:28-130): spatial plaquettes on the boundary
    slices count half."""
    T = o.dims[0]
    if mu == 0 or nu == 0:
        return 1.0
    if bc == BC_OPEN and x_t in (0, T - 1):
        return 0.5
    if bc == BC_SF and x_t in (0, T - 2):
        return 0.5
    return 1.0


def volume(o, bc):
    T, v3 = o.dims[0], o.V // o.dims[0]
    return {BC_OPEN: (T - 1) * v3, BC_SF: (T - 2) * v3}.get(bc, o.V)


@BCS
def test_plaquette(model, nc, bc):
    o, rng = make(model, nc, bc=bc)
    U = mat_to_complex(o.get_gauge(), nc).reshape(o.V, 4, nc, nc)
    crd, fw, _ = eo_geometry(o.dims)
    s = 0.0
    for x in range(o.V):
        for mu in range(4):
            for nu in range(mu):
                P = U[x, mu] @ U[fw[x, mu], nu] @ U[fw[x, nu], mu].conj().T @ U[x, nu].conj().T
                s += weight(o, bc, crd[x][0], mu, nu) * np.trace(P).real
    assert abs(o.plaquette() - s / (nc * 6 * volume(o, bc))) < 1e-14
    if bc == BC_PERIODIC:
        o.set_gauge(unit_gauge(o.V, nc))
        assert abs(o.plaquette() - 1.0) < 1e-15
    elif bc == BC_OPEN:
        # unit links with the open boundary: every weighted plaquette not
        # touching the zero links is N; those touching them vanish
        o.set_gauge(unit_gauge(o.V, nc))
        o.apply_bcs_on_gauge()
        T, v3 = o.dims[0], o.V // o.dims[0]
        expect = (3 * (T - 1) * v3 + 3 * (T - 2) * v3 + 2 * 3 * 0.5 * v3) / (6.0 * (T - 1) * v3)
        assert abs(o.plaquette() - expect) < 1e-15


@BCS
def test_gauge_force_is_minus_gradient(model, nc, bc):
    o, rng = make(model, nc, seed=8, bc=bc)
    beta = 5.6
    links = o.get_gauge()
    o.set_momenta(np.zeros((4 * o.V, nc * nc - 1)))
    o.wilson_gauge_force(1.0, beta)
    F = o.get_momenta()
    Q = rng.normal(size=F.shape)
    Q[fixed_links(o, bc)] = 0.0  # the boundary links do not move
    U = mat_to_complex(links, nc)
    XQ = np.array([alg_to_mat(q, nc) for q in Q])

    def S(tau):
        o.set_gauge(complex_to_mat(np.array([expm(tau * x) @ u for x, u in zip(XQ, U)])))
        return o.gauge_action(beta)

    h = 1e-4
    dS = (8 * (S(h) - S(-h)) - (S(2 * h) - S(-2 * h))) / (12 * h)
    want = -TF * np.sum(Q * F)
    scale = TF * np.sum(np.abs(Q * F))
    assert abs(dS - want) < 1e-8 * scale


@MODELS
def test_update_links_is_matrix_exponential(model, nc):
    o, rng = make(model, nc, seed=2)
    P = rng.normal(size=(4 * o.V, nc * nc - 1))
    o.set_momenta(P)
    U0 = mat_to_complex(o.get_gauge(), nc)
    o.update_links(0.13)
    U1 = mat_to_complex(o.get_gauge(), nc)
    for k in range(0, 4 * o.V, 37):
        assert np.abs(U1[k] - expm(0.13 * alg_to_mat(P[k], nc)) @ U0[k]).max() < 1e-13
        assert abs(np.linalg.det(U1[k]) - 1.0) < 1e-13


def o2mn(o, beta, length, steps):
    """o2mn_multilevel with a single level (integrator.cpp:42-63)."""
    a = 0.1931833275037836
    dt = length / steps
    for n in range(steps):
        o.wilson_gauge_force(a * dt if n == 0 else 2 * a * dt, beta)
        o.update_links(dt / 2)
        o.wilson_gauge_force((1 - 2 * a) * dt, beta)
        o.update_links(dt / 2)
    o.wilson_gauge_force(a * dt, beta)


def hamiltonian(o, beta):
    return o.kinetic() + o.gauge_action(beta)


@BCS
def test_o2mn_energy_conservation_scales_as_dt2_and_is_reversible(model, nc, bc):
    o, rng = make(model, nc, seed=3, bc=bc)
    beta = 5.6
    U0 = o.get_gauge()
    P0 = rng.normal(size=(4 * o.V, nc * nc - 1))
    P0[fixed_links(o, bc)] = 0.0  # as the first hamiltonian() leaves them
    dH = []
    for steps in (4, 8):
        o.set_gauge(U0)
        o.set_momenta(P0)
        H0 = hamiltonian(o, beta)
        o2mn(o, beta, 0.5, steps)
        dH.append(abs(hamiltonian(o, beta) - H0))
    ratio = dH[0] / dH[1]
    assert 3.0 < ratio < 5.5, (dH, ratio)
    # reversibility: flip the momenta and integrate back
    o.set_momenta(-o.get_momenta())
    o2mn(o, beta, 0.5, 8)
    o.kinetic()  # the momentum BCs, as hamiltonian() applies them
    assert np.abs(o.get_gauge() - U0).max() < 1e-11
    assert np.abs(o.get_momenta() + P0).max() < 1e-10
    # the fixed links never moved
    fixed = fixed_links(o, bc)
    assert np.array_equal(o.get_gauge()[fixed], U0[fixed])


# ---------------------------------------------------------------------------
# Symanzik-improved gauge action
# ---------------------------------------------------------------------------
C0_LW = 5.0 / 3.0  # tree-level Luscher-Weisz: c1 = (1 - c0) / 8 = -1/12


def rect_weight(o, bc, t, a, b):
    """rect_weight(id, a, b) of bcs.cpp:30-31 (open: the plaquette weights)
    and 74-120 (SF) for the rectangle with short side a, long side b based at
    time slice t."""
    T = o.dims[0]
    if bc == BC_OPEN:
        return weight(o, bc, t, a, b)
    if bc == BC_SF:
        if t in (0, T - 2) and 1 <= a < b:
            return 0.0
        if t in (0, T - 3) and a == 0 and b >= 1:
            return 1.5
    return 1.0


def loop_sum(o, bc, h, w):
    """sum over sites and planes mu > nu of the weighted Re Tr of the h x w
    loop (h links along mu, then w along nu, back along mu, back along nu),
    from explicit matrix products"""
    nc = o.nc
    U = mat_to_complex(o.get_gauge(), nc).reshape(o.V, 4, nc, nc)
    crd, fw, _ = eo_geometry(o.dims)
    s = 0.0
    for x in range(o.V):
        for mu in range(1, 4):
            for nu in range(mu):
                P = np.eye(nc, dtype=complex)
                y = x
                for d, n in ((mu, h), (nu, w)):
                    for _ in range(n):
                        P = P @ U[y, d]
                        y = fw[y, d]
                Q = np.eye(nc, dtype=complex)
                y = x
                for d, n in ((nu, w), (mu, h)):
                    for _ in range(n):
                        Q = Q @ U[y, d]
                        y = fw[y, d]
                tr = np.trace(P @ Q.conj().T).real
                t = crd[x][0]
                if (h, w) == (1, 1):
                    tr *= weight(o, bc, t, mu, nu)
                elif (h, w) == (1, 2):
                    tr *= rect_weight(o, bc, t, mu, nu)
                elif (h, w) == (2, 1):
                    tr *= rect_weight(o, bc, t, nu, mu)
                s += tr
    return s / (nc * 6 * volume(o, bc))


@BCS
def test_avr_wilson_against_explicit_loops(model, nc, bc):
    o, rng = make(model, nc, seed=21, bc=bc)
    for h, w in ((1, 1), (1, 2), (2, 1), (2, 2)):
        assert abs(o.avr_wilson(h, w) - loop_sum(o, bc, h, w)) < 1e-14, (h, w)
    assert abs(o.avr_wilson(1, 1) - o.plaquette()) < 1e-15
    if bc == BC_PERIODIC:
        o.set_gauge(unit_gauge(o.V, nc))
        for h, w in ((1, 2), (2, 1), (3, 2)):
            assert abs(o.avr_wilson(h, w) - 1.0) < 1e-15


@BCS
def test_improved_gauge_force_is_minus_gradient(model, nc, bc):
    o, rng = make(model, nc, seed=23, bc=bc)
    beta, c0 = 4.2, C0_LW
    c1 = (1.0 - c0) / 8.0
    links = o.get_gauge()
    o.set_momenta(np.zeros((4 * o.V, nc * nc - 1)))
    o.improved_gauge_force(1.0, beta, c0, c1)
    F = o.get_momenta()
    Q = rng.normal(size=F.shape)
    Q[fixed_links(o, bc)] = 0.0
    U = mat_to_complex(links, nc)
    XQ = np.array([alg_to_mat(q, nc) for q in Q])

    def S(tau):
        o.set_gauge(complex_to_mat(np.array([expm(tau * x) @ u for x, u in zip(XQ, U)])))
        return o.improved_gauge_action(beta, c0)

    h = 1e-4
    dS = (8 * (S(h) - S(-h)) - (S(2 * h) - S(-2 * h))) / (12 * h)
    want = -TF * np.sum(Q * F)
    scale = TF * np.sum(np.abs(Q * F))
    assert abs(dS - want) < 1e-8 * scale
    # the rectangles alone (c0 = 0) are the gradient of the rectangle sum
    o.set_gauge(links)
    o.set_momenta(np.zeros_like(F))
    o.improved_gauge_force(1.0, beta, 0.0, 1.0)
    R = o.get_momenta()

    def Sr(tau):
        o.set_gauge(complex_to_mat(np.array([expm(tau * x) @ u for x, u in zip(XQ, U)])))
        return -beta * (o.avr_wilson(2, 1) + o.avr_wilson(1, 2)) * 6.0 * volume(o, bc)

    dS = (8 * (Sr(h) - Sr(-h)) - (Sr(2 * h) - Sr(-2 * h))) / (12 * h)
    assert abs(dS + TF * np.sum(Q * R)) < 1e-8 * TF * np.sum(np.abs(Q * R))


@MODELS
def test_improved_force_with_c0_one_is_wilson_force(model, nc):
    o, rng = make(model, nc, seed=25)
    P = rng.normal(size=(4 * o.V, nc * nc - 1))
    o.set_momenta(P)
    o.wilson_gauge_force(0.3, 5.6)
    want = o.get_momenta()
    o.set_momenta(P)
    o.improved_gauge_force(0.3, 5.6, 1.0, 0.0)
    assert np.abs(o.get_momenta() - want).max() < 1e-14


@pytest.mark.parametrize("model,nc,bc", [("su3_fund", 3, BC_PERIODIC), ("su2_fund", 2, BC_SF)])
def test_improved_o2mn_scales_as_dt2_and_is_reversible(model, nc, bc):
    o, rng = make(model, nc, seed=27, bc=bc)
    beta, c0 = 4.2, C0_LW
    c1 = (1.0 - c0) / 8.0
    a = 0.1931833275037836

    def force(dt):
        o.improved_gauge_force(dt, beta, c0, c1)

    def traj(steps, length=0.5):
        dt = length / steps
        for n in range(steps):
            force(a * dt if n == 0 else 2 * a * dt)
            o.update_links(dt / 2)
            force((1 - 2 * a) * dt)
            o.update_links(dt / 2)
        force(a * dt)

    def ham():
        return o.kinetic() + o.improved_gauge_action(beta, c0)

    U0 = o.get_gauge()
    P0 = rng.normal(size=(4 * o.V, nc * nc - 1))
    P0[fixed_links(o, bc)] = 0.0
    dH = []
    for steps in (4, 8):
        o.set_gauge(U0)
        o.set_momenta(P0)
        H0 = ham()
        traj(steps)
        dH.append(abs(ham() - H0))
    ratio = dH[0] / dH[1]
    assert 3.0 < ratio < 5.5, (dH, ratio)
    o.set_momenta(-o.get_momenta())
    traj(8)
    o.kinetic()
    assert np.abs(o.get_gauge() - U0).max() < 1e-11
    assert np.abs(o.get_momenta() + P0).max() < 1e-10
