"""Pins the oracle's gauge molecular dynamics (library/forces.cpp:262-283,
staples.cpp:82-136, wilsonloop.cpp:58-133, update.cpp:201-247,
integrator.cpp:42-63, suNa.cpp:113-134):
  - plaquette of the unit field is 1, and agrees with a numpy loop;
  - the Wilson gauge force is -(1/TF) times the gradient of
    S_g = -beta * plaq * 6V (finite differences along random directions);
  - the link update equals scipy's matrix exponential;
  - a 2nd-order minimal-norm trajectory conserves H up to O(dt^2) and is
    reversible.
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
