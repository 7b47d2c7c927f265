"""Pins the CPU oracle's Wilson-clover operator against textbook checks.

Each check is chosen so that a plausible slip (a wrong projector sign, a
swapped neighbour, a misplaced boundary sign, a transposed clover block)
fails it.  None of them retypes the oracle's own per-direction formulas.
"""
import numpy as np
import pytest

from fields import (GAMMA, GAMMA5, complex_to_spinor, dense_hop, eo_geometry, gaussian_spinor, mat_to_complex,
                    random_gauge, spinor_to_complex, unit_gauge, complex_to_mat)
from oracle import BC_ANTIPERIODIC, BC_OPEN, BC_PERIODIC, BC_SF, Oracle

DIMS = (4, 4, 4, 4)


def make(model="su3_fund", dims=DIMS, bc=BC_ANTIPERIODIC, csw=1.0, seed=1, cf=1.0):
    o = Oracle(model).init(dims, bc=bc, csw=csw, cf=cf)
    rng = np.random.default_rng(seed)
    o.set_gauge(random_gauge(rng, o.V, o.nc))
    if bc in (BC_OPEN, BC_SF):
        o.apply_bcs_on_gauge()
    o.represent()
    return o, rng


def test_gamma_basis_clifford():
    one = np.eye(4)
    for mu in range(4):
        assert np.allclose(GAMMA[mu], GAMMA[mu].conj().T)
        for nu in range(4):
            ac = GAMMA[mu] @ GAMMA[nu] + GAMMA[nu] @ GAMMA[mu]
            assert np.allclose(ac, 2 * one * (mu == nu))
    assert np.allclose(GAMMA[0] @ GAMMA[1] @ GAMMA[2] @ GAMMA[3], GAMMA5)


def test_geometry_matches_eo_formula():
    o = Oracle("su3_fund").init((4, 2, 6, 4))
    coords, fw, bk = eo_geometry(o.dims)
    assert np.array_equal(o.coords(), coords)
    for i in range(0, o.V, 7):
        for mu in range(4):
            assert o.neighbour(i, mu, 1) == fw[i, mu]
            assert o.neighbour(i, mu, -1) == bk[i, mu]
    # even sites first
    assert np.all(coords[: o.VH].sum(axis=1) % 2 == 0)
    assert np.all(coords[o.VH:].sum(axis=1) % 2 == 1)


@pytest.mark.parametrize("bc", [BC_PERIODIC, BC_ANTIPERIODIC, BC_OPEN])
def test_hop_matches_dense_gamma_formula(bc):
    o, rng = make(bc=bc, dims=(4, 4, 2, 6))
    # boundary sign placement: represented U_0 at t = 0 flipped (bcs.cpp:264-272)
    coords, _, _ = eo_geometry(o.dims)
    g = mat_to_complex(o.get_gauge(), 3).reshape(o.V, 4, 3, 3)
    r = mat_to_complex(o.get_repr(), 3).reshape(o.V, 4, 3, 3)
    expect = g.copy()
    if bc == BC_ANTIPERIODIC:
        expect[coords[:, 0] == 0, 0] *= -1
    assert np.array_equal(r, expect)
    psi = gaussian_spinor(rng, o.V, 3)
    if bc == BC_OPEN:
        o.apply_bcs_on_spinor(psi)
    ref = dense_hop(o, o.get_repr(), psi)
    if bc == BC_OPEN:
        t = coords[:, 0]
        ref[(t == 0) | (t == o.dims[0] - 1)] = 0
    d = np.zeros_like(psi)
    o.Dphi_eo(d, psi)
    assert np.abs(d[: o.VH] - ref[: o.VH]).max() < 1e-13
    assert np.all(d[o.VH:] == 0)
    d2 = np.zeros_like(psi)
    o.Dphi_oe(d2, psi)
    assert np.abs(d2[o.VH:] - ref[o.VH:]).max() < 1e-13


def test_free_field_plane_wave():
    """U = 1, periodic: the hop acting on exp(i p.x) u gives
    (-sum cos p_mu + i sum gamma_mu sin p_mu) exp(i p.x) u."""
    dims = (4, 4, 4, 8)
    o = Oracle("su3_fund").init(dims, bc=BC_PERIODIC, csw=0.0)
    o.set_gauge(unit_gauge(o.V, 3))
    o.represent()
    coords, _, _ = eo_geometry(dims)
    rng = np.random.default_rng(7)
    n = np.array([1, 3, 2, 5])
    p = 2 * np.pi * n / np.array(dims)
    u = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
    phase = np.exp(1j * coords @ p)
    psi = complex_to_spinor(phase[:, None, None] * u[None])
    d = np.zeros_like(psi)
    o.Dphi_eo(d, psi)
    o.Dphi_oe(d, psi)
    M = -np.cos(p).sum() * np.eye(4) + 1j * sum(GAMMA[mu] * np.sin(p[mu]) for mu in range(4))
    want = complex_to_spinor(phase[:, None, None] * (M @ u)[None])
    assert np.abs(d - want).max() < 1e-13


def test_clover_term_is_sigma_F():
    """Per site, the clover matrix applied by Dphi_ee (less the mass) equals
    -(c_sw/16) sum_{mu<nu} (1/2)[g_mu, g_nu] (x) (Q_mu_nu - Q_mu_nu^+) with
    Q the sum of the four plaquette leaves."""
    csw = 1.7
    o, rng = make(bc=BC_PERIODIC, csw=csw, dims=(4, 2, 4, 4))
    _, fw, bk = eo_geometry(o.dims)
    U = mat_to_complex(o.get_gauge(), 3).reshape(o.V, 4, 3, 3)
    dag = lambda m: m.conj().T

    def Q(x, mu, nu):
        a, b = fw[x, mu], fw[x, nu]
        q = U[x, mu] @ U[a, nu] @ dag(U[b, mu]) @ dag(U[x, nu])
        a = bk[x, mu]
        b = fw[a, nu]
        q = q + U[x, nu] @ dag(U[b, mu]) @ dag(U[a, nu]) @ U[a, mu]
        a, b = bk[x, mu], bk[x, nu]
        c = bk[a, nu]
        q = q + dag(U[a, mu]) @ dag(U[c, nu]) @ U[c, mu] @ U[b, nu]
        a = bk[x, nu]
        b = fw[a, mu]
        q = q + dag(U[a, nu]) @ U[a, mu] @ U[b, nu] @ dag(U[x, mu])
        return q

    mass = -0.3
    psi = gaussian_spinor(rng, o.V, 3)
    d = np.zeros_like(psi)
    o.Dphi_ee(mass, d, psi)
    o.Dphi_oo(mass, d, psi)
    P = spinor_to_complex(psi, 3)
    D = spinor_to_complex(d, 3)
    for x in range(0, o.V, 5):
        M = sum(np.kron(0.5 * (GAMMA[m] @ GAMMA[n] - GAMMA[n] @ GAMMA[m]), Q(x, m, n) - dag(Q(x, m, n)))
                for m in range(4) for n in range(m + 1, 4))
        M = -(csw / 16.0) * M + (4 + mass) * np.eye(12)
        want = (M @ P[x].reshape(12)).reshape(4, 3)
        assert np.abs(D[x] - want).max() < 1e-13


def _dense_operator(o, mass, apply):
    """Columns of a linear map on the even half, by unit vectors."""
    nh = o.VH * 4 * o.dr
    cols = []
    for k in range(nh):
        e = np.zeros((o.V, 4, o.dr), dtype=complex)
        e.reshape(-1)[k] = 1.0
        cols.append(apply(e))
    return np.stack(cols, axis=1)


def test_schur_operator_brute_force():
    """Dphi = D_ee - D_eo D_oo^-1 D_oe built densely from the dense hop and
    dense clover blocks, inverted by numpy."""
    o, rng = make(dims=(2, 2, 2, 4), csw=1.3)
    mass = -0.4
    V, VH, dr = o.V, o.VH, 3
    n = V * 4 * dr
    # full dense Wilson-clover matrix from the dense hop + the clover of Dphi_ee/oo
    Dh = np.zeros((n, n), dtype=complex)
    Dc = np.zeros((n, n), dtype=complex)
    for k in range(n):
        e = np.zeros((V, 4, dr), dtype=complex)
        e.reshape(-1)[k] = 1.0
        s = complex_to_spinor(e)
        Dh[:, k] = spinor_to_complex(dense_hop(o, o.get_repr(), s), dr).reshape(-1)
        c = np.zeros_like(s)
        o.Dphi_ee(mass, c, s.copy())
        o.Dphi_oo(mass, c, s.copy())
        Dc[:, k] = spinor_to_complex(c, dr).reshape(-1)
    ne = VH * 4 * dr
    E, O = slice(0, ne), slice(ne, n)
    S = Dc[E, E] + Dh[E, E] - (Dh[E, O] @ np.linalg.solve(Dc[O, O] + Dh[O, O], Dh[O, E]))
    assert np.abs(Dh[E, E]).max() == 0 and np.abs(Dc[E, O]).max() == 0
    psi = gaussian_spinor(rng, V, dr)
    psi[VH:] = 0
    want = S @ spinor_to_complex(psi, dr).reshape(-1)[:ne]
    d = np.zeros_like(psi)
    src = psi.copy()
    o.Dphi(mass, d, src)
    got = spinor_to_complex(d, dr).reshape(-1)[:ne]
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-13
    # side effect of dirac.cpp:743: the odd half of the input now holds -D_oe psi_e
    side = spinor_to_complex(src, dr).reshape(-1)[ne:]
    assert np.abs(side + Dh[O, E] @ spinor_to_complex(psi, dr).reshape(-1)[:ne]).max() < 1e-13


@pytest.mark.parametrize("model", ["su3_fund", "su2_adj", "su3_sym"])
def test_gamma5_hermiticity(model):
    """<u, H v> = <H u, v> (checks/dirac.cpp:7-51)."""
    o, rng = make(model=model, csw=1.0)
    mass = -0.5
    u = gaussian_spinor(rng, o.V, o.dr)
    v = gaussian_spinor(rng, o.V, o.dr)
    hu, hv = np.zeros_like(u), np.zeros_like(v)
    o.Hphi(mass, hu, u.copy())
    o.Hphi(mass, hv, v.copy())
    a = o.product(u, hv)
    b = o.product(hu, v)
    assert abs(a - b) / abs(a) < 1e-13


def _gauge_transform(o, rng, links):
    """U'(x,mu) = g(x) U(x,mu) g(x+mu)^+ with random SU(N) g."""
    from fields import reunitarize
    _, fw, _ = eo_geometry(o.dims)
    n = o.nc
    g = reunitarize(rng.normal(size=(o.V, n, n)) + 1j * rng.normal(size=(o.V, n, n)))
    U = mat_to_complex(links, n).reshape(o.V, 4, n, n)
    Up = np.einsum("vab,vmbc->vmac", g, U)
    for mu in range(4):
        Up[:, mu] = Up[:, mu] @ np.conj(np.transpose(g[fw[:, mu]], (0, 2, 1)))
    return g, complex_to_mat(Up.reshape(-1, n, n))


def test_gauge_covariance():
    """D[U^g] g psi = g D[U] psi (checks/dirac.cpp:147-189), fundamental."""
    o, rng = make(csw=1.0)
    mass = -0.5
    links = o.get_gauge()
    psi = gaussian_spinor(rng, o.V, 3)
    psi[o.VH:] = 0
    d0 = np.zeros_like(psi)
    o.Dphi(mass, d0, psi.copy())
    g, lp = _gauge_transform(o, rng, links)
    o.set_gauge(lp)
    o.represent()
    gpsi = complex_to_spinor(np.einsum("vab,vsb->vsa", g, spinor_to_complex(psi, 3)))
    d1 = np.zeros_like(psi)
    o.Dphi(mass, d1, gpsi)
    want = complex_to_spinor(np.einsum("vab,vsb->vsa", g, spinor_to_complex(d0, 3)))
    assert np.abs(d1[: o.VH] - want[: o.VH]).max() < 1e-12


@pytest.mark.parametrize("bc", [BC_PERIODIC, BC_OPEN, BC_SF])
def test_clover_inverse_and_logdet(bc):
    """D_oo^-1 D_oo u = u (checks/dirac.cpp:54-90) and log det D_oo against
    numpy on the dense per-site blocks."""
    o, rng = make(bc=bc, csw=1.2, cf=1.3, dims=(6, 4, 2, 4))
    mass = -0.6
    u = gaussian_spinor(rng, o.V, 3)
    a = np.zeros_like(u)
    b = np.zeros_like(u)
    o.Dphi_oo(mass, a, u)
    o.Dphi_oo_inv(mass, b, a)
    assert np.abs(b[o.VH:] - u[o.VH:]).max() < 1e-13
    # dense log det over odd sites: D_oo is block diagonal per site
    ld = 0.0
    for x in range(o.VH, o.V):
        cols = []
        for k in range(12):
            e = np.zeros((o.V, 4, 3), dtype=complex)
            e[x].reshape(-1)[k] = 1
            s = complex_to_spinor(e)
            c = np.zeros_like(s)
            o.Dphi_oo(mass, c, s)
            cols.append(spinor_to_complex(c, 3)[x].reshape(-1))
        sign, l = np.linalg.slogdet(np.stack(cols, axis=1))
        ld += l
    assert abs(o.clover_logdet(mass) - ld) < 1e-10 * abs(ld)
