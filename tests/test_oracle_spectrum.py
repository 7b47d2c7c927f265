"""Pins the spectroscopy restatement: the gamma tables of mesons.cpp are
gamma5 * Gamma in the (pinned) chiral basis; the EO propagator solves the
full-lattice Dirac equation; the point-source pion correlator is the
closed form sum |S|^2."""
import numpy as np

from fields import complex_to_spinor, eo_geometry, gaussian_spinor, random_gauge, spinor_to_complex
from oracle import BC_ANTIPERIODIC, Oracle
from spectrum_ref import TABLES, contract, oracle_propagator, table_matrix


def gammas(o):
    g5 = np.zeros((4, 4), dtype=complex)
    for s in range(4):
        c = np.zeros((o.V, 4, o.dr), dtype=complex)
        c[0, s, 0] = 1.0
        sp = complex_to_spinor(c)
        o.g5(sp)
        g5[:, s] = spinor_to_complex(sp, o.dr)[0, :, 0]
    G = [np.eye(4) - g5 @ o.spin_matrix(0, mu) for mu in range(4)]
    return g5, G


def test_tables_are_gamma5_gamma():
    o = Oracle("su3_fund").init((2, 2, 2, 4), csw=0.0)
    g5, G = gammas(o)
    I = np.eye(4)
    gam = {"id": I, "g1": G[1], "g2": G[2], "g3": G[3], "g5": g5, "g5g1": g5 @ G[1], "g5g2": g5 @ G[2],
           "g5g3": g5 @ G[3], "g0g1": G[0] @ G[1], "g0g2": G[0] @ G[2], "g0g3": G[0] @ G[3], "g0g5": G[0] @ g5,
           "g0g5g1": G[0] @ g5 @ G[1], "g0g5g2": G[0] @ g5 @ G[2], "g0g5g3": G[0] @ g5 @ G[3]}
    for name in TABLES:
        assert np.allclose(table_matrix(name), g5 @ gam[name]), name


def full_dirac(o, mass, x):
    """[[D_ee, D_eo], [D_oe, D_oo]] x from the oracle's parity blocks."""
    out = np.zeros_like(x)
    a, b = np.zeros_like(x), np.zeros_like(x)
    o.Dphi_ee(mass, a, x.copy())
    o.Dphi_eo(b, x.copy())
    out[: o.VH] = a[: o.VH] + b[: o.VH]
    a, b = np.zeros_like(x), np.zeros_like(x)
    o.Dphi_oo(mass, a, x.copy())
    o.Dphi_oe(b, x.copy())
    out[o.VH:] = a[o.VH:] + b[o.VH:]
    return out


def test_propagator_solves_full_dirac_equation_and_pion_closed_form():
    o = Oracle("su3_fund").init((4, 2, 2, 4), bc=BC_ANTIPERIODIC, csw=1.0)
    rng = np.random.default_rng(3)
    o.set_gauge(random_gauge(rng, o.V, 3))
    o.represent()
    mass = -0.2
    b = gaussian_spinor(rng, o.V, 3)
    x = oracle_propagator(o, mass, 1e-26, b)
    r = full_dirac(o, mass, x) - b
    assert np.sum(r * r) / np.sum(b * b) < 1e-22
    # point sources at (0, 0, 0, 0): the g5 channel is sum |S|^2 per slice
    props = []
    for n in range(12):
        s = np.zeros((o.V, 24))
        spin, col = divmod(n, 3)
        s[0, col * 4 + spin] = 1.0
        props.append(oracle_propagator(o, mass, 1e-26, s))
    coords, _, _ = eo_geometry(o.dims)
    t_of = coords[:, 0]
    C = contract(props, 3, o.dims[0], t_of, 0, point=True)
    want = np.zeros(o.dims[0])
    for p in props:
        np.add.at(want, t_of, np.sum(p * p, axis=1))
    assert np.allclose(C[4], want, rtol=1e-12, atol=0)
    assert np.all(C[4] > 0)
