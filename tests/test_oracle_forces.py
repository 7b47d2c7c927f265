"""Pins the oracle's fermion forces (library/forces.cpp, clover.cpp:170-279).

The molecular-dynamics equations conserve H = (TF/2) sum P^2 + S, with
U -> exp(tau X_Q) U and P -> P + dt F.  So for any algebra field Q the force
F that the reference's update functions add to the momenta must satisfy
    dS/dtau (U -> exp(tau X_Q) U) = -TF sum_{x,mu,a} Q_a F_a.
The action is the two-flavour HMC monomial S = |Dphi^-1 pf|^2 - 2 log det D_oo
(monomials/hmc.cpp:40-65); the forces are assembled exactly as
MonomialHMC::update (monomials/hmc.cpp:67-93) inside begin/end_fermion_force
(monomials/list.cpp:11-30).  The derivative is taken by central differences,
so a wrong sign, index or factor in any force term fails the test.
"""
import numpy as np
import pytest
from scipy.linalg import expm

from fields import complex_to_mat, gaussian_spinor, mat_to_complex, random_gauge
from oracle import BC_ANTIPERIODIC, BC_PERIODIC, Oracle

TF = 2.0


def alg_to_mat(av, n):
    """suNg(suNa) of library/suNg.cpp:15-58, written from the generator
    definition (off-diagonal i(e_ij + e_ji), e_ij - e_ji; diagonal
    sqrt(2/(i(i+1))) diag(1..1, -i, 0..))."""
    m = np.zeros((n, n), dtype=complex)
    k = 0
    for i in range(n):
        for j in range(n):
            if i > j:
                m[i, j] += 1j * av[k]
                m[j, i] += 1j * av[k]
                k += 1
            elif i < j:
                m[i, j] += av[k]
                m[j, i] -= av[k]
                k += 1
            elif i > 0:
                v = np.sqrt(2.0 / (i * (i + 1)))
                for q in range(i):
                    m[q, q] += 1j * v * av[k]
                m[i, i] -= 1j * i * v * av[k]
                k += 1
    return m


def hmc_force(o, mass, pf, prec=1e-28):
    """Forces of MonomialHMC::update with dt = 1 on zero momenta."""
    o.set_momenta(np.zeros((4 * o.V, o.nc ** 2 - 1)))
    o.begin_fermion_force()
    a = np.zeros_like(pf)
    o.bicgstab(0, mass, prec, a, pf.copy())
    o.g5(a)
    b = np.zeros_like(pf)
    o.bicgstab(0, mass, prec, b, a.copy())
    o.wilson_fermion_force_q(1.0, mass, 1.0, b)
    o.clover_logdet_force(mass, 2.0)
    o.end_fermion_force(1.0)
    return o.get_momenta()


@pytest.mark.parametrize("model,csw,bc", [("su3_fund", 1.0, BC_ANTIPERIODIC), ("su3_fund", 0.0, BC_PERIODIC),
                                          ("su2_adj", 1.3, BC_ANTIPERIODIC), ("su3_sym", 1.0, BC_PERIODIC)])
def test_force_is_minus_gradient_of_action(model, csw, bc):
    o = Oracle(model).init((4, 4, 2, 4), bc=bc, csw=csw)
    rng = np.random.default_rng(21)
    n = o.nc
    links = random_gauge(rng, o.V, n)
    o.set_gauge(links)
    o.represent()
    mass = -0.2
    pf = gaussian_spinor(rng, o.V, o.dr)
    pf[o.VH:] = 0
    F = hmc_force(o, mass, pf)
    Q = rng.normal(size=F.shape)
    U = mat_to_complex(links, n)
    XQ = np.array([alg_to_mat(q, n) for q in Q])

    def S(tau):
        Ut = np.array([expm(tau * x) @ u for x, u in zip(XQ, U)])
        o.set_gauge(complex_to_mat(Ut))
        o.represent()
        return o.hmc_action(mass, 1e-28, pf)

    h = 1e-4
    # fourth-order central difference; the scale of the individual terms
    # (not their sum, which cancels) sets the tolerance
    dS = (8 * (S(h) - S(-h)) - (S(2 * h) - S(-2 * h))) / (12 * h)
    want = -TF * np.sum(Q * F)
    scale = TF * np.sum(np.abs(Q * F))
    assert abs(dS - want) < 1e-8 * scale, (dS, want, scale)
    assert abs(want) > 1e-3 * scale


def test_spin_matrices_match_reference_forms():
    """g5_one_minus_gamma(1) and g5_sigma(0,3) transcribed component-wise from
    library/spinor.cpp:440-500 and 340-352 against the dense forms."""
    o = Oracle("su3_fund").init((2, 2, 2, 2))
    rng = np.random.default_rng(2)
    psi = rng.normal(size=4) + 1j * rng.normal(size=4)
    re, im = psi.real, psi.imag
    # g5_one_minus_gamma(1): spinor.cpp:456-467
    want = np.array([re[0] - im[3] + 1j * (im[0] + re[3]),
                     re[1] - im[2] + 1j * (im[1] + re[2]),
                     -re[2] - im[1] + 1j * (-im[2] + re[1]),
                     -re[3] - im[0] + 1j * (-im[3] + re[0])])
    assert np.allclose(o.spin_matrix(0, 1) @ psi, want)
    # g5_sigma(0,3): spinor.cpp:340-352
    want = np.array([im[0] - 1j * re[0], -im[1] + 1j * re[1], im[2] - 1j * re[2], -im[3] + 1j * re[3]])
    assert np.allclose(o.spin_matrix(1, 0, 3) @ psi, want)


def test_rhmc_force_is_minus_gradient_of_rational_action():
    """MonomialRHMC (monomials/rhmc.cpp:65-106): with the rational function
    r(x) = a0 + sum_k res_k / (x + pole_k) the action
    S = <pf, r(H^2) pf> - log det D_oo is a rational function of the links,
    and the force sum_k force_q(res_k, (H^2 + pole_k)^-1 pf) plus
    clover_logdet_force(mass, 1) must be exactly its gradient."""
    from oracle import MVM_HPHI_SQ
    yh = pytest.importorskip("yahmc_amd")
    o = Oracle("su3_fund").init((4, 4, 2, 4), bc=BC_ANTIPERIODIC, csw=1.0)
    rng = np.random.default_rng(23)
    links = random_gauge(rng, o.V, 3)
    o.set_gauge(links)
    o.represent()
    mass, prec = -0.2, 1e-28
    r = yh.rational_find(-1, 2, 1e-8, 1e-3, 64.0)
    poles = np.array(r.pole[: r.order])
    pf = gaussian_spinor(rng, o.V, o.dr)
    pf[o.VH:] = 0

    o.set_momenta(np.zeros((4 * o.V, 8)))
    o.begin_fermion_force()
    _, xs = o.cg_mshift(MVM_HPHI_SQ, mass, prec, poles, pf.copy())
    for k in range(r.order):
        o.wilson_fermion_force_q(1.0, mass, r.residue[k], xs[k].copy())
    o.clover_logdet_force(mass, 1.0)
    o.end_fermion_force(1.0)
    F = o.get_momenta()
    Q = rng.normal(size=F.shape)
    U = mat_to_complex(links, 3)
    XQ = np.array([alg_to_mat(q, 3) for q in Q])

    def S(tau):
        o.set_gauge(complex_to_mat(np.array([expm(tau * x) @ u for x, u in zip(XQ, U)])))
        o.represent()
        _, ys = o.cg_mshift(MVM_HPHI_SQ, mass, prec, poles, pf.copy())
        rp = r.a0 * pf + sum(r.residue[k] * ys[k] for k in range(r.order))
        return o.product_re(pf, rp) - o.clover_logdet(mass)

    h = 1e-4
    dS = (8 * (S(h) - S(-h)) - (S(2 * h) - S(-2 * h))) / (12 * h)
    want = -TF * np.sum(Q * F)
    scale = TF * np.sum(np.abs(Q * F))
    assert abs(dS - want) < 1e-8 * scale, (dS, want, scale)
    assert abs(want) > 1e-3 * scale
