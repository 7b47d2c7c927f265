"""A short HMC trajectory with the Wilson gauge monomial on level 0 and the
two-flavour EO clover fermion monomial on level 1 (o2mn on both levels,
library/integrator.cpp:42-63, monomials/hmc.cpp, list.cpp), driven through
the device kernels by yahmc_amd.hmc, against the same trajectory restated
on the oracle (pseudofermion pf = D eta, force from (H^2)^-1 g5 pf, final
action by BiCGstab as monomials/hmc.cpp:42-65 does).  Same initial links,
momenta and eta on both sides; Metropolis off so that the comparison does
not depend on the accept step.  Checks dH, the plaquette and the final
links and momenta."""
import numpy as np
import pytest

from fields import random_gauge
from oracle import BC_ANTIPERIODIC, MVM_HPHI_SQ, Oracle

pytestmark = pytest.mark.gpu
yh = pytest.importorskip("yahmc_amd")
from yahmc_amd.hmc import HMC, FermionMonomial, GaugeMonomial  # noqa: E402

A = 0.1931833275037836


class FixedRng:
    """Hands the trajectory the momenta and eta chosen by the test."""

    def __init__(self, P, eta):
        self.P, self.eta = P, eta

    def normal(self, loc, scale, size):
        if size == self.P.shape:
            return self.P
        return self.eta[: size[0]]

    def uniform(self):
        return 0.0


def oracle_trajectory(o, links, P, eta, beta, mass, prec, length, nf, ng):
    o.set_gauge(links)
    o.represent()
    o.set_momenta(P)
    la = float(np.sum(eta * eta)) - 2.0 * o.clover_logdet(mass)
    pf = np.zeros_like(eta)
    o.Dphi(mass, pf, eta.copy())
    g5pf = pf.copy()
    o.g5(g5pf)
    h0 = o.kinetic() + o.gauge_action(beta) + la

    def momenta(dt, level):
        if level == 0:
            o.wilson_gauge_force(dt, beta)
            return
        o.begin_fermion_force()
        y = np.zeros_like(eta)
        o.cg(MVM_HPHI_SQ, mass, prec, y, g5pf.copy())
        o.wilson_fermion_force_q(dt, mass, 1.0, y)
        o.clover_logdet_force(mass, 2.0)
        o.end_fermion_force(dt)

    def fields(dt, level):
        if level == 0:
            o.update_links(dt)
        else:
            integrate(dt, level - 1)
            o.represent()

    def integrate(length, level):
        steps = ng if level == 0 else nf
        dt = length / steps
        for n in range(steps):
            momenta(A * dt if n == 0 else 2 * A * dt, level)
            fields(dt / 2, level)
            momenta((1 - 2 * A) * dt, level)
            fields(dt / 2, level)
        momenta(A * dt, level)

    integrate(length, 1)
    o.represent()
    h1 = o.kinetic() + o.gauge_action(beta) + o.hmc_action(mass, prec, pf.copy())
    return h1 - h0, o.plaquette(), o.get_gauge(), o.get_momenta()


@pytest.mark.parametrize("model,nc,rep", [("su3_fund", 3, "fundamental"), ("su3_sym", 3, "symmetric")])
def test_trajectory_matches_oracle(model, nc, rep):
    print(f"\n[case] hmc trajectory {model}", flush=True)
    dims = (4, 4, 4, 8)
    beta, mass, prec, length, nf, ng = 5.6, -0.2, 1e-26, 0.3, 3, 2
    o = Oracle(model).init(dims, bc=BC_ANTIPERIODIC, csw=1.0)
    rng = np.random.default_rng(17)
    links = random_gauge(rng, o.V, nc)
    P = rng.normal(0.0, np.sqrt(0.5), size=(4 * o.V, nc * nc - 1))
    eta = np.zeros((o.V, 8 * o.dr))
    eta[: o.VH] = rng.normal(0.0, np.sqrt(0.5), size=(o.VH, 8 * o.dr))
    dh_o, plaq_o, links_o, mom_o = oracle_trajectory(o, links, P, eta, beta, mass, prec, length, nf, ng)

    lat = yh.Lattice(dims, nc=nc, repr=rep, bc="antiperiodic", csw=1.0)
    lat.set_gauge(links)
    fm = FermionMonomial(mass, prec=prec, level=1)
    hmc = HMC(lat, [GaugeMonomial(beta), fm], [("o2mn", ng), ("o2mn", nf)])
    rec = hmc.trajectory(FixedRng(P, eta[: o.VH]), length=length, metropolis=False)
    print(f"dH device {rec['dH']:.10e} oracle {dh_o:.10e}; plaquette {rec['plaquette']:.12f} {plaq_o:.12f}; "
          f"CG iterations {fm.iterations}")
    # links and momenta before the final reunitarisation agree to solver precision
    assert abs(rec["dH"] - dh_o) < 1e-7 * max(1.0, abs(dh_o))
    assert abs(rec["plaquette"] - plaq_o) < 1e-10
    assert np.abs(lat.get_momenta() - mom_o).max() < 1e-9
    lat_links = lat.get_gauge()
    o.set_gauge(links_o)
    o.reunitarize()
    assert np.abs(lat_links - o.get_gauge()).max() < 1e-10
