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
from oracle import BC_ANTIPERIODIC, MVM_HPHI, MVM_HPHI_SQ, Oracle

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


@pytest.mark.parametrize("model,nc,rep", [("su3_fund", 3, "fundamental"), ("su3_sym", 3, "symmetric"),
                                          ("su3_adj", 3, "adjoint")])
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


class OracleHasenbusch:
    """MonomialHB (monomials/hasenbusch.cpp, CLOVER_TERM_EO branch) on the
    oracle, inverses by auto_inverter as the reference does."""

    def __init__(self, o, mass, dm, prec):
        self.o, self.om, self.sm, self.prec = o, mass, mass + dm, prec

    def reset(self, eta):
        o = self.o
        self.la = float(np.sum(eta * eta))
        a = np.zeros_like(eta)
        o.Hphi(self.om, a, eta.copy())
        self.pf = np.zeros_like(eta)
        o.auto_inverter(MVM_HPHI, self.sm, self.prec, self.pf, a)
        return self.la

    def action(self):
        o = self.o
        a, b = np.zeros_like(self.pf), np.zeros_like(self.pf)
        o.Hphi(self.sm, a, self.pf.copy())
        o.auto_inverter(MVM_HPHI, self.om, self.prec, b, a)
        return o.sqnorm(b)

    def update(self, dt):
        o = self.o
        a, y = np.zeros_like(self.pf), np.zeros_like(self.pf)
        o.Hphi(self.sm, a, self.pf.copy())
        o.auto_inverter(MVM_HPHI_SQ, self.om, self.prec, y, a)
        o.wilson_fermion_force_q(dt, self.om, 1.0, y.copy())
        o.wilson_fermion_force(dt, self.sm, -1.0, self.pf, y)


def test_hasenbusch_trajectory_matches_oracle():
    """Gauge + two-flavour monomial at m + dm + Hasenbusch ratio (m, dm) on
    level 1: dH and plaquette against the oracle restatement."""
    print("\n[case] hmc + hasenbusch trajectory", flush=True)
    from oracle import MVM_HPHI  # noqa: F401
    from yahmc_amd.hmc import HasenbuschMonomial
    dims = (4, 4, 4, 8)
    beta, mass, dm, prec, length, nf, ng = 5.6, -0.25, 0.2, 1e-26, 0.3, 2, 1
    o = Oracle("su3_fund").init(dims, bc=BC_ANTIPERIODIC, csw=1.0)
    rng = np.random.default_rng(29)
    links = random_gauge(rng, o.V, 3)
    P = rng.normal(0.0, np.sqrt(0.5), size=(4 * o.V, 8))
    eta1 = np.zeros((o.V, 24))
    eta1[: o.VH] = rng.normal(0.0, np.sqrt(0.5), size=(o.VH, 24))
    eta2 = np.zeros((o.V, 24))
    eta2[: o.VH] = rng.normal(0.0, np.sqrt(0.5), size=(o.VH, 24))

    # oracle: HMC monomial at m + dm and Hasenbusch (m, dm)
    o.set_gauge(links)
    o.represent()
    o.set_momenta(P)
    hm = mass + dm
    la = float(np.sum(eta1 * eta1)) - 2.0 * o.clover_logdet(hm)
    pf = np.zeros_like(eta1)
    o.Dphi(hm, pf, eta1.copy())
    g5pf = pf.copy()
    o.g5(g5pf)
    hb = OracleHasenbusch(o, mass, dm, prec)
    h0 = o.kinetic() + o.gauge_action(beta) + la + hb.reset(eta2)

    def momenta(dt, level):
        if level == 0:
            o.wilson_gauge_force(dt, beta)
            return
        o.begin_fermion_force()
        y = np.zeros_like(eta1)
        o.cg(MVM_HPHI_SQ, hm, prec, y, g5pf.copy())
        o.wilson_fermion_force_q(dt, hm, 1.0, y)
        o.clover_logdet_force(hm, 2.0)
        hb.update(dt)
        o.end_fermion_force(dt)

    def integrate(length, level):
        steps = ng if level == 0 else nf
        dt = length / steps
        for n in range(steps):
            momenta(A * dt if n == 0 else 2 * A * dt, level)
            fields(dt / 2, level)
            momenta((1 - 2 * A) * dt, level)
            fields(dt / 2, level)
        momenta(A * dt, level)

    def fields(dt, level):
        if level == 0:
            o.update_links(dt)
        else:
            integrate(dt, level - 1)
            o.represent()

    integrate(length, 1)
    o.represent()
    h1 = o.kinetic() + o.gauge_action(beta) + o.hmc_action(hm, prec, pf.copy()) + hb.action()
    dh_o = h1 - h0

    class Rng:
        def __init__(self):
            self.k = 0

        def normal(self, loc, scale, size):
            if size == P.shape:
                return P
            self.k += 1
            return (eta1 if self.k == 1 else eta2)[: size[0]]

        def uniform(self):
            return 0.0

    lat = yh.Lattice(dims, nc=3, repr="fundamental", bc="antiperiodic", csw=1.0)
    lat.set_gauge(links)
    hmc = HMC(lat, [GaugeMonomial(beta), FermionMonomial(hm, prec=prec), HasenbuschMonomial(mass, dm, prec=prec)],
              [("o2mn", ng), ("o2mn", nf)])
    rec = hmc.trajectory(Rng(), length=length, metropolis=False)
    print(f"dH device {rec['dH']:.10e} oracle {dh_o:.10e}")
    assert abs(rec["dH"] - dh_o) < 1e-7 * max(1.0, abs(dh_o))
    assert abs(rec["plaquette"] - o.plaquette()) < 1e-10


def test_evlimits_bracket_dense_spectrum():
    """evlimits (library/evlimit.cpp): power and inverse power iteration on
    H^2 bracket the dense spectrum of a tiny lattice from inside, within
    10% at the reference's stopping rule (relative change below 1e-3)."""
    print("\n[case] evlimits", flush=True)
    from fields import complex_to_spinor, spinor_to_complex
    from yahmc_amd.hmc import evlimits
    dims = (2, 2, 2, 4)
    o = Oracle("su3_fund").init(dims, bc=BC_ANTIPERIODIC, csw=1.0)
    rng = np.random.default_rng(41)
    links = random_gauge(rng, o.V, 3)
    o.set_gauge(links)
    o.represent()
    mass = -0.2
    cols = []
    for k in range(o.VH * 12):
        e = np.zeros((o.V, 4, 3), dtype=complex)
        e[: o.VH].reshape(-1)[k] = 1.0
        h = np.zeros((o.V, 24))
        o.Hphi_sq(mass, h, complex_to_spinor(e))
        cols.append(spinor_to_complex(h, 3)[: o.VH].reshape(-1))
    w = np.linalg.eigvalsh(np.stack(cols, axis=1))
    lat = yh.Lattice(dims, nc=3, repr="fundamental", bc="antiperiodic", csw=1.0)
    lat.set_gauge(links)
    lo, hi = evlimits(lat, mass, np.random.default_rng(3))
    print(f"evlimits [{lo:.6e}, {hi:.6e}], dense [{w.min():.6e}, {w.max():.6e}]")
    # |H^2 b| <= lambda_max |b| and 1/|H^-2 b| >= lambda_min |b|: the
    # estimates approach the ends from inside
    assert 0.9 * w.max() < hi <= w.max() * (1 + 1e-9)
    assert w.min() * (1 - 1e-4) <= lo < 1.1 * w.min()


class OracleRHMC:
    """MonomialRHMC (monomials/rhmc.cpp:30-106) on the oracle: rational_calc
    (rational.cpp:134-166) and the force from the oracle's multishift CG,
    with the rational approximations the device monomial chose."""

    def __init__(self, o, mass, rpf, rfc, prec):
        self.o, self.m, self.rpf, self.rfc, self.prec = o, mass, rpf, rfc, prec

    def _calc(self, r, s):
        _, xs = self.o.cg_mshift(MVM_HPHI_SQ, self.m, self.prec, np.array(r.pole[: r.order]), s.copy())
        return r.a0 * s + sum(r.residue[k] * xs[k] for k in range(r.order))

    def reset(self, eta):
        self.pf = self._calc(self.rpf, eta)
        return float(np.sum(eta * eta)) - self.o.clover_logdet(self.m)

    def action(self):
        return self.o.product_re(self.pf, self._calc(self.rfc, self.pf)) - self.o.clover_logdet(self.m)

    def update(self, dt):
        n = self.rfc.order
        _, xs = self.o.cg_mshift(MVM_HPHI_SQ, self.m, self.prec, np.array(self.rfc.pole[:n]), self.pf.copy())
        for k in range(n):
            self.o.wilson_fermion_force_q(dt, self.m, self.rfc.residue[k], xs[k].copy())
        self.o.clover_logdet_force(self.m, 1.0)


def test_rhmc_trajectory_matches_oracle():
    """Gauge + one-flavour RHMC monomial (x^(1/4) heatbath, x^(-1/2) action
    and force, evlimits range): dH, plaquette and momenta against the
    oracle restatement; the heatbath action |eta|^2 agrees with
    <pf, r_fc(H^2) pf> to the rational precision."""
    print("\n[case] rhmc trajectory", flush=True)
    from yahmc_amd.hmc import RHMCMonomial
    dims = (4, 4, 4, 8)
    beta, mass, prec, length, nf, ng = 5.6, -0.2, 1e-24, 0.3, 2, 1
    o = Oracle("su3_fund").init(dims, bc=BC_ANTIPERIODIC, csw=1.0)
    rng = np.random.default_rng(31)
    links = random_gauge(rng, o.V, 3)
    P = rng.normal(0.0, np.sqrt(0.5), size=(4 * o.V, 8))
    eta = np.zeros((o.V, 24))
    eta[: o.VH] = rng.normal(0.0, np.sqrt(0.5), size=(o.VH, 24))

    lat = yh.Lattice(dims, nc=3, repr="fundamental", bc="antiperiodic", csw=1.0)
    lat.set_gauge(links)
    rm = RHMCMonomial(mass, rap_prec=1e-10, inv_prec=prec)
    hmc = HMC(lat, [GaugeMonomial(beta), rm], [("o2mn", ng), ("o2mn", nf)])
    rec = hmc.trajectory(FixedRng(P, eta[: o.VH]), length=length, metropolis=False)
    print(f"evlimits x (0.95, 1.05) {rm.limits}; orders pf {rm.rpf.order} fc {rm.rfc.order}; "
          f"multishift iterations {rm.iterations}")

    o.set_gauge(links)
    o.represent()
    o.set_momenta(P)
    orh = OracleRHMC(o, mass, rm.rpf, rm.rfc, prec)
    la = orh.reset(eta)
    # heatbath: <pf, r_fc pf> = <eta, r_pf r_fc r_pf eta> = |eta|^2 up to the rational errors
    s0 = orh.action()
    assert abs(s0 - la) < 3e-10 * float(np.sum(eta * eta))
    h0 = o.kinetic() + o.gauge_action(beta) + la

    def momenta(dt, level):
        if level == 0:
            o.wilson_gauge_force(dt, beta)
            return
        o.begin_fermion_force()
        orh.update(dt)
        o.end_fermion_force(dt)

    def integrate(length, level):
        steps = ng if level == 0 else nf
        dt = length / steps
        for n in range(steps):
            momenta(A * dt if n == 0 else 2 * A * dt, level)
            fields(dt / 2, level)
            momenta((1 - 2 * A) * dt, level)
            fields(dt / 2, level)
        momenta(A * dt, level)

    def fields(dt, level):
        if level == 0:
            o.update_links(dt)
        else:
            integrate(dt, level - 1)
            o.represent()

    integrate(length, 1)
    o.represent()
    mom_o = o.get_momenta()
    dh_o = o.kinetic() + o.gauge_action(beta) + orh.action() - h0
    print(f"dH device {rec['dH']:.10e} oracle {dh_o:.10e}")
    assert abs(rec["dH"] - dh_o) < 1e-7 * max(1.0, abs(dh_o))
    assert abs(rec["plaquette"] - o.plaquette()) < 1e-10
    assert np.abs(lat.get_momenta() - mom_o).max() < 1e-9
