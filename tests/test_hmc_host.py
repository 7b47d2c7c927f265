"""Host logic of the HMC monomials (yahmc_amd/hmc.py) against a recording
stand-in for the lattice, with no device: the gauge monomial's action and
force for the Wilson and the improved action (monomials/gauge.cpp:8-49),
and the two-flavour monomial's MRE history (monomials/hmc.cpp:16-17,
36-37, 75-85): reset with the pseudofermion, guess before and store after
every solve, none with mre_past = 0."""
import numpy as np
import pytest

from yahmc_amd.hmc import FermionMonomial, GaugeMonomial


class Rec:
    """Records every call as (name, args)."""

    def __init__(self, log, name):
        self.log, self.name = log, name

    def __getattr__(self, attr):
        def call(*a, **k):
            self.log.append((f"{self.name}.{attr}", a))
            return None
        return call


class FakeLattice:
    V, dr, csw, gauge_volume = 16, 3, 1.0, 16

    def __init__(self):
        self.log = []
        self.loops = {(1, 1): 0.6, (2, 1): 0.4, (1, 2): 0.3}

    def plaquette(self):
        return self.loops[(1, 1)]

    def avr_wilson(self, h, w):
        self.log.append(("avr_wilson", (h, w)))
        return self.loops[(h, w)]

    def wilson_gauge_force(self, *a):
        self.log.append(("wilson_gauge_force", a))

    def improved_gauge_force(self, *a):
        self.log.append(("improved_gauge_force", a))

    def field(self, prec):
        return Rec(self.log, f"field{sum(1 for e in self.log if e[0] == 'field')}")

    def mre(self, npast, mvm, mass):
        self.log.append(("mre", (npast, mvm, mass)))
        return Rec(self.log, "mre")

    def clover_logdet(self, m):
        return 0.25

    def Dphi(self, *a):
        self.log.append(("Dphi", ()))

    def spinor_g5(self, *a):
        pass

    def inv_par(self, *a, **k):
        return ("par", a, k)

    def cg_mixed(self, par, d, s):
        self.log.append(("cg_mixed", ()))
        return 7, 1

    def spinor_product_re(self, a, b):
        return 1.5

    def wilson_fermion_force_q(self, *a):
        self.log.append(("force_q", ()))

    def clover_logdet_force(self, *a):
        pass


def test_gauge_monomial_wilson_and_improved():
    lat = FakeLattice()
    w = GaugeMonomial(5.6)
    assert not w.improved
    assert w.action(lat) == pytest.approx(-5.6 * 0.6 * 6 * 16)
    w.update(lat, 0.1)
    assert lat.log[-1] == ("wilson_gauge_force", (0.1, 5.6))
    c0 = 5.0 / 3.0
    imp = GaugeMonomial(4.2, c0=c0)
    assert imp.improved and imp.c1 == pytest.approx(-1.0 / 12.0)
    want = -4.2 * (c0 * 0.6 + (1 - c0) / 8 * (0.4 + 0.3)) * 6 * 16
    assert imp.action(lat) == pytest.approx(want)
    imp.update(lat, 0.2)
    assert lat.log[-1] == ("improved_gauge_force", (0.2, 4.2, c0, (1 - c0) / 8))
    # |c1| <= 1e-8 is the Wilson action (gauge.cpp:11)
    assert not GaugeMonomial(5.6, c0=1.0 + 1e-9).improved


@pytest.mark.parametrize("npast", [0, 1, 3])
def test_fermion_monomial_mre_history(npast):
    lat = FakeLattice()
    fm = FermionMonomial(-0.5, prec=1e-20, mre_past=npast)
    fm.reset(lat, np.random.default_rng(1))
    names = [e[0] for e in lat.log]
    assert ("mre" in names) == (npast > 0)
    if npast:
        assert lat.log[names.index("mre")][1] == (npast, "Hphi_sq", -0.5)
        # the history restarts with the pseudofermion, after it is built
        assert names.index("mre.reset") > names.index("Dphi")
    lat.log.clear()
    fm.update(lat, 0.1)
    fm.action(lat)  # the first action returns the heatbath value, no solve
    fm.action(lat)
    names = [e[0] for e in lat.log]
    solves = [i for i, n in enumerate(names) if n == "cg_mixed"]
    assert len(solves) == 2 and fm.iterations == [7, 7]
    for i in solves:
        if npast:
            assert names[i - 1] == "mre.guess" and names[i + 1] == "mre.store"
        else:
            assert "mre.guess" not in names and names[i - 1].endswith(".zero")
