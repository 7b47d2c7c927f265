"""Gauge molecular dynamics on the device against the oracle
(library/forces.cpp:262-310, update.cpp:201-320, wilsonloop.cpp:58-133,
integrator.cpp:42-63): plaquette, Wilson loops, Wilson and improved gauge
forces, link update, reunitarisation, momentum energy, whole o2mn
trajectories, and the represented operator after the links moved."""
import numpy as np
import pytest

from fields import gaussian_spinor, random_gauge
from oracle import BC_ANTIPERIODIC, BC_OPEN, BC_PERIODIC, BC_SF, Oracle

pytestmark = pytest.mark.gpu
yh = pytest.importorskip("yahmc_amd")

MODELS = [("su3_fund", 3, "fundamental"), ("su2_fund", 2, "fundamental"), ("su4_asym", 4, "antisymmetric")]


def setup(model, nc, rep, dims=(4, 4, 4, 8), seed=5, bc="antiperiodic", csw=1.0):
    obc = BC_ANTIPERIODIC if bc == "antiperiodic" else BC_PERIODIC
    o = Oracle(model).init(dims, bc=obc, csw=csw)
    rng = np.random.default_rng(seed)
    links = random_gauge(rng, o.V, nc)
    o.set_gauge(links)
    lat = yh.Lattice(dims, nc=nc, repr=rep, bc=bc, csw=csw)
    lat.set_gauge(links)
    return o, lat, rng


@pytest.mark.parametrize("model,nc,rep", MODELS)
def test_plaquette_and_force(model, nc, rep):
    print(f"\n[case] plaquette/gauge force {model}", flush=True)
    o, lat, rng = setup(model, nc, rep)
    assert abs(lat.plaquette() - o.plaquette()) < 1e-14
    P = rng.normal(size=(4 * o.V, nc * nc - 1))
    o.set_momenta(P)
    lat.set_momenta(P)
    o.wilson_gauge_force(0.37, 5.6)
    lat.wilson_gauge_force(0.37, 5.6)
    got, want = lat.get_momenta(), o.get_momenta()
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    assert abs(lat.momenta_energy() - o.kinetic()) < 1e-12 * o.kinetic()


@pytest.mark.parametrize("model,nc,rep", MODELS)
def test_update_links_and_reunitarize(model, nc, rep):
    print(f"\n[case] update_links {model}", flush=True)
    o, lat, rng = setup(model, nc, rep, seed=11)
    P = rng.normal(size=(4 * o.V, nc * nc - 1))
    o.set_momenta(P)
    lat.set_momenta(P)
    o.update_links(0.21)
    lat.update_links(0.21)
    assert np.abs(lat.get_gauge() - o.get_gauge()).max() < 1e-13
    # perturb away from the group, then reunitarise on both sides
    bent = o.get_gauge() * (1.0 + 1e-3 * rng.normal(size=(4 * o.V, 2 * nc * nc)))
    o.set_gauge(bent)
    lat.set_gauge(bent)
    o.reunitarize()
    lat.reunitarize()
    assert np.abs(lat.get_gauge() - o.get_gauge()).max() < 1e-13


def test_o2mn_trajectory_matches_oracle_and_operator_follows():
    """A five-step o2mn gauge trajectory: dH and the final links agree with
    the oracle; after represent_gauge_field the device Dirac operator is the
    oracle's on the moved links."""
    print("\n[case] o2mn trajectory su3_fund", flush=True)
    o, lat, rng = setup("su3_fund", 3, "fundamental", seed=7)
    beta, mass = 5.6, -0.3
    P = rng.normal(size=(4 * o.V, 8))
    o.set_momenta(P)
    lat.set_momenta(P)

    def ham(x):
        return x.momenta_energy() - beta * x.plaquette() * 6 * o.V if x is lat else \
            x.kinetic() + x.gauge_action(beta)

    h0d, h0o = ham(lat), ham(o)
    assert abs(h0d - h0o) < 1e-10 * abs(h0o)
    a, steps, length = 0.1931833275037836, 5, 0.5
    dt = length / steps
    for x in (o, lat):
        for n in range(steps):
            x.wilson_gauge_force(a * dt if n == 0 else 2 * a * dt, beta)
            x.update_links(dt / 2)
            x.wilson_gauge_force((1 - 2 * a) * dt, beta)
            x.update_links(dt / 2)
        x.wilson_gauge_force(a * dt, beta)
    dHd, dHo = ham(lat) - h0d, ham(o) - h0o
    print(f"dH device {dHd:.6e} oracle {dHo:.6e}")
    assert abs(dHd - dHo) < 1e-9 * abs(h0o)
    assert np.abs(lat.get_gauge() - o.get_gauge()).max() < 1e-12
    assert np.abs(lat.get_momenta() - o.get_momenta()).max() < 1e-11
    # the operator follows the moved links
    lat.represent_gauge_field()
    o.represent()
    s = gaussian_spinor(rng, o.V, o.dr)
    want = np.zeros_like(s)
    o.Dphi(mass, want, s.copy())
    fs, fd = lat.field("f64").upload(s), lat.field("f64")
    lat.Dphi(mass, fd, fs)
    got = fd.download()
    assert np.abs(got[: o.VH] - want[: o.VH]).max() < 1e-12 * np.abs(want).max()


def setup_bc(model, nc, rep, bc, dims=(6, 4, 4, 4), seed=13):
    """Random links with the gauge-field BCs applied by the oracle
    (bcs.cpp:201-262) and handed to the device as they are."""
    obc = {"open": BC_OPEN, "sf": BC_SF}[bc]
    o = Oracle(model).init(dims, bc=obc, csw=1.0, cf=1.1)
    rng = np.random.default_rng(seed)
    o.set_gauge(random_gauge(rng, o.V, nc))
    o.apply_bcs_on_gauge()
    links = o.get_gauge()
    lat = yh.Lattice(dims, nc=nc, repr=rep, bc=bc, csw=1.0, cf=1.1)
    lat.set_gauge(links)
    return o, lat, rng


BC_CASES = [("su3_fund", 3, "fundamental", "open"), ("su3_fund", 3, "fundamental", "sf"),
            ("su2_fund", 2, "fundamental", "sf"), ("su4_asym", 4, "antisymmetric", "open")]


@pytest.mark.parametrize("model,nc,rep,bc", BC_CASES)
def test_boundary_gauge_md_matches_oracle(model, nc, rep, bc):
    """Open / SF gauge boundaries (bcs.cpp:11-262): weighted plaquette and
    staples, momentum BCs before the energy and the link update, fixed links
    unchanged by the update; the device applies the gauge BCs to a cold
    field as the oracle does."""
    print(f"\n[case] boundary gauge MD {model} {bc}", flush=True)
    o, lat, rng = setup_bc(model, nc, rep, bc)
    assert abs(lat.plaquette() - o.plaquette()) < 1e-14
    P = rng.normal(size=(4 * o.V, nc * nc - 1))
    o.set_momenta(P)
    lat.set_momenta(P)
    o.wilson_gauge_force(0.37, 5.6)
    lat.wilson_gauge_force(0.37, 5.6)
    want = o.get_momenta()
    assert np.abs(lat.get_momenta() - want).max() < 1e-13 * np.abs(want).max()
    e_dev, e_or = lat.momenta_energy(), o.kinetic()   # both zero the fixed links' momenta
    assert abs(e_dev - e_or) < 1e-12 * e_or
    assert np.array_equal(lat.get_momenta() == 0, o.get_momenta() == 0)
    U0 = o.get_gauge()
    o.update_links(0.21)
    lat.update_links(0.21)
    U1 = o.get_gauge()
    assert np.abs(lat.get_gauge() - U1).max() < 1e-13
    # the fixed links did not move
    zero = np.all(o.get_momenta() == 0, axis=1)
    assert np.count_nonzero(zero) > 0
    assert np.abs(U1[zero] - U0[zero]).max() < 1e-15
    # cold start: the device sets the boundary links the oracle's way
    cold = np.zeros((4 * o.V, 2 * nc * nc))
    for i in range(nc):
        cold[:, i * nc + i] = 1.0
    o.set_gauge(cold)
    o.apply_bcs_on_gauge()
    lat.set_gauge(cold)
    lat.apply_bcs_on_gauge_field()
    assert np.abs(lat.get_gauge() - o.get_gauge()).max() < 1e-15
    assert abs(lat.plaquette() - o.plaquette()) < 1e-14


def test_open_boundary_o2mn_trajectory_matches_oracle():
    """A five-step o2mn gauge trajectory with open boundaries: dH, links and
    momenta agree with the oracle."""
    print("\n[case] o2mn trajectory su3_fund open", flush=True)
    o, lat, rng = setup_bc("su3_fund", 3, "fundamental", "open", seed=17)
    beta = 5.6
    P = rng.normal(size=(4 * o.V, 8))
    o.set_momenta(P)
    lat.set_momenta(P)

    def ham(x):
        return x.momenta_energy() - beta * x.plaquette() * 6 * lat.gauge_volume if x is lat else \
            x.kinetic() + x.gauge_action(beta)

    h0d, h0o = ham(lat), ham(o)
    assert abs(h0d - h0o) < 1e-10 * abs(h0o)
    a, steps, length = 0.1931833275037836, 5, 0.5
    dt = length / steps
    for x in (o, lat):
        for n in range(steps):
            x.wilson_gauge_force(a * dt if n == 0 else 2 * a * dt, beta)
            x.update_links(dt / 2)
            x.wilson_gauge_force((1 - 2 * a) * dt, beta)
            x.update_links(dt / 2)
        x.wilson_gauge_force(a * dt, beta)
    dHd, dHo = ham(lat) - h0d, ham(o) - h0o
    print(f"dH device {dHd:.6e} oracle {dHo:.6e}")
    assert abs(dHd - dHo) < 1e-9 * abs(h0o)
    assert np.abs(lat.get_gauge() - o.get_gauge()).max() < 1e-12
    assert np.abs(lat.get_momenta() - o.get_momenta()).max() < 1e-11


IMPROVED_CASES = [("su3_fund", 3, "fundamental", "antiperiodic"), ("su2_fund", 2, "fundamental", "periodic"),
                  ("su4_asym", 4, "antisymmetric", "antiperiodic"), ("su3_fund", 3, "fundamental", "open"),
                  ("su3_fund", 3, "fundamental", "sf"), ("su2_fund", 2, "fundamental", "sf")]
C0_LW = 5.0 / 3.0  # tree-level Luscher-Weisz, c1 = -1/12


@pytest.mark.parametrize("model,nc,rep,bc", IMPROVED_CASES)
def test_improved_gauge_force_and_loops_match_oracle(model, nc, rep, bc):
    """avr_wilson(h, w) (wilsonloop.cpp:58-133) and improved_gauge_force
    (forces.cpp:285-310, staples.cpp:7-80) with the rectangle weights of the
    open and SF boundaries (bcs.cpp:30-31, 72-120); the gauge monomial with
    c0 != 1 (monomials/gauge.cpp) uses both."""
    print(f"\n[case] improved gauge force {model} {bc}", flush=True)
    from yahmc_amd.hmc import GaugeMonomial
    if bc in ("open", "sf"):
        o, lat, rng = setup_bc(model, nc, rep, bc, seed=29)
    else:
        o, lat, rng = setup(model, nc, rep, seed=29, bc=bc)
    for h, w in ((1, 1), (1, 2), (2, 1), (2, 2), (3, 1)):
        assert abs(lat.avr_wilson(h, w) - o.avr_wilson(h, w)) < 1e-14, (h, w)
    assert abs(lat.avr_wilson(1, 1) - lat.plaquette()) < 1e-15
    beta, c0 = 4.2, C0_LW
    c1 = (1.0 - c0) / 8.0
    P = rng.normal(size=(4 * o.V, nc * nc - 1))
    o.set_momenta(P)
    lat.set_momenta(P)
    o.improved_gauge_force(0.37, beta, c0, c1)
    lat.improved_gauge_force(0.37, beta, c0, c1)
    want = o.get_momenta()
    assert np.abs(lat.get_momenta() - want).max() < 1e-13 * np.abs(want).max()
    # the monomial: action and update
    mono = GaugeMonomial(beta, c0=c0)
    assert mono.improved
    assert abs(mono.action(lat) - o.improved_gauge_action(beta, c0)) < 1e-12 * abs(o.improved_gauge_action(beta, c0))
    lat.set_momenta(P)
    mono.update(lat, 0.37)
    assert np.abs(lat.get_momenta() - want).max() < 1e-13 * np.abs(want).max()
    with pytest.raises(yh.YahmcError):
        lat.avr_wilson(0, 1)


def test_improved_sf_o2mn_trajectory_matches_oracle():
    """A five-step o2mn trajectory of the improved gauge action with SF
    boundaries: dH, links and momenta agree with the oracle."""
    print("\n[case] improved o2mn trajectory su3_fund sf", flush=True)
    o, lat, rng = setup_bc("su3_fund", 3, "fundamental", "sf", seed=31)
    beta, c0 = 4.2, C0_LW
    c1 = (1.0 - c0) / 8.0
    P = rng.normal(size=(4 * o.V, 8))
    o.set_momenta(P)
    lat.set_momenta(P)

    def ham(x):
        if x is lat:
            v1 = lat.avr_wilson(2, 1) + lat.avr_wilson(1, 2)
            return x.momenta_energy() - beta * (c0 * x.plaquette() + c1 * v1) * 6 * lat.gauge_volume
        return x.kinetic() + x.improved_gauge_action(beta, c0)

    h0d, h0o = ham(lat), ham(o)
    assert abs(h0d - h0o) < 1e-10 * abs(h0o)
    a, steps, length = 0.1931833275037836, 5, 0.5
    dt = length / steps
    for x in (o, lat):
        for n in range(steps):
            x.improved_gauge_force(a * dt if n == 0 else 2 * a * dt, beta, c0, c1)
            x.update_links(dt / 2)
            x.improved_gauge_force((1 - 2 * a) * dt, beta, c0, c1)
            x.update_links(dt / 2)
        x.improved_gauge_force(a * dt, beta, c0, c1)
    dHd, dHo = ham(lat) - h0d, ham(o) - h0o
    print(f"dH device {dHd:.6e} oracle {dHo:.6e}")
    assert abs(dHd - dHo) < 1e-9 * abs(h0o)
    assert np.abs(lat.get_gauge() - o.get_gauge()).max() < 1e-12
    assert np.abs(lat.get_momenta() - o.get_momenta()).max() < 1e-11


def test_sf_needs_nc_up_to_3():
    lat = yh.Lattice((6, 4, 4, 4), nc=4, repr="antisymmetric", bc="sf", csw=0.0)
    rng = np.random.default_rng(1)
    lat.set_gauge(random_gauge(rng, lat.V, 4))
    with pytest.raises(yh.YahmcError) as e:
        lat.plaquette()
    assert e.value.code == -6


def test_config_save_load_roundtrip(tmp_path):
    """gauge_field_save / gauge_field_load through the device: a field moved
    by an MD step, saved (native and HiRep) and loaded into a fresh lattice,
    gives the same plaquette (stored = recomputed) and the same operator."""
    print("\n[case] config IO round trip", flush=True)
    o, lat, rng = setup("su3_fund", 3, "fundamental", seed=19)
    lat.set_momenta(rng.normal(0.0, np.sqrt(0.5), size=(4 * o.V, 8)))
    lat.update_links(0.05)
    lat.represent_gauge_field()
    s = gaussian_spinor(rng, o.V, 3)
    fs = lat.field("f64").upload(s)
    d0 = lat.field("f64")
    lat.Dphi(-0.3, d0, fs)
    for fmt, name in (("native", "c.bin"), ("hirep", "run_n7")):
        lat.save_config(tmp_path / name, format=fmt, num_cnfg=7, num_accept=5)
        lat2 = yh.Lattice(o.dims, nc=3, repr="fundamental", bc="antiperiodic", csw=1.0)
        info, dplaq = lat2.load_config(tmp_path / name, format=fmt)
        assert info["num_cnfg"] == 7 and dplaq < 1e-15
        assert np.array_equal(lat2.get_gauge(), lat.get_gauge())
        d1 = lat2.field("f64")
        lat2.Dphi(-0.3, d1, lat2.field("f64").upload(s))
        assert np.array_equal(d1.download()[: o.VH], d0.download()[: o.VH])
    lat3 = yh.Lattice((4, 4, 4, 4), nc=3, repr="fundamental", bc="antiperiodic", csw=1.0)
    with pytest.raises(yh.YahmcError):
        lat3.load_config(tmp_path / "c.bin")
