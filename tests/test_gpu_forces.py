"""GPU parity of the fermion forces against the oracle (whose forces are
pinned by finite differences of the action, tests/test_oracle_forces.py).

The sequence is MonomialHMC::update's (monomials/hmc.cpp:67-93) inside
begin/end_fermion_force (monomials/list.cpp:11-30), starting from random
momenta so the accumulation is checked too.  Tolerance: 1e-12 relative on
the momenta and on the clover force matrices.
"""
import numpy as np
import pytest

from fields import gaussian_spinor, random_gauge
from oracle import BC_ANTIPERIODIC, BC_OPEN, BC_PERIODIC, Oracle

pytestmark = pytest.mark.gpu
yh = pytest.importorskip("yahmc_amd")

MODELS = {"su3_fund": (3, "fundamental"), "su2_adj": (2, "adjoint"), "su3_sym": (3, "symmetric"),
          "su2_fund": (2, "fundamental"), "su3_adj": (3, "adjoint"), "su3_asym": (3, "antisymmetric"),
          "su4_asym": (4, "antisymmetric")}
BC_NAMES = {BC_PERIODIC: "periodic", BC_ANTIPERIODIC: "antiperiodic", BC_OPEN: "open"}


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("model,csw,bc", [("su3_fund", 1.0, BC_ANTIPERIODIC), ("su3_fund", 0.0, BC_PERIODIC),
                                          ("su2_adj", 1.2, BC_ANTIPERIODIC), ("su3_sym", 1.0, BC_ANTIPERIODIC),
                                          ("su2_fund", 1.0, BC_PERIODIC), ("su3_adj", 1.0, BC_ANTIPERIODIC),
                                          ("su3_asym", 1.1, BC_PERIODIC), ("su4_asym", 1.0, BC_ANTIPERIODIC),
                                          ("su3_fund", 1.0, BC_OPEN)])
def test_hmc_force_sequence(model, csw, bc):
    print(f"\n[case] forces model={model} csw={csw} bc={bc}", flush=True)
    nc, rep = MODELS[model]
    dims = (6, 4, 4, 4) if bc == BC_OPEN else (4, 4, 4, 4)
    o = Oracle(model).init(dims, bc=bc, csw=csw)
    rng = np.random.default_rng(5)
    links = random_gauge(rng, o.V, nc)
    o.set_gauge(links)
    o.represent()
    lat = yh.Lattice(dims, nc=nc, repr=rep, bc=BC_NAMES[bc], csw=csw)
    lat.set_gauge(links)
    mass, dt = -0.3, 0.7
    ng = nc * nc - 1
    mom0 = rng.normal(size=(4 * o.V, ng))
    o.set_momenta(mom0)
    lat.set_momenta(mom0)
    assert np.array_equal(lat.get_momenta(), mom0)
    # a solution-like field on the even sites, arbitrary odd halves
    b = gaussian_spinor(rng, o.V, o.dr)
    o.begin_fermion_force()
    lat.begin_fermion_force()
    fb = lat.field("f64").upload(b)
    ob = b.copy()
    o.wilson_fermion_force_q(dt, mass, 1.3, ob)
    lat.wilson_fermion_force_q(dt, mass, 1.3, fb)
    assert rel(lat.get_momenta() - mom0, o.get_momenta() - mom0) < 1e-12
    # odd half of the argument now holds D_oo^-1 D_oe b_e (forces.cpp:182-188)
    assert rel(fb.download()[o.VH:], ob[o.VH:]) < 1e-12
    if csw != 0.0:
        assert rel(lat.get_clover_force(), o.get_clover_force()) < 1e-12
    o.clover_logdet_force(mass, 2.0)
    lat.clover_logdet_force(mass, 2.0)
    if csw != 0.0:
        assert rel(lat.get_clover_force(), o.get_clover_force()) < 1e-12
    o.end_fermion_force(dt)
    lat.end_fermion_force(dt)
    assert rel(lat.get_momenta() - mom0, o.get_momenta() - mom0) < 1e-12


def test_clover_force_core_work_orders_agree():
    """The clover-force core gives the same momenta (bit for bit) whichever
    order its workgroups take the (site block, mu) items in."""
    print("\n[case] clover-force core work orders", flush=True)
    nc, rep = MODELS["su3_sym"]
    dims = (4, 4, 4, 8)
    lat = yh.Lattice(dims, nc=nc, repr=rep, bc="antiperiodic", csw=1.0)
    rng = np.random.default_rng(11)
    V = lat.V
    lat.set_gauge(random_gauge(rng, V, nc))
    a = lat.field("f64").upload(gaussian_spinor(rng, V, 6))
    b = lat.field("f64").upload(gaussian_spinor(rng, V, 6))
    mom0 = rng.normal(size=(4 * V, nc * nc - 1))
    out = []
    for order in (-1, 0, 1, 4):
        lat.tune(yh.TUNE_FORCE_ORDER, order)
        lat.set_momenta(mom0)
        lat.begin_fermion_force()
        lat.clover_fermion_force(1.0, a, b)
        lat.end_fermion_force(0.3)
        out.append(lat.get_momenta())
    assert np.abs(out[0] - mom0).max() > 0
    for m in out[1:]:
        assert np.array_equal(m, out[0])
