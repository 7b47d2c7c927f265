"""Operator-application counter, FP64 12-real links, and contexts of
different models alive at the same time.

mvm_get: the reference counts every dirac_kernel call and reports half of
that count in the EO build (library/dirac.cpp:9-23, 233), so one Dphi, Hphi
or Dphi_oe + Dphi_eo pair counts 1 and one Hphi_sq counts 2.  A solver's
count therefore follows from its iteration count:
  cg_inverter (cg.cpp:132-202)        initial residual + one H^2 per iteration
  bicgstab_inverter (bicgstab.cpp)    initial residual + two Dphi per iteration
  cg_mshift (cg.cpp:11-130)           one H^2 per iteration, zero guess
The oracle restates the same counter; both sides are checked against these
formulas, and the device count must not include the launches that a solver
queues past convergence."""
import numpy as np
import pytest

from fields import gaussian_spinor, random_gauge
from oracle import BC_ANTIPERIODIC, BC_PERIODIC, MVM_HPHI_SQ, Oracle

pytestmark = pytest.mark.gpu
yh = pytest.importorskip("yahmc_amd")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def setup(model="su3_fund", nc=3, rep="fundamental", dims=(4, 4, 4, 8), seed=21, csw=1.0, bc=BC_ANTIPERIODIC):
    print(f"\n[case] model={model} dims={dims} seed={seed}", flush=True)
    o = Oracle(model).init(dims, bc=bc, csw=csw)
    rng = np.random.default_rng(seed)
    links = random_gauge(rng, o.V, nc)
    o.set_gauge(links)
    o.represent()
    lat = yh.Lattice(dims, nc=nc, repr=rep, bc="antiperiodic" if bc == BC_ANTIPERIODIC else "periodic", csw=csw)
    lat.set_gauge(links)
    return o, lat, rng, links


def even_source(o, rng):
    s = gaussian_spinor(rng, o.V, o.dr)
    s[o.VH:] = 0
    return s


def test_mvm_count_of_operators():
    o, lat, rng, _ = setup()
    s = even_source(o, rng)
    fs, fd = lat.field("f64").upload(s), lat.field("f64")
    lat.mvm_reset()
    lat.Dphi(-0.5, fd, fs)
    assert lat.mvm_get() == 1
    lat.Hphi(-0.5, fd, fs)
    assert lat.mvm_get() == 2
    lat.Hphi_sq(-0.5, fd, fs)
    assert lat.mvm_get() == 4
    lat.Dphi_oe(fd, fs)
    lat.Dphi_eo(fd, fs)
    assert lat.mvm_get() == 5
    o.mvm_reset()
    d = np.zeros_like(s)
    o.Dphi(-0.5, d, s.copy())
    o.Hphi(-0.5, d, s.copy())
    o.Hphi_sq(-0.5, d, s.copy())
    o.Dphi_oe(d, s.copy())
    o.Dphi_eo(d, s.copy())
    assert o.mvm_get() == 5


@pytest.mark.parametrize("prec", [1e-20, 1e-26])
def test_mvm_count_of_cg(prec):
    """CG on H^2 from a zero guess: 2 (n + 1) for n iterations, on both
    sides, whatever n is modulo the device's 16-iteration chunks."""
    o, lat, rng, _ = setup()
    mass = -0.5
    b = even_source(o, rng)
    o.mvm_reset()
    it_or, _ = o.cg(MVM_HPHI_SQ, mass, prec, np.zeros_like(b), b)
    assert o.mvm_get() == 2 * (it_or + 1)
    lat.mvm_reset()
    it = lat.cg_inverter(lat.inv_par(mass, prec, "Hphi_sq"), lat.field("f64"), lat.field("f64").upload(b))
    print(f"iterations gpu={it} oracle={it_or}, mvm gpu={lat.mvm_get()}")
    assert lat.mvm_get() == 2 * (it + 1)
    assert abs(it - it_or) <= 2


def test_mvm_count_of_bicgstab_mshift_and_mixed_cg():
    o, lat, rng, _ = setup()
    mass = -0.5
    b = even_source(o, rng)
    fb = lat.field("f64").upload(b)
    lat.mvm_reset()
    it = lat.bicgstab_inverter(lat.inv_par(mass, 1e-20, "Dphi"), lat.field("f64"), fb)
    assert lat.mvm_get() == 1 + it  # `it` is twice the iterations
    o.mvm_reset()
    it_or, _ = o.bicgstab(0, mass, 1e-20, np.zeros_like(b), b)
    assert o.mvm_get() == 1 + it_or
    shifts = [0.3, 0.05, 1.2]
    lat.mvm_reset()
    it = lat.cg_mshift(lat.inv_par(mass, 1e-20, "Hphi_sq"), shifts, [lat.field("f64") for _ in shifts], fb)
    assert lat.mvm_get() == 2 * it
    o.mvm_reset()
    it_or, _ = o.cg_mshift(MVM_HPHI_SQ, mass, 1e-20, shifts, b)
    assert o.mvm_get() == 2 * it_or
    # batched: every right-hand side counts its own iterations
    b2 = even_source(o, rng)
    srcs = [fb, lat.field("f64").upload(b2)]
    dsts = [[lat.field("f64") for _ in shifts] for _ in srcs]
    lat.mvm_reset()
    its = lat.cg_mshift_batched(lat.inv_par(mass, 1e-20, "Hphi_sq"), shifts, dsts, srcs)
    assert lat.mvm_get() == 2 * sum(its)
    # mixed precision: FP64 initial residual, one sloppy H^2 per iteration,
    # one FP64 H^2 per reliable update
    lat.mvm_reset()
    it, upd = lat.cg_mixed(lat.inv_par(mass, 1e-20, "Hphi_sq", sloppy="f16", reliable_every=50),
                           lat.field("f64"), fb)
    assert lat.mvm_get() == 2 * (1 + it + upd)


def test_fp64_12_real_links():
    """FP64 operators reading 12-real SU(3) links (third row conj(r0 x r1))
    against the oracle's 18-real operator: the reconstruction is exact up to
    rounding for SU(3) links."""
    o, lat, rng, _ = setup(dims=(4, 4, 4, 8))
    lat.tune(yh.TUNE_RECON12_F64, 1)
    mass = -0.75
    s = even_source(o, rng)
    o.apply_bcs_on_spinor(s)
    fs, fd = lat.field("f64").upload(s), lat.field("f64")
    want = np.zeros_like(s)
    o.Dphi_eo(want, s.copy())
    lat.Dphi_eo(fd, fs)
    assert rel(fd.download()[: o.VH], want[: o.VH]) < 1e-13
    for name in ("Dphi", "Hphi", "Hphi_sq"):
        want = np.zeros_like(s)
        getattr(o, name)(mass, want, s.copy())
        getattr(lat, name)(mass, fd, lat.field("f64").upload(s))
        assert rel(fd.download()[: o.VH], want[: o.VH]) < 1e-13, name
    prec = 1e-24
    fx = lat.field("f64")
    it = lat.cg_inverter(lat.inv_par(mass, prec, "Hphi_sq"), fx, lat.field("f64").upload(s))
    x = fx.download()
    ax = np.zeros_like(x)
    o.Hphi_sq(mass, ax, x.copy())
    r = s[: o.VH] - ax[: o.VH]
    assert np.sum(r * r) / np.sum(s[: o.VH] ** 2) < prec
    lat.tune(yh.TUNE_RECON12_F64, 0)
    with pytest.raises(yh.YahmcError):
        lat2 = yh.Lattice((4, 4, 4, 4), nc=2, repr="fundamental", bc="antiperiodic", csw=1.0)
        lat2.tune(yh.TUNE_RECON12_F64, 1)


def test_two_live_contexts_of_different_models():
    """An SU(2) fundamental context stays correct while an SU(3) sextet
    context (other representation basis and generators) is created and
    used: represented links, clover term, fermion forces, gauge force and
    link update of the first one against the oracle afterwards."""
    o2, lat2, rng, links2 = setup("su2_fund", 2, "fundamental", dims=(4, 4, 4, 4), seed=31, bc=BC_PERIODIC)
    o6, lat6, _, links6 = setup("su3_sym", 3, "symmetric", dims=(4, 4, 4, 4), seed=32, bc=BC_PERIODIC)
    assert rel(lat6.gauge_repr(), o6.get_repr()) < 1e-13
    # the SU(2) context again, after the sextet context was set up
    lat2.set_gauge(links2)
    assert rel(lat2.gauge_repr(), o2.get_repr()) < 1e-13
    assert rel(lat2.clover_term(), o2.get_clover()) < 1e-13
    mass = -0.5
    x = gaussian_spinor(rng, o2.V, o2.dr)
    y = gaussian_spinor(rng, o2.V, o2.dr)
    P = rng.normal(size=(4 * o2.V, 3))
    o2.set_momenta(P)
    lat2.set_momenta(P)
    o2.begin_fermion_force()
    lat2.begin_fermion_force()
    o2.wilson_fermion_force(0.1, mass, 1.0, x.copy(), y.copy())
    fx, fy = lat2.field("f64").upload(x), lat2.field("f64").upload(y)
    lat2.wilson_fermion_force(0.1, mass, 1.0, fx, fy)
    o2.end_fermion_force(0.1)
    lat2.end_fermion_force(0.1)
    assert rel(lat2.get_momenta(), o2.get_momenta()) < 1e-12
    o2.wilson_gauge_force(0.2, 2.2)
    lat2.wilson_gauge_force(0.2, 2.2)
    assert rel(lat2.get_momenta(), o2.get_momenta()) < 1e-12
    o2.update_links(0.05)
    lat2.update_links(0.05)
    assert rel(lat2.get_gauge(), o2.get_gauge()) < 1e-13
    # and the sextet context is still right too
    P6 = rng.normal(size=(4 * o6.V, 8))
    o6.set_momenta(P6)
    lat6.set_momenta(P6)
    o6.wilson_gauge_force(0.2, 5.6)
    lat6.wilson_gauge_force(0.2, 5.6)
    assert rel(lat6.get_momenta(), o6.get_momenta()) < 1e-12
    lat6.close()
    lat2.close()
