"""MINRES (library/minres.cpp), auto_inverter (library/inverter.cpp) and
the MRE initial guess (library/mre.cpp) on the device against the oracle."""
import numpy as np
import pytest

from fields import gaussian_spinor, random_gauge
from oracle import BC_ANTIPERIODIC, MVM_DPHI, MVM_HPHI, MVM_HPHI_SQ, Oracle, OracleMRE

pytestmark = pytest.mark.gpu
yh = pytest.importorskip("yahmc_amd")


def setup(model="su3_fund", nc=3, rep="fundamental", dims=(4, 4, 4, 4), seed=31):
    o = Oracle(model).init(dims, bc=BC_ANTIPERIODIC, csw=1.0)
    rng = np.random.default_rng(seed)
    links = random_gauge(rng, o.V, nc)
    o.set_gauge(links)
    o.represent()
    lat = yh.Lattice(dims, nc=nc, repr=rep, bc="antiperiodic", csw=1.0)
    lat.set_gauge(links)
    return o, lat, rng


def source(o, rng):
    b = gaussian_spinor(rng, o.V, o.dr)
    b[o.VH:] = 0
    return b


@pytest.mark.parametrize("model,nc,rep", [("su3_fund", 3, "fundamental"), ("su2_adj", 2, "adjoint")])
@pytest.mark.parametrize("mvm,name", [(MVM_HPHI, "Hphi"), (MVM_HPHI_SQ, "Hphi_sq")])
def test_minres_matches_oracle(model, nc, rep, mvm, name):
    print(f"\n[case] minres {model} {name}", flush=True)
    o, lat, rng = setup(model, nc, rep)
    mass, prec = -0.2, 1e-24
    b = source(o, rng)
    xo = np.zeros_like(b)
    it_o, _ = o.minres(mvm, mass, prec, xo, b.copy())
    fx = lat.field("f64")
    fx.zero()
    it = lat.minres_inverter(lat.inv_par(mass, prec, name), fx, lat.field("f64").upload(b))
    x = fx.download()
    print(f"iterations device {it} oracle {it_o}")
    assert abs(it - it_o) <= 2
    assert np.abs(x[: o.VH] - xo[: o.VH]).max() < 1e-10 * np.abs(xo).max()


def test_auto_inverter_matches_oracle():
    print("\n[case] auto_inverter", flush=True)
    o, lat, rng = setup(seed=33)
    mass, prec = -0.2, 1e-24
    b = source(o, rng)
    o.lib.or_zero_auto_scratch()
    for name, mvm in (("Hphi_sq", MVM_HPHI_SQ), ("Hphi", MVM_HPHI), ("Dphi", MVM_DPHI)):
        xo = np.zeros_like(b)
        bo = b.copy()
        it_o = o.auto_inverter(mvm, mass, prec, xo, bo)
        fb = lat.field("f64").upload(b)
        fx = lat.field("f64")
        fx.zero()
        it = lat.auto_inverter(lat.inv_par(mass, prec, name), fx, fb)
        x = fx.download()
        print(f"{name}: iterations device {it} oracle {it_o}")
        assert abs(it - it_o) <= 4
        assert np.abs(x[: o.VH] - xo[: o.VH]).max() < 1e-9 * np.abs(xo).max()
        assert np.array_equal(fb.download()[: o.VH], b[: o.VH])  # source restored


@pytest.mark.parametrize("mvm,name", [(MVM_DPHI, "Dphi"), (MVM_HPHI, "Hphi"), (MVM_HPHI_SQ, "Hphi_sq")])
def test_mre_guess_matches_oracle_and_saves_iterations(mvm, name):
    print(f"\n[case] mre {name}", flush=True)
    o, lat, rng = setup(seed=35)
    mass = -0.2
    mre_o = OracleMRE(o, 5, mvm, mass)
    mre = lat.mre(5, name, mass)
    for k in range(6):  # more than max: the oldest drops out
        v = source(o, rng)
        mre_o.store(v)
        mre.store(lat.field("f64").upload(v))
    b = source(o, rng)
    want = mre_o.guess(b)
    fg = lat.field("f64")
    mre.guess(fg, lat.field("f64").upload(b))
    got = fg.download()
    assert np.abs(got[: o.VH] - want[: o.VH]).max() < 1e-10 * np.abs(want).max()
    if name != "Dphi":
        return
    # a slowly varying sequence of right-hand sides: BiCGstab from the MRE
    # guess needs fewer iterations than from zero
    mre.reset()
    base = source(o, rng)
    par = lat.inv_par(mass, 1e-20, "Dphi")
    its_mre, its_zero = [], []
    for k in range(6):
        bk = base + 0.01 * source(o, rng)
        fb = lat.field("f64").upload(bk)
        x0 = lat.field("f64")
        x0.zero()
        its_zero.append(lat.bicgstab_inverter(par, x0, fb))
        x1 = lat.field("f64")
        mre.guess(x1, fb)
        its_mre.append(lat.bicgstab_inverter(par, x1, fb))
        mre.store(x1)
    print(f"BiCGstab iterations from zero {its_zero}, from the MRE guess {its_mre}")
    assert sum(its_mre[2:]) < sum(its_zero[2:])
