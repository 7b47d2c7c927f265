"""SU(3) adjoint fermions (D_R = 8) through the solvers: FP64 CG against the
oracle's cg_inverter, the mixed-precision CGs and the multishift CG to their
FP64 true residuals with the oracle's operator (library/cg.cpp:11-202,
library/repr.cpp:160-261 for the representation)."""
import numpy as np
import pytest

from fields import gaussian_spinor, random_gauge
from oracle import BC_ANTIPERIODIC, MVM_HPHI_SQ, Oracle

pytestmark = pytest.mark.gpu
yh = pytest.importorskip("yahmc_amd")


def setup(dims=(4, 4, 4, 4), seed=23):
    print(f"\n[case] su3_adj dims={dims} seed={seed}", flush=True)
    o = Oracle("su3_adj").init(dims, bc=BC_ANTIPERIODIC, csw=1.0)
    rng = np.random.default_rng(seed)
    links = random_gauge(rng, o.V, 3)
    o.set_gauge(links)
    o.represent()
    lat = yh.Lattice(dims, nc=3, repr="adjoint", bc="antiperiodic", csw=1.0)
    lat.set_gauge(links)
    b = gaussian_spinor(rng, o.V, 8)
    b[o.VH:] = 0
    return o, lat, b


def true_res(o, mass, x, b, shift=0.0):
    ax = np.zeros_like(x)
    o.Hphi_sq(mass, ax, x.copy())
    r = b[: o.VH] - ax[: o.VH] - shift * x[: o.VH]
    return float(np.sum(r * r) / np.sum(b[: o.VH] ** 2))


def test_cg_matches_oracle():
    o, lat, b = setup()
    mass, prec = -0.5, 1e-26
    x_or = np.zeros_like(b)
    it_or, _ = o.cg(MVM_HPHI_SQ, mass, prec, x_or, b)
    fb, fx = lat.field("f64").upload(b), lat.field("f64")
    it = lat.cg_inverter(lat.inv_par(mass, prec, "Hphi_sq"), fx, fb)
    x = fx.download()
    assert abs(it - it_or) <= 2
    assert np.abs(x[: o.VH] - x_or[: o.VH]).max() / np.abs(x_or[: o.VH]).max() < 1e-10
    assert true_res(o, mass, x, b) < prec


@pytest.mark.parametrize("sloppy", ["f32", "f16"])
def test_mixed_cg_true_residual(sloppy):
    o, lat, b = setup(seed=24)
    mass, prec = -0.5, 1e-22
    fb, fx = lat.field("f64").upload(b), lat.field("f64")
    it, upd = lat.cg_mixed(lat.inv_par(mass, prec, "Hphi_sq", sloppy=sloppy, reliable_every=50), fx, fb)
    assert upd >= 1
    assert true_res(o, mass, fx.download(), b) < prec


def test_multishift_true_residuals():
    o, lat, b = setup(seed=25)
    mass, prec = -0.5, 1e-24
    shifts = [0.05, 0.4, 1.5]
    fb = lat.field("f64").upload(b)
    xs = [lat.field("f64") for _ in shifts]
    lat.cg_mshift(lat.inv_par(mass, prec, "Hphi_sq"), shifts, xs, fb)
    for s, fx in zip(shifts, xs):
        assert true_res(o, mass, fx.download(), b, s) < 1e-22
