"""yh_cg_noalloc: the CG entry point that allocates nothing (its temporaries
live in the library's per-GPU scratch array).  Checked against yh_cg on the
same inputs, for two calls in a row, the second on a smaller lattice (the
array keeps the first call's contents), and for the size limit."""
import numpy as np
import pytest

from fields import gaussian_spinor, random_gauge
from oracle import BC_ANTIPERIODIC, Oracle

pytestmark = pytest.mark.gpu
yh = pytest.importorskip("yahmc_amd")


def solve_both(dims, seed, mvm, mass=-0.75, prec=1e-24):
    print(f"\n[case] cg_noalloc dims={dims} seed={seed} mvm={mvm}", flush=True)
    rng = np.random.default_rng(seed)
    V = int(np.prod(dims))
    links = random_gauge(rng, V, 3)
    lat = yh.Lattice(dims, nc=3, repr="fundamental", bc="antiperiodic", csw=1.0)
    lat.set_gauge(links)
    s = gaussian_spinor(rng, V, 3)
    s[V // 2:] = 0
    fb = lat.field("f64").upload(s)
    par = lat.inv_par(mass, prec, mvm)
    fx_ref = lat.field("f64")
    it_ref = lat.cg_inverter(par, fx_ref, fb)
    fx = lat.field("f64")
    it = lat.cg_noalloc(par, fx, fb)
    return lat, it_ref, it, fx_ref.download(), fx.download(), links, s


def test_noalloc_cg_matches_cg_twice_second_smaller():
    first = solve_both((8, 8, 8, 8), 31, "Hphi_sq")
    second = solve_both((4, 4, 4, 8), 32, "Hphi_sq")
    for lat, it_ref, it, x_ref, x, links, s in (first, second):
        VH = lat.VH
        assert it == it_ref
        # the same kernels in the same order on the same inputs: identical iterates
        assert np.array_equal(x[:VH], x_ref[:VH])
    # and the solution is the oracle's: true residual with the oracle operator
    lat, it_ref, it, x_ref, x, links, s = second
    o = Oracle("su3_fund").init((4, 4, 4, 8), bc=BC_ANTIPERIODIC, csw=1.0)
    o.set_gauge(links)
    o.represent()
    ax = np.zeros_like(x)
    o.Hphi_sq(-0.75, ax, x.copy())
    VH = lat.VH
    r = s[:VH] - ax[:VH]
    assert float(np.sum(r * r) / np.sum(s[:VH] ** 2)) < 1e-24


def test_noalloc_cg_rejects_a_lattice_beyond_the_scratch():
    dims = (32, 16, 16, 16)  # 5 x 64 x 3 x VHpad bytes > 32 MiB
    print(f"\n[case] cg_noalloc size limit dims={dims}", flush=True)
    rng = np.random.default_rng(5)
    V = int(np.prod(dims))
    lat = yh.Lattice(dims, nc=3, repr="fundamental", bc="antiperiodic", csw=1.0)
    lat.set_gauge(random_gauge(rng, V, 3))
    fb, fx = lat.field("f64"), lat.field("f64")
    with pytest.raises(yh.YahmcError) as ei:
        lat.cg_noalloc(lat.inv_par(-0.75, 1e-20, "Hphi_sq"), fx, fb)
    assert "scratch" in str(ei.value)
