"""LDS staging of the hop's y / z neighbour spinors (YH_TUNE_HOP_STAGE).

The staged kernels copy the window of input spinors a workgroup's y and z
neighbours come from (plus the two y-wrap rows of its (t, x) plane) into
LDS and read those neighbours there; the arithmetic is unchanged, so every
operator must equal the unstaged one bit for bit, and the oracle's to the
FP64 bar.  Lattices with a (t, x) plane of a multiple of 256 sites, where
the staged kernels run: L_Z = 32 (window halo 17) and L_Z = 64 (halo 33)."""
import numpy as np
import pytest

from test_gpu_parity import BC_ANTIPERIODIC, BC_OPEN, rel, setup, src_field

pytestmark = pytest.mark.gpu
yh = pytest.importorskip("yahmc_amd")


@pytest.mark.parametrize("model,dims,bc", [("su3_fund", (4, 4, 16, 32), BC_ANTIPERIODIC),
                                           ("su3_fund", (4, 2, 8, 64), BC_OPEN),
                                           ("su2_fund", (4, 4, 16, 32), BC_ANTIPERIODIC),
                                           ("su2_adj", (4, 2, 16, 32), BC_ANTIPERIODIC)])
def test_staged_hop_matches_unstaged_and_oracle(model, dims, bc):
    o, lat, rng = setup(model, dims=dims, bc=bc, csw=1.0, cf=1.1)
    mass = -0.75
    s = src_field(o, rng)
    # FP64 against the oracle, staged
    lat.tune(yh.TUNE_HOP_STAGE, 7)
    fs, fd = lat.field("f64").upload(s), lat.field("f64")
    for name, part in (("Dphi_eo", slice(0, o.VH)), ("Dphi_oe", slice(o.VH, o.V))):
        want = np.zeros_like(s)
        getattr(o, name)(want, s)
        getattr(lat, name)(fd, fs)
        assert rel(fd.download()[part], want[part]) < 1e-13, name
    want = np.zeros_like(s)
    o.Hphi_sq(mass, want, s.copy())
    lat.Hphi_sq(mass, fd, fs)
    assert rel(fd.download()[: o.VH], want[: o.VH]) < 1e-13
    # every precision: staged == unstaged, bit for bit
    for prec in ("f64", "f32", "f16"):
        fd = lat.field(prec)
        outs = []
        for stage in (0, 1):
            lat.tune(yh.TUNE_HOP_STAGE, 7 if stage else 0)
            # a fresh input each time: Dphi writes -D_oe src into its odd half
            fs = lat.field(prec).upload(s)
            got = []
            lat.Dphi_eo(fd, fs)
            got.append(fd.download()[: o.VH])
            lat.Dphi(mass, fd, fs)
            got.append(fd.download()[: o.VH])
            lat.Hphi_sq(mass, fd, fs)
            got.append(fd.download()[: o.VH])
            outs.append(got)
        for a, b in zip(outs[0], outs[1]):
            assert np.array_equal(a, b), prec
    lat.tune(yh.TUNE_HOP_STAGE, 3)


def test_staged_cg_matches_unstaged():
    o, lat, rng = setup("su3_fund", dims=(4, 4, 16, 32), bc=BC_ANTIPERIODIC)
    s = src_field(o, rng, even_only=True)
    fb = lat.field("f64").upload(s)
    sols = []
    for stage in (0, 1):
        lat.tune(yh.TUNE_HOP_STAGE, 7 if stage else 0)
        fx = lat.field("f64")
        it = lat.cg_inverter(lat.inv_par(-0.75, 1e-24, "Hphi_sq"), fx, fb)
        sols.append((it, fx.download()[: o.VH]))
        fx2 = lat.field("f64")
        it2, upd = lat.cg_mixed(lat.inv_par(-0.75, 1e-24, "Hphi_sq", sloppy="f16", reliable_every=50), fx2, fb)
        sols.append((it2, fx2.download()[: o.VH]))
    lat.tune(yh.TUNE_HOP_STAGE, 3)
    assert sols[0][0] == sols[2][0] and np.array_equal(sols[0][1], sols[2][1])
    assert sols[1][0] == sols[3][0] and np.array_equal(sols[1][1], sols[3][1])
