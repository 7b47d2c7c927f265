"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded gauge field and sources.

Tolerances: FP64 operator outputs within 1e-13 relative (max |diff| / max
|ref|); representation and clover term within 1e-13; FP32 operators within
2e-6; 16-bit fixed point within 2e-4; solvers to the reference's own squared relative
residual (1e-28 for CG / BiCGstab / multishift, checks/inverters.cpp:7).
"""
import numpy as np
import pytest

from fields import gaussian_spinor, random_gauge
from oracle import BC_ANTIPERIODIC, BC_OPEN, BC_PERIODIC, BC_SF, MVM_HPHI_SQ, Oracle

pytestmark = pytest.mark.gpu

yh = pytest.importorskip("yahmc_amd")

MODELS = {"su3_fund": (3, "fundamental"), "su2_fund": (2, "fundamental"), "su2_adj": (2, "adjoint"),
          "su3_sym": (3, "symmetric"), "su3_asym": (3, "antisymmetric"), "su4_asym": (4, "antisymmetric"),
          "su3_adj": (3, "adjoint")}
BCN = {BC_PERIODIC: "periodic", BC_ANTIPERIODIC: "antiperiodic", BC_OPEN: "open", BC_SF: "sf"}


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def setup(model="su3_fund", dims=(4, 4, 4, 4), bc=BC_ANTIPERIODIC, csw=1.0, cf=1.0, seed=11):
    print(f"\n[case] model={model} dims={dims} bc={BCN[bc]} csw={csw} seed={seed}", flush=True)
    nc, rep = MODELS[model]
    o = Oracle(model).init(dims, bc=bc, csw=csw, cf=cf)
    rng = np.random.default_rng(seed)
    links = random_gauge(rng, o.V, nc)
    o.set_gauge(links)
    if bc in (BC_OPEN, BC_SF):
        o.apply_bcs_on_gauge()
        links = o.get_gauge()
    o.represent()
    lat = yh.Lattice(dims, nc=nc, repr=rep, bc=BCN[bc], csw=csw, cf=cf)
    lat.set_gauge(links)
    return o, lat, rng


def src_field(o, rng, even_only=False):
    s = gaussian_spinor(rng, o.V, o.dr)
    if even_only:
        s[o.VH:] = 0
    o.apply_bcs_on_spinor(s)
    return s


@pytest.mark.parametrize("model", list(MODELS))
def test_represent_and_clover_term(model):
    o, lat, _ = setup(model, bc=BC_ANTIPERIODIC)
    assert rel(lat.gauge_repr(), o.get_repr()) < 1e-13
    assert rel(lat.clover_term(), o.get_clover()) < 1e-13


@pytest.mark.parametrize("bc", [BC_PERIODIC, BC_ANTIPERIODIC, BC_OPEN, BC_SF])
def test_hop_eo_oe(bc):
    o, lat, rng = setup("su3_fund", dims=(6, 4, 4, 4), bc=bc)
    s = src_field(o, rng)
    fs = lat.field("f64").upload(s)
    fd = lat.field("f64")
    want = np.zeros_like(s)
    o.Dphi_eo(want, s)
    lat.Dphi_eo(fd, fs)
    got = fd.download()
    assert rel(got[: o.VH], want[: o.VH]) < 1e-13
    want = np.zeros_like(s)
    o.Dphi_oe(want, s)
    lat.Dphi_oe(fd, fs)
    got = fd.download()
    assert rel(got[o.VH:], want[o.VH:]) < 1e-13


@pytest.mark.parametrize("bc", [BC_ANTIPERIODIC, BC_OPEN])
def test_clover_apply_and_inverse(bc):
    o, lat, rng = setup("su3_fund", dims=(6, 4, 4, 4), bc=bc, cf=1.25)
    mass = -0.75
    s = src_field(o, rng)
    fs = lat.field("f64").upload(s)
    fd = lat.field("f64")
    for name, part in (("Dphi_ee", slice(0, o.VH)), ("Dphi_oo", slice(o.VH, o.V)), ("Dphi_oo_inv", slice(o.VH, o.V))):
        want = np.zeros_like(s)
        getattr(o, name)(mass, want, s.copy())
        getattr(lat, name)(mass, fd, fs)
        got = fd.download()
        assert rel(got[part], want[part]) < 1e-13, name
    ld = lat.clover_logdet(mass)
    assert abs(ld - o.clover_logdet(mass)) < 1e-11 * abs(ld)


@pytest.mark.parametrize("model,bc", [("su3_fund", BC_ANTIPERIODIC), ("su3_fund", BC_PERIODIC), ("su3_fund", BC_OPEN),
                                      ("su3_fund", BC_SF), ("su2_fund", BC_ANTIPERIODIC), ("su2_adj", BC_ANTIPERIODIC),
                                      ("su3_sym", BC_ANTIPERIODIC), ("su3_asym", BC_OPEN),
                                      ("su4_asym", BC_ANTIPERIODIC), ("su3_adj", BC_ANTIPERIODIC), ("su3_adj", BC_OPEN)])
def test_schur_operators(model, bc):
    o, lat, rng = setup(model, dims=(4, 4, 4, 4), bc=bc, csw=1.0, cf=1.1)
    mass = -0.75
    s = src_field(o, rng, even_only=True)
    for name in ("Dphi", "Hphi", "Hphi_sq"):
        fs = lat.field("f64").upload(s)
        fd = lat.field("f64")
        want = np.zeros_like(s)
        ssrc = s.copy()
        getattr(o, name)(mass, want, ssrc)
        getattr(lat, name)(mass, fd, fs)
        got = fd.download()
        assert rel(got[: o.VH], want[: o.VH]) < 1e-13, name
        if name != "Hphi_sq":
            side = fs.download()
            # odd half of the input holds -D_oe src_e (library/dirac.cpp:743)
            assert rel(side[o.VH:], ssrc[o.VH:]) < 1e-13, name + " side effect"


def test_config1_8x4_operators():
    """8^4, SU(3) fundamental, c_sw = 1, m = -0.75, antiperiodic time."""
    o, lat, rng = setup("su3_fund", dims=(8, 8, 8, 8), bc=BC_ANTIPERIODIC, csw=1.0, seed=32244)
    mass = -0.75
    s = src_field(o, rng, even_only=True)
    fs = lat.field("f64").upload(s)
    fd = lat.field("f64")
    want = np.zeros_like(s)
    o.Dphi_eo(want, s.copy())
    lat.Dphi_eo(fd, fs)
    assert rel(fd.download()[: o.VH], want[: o.VH]) < 1e-13
    want = np.zeros_like(s)
    o.Dphi(mass, want, s.copy())
    lat.Dphi(mass, fd, lat.field("f64").upload(s))
    assert rel(fd.download()[: o.VH], want[: o.VH]) < 1e-13


# 16-bit: fixed-point spinors, links and clover resolve 1.5e-5 of each scale
@pytest.mark.parametrize("prec,tol", [("f32", 2e-6), ("f16", 2e-4)])
def test_low_precision_operators(prec, tol):
    o, lat, rng = setup("su3_fund", dims=(4, 4, 4, 8), bc=BC_ANTIPERIODIC)
    mass = -0.75
    s = src_field(o, rng, even_only=True)
    want = np.zeros_like(s)
    o.Hphi(mass, want, s.copy())
    fs = lat.field(prec).upload(s)
    fd = lat.field(prec)
    lat.Hphi(mass, fd, fs)
    assert rel(fd.download()[: o.VH], want[: o.VH]) < tol
    want = np.zeros_like(s)
    o.Dphi_eo(want, s)
    lat.Dphi_eo(fd, lat.field(prec).upload(s))
    assert rel(fd.download()[: o.VH], want[: o.VH]) < tol


@pytest.mark.parametrize("model", [m for m in MODELS if m != "su3_fund"])
def test_f16_operators_all_representations(model):
    """16-bit operators of every representation (6-dimensional ones included)
    against the oracle's FP64 operator, open and antiperiodic time."""
    o, lat, rng = setup(model, dims=(4, 4, 4, 6), bc=BC_OPEN if model == "su3_asym" else BC_ANTIPERIODIC)
    mass = -0.6
    s = src_field(o, rng, even_only=True)
    want = np.zeros_like(s)
    o.Hphi(mass, want, s.copy())
    fd = lat.field("f16")
    lat.Hphi(mass, fd, lat.field("f16").upload(s))
    assert rel(fd.download()[: o.VH], want[: o.VH]) < 2e-4
    want = np.zeros_like(s)
    o.Dphi_eo(want, s)
    lat.Dphi_eo(fd, lat.field("f16").upload(s))
    assert rel(fd.download()[: o.VH], want[: o.VH]) < 2e-4


def test_f16_storage_round_trip():
    """16-bit fixed point: each component is stored as round(32768 c / m) in
    two's complement, m the site's largest |component|, so the round trip is
    within 2^-16 m (2^-15 m at +m, which clamps to 32767) and exact at -m and 0."""
    o, lat, rng = setup("su3_fund", dims=(4, 4, 4, 4))
    s = src_field(o, rng)
    s[3] = 0.0                       # all-zero site
    s[5, 7] = -np.abs(s[5]).max() * 1.5  # the site's extreme is negative: stored as -32768
    s[9, 2] = np.abs(s[9]).max() * 2.0   # positive extreme: clamps to 32767/32768
    s[11] *= 1e-20                   # tiny site scale
    back = lat.field("f16").upload(s).download()
    m = np.abs(s).max(axis=1, keepdims=True)
    # a site's largest component clamps when it is positive (q = 32767)
    bound = np.where(s == m, m * (2.0 ** -15 + 2e-7), m * (2.0 ** -16 + 2e-7))
    err = np.abs(back - s)
    assert np.all(err <= bound), np.max(err / np.maximum(m, 1e-300))
    assert err[9, 2] > m[9, 0] * 2.0 ** -16  # the clamp is visible
    assert np.all(back[3] == 0.0)
    assert back[5, 7] == np.float32(s[5, 7])


def test_cg_mixed_f16_sextet_true_residual():
    """Mixed CG with a 16-bit sloppy operator on SU(3) two-index symmetric
    fermions reaches the FP64 true residual (prec 1e-20)."""
    o, lat, rng = setup("su3_sym", dims=(6, 6, 6, 6), bc=BC_ANTIPERIODIC, csw=1.0, seed=9901)
    mass, prec = -0.4, 1e-20
    b = src_field(o, rng, even_only=True)
    fx = lat.field("f64")
    it, upd = lat.cg_mixed(lat.inv_par(mass, prec, "Hphi_sq", sloppy="f16", reliable_every=50, max_iter=5000),
                           fx, lat.field("f64").upload(b))
    r = true_res(o, MVM_HPHI_SQ, mass, fx.download(), b)
    print(f"sextet mixed f16: iterations={it} updates={upd} true residual^2={r:.3e}")
    assert r < prec


def test_linalg():
    o, lat, rng = setup("su3_fund")
    a = src_field(o, rng)
    b = src_field(o, rng)
    fa = lat.field("f64").upload(a)
    fb = lat.field("f64").upload(b)
    assert abs(lat.spinor_sqnorm(fa) - o.sqnorm(a)) < 1e-12 * o.sqnorm(a)
    p = lat.spinor_product(fa, fb)
    q = o.product(a, b)
    assert abs(p - q) < 1e-12 * abs(q)
    assert abs(lat.spinor_product_re(fa, fb) - o.product_re(a, b)) < 1e-12 * abs(q)
    lat.spinor_mulc_add_assign(fa, complex(0.3, -0.7), fb)
    want = a.copy()
    ca = (a[:, :12] + 1j * a[:, 12:]) + complex(0.3, -0.7) * (b[:, :12] + 1j * b[:, 12:])
    want[:, :12], want[:, 12:] = ca.real, ca.imag
    assert rel(fa.download()[: o.VH], want[: o.VH]) < 1e-15
    lat.spinor_g5(fb)
    o.g5(b)
    assert rel(fb.download()[: o.VH], b[: o.VH]) == 0


def true_res(o, mvm, mass, x, b):
    ax = np.zeros_like(x)
    o.mvm(mvm, mass, ax, x.copy())
    r = b[: o.VH] - ax[: o.VH]
    return np.sum(r * r) / np.sum(b[: o.VH] ** 2)


def test_cg_config1():
    """Config 1: 8^4, m = -0.75, c_sw = 1, CG on H^2 to relative residual
    1e-14 (prec 1e-28), against the oracle's cg_inverter."""
    o, lat, rng = setup("su3_fund", dims=(8, 8, 8, 8), bc=BC_ANTIPERIODIC, csw=1.0, seed=32244)
    mass, prec = -0.75, 1e-28
    b = src_field(o, rng, even_only=True)
    x_or = np.zeros_like(b)
    it_or, _ = o.cg(MVM_HPHI_SQ, mass, prec, x_or, b)
    fb = lat.field("f64").upload(b)
    fx = lat.field("f64")
    it = lat.cg_inverter(lat.inv_par(mass, prec, "Hphi_sq"), fx, fb)
    x = fx.download()
    print(f"iterations gpu={it} oracle={it_or}")
    assert abs(it - it_or) <= 2
    assert true_res(o, MVM_HPHI_SQ, mass, x, b) < 1e-27
    assert rel(x[: o.VH], x_or[: o.VH]) < 1e-12


def test_bicgstab_and_mshift():
    o, lat, rng = setup("su3_fund", dims=(4, 4, 4, 4), bc=BC_ANTIPERIODIC, csw=1.0)
    mass, prec = -0.2, 1e-28
    b = src_field(o, rng, even_only=True)
    fb = lat.field("f64").upload(b)
    fx = lat.field("f64")
    it = lat.bicgstab_inverter(lat.inv_par(mass, prec, "Dphi"), fx, fb)
    it_or, _ = o.bicgstab(0, mass, prec, np.zeros_like(b), b)
    assert it % 2 == 0 and abs(it - it_or) <= 4
    assert true_res(o, 0, mass, fx.download(), b) < 1e-27
    shifts = [0.5, 0.1, 0.05, 1.05]
    xs = [lat.field("f64") for _ in shifts]
    it = lat.cg_mshift(lat.inv_par(mass, prec, "Hphi_sq"), shifts, xs, fb)
    it_or, xs_or = o.cg_mshift(MVM_HPHI_SQ, mass, prec, shifts, b)
    assert abs(it - it_or) <= 2
    for s, f, xo in zip(shifts, xs, xs_or):
        x = f.download()
        ax = np.zeros_like(x)
        o.Hphi_sq(mass, ax, x.copy())
        r = b[: o.VH] - ax[: o.VH] - s * x[: o.VH]
        assert np.sum(r * r) / np.sum(b[: o.VH] ** 2) < 1e-27
        assert rel(x[: o.VH], xo[: o.VH]) < 1e-11


def test_half_links_under_fp32_spinors():
    """YH_TUNE_HALF_LINKS: the FP32 operator reading FP16 12-real links agrees
    with the oracle to FP16 link accuracy, and the mixed CG built on it still
    reaches the FP64 true residual."""
    o, lat, rng = setup("su3_fund", dims=(8, 8, 8, 8), bc=BC_ANTIPERIODIC, csw=1.0, seed=32244)
    mass, prec = -0.75, 1e-20
    lat.tune(yh.TUNE_HALF_LINKS, 1)
    s = src_field(o, rng, even_only=True)
    want = np.zeros_like(s)
    o.Hphi(mass, want, s.copy())
    fd = lat.field("f32")
    lat.Hphi(mass, fd, lat.field("f32").upload(s))
    err = rel(fd.download()[: o.VH], want[: o.VH])
    assert 1e-6 < err < 5e-3  # the FP16 links are in use, to their accuracy
    fb = lat.field("f64").upload(s)
    fx = lat.field("f64")
    it, upd = lat.cg_mixed(lat.inv_par(mass, prec, "Hphi_sq", sloppy="f32", max_iter=5000), fx, fb)
    r = true_res(o, MVM_HPHI_SQ, mass, fx.download(), s)
    print(f"half links: operator rel err {err:.2e}; mixed CG iterations={it} updates={upd} true residual^2={r:.3e}")
    assert r < prec
    lat.tune(yh.TUNE_HALF_LINKS, 0)
    lat.Hphi(mass, fd, lat.field("f32").upload(s))
    assert rel(fd.download()[: o.VH], want[: o.VH]) < 2e-6


@pytest.mark.parametrize("sloppy", ["f32", "f16"])
def test_cg_mixed_true_residual(sloppy):
    """Mixed-precision CG reaches the FP64 true residual 1e-10 (prec 1e-20),
    checked with the oracle's FP64 operator on the returned solution."""
    o, lat, rng = setup("su3_fund", dims=(8, 8, 8, 8), bc=BC_ANTIPERIODIC, csw=1.0, seed=32244)
    mass, prec = -0.75, 1e-20
    b = src_field(o, rng, even_only=True)
    fb = lat.field("f64").upload(b)
    fx = lat.field("f64")
    it, upd = lat.cg_mixed(lat.inv_par(mass, prec, "Hphi_sq", sloppy=sloppy, reliable_every=50, max_iter=5000, log=1), fx, fb)
    x = fx.download()
    r = true_res(o, MVM_HPHI_SQ, mass, x, b)
    print(f"mixed {sloppy}: iterations={it} updates={upd} true residual^2={r:.3e}")
    assert r < prec


def test_error_paths():
    o, lat, rng = setup("su3_fund")
    with pytest.raises((AssertionError, yh.YahmcError)):
        lat.field("f64").upload(np.zeros((o.V, 12)))  # wrong width is caught on the host
    f64 = lat.field("f64")
    f32 = lat.field("f32")
    with pytest.raises(yh.YahmcError):
        lat.Dphi_eo(f64, f32)  # precision mismatch
    with pytest.raises(yh.YahmcError):
        lat.cg_inverter(lat.inv_par(-0.5, 1e-20, "Hphi_sq", max_iter=16), f32, f32)
    b = src_field(o, rng, even_only=True)
    with pytest.raises(yh.YahmcError) as e:
        lat.cg_inverter(lat.inv_par(-0.5, 1e-30, "Hphi_sq", max_iter=16), lat.field("f64"), lat.field("f64").upload(b))
    assert e.value.code == -4
