"""Device-resident multishift CG, batched over right-hand sides, and the
rational approximation applied through it (rational_calc,
library/rational.cpp:134-166), checked against the oracle and against exact
matrix functions of a dense H^2 on a tiny lattice."""
import numpy as np
import pytest

from fields import gaussian_spinor, random_gauge, spinor_to_complex
from oracle import BC_ANTIPERIODIC, MVM_HPHI_SQ, Oracle

pytestmark = pytest.mark.gpu
yh = pytest.importorskip("yahmc_amd")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def setup(model, nc, rep, dims, seed=3, csw=1.0):
    o = Oracle(model).init(dims, bc=BC_ANTIPERIODIC, csw=csw)
    rng = np.random.default_rng(seed)
    links = random_gauge(rng, o.V, nc)
    o.set_gauge(links)
    o.represent()
    lat = yh.Lattice(dims, nc=nc, repr=rep, bc="antiperiodic", csw=csw)
    lat.set_gauge(links)
    return o, lat, rng


@pytest.mark.parametrize("model,nc,rep,nrhs", [("su3_fund", 3, "fundamental", 3), ("su2_adj", 2, "adjoint", 4)])
def test_batched_multishift_matches_single_rhs_oracle(model, nc, rep, nrhs):
    """Batched solves against the oracle's cg_mshift per right-hand side.
    With the multi-RHS hop (batches of 4 and 2, the rest one by one) the
    <p, H^2 p> reductions group their sums differently, so the iterates agree
    with the single-RHS solver to rounding; without it (TUNE_MULTI_RHS 0)
    they are the single-RHS iterates bit for bit."""
    print(f"\n[case] batched mshift {model} nrhs={nrhs}", flush=True)
    o, lat, rng = setup(model, nc, rep, (4, 4, 4, 4))
    mass, prec = -0.2, 1e-24
    shifts = [0.5, 0.1, 0.05, 1.05, 0.01]
    srcs = []
    for j in range(nrhs):
        b = gaussian_spinor(rng, o.V, o.dr)
        b[o.VH:] = 0
        srcs.append(b)
    fsrc = [lat.field("f64").upload(b) for b in srcs]
    par = lat.inv_par(mass, prec, "Hphi_sq")
    batched = {}
    for multi in (1, 0):
        lat.tune(yh.TUNE_MULTI_RHS, multi)
        dst = [[lat.field("f64") for _ in shifts] for _ in srcs]
        its = lat.cg_mshift_batched(par, shifts, dst, fsrc)
        batched[multi] = (its, [[f.download() for f in row] for row in dst])
    lat.tune(yh.TUNE_MULTI_RHS, 1)
    for j, b in enumerate(srcs):
        it_or, xs_or = o.cg_mshift(MVM_HPHI_SQ, mass, prec, shifts, b)
        single = [lat.field("f64") for _ in shifts]
        it1 = lat.cg_mshift(par, shifts, single, fsrc[j])
        assert batched[0][0][j] == it1
        assert abs(batched[1][0][j] - it1) <= 1 and abs(it1 - it_or) <= 2
        for s, x_mr, x_1, f1, xo in zip(shifts, batched[1][1][j], batched[0][1][j], single, xs_or):
            x_single = f1.download()
            assert np.array_equal(x_1, x_single)
            assert rel(x_mr[: o.VH], x_single[: o.VH]) < 1e-12
            assert rel(x_mr[: o.VH], xo[: o.VH]) < 1e-10
            ax = np.zeros_like(x_mr)
            o.Hphi_sq(mass, ax, x_mr.copy())
            r = b[: o.VH] - ax[: o.VH] - s * x_mr[: o.VH]
            assert np.sum(r * r) / np.sum(b[: o.VH] ** 2) < 1e-23


def test_rational_calc_inverse_square_root():
    """r(H^2) b with a Zolotarev x^(-1/2) equals (H^2)^(-1/2) b from the
    eigen-decomposition of the dense H^2 (tiny lattice), to the stated
    precision of the approximation."""
    dims = (2, 2, 2, 4)
    o, lat, rng = setup("su3_fund", 3, "fundamental", dims, seed=9)
    mass = -0.2
    VH, n = o.VH, o.VH * 12
    cols = []
    for k in range(n):
        e = np.zeros((o.V, 4, 3), dtype=complex)
        e[:VH].reshape(-1)[k] = 1.0
        from fields import complex_to_spinor
        s = complex_to_spinor(e)
        h = np.zeros_like(s)
        o.Hphi_sq(mass, h, s)
        cols.append(spinor_to_complex(h, 3)[:VH].reshape(-1))
    H2 = np.stack(cols, axis=1)
    assert np.abs(H2 - H2.conj().T).max() < 1e-12
    w, U = np.linalg.eigh(H2)
    lo, hi = w.min() * 0.9, w.max() * 1.1
    rat = yh.rational_find(-1, 2, 1e-11, lo, hi)
    b = gaussian_spinor(rng, o.V, 3)
    b[VH:] = 0
    bc = spinor_to_complex(b, 3)[:VH].reshape(-1)
    want = U @ ((U.conj().T @ bc) / np.sqrt(w))
    fb = lat.field("f64").upload(b)
    fd = lat.field("f64")
    it = lat.rational_calc(rat, lat.inv_par(mass, 1e-26, "Hphi_sq"), fd, fb)
    got = spinor_to_complex(fd.download(), 3)[:VH].reshape(-1)
    print(f"order={rat.order} prec={rat.prec:.2e} iterations={it} range=[{lo:.3e}, {hi:.3e}]")
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 5 * rat.prec + 1e-12


def test_graph_replay_matches_eager():
    """The multishift chunk captured in a HIP graph and replayed gives the
    iterates of the eager launches bit for bit, over two solves in a row
    with other device work and host synchronisation in between."""
    print("\n[case] mshift graph replay", flush=True)
    o, lat, rng = setup("su3_fund", 3, "fundamental", (4, 4, 4, 8), seed=21)
    mass, prec = -0.2, 1e-24
    shifts = [0.7, 0.05, 0.002]
    srcs = []
    for _ in range(2):
        b = gaussian_spinor(rng, o.V, o.dr)
        b[o.VH:] = 0
        srcs.append(lat.field("f64").upload(b))
    results = {}
    for graphs in (0, 1):
        lat.tune(yh.TUNE_GRAPHS, graphs)
        runs = []
        for rep in range(2):
            dst = [[lat.field("f64") for _ in shifts] for _ in srcs]
            its = lat.cg_mshift_batched(lat.inv_par(mass, prec, "Hphi_sq"), shifts, dst, srcs)
            # eager device work and a synchronisation between the replays
            t = lat.field("f64")
            lat.Hphi_sq(mass, t, srcs[0])
            lat.synchronize()
            runs.append((its, [[f.download() for f in row] for row in dst]))
        assert runs[0][0] == runs[1][0]
        for r0, r1 in zip(runs[0][1], runs[1][1]):
            for a, b in zip(r0, r1):
                assert np.array_equal(a, b)
        results[graphs] = runs[0]
        print(f"graphs={graphs}: iterations {runs[0][0]}")
    assert max(results[1][0]) > 16  # more than one chunk: the graph was replayed
    assert results[0][0] == results[1][0]
    for r0, r1 in zip(results[0][1], results[1][1]):
        for a, b in zip(r0, r1):
            assert np.array_equal(a, b)


def test_autotune_keeps_results():
    """The autotuner only changes the workgroup -> XCD order: the hop output
    is bit-identical before and after, and every chunk it picks is valid."""
    print("\n[case] autotune", flush=True)
    o, lat, rng = setup("su3_fund", 3, "fundamental", (8, 8, 8, 8), seed=23)
    s = gaussian_spinor(rng, o.V, 3)
    s[: o.VH] = 0
    fs = lat.field("f64").upload(s)
    d0 = lat.field("f64")
    lat.Dphi_eo(d0, fs)
    before = d0.download()
    chosen = lat.autotune(reps=5)
    print(f"chosen XCD chunks {chosen}")
    assert ("f64", 18) in chosen and ("f32", 12) in chosen
    assert all(v >= 0 for v in chosen.values())
    d1 = lat.field("f64")
    lat.Dphi_eo(d1, fs)
    assert np.array_equal(d1.download(), before)
