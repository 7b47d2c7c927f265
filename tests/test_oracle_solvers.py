"""Pins the oracle's Krylov solvers with the reference's own self-checks
(checks/inverters.cpp:10-191): true residual below the requested squared
relative precision, mass -0.2, shifts {0.5, 0.1, 0.05, 1.05}."""
import numpy as np
import pytest

from fields import gaussian_spinor, random_gauge
from oracle import MVM_DPHI, MVM_HPHI, MVM_HPHI_SQ, Oracle

PREC = 1e-28
MASS = -0.2


@pytest.fixture(scope="module")
def orc():
    o = Oracle("su3_fund").init((4, 4, 4, 4), csw=1.0)
    rng = np.random.default_rng(5)
    o.set_gauge(random_gauge(rng, o.V, 3))
    o.represent()
    src = gaussian_spinor(rng, o.V, 3)
    src[o.VH:] = 0
    return o, src


def _true_res(o, mvm, mass, x, b):
    ax = np.zeros_like(x)
    o.mvm(mvm, mass, ax, x.copy())
    r = b[: o.VH] - ax[: o.VH]
    return np.sum(r * r) / np.sum(b[: o.VH] ** 2)


def test_cg_hphi_sq(orc):
    o, b = orc
    x = np.zeros_like(b)
    it, res = o.cg(MVM_HPHI_SQ, MASS, PREC, x, b)
    assert it > 0 and res < PREC
    assert _true_res(o, MVM_HPHI_SQ, MASS, x, b) < 1e-26
    # restart from the solution: already converged -> 0 iterations (cg.cpp:165-169)
    it2, _ = o.cg(MVM_HPHI_SQ, MASS, 1e-20, x, b)
    assert it2 == 0


def test_bicgstab_dphi(orc):
    o, b = orc
    x = np.zeros_like(b)
    it, res = o.bicgstab(MVM_DPHI, MASS, PREC, x, b)
    assert it > 0 and it % 2 == 0 and res < PREC
    assert _true_res(o, MVM_DPHI, MASS, x, b) < 1e-26


def test_auto_inverter_hphi_sq_equals_cg(orc):
    o, b = orc
    x1 = np.zeros_like(b)
    o.cg(MVM_HPHI_SQ, MASS, PREC, x1, b)
    x2 = np.zeros_like(b)
    src = b.copy()
    o.auto_inverter(MVM_HPHI_SQ, MASS, PREC, x2, src)
    assert np.array_equal(src[: o.VH], b[: o.VH])
    assert np.abs(x1[: o.VH] - x2[: o.VH]).max() < 1e-10 * np.abs(x1).max()


def test_cg_mshift(orc):
    o, b = orc
    shifts = np.array([0.5, 0.1, 0.05, 1.05])
    it, xs = o.cg_mshift(MVM_HPHI_SQ, MASS, PREC, shifts, b)
    assert it > 0
    for s, x in zip(shifts, xs):
        ax = np.zeros_like(x)
        o.Hphi_sq(MASS, ax, x.copy())
        r = b[: o.VH] - ax[: o.VH] - s * x[: o.VH]
        assert np.sum(r * r) / np.sum(b[: o.VH] ** 2) < 1e-26


def make_solver_lattice(seed=7):
    o = Oracle("su3_fund").init((4, 4, 4, 4), csw=1.0)
    rng = np.random.default_rng(seed)
    o.set_gauge(random_gauge(rng, o.V, 3))
    o.represent()
    return o, rng


def test_minres_true_residual():
    """minres_inverter (library/minres.cpp) on Hphi and Hphi_sq reaches the
    requested true residual (checks/inverters.cpp tolerance style)."""
    o, rng = make_solver_lattice()
    mass = -0.2
    for mvm in (MVM_HPHI, MVM_HPHI_SQ):
        b = gaussian_spinor(rng, o.V, o.dr)
        b[o.VH:] = 0
        x = np.zeros_like(b)
        it, res = o.minres(mvm, mass, 1e-24, x, b.copy())
        ax = np.zeros_like(b)
        o.mvm(mvm, mass, ax, x.copy())
        r = b[: o.VH] - ax[: o.VH]
        assert it > 0 and np.sum(r * r) / np.sum(b[: o.VH] ** 2) < 1e-23
    # an exact guess returns 0 iterations
    x2 = x.copy()
    it2, _ = o.minres(MVM_HPHI_SQ, mass, 1e-20, x2, b.copy())
    assert it2 == 0


def test_mre_guess_is_galerkin_solution():
    """The MRE guess x0 from past solutions s_k satisfies the Galerkin
    condition <s_j, b - M x0> = 0 for every stored s_j, and reproduces b's
    solution exactly when b = M (sum c_k s_k)."""
    from oracle import OracleMRE, spinor_to_act
    o, rng = make_solver_lattice(seed=8)
    mass = -0.2
    mre = OracleMRE(o, 5, MVM_DPHI, mass)
    past = []
    for k in range(4):
        v = gaussian_spinor(rng, o.V, o.dr)
        v[o.VH:] = 0
        mre.store(v)
        past.append(v)
    b = gaussian_spinor(rng, o.V, o.dr)
    b[o.VH:] = 0
    x0 = mre.guess(b)
    mx = np.zeros_like(x0)
    o.mvm(MVM_DPHI, mass, mx, x0.copy())
    r = spinor_to_act(o, b - mx)[: o.VH].reshape(-1)
    for v in past:
        vc = spinor_to_act(o, v)[: o.VH].reshape(-1)
        assert abs(np.vdot(vc, r)) < 1e-10 * np.linalg.norm(vc) * np.linalg.norm(r)
    coef = rng.normal(size=4) + 1j * rng.normal(size=4)
    want = sum(c * spinor_to_act(o, v) for c, v in zip(coef, past))
    from oracle import act_to_spinor
    ws = act_to_spinor(o, want)
    b2 = np.zeros_like(ws)
    o.mvm(MVM_DPHI, mass, b2, ws.copy())
    b2[o.VH:] = 0
    x1 = mre.guess(b2)
    assert np.abs(x1[: o.VH] - ws[: o.VH]).max() < 1e-10 * np.abs(ws).max()
