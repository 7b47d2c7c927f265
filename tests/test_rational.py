"""Rational approximations (CPU, host code of libyahmc.so).

The reference picks x^(-1/2) and x^(-1/4) approximations from a table
(library/rational.cpp:56-132, include/approx.h).  The project generates
them: Zolotarev's closed form for x^(+-1/2) and the Remez exchange for
other powers.  Pinned here by
  - the elliptic functions against scipy.special (ellipj, ellipk),
  - the reference table itself (tests/golden/rational_table_sample.json,
    extracted from include/approx.h by tests/golden/make_rational_fixture.py):
    at the order and interval of each sampled entry the generated
    approximation has the table's precision, and for two entries the same
    a0, roots and poles (the minimax approximation is unique),
  - the equioscillation theorem: the error of the Remez result alternates
    in sign at 2n + 2 extrema of equal size (so, by de la Vallee Poussin's
    bound, no approximation of the order does better),
  - the partial-fraction residues reproducing the product form,
  - the maximum relative error on a dense grid (long double) for the
    12-pole x^(-1/2) on [1e-4, 8] of the RHMC workload.
"""
import json
import os

import numpy as np
import pytest
from scipy.special import ellipj, ellipkm1

yh = pytest.importorskip("yahmc_amd")
HERE = os.path.dirname(os.path.abspath(__file__))


def test_elliptic_functions_against_scipy():
    L = yh.lib()
    for kc in (0.9, 0.3, 1e-2, 1e-4, 3.5e-3):
        # ellipkm1 takes 1 - m = kc^2 directly (no cancellation near m = 1)
        K = ellipkm1(kc * kc)
        assert abs(L.yh_ellip_k(kc) - K) < 1e-13 * K
    for kc in (0.9, 0.3, 0.05):
        m = 1.0 - kc * kc
        for u in (0.1, 0.7, 1.3, 0.95 * ellipkm1(kc * kc)):
            sn = ellipj(u, m)[0]
            assert abs(L.yh_jacobi_sn(u, kc) - sn) < 1e-12


def max_rel_err(r, power, lo, hi, n=200001):
    x = np.exp(np.linspace(np.log(lo), np.log(hi), n))
    v = np.array([r(xx) for xx in x[::50]])
    return np.max(np.abs(v / x[::50] ** power - 1))


def test_twelve_poles_on_rhmc_interval():
    r = yh.rational_zolotarev(12, -1, 2, 1e-4, 8.0)
    err = max_rel_err(r, -0.5, 1e-4, 8.0)
    assert err < 1.2 * r.prec and r.prec < 1e-6
    assert all(r.pole[i] > 0 for i in range(12))
    # partial fractions reproduce the product form
    xs = np.array([1e-4, 3e-3, 0.1, 2.0, 8.0])
    prod = np.array([r.a0 * np.prod([(x + r.root[i]) / (x + r.pole[i]) for i in range(12)]) for x in xs])
    pf = np.array([r(x) for x in xs])
    assert np.allclose(pf, prod, rtol=1e-10)


def test_positive_power_swaps_roots_and_poles():
    rm = yh.rational_zolotarev(10, -1, 2, 1e-3, 10.0)
    rp = yh.rational_zolotarev(10, 1, 2, 1e-3, 10.0)
    assert max_rel_err(rp, 0.5, 1e-3, 10.0) < 1.2 * rm.prec
    for x in (1e-3, 0.5, 10.0):
        assert abs(rp(x) * rm(x) - 1) < 1e-12


def ld_rel_err(r, power, lo, hi, n=20001):
    """max |r(x) / x^power - 1| on a log grid, product form in long double"""
    L = np.longdouble
    x = np.exp(np.linspace(np.log(L(lo)), np.log(L(hi)), n))
    v = np.full_like(x, L(r.a0))
    for i in range(r.order):
        v *= (x + L(r.root[i])) / (x + L(r.pole[i]))
    return v / x ** L(power) - 1


def table():
    return json.load(open(os.path.join(HERE, "golden", "rational_table_sample.json")))


def test_generated_precision_equals_reference_table():
    data = table()
    assert len(data["entries"]) >= 40 and {e["d"] for e in data["entries"]} == {2, 4}
    for e in data["entries"]:
        r = yh.rational_remez(e["order"], e["n"], e["d"], e["min"], e["max"])
        # both are the minimax approximation: equal errors, up to what long
        # double resolves at the 1e-15 level
        tol = 1e-4 if e["prec"] > 1e-13 else 2e-2
        assert abs(r.prec / e["prec"] - 1) < tol, (e, r.prec)
        # independent check on a dense grid; rounding the 2n + 1 coefficients
        # to doubles moves r(x) by up to (2n + 1) * 1.1e-16 relative (the
        # table's stored coefficients have the same limit)
        err = np.max(np.abs(ld_rel_err(r, e["n"] / e["d"], e["min"], e["max"])))
        assert err <= e["prec"] * 1.001 + 1.2e-16 * (2 * e["order"] + 1), (e, err)
        if e["d"] == 2 and e["prec"] > 1e-13:
            z = yh.rational_zolotarev(e["order"], e["n"], e["d"], e["min"], e["max"])
            assert abs(z.prec / e["prec"] - 1) < 1e-3, (e, z.prec)


@pytest.mark.parametrize("which", [0, 1])
def test_coefficients_equal_reference_table(which):
    c = table()["coefficients"][which]
    r = yh.rational_remez(c["order"], c["n"], c["d"], c["min"], c["max"])
    assert abs(r.a0 / c["a0"] - 1) < 1e-7
    got_r, got_p = np.sort(np.array(r.root[: r.order])), np.sort(np.array(r.pole[: r.order]))
    assert np.max(np.abs(got_r / np.sort(c["root"]) - 1)) < 1e-6
    assert np.max(np.abs(got_p / np.sort(c["pole"]) - 1)) < 1e-6


@pytest.mark.parametrize("order,n,d,lo,hi", [(12, -1, 4, 1e-4, 8.0), (8, -3, 4, 1e-3, 10.0), (16, -1, 3, 1e-5, 64.0)])
def test_remez_equioscillates(order, n, d, lo, hi):
    r = yh.rational_remez(order, n, d, lo, hi)
    e = ld_rel_err(r, n / d, lo, hi, n=200001)
    E = np.max(np.abs(e))
    assert abs(E / r.prec - 1) < 1e-6
    # extrema of the error: one per sign run; all 2n + 2 of them of size E
    sign = np.sign(e)
    runs = np.split(np.abs(e), np.flatnonzero(np.diff(sign)) + 1)
    peaks = np.array([x.max() for x in runs])
    assert len(peaks) == 2 * order + 2
    assert peaks.min() > E * (1 - 1e-6)


def test_quarter_powers_and_find():
    r = yh.rational_find(1, 4, 1e-10, 1e-4, 8.0)
    assert r.prec < 1e-10
    assert yh.rational_remez(r.order - 1, 1, 4, 1e-4, 8.0).prec >= 1e-10
    rm = yh.rational_remez(r.order, -1, 4, 1e-4, 8.0)
    for x in (1e-4, 0.3, 8.0):
        # positive power = reciprocal of the negative one (rational.cpp:111-122)
        assert abs(r(x) * rm(x) - 1) < 1e-13
    assert np.max(np.abs(ld_rel_err(r, 0.25, 1e-4, 8.0))) < 1.001 * r.prec
    with pytest.raises(yh.YahmcError):
        yh.rational_remez(10, -5, 4, 1e-4, 8.0)


def test_find_picks_lowest_order():
    r = yh.rational_find(-1, 2, 1e-10, 1e-4, 8.0)
    assert r.prec < 1e-10
    lower = yh.rational_zolotarev(r.order - 1, -1, 2, 1e-4, 8.0)
    assert lower.prec >= 1e-10
