"""Rational approximations (CPU, host code of libyahmc.so).

The reference picks x^(-1/2) approximations from a table
(library/rational.cpp:56-132, include/approx.h).  The project generates
Zolotarev's optimal approximation instead; it is pinned here by
  - the elliptic functions against scipy.special (ellipj, ellipk),
  - optimality: at the order and interval of reference table entries the
    generated approximation is at least as accurate as the table's stated
    precision (tests/golden/rational_table_sample.json, extracted from
    include/approx.h by tests/golden/make_rational_fixture.py),
  - the partial-fraction residues reproducing the product form,
  - the maximum relative error on a dense grid for the 12-pole x^(-1/2) on
    [1e-4, 8] of the RHMC workload.
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


def test_at_least_as_accurate_as_reference_table():
    data = json.load(open(os.path.join(HERE, "golden", "rational_table_sample.json")))
    assert len(data["entries"]) >= 4
    for e in data["entries"]:
        r = yh.rational_zolotarev(e["order"], -1, 2, e["min"], e["max"])
        err = max_rel_err(r, -0.5, e["min"], e["max"])
        # Zolotarev is optimal: equal to the table up to the FP64 evaluation
        # noise of a degree-n product (~n ulp)
        assert err <= e["prec"] * 1.002 + 1e-15 * e["order"], (e, err)


def test_find_picks_lowest_order():
    r = yh.rational_find(-1, 2, 1e-10, 1e-4, 8.0)
    assert r.prec < 1e-10
    lower = yh.rational_zolotarev(r.order - 1, -1, 2, 1e-4, 8.0)
    assert lower.prec >= 1e-10
