"""Pins the oracle's reunitarize (library/suNg.cpp:171-359) on matrices that
are far from the group, for N = 2, 3 and 4, by what the procedure fixes
independently of how it is coded:
  - the result is unitary (U U^+ = 1) with det U = 1;
  - row 0 is input row 0 normalised, and row k lies in the span of input
    rows 0..k (Gram-Schmidt), which with the two properties above fixes the
    result up to a phase per row;
  - the phases: rows 0..N-2 are the Gram-Schmidt rows themselves (the
    R factor of a QR decomposition of the transposed input has a positive
    real diagonal), the last row carries exp(-i arg det);
  - it is idempotent on SU(N).
The reference is computed with numpy's QR (LAPACK Householder), a different
algorithm from the oracle's row-by-row procedure, so a wrong conjugate in
the projection, a sign error in the N = 2 row or a missing determinant
phase for N = 4 each fails a check."""
import numpy as np
import pytest

from fields import complex_to_mat, mat_to_complex
from oracle import BC_PERIODIC, Oracle

MODELS = [("su2_fund", 2), ("su3_fund", 3), ("su4_asym", 4)]


def reference_reunitarize(a):
    """Gram-Schmidt of the rows by QR of a^H, phases fixed as above."""
    n = a.shape[-1]
    q, r = np.linalg.qr(a.conj().T)
    ph = np.diag(r) / np.abs(np.diag(r))
    q = q * ph[None, :]                 # positive real diagonal of R
    u = q.conj().T                      # rows = orthonormalised input rows
    d = np.linalg.det(u)
    u[n - 1] *= np.conj(d) / abs(d)
    return u


def oracle_reunitarize(model, mats):
    o = Oracle(model).init((2, 2, 2, 2), bc=BC_PERIODIC, csw=0.0)
    assert mats.shape[0] == 4 * o.V
    o.set_gauge(complex_to_mat(mats))
    o.reunitarize()
    return mat_to_complex(o.get_gauge(), o.nc)


@pytest.mark.parametrize("model,nc", MODELS)
def test_reunitarize_off_group(model, nc):
    rng = np.random.default_rng(100 + nc)
    a = rng.normal(size=(64, nc, nc)) + 1j * rng.normal(size=(64, nc, nc))
    a[::3] *= 7.5                      # scale does not matter
    u = oracle_reunitarize(model, a)
    eye = np.eye(nc)
    for k in range(64):
        assert np.abs(u[k] @ u[k].conj().T - eye).max() < 1e-14
        assert abs(np.linalg.det(u[k]) - 1.0) < 1e-14
        # row 0 parallel to input row 0, with a positive real factor
        c = np.vdot(u[k][0], a[k][0])
        assert abs(c.imag) < 1e-12 * abs(c) and c.real > 0
        # row j in the span of input rows 0..j
        for j in range(1, nc):
            basis = a[k][: j + 1].T
            coef, *_ = np.linalg.lstsq(basis, u[k][j], rcond=None)
            assert np.abs(basis @ coef - u[k][j]).max() < 1e-13
        assert np.abs(u[k] - reference_reunitarize(a[k])).max() < 1e-13


@pytest.mark.parametrize("model,nc", MODELS)
def test_reunitarize_idempotent_on_group(model, nc):
    rng = np.random.default_rng(200 + nc)
    a = rng.normal(size=(64, nc, nc)) + 1j * rng.normal(size=(64, nc, nc))
    su = np.array([reference_reunitarize(x) for x in a])
    u = oracle_reunitarize(model, su)
    assert np.abs(u - su).max() < 1e-14


@pytest.mark.parametrize("model,nc", MODELS)
def test_reunitarize_near_group(model, nc):
    """A small perturbation of an SU(N) matrix moves back to the group and
    stays close (the use inside update_links, update.cpp:228-247)."""
    rng = np.random.default_rng(300 + nc)
    a = rng.normal(size=(64, nc, nc)) + 1j * rng.normal(size=(64, nc, nc))
    su = np.array([reference_reunitarize(x) for x in a])
    eps = 1e-6 * (rng.normal(size=su.shape) + 1j * rng.normal(size=su.shape))
    u = oracle_reunitarize(model, su + eps)
    assert np.abs(u - su).max() < 1e-5
    for k in range(64):
        assert abs(np.linalg.det(u[k]) - 1.0) < 1e-14
        assert np.abs(u[k] - reference_reunitarize(su[k] + eps[k])).max() < 1e-13
