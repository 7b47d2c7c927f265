"""Pins the oracle's representation map (library/repr.cpp) and generators.

The oracle writes R(U) as the action of U on the representation basis; here
it is checked against (1) group properties (checks/group.cpp:5-98) and (2)
the reference's explicit per-entry expansion (repr.cpp:115-261), which is a
different formula for the same numbers.
"""
import numpy as np
import pytest

from fields import complex_to_mat, mat_to_complex, reunitarize
from oracle import Oracle

SQ2 = np.sqrt(2.0)
REPR_TRACE = {0: lambda n: 2, 1: lambda n: 4 * n, 2: lambda n: 2 * (n + 2), 3: lambda n: 2 * (n - 2)}


def rand_su(rng, n, k=1):
    return reunitarize(rng.normal(size=(k, n, n)) + 1j * rng.normal(size=(k, n, n)))


def R(o, u):
    return mat_to_complex(o.represent_link(complex_to_mat(u[None]).ravel())[None], o.dr)[0]


@pytest.mark.parametrize("model", ["su3_fund", "su2_fund", "su2_adj", "su3_adj", "su3_sym", "su3_asym", "su4_asym"])
def test_homomorphism_and_orthogonality(model):
    o = Oracle(model).init((2, 2, 2, 2))
    rng = np.random.default_rng(3)
    for _ in range(4):
        u, v = rand_su(rng, o.nc, 2)
        assert np.abs(R(o, u @ v) - R(o, u) @ R(o, v)).max() < 1e-14
        assert np.abs(R(o, u.conj().T) - R(o, u).conj().T).max() < 1e-14
    f, r = o.generators()
    ng = o.nc ** 2 - 1
    F = mat_to_complex(f, o.nc)
    Rg = mat_to_complex(r, o.dr)
    for a in range(ng):
        for b in range(ng):
            assert abs(np.trace(F[a] @ F[b]) + 2 * (a == b)) < 1e-14
            assert abs(np.trace(Rg[a] @ Rg[b]) + REPR_TRACE[o.repr](o.nc) * (a == b)) < 1e-13
    # generators are the derivative of the map: R(exp(eps T)) ~ 1 + eps T^R
    from scipy.linalg import expm
    eps = 1e-6
    for a in range(ng):
        d = (R(o, expm(eps * F[a])) - R(o, expm(-eps * F[a]))) / (2 * eps)
        assert np.abs(d - Rg[a]).max() < 1e-8


def _ref_sym(U, n):
    """repr.cpp:130-158, transcribed entry by entry."""
    pairs = [(i, j) for i in range(n) for j in range(i + 1)]
    out = np.zeros((len(pairs), len(pairs)), dtype=complex)
    for a, (i, j) in enumerate(pairs):
        for b, (k, l) in enumerate(pairs):
            if i == j and k == l:
                out[a, b] = U[i, k] * U[j, l]
            elif i == j or k == l:
                out[a, b] = SQ2 * U[i, k] * U[j, l]
            else:
                out[a, b] = U[i, k] * U[j, l] + U[i, l] * U[j, k]
    return out


def _ref_asym(U, n):
    """repr.cpp:115-128."""
    pairs = [(i, j) for i in range(1, n) for j in range(i)]
    out = np.zeros((len(pairs), len(pairs)), dtype=complex)
    for a, (i, j) in enumerate(pairs):
        for b, (k, l) in enumerate(pairs):
            out[a, b] = U[i, k] * U[j, l] - U[i, l] * U[j, k]
    return out


def _ref_adj(U, n):
    """repr.cpp:160-261 (real adjoint matrix), transcribed."""
    dim = n * n - 1
    # diagonal basis entries of the reference's adjoint basis (repr.cpp:54-61):
    # basis[m] = (-i/sqrt2) suNg(e_m); on diagonal generators this is real
    diag = {}
    m = 0
    for i in range(n):
        for j in range(n):
            if i != j:
                m += 1
            elif i > 0:
                v = np.sqrt(2.0 / (i * (i + 1)))
                d = np.zeros(n)
                d[:i] = v / SQ2
                d[i] = -i * v / SQ2
                diag[m] = d
                m += 1
    re, im = U.real, U.imag
    out = np.zeros((dim, dim))
    nn = 0
    for i in range(n):
        for j in range(n):
            if i == j and i == 0:
                continue
            mm = 0
            for a in range(n):
                for b in range(n):
                    if a == b and a == 0:
                        continue
                    if i > j:
                        if a > b:
                            val = re[j, b] * re[i, a] + re[j, a] * re[i, b] + im[j, b] * im[i, a] + im[j, a] * im[i, b]
                        elif a < b:
                            val = re[i, b] * im[j, a] + re[j, b] * im[i, a] - re[i, a] * im[j, b] - re[j, a] * im[i, b]
                        else:
                            val = sum(SQ2 * diag[mm][k] * (re[j, k] * re[i, k] + im[j, k] * im[i, k]) for k in range(a + 1))
                    elif i < j:
                        if a > b:
                            val = re[i, b] * im[j, a] - re[j, b] * im[i, a] + re[i, a] * im[j, b] - re[j, a] * im[i, b]
                        elif a < b:
                            val = re[j, b] * re[i, a] - re[j, a] * re[i, b] + im[j, b] * im[i, a] - im[j, a] * im[i, b]
                        else:
                            val = sum(SQ2 * diag[mm][k] * (re[i, k] * im[j, k] - re[j, k] * im[i, k]) for k in range(a + 1))
                    else:
                        if a > b:
                            val = sum(SQ2 * diag[nn][k] * (re[k, b] * re[k, a] + im[k, b] * im[k, a]) for k in range(i + 1))
                        elif a < b:
                            val = sum(SQ2 * diag[nn][k] * (re[k, b] * im[k, a] - im[k, b] * re[k, a]) for k in range(i + 1))
                        else:
                            val = sum(diag[nn][x] * diag[mm][y] * (re[x, y] ** 2 + im[x, y] ** 2)
                                      for x in range(i + 1) for y in range(a + 1))
                    out[nn, mm] = val
                    mm += 1
            nn += 1
    return out


@pytest.mark.parametrize("model,fn", [("su3_sym", _ref_sym), ("su3_asym", _ref_asym), ("su4_asym", _ref_asym),
                                      ("su2_adj", _ref_adj), ("su3_adj", _ref_adj)])
def test_matches_reference_entrywise_expansion(model, fn):
    o = Oracle(model).init((2, 2, 2, 2))
    rng = np.random.default_rng(11)
    for u in rand_su(rng, o.nc, 3):
        assert np.abs(R(o, u) - fn(u, o.nc)).max() < 1e-14
