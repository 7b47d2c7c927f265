"""Seeded synthetic inputs and dense textbook references used by the tests.

Everything here is written from the mathematics (gamma algebra, Wilson and
clover operators in dense form) so that it can pin the oracle and the HIP
path independently of either.
"""
import numpy as np

# Chiral gamma basis as the reference applies it per spinor component
# (library/spinor.cpp:191-264): (gamma_mu psi)_a = sum_b G[mu][a, b] psi_b.
GAMMA = np.zeros((4, 4, 4), dtype=complex)
GAMMA[0][0, 2] = GAMMA[0][1, 3] = GAMMA[0][2, 0] = GAMMA[0][3, 1] = -1
GAMMA[1][0, 3] = GAMMA[1][1, 2] = -1j
GAMMA[1][2, 1] = GAMMA[1][3, 0] = 1j
GAMMA[2][0, 3] = -1
GAMMA[2][1, 2] = 1
GAMMA[2][2, 1] = 1
GAMMA[2][3, 0] = -1
GAMMA[3][0, 2] = -1j
GAMMA[3][1, 3] = 1j
GAMMA[3][2, 0] = 1j
GAMMA[3][3, 1] = -1j
GAMMA5 = np.diag([1, 1, -1, -1]).astype(complex)


def gaussian_spinor(rng, V, dr, parity_sites=None):
    """Complex Gaussian spinor field, sigma = 1/sqrt(2) per real part, in the
    reference layout (V, 8*dr): re[dr][4] then im[dr][4]."""
    s = rng.normal(0.0, np.sqrt(0.5), size=(V, 8 * dr))
    if parity_sites is not None:
        mask = np.zeros(V, dtype=bool)
        mask[parity_sites] = True
        s[~mask] = 0.0
    return s


def spinor_to_complex(s, dr):
    """(V, 8*dr) reference layout -> complex array (V, 4 spins, dr colours)."""
    V = s.shape[0]
    re = s[:, : 4 * dr].reshape(V, dr, 4)
    im = s[:, 4 * dr:].reshape(V, dr, 4)
    return (re + 1j * im).transpose(0, 2, 1)


def complex_to_spinor(c):
    V, _, dr = c.shape
    t = c.transpose(0, 2, 1).reshape(V, 4 * dr)
    return np.concatenate([t.real, t.imag], axis=1).copy()


def mat_to_complex(m, n):
    """(K, 2*n*n) re-plane / im-plane row-major -> complex (K, n, n)."""
    return (m[:, : n * n] + 1j * m[:, n * n:]).reshape(-1, n, n)


def complex_to_mat(c):
    K, n, _ = c.shape
    f = c.reshape(K, n * n)
    return np.concatenate([f.real, f.imag], axis=1).copy()


def reunitarize(m):
    """Gram-Schmidt unitarisation of complex (K, N, N) matrices following
    the reference's row-by-row procedure (library/suNg.cpp:171-359): for
    N = 2 the second row is (-conj b, conj a), for N = 3 the third row is
    conj(r0 x r1); otherwise plain Gram-Schmidt followed by removal of the
    determinant phase from the last row."""
    m = m.copy()
    n = m.shape[-1]
    r0 = m[:, 0, :] / np.linalg.norm(m[:, 0, :], axis=1, keepdims=True)
    if n == 2:
        out = np.empty_like(m)
        out[:, 0] = r0
        out[:, 1, 0] = -np.conj(r0[:, 1])
        out[:, 1, 1] = np.conj(r0[:, 0])
        return out
    if n == 3:
        r1 = m[:, 1, :]
        r1 = r1 - np.sum(np.conj(r0) * r1, axis=1, keepdims=True) * r0
        r1 = r1 / np.linalg.norm(r1, axis=1, keepdims=True)
        r2 = np.conj(np.cross(r0, r1))
        return np.stack([r0, r1, r2], axis=1)
    rows = [r0]
    for i in range(1, n):
        r = m[:, i, :]
        for q in rows:
            r = r - np.sum(np.conj(q) * r, axis=1, keepdims=True) * q
        rows.append(r / np.linalg.norm(r, axis=1, keepdims=True))
    out = np.stack(rows, axis=1)
    det = np.linalg.det(out)
    out[:, -1, :] *= np.conj(det / np.abs(det))[:, None]
    return out


def random_gauge(rng, V, nc):
    """Seeded random SU(N) field, links (4V, 2*nc*nc) in the reference
    layout: Gram-Schmidt unitarisation of i.i.d. complex N(0,1) matrices."""
    z = rng.normal(size=(4 * V, nc, nc)) + 1j * rng.normal(size=(4 * V, nc, nc))
    return complex_to_mat(reunitarize(z))


def unit_gauge(V, nc):
    e = np.broadcast_to(np.eye(nc, dtype=complex), (4 * V, nc, nc)).copy()
    return complex_to_mat(e)


def eo_geometry(dims):
    """Single-process EO site order, id = (parity*V + lex)/2 with
    lex = ((t*X + x)*Y + y)*Z + z (library/geometry.cpp:756-878), and the
    periodic forward/backward neighbour tables (geometry.cpp:1161-1187)."""
    T, X, Y, Z = dims
    V = T * X * Y * Z
    t, x, y, z = np.meshgrid(np.arange(T), np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    c = np.stack([t.ravel(), x.ravel(), y.ravel(), z.ravel()], axis=1)

    def idx(cc):
        lex = ((cc[:, 0] * X + cc[:, 1]) * Y + cc[:, 2]) * Z + cc[:, 3]
        par = cc.sum(axis=1) % 2
        return (par * V + lex) // 2

    ids = idx(c)
    coords = np.zeros((V, 4), dtype=int)
    coords[ids] = c
    fw = np.zeros((V, 4), dtype=int)
    bk = np.zeros((V, 4), dtype=int)
    for mu in range(4):
        cf = c.copy()
        cf[:, mu] = (cf[:, mu] + 1) % dims[mu]
        cb = c.copy()
        cb[:, mu] = (cb[:, mu] - 1) % dims[mu]
        fw[ids, mu] = idx(cf)
        bk[ids, mu] = idx(cb)
    return coords, fw, bk


def dense_hop(orc, links_repr, psi, sign=1.0):
    """Dense-gamma Wilson hopping term on ALL sites:
        out(x) = sign * (-1/2) sum_mu [(1 - g_mu) U(x,mu) psi(x+mu)
                                      + (1 + g_mu) U(x-mu,mu)^+ psi(x-mu)]
    using the represented links as stored (BC signs included).  Written from
    the formula, not from the per-direction projections."""
    dr = orc.dr
    V = orc.V
    U = mat_to_complex(links_repr, dr).reshape(V, 4, dr, dr)
    P = spinor_to_complex(psi, dr)
    out = np.zeros_like(P)
    _, fw, bk = eo_geometry(orc.dims)
    one = np.eye(4)
    for mu in range(4):
        pm = one - GAMMA[mu]
        pp = one + GAMMA[mu]
        f = np.einsum("vij,vsj->vsi", U[:, mu], P[fw[:, mu]])
        f = np.einsum("ab,vbi->vai", pm, f)
        b = np.einsum("vji,vsj->vsi", np.conj(U[bk[:, mu], mu]), P[bk[:, mu]])
        b = np.einsum("ab,vbi->vai", pp, b)
        out += f + b
    return complex_to_spinor(-0.5 * sign * out)
