"""Seeded synthetic inputs (host side, numpy): random SU(N) gauge fields in
the reference's gfield_fund layout and Gaussian spinor sources.

gauge_field_set_random (library/gaugefield.cpp:181-197) draws each link from
the reference's GSL stream; that stream is not reproducible here, so this
module uses numpy's PCG64 with the same construction: an i.i.d. complex
Gaussian N x N matrix unitarised row by row (library/suNg.cpp:171-359), which
for N = 3 leaves det = 1.
"""
import numpy as np


def _reunitarize(m):
    n = m.shape[-1]
    r0 = m[:, 0, :] / np.linalg.norm(m[:, 0, :], axis=1, keepdims=True)
    if n == 2:
        out = np.empty_like(m)
        out[:, 0] = r0
        out[:, 1, 0] = -np.conj(r0[:, 1])
        out[:, 1, 1] = np.conj(r0[:, 0])
        return out
    r1 = m[:, 1, :] - np.sum(np.conj(r0) * m[:, 1, :], axis=1, keepdims=True) * r0
    r1 /= np.linalg.norm(r1, axis=1, keepdims=True)
    if n == 3:
        return np.stack([r0, r1, np.conj(np.cross(r0, r1))], axis=1)
    rows = [r0, r1]
    for i in range(2, n):
        r = m[:, i, :]
        for q in rows:
            r = r - np.sum(np.conj(q) * r, axis=1, keepdims=True) * q
        rows.append(r / np.linalg.norm(r, axis=1, keepdims=True))
    out = np.stack(rows, axis=1)
    det = np.linalg.det(out)
    out[:, -1, :] *= np.conj(det / np.abs(det))[:, None]
    return out


def random_gauge(dims, nc=3, seed=32244, chunk=1 << 20):
    """(4V, 2*nc*nc) float64 links, suNg re-plane then im-plane."""
    V = int(np.prod(dims))
    rng = np.random.default_rng(seed)
    out = np.empty((4 * V, 2 * nc * nc))
    for s in range(0, 4 * V, chunk):
        n = min(chunk, 4 * V - s)
        z = rng.normal(size=(n, nc, nc)) + 1j * rng.normal(size=(n, nc, nc))
        u = _reunitarize(z).reshape(n, nc * nc)
        out[s:s + n, : nc * nc] = u.real
        out[s:s + n, nc * nc:] = u.imag
    return out


def gaussian_source(dims, dr=3, seed=32245, parity="even"):
    """(V, 8*dr) complex Gaussian spinor, sigma = 1/sqrt 2 per real part
    (include/random.h:15), zero on the other parity."""
    V = int(np.prod(dims))
    rng = np.random.default_rng(seed)
    s = np.zeros((V, 8 * dr))
    if parity == "even":
        s[: V // 2] = rng.normal(0.0, np.sqrt(0.5), size=(V // 2, 8 * dr))
    else:
        s[:] = rng.normal(0.0, np.sqrt(0.5), size=(V, 8 * dr))
    return s
