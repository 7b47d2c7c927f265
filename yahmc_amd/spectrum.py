"""Spectroscopy over the device kernels: point and stochastic-wall sources
(observables/propagator.cpp:22-73), propagators (calc_propagator, through
yh_calc_propagator) and the 16 meson correlators (observables/mesons.cpp:
174-299, contractions by yh_meson_contract).

Channel names and print signs follow print_correlator (mesons.cpp:281-297).
The reference prints the arrays of g0g5g2 under the label "g0g5g1" and of
g0g5g3 under "g0g5g2"; here every channel is returned under its own name.
Random numbers (time slice, Z2 noise) come from the caller's numpy
generator; the reference draws them from GSL."""
import numpy as np

CHANNELS = ("id", "g1", "g2", "g3", "g5", "g5g1", "g5g2", "g5g3", "g0g1", "g0g2", "g0g3", "g0g5", "g0g5g1", "g0g5g2",
            "g0g5g3", "g5_g0g5")
SIGNS = (-1, -1, -1, -1, +1, -1, -1, -1, -1, -1, -1, +1, +1, +1, +1, +1)
WALL, POINT = 1, 2


def site_index(dims, t, x, y, z):
    """EO index of (t, x, y, z) (geometry.cpp:756-878)."""
    T, X, Y, Z = dims
    V = T * X * Y * Z
    lex = ((t * X + x) * Y + y) * Z + z
    return ((t + x + y + z) & 1) * (V // 2) + lex // 2


def point_source(lat, n, tau):
    """create_point_source (propagator.cpp:61-73): unit component
    (spin n // D_R, colour n % D_R) at (tau, 0, 0, 0)."""
    s = np.zeros((lat.V, 8 * lat.dr))
    spin, col = divmod(n, lat.dr)
    s[site_index(lat.dims, tau, 0, 0, 0), col * 4 + spin] = 1.0
    return s


def wall_source(lat, n, tau, prev, rng):
    """create_semwall_source (propagator.cpp:22-59): Z2 noise (+-1/sqrt 2,
    real and imaginary parts) on time slice tau in spin 0 for n = 0; for
    n > 0 the previous source with its spin n - 1 moved to spin n."""
    dr = lat.dr
    if n == 0:
        s = np.zeros((lat.V, 8 * dr))
        T, X, Y, Z = lat.dims
        for x in range(X):
            for y in range(Y):
                for z in range(Z):
                    i = site_index(lat.dims, tau, x, y, z)
                    z2 = np.where(rng.uniform(size=2 * dr) < 0.5, 1.0, -1.0) / np.sqrt(2.0)
                    for c in range(dr):
                        s[i, c * 4] = z2[2 * c]
                        s[i, 4 * dr + c * 4] = z2[2 * c + 1]
        return s
    s = prev.copy()
    for c in range(dr):
        for part in (0, 4 * dr):
            s[:, part + c * 4 + n] = s[:, part + c * 4 + n - 1]
            s[:, part + c * 4 + n - 1] = 0.0
    return s


def meson_measure(lat, mass, prec, hits, method, rng, taus=None):
    """meson_measure (mesons.cpp:219-299): for each hit a random source time
    slice (or taus[k]), 4 (wall) or 4 D_R (point) propagators, contractions;
    returns {channel: correlator} with print_correlator's sign and
    normalisation 1 / (norm * hits), norm = spatial volume (wall) or 1."""
    T = lat.dims[0]
    ninv = 4 if method == WALL else 4 * lat.dr
    norm = lat.V // T if method == WALL else 1
    acc = np.zeros((16, T))
    props = [lat.field("f64") for _ in range(ninv)]
    src_f = lat.field("f64")
    par = lat.inv_par(mass, prec, "Dphi")
    iters = []
    for h in range(hits):
        tau = int(rng.uniform() * T) if taus is None else int(taus[h])
        src = None
        for n in range(ninv):
            src = wall_source(lat, n, tau, src, rng) if method == WALL else point_source(lat, n, tau)
            src_f.upload(src)
            iters.append(lat.calc_propagator(par, props[n], src_f))
        acc += lat.meson_contract(method, props, tau)
    out = {name: sign * acc[k] / (norm * hits) for k, (name, sign) in enumerate(zip(CHANNELS, SIGNS))}
    out["_iterations"] = iters
    return out
