"""Test-side restatement of the spectroscopy pieces for checking: the
gamma5 * Gamma tables of observables/mesons.cpp:10-26 as dense matrices,
the EO propagator of observables/propagator.cpp:75-122 on the oracle, and
the contraction formulas of mesons.cpp:55-110 with numpy."""
import numpy as np

from fields import spinor_to_complex
from oracle import MVM_DPHI

TABLES = {
    "id": ([0, 1, 2, 3], [1, 1, -1, -1]), "g1": ([3, 2, 1, 0], [-1j] * 4), "g2": ([3, 2, 1, 0], [-1, 1, -1, 1]),
    "g3": ([2, 3, 0, 1], [-1j, 1j, -1j, 1j]), "g5": ([0, 1, 2, 3], [1, 1, 1, 1]),
    "g5g1": ([3, 2, 1, 0], [-1j, -1j, 1j, 1j]), "g5g2": ([3, 2, 1, 0], [-1, 1, 1, -1]),
    "g5g3": ([2, 3, 0, 1], [-1j, 1j, 1j, -1j]), "g0g1": ([1, 0, 3, 2], [-1j] * 4),
    "g0g2": ([1, 0, 3, 2], [-1, 1, -1, 1]), "g0g3": ([0, 1, 2, 3], [-1j, 1j, -1j, 1j]),
    "g0g5": ([2, 3, 0, 1], [1, 1, 1, 1]), "g0g5g1": ([1, 0, 3, 2], [1j, 1j, -1j, -1j]),
    "g0g5g2": ([1, 0, 3, 2], [1, -1, -1, 1]), "g0g5g3": ([0, 1, 2, 3], [1j, -1j, -1j, 1j]),
}
CHANNELS = [(n, n) for n in TABLES] + [("g0g5", "g5")]


def table_matrix(name):
    col, val = TABLES[name]
    M = np.zeros((4, 4), dtype=complex)
    for a in range(4):
        M[a, col[a]] = val[a]
    return M


def oracle_propagator(o, mass, prec, b):
    """calc_propagator_eo with the clover term (propagator.cpp:86-122)."""
    b = b.copy()
    t = np.zeros_like(b)
    u = np.zeros_like(b)
    o.Dphi_oo_inv(mass, u, b)            # u_o = D_oo^-1 b_o
    o.Dphi_eo(t, u)                      # t_e = D_eo u_o
    src = np.zeros_like(b)
    src[: o.VH] = b[: o.VH] - t[: o.VH]
    x = np.zeros_like(b)
    o.lib.or_zero_auto_scratch()
    o.auto_inverter(MVM_DPHI, mass, prec, x, src)
    t2 = np.zeros_like(b)
    o.Dphi_oe(t2, x)                     # t2_o = D_oe x_e
    r = np.zeros_like(b)
    r[o.VH:] = b[o.VH:] - t2[o.VH:]
    xo = np.zeros_like(b)
    o.Dphi_oo_inv(mass, xo, r)
    x[o.VH:] = xo[o.VH:]
    return x


def contract(props, dr, T, t_of_site, tau, point):
    """16 x T correlators, dense-matrix form of mesons.cpp:55-110."""
    S = np.array([spinor_to_complex(p, dr) for p in props])  # [n][site][spin][colour]
    out = np.zeros((16, T))
    for k, (n1, n2) in enumerate(CHANNELS):
        M1, M2 = table_matrix(n1), table_matrix(n2)
        if point:
            Sn = S.reshape(4, dr, S.shape[1], 4, dr)      # [c][x][site][spin][y]
            val = np.einsum("ab,cd,cxsby,dxsay->s", M1, M2, Sn, np.conj(Sn)).real
        else:
            val = np.einsum("ab,cd,csbx,dsax->s", M1, M2, S, np.conj(S)).real
        for site, v in enumerate(val):
            out[k, (t_of_site[site] - tau) % T] += v
    return out
