"""Pins the oracle's parallel-transport convolution layer and the network
built from it (the learned preconditioner; no reference counterpart):
  - with Wilson-hop spin weights a one-channel layer IS the hop term of the
    (already pinned) Dirac operator, which fixes the path / link direction
    conventions;
  - the adjoint layer satisfies <g, L h> = <L^+ g, h>;
  - the layer is gauge covariant;
  - the weight gradient equals the (exact, linear) directional derivative
    Re<g, L(dW) h>;
  - the network loss gradient matches central finite differences in FP64;
  - the preconditioner A A^+ is Hermitian and positive."""
import numpy as np
import pytest

from fields import complex_to_spinor, gaussian_spinor, random_gauge, spinor_to_complex
from oracle import BC_ANTIPERIODIC, Oracle, OracleNet, act_to_spinor, spinor_to_act
from test_oracle_operator import _gauge_transform


def make(model="su3_fund", nc=3, dims=(4, 4, 4, 4), seed=1):
    o = Oracle(model).init(dims, bc=BC_ANTIPERIODIC, csw=1.0)
    rng = np.random.default_rng(seed)
    o.set_gauge(random_gauge(rng, o.V, nc))
    o.represent()
    return o, rng


def cgauss(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def hop_weights(o):
    """W_{+mu} = -1/2 (1 - gamma_mu), W_{-mu} = -1/2 (1 + gamma_mu); the
    oracle gives gamma5 (1 - gamma_mu), and gamma5 is read off its g5."""
    e = np.eye(4)
    g5 = np.zeros((4, 4), dtype=complex)
    for s in range(4):
        c = np.zeros((o.V, 4, o.dr), dtype=complex)
        c[0, s, 0] = 1.0
        sp = complex_to_spinor(c)
        o.g5(sp)
        g5[:, s] = spinor_to_complex(sp, o.dr)[0, :, 0]
    W = np.zeros((9, 1, 1, 4, 4), dtype=complex)
    for mu in range(4):
        one_minus = g5 @ o.spin_matrix(0, mu)
        W[1 + mu, 0, 0] = -0.5 * one_minus
        W[5 + mu, 0, 0] = -0.5 * (2 * e - one_minus)
    return W


@pytest.mark.parametrize("model,nc", [("su3_fund", 3), ("su2_adj", 2), ("su3_sym", 3)])
def test_hop_weights_reproduce_dirac_hop(model, nc):
    o, rng = make(model, nc)
    W = hop_weights(o)
    s = gaussian_spinor(rng, o.V, o.dr)
    s[: o.VH] = 0  # odd input
    want = np.zeros_like(s)
    o.Dphi_eo(want, s.copy())
    got = act_to_spinor(o, o.ptc_layer(W, spinor_to_act(o, s)))
    assert np.abs(got[: o.VH] - want[: o.VH]).max() < 1e-13


def test_adjoint_identity():
    o, rng = make(seed=4)
    W = cgauss(rng, 9, 3, 2, 4, 4)
    h = cgauss(rng, o.V, 2, 4, 3)
    g = cgauss(rng, o.V, 3, 4, 3)
    lhs = np.vdot(g, o.ptc_layer(W, h))
    rhs = np.vdot(o.ptc_layer(W, g, adjoint=True), h)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_gauge_covariance():
    o, rng = make(seed=6)
    W = cgauss(rng, 9, 2, 2, 4, 4)
    h = cgauss(rng, o.V, 2, 4, 3)
    out = o.ptc_layer(W, h)
    g, lp = _gauge_transform(o, rng, o.get_gauge())
    o.set_gauge(lp)
    o.represent()
    hg = np.einsum("vab,vcsb->vcsa", g, h)
    got = o.ptc_layer(W, hg)
    assert np.abs(got - np.einsum("vab,vcsb->vcsa", g, out)).max() < 1e-12


def test_wgrad_is_directional_derivative():
    o, rng = make(seed=9)
    cin, cout = 2, 3
    h = cgauss(rng, o.V, cin, 4, 3)
    g = cgauss(rng, o.V, cout, 4, 3)
    G = o.ptc_wgrad(g, h)
    for trial in range(3):
        dW = cgauss(rng, 9, cout, cin, 4, 4)
        want = np.vdot(g, o.ptc_layer(dW, h)).real
        got = np.sum(np.conj(G) * dW).real
        assert abs(got - want) < 1e-11 * np.sum(np.abs(G) * np.abs(dW))
    # one weight at a time, real and imaginary directions
    for k in [(0, 0, 0, 1, 2), (3, 2, 1, 0, 0), (8, 1, 1, 3, 3)]:
        for ph in (1.0, 1j):
            dW = np.zeros((9, cout, cin, 4, 4), dtype=complex)
            dW[k] = ph
            want = np.vdot(g, o.ptc_layer(dW, h)).real
            got = (G[k].real if ph == 1.0 else G[k].imag)
            assert abs(got - want) < 1e-11 * max(1.0, abs(want))


def random_net(o, rng, chans=(1, 4, 4, 1), scale=0.3):
    Ws = []
    for ci, co in zip(chans[:-1], chans[1:]):
        W = scale * cgauss(rng, 9, co, ci, 4, 4) / np.sqrt(9 * ci * 4)
        Ws.append(W)
    return OracleNet(o, Ws)


def test_network_loss_gradient_finite_differences():
    o, rng = make(dims=(4, 4, 2, 4), seed=12)
    net = random_net(o, rng)
    b = spinor_to_act(o, gaussian_spinor(rng, o.V, 3))
    x = spinor_to_act(o, gaussian_spinor(rng, o.V, 3))
    b[o.VH:] = 0
    x[o.VH:] = 0
    loss, grad = net.loss_and_grad(b, x)
    dWs = [cgauss(rng, *W.shape) for W in net.W]
    want = sum(np.sum(np.conj(g) * d).real for g, d in zip(grad, dWs))
    W0 = [W.copy() for W in net.W]

    def L(t):
        net.W = [W + t * d for W, d in zip(W0, dWs)]
        v, _ = net.loss_and_grad(b, x)
        return v

    h = 1e-4
    fd = (8 * (L(h) - L(-h)) - (L(2 * h) - L(-2 * h))) / (12 * h)
    net.W = W0
    assert abs(fd - want) < 1e-7 * abs(want), (fd, want)


def test_preconditioner_hermitian_positive():
    o, rng = make(seed=13)
    net = random_net(o, rng)
    a = spinor_to_act(o, gaussian_spinor(rng, o.V, 3))
    b = spinor_to_act(o, gaussian_spinor(rng, o.V, 3))
    a[o.VH:] = 0
    b[o.VH:] = 0
    lhs = np.vdot(a, net.precondition(b))
    rhs = np.vdot(net.precondition(a), b)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    assert np.vdot(a, net.precondition(a)).real > 0
    assert np.abs(net.precondition(a)[o.VH:]).max() == 0
