"""Pure-gauge molecular dynamics driven through the device kernels:
the 2nd-order minimal-norm integrator (library/integrator.cpp:42-63) with
the Wilson gauge force on one level, the trajectory of library/update.cpp:
284-339 (momentum refresh, Hamiltonian, integration, Metropolis, reunitarise
or restore), and the smoothing used to prepare thermalised-like fields.

Momenta are drawn on the host (numpy PCG64) with the reference's
normalisation: each algebra component N(0, 1/2), so that the kinetic term
TF/2 sum P^2 (TF = 2, update.cpp:201-216) is the Gaussian weight."""
import numpy as np

O2MN_LAMBDA = 0.1931833275037836  # integrator.cpp:44


def o2mn_gauge(lat, beta, length, steps):
    """o2mn_multilevel with the gauge monomial alone (integrator.cpp:42-63)."""
    a = O2MN_LAMBDA
    dt = length / steps
    for n in range(steps):
        lat.wilson_gauge_force(a * dt if n == 0 else 2 * a * dt, beta)
        lat.update_links(dt / 2)
        lat.wilson_gauge_force((1 - 2 * a) * dt, beta)
        lat.update_links(dt / 2)
    lat.wilson_gauge_force(a * dt, beta)


def hamiltonian(lat, beta):
    """TF/2 sum P^2 - beta * plaq * 6 volume (update.cpp:201-226 with the
    Wilson gauge action; volume = V with periodic gauge boundaries)."""
    return lat.momenta_energy() - beta * lat.plaquette() * 6.0 * lat.gauge_volume


def gauge_trajectory(lat, beta, rng, length=1.0, steps=10, metropolis=True):
    """One pure-gauge HMC trajectory (update.cpp:284-339); returns dH,
    acceptance and the plaquette afterwards.  Without Metropolis every
    trajectory is kept (MD-only evolution)."""
    ng = lat.nc * lat.nc - 1
    U0 = lat.get_gauge() if metropolis else None
    lat.set_momenta(rng.normal(0.0, np.sqrt(0.5), size=(4 * lat.V, ng)))
    h0 = hamiltonian(lat, beta)
    o2mn_gauge(lat, beta, length, steps)
    dh = hamiltonian(lat, beta) - h0
    accepted = True
    if metropolis:
        accepted = bool(dh < 0 or rng.uniform() < np.exp(-dh))
    if accepted:
        lat.reunitarize()
    else:
        lat.set_gauge(U0)
    lat.represent_gauge_field()
    return {"dH": float(dh), "accepted": accepted, "plaquette": lat.plaquette()}


def cold_start(lat):
    """All links = 1 (gaugefield.cpp cold start), then the gauge-field
    boundary conditions (gaugefield.cpp:193)."""
    n = lat.nc
    links = np.zeros((4 * lat.V, 2 * n * n))
    for i in range(n):
        links[:, i * n + i] = 1.0
    lat.set_gauge(links)
    lat.apply_bcs_on_gauge_field()
    lat.represent_gauge_field()


def smooth(lat, beta=5.6, trajectories=10, length=1.0, steps=10, seed=32246, metropolis=False, cold=True):
    """Prepare a thermalised-like field: cold start, then `trajectories`
    pure-gauge trajectories.  From a cold start at beta = 5.6 the first
    trajectories carry dH of O(V) and Metropolis would reject them, so the
    default keeps every trajectory (MD-only); the log records dH and the
    plaquette of each."""
    rng = np.random.default_rng(seed)
    if cold:
        cold_start(lat)
    return [gauge_trajectory(lat, beta, rng, length, steps, metropolis) for _ in range(trajectories)]
