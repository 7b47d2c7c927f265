"""Hybrid Monte Carlo driven through the device kernels.

Monomials (library/monomials/gauge.cpp, hmc.cpp, list.cpp), the nested
integrators o2lf / o2mn / o4mn (library/integrator.cpp:6-96) and the
trajectory with its Metropolis step (library/update.cpp:284-339).  Level 0
is the innermost integrator and carries the gauge monomial; fermion
monomials sit on higher levels, whose field update runs the next lower
integrator and then re-represents the links (set_gauge_field_state,
library/repr.cpp:282-298).

The two-flavour monomial solves with CG on H^2 instead of the reference's
two BiCGstab solves on D (MonomialHMC::update, hmc.cpp:67-93): the force
needs Y = D^-1 g5 D^-1 pf = (H^2)^-1 g5 pf, and the action
|D^-1 pf|^2 = <g5 pf, (H^2)^-1 g5 pf>, both from the same solve (FP32
reliable-update CG reaching the FP64 true residual `prec`).

Random numbers (momenta, pseudofermion sources) are drawn on the host with
numpy and the reference's normalisation (N(0, 1/2) per real component,
include/random.h:15); the GSL generator itself is not reproduced."""
import numpy as np

from .md import cold_start  # noqa: F401  (re-exported for drivers)

O2MN = 0.1931833275037836                      # integrator.cpp:44
O4MN = (0.08398315262876693, 0.2539785108410595, 0.6822365335719091, -0.03230286765269967)  # :67-70


class GaugeMonomial:
    """Wilson plaquette action, or with c0 != 1 the Symanzik-improved action
    with rectangles weighted c1 = (1 - c0) / 8 (MonomialGauge, gauge.cpp)."""

    def __init__(self, beta, level=0, c0=1.0):
        if level != 0:
            raise ValueError("the gauge monomial must be on integrator level 0 (update.cpp)")
        self.beta = beta
        self.level = 0
        self.c0 = c0
        self.c1 = (1.0 - c0) / 8.0
        self.improved = abs(self.c1) > 1e-8  # gauge.cpp:11

    def reset(self, lat, rng):
        pass

    def action(self, lat):
        if self.improved:
            v1 = lat.avr_wilson(2, 1) + lat.avr_wilson(1, 2)
            return -self.beta * (self.c0 * lat.plaquette() + self.c1 * v1) * 6.0 * lat.gauge_volume
        return -self.beta * lat.plaquette() * 6.0 * lat.gauge_volume

    def update(self, lat, dt):
        if self.improved:
            lat.improved_gauge_force(dt, self.beta, self.c0, self.c1)
        else:
            lat.wilson_gauge_force(dt, self.beta)


class FermionMonomial:
    """Two flavours of EO Wilson-clover fermions (MonomialHMC, hmc.cpp)."""

    # reliable updates of the monomial's mixed CG: an FP64 update whenever
    # the sloppy residual has dropped by DELTA, and at the latest every
    # RELIABLE_EVERY iterations.  Each update is an FP64 H^2 (about twelve
    # 16-bit iterations at D_R = 6).  On the sextet trajectory DELTA = 0.01
    # gives the iterations and dH of 0.1 with half the updates (2.36 -> 2.12 s;
    # 0.001 costs extra iterations, 1e-4 breaks the 16-bit recurrence down);
    # on slowly converging systems the 16-bit recurrence drifts before it
    # gains two decades (32^4 fundamental: 1,344 iterations with no cap
    # against 783 with updates every 50), hence the cap.
    DELTA = 0.01
    RELIABLE_EVERY = 50
    # past solutions in the chronological (MRE) initial guess of the solves
    # (mre_par of MonomialHMC, hmc.cpp:16-17, 75-85).  Each costs an FP64
    # H^2 per guess; on the sextet trajectory one saves 9 % of the CG
    # iterations (2.12 -> 2.03 s), more cost more products than they save.
    MRE_PAST = 1

    def __init__(self, mass, dm=0.0, prec=1e-24, level=1, sloppy="f32", max_iter=100000, keep_pairs=0,
                 delta=DELTA, reliable_every=RELIABLE_EVERY, mre_past=MRE_PAST):
        if level == 0:
            raise ValueError("fermion monomials cannot be on integrator level 0 (update.cpp)")
        self.omass = mass
        self.smass = mass + dm
        self.prec = prec
        self.level = level
        self.sloppy = sloppy
        self.max_iter = max_iter
        self.delta = delta
        self.reliable_every = reliable_every
        self.mre_past = mre_past
        self.mre = None
        self.pf = None
        self.la = None
        self.iterations = []
        # (source, solution) pairs of the monomial's own solves, the training
        # data of the learned preconditioner (the last keep_pairs solves, each
        # solved with the links of its MD step)
        self.keep_pairs = keep_pairs
        self.pairs = []

    def _fields(self, lat):
        if self.pf is None:
            self.pf = lat.field("f64")
            self.g5pf = lat.field("f64")
            self.y = lat.field("f64")
            self.eta = lat.field("f64")
            if self.mre_past > 0:
                self.mre = lat.mre(self.mre_past, "Hphi_sq", self.smass)

    def reset(self, lat, rng):
        """pf = D eta, eta Gaussian on the even sites; the first action call
        returns |eta|^2 - 2 log det D_oo (hmc.cpp:27-40)."""
        self._fields(lat)
        eta = np.zeros((lat.V, 8 * lat.dr))
        eta[: lat.V // 2] = rng.normal(0.0, np.sqrt(0.5), size=(lat.V // 2, 8 * lat.dr))
        self.eta.upload(eta)
        self.la = float(np.sum(eta * eta))
        if lat.csw != 0.0:
            self.la -= 2.0 * lat.clover_logdet(self.omass)
        lat.Dphi(self.smass, self.pf, self.eta)
        self.g5pf.copy_from(self.pf)
        lat.spinor_g5(self.g5pf)
        if self.mre is not None:
            self.mre.reset()  # hmc.cpp:36-37

    def _solve(self, lat):
        """y = (H^2)^-1 g5 pf from the MRE extrapolation of the past
        solutions, which the new one joins (hmc.cpp:75-85).  The action at
        the end of a trajectory solves the system of the last force step
        (no link update in between), so its guess is already the solution
        and its CG stops after the first true-residual check."""
        if self.mre is not None:
            self.mre.guess(self.y, self.g5pf)
        else:
            self.y.zero()
        par = lat.inv_par(self.smass, self.prec, "Hphi_sq", sloppy=self.sloppy, max_iter=self.max_iter,
                          delta=self.delta, reliable_every=self.reliable_every)
        it, _ = lat.cg_mixed(par, self.y, self.g5pf)
        self.iterations.append(it)
        if self.mre is not None:
            self.mre.store(self.y)
        if self.keep_pairs:
            b, x = lat.field("f64"), lat.field("f64")
            b.copy_from(self.g5pf)
            x.copy_from(self.y)
            self.pairs.append((b, x))
            while len(self.pairs) > self.keep_pairs:
                for f in self.pairs.pop(0):
                    f.free()
        return self.y

    def action(self, lat):
        if self.la is not None:
            v, self.la = self.la, None
            return v
        y = self._solve(lat)
        v = lat.spinor_product_re(self.g5pf, y)
        if lat.csw != 0.0:
            v -= 2.0 * lat.clover_logdet(self.omass)
        return v

    def update(self, lat, dt):
        y = self._solve(lat)
        lat.wilson_fermion_force_q(dt, self.smass, 1.0, y)
        if lat.csw != 0.0:
            lat.clover_logdet_force(self.omass, 2.0)


class HasenbuschMonomial:
    """Mass-preconditioning ratio det(H(m)^2) / det(H(m + dm)^2) for EO
    clover fermions (MonomialHB, monomials/hasenbusch.cpp:8-96, the
    CLOVER_TERM_EO branch).  Inverses of Hphi are taken as
    H^-1 b = (H^2)^-1 H b with the mixed-precision CG on H^2."""

    def __init__(self, mass, dm, prec=1e-24, level=1, sloppy="f32", max_iter=100000, delta=FermionMonomial.DELTA,
                 reliable_every=FermionMonomial.RELIABLE_EVERY):
        if level == 0:
            raise ValueError("fermion monomials cannot be on integrator level 0 (update.cpp)")
        self.omass = mass
        self.smass = mass + dm
        self.prec = prec
        self.level = level
        self.sloppy = sloppy
        self.max_iter = max_iter
        self.delta = delta
        self.reliable_every = reliable_every
        self.pf = None
        self.la = -1.0
        self.iterations = []

    def _fields(self, lat):
        if self.pf is None:
            self.pf, self.a, self.b, self.eta = (lat.field("f64") for _ in range(4))

    def _h2_inverse(self, lat, mass, dst, src):
        dst.zero()
        par = lat.inv_par(mass, self.prec, "Hphi_sq", sloppy=self.sloppy, max_iter=self.max_iter, delta=self.delta,
                          reliable_every=self.reliable_every)
        it, _ = lat.cg_mixed(par, dst, src)
        self.iterations.append(it)

    def _h_inverse(self, lat, mass, dst, src):
        """dst = Hphi(mass)^-1 src = (H^2)^-1 H src"""
        lat.Hphi(mass, self.b, src)
        self._h2_inverse(lat, mass, dst, self.b)

    def reset(self, lat, rng):
        """pf = H(m + dm)^-1 H(m) eta (hasenbusch.cpp:29-42)."""
        self._fields(lat)
        eta = np.zeros((lat.V, 8 * lat.dr))
        eta[: lat.V // 2] = rng.normal(0.0, np.sqrt(0.5), size=(lat.V // 2, 8 * lat.dr))
        self.eta.upload(eta)
        self.la = float(np.sum(eta * eta))
        lat.Hphi(self.omass, self.a, self.eta)
        self._h_inverse(lat, self.smass, self.pf, self.a)

    def action(self, lat):
        """|H(m)^-1 H(m + dm) pf|^2 (hasenbusch.cpp:44-68)."""
        if self.la > 0:
            v, self.la = self.la, -1.0
            return v
        lat.Hphi(self.smass, self.eta, self.pf)
        self._h_inverse(lat, self.omass, self.a, self.eta)
        return lat.spinor_sqnorm(self.a)

    def update(self, lat, dt):
        """hasenbusch.cpp:70-83: Y = H(m)^-2 H(m + dm) pf; force_q(m, +1, Y)
        and force(m + dm, -1, pf, Y)."""
        lat.Hphi(self.smass, self.eta, self.pf)
        self._h2_inverse(lat, self.omass, self.a, self.eta)
        lat.wilson_fermion_force_q(dt, self.omass, 1.0, self.a)
        lat.wilson_fermion_force(dt, self.smass, -1.0, self.pf, self.a)


class RHMCMonomial:
    """One flavour of EO Wilson-clover fermions by the rational
    approximation (MonomialRHMC, monomials/rhmc.cpp:10-106):
    det(H^2)^(1/2) with the heatbath pf = r_pf(H^2) eta, r_pf ~ x^(1/4),
    the action <pf, r_fc(H^2) pf> and the force
    sum_n res_n force_q((H^2 + pole_n)^-1 pf), r_fc ~ x^(-1/2), both
    approximations found on the evlimits range widened by 5 % either side
    (rhmc.cpp:30-63).  The clover term adds -log det D_oo once (one flavour).
    Rational approximations are generated rather than looked up
    (yahmc_amd.rational_find); the shifted solves are one multishift CG."""

    def __init__(self, mass, rap_prec=1e-10, inv_prec=1e-24, level=1, max_iter=0):
        if level == 0:
            raise ValueError("fermion monomials cannot be on integrator level 0 (update.cpp)")
        self.mass = mass
        self.rap_prec = rap_prec
        self.inv_prec = inv_prec
        self.level = level
        self.max_iter = max_iter
        self.pf = None
        self.out = []
        self.la = 0.0
        self.rpf = self.rfc = None
        self.iterations = []
        self.limits = None

    def _par(self, lat):
        return lat.inv_par(self.mass, self.inv_prec, "Hphi_sq", max_iter=self.max_iter)

    def reset(self, lat, rng):
        from . import rational_find
        if self.pf is None:
            self.pf, self.tmp = lat.field("f64"), lat.field("f64")
        lo, hi = evlimits(lat, self.mass, rng)
        lo, hi = 0.95 * lo, 1.05 * hi
        self.limits = (lo, hi)
        self.rpf = rational_find(1, 4, self.rap_prec, lo, hi)
        self.rfc = rational_find(-1, 2, self.rap_prec, lo, hi)
        self.tmp.upload(_gaussian_even(lat, rng))
        self.la = lat.spinor_sqnorm(self.tmp)
        self.iterations.append(lat.rational_calc(self.rpf, self._par(lat), self.pf, self.tmp))
        if lat.csw != 0.0:
            self.la -= lat.clover_logdet(self.mass)
        while len(self.out) < self.rfc.order:
            self.out.append(lat.field("f64"))

    def action(self, lat):
        if self.la != 0:
            v, self.la = self.la, 0.0
            return v
        self.iterations.append(lat.rational_calc(self.rfc, self._par(lat), self.tmp, self.pf))
        v = lat.spinor_product_re(self.pf, self.tmp)
        if lat.csw != 0.0:
            v -= lat.clover_logdet(self.mass)
        return v

    def update(self, lat, dt):
        n = self.rfc.order
        outs = self.out[:n]
        self.iterations.append(lat.cg_mshift(self._par(lat), np.array(self.rfc.pole[:n]), outs, self.pf))
        for k in range(n):
            lat.wilson_fermion_force_q(dt, self.mass, self.rfc.residue[k], outs[k])
        if lat.csw != 0.0:
            lat.clover_logdet_force(self.mass, 1.0)


def largest_eigenvalue(lat, mass, rng, mvm="Hphi_sq", tol=1e-3):
    """Power iteration (library/evlimit.cpp:52-84)."""
    b = lat.field("f64").upload(_gaussian_even(lat, rng))
    t = lat.field("f64")
    op = getattr(lat, mvm)
    eig, old, rel = 0.0, 0.0, 1.0
    while rel > tol:
        op(mass, t, b)
        nrm = np.sqrt(lat.spinor_sqnorm(t))
        old, eig = eig, nrm
        b.zero()
        lat.spinor_mulr_add_assign(b, 1.0 / nrm, t)
        rel = abs((eig - old) / eig)
    return eig


def smallest_eigenvalue(lat, mass, rng, mvm="Hphi_sq", tol=1e-3):
    """Inverse power iteration with auto_inverter at precision 1e-6
    (library/evlimit.cpp:9-50)."""
    b = lat.field("f64").upload(_gaussian_even(lat, rng))
    t = lat.field("f64")
    par = lat.inv_par(mass, 1e-6, mvm)
    eig, old, rel = 0.0, 0.0, 1.0
    while rel > tol:
        lat.auto_inverter(par, t, b)
        nrm = np.sqrt(lat.spinor_sqnorm(t))
        old, eig = eig, nrm
        b.zero()
        lat.spinor_mulr_add_assign(b, 1.0 / nrm, t)
        rel = abs((eig - old) / eig)
    return 1.0 / eig


def evlimits(lat, mass, rng, mvm="Hphi_sq"):
    """Spectral range [min, max] of mvm (evlimit.cpp:86-91)."""
    return smallest_eigenvalue(lat, mass, rng, mvm), largest_eigenvalue(lat, mass, rng, mvm)


def _gaussian_even(lat, rng):
    s = np.zeros((lat.V, 8 * lat.dr))
    s[: lat.V // 2] = rng.normal(0.0, np.sqrt(0.5), size=(lat.V // 2, 8 * lat.dr))
    return s


class HMC:
    """integrators: [(type, steps)] from level 0 (innermost) upwards, type in
    {"o2lf", "o2mn", "o4mn"}; monomials carry their level."""

    def __init__(self, lat, monomials, integrators):
        self.lat = lat
        self.monomials = list(monomials)
        self.integrators = list(integrators)
        self.dirty = False
        for m in self.monomials:
            if m.level >= len(self.integrators):
                raise ValueError(f"monomial level {m.level} has no integrator")

    # ---- update_momenta / update_fields (integrator.cpp:6-17, list.cpp) --
    def _momenta(self, dt, level):
        lat = self.lat
        if level:
            lat.begin_fermion_force()
        for m in self.monomials:
            if m.level == level:
                m.update(lat, dt)
        if level:
            lat.end_fermion_force(dt)

    def _fields(self, dt, level):
        if level == 0:
            self.lat.update_links(dt)
            self.dirty = True
        else:
            self._integrate(dt, level - 1)
            if self.dirty:
                self.lat.represent_gauge_field()
                self.dirty = False

    def _integrate(self, length, level):
        kind, steps = self.integrators[level]
        dt = length / steps
        M, F = self._momenta, self._fields
        if kind == "o2lf":
            for n in range(steps):
                M(dt / 2 if n == 0 else dt, level)
                F(dt, level)
            M(dt / 2, level)
        elif kind == "o2mn":
            a = O2MN
            for n in range(steps):
                M(a * dt if n == 0 else 2 * a * dt, level)
                F(dt / 2, level)
                M((1 - 2 * a) * dt, level)
                F(dt / 2, level)
            M(a * dt, level)
        elif kind == "o4mn":
            a, b, c, d = O4MN
            for n in range(steps):
                M(a * dt if n == 0 else 2 * a * dt, level)
                F(b * dt, level)
                M(c * dt, level)
                F(d * dt, level)
                M((0.5 - a - c) * dt, level)
                F((1 - 2 * (b + d)) * dt, level)
                M((0.5 - a - c) * dt, level)
                F(d * dt, level)
                M(c * dt, level)
                F(b * dt, level)
            M(a * dt, level)
        else:
            raise ValueError(f"unknown integrator {kind}")

    def hamiltonian(self):
        """TF/2 sum P^2 + sum of monomial actions (update.cpp:201-226)."""
        return self.lat.momenta_energy() + sum(m.action(self.lat) for m in self.monomials)

    def trajectory(self, rng, length=1.0, metropolis=True):
        """One trajectory (update.cpp:284-339): refresh momenta and
        pseudofermions, integrate, Metropolis on dH, then reunitarise or
        restore the links."""
        lat = self.lat
        U0 = lat.get_gauge()
        ng = lat.nc * lat.nc - 1
        lat.set_momenta(rng.normal(0.0, np.sqrt(0.5), size=(4 * lat.V, ng)))
        for m in self.monomials:
            m.reset(lat, rng)
        h0 = self.hamiltonian()
        self._integrate(length, len(self.integrators) - 1)
        if self.dirty:
            lat.represent_gauge_field()
            self.dirty = False
        dh = self.hamiltonian() - h0
        accepted = True
        if metropolis:
            accepted = bool(dh < 0 or rng.uniform() < np.exp(-dh))
        if accepted:
            lat.reunitarize()
        else:
            lat.set_gauge(U0)
        lat.represent_gauge_field()
        return {"dH": float(dh), "accepted": accepted, "plaquette": lat.plaquette()}
