"""Benchmark: EO Wilson-clover Dslash bandwidth and CG time-to-solution on
one MI355X (32^4, SU(3) fundamental clover, c_sw = 1, m = -0.75,
antiperiodic time).

A "step" is one hop D_eo over the V/2 = 524,288 even sites of the 32^4
lattice with all inputs resident in HBM.  The headline hop is the
reference's own Dslash arithmetic: FP64 spinors and full 18-real links
(library/dirac.cpp:227-704).  `value` is the achieved HBM bandwidth of that
kernel under the compulsory-traffic model (each input spinor, each link and
each output spinor moves once per hop), measured with HIP events on the
stream the kernel runs on.  The line also carries the reduced-precision hops
(FP32 and FP16 storage, 12-real links), the CG time-to-solution to an FP64 true residual of 1e-10
(plain FP64 CG; mixed FP32 / FP16 reliable-update CG; flexible CG with the
learned gauge-covariant preconditioner, trained on the device on (source,
solution) pairs first), the roofline object and a CPU baseline (the oracle
on the host cores).

The gauge field is prepared as the 32^4 workload specifies: seeded cold
start, then 10 pure-gauge o2mn trajectories (tau = 1, 10 steps) at
beta = 5.6, every trajectory kept (MD-only: from the cold start the first
trajectories have dH of O(V) and Metropolis would reject them).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DIMS = (32, 32, 32, 32)
MASS = -0.75
CSW = 1.0
# compulsory bytes per output site of one hop, D_R = 3:
#   1 input spinor + 8 links + 1 output spinor
BYTES_PER_SITE = {
    ("f64", 0): 192 + 8 * 144 + 192,   # 1536
    ("f64", 1): 192 + 8 * 96 + 192,    # 1152
    ("f32", 1): 96 + 8 * 48 + 96,      # 576
    ("f16", 1): 52 + 8 * 24 + 52,      # 296 (48 B half spinor + 4 B scale)
    ("f32", 2): 96 + 8 * 24 + 96,      # 384 (FP32 spinors, FP16 12-real links)
}
# naive no-reuse model (every neighbour spinor and link read for every
# output site): 8 spinors + 8 links + 1 output
BYTES_PER_SITE_NAIVE = {
    ("f64", 0): 8 * 192 + 8 * 144 + 192,   # 2880
    ("f64", 1): 8 * 192 + 8 * 96 + 192,    # 2496
    ("f32", 1): 8 * 96 + 8 * 48 + 96,      # 1248
    ("f16", 1): 8 * 52 + 8 * 24 + 52,      # 660
    ("f32", 2): 8 * 96 + 8 * 24 + 96,      # 1056
}
FLOP_PER_SITE = 1320
# HBM-side bytes per launch of each hop variant, from two rocprofv3 --pmc
# passes (FETCH_SIZE, WRITE_SIZE) scaled by the same counters on the copy
# kernel; see profiles/hop_f64r18_32x4_traffic.md and hop_f32r12_32x4_traffic.md.
# None = not measured.
TRAFFIC_PER_LAUNCH = {"f64_r18": 903698562, "f32_r12": 386920650}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(threads, budget_s=12.0):
    """The oracle's hop (plain C + OpenMP, FP64, 18-real links) on a 16^4
    sample of the same workload, timed on the host cores."""
    import oracle
    from yahmc_amd.synthetic import gaussian_source, random_gauge
    dims = (16, 16, 16, 16)
    o = oracle.Oracle("su3_fund").init(dims, bc=oracle.BC_ANTIPERIODIC, csw=0.0)
    o.set_threads(threads)
    o.set_gauge(random_gauge(dims, 3, seed=1))
    o.represent()
    s = gaussian_source(dims, 3, seed=2, parity="full")
    d = np.zeros_like(s)
    o.Dphi_eo(d, s)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        o.Dphi_eo(d, s)
        n += 1
    dt = (time.perf_counter() - t0) / n
    vh = o.VH
    return {"value": round(BYTES_PER_SITE[("f64", 0)] * vh / dt / 1e9, 3), "unit": "GB/s", "cores": threads,
            "kind": "port", "cpu": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"oracle Dphi_eo (FP64, 18-real links) on a 16^4 lattice, {n} applications in {budget_s:.0f} s, "
                      f"{dt * 1e3:.2f} ms each; same compulsory-bytes model as the GPU value"}


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform's."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def true_residual(lat, fx, fb, mass=MASS):
    ax = lat.field("f64")
    lat.Hphi_sq(mass, ax, fx)
    lat.spinor_mulr_add_assign(ax, -1.0, fb)
    res = lat.spinor_sqnorm(ax) / lat.spinor_sqnorm(fb)
    ax.free()
    return float(np.sqrt(res))


def mixed_half_links(lat, fx, fb, prec):
    """Mixed CG whose FP32 sloppy operator reads FP16 12-real links."""
    import yahmc_amd as yh
    lat.tune(yh.TUNE_HALF_LINKS, 1)
    try:
        return lat.cg_mixed(lat.inv_par(MASS, prec, "Hphi_sq", sloppy="f32", max_iter=20000), fx, fb)
    finally:
        lat.tune(yh.TUNE_HALF_LINKS, 0)


def learned_solves(yh, lat, fb, prec, train_steps):
    """Train gauge-covariant networks on the device on (source, solution)
    pairs, then time the flexible CG on H^2 with M = A A^+ (FP32 inner,
    FP64 reliable updates) against the same solver without it.  The pairs
    are (H^2 v, v) for Gaussian v: exact solutions that cost one operator
    application each (training on FP64 CG solutions of Gaussian sources is
    supported by the same call but weights the lowest modes, which a local
    network cannot represent)."""
    from yahmc_amd.synthetic import gaussian_source
    out = {}
    t0 = time.perf_counter()
    bs, xs = [], []
    for k in range(8):
        x = lat.field("f64").upload(gaussian_source(DIMS, 3, seed=32300 + k))
        b = lat.field("f64")
        lat.Hphi_sq(MASS, b, x)
        bs.append(b)
        xs.append(x)
    t_pairs = time.perf_counter() - t0
    nets = {}
    for name, chans in (("learned_f32", (1, 8, 8, 1)), ("learned_small_f32", (1, 4, 1))):
        net = yh.Net(lat, chans)
        rng = np.random.default_rng(32400)
        net.init_weights(rng)
        lat.synchronize()
        t0 = time.perf_counter()
        hist = net.train(bs, xs, steps=train_steps, batch=4, lr=2e-3, rng=rng)
        lat.synchronize()
        t_train = time.perf_counter() - t0
        nets[name] = (net, {"pairs": len(bs), "pair_kind": "(H^2 v, v), v Gaussian", "pair_seconds": round(t_pairs, 2),
                            "steps": train_steps, "batch": 4, "train_seconds": round(t_train, 2),
                            "loss_first10": round(float(np.mean(hist[:10])), 4),
                            "loss_last10": round(float(np.mean(hist[-10:])), 4), "channels": list(chans)})
        log(f"network training {name}: {nets[name][1]}")
    for name, n in [("fcg_plain_f32", None)] + [(k, v[0]) for k, v in nets.items()]:
        fx = lat.field("f64")
        lat.synchronize()
        t = time.perf_counter()
        it, upd = lat.cg_net(lat.inv_par(MASS, prec, "Hphi_sq", max_iter=20000), n, fx, fb)
        lat.synchronize()
        dt = time.perf_counter() - t
        out[name] = {"seconds": round(dt, 4), "iterations": it, "reliable_updates": upd,
                     "true_rel_residual": true_residual(lat, fx, fb)}
        log(f"cg {name}: {out[name]}")
        fx.free()
    for k, (net, train) in nets.items():
        out[k]["training"] = train
        out[k]["iteration_ratio_vs_fcg_plain"] = round(out["fcg_plain_f32"]["iterations"] / max(1, out[k]["iterations"]), 2)
        net.free()
    for f in bs + xs:
        f.free()
    return out


def learned_protocol(yh, dims, ntrain=4, nheld=2, nsrc=4, steps=120, mass=-0.76, pairs="hsq", gcr=False, npairs=8):
    """The learned-preconditioner evaluation protocol, reduced in size:
    `ntrain` training and `nheld` held-out configurations (each a cold start
    smoothed by 10 pure-gauge trajectories with its own seed), m = -0.76,
    a 1-8-8-1 network trained on the device config by config on (H^2 v, v)
    pairs of each training configuration, then on every held-out
    configuration `nsrc` Gaussian sources solved one after another to an FP64
    true residual of 1e-10 by (a) mixed FP32 CG with the MRE guess of the
    previous solutions (mre_past = 5) and (b) the learned solve (flexible CG,
    or flexible GCR with `gcr`, with the network from its own MRE guess,
    mixed-CG fallback).  Training pairs: `hsq` = (H^2 v, v) for Gaussian v
    (exact, one operator application each); `solution` = (Gaussian source,
    its solution to an FP64 true residual of 1e-10 by the mixed FP32 CG), the
    workload's kind."""
    from yahmc_amd.md import smooth
    from yahmc_amd.synthetic import gaussian_source
    t_all = time.perf_counter()
    lat = yh.Lattice(dims, nc=3, repr="fundamental", bc="antiperiodic", csw=CSW)
    cfgs = []
    for k in range(ntrain + nheld):
        smooth(lat, beta=5.6, trajectories=10, length=1.0, steps=10, seed=32246 + 1000 * (k + 1))
        cfgs.append((lat.get_gauge(), lat.plaquette()))
        log(f"protocol: configuration {k} smoothed, plaquette {cfgs[-1][1]:.5f}")
    net = yh.Net(lat, (1, 8, 8, 1))
    rng = np.random.default_rng(32500)
    net.init_weights(rng)
    hist = []
    t0 = time.perf_counter()
    for k in range(ntrain):
        lat.set_gauge(cfgs[k][0])
        if pairs == "hsq":
            xs = [lat.field("f64").upload(gaussian_source(dims, 3, seed=32600 + 100 * k + j)) for j in range(npairs)]
            bs = []
            for x in xs:
                b = lat.field("f64")
                lat.Hphi_sq(mass, b, x)
                bs.append(b)
        else:
            bs = [lat.field("f64").upload(gaussian_source(dims, 3, seed=32600 + 100 * k + j)) for j in range(npairs)]
            xs = []
            for j, b in enumerate(bs):
                x = lat.field("f64")
                it, _ = lat.cg_mixed(lat.inv_par(mass, 1e-20, "Hphi_sq", sloppy="f32", max_iter=50000), x, b)
                xs.append(x)
                log(f"protocol: configuration {k} pair {j}: {it} CG iterations")
        hist += net.train(bs, xs, steps=steps // ntrain, batch=4, lr=2e-3, rng=rng)
        log(f"protocol: trained on configuration {k}, loss {hist[-1]:.4f}")
        for f in bs + xs:
            f.free()
    lat.synchronize()
    t_train = time.perf_counter() - t0
    par = lat.inv_par(mass, 1e-20, "Hphi_sq", sloppy="f32")
    lat.tune(yh.TUNE_LEARNED_GCR, 1 if gcr else 0)
    res = {"mixed_cg_mre5": {"iterations": 0, "seconds": 0.0, "max_true_rel_residual": 0.0},
           "learned_fcg_mre5": {"iterations": 0, "seconds": 0.0, "max_true_rel_residual": 0.0, "fallbacks": 0}}
    for k in range(ntrain, ntrain + nheld):
        lat.set_gauge(cfgs[k][0])
        for name, use_net in (("mixed_cg_mre5", None), ("learned_fcg_mre5", net)):
            mre = yh.MRE(lat, 5, "Hphi_sq", mass)
            for j in range(nsrc):
                fb = lat.field("f64").upload(gaussian_source(dims, 3, seed=32700 + 100 * k + j))
                fx = lat.field("f64")
                lat.synchronize()
                t = time.perf_counter()
                it, _, fell = lat.cg_learned(par, use_net, mre, fx, fb)
                lat.synchronize()
                r = res[name]
                r["seconds"] += time.perf_counter() - t
                r["iterations"] += it
                r["max_true_rel_residual"] = max(r["max_true_rel_residual"], true_residual(lat, fx, fb, mass))
                if use_net is not None:
                    r["fallbacks"] += int(fell)
                log(f"protocol: held-out configuration {k} source {j} {name}: {it} iterations")
                fb.free()
                fx.free()
            mre.free()
    for r in res.values():
        r["seconds"] = round(r["seconds"], 3)
    out = {"lattice": list(dims), "mass": mass, "train_configs": ntrain, "held_out_configs": nheld,
           "sources_per_config": nsrc, "plaquettes": [round(c[1], 5) for c in cfgs],
           "learned_solver": "flexible GCR" if gcr else "flexible CG",
           "network": [1, 8, 8, 1], "training": {"pairs_per_config": npairs,
                                                  "pair_kind": "(H^2 v, v), v Gaussian" if pairs == "hsq" else
                                                  "(Gaussian source, solution to FP64 true residual 1e-10)",
                                                  "steps": len(hist), "seconds": round(t_train, 2),
                                                  "loss_first10": round(float(np.mean(hist[:10])), 4),
                                                  "loss_last10": round(float(np.mean(hist[-10:])), 4)},
           "held_out": res, "seconds_total": round(time.perf_counter() - t_all, 1)}
    net.free()
    lat.close()
    log(f"learned protocol: {out}")
    return out


SLOPPY_NAME = {"f32": "FP32", "f16": "16-bit fixed point"}


def hmc_trajectory(yh, sloppy="f16"):
    """HMC trajectories on 24^3 x 32 with two flavours of sextet (two-index
    symmetric SU(3)) clover fermions: tau = 1, o2mn with the Wilson gauge
    monomial on level 0 (1 step) and the fermion monomial on level 1,
    mixed-precision CG (16-bit fixed-point sloppy operator by default) to an
    FP64 true residual of 1e-12, FP64 link update and Hamiltonian; beta = 5.6
    and m = -0.75 (values the workload leaves open).  The seeded random SU(3)
    field is first smoothed as the 32^4 field is (cold start, 10 pure-gauge
    trajectories at beta = 5.6).  The timed trajectory has 10 fermion steps;
    the same trajectory (same momenta and pseudofermions) is then repeated
    from the same field with 20 and 40 steps, and o2mn's dH ~ dt^2 makes the
    ratios of dH at 10 and 20 and at 20 and 40 steps approach 4."""
    from yahmc_amd.hmc import HMC, FermionMonomial, GaugeMonomial
    from yahmc_amd.md import smooth
    dims = (32, 24, 24, 24)
    lat = yh.Lattice(dims, nc=3, repr="symmetric", bc="antiperiodic", csw=1.0)
    t0 = time.perf_counter()
    smooth(lat, beta=5.6, trajectories=10, length=1.0, steps=10, seed=32246)
    lat.represent_gauge_field()
    smooth_s = time.perf_counter() - t0
    U = lat.get_gauge()
    plaq0 = lat.plaquette()
    runs = {}
    for nsteps in (10, 20, 40):
        lat.set_gauge(U)
        fm = FermionMonomial(-0.75, prec=1e-24, level=1, sloppy=sloppy)
        hmc = HMC(lat, [GaugeMonomial(5.6), fm], [("o2mn", 1), ("o2mn", nsteps)])
        lat.synchronize()
        t0 = time.perf_counter()
        rec = hmc.trajectory(np.random.default_rng(32247), length=1.0, metropolis=(nsteps == 10))
        lat.synchronize()
        runs[nsteps] = {"seconds": round(time.perf_counter() - t0, 3), "dH": rec["dH"], "accepted": rec["accepted"],
                        "plaquette_after": round(rec["plaquette"], 6), "cg_solves": len(fm.iterations),
                        "cg_iterations_total": int(sum(fm.iterations))}
    out = {"lattice": list(dims), "repr": "sextet", "beta": 5.6, "mass": -0.75, "csw": 1.0,
           "start": {"smoothing": "cold start + 10 pure-gauge trajectories, beta 5.6, 10 o2mn steps",
                     "seconds": round(smooth_s, 2), "plaquette": round(plaq0, 6)},
           "integrator": "o2mn, gauge level 0 x 1 step, fermion level 1 x 10 steps", "tau": 1.0,
           "seconds": runs[10]["seconds"], "dH": runs[10]["dH"], "accepted": runs[10]["accepted"],
           "plaquette_after": runs[10]["plaquette_after"], "cg_solves": runs[10]["cg_solves"],
           "cg_iterations_total": runs[10]["cg_iterations_total"],
           "fermion_steps_20": runs[20], "fermion_steps_40": runs[40],
           "dH_ratio_10_over_20": round(runs[10]["dH"] / runs[20]["dH"], 3) if runs[20]["dH"] != 0 else None,
           "dH_ratio_20_over_40": round(runs[20]["dH"] / runs[40]["dH"], 3) if runs[40]["dH"] != 0 else None,
           "cg": f"mixed reliable-update CG, sloppy {SLOPPY_NAME[sloppy]}, FP64 true residual 1e-12, "
                 f"update when the residual drops by {FermionMonomial.DELTA} or every "
                 f"{FermionMonomial.RELIABLE_EVERY} iterations, MRE guess from "
                 f"{FermionMonomial.MRE_PAST} past solution(s)"}
    log(f"hmc trajectory: {out}")
    lat.close()
    return out


def force_core(yh, reps=3):
    """The clover-force core alone (for profiling): 24^3 x 32 sextet, seeded
    random SU(3) links, clover force matrices from two Gaussian fields, then
    `reps` end_fermion_force calls, each one k_force_clover_x launch."""
    from yahmc_amd.synthetic import gaussian_source, random_gauge
    dims = (32, 24, 24, 24)
    lat = yh.Lattice(dims, nc=3, repr="symmetric", bc="antiperiodic", csw=1.0)
    lat.set_gauge(random_gauge(dims, 3, seed=32250))
    lat.set_momenta(np.zeros((4 * lat.V, 8)))
    a = lat.field("f64").upload(gaussian_source(dims, 6, seed=32251, parity="both"))
    b = lat.field("f64").upload(gaussian_source(dims, 6, seed=32252, parity="both"))
    ts = []
    for _ in range(reps):
        lat.begin_fermion_force()
        lat.clover_fermion_force(1.0, a, b)
        lat.synchronize()
        t0 = time.perf_counter()
        lat.end_fermion_force(0.1)
        lat.synchronize()
        ts.append(time.perf_counter() - t0)
    lat.close()
    return {"lattice": list(dims), "repr": "sextet", "end_fermion_force_ms": [round(t * 1e3, 3) for t in ts]}


def rhmc_multishift(yh):
    """The RHMC solve of the SU(2) adjoint workload: 24^3 x 48, clover,
    m = -1.0; 12-pole Zolotarev x^(-1/2) on [1e-4, 8]; 4 Gaussian
    pseudofermion sources through one batched multishift solve to
    per-shift tolerance 1e-10 (prec 1e-20), FP64 partial-fraction sum.
    Timed with the chunk replayed from a HIP graph and launched eagerly, and
    with the operator applied to the four right-hand sides per hop launch
    (each link and clover block loaded once for all of them) or one
    right-hand side per launch."""
    from yahmc_amd.synthetic import gaussian_source, random_gauge
    dims = (48, 24, 24, 24)
    lat = yh.Lattice(dims, nc=2, repr="adjoint", bc="antiperiodic", csw=1.0)
    lat.set_gauge(random_gauge(dims, 2, seed=32244))
    tuned = lat.autotune(reps=10)
    rat = yh.rational_find(-1, 2, 1e-10, 1e-4, 8.0)
    if rat.order > 12:
        rat = yh.rational_zolotarev(12, -1, 2, 1e-4, 8.0)
    shifts = [rat.pole[i] for i in range(rat.order)]
    srcs = [lat.field("f64").upload(gaussian_source(dims, lat.dr, seed=32247 + k)) for k in range(4)]
    out = {"lattice": list(dims), "repr": "SU(2) adjoint", "mass": -1.0, "csw": 1.0,
           "rational": {"order": rat.order, "range": [1e-4, 8.0], "max_rel_error": rat.prec},
           "rhs": len(srcs), "prec": 1e-20, "xcd_chunks": {f"{k[0]}_r{k[1]}": v for k, v in tuned.items()}}
    for name, graphs, multi in (("graph", 1, 1), ("eager", 0, 1), ("graph_one_rhs_per_hop", 1, 0)):
        lat.tune(yh.TUNE_GRAPHS, graphs)
        lat.tune(yh.TUNE_MULTI_RHS, multi)
        dst = [[lat.field("f64") for _ in shifts] for _ in srcs]
        lat.synchronize()
        t0 = time.perf_counter()
        its = lat.cg_mshift_batched(lat.inv_par(-1.0, 1e-20, "Hphi_sq"), shifts, dst, srcs)
        lat.synchronize()
        dt = time.perf_counter() - t0
        out[name] = {"seconds": round(dt, 3), "iterations": its}
        if name == "graph":
            # per-shift FP64 true residual |b - (H^2 + s) x| / |b| of every solution
            worst = 0.0
            for j, row in enumerate(dst):
                for s, f in zip(shifts, row):
                    ax = lat.field("f64")
                    lat.Hphi_sq(-1.0, ax, f)
                    lat.spinor_mulr_add_assign(ax, s, f)
                    lat.spinor_mulr_add_assign(ax, -1.0, srcs[j])
                    worst = max(worst, float(np.sqrt(lat.spinor_sqnorm(ax) / lat.spinor_sqnorm(srcs[j]))))
                    ax.free()
            out[name]["max_true_rel_residual"] = worst
        for row in dst:
            for f in row:
                f.free()
    lat.tune(yh.TUNE_MULTI_RHS, 1)
    # the partial-fraction application r(H^2) b for one source, FP64 sum
    fd = lat.field("f64")
    t0 = time.perf_counter()
    it = lat.rational_calc(rat, lat.inv_par(-1.0, 1e-20, "Hphi_sq"), fd, srcs[0])
    lat.synchronize()
    out["rational_calc_one_rhs"] = {"seconds": round(time.perf_counter() - t0, 3), "iterations": it}
    log(f"rhmc multishift: {out}")
    lat.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--prec", default="f64", choices=["f64", "f32", "f16"], help="precision of the headline hop")
    ap.add_argument("--skip-cg", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--only-hop", action="store_true", help="time the headline hop only (for profiling)")
    ap.add_argument("--no-smooth", action="store_true", help="keep the hot random field (no gauge trajectories)")
    ap.add_argument("--skip-learned", action="store_true")
    ap.add_argument("--train-steps", type=int, default=120)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default XCD chunk instead of tuning it")
    ap.add_argument("--xcd-chunk", type=int, default=None,
                    help="fix the XCD chunk of every hop instead of tuning it (profiling runs)")
    ap.add_argument("--skip-hmc", action="store_true")
    ap.add_argument("--skip-rhmc", action="store_true")
    ap.add_argument("--only-rhmc", action="store_true", help="the batched multishift workload only")
    ap.add_argument("--only-hmc", action="store_true", help="the sextet HMC trajectory only")
    ap.add_argument("--only-force", action="store_true", help="the sextet clover-force core only (for profiling)")
    ap.add_argument("--hmc-sloppy", default="f16", choices=["f32", "f16"], help="sloppy precision of the HMC solves")
    ap.add_argument("--skip-protocol", action="store_true", help="skip the held-out learned-preconditioner protocol")
    ap.add_argument("--only-protocol", action="store_true", help="the learned-preconditioner protocol only")
    ap.add_argument("--protocol-pairs", default="hsq", choices=["hsq", "solution"],
                    help="training pairs of the protocol: (H^2 v, v) or (Gaussian source, solution)")
    ap.add_argument("--protocol-gcr", action="store_true", help="the protocol's learned solve is flexible GCR")
    ap.add_argument("--protocol-sizes", default="4,2,4,8",
                    help="training configurations, held-out configurations, sources per held-out configuration, "
                         "pairs per training configuration")
    ap.add_argument("--protocol-dims", default="32,16,16,16",
                    help="lattice T,X,Y,Z of the learned-preconditioner protocol (the workload's is 64,32,32,32)")
    ap.add_argument("--only-learned", action="store_true",
                    help="smoothing, network training and the two flexible-CG solves only (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU runs are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import yahmc_amd as yh
    from yahmc_amd.synthetic import gaussian_source, random_gauge

    if args.only_hmc:
        print(json.dumps(hmc_trajectory(yh, args.hmc_sloppy)))
        return
    if args.only_force:
        print(json.dumps(force_core(yh)))
        return
    if args.only_rhmc:
        print(json.dumps(rhmc_multishift(yh)))
        return
    if args.only_protocol:
        nt, nh, ns, npr = (int(v) for v in args.protocol_sizes.split(","))
        print(json.dumps(learned_protocol(yh, tuple(int(v) for v in args.protocol_dims.split(",")), ntrain=nt, nheld=nh,
                                          nsrc=ns, npairs=npr, pairs=args.protocol_pairs, gcr=args.protocol_gcr)))
        return
    t0 = time.perf_counter()
    links = random_gauge(DIMS, 3, seed=32244)
    src = gaussian_source(DIMS, 3, seed=32245, parity="full")
    log(f"synthetic inputs {time.perf_counter() - t0:.1f}s")
    lat = yh.Lattice(DIMS, nc=3, repr="fundamental", bc="antiperiodic", csw=CSW)
    lat.set_gauge(links)
    log("gauge field, representation and clover term built on device")
    smoothing = None
    if not args.no_smooth and not args.only_hop:
        from yahmc_amd.md import smooth
        t0 = time.perf_counter()
        traj = smooth(lat, beta=5.6, trajectories=10, length=1.0, steps=10, seed=32246)
        smoothing = {"start": "cold", "beta": 5.6, "trajectories": 10, "tau": 1.0, "md_steps": 10,
                     "integrator": "o2mn", "metropolis": False,
                     "dH": [round(t["dH"], 3) for t in traj], "plaquette": [round(t["plaquette"], 5) for t in traj],
                     "seconds": round(time.perf_counter() - t0, 2)}
        log(f"smoothing: {smoothing}")

    tuned = None
    if args.xcd_chunk is not None:
        lat.tune(yh.TUNE_XCD_CHUNK, args.xcd_chunk)
        tuned = {"fixed": args.xcd_chunk}
    elif not args.no_autotune:
        t0 = time.perf_counter()
        tuned = {f"{k[0]}_r{k[1]}": v for k, v in lat.autotune(reps=20).items()}
        log(f"autotuned XCD chunks {tuned} ({time.perf_counter() - t0:.1f}s)")
    hop = {}
    variants = [(args.prec, 0 if args.prec == "f64" else 1)] if args.only_hop else \
        [("f64", 0), ("f64", 1), ("f32", 1), ("f32", 2), ("f16", 1)]
    if args.only_learned:
        fb = lat.field("f64").upload(gaussian_source(DIMS, 3, seed=32246))
        print(json.dumps(learned_solves(yh, lat, fb, 1e-20, args.train_steps)))
        return
    for prec, recon in variants:
        fs = lat.field(prec).upload(src)
        fd = lat.field(prec)
        lat.bench_hop(prec, recon, max(1, args.warmup), fd, fs)
        ms = lat.bench_hop(prec, recon, args.steps, fd, fs)
        b = BYTES_PER_SITE[(prec, recon)] * lat.VH
        hop[f"{prec}_r{ {0: '18', 1: '12', 2: '12h'}[recon] }"] = {
            "us_per_hop": round(ms * 1e3, 2), "GB_s": round(b / (ms * 1e-3) / 1e9, 1),
            "frac_hbm": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3),
            "GFLOP_s": round(FLOP_PER_SITE * lat.VH / (ms * 1e-3) / 1e9, 1), "bytes_per_site": BYTES_PER_SITE[(prec, recon)],
            "GB_s_naive_model": round(BYTES_PER_SITE_NAIVE[(prec, recon)] * lat.VH / (ms * 1e-3) / 1e9, 1)}
        log(f"hop {prec} recon{ {0: '18', 1: '12', 2: '12h'}[recon] }: {hop[list(hop)[-1]]}")
        fs.free()
        fd.free()
    key = f"{args.prec}_r{18 if args.prec == 'f64' else 12}"
    head = hop[key]
    copy_ms = lat.bench_copy(1 << 30, 20)
    copy_gbs = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    log(f"copy kernel: {copy_gbs:.0f} GB/s")

    cg = {}
    if not args.skip_cg and not args.only_hop:
        b = gaussian_source(DIMS, 3, seed=32246)
        fb = lat.field("f64").upload(b)
        prec = 1e-20
        for name, fn in (("fp64", lambda fx: (lat.cg_inverter(lat.inv_par(MASS, prec, "Hphi_sq", max_iter=20000), fx, fb), 0)),
                         ("mixed_f32", lambda fx: lat.cg_mixed(lat.inv_par(MASS, prec, "Hphi_sq", sloppy="f32", max_iter=20000), fx, fb)),
                         ("mixed_f16", lambda fx: lat.cg_mixed(lat.inv_par(MASS, prec, "Hphi_sq", sloppy="f16", max_iter=20000,
                                                                           reliable_every=50), fx, fb)),
                         ("mixed_f32_h16links", lambda fx: mixed_half_links(lat, fx, fb, prec)),
                         ("bicgstab_fp64", lambda fx: (lat.bicgstab_inverter(lat.inv_par(MASS, prec, "Hphi_sq", max_iter=20000), fx, fb), 0)),
                         ("bicgstab_mixed_f32", lambda fx: lat.bicgstab_mixed(lat.inv_par(MASS, prec, "Hphi_sq", max_iter=20000), fx, fb))):
            fx = lat.field("f64")
            lat.synchronize()
            t = time.perf_counter()
            it, upd = fn(fx)
            lat.synchronize()
            dt = time.perf_counter() - t
            # FP64 true residual with the FP64 GPU operator
            cg[name] = {"seconds": round(dt, 4), "iterations": it, "reliable_updates": upd,
                        "true_rel_residual": true_residual(lat, fx, fb)}
            if name == "mixed_f16":
                cg[name]["reliable_every"] = 50
            if name.startswith("bicgstab"):
                cg[name]["note"] = "BiCGstab on H^2 (iterations = 2 x BiCGstab steps, as bicgstab.cpp returns them)"
            log(f"cg {name}: {cg[name]}")
            fx.free()
        if not args.skip_learned:
            cg.update(learned_solves(yh, lat, fb, prec, args.train_steps))
        if "fp64" in cg:
            for k in ("mixed_f32", "mixed_f16", "mixed_f32_h16links", "bicgstab_fp64", "bicgstab_mixed_f32", "fcg_plain_f32",
                      "learned_f32", "learned_small_f32"):
                if k in cg:
                    cg[k]["speedup_vs_fp64"] = round(cg["fp64"]["seconds"] / cg[k]["seconds"], 2)

    hmc = rhmc = protocol = None
    if not args.only_hop:
        lat.close()
        if not args.skip_hmc:
            hmc = hmc_trajectory(yh, args.hmc_sloppy)
        if not args.skip_rhmc:
            rhmc = rhmc_multishift(yh)
        if not args.skip_protocol and not args.skip_learned:
            protocol = learned_protocol(yh, tuple(int(v) for v in args.protocol_dims.split(",")))

    cpu = None
    if not args.skip_cpu and not args.only_hop:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(threads)
        log(f"cpu baseline: {cpu}")

    ms = head["us_per_hop"] / 1e3
    algo_bytes = BYTES_PER_SITE[(args.prec, 0 if args.prec == "f64" else 1)] * lat.VH
    line = {
        "metric": "Dslash GB/s (% HBM peak); CG time to 1e-10, mixed+learned vs FP64, 32^4",
        "value": head["GB_s"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32", "f16": "int16-fixed-point-storage/f32-arith"}[args.prec],
        "data": "synthetic (seeded SU(3) field smoothed by pure-gauge trajectories, Gaussian sources)"
                if smoothing else "synthetic (seeded random SU(3) links, Gaussian source)",
        "config": {"workload": f"hop D_eo, {args.prec} recon-{18 if args.prec == 'f64' else 12}, 32^4 SU(3) fundamental "
                               "clover, antiperiodic time",
                   "lattice": list(DIMS), "sites_per_step": lat.VH, "mass": MASS, "csw": CSW,
                   "gauge": smoothing if smoothing else "random SU(3) links (hot), not smoothed",
                   "xcd_chunk": tuned if tuned else "default (1/8 of a t-slice)"},
        "roofline": {"bound": "hbm", "achieved": head["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(head["GB_s"] / HBM_PEAK_GBS, 4), "traffic": TRAFFIC_PER_LAUNCH.get(key),
                     "algorithmic_bytes_per_launch": algo_bytes},
        "cpu_baseline": cpu,
        "dslash": hop,
        "copy_kernel_GB_s": round(copy_gbs, 1),
        "cg_time_to_1e-10": cg,
        "hmc_trajectory": hmc,
        "rhmc_multishift": rhmc,
        "learned_protocol": protocol,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
