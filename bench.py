"""Particle-step throughput of the MI355X LBVH + SPH step on the 1,048,576-particle two-body impact.

Workload (one GPU): the seeded two-body giant impact (917,504-particle target + 131,072-particle impactor,
30 % iron / 70 % granite, h for 48 neighbours, Barnes-Hut theta = 0.5, Plummer softening 0.01, AV 1/2),
plain FP32, adaptive Courant dt at CFL 0.3 computed on the device. A step is one full KDK step: tree build,
gravity, density + EOS and SPH force twice (eval -> half-kick + drift -> eval -> half-kick), as the
reference's CPU step and GPU driver do. Inputs are resident in HBM before the timed region.

Prints one JSON line: throughput (particle-steps/s), per-phase ms, roofline of the dominant kernel
(gravity traversal, flops from counted interactions, duration from HIP events on the handle's stream),
energy / angular-momentum drift over the timed run (exact f64 PE on the device), and the CPU baseline
(the oracle's single-threaded restatement of HydroBody::step, one force evaluation of the same scene), plus
the mixed-precision form and the evaluation-reuse variant of the same workload and the adjoint/forward cost
of the reverse pass on the 262,144-particle differentiation workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 peak (packed FP32: two lanes' worth per VALU op)
FP32_NONPACKED_TFLOPS = 78.6  # the same chip without packed FP32, which this build's HIP code does not emit
HBM_PEAK_GBS = 8000.0
DE_BOUND = 1e-4  # the energy bound the headline metric names


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--precision", choices=["fp32", "mixed"], default="fp32")
    ap.add_argument("--reuse", action="store_true", help="reuse tree/gravity/density between evals at equal positions")
    ap.add_argument("--gravity", choices=["tree", "direct"], default="tree")
    ap.add_argument("--quadrupole", action="store_true")
    ap.add_argument("--leaf", type=int, default=1)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--group-extent", type=float, default=4.0)
    ap.add_argument("--ppl", type=int, default=1, help="gravity walk particles per lane")
    ap.add_argument("--expand", type=int, default=4, help="SPH walks take subtrees of <= this many particles whole")
    ap.add_argument("--cfl", type=float, default=0.3)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward cost measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-instance ensemble measurement")
    ap.add_argument("--no-4m", action="store_true", help="skip the plain / mixed throughput at N = 4,194,304")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--adjoint-n", type=int, default=262144)
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=20, help="checkpoint interval of the reverse pass")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint/forward cost measurement and print its object (for profiling it)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-precision-study", action="store_true",
                    help="skip the precision-attributable energy error of the plain and mixed forms")
    ap.add_argument("--study-steps", type=int, default=200, help="KDK steps of the precision study")
    ap.add_argument("--drift-study", action="store_true",
                    help="run only the energy-drift decomposition (time step, gravity, clamp) and print it")
    ap.add_argument("--no-energy", action="store_true", help="skip the O(N^2) energy diagnostics")
    ap.add_argument("--variants", default="mixed,reuse",
                    help="comma list of extra timed runs of the same workload: mixed (precision form), reuse "
                         "(tree/gravity/density carried from the step's second evaluation to the next step's "
                         "first, bit-identical results); empty for none")
    return ap.parse_args()


def lscpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measured_traffic():
    """HBM bytes per gravity launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by the
    factors measured on plain streaming kernels), and the commit whose code those passes measured; (None,
    None) when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_gravity_tree_two_body_1m_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d["traffic_bytes_per_launch"], d.get("measured_on")
    except (OSError, KeyError, ValueError):
        return None, None


def host_cores():
    """Logical CPUs this process may run on (the box's share), and the machine's total."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return avail, os.cpu_count()


def cpu_c1_run(steps=100, cfl=0.3):
    """The CPU path's own workload: the oracle's restatement of HydroBody::step (hydrostatic.rs:338-480,
    octree gravity theta 0.5, f64) on the 8192-particle two-body scene for `steps` KDK steps at its Courant
    dt, 1 thread, with the time split into density, gravity and SPH force + du/dt."""
    from botheadstudios_amd import scene
    from oracle import oracle
    oracle.set_threads(1)
    sc = scene.two_body_8k()
    pos, vel, mass, u, h, mat = scene.f64_state(sc.particles)
    hb = oracle.Hydro(pos, vel, mass, u, h, mat, sc.eos64, sc.softening)
    t0 = time.perf_counter()
    hb.compute_density()
    for _ in range(steps):
        hb.step(hb.courant_dt(cfl))
    wall = time.perf_counter() - t0
    ph = hb.phase_ms()
    n = len(mass)
    return {"workload": f"two-body impact N={n}, {steps} KDK steps at CFL {cfl}, f64, octree theta 0.5",
            "wall_s": wall, "value": n * steps / wall, "unit": "particle-steps/s", "cores": 1,
            "ms_per_step": {k: v / steps for k, v in ph.items()}, "ms_per_step_total": wall * 1e3 / steps}


def cpu_baseline(sc):
    """One force evaluation (density + forces_and_dudt = half a KDK step) of the oracle's restatement of
    HydroBody (hydrostatic.rs:338-446, octree gravity theta 0.5) on the same scene, 1 thread; plus the CPU
    path's own 100-step run on the 8192-particle scene."""
    from botheadstudios_amd import scene
    from oracle import oracle
    oracle.set_threads(1)
    pos, vel, mass, u, h, mat = scene.f64_state(sc.particles)
    hb = oracle.Hydro(pos, vel, mass, u, h, mat, sc.eos64, sc.softening)
    t0 = time.perf_counter()
    hb.compute_density()
    hb.forces()
    dt = time.perf_counter() - t0
    ph = hb.phase_ms()
    hb.close()
    n = len(mass)
    avail, total = host_cores()
    # the reference GPU walk's own work (per-particle MAC, bh_gravity.wgsl:337-394) on a seeded sample
    oracle.set_threads(avail)
    sample = np.random.default_rng(0x16).choice(n, 16384, replace=False)
    counts = oracle.lbvh_walk_counts(sc.particles["pos"], sc.particles["mass"], sc.softening, sample)
    oracle.set_threads(1)
    return {"value": n / (2.0 * dt), "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"one force evaluation (density + forces_and_dudt, half a KDK step) of the N={n} scene, "
                      f"{dt:.1f} s on 1 thread of {lscpu_model()}",
            "host": {"cpu_model": lscpu_model(), "logical_cpus_available": avail, "logical_cpus_machine": total},
            "eval_ms": ph, "c1_run": cpu_c1_run(),
            "reference_walk": {"sample": "16,384 particles of the start state (seed 0x16)",
                               "cells_per_particle": float(counts[:, 0].mean()),
                               "particles_per_particle": float(counts[:, 1].mean())}}


def time_disk_view(s, reps=5):
    """Wall time of one blocking on-device disk measurement (DiskView::of of every instance with its
    clumps) plus the clump query of the first instance, averaged over `reps`."""
    s.disk_view()
    t0 = time.perf_counter()
    for _ in range(reps):
        views = s.disk_view()
        clumps = s.disk_clumps(0, 2)
    ms = (time.perf_counter() - t0) * 1e3 / reps
    return ms, views, clumps


def measure_cpu(rec):
    """The oracle's DiskView::of and DiskView::clumps (reference self energy) on the same read-back state,
    1 thread."""
    from oracle import oracle
    oracle.set_threads(1)
    t0 = time.perf_counter()
    v, idx = oracle.disk_view(rec)
    if len(idx) >= 2:
        p = rec["pos"][idx].astype(np.float64)
        oracle.find_clumps(p, rec["vel"][idx].astype(np.float64), rec["mass"][idx].astype(np.float64),
                           np.maximum(rec["rho"][idx], np.float32(1.0)).astype(np.float64), 2.0 * v["mean_h"],
                           6.674e-11, 1.0e4, [v["com_x"], v["com_y"], v["com_z"]], v["m_remnant"], v["r_remnant"])
    return (time.perf_counter() - t0) * 1e3


def run_debris(n):
    """The disk measurement on a synthetic post-impact debris field of n particles (a percolating diffuse
    disk plus cold clumps), so friends-of-friends, clump statistics and the self-energy sums run at scale."""
    from botheadstudios_amd import isph, materials, scene
    rec = scene.debris_field(n)
    with isph.GpuSph(n) as s:
        s.upload(rec, materials.eos_table(("granite",)), 1.0e4)
        ms, views, clumps = time_disk_view(s)
    d = views[0]
    return {"workload": f"DiskView and clumps of a synthetic post-impact debris field, N={n} (central body 80 % "
                        f"of the particles, orbiting disk with cold 50-particle clumps, plunging and escaping "
                        f"debris)", "gpu_ms": ms, "n_disk": d.n_disk, "n_clumps": d.n_clumps,
            "n_bound": d.n_bound, "largest_clump_members": max((c.n_members for c in clumps), default=0)}, rec


def run_gpu(sc, args, precision, steps, warmup, reuse, measure=False):
    from botheadstudios_amd import isph
    rec = sc.particles
    n = len(rec)
    s = isph.GpuSph(n)
    s.upload(rec, sc.eos, sc.softening)
    s.set_eos_f64(sc.eos64)
    s.set_gravity(isph.GRAVITY_TREE if args.gravity == "tree" else isph.GRAVITY_DIRECT, args.theta, args.leaf,
                  args.quadrupole)
    s.set_precision(isph.PRECISION_MIXED if precision == "mixed" else isph.PRECISION_FP32)
    s.set_walk(args.group_extent, args.ppl, args.expand)
    s.set_reuse(reuse)
    s.step_cfl(args.cfl, warmup)
    s.synchronize()
    d0 = None if args.no_energy else s.diagnostics(with_pe=True)
    t_sim0, _ = s.sim_time()
    s.reset_stats()
    s.set_timing(True)
    s.synchronize()
    t0 = time.perf_counter()
    s.step_cfl(args.cfl, steps)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    phases = s.phase_times()
    stats = s.kernel_stats()
    stats.update(s.walk_stats())
    clamp = s.clamp_energy()
    s.set_timing(False)
    t_sim1, dt_last = s.sim_time()
    d1 = None if args.no_energy else s.diagnostics(with_pe=True)
    out = {"elapsed": elapsed, "phases": phases, "stats": stats, "sim_time": t_sim1 - t_sim0, "dt_last": dt_last}
    if measure:
        def disk():
            ms, views, clumps = time_disk_view(s)
            d = views[0]
            return {"workload": f"DiskView of the N={n} state after the timed steps, with its clumps",
                    "gpu_ms": ms, "n_live": d.n_live, "n_disk": d.n_disk, "n_clumps": d.n_clumps,
                    "disk_mass": d.disk_mass(), "r_remnant": d.r_remnant}

        out["measure"] = _leg(disk)  # after the timed steps: a failure here keeps the headline
        out["measure_rec"] = None if "error" in out["measure"] else s.readback_sync()
    s.close()
    if d0 is not None:
        e0 = d0["ke"] + d0["ie"] + d0["pe"]
        e1 = d1["ke"] + d1["ie"] + d1["pe"]
        l0 = np.linalg.norm(d0["L"])
        out["dE_rel"] = abs(e1 - e0) / abs(e0)
        out["dL_rel"] = float(np.linalg.norm(d1["L"] - d0["L"]) / l0)
        out["clamp_energy_rel"] = clamp / abs(e0)
    return out


def run_adjoint(args):
    """Adjoint/forward cost on the differentiation workload: the two-body scene at N = 262,144, mixed form,
    LBVH gravity, 200 KDK steps at the device Courant dt, reverse pass with a checkpoint every 20 steps (each
    segment recomputed from its checkpoint). Loss: smoothed mass outside 2.9 target radii whose specific
    angular momentum exceeds the circular value. Also times the plain FP32 forward of the same 200 steps."""
    from botheadstudios_amd import isph, scene
    sc = scene.two_body(args.adjoint_n)
    n = len(sc.particles)
    loss = isph.Loss.make(scene.L_U, r_cut=2.9, k_r=20.0, k_j=10.0)
    with isph.GpuSph(n) as s:
        s.upload(sc.particles, sc.eos, sc.softening)
        s.set_eos_f64(sc.eos64)
        s.set_precision(isph.PRECISION_MIXED)
        s.set_gravity(isph.GRAVITY_TREE, args.theta, 1)
        s.step_cfl(args.cfl, 1)  # warm-up of the kernels (the reverse pass starts from this state)
        s.synchronize()
        t0 = time.perf_counter()
        out = s.vjp_run(loss, args.adjoint_steps, args.adjoint_every, cfl=args.cfl)
        wall = time.perf_counter() - t0
        dts = out["dts"]
    with isph.GpuSph(n) as s:
        s.upload(sc.particles, sc.eos, sc.softening)
        s.set_gravity(isph.GRAVITY_TREE, args.theta, 1)
        s.set_dt(float(dts[0]), 1.0)
        s.encode_kdk(1)
        s.synchronize()
        t0 = time.perf_counter()
        for dt in dts:
            s.set_dt(float(dt), 1.0)
            s.encode_kdk(1)
        s.synchronize()
        plain_ms = (time.perf_counter() - t0) * 1e3
    fwd, rev = float(out["ms"][0]), float(out["ms"][1])
    g = out["deos"][0]
    return {"workload": f"two-body impact N={n}, mixed form, LBVH theta={args.theta}, {args.adjoint_steps} KDK steps "
                        f"at CFL {args.cfl}, checkpoint every {args.adjoint_every}",
            "loss": out["loss"], "forward_ms": fwd, "reverse_ms": rev, "wall_s": wall,
            "adjoint_over_forward": rev / fwd, "plain_fp32_forward_ms": plain_ms,
            "adjoint_over_plain_forward": rev / plain_ms,
            "granite_dloss_d_a_b_A_B": [float(x) for x in g]}


def _energy(d):
    return d["ke"] + d["ie"] + d["pe"]


def run_precision_study(args, sc):
    """Precision-attributable energy error: the plain and mixed forms run the same discrete step as the f64
    reference form (LBVH theta, leaf 1, monopole, one particle per lane, the reference run's own Courant dt
    sequence) from the same state. A form's energy minus the reference's after the run is what that form's
    rounding costs, relative to |E0|; the same for the angular momentum about the COM."""
    from botheadstudios_amd import isph
    n = len(sc.particles)
    steps = args.study_steps

    def handle(form):
        s = isph.GpuSph(n)
        s.upload(sc.particles, sc.eos, sc.softening)
        s.set_eos_f64(sc.eos64)
        s.set_gravity(isph.GRAVITY_TREE, args.theta, 1)
        s.set_walk(args.group_extent, 1, args.expand)
        s.set_precision(form)
        return s

    res = {}
    with handle(isph.PRECISION_F64) as s:
        d0 = s.diagnostics(with_pe=True)
        s.reset_stats()
        dts = []
        t0 = time.perf_counter()
        for _ in range(steps):
            s.step_cfl(args.cfl, 1)
            dts.append(s.sim_time()[1])
        ref_ms = (time.perf_counter() - t0) * 1e3 / steps
        d1 = s.diagnostics(with_pe=True)
        res["f64"] = (d0, d1, s.clamp_energy())
    for name, form in (("fp32", isph.PRECISION_FP32), ("mixed", isph.PRECISION_MIXED)):
        with handle(form) as s:
            d0 = s.diagnostics(with_pe=True)
            s.reset_stats()
            for dt in dts:
                s.set_dt(dt, 1.0)
                s.encode_kdk(1)
            d1 = s.diagnostics(with_pe=True)
            res[name] = (d0, d1, s.clamp_energy())
    e0 = _energy(res["f64"][0])
    l0 = np.linalg.norm(res["f64"][0]["L"])
    e_ref, l_ref = _energy(res["f64"][1]), res["f64"][1]["L"]
    out = {"workload": f"two-body impact N={n}, LBVH theta={args.theta} monopole, {steps} KDK steps at the f64 "
                       f"reference run's Courant dt (CFL {args.cfl}), all forms from the same state",
           "steps": steps, "reference_ms_per_step": ref_ms,
           "dE_rel": {k: (_energy(v[1]) - _energy(v[0])) / abs(e0) for k, v in res.items()},
           "clamp_energy_rel": {k: v[2] / abs(e0) for k, v in res.items()}}
    for name in ("fp32", "mixed"):
        d1 = res[name][1]
        out[f"{name}_precision_dE"] = abs(_energy(d1) - e_ref) / abs(e0)
        out[f"{name}_precision_dL"] = float(np.linalg.norm(d1["L"] - l_ref) / l0)
    out["precision_dE_ratio_fp32_over_mixed"] = out["fp32_precision_dE"] / max(out["mixed_precision_dE"], 1e-300)
    out["precision_dL_ratio_fp32_over_mixed"] = out["fp32_precision_dL"] / max(out["mixed_precision_dL"], 1e-300)
    return out


def run_drift_study(args, sc):
    """Where the energy drift of the bench workload comes from: the plain form at CFL 0.3, 0.1 and 0.03 over
    about the same simulated time (steps scaled by 0.3 / CFL), with the clamp's added energy and the KE / IE /
    PE split, on the monopole tree, on the quadrupole tree and (up to 262,144 particles) on the exact sum."""
    from botheadstudios_amd import isph
    n = len(sc.particles)
    rows = []
    cases = [(0.3, "tree"), (0.1, "tree"), (0.03, "tree"), (0.3, "quadrupole"), (0.1, "quadrupole"),
             (0.03, "quadrupole")]
    if n <= 262144:  # the exact sum at this size takes seconds per step at most
        cases += [(0.3, "direct"), (0.1, "direct")]
    for cfl, grav in cases:
        with isph.GpuSph(n) as s:
            s.upload(sc.particles, sc.eos, sc.softening)
            s.set_eos_f64(sc.eos64)
            s.set_gravity(isph.GRAVITY_DIRECT if grav == "direct" else isph.GRAVITY_TREE, args.theta, 1,
                          grav == "quadrupole")
            d0 = s.diagnostics(with_pe=True)
            s.reset_stats()
            steps = int(round(args.steps * 0.3 / cfl))
            s.step_cfl(cfl, steps)
            d1 = s.diagnostics(with_pe=True)
            t, _ = s.sim_time()
            e0 = _energy(d0)
            rows.append({"cfl": cfl, "gravity": grav, "steps": steps, "sim_time_s": t,
                         "dE_rel": (_energy(d1) - e0) / abs(e0), "clamp_energy_rel": s.clamp_energy() / abs(e0),
                         "dKE_rel": (d1["ke"] - d0["ke"]) / abs(e0), "dIE_rel": (d1["ie"] - d0["ie"]) / abs(e0),
                         "dPE_rel": (d1["pe"] - d0["pe"]) / abs(e0)})
            print(json.dumps(rows[-1]), flush=True)
    return {"workload": f"two-body impact N={n}, plain form, theta={args.theta}", "runs": rows}


def _impact_gradient(fam, loss, steps, params, checkpoint_every, cfl, theta, dts=None):
    """Loss after `steps` KDK steps of the two-body scene at `params` = (angle deg, speed / v_esc, mass
    ratio), mixed form, LBVH theta; and its gradient with respect to the three parameters (the reverse pass's
    start-state gradient projected on the scene's tangents). Returns (loss, grad[3], dts, ms2)."""
    from botheadstudios_amd import isph, materials, scene
    rec, pos, vel = fam.state(*params)
    n = len(rec)
    with isph.GpuSph(n) as s:
        s.upload(rec, materials.eos_table(), fam.softening * scene.L_U)
        s.set_eos_f64(materials.eos_f64())
        s.set_precision(isph.PRECISION_MIXED)
        s.set_state_f64(pos, vel)
        s.set_gravity(isph.GRAVITY_TREE, theta, 1)
        out = s.vjp_run(loss, steps, checkpoint_every, cfl=cfl, dts=dts)
    g = np.array([(out["dpos"] * dp).sum() + (out["dvel"] * dv).sum() + (out["dmass"] * dm).sum()
                  for dp, dv, dm in fam.tangents(*params)])
    return out["loss"], g, out["dts"], out["ms"]


def fit_along_line(evaluate, p0, scale, target, iterations, first=None):
    """Fit a scalar loss to `target` over parameters p (`evaluate(p)` -> (loss, gradient)). The first step is
    Gauss-Newton on the gradient at p0, scaled by the parameters' nominal sizes `scale`: the minimum-norm step
    that moves the linearised loss by target - L(p0), p = p0 + t d with d = s^2 g / sum (s g)^2 and t = target -
    L(p0). A gradient that holds the tree walk's decisions fixed misjudges the size of the response over a
    long horizon (by tens of percent), so the later steps stay on that line and solve L(p0 + t d) = target in
    t from the measured losses alone: secant steps (at most four times the last step's length) until the
    target is bracketed, then regula falsi in the bracket with the Illinois halving. A positive loss (a mass)
    is matched in log L - log target, which makes an exponential response linear in t. Returns the history:
    params, loss and t of every evaluation."""
    l0, g0 = first if first is not None else evaluate(p0)
    p0 = np.asarray(p0, dtype=np.float64)
    scale = np.asarray(scale, dtype=np.float64)
    target = float(target)
    hist = [{"params": [float(x) for x in p0], "loss_kg": float(l0), "t": 0.0}]
    sg = scale * np.asarray(g0, dtype=np.float64)
    denom = float(sg @ sg)
    if not (denom > 0.0) or iterations < 1:
        return hist
    d = scale * sg / denom
    use_log = target > 0.0 and float(l0) > 0.0

    def resid(loss):
        return float(np.log(max(loss, 1e-12 * target) / target)) if use_log else loss - target

    r = resid(float(l0))
    lo = (0.0, r) if r < 0.0 else None  # bracket ends (t, residual) below and above the target
    hi = (0.0, r) if r > 0.0 else None
    prev, side = (0.0, r), 0
    t = target - float(l0)
    for _ in range(iterations):
        loss = float(evaluate(p0 + t * d)[0])
        hist.append({"params": [float(x) for x in p0 + t * d], "loss_kg": loss, "t": t})
        if not np.isfinite(loss):
            break
        r = resid(loss)
        if r == 0.0:
            break
        if r < 0.0:
            if side < 0 and hi is not None:
                hi = (hi[0], 0.5 * hi[1])
            lo, side = (t, r), -1
        else:
            if side > 0 and lo is not None:
                lo = (lo[0], 0.5 * lo[1])
            hi, side = (t, r), 1
        if lo is not None and hi is not None:
            t_next = lo[0] - lo[1] * (hi[0] - lo[0]) / (hi[1] - lo[1])
        else:
            step = t - prev[0]
            slope = (r - prev[1]) / step if step != 0.0 else 0.0
            move = -r / slope if slope != 0.0 else step
            t_next = t + float(np.clip(move, -4.0 * abs(step), 4.0 * abs(step)))
        prev = (t, r)
        t = t_next
    return hist


def run_adjoint_outcome(args, steps=800, fit_steps=400, iterations=3):
    """The differentiation workload where the impact has an outcome. (1) The reverse pass over `steps` KDK
    steps from the standard start (the ejecta have left 2.9 target radii by then): loss and cost. (2) A fit of
    (angle, speed, mass ratio) over `fit_steps` steps toward twice the nominal loss (`fit_along_line`: a
    Gauss-Newton step on the reverse pass's gradient, then a bracketed secant search along it), with the dt
    sequence of the first evaluation held fixed."""
    from botheadstudios_amd import isph, scene
    n = args.adjoint_n
    fam = scene.ImpactFamily(n - n // 8, n // 8, 0xC5)
    loss = isph.Loss.make(scene.L_U, r_cut=2.9, k_r=20.0, k_j=10.0)
    p0 = np.array([45.0, 1.2, 0.11])
    l_long, g_long, _, ms = _impact_gradient(fam, loss, steps, p0, args.adjoint_every, args.cfl, args.theta)
    out = {"workload": f"two-body impact N={n}, mixed form, LBVH theta={args.theta}, {steps} KDK steps at CFL "
                       f"{args.cfl}, checkpoint every {args.adjoint_every}",
           "loss_kg": l_long, "loss_over_total_mass": None, "forward_ms": float(ms[0]), "reverse_ms": float(ms[1]),
           "adjoint_over_forward": float(ms[1] / ms[0]),
           "grad_angle_speed_ratio": [float(x) for x in g_long]}
    rec, _, _ = fam.state(*p0)
    out["loss_over_total_mass"] = l_long / float(rec["mass"].astype(np.float64).sum())
    scale = np.array([45.0, 1.2, 0.11])
    l, g, dts, _ = _impact_gradient(fam, loss, fit_steps, p0, args.adjoint_every, args.cfl, args.theta)
    target = 2.0 * l

    def evaluate(q):
        lq, gq, _, _ = _impact_gradient(fam, loss, fit_steps, q, args.adjoint_every, args.cfl, args.theta, dts)
        return lq, gq

    hist = fit_along_line(evaluate, p0, scale, target, iterations, first=(l, g))
    out["fit"] = {"steps": fit_steps, "target_kg": target, "iterations": hist,
                  "final_over_target": hist[-1]["loss_kg"] / target}
    return out


def run_4m(args, steps=10):
    """Plain vs mixed throughput on the 4,194,304-particle two-body impact (the precision form's size)."""
    from botheadstudios_amd import isph, scene
    sc = scene.two_body(4194304, seed=0xC3)
    n = len(sc.particles)
    res = {}
    for name, prec in (("fp32", isph.PRECISION_FP32), ("mixed", isph.PRECISION_MIXED)):
        with isph.GpuSph(n) as s:
            s.upload(sc.particles, sc.eos, sc.softening)
            s.set_eos_f64(sc.eos64)
            s.set_precision(prec)
            s.step_cfl(args.cfl, 2)
            s.synchronize()
            t0 = time.perf_counter()
            s.step_cfl(args.cfl, steps)
            s.synchronize()
            res[name] = n * steps / (time.perf_counter() - t0)
    return {"workload": f"two-body impact N={n}, LBVH theta={args.theta}, {steps} KDK steps at CFL {args.cfl}",
            "fp32_value": res["fp32"], "mixed_value": res["mixed"], "mixed_over_fp32": res["mixed"] / res["fp32"],
            "unit": "particle-steps/s"}


def run_batch(args):
    """All instances as one ensemble (mixed form, captured step graph, autotuned walk shape) against the same
    instances run one after another in the plain form, same number of steps each."""
    import tempfile
    from botheadstudios_amd import isph
    from botheadstudios_amd import scene
    scs = scene.impact_ensemble(args.batch_count)
    total = sum(len(sc.particles) for sc in scs)
    cache = os.path.join(tempfile.gettempdir(), "isph_autotune.txt")
    with isph.GpuSph(total) as s:
        s.upload_ensemble([sc.particles for sc in scs], scs[0].eos, scs[0].softening)
        s.set_eos_f64(scs[0].eos64)
        s.set_precision(isph.PRECISION_MIXED)
        s.set_graph(True)
        from_cache, tun = s.autotune(args.cfl, 3, cache)
        s.step_cfl(args.cfl, 2)
        s.synchronize()
        t0 = time.perf_counter()
        s.step_cfl(args.cfl, args.batch_steps)
        s.synchronize()
        ens = time.perf_counter() - t0
        measure_ms, views, _ = time_disk_view(s)
    seq = 0.0
    for sc in scs:
        n = len(sc.particles)
        with isph.GpuSph(n) as s:
            s.upload(sc.particles, sc.eos, sc.softening)
            s.set_eos_f64(sc.eos64)
            s.step_cfl(args.cfl, 1)
            s.synchronize()
            t0 = time.perf_counter()
            s.step_cfl(args.cfl, args.batch_steps)
            s.synchronize()
            seq += time.perf_counter() - t0
    return {"workload": f"{len(scs)} two-body impacts, N log-uniform in [16384, 262144] (total {total}), mixed form, "
                        f"{args.batch_steps} KDK steps each at CFL {args.cfl}",
            "instances": len(scs), "particles": total, "steps": args.batch_steps,
            "value": total * args.batch_steps / ens, "unit": "particle-steps/s", "ms_per_step": ens * 1e3 / args.batch_steps,
            "sequential_plain_value": total * args.batch_steps / seq,
            "speedup_vs_sequential_plain": seq / ens,
            "disk_view_all_instances_ms": measure_ms, "disk_members_total": int(sum(v.n_disk for v in views)),
            "tuning": {"leaf_size": tun.leaf_size, "group_extent": tun.group_extent,
                       "grav_particles_per_lane": tun.grav_particles_per_lane, "sph_expand": tun.sph_expand,
                       "from_cache": from_cache}}


def _leg(fn, *a):
    """An optional leg of the line: its result, or the error it raised, so that the headline measurement is
    still printed (the traceback goes to stderr)."""
    try:
        return fn(*a)
    except Exception as e:  # noqa: BLE001 - reported in the line
        import traceback
        traceback.print_exc(file=sys.stderr)
        return {"error": f"{type(e).__name__}: {e}"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the particle step here is a single-GPU build (got --gpus {args.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.adjoint_only:
        print(json.dumps({"metric": "adjoint/forward cost", "adjoint": run_adjoint(args)}), flush=True)
        return
    from botheadstudios_amd import scene

    sc = scene.two_body_1m() if args.n == 1048576 else scene.two_body(args.n)
    n = len(sc.particles)
    if args.drift_study:
        print(json.dumps({"metric": "energy drift decomposition", "drift": run_drift_study(args, sc)}), flush=True)
        return
    r = run_gpu(sc, args, args.precision, args.steps, args.warmup, args.reuse, measure=True)
    st = r["stats"]
    launches = max(st["gravity_launches"], 1.0)
    flops = st["gravity_flops"]
    grav_s = st["gravity_ms"] / 1e3
    achieved = flops / grav_s / 1e12 if grav_s > 0 else 0.0
    value = n * args.steps / r["elapsed"]
    traffic, traffic_code = measured_traffic() if (n == 1048576 and args.precision == "fp32") else (None, None)
    line = {
        "metric": "particle-steps/s @N=1M (plain FP32 vs mixed, dE<1e-4); adjoint/forward cost",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["elapsed"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f32-pairs/f64-state",
        "data": "synthetic",
        "config": {"workload": f"two-body giant impact, N={n} ({sc.n_target} target + {sc.n_impactor} impactor), "
                               + (f"LBVH Barnes-Hut theta={args.theta}, leaf={args.leaf}, "
                                  f"{'quadrupole' if args.quadrupole else 'monopole'}, "
                                  if args.gravity == "tree" else "direct-sum gravity, ")
                               + f"SPH 48 neighbours, Tillotson granite/iron, KDK at CFL {args.cfl}",
                   "n": n, "precision": args.precision, "reuse": bool(args.reuse)},
        "roofline": {"bound": "valu", "engine": "valu-fp32", "kernel": "k_gravity_tree",
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "peak_nonpacked": FP32_NONPACKED_TFLOPS, "frac_nonpacked": achieved / FP32_NONPACKED_TFLOPS,
                     "traffic": traffic,
                     "traffic_measured_on": traffic_code,
                     "flops_per_launch": flops / launches,
                     "avg_launch_ms": st["gravity_ms"] / launches},
        "phases_ms_per_step": {k: v / args.steps for k, v in r["phases"].items()},
        "interactions_per_particle_eval": (st["cell_interactions"] + st["particle_interactions"] +
                                           st["quad_interactions"]) / (launches * n),
        "sim_time_s": r["sim_time"],
        "walk": {k: v for k, v in st.items()},
    }
    if "dE_rel" in r:
        line["dE_rel"] = r["dE_rel"]
        line["dL_rel"] = r["dL_rel"]
        line["dE_ok"] = bool(r["dE_rel"] < DE_BOUND)
        line["clamp_energy_rel"] = r["clamp_energy_rel"]
    line["measure"] = {"impact_state": r["measure"]}
    debris = _leg(run_debris, n)
    debris_rec = None
    if isinstance(debris, tuple):
        debris, debris_rec = debris
    line["measure"]["debris_field"] = debris
    variants = [x for x in args.variants.split(",") if x]
    for v in variants:
        if v not in ("mixed", "reuse"):
            sys.exit(f"unknown variant {v!r}")

    def variant(v):
        prec = "mixed" if v == "mixed" else args.precision
        rv = run_gpu(sc, args, prec, args.steps, args.warmup, args.reuse or v == "reuse")
        return {"value": n * args.steps / rv["elapsed"], "ms_per_step": rv["elapsed"] * 1e3 / args.steps,
                "ratio_to_value": r["elapsed"] / rv["elapsed"],
                "phases_ms_per_step": {k: x / args.steps for k, x in rv["phases"].items()},
                "dE_rel": rv.get("dE_rel"), "dL_rel": rv.get("dL_rel"),
                "dE_ok": None if rv.get("dE_rel") is None else bool(rv["dE_rel"] < DE_BOUND)}

    for v in variants:
        line[v] = _leg(variant, v)
    if not args.no_adjoint:
        line["adjoint"] = _leg(run_adjoint, args)
        if "error" not in line["adjoint"]:
            line["adjoint"]["outcome"] = _leg(run_adjoint_outcome, args)
    if not args.no_batch:
        line["batch"] = _leg(run_batch, args)
    if not args.no_4m:
        line["precision_4m"] = _leg(run_4m, args)
    if not args.no_precision_study:
        line["precision_study"] = _leg(run_precision_study, args, sc)
    if not args.no_cpu:
        line["cpu_baseline"] = _leg(cpu_baseline, sc)
        if "error" not in line["cpu_baseline"]:
            # the roofline on the reference walk's interaction count (its per-particle MAC needs fewer
            # interactions than the group MAC, which accepts a node only for the whole walk group)
            rw = line["cpu_baseline"]["reference_walk"]
            ref_inter = rw["cells_per_particle"] + rw["particles_per_particle"]
            walk_inter = line["interactions_per_particle_eval"]
            rf = line["roofline"]
            rf["reference_walk_interactions_per_particle"] = ref_inter
            rf["walk_interactions_per_particle"] = walk_inter
            rf["achieved_reference_walk"] = rf["achieved"] * ref_inter / walk_inter
            rf["frac_reference_walk"] = rf["achieved_reference_walk"] / FP32_PEAK_TFLOPS
        if r["measure_rec"] is not None:
            line["measure"]["impact_state"]["cpu_oracle_ms"] = _leg(measure_cpu, r["measure_rec"])
        if debris_rec is not None:
            line["measure"]["debris_field"]["cpu_oracle_ms"] = _leg(measure_cpu, debris_rec)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
