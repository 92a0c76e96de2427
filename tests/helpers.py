"""Shared generators (the reference tests' own seeds and RNGs) and comparison metrics."""
import numpy as np

from botheadstudios_amd import materials, scene

G = 6.674e-11


def splitmix_stream(seed):
    """bhtree.rs:280-286 / neighbors.rs:98-104 splitmix64 -> [0,1) doubles, sequential."""
    rng = scene.SplitMix64(seed)
    return lambda k: rng.uniform(k)


def splitmix_cloud(n, seed, box=1.0e6):
    """gpu-bh-verify build_cloud (main.rs:124-137): per body x, y, z then its mass."""
    u = splitmix_stream(seed)(4 * n).reshape(n, 4)
    pos = u[:, :3] * box
    mass = 1.0e18 * (0.5 + u[:, 3])
    return pos, mass


def splitmix_cloud_bhtree(n, seed, box=1.0e6):
    """bhtree.rs:359-364: all positions (x, y, z per body) first, then all masses."""
    u = splitmix_stream(seed)
    pos = u(3 * n).reshape(n, 3) * box
    mass = 1.0e18 * (0.5 + u(n))
    return pos, mass


def xorshift_debris(n, seed):
    """gpu_gravity.rs:751-768 debris_cloud: xorshift64, normalize(rng3) * 1.7e6 * (0.2 + |rng|)."""
    s = seed & 0xFFFFFFFFFFFFFFFF
    out = np.zeros((n, 3))

    def rng():
        nonlocal s
        s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
        s ^= s >> 7
        s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
        return (s >> 40) / float(1 << 24) - 0.5

    for i in range(n):
        v = np.array([rng(), rng(), rng()])
        nv = np.linalg.norm(v)
        v = v / nv if nv > 0 else np.zeros(3)
        out[i] = v * (1.7e6 * (0.2 + abs(rng())))
    return out, np.full(n, 5.0e19)


def particles_from(pos, mass, h=1.0, u=0.0, vel=None, mat=0):
    n = len(mass)
    rec = np.zeros(n, dtype=scene.PARTICLE_DTYPE)
    rec["pos"] = np.asarray(pos, np.float32)
    rec["mass"] = np.asarray(mass, np.float32)
    rec["h"] = h
    rec["u"] = u
    rec["mat"] = mat
    if vel is not None:
        rec["vel"] = np.asarray(vel, np.float32)
    return rec


def rms_rel(a, b):
    """gpu-bh-verify compare (main.rs:163-180): ||sum err|| / ||sum ref|| and the worst particle."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    e = ((a - b) ** 2).sum(axis=1)
    r = (b ** 2).sum(axis=1)
    return float(np.sqrt(e.sum() / max(r.sum(), 1e-300))), float(np.max(np.sqrt(e) / np.maximum(np.sqrt(r), 1e-30)))


def rel_err_floor(x, ref, floor_frac=1e-3):
    """Per-particle relative error with denominator max(|ref|, floor_frac * RMS(|ref|)) (vectors or scalars)."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    if x.ndim == 1:
        x = x[:, None]
        ref = ref[:, None]
    num = np.sqrt(((x - ref) ** 2).sum(axis=1))
    mag = np.sqrt((ref ** 2).sum(axis=1))
    den = np.maximum(mag, floor_frac * np.sqrt(np.mean(mag ** 2)))
    den = np.maximum(den, 1e-300)
    return num / den


def with_random_velocities(sc, amp_frac=0.05, seed=0x5EED):
    """Add seeded random velocities (a fraction of the target sound speed scale) so the artificial
    viscosity and du/dt are exercised (sph-verify build_config gives +/-200 m/s, main.rs:279)."""
    rng = scene.SplitMix64(seed)
    v = (rng.uniform(3 * len(sc.particles)).reshape(-1, 3) * 2 - 1) * 300.0
    rec = sc.particles.copy()
    rec["vel"] = (rec["vel"].astype(np.float64) + v).astype(np.float32)
    return rec


def hydro_from(rec, eos64, softening, gravity="none"):
    from oracle import oracle
    pos, vel, mass, u, h, mat = scene.f64_state(rec)
    hb = oracle.Hydro(pos, vel, mass, u, h, mat, eos64, softening)
    hb.configure(theta=0.5, av_alpha=1.0, av_beta=2.0, gravity=gravity)
    return hb


def new_sphere(eos10, total_mass, temp_k, specific_heat, n):
    """HydroBody::new_sphere (hydrostatic.rs:78-101): n equal-mass particles on a Fibonacci-sphere spiral at
    uniform number density, u = c T, h = smoothing_for(m, rho0), softening half a mean spacing, rho = rho0.
    Returns (pos, vel, mass, u, h, softening) in f64."""
    from botheadstudios_amd import bodies
    rho0 = eos10[0]
    r0 = np.cbrt(total_mass / (4.0 / 3.0 * np.pi * rho0))
    m_i = total_mass / n
    i = np.arange(n, dtype=np.float64)
    pos = bodies.fib_dir(i, n, 0.0) * (r0 * np.cbrt((i + 0.5) / n))[:, None]
    h = np.full(n, bodies.smoothing_for(m_i, rho0))
    return (pos, np.zeros((n, 3)), np.full(n, m_i), np.full(n, specific_heat * temp_k), h,
            0.5 * np.cbrt(m_i / rho0))


def head_on_scene():
    """The initial state of `a_head_on_collision_conserves_energy_and_shock_heats` (hydrostatic.rs:1354-1396):
    two identical 600-particle basalt spheres of radius 400 km at 300 K, each relaxed for 1500 damped steps
    (dt = relax_dt(0.2), damping 0.94) on the f64 CPU step, then placed at x = -/+2.2 r0 with their x
    velocities set to +/-1.5 km/s (their relaxed y, z velocities kept). Both bodies are built and relaxed
    identically, so one relaxation serves both. Returns (pos, vel, mass, u, h, mat, eos64, softening)."""
    from botheadstudios_amd import materials
    from oracle import oracle
    eos = np.array([materials.BASALT], np.float64)
    r0 = 4.0e5
    m_body = 4.0 / 3.0 * np.pi * r0 ** 3 * eos[0][0]
    pos, vel, mass, u, h, soft = new_sphere(eos[0], m_body, 300.0, materials.SPECIFIC_HEAT["basalt"], 600)
    mat = np.zeros(600, np.uint32)
    hb = oracle.Hydro(pos, vel, mass, u, h, mat, eos, soft)
    # relax_dt (hydrostatic.rs:483-492): cfl * min h / max(c(rho0, u0), 1)
    dt_relax = 0.2 * h.min() / max(np.sqrt(oracle.till_c2(eos[0], eos[0][0], u[0])), 1.0)
    hb.relax(dt_relax, 0.94, 1500)
    p, v, _, _ = hb.state()
    hb.close()
    sep, v_app = 2.2 * r0, 1500.0
    pa, va = p.copy(), v.copy()
    pa[:, 0] -= sep
    va[:, 0] = v_app
    pb, vb = p.copy(), v.copy()
    pb[:, 0] += sep
    vb[:, 0] = -v_app
    cat = np.concatenate
    return (cat([pa, pb]), cat([va, vb]), cat([mass, mass]), cat([u, u]), cat([h, h]), cat([mat, mat]), eos, soft)
