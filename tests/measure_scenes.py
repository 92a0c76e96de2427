"""Inputs for the measurement tests: the reference's clump fixtures and a synthetic post-impact field.

`cold_blob` / `spinning_blob` follow the fixtures of accretion.rs:540-562 and :962-974 (golden-angle
spiral points on shells, cold or rigidly rotating). `disk_field` builds a particle field with a central
body, an orbiting disk holding a few tight cold clumps, plunging and escaping matter, and retired (mass 0)
and non-finite particles, as 48-byte particle records.
"""
import numpy as np

from botheadstudios_amd import scene

G = 6.674e-11  # orbit.rs:18


def cold_blob(center, bulk, radius, n, m_i, rho):
    golden = np.pi * (3.0 - np.sqrt(5.0))
    i = np.arange(n, dtype=np.float64)
    rr = radius * np.cbrt((i + 0.5) / n)
    y = 1.0 - 2.0 * (i + 0.5) / n
    rad = np.sqrt(np.maximum(1.0 - y * y, 0.0))
    th = golden * i
    p = np.asarray(center, np.float64) + np.stack([np.cos(th) * rad * rr, y * rr, np.sin(th) * rad * rr], 1)
    v = np.tile(np.asarray(bulk, np.float64), (n, 1))
    return p, v, np.full(n, float(m_i)), np.full(n, float(rho))


def spinning_blob(center, bulk, omega, radius, n, m_i, rho):
    p, _, m, r = cold_blob(center, bulk, radius, n, m_i, rho)
    v = np.asarray(bulk, np.float64) + np.cross(np.asarray(omega, np.float64), p - np.asarray(center, np.float64))
    return p, v, m, r


def records(pos, vel, mass, h, rho, prov, u=1.0e5):
    rec = np.zeros(len(mass), dtype=scene.PARTICLE_DTYPE)
    rec["pos"] = pos
    rec["vel"] = vel
    rec["mass"] = mass
    rec["h"] = h
    rec["rho"] = rho
    rec["prov"] = prov
    rec["u"] = u
    return rec


def disk_field(seed, n_core=3000, n_disk=1200, n_clumps=6, clump_n=40, n_plunge=300, n_esc=200, n_dead=7,
               offset=(0.0, 0.0, 0.0), drift=(0.0, 0.0, 0.0)):
    """A central ball (R = 6.4e6 m, M = 6e24 kg) with orbiting, clumped, plunging and escaping matter."""
    rng = np.random.default_rng(seed)
    R, M = 6.4e6, 6.0e24
    mu = G * M
    pos, vel, mass, h, rho, prov = [], [], [], [], [], []

    def add(p, v, m, hh, rr, pv):
        pos.append(p), vel.append(v), mass.append(m), h.append(hh), rho.append(rr), prov.append(pv)

    # central body: uniform ball, slow random motion
    d = rng.normal(size=(n_core, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    p = d * (R * np.cbrt(rng.random(n_core)))[:, None]
    add(p, rng.normal(scale=30.0, size=(n_core, 3)), np.full(n_core, M / n_core), np.full(n_core, 3.0e5),
        rng.uniform(3000, 5000, n_core), np.zeros(n_core, np.uint32))
    m_small = 1.0e20

    def ring(n, rmin, rmax, speed_scale, thick=0.05):
        r = rng.uniform(rmin, rmax, n)
        phi = rng.uniform(0, 2 * np.pi, n)
        z = rng.normal(scale=thick, size=n) * r
        p = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
        vc = np.sqrt(mu / r) * speed_scale
        v = np.stack([-vc * np.sin(phi), vc * np.cos(phi), np.zeros(n)], 1)
        return p, v + rng.normal(scale=50.0, size=(n, 3))

    # orbiting disk (bound, perigee outside the remnant)
    p, v = ring(n_disk, 2.0 * R, 6.0 * R, rng.uniform(0.95, 1.1, n_disk))
    add(p, v, np.full(n_disk, m_small), np.full(n_disk, 1.2e5), rng.uniform(2000, 3500, n_disk),
        rng.integers(0, 2, n_disk).astype(np.uint32))
    # cold clumps on circular orbits in the disk
    for k in range(n_clumps):
        r = (2.5 + 0.6 * k) * R
        phi = 2 * np.pi * k / n_clumps
        c = np.array([r * np.cos(phi), r * np.sin(phi), 0.0])
        vc = np.sqrt(mu / r) * np.array([-np.sin(phi), np.cos(phi), 0.0])
        bp, bv, bm, br = cold_blob(c, vc, 1.5e5, clump_n, 5.0e20 * (1 + k), 3000.0)
        add(bp, bv, bm, np.full(clump_n, 1.0e5), br, np.full(clump_n, k % 2, np.uint32))
    # plunging (bound, perigee inside the remnant)
    p, v = ring(n_plunge, 1.5 * R, 4.0 * R, 0.2)
    add(p, v, np.full(n_plunge, m_small), np.full(n_plunge, 1.2e5), np.full(n_plunge, 2500.0),
        rng.integers(0, 2, n_plunge).astype(np.uint32))
    # escaping
    p, v = ring(n_esc, 2.0 * R, 5.0 * R, 2.0)
    add(p, v, np.full(n_esc, m_small), np.full(n_esc, 1.2e5), np.full(n_esc, 2500.0),
        rng.integers(0, 2, n_esc).astype(np.uint32))
    pos, vel = np.concatenate(pos), np.concatenate(vel)
    mass, h, rho, prov = map(np.concatenate, (mass, h, rho, prov))
    pos = pos + np.asarray(offset)
    vel = vel + np.asarray(drift)
    perm = rng.permutation(len(mass))
    rec = records(pos[perm], vel[perm], mass[perm], h[perm], rho[perm], prov[perm])
    # retired and non-finite particles are not matter (gpu_sph.rs:697-707)
    dead = rng.choice(len(rec), n_dead, replace=False)
    rec["mass"][dead[:3]] = 0.0
    rec["pos"][dead[3], 1] = np.nan
    rec["vel"][dead[4], 2] = np.inf
    rec["mass"][dead[5]] = np.nan
    rec["pos"][dead[6], 0] = -np.inf
    return rec
