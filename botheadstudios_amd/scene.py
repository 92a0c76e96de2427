"""Seeded two-body giant-impact scenes in the reference's `SphParticle` layout.

The workload: a target and an impactor, each a 30 % iron core / 70 % granite mantle body filled on a
seeded jittered cubic lattice (jitter uniform in +/-0.1 spacing), equal particle mass within a body,
smoothing length set for 48 neighbours inside the kernel support (support radius h, q < 1 as in
`sph_step.wgsl:73-86`), u0 = 1e-3, Plummer softening 0.01 and the impactor placed the way the reference's
offline driver places it (`tools/impact-run/src/main.rs:548-564`): 1.6x contact distance along +x,
impact parameter (R_t + R_i) sin(angle) along +y, velocity -v x with v a multiple of the mutual escape
speed sqrt(2 G (M_t + M_i) / (R_t + R_i)).

Sizes are given in G = 1 code units (R_target = 1, M_target = 1) and converted to SI with a fixed unit
map, because the kernels carry the reference's SI constants (G = 6.674e-11 and the SI epsilons of
`sph_step.wgsl:88,100,272,370`):
  L_u = 6.371e6 m, rho_u = 1 / (0.3/7850 + 0.7/2680) kg/m^3, M_u = rho_u 4/3 pi L_u^3,
  T_u = sqrt(L_u^3 / (G M_u)), V_u = L_u / T_u, specific energy unit V_u^2.

Random numbers come from splitmix64 (the generator of `bhtree.rs:280-286`), drawn as a counter stream so
the numpy form equals sequential calls.
"""
from dataclasses import dataclass

import numpy as np

from . import materials

G_SI = 6.674e-11  # orbit.rs:18
L_U = 6.371e6
RHO_U = 1.0 / (0.3 / materials.IRON[0] + 0.7 / materials.GRANITE[0])
M_U = RHO_U * 4.0 / 3.0 * np.pi * L_U ** 3
T_U = float(np.sqrt(L_U ** 3 / (G_SI * M_U)))
V_U = L_U / T_U
E_U = V_U * V_U

PARTICLE_DTYPE = np.dtype([
    ("pos", "<f4", (3,)), ("h", "<f4"),
    ("vel", "<f4", (3,)), ("u", "<f4"),
    ("mass", "<f4"), ("mat", "<u4"), ("rho", "<f4"), ("prov", "<u4"),
])
assert PARTICLE_DTYPE.itemsize == 48

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """splitmix64 stream producing doubles in [0, 1) exactly like `bhtree.rs:280-286`."""

    def __init__(self, seed):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def uniform(self, k):
        with np.errstate(over="ignore"):
            steps = np.arange(1, k + 1, dtype=np.uint64)
            z = self.state + steps * _GOLD
            self.state = self.state + np.uint64(k) * _GOLD
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return (z >> np.uint64(11)).astype(np.float64) / float(1 << 53)


@dataclass
class Scene:
    """A generated scene: the 48-byte particle records plus the f64 state they were rounded from."""
    particles: np.ndarray      # PARTICLE_DTYPE, SI, f32
    pos64: np.ndarray          # (N,3) f64 SI
    vel64: np.ndarray          # (N,3) f64 SI
    mass64: np.ndarray
    h64: np.ndarray
    u64: np.ndarray
    mat: np.ndarray
    eos: np.ndarray            # EOS_DTYPE table (granite, iron)
    eos64: np.ndarray          # (n_mat, 10) f64
    softening: float           # SI
    n_target: int
    n_impactor: int


def _lattice_region(n, r_in, r_out, rng):
    """n points of a jittered cubic lattice inside the shell r_in <= |p| < r_out (code units)."""
    vol = 4.0 / 3.0 * np.pi * (r_out ** 3 - r_in ** 3)
    s = (vol / n) ** (1.0 / 3.0)
    while True:
        m = int(np.ceil(r_out / s)) + 1
        ax = (np.arange(-m, m, dtype=np.float64) + 0.5) * s
        gx, gy, gz = np.meshgrid(ax, ax, ax, indexing="ij")
        pts = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
        jit = (rng.uniform(pts.size).reshape(pts.shape) * 2.0 - 1.0) * (0.1 * s)
        pts = pts + jit
        r = np.sqrt((pts * pts).sum(axis=1))
        keep = pts[(r >= r_in) & (r < r_out)]
        if keep.shape[0] >= n:
            break
        s *= 0.99
    key = rng.uniform(keep.shape[0])
    sel = np.argsort(key, kind="stable")[:n]
    return keep[sel], s


def make_body(n, radius, mass, core_frac, seed, nbr=48.0):
    """A differentiated body in code units: (pos, mass_i, h, mat) with the core first."""
    rng = SplitMix64(seed)
    rho_fe, rho_gr = materials.IRON[0], materials.GRANITE[0]
    vfrac_core = (core_frac / rho_fe) / (core_frac / rho_fe + (1.0 - core_frac) / rho_gr)
    r_core = radius * vfrac_core ** (1.0 / 3.0)
    n_core = int(round(core_frac * n))
    n_mantle = n - n_core
    m_i = mass / n
    v_core = 4.0 / 3.0 * np.pi * r_core ** 3
    v_mantle = 4.0 / 3.0 * np.pi * radius ** 3 - v_core
    rho_core = core_frac * mass / v_core
    rho_mantle = (1.0 - core_frac) * mass / v_mantle
    p_core, _ = _lattice_region(n_core, 0.0, r_core, rng)
    p_mantle, _ = _lattice_region(n_mantle, r_core, radius, rng)
    pos = np.concatenate([p_core, p_mantle])
    h_core = (3.0 * nbr * m_i / (4.0 * np.pi * rho_core)) ** (1.0 / 3.0)
    h_mantle = (3.0 * nbr * m_i / (4.0 * np.pi * rho_mantle)) ** (1.0 / 3.0)
    h = np.concatenate([np.full(n_core, h_core), np.full(n_mantle, h_mantle)])
    mat = np.concatenate([np.full(n_core, materials.MAT_IRON, np.uint32),
                          np.full(n_mantle, materials.MAT_GRANITE, np.uint32)])
    rho = np.concatenate([np.full(n_core, rho_core), np.full(n_mantle, rho_mantle)])
    pos -= pos.mean(axis=0)  # equal masses: COM-centred
    return pos, np.full(n, m_i), h, mat, rho


def impact_scene(n_target, n_impactor, seed, r_t=1.0, m_t=1.0, r_i=0.5, m_i=0.11, core_t=0.3, core_i=0.3,
                 angle_deg=45.0, v_over_vesc=1.2, u0=1e-3, softening=0.01):
    """The two-body impact scene, converted to SI and rounded into `SphParticle` records."""
    pt, mt, ht, matt, rhot = make_body(n_target, r_t, m_t, core_t, seed ^ 0x7A1)
    pi_, mi_, hi_, mati, rhoi = make_body(n_impactor, r_i, m_i, core_i, seed ^ 0x1A2)
    contact = r_t + r_i
    v = v_over_vesc * np.sqrt(2.0 * (m_t + m_i) / contact)
    b = contact * np.sin(np.deg2rad(angle_deg))
    pi_ = pi_ + np.array([1.6 * contact, b, 0.0])
    pos = np.concatenate([pt, pi_])
    vel = np.zeros_like(pos)
    vel[n_target:, 0] = -v
    mass = np.concatenate([mt, mi_])
    h = np.concatenate([ht, hi_])
    mat = np.concatenate([matt, mati])
    rho = np.concatenate([rhot, rhoi])
    n = n_target + n_impactor
    u = np.full(n, u0)

    pos_si = pos * L_U
    vel_si = vel * V_U
    mass_si = mass * M_U
    h_si = h * L_U
    u_si = u * E_U
    rho_si = rho * (M_U / L_U ** 3)

    rec = np.zeros(n, dtype=PARTICLE_DTYPE)
    rec["pos"] = pos_si.astype(np.float32)
    rec["vel"] = vel_si.astype(np.float32)
    rec["h"] = h_si.astype(np.float32)
    rec["u"] = u_si.astype(np.float32)
    rec["mass"] = mass_si.astype(np.float32)
    rec["mat"] = mat
    rec["rho"] = rho_si.astype(np.float32)
    rec["prov"] = np.concatenate([np.zeros(n_target, np.uint32), np.ones(n_impactor, np.uint32)])
    return Scene(particles=rec, pos64=pos_si, vel64=vel_si, mass64=mass_si, h64=h_si, u64=u_si, mat=mat,
                 eos=materials.eos_table(), eos64=materials.eos_f64(), softening=softening * L_U,
                 n_target=n_target, n_impactor=n_impactor)


# The workloads, named by what they are.
def two_body_8k(seed=0xC1):
    """7168-particle target + 1024-particle impactor (N = 8192)."""
    return impact_scene(7168, 1024, seed)


def two_body_1m(seed=0xC2):
    """917,504-particle target + 131,072-particle impactor (N = 1,048,576)."""
    return impact_scene(917504, 131072, seed)


def two_body(n, seed=0xC5):
    """A scaled two-body scene with the same 7:1 target:impactor split as the 8k and 1M scenes."""
    n_i = n // 8
    return impact_scene(n - n_i, n_i, seed)


def impact_ensemble(count=64, seed=0xC4):
    """The batching workload: `count` independent impacts with N log-uniform in [16384, 262144] (7:1
    target:impactor), impact angle U(30, 60) deg, speed U(1.0, 1.4) v_esc, impactor / target mass ratio
    U(0.05, 0.15) and target core mass fraction U(0.25, 0.35); seeded (numpy PCG64 for the draws, each
    instance's lattice from its own splitmix64 seed)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(16384), np.log(262144), count)).astype(np.int64)
    out = []
    for k in range(count):
        n = int(sizes[k])
        ni = n // 8
        out.append(impact_scene(n - ni, ni, seed + 1 + k, m_i=float(rng.uniform(0.05, 0.15)),
                                core_t=float(rng.uniform(0.25, 0.35)), core_i=0.3,
                                angle_deg=float(rng.uniform(30.0, 60.0)), v_over_vesc=float(rng.uniform(1.0, 1.4))))
    return out


def f64_state(rec):
    """Widen `SphParticle` records to the f64 arrays the CPU path takes (bit-identical start state)."""
    return (rec["pos"].astype(np.float64), rec["vel"].astype(np.float64), rec["mass"].astype(np.float64),
            rec["u"].astype(np.float64), rec["h"].astype(np.float64), rec["mat"].astype(np.uint32))


class ImpactFamily:
    """The two-body scene as a smooth function of the impact parameters, for the differentiable step.

    Bodies are generated once (as in `impact_scene`); `state(angle_deg, v_over_vesc, mass_ratio)` places the
    impactor and sets its velocity and particle masses, and `tangents(...)` gives the derivative of the f64
    SI state with respect to each parameter. Smoothing lengths stay those of the nominal impactor (mass
    ratio 0.11), so that the state is smooth in the mass ratio. At the nominal parameters the state equals
    `impact_scene`'s.
    """

    def __init__(self, n_target, n_impactor, seed, r_t=1.0, m_t=1.0, r_i=0.5, m_i=0.11, core=0.3, u0=1e-3,
                 softening=0.01, separation=1.6):
        self.pt, self.mt, self.ht, self.matt, _ = make_body(n_target, r_t, m_t, core, seed ^ 0x7A1)
        self.pi, _, self.hi, self.mati, _ = make_body(n_impactor, r_i, m_i, core, seed ^ 0x1A2)
        self.r_t, self.m_t, self.r_i = r_t, m_t, r_i
        self.n_target, self.n_impactor = n_target, n_impactor
        self.u0, self.softening = u0, softening
        self.separation = separation  # start distance along x in contact distances (impact-run: 1.6)

    def _v(self, speed, ratio):
        return speed * np.sqrt(2.0 * (self.m_t + ratio * self.m_t) / (self.r_t + self.r_i))

    def arrays(self, angle_deg=45.0, v_over_vesc=1.2, mass_ratio=0.11):
        """The f64 SI state: pos (n,3), vel (n,3), mass, h, u, mat."""
        contact = self.r_t + self.r_i
        b = contact * np.sin(np.deg2rad(angle_deg))
        pos = np.concatenate([self.pt, self.pi + np.array([self.separation * contact, b, 0.0])]) * L_U
        vel = np.zeros_like(pos)
        vel[self.n_target:, 0] = -self._v(v_over_vesc, mass_ratio) * V_U
        mass = np.concatenate([self.mt, np.full(self.n_impactor, mass_ratio * self.m_t / self.n_impactor)]) * M_U
        h = np.concatenate([self.ht, self.hi]) * L_U
        mat = np.concatenate([self.matt, self.mati])
        u = np.full(len(mass), self.u0 * E_U)
        return pos, vel, mass, h, u, mat

    def state(self, angle_deg=45.0, v_over_vesc=1.2, mass_ratio=0.11):
        """(records, pos, vel): the 48-byte records (f32) and the f64 position / velocity state."""
        pos, vel, mass, h, u, mat = self.arrays(angle_deg, v_over_vesc, mass_ratio)
        n = len(mass)
        rec = np.zeros(n, dtype=PARTICLE_DTYPE)
        rec["pos"] = pos.astype(np.float32)
        rec["vel"] = vel.astype(np.float32)
        rec["h"] = h.astype(np.float32)
        rec["u"] = np.float32(self.u0 * E_U)
        rec["mass"] = mass.astype(np.float32)
        rec["mat"] = mat
        rec["prov"] = np.concatenate([np.zeros(self.n_target, np.uint32), np.ones(self.n_impactor, np.uint32)])
        return rec, pos, vel

    def tangents(self, angle_deg=45.0, v_over_vesc=1.2, mass_ratio=0.11):
        """d(pos, vel, mass)/d(angle in degrees, v_over_vesc, mass_ratio), each (n,3), (n,3), (n,)."""
        n, nt = self.n_target + self.n_impactor, self.n_target
        contact = self.r_t + self.r_i
        out = []
        dpos = np.zeros((n, 3))
        dpos[nt:, 1] = contact * np.cos(np.deg2rad(angle_deg)) * np.pi / 180.0 * L_U
        out.append((dpos, np.zeros((n, 3)), np.zeros(n)))
        dvel = np.zeros((n, 3))
        dvel[nt:, 0] = -self._v(1.0, mass_ratio) * V_U
        out.append((np.zeros((n, 3)), dvel, np.zeros(n)))
        dvel = np.zeros((n, 3))
        dvel[nt:, 0] = -v_over_vesc * (self.m_t / contact) / np.sqrt(2.0 * (self.m_t + mass_ratio * self.m_t) / contact) * V_U
        dm = np.zeros(n)
        dm[nt:] = self.m_t / self.n_impactor * M_U
        out.append((np.zeros((n, 3)), dvel, dm))
        return out


def debris_field(n, seed=0xD15C, clump_size=50):
    """A synthetic post-impact field for the measurement workloads (SI, granite): a central body holding
    95 % of the mass in 80 % of the particles, and the rest as an orbiting debris disk (a third of it in
    cold clumps of `clump_size` particles on circular orbits, the rest diffuse and near-circular), plus
    plunging and escaping debris. Returns the particle records."""
    rng = np.random.default_rng(seed)
    R, M = L_U, 6.0e24
    mu = G_SI * M
    n_core = int(0.8 * n)
    n_deb = n - n_core
    m_core = 0.95 * M / n_core
    m_deb = 0.05 * M / n_deb
    pos = np.empty((n, 3))
    vel = np.empty((n, 3))
    d = rng.normal(size=(n_core, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pos[:n_core] = d * (R * np.cbrt(rng.random(n_core)))[:, None]
    vel[:n_core] = rng.normal(scale=30.0, size=(n_core, 3))
    n_cl = (n_deb // 3) // clump_size
    k = n_core
    golden = np.pi * (3.0 - np.sqrt(5.0))
    i = np.arange(clump_size, dtype=np.float64)
    rr = 0.02 * R * np.cbrt((i + 0.5) / clump_size)
    y = 1.0 - 2.0 * (i + 0.5) / clump_size
    shell = np.stack([np.cos(golden * i) * np.sqrt(1 - y * y) * rr, y * rr, np.sin(golden * i) * np.sqrt(1 - y * y) * rr], 1)
    r_c = rng.uniform(2.0 * R, 8.0 * R, n_cl)
    ph = rng.uniform(0, 2 * np.pi, n_cl)
    for c in range(n_cl):
        ctr = np.array([r_c[c] * np.cos(ph[c]), r_c[c] * np.sin(ph[c]), 0.0])
        vc = np.sqrt(mu / r_c[c]) * np.array([-np.sin(ph[c]), np.cos(ph[c]), 0.0])
        pos[k:k + clump_size] = ctr + shell
        vel[k:k + clump_size] = vc
        k += clump_size
    rest = n - k
    kind = rng.random(rest)  # 0.8 diffuse disk, 0.1 plunging, 0.1 escaping
    r = rng.uniform(1.5 * R, 8.0 * R, rest)
    ph = rng.uniform(0, 2 * np.pi, rest)
    pos[k:] = np.stack([r * np.cos(ph), r * np.sin(ph), rng.normal(scale=0.05, size=rest) * r], 1)
    speed = np.sqrt(mu / r) * np.where(kind < 0.8, rng.uniform(0.95, 1.1, rest), np.where(kind < 0.9, 0.2, 2.0))
    vel[k:] = np.stack([-speed * np.sin(ph), speed * np.cos(ph), np.zeros(rest)], 1)
    mass = np.concatenate([np.full(n_core, m_core), np.full(n - n_core, m_deb)])
    h = np.concatenate([np.full(n_core, 2.0 * np.cbrt(m_core / 3000.0)), np.full(n - n_core, 2.0 * np.cbrt(m_deb / 2700.0))])
    rec = np.zeros(n, dtype=PARTICLE_DTYPE)
    perm = rng.permutation(n)
    rec["pos"] = pos[perm]
    rec["vel"] = vel[perm]
    rec["mass"] = mass[perm]
    rec["h"] = h[perm]
    rec["u"] = 1.0e5
    rec["rho"] = 2700.0
    rec["prov"] = (rng.random(n) < 0.3).astype(np.uint32)
    return rec
