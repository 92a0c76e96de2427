"""Layered bodies, their particalization and the N-body SPH assembly: the scene builders in front of the
particle step, on the host in f64 as the engine builds them on the CPU.

  planet.rs:26-130         LayeredBody: layers, radius, temperature_at, enclosed / total mass
  hydrostatic.rs:32-38     fib_dir; :71-73 smoothing_for; :211-264 particalize; :483-492 relax_dt
  hydrostatic.rs:57-68     Cap; :291-335 particalize_cap (a surface patch of the planet, the rest a bulk)
  eos.rs:52-96             Tillotson pressure and the central-difference sound speed (relax_dt needs it)
  gpu_sph.rs:98-113        SphEos::from_tillotson; :1010-1060 SphAssembly (materials deduped by rho0)
  gpu_sph.rs:264-292       build_far_apart_n; :419-441 far_apart_pair
  gpu_sph.rs:451-503       body_bulk (robust centre, mass and 98 %-mass radius)
  gpu_sph.rs:505-549       BodyPlacement, assemble_from_relaxed_n
  gpu_sph.rs:602-653       assemble_from_relaxed_at / assemble_from_relaxed_with
  gpu_sph.rs:293-330       build_cap_relax; :551-591 assemble_cap_impact
  accretion.rs:1418-1423   resolution_distance (the 1 % tidal hand-off distance)
  terra/world_def.rs:317-420  ImpactDef and its defaults (the Birth-of-the-Moon world, worlds/birth/world.json)

The body definitions are the layer tables (material, outer radius, in-situ density, geotherm) of the
engine's assets/bodies/{proto-earth,theia,earth,moon}.json. `birth_of_the_moon` and `moon_drop` run the
engine's two whole-body collision scenes end to end on the device stepper: staging far apart, damped
relaxation with artificial viscosity off, assembly on the collision geometry, the impact, and the on-device
disk measurement. `moon_on_earth_cap` runs the engine's resolution-on-demand branch (lib.rs:1950-1996,
3711-3745): only a cap of the Earth around the impact site is particles, the rest is the abstract bulk.
"""
from dataclasses import dataclass, field

import numpy as np

from . import materials
from .scene import PARTICLE_DTYPE

G = 6.674e-11  # orbit.rs:18
FOUR_THIRDS_PI = 4.0 / 3.0 * np.pi
MAX_MATERIALS = 16  # gpu_sph.rs:156


# ---- bodies (planet.rs) --------------------------------------------------------------------------------
@dataclass(frozen=True)
class Layer:
    material: str
    outer_r: float   # m
    density: float   # in-situ mean density, kg/m^3
    t_inner: float   # K at the inner boundary
    t_outer: float   # K at the outer boundary


@dataclass
class LayeredBody:
    layers: list
    name: str = ""

    def radius(self):  # planet.rs:86-88
        return self.layers[-1].outer_r if self.layers else 0.0

    def temperature_at(self, r):  # planet.rs:99-109
        inner = 0.0
        for l in self.layers:
            if r <= l.outer_r:
                f = min(max((r - inner) / max(l.outer_r - inner, 1.0), 0.0), 1.0)
                return l.t_inner + (l.t_outer - l.t_inner) * f
            inner = l.outer_r
        return self.layers[-1].t_outer if self.layers else 0.0

    def enclosed_mass(self, r):  # planet.rs:112-127
        m, inner = 0.0, 0.0
        for l in self.layers:
            ro = min(l.outer_r, r)
            if ro > inner:
                m += l.density * (4.0 / 3.0) * np.pi * (ro ** 3 - inner ** 3)
            inner = l.outer_r
            if inner >= r:
                break
        return m

    def total_mass(self):  # planet.rs:129-131
        return self.enclosed_mass(self.radius())


# assets/bodies/*.json layer tables: (material, outer_r m, density kg/m^3, t_inner K, t_outer K)
_BODY_LAYERS = {
    "proto-earth": [("iron", 1181276.0, 12900.0, 5900.0, 5800.0), ("iron", 2272614.0, 11900.0, 5800.0, 5150.0),
                    ("iron", 3365404.0, 10600.0, 5150.0, 4150.0),
                    ("peridotite", 6137026.0, 4500.0, 3200.0, 2200.0),
                    ("peridotite", 6161203.0, 2900.0, 2200.0, 1900.0)],
    "theia": [("iron", 1800000.0, 7200.0, 2600.0, 2400.0), ("peridotite", 3390000.0, 3400.0, 1800.0, 300.0)],
    "earth": [("iron", 1221500.0, 12900.0, 5900.0, 5800.0), ("iron", 2350000.0, 11900.0, 5800.0, 5150.0),
              ("iron", 3480000.0, 10600.0, 5150.0, 4150.0), ("peridotite", 6346000.0, 4500.0, 2900.0, 1600.0),
              ("basalt", 6371000.0, 2900.0, 900.0, 288.0)],
    "moon": [("iron", 240000.0, 7800.0, 1800.0, 1820.0), ("iron", 330000.0, 7000.0, 2000.0, 1960.0),
             ("peridotite", 1697000.0, 3350.0, 1800.0, 500.0), ("basalt", 1737000.0, 2900.0, 500.0, 250.0)],
}


def body(name):
    """planet::body (planet.rs:309-320): a defined body by id."""
    if name not in _BODY_LAYERS:
        raise KeyError(f"unknown body definition {name!r}")
    return LayeredBody([Layer(*l) for l in _BODY_LAYERS[name]], name)


# ---- Tillotson on the host (eos.rs:52-96) -------------------------------------------------------------
def till_pressure(t, rho, u):
    rho0, a, b, A, B, E0, Eiv, Ecv, alpha, beta = t
    rho = max(rho, 1.0e-9)
    eta = rho / rho0
    mu = eta - 1.0
    omega = u / (E0 * eta * eta) + 1.0
    pc = (a + b / omega) * rho * u + A * mu + B * mu * mu
    if eta >= 1.0 or u <= Eiv:
        return pc
    z = rho0 / rho - 1.0
    pe = a * rho * u + (b * rho * u / omega + A * mu * np.exp(-beta * z)) * np.exp(-alpha * z * z)
    if u >= Ecv:
        return pe
    return ((u - Eiv) * pe + (Ecv - u) * pc) / (Ecv - Eiv)


def till_sound_speed_sq(t, rho, u):
    rho = max(rho, 1.0e-9)
    dr = rho * 1.0e-4
    du = abs(u) * 1.0e-4 + 1.0
    dp_drho = (till_pressure(t, rho + dr, u) - till_pressure(t, rho - dr, u)) / (2.0 * dr)
    dp_du = (till_pressure(t, rho, u + du) - till_pressure(t, rho, u - du)) / (2.0 * du)
    p = till_pressure(t, rho, u)
    return max(dp_drho + p / (rho * rho) * dp_du, 0.0)


def tillotson_for(material):
    """Tillotson::for_material (eos.rs:169-171), basalt when the catalogue has no block."""
    return materials.CATALOGUE.get(material, materials.BASALT)


# ---- particalization (hydrostatic.rs) -------------------------------------------------------------------
def _round_half_away(x):
    return float(np.floor(x + 0.5)) if x >= 0 else -float(np.floor(-x + 0.5))


def fib_dir(i, n, offset):  # hydrostatic.rs:32-38
    golden = np.pi * (3.0 - np.sqrt(5.0))
    y = 1.0 - 2.0 * (i + 0.5) / n
    r = np.sqrt(np.maximum(1.0 - y * y, 0.0))
    th = golden * i + offset
    return np.stack([np.cos(th) * r, y, np.sin(th) * r], -1)


def smoothing_for(m, rho0):  # hydrostatic.rs:71-73
    return 2.0 * np.cbrt(m / rho0)


@dataclass
class HydroBody:
    """The particalized body (hydrostatic.rs:42-56): f64 state and each particle's material."""
    pos: np.ndarray
    vel: np.ndarray
    mass: np.ndarray
    u: np.ndarray
    h: np.ndarray
    rho: np.ndarray
    eos: list            # Tillotson tuple per particle
    softening: float
    materials: list = field(default_factory=list)  # material name per particle

    def relax_dt(self, cfl):  # hydrostatic.rs:483-492
        min_h = float(np.min(self.h))
        u0 = float(self.u[0]) if len(self.u) else 1.0
        c_max = 1.0
        for t in {tuple(e) for e in self.eos}:
            c_max = max(c_max, np.sqrt(till_sound_speed_sq(t, t[0], u0)))
        return cfl * min_h / max(c_max, 1.0)


def particalize(matter, resolution):
    """HydroBody::particalize (hydrostatic.rs:211-264): equal-mass particles, layer by layer, equal-volume
    radii on a per-layer Fibonacci spiral, each with its layer's own material, density and geotherm."""
    total = matter.total_mass()
    m_i = max(total / max(resolution, 1), 1.0)
    pos, mass, eos, h, u, rho, names = [], [], [], [], [], [], []
    inner = 0.0
    for li, layer in enumerate(matter.layers):
        ri3, ro3 = inner * inner * inner, layer.outer_r * layer.outer_r * layer.outer_r
        m_layer = layer.density * FOUR_THIRDS_PI * (ro3 - ri3)
        n = int(max(_round_half_away(m_layer / m_i), 1.0))
        t = tillotson_for(layer.material)
        sh = materials.SPECIFIC_HEAT.get(layer.material, 1000.0)
        h_i = smoothing_for(m_i, layer.density)
        i = np.arange(n, dtype=np.float64)
        rr = np.cbrt(ri3 + (ro3 - ri3) * (i + 0.5) / n)
        pos.append(fib_dir(i, n, li * 0.7) * rr[:, None])
        mass.append(np.full(n, m_i))
        eos += [t] * n
        names += [layer.material] * n
        h.append(np.full(n, h_i))
        u.append(np.array([sh * matter.temperature_at(r) for r in rr]))
        rho.append(np.full(n, layer.density))
        inner = layer.outer_r
    h = np.concatenate(h)
    pos = np.concatenate(pos)
    return HydroBody(pos=pos, vel=np.zeros_like(pos), mass=np.concatenate(mass), u=np.concatenate(u), h=h,
                     rho=np.concatenate(rho), eos=eos, softening=0.25 * float(np.min(h)), materials=names)


@dataclass
class Cap:
    """hydrostatic.rs:57-68: a particalized surface patch of a planet, the radius of the floor it rests on
    and the mass of the un-resolved remainder (the bulk's Gauss gravity source)."""
    body: HydroBody
    r_core: float
    bulk_mass: float


def particalize_cap(matter, site_dir, cap_radius, resolution):
    """HydroBody::particalize_cap (hydrostatic.rs:291-335): particalize the whole body finely enough that
    the ball of `cap_radius` about the surface point along `site_dir` keeps about `resolution` particles
    (that ball is about half inside the planet, volume 2/3 pi R^3; the whole-body count is clamped to
    [resolution, 400 000]), then keep only the particles inside the ball. The floor sits at the deepest
    kept particle; the bulk mass is the planet less the cap."""
    r_surf = matter.radius()
    d = np.asarray(site_dir, np.float64)
    nd = float(np.linalg.norm(d))
    site = d / nd * r_surf if nd > 0.0 else np.zeros(3)
    v_planet = FOUR_THIRDS_PI * r_surf ** 3
    v_cap = (2.0 / 3.0) * np.pi * cap_radius ** 3
    full_n = int(min(max(np.ceil(resolution * (v_planet / max(v_cap, 1.0))), float(resolution)), 400000.0))
    full = particalize(matter, full_n)
    keep = np.nonzero(np.linalg.norm(full.pos - site, axis=1) < cap_radius)[0]
    body_ = HydroBody(pos=full.pos[keep], vel=np.zeros((len(keep), 3)), mass=full.mass[keep], u=full.u[keep],
                      h=full.h[keep], rho=full.rho[keep], eos=[full.eos[i] for i in keep],
                      softening=full.softening, materials=[full.materials[i] for i in keep])
    cap_mass = float(body_.mass.sum())
    r_core = float(np.linalg.norm(body_.pos, axis=1).min()) if len(keep) else np.inf
    return Cap(body=body_, r_core=r_core if np.isfinite(r_core) else r_surf,
               bulk_mass=max(matter.total_mass() - cap_mass, 0.0))


def resolution_distance(m1, r1, m2, tidal_fraction):
    """accretion.rs:1418-1423: the separation at which body 2's tidal stress across body 1 reaches
    `tidal_fraction` of body 1's surface gravity (0 for non-physical inputs)."""
    if m1 <= 0.0 or r1 <= 0.0 or m2 <= 0.0 or tidal_fraction <= 0.0:
        return 0.0
    return r1 * np.cbrt(2.0 * (m2 / m1) / tidal_fraction)


# ---- the SPH assembly (gpu_sph.rs) ---------------------------------------------------------------------
def eos_record(t):
    """SphEos::from_tillotson (gpu_sph.rs:98-113)."""
    rec = np.zeros(1, dtype=materials.EOS_DTYPE)
    for k, v in zip(materials.EOS_DTYPE.names[:10], t):
        rec[0][k] = v
    return rec[0]


class SphAssembly:
    """gpu_sph.rs:1010-1060: particles of many bodies and one material table deduped by reference density."""

    def __init__(self):
        self.parts = []
        self.eos = []   # EOS_DTYPE records
        self.names = []

    def material_index(self, t, name):
        rec = eos_record(t)
        for k, e in enumerate(self.eos):
            if e["rho0"] == rec["rho0"]:
                return k
        if len(self.eos) < MAX_MATERIALS:
            self.eos.append(rec)
            self.names.append(name)
            return len(self.eos) - 1
        return MAX_MATERIALS - 1

    def add_body(self, hb, prov, offset, vel):
        n = len(hb.mass)
        rec = np.zeros(n, dtype=PARTICLE_DTYPE)
        rec["pos"] = (hb.pos + np.asarray(offset, np.float64)).astype(np.float32)
        rec["h"] = hb.h.astype(np.float32)
        rec["vel"] = np.broadcast_to(np.asarray(vel, np.float64), (n, 3)).astype(np.float32)
        rec["u"] = hb.u.astype(np.float32)
        rec["mass"] = hb.mass.astype(np.float32)
        rec["mat"] = [self.material_index(t, nm) for t, nm in zip(hb.eos, hb.materials)]
        rec["rho"] = hb.rho.astype(np.float32)
        rec["prov"] = prov
        self.parts.append(rec)

    @property
    def particles(self):
        return np.concatenate(self.parts) if self.parts else np.zeros(0, dtype=PARTICLE_DTYPE)

    def eos_table(self):
        return np.array(self.eos, dtype=materials.EOS_DTYPE)

    def eos_f64(self):
        return np.array([[float(e[k]) for k in materials.EOS_DTYPE.names[:10]] for e in self.eos])


def build_far_apart_n(bodies, separation):
    """gpu_sph.rs:264-292: particalize each (matter, resolution) and place the bodies on +x, adjacent
    spacing `separation` x the largest diameter. Returns (particles, eos table, softening, relax dt, the
    assembly)."""
    hbs = [particalize(m, res) for m, res in bodies]
    r_max = max(m.radius() for m, _ in bodies)
    spacing = separation * 2.0 * r_max
    asm = SphAssembly()
    for k, hb in enumerate(hbs):
        asm.add_body(hb, k, (k * spacing, 0.0, 0.0), (0.0, 0.0, 0.0))
    softening = np.float32(min(hb.softening for hb in hbs))
    relax_dt = np.float32(min(hb.relax_dt(0.2) for hb in hbs))
    return asm.particles, asm.eos_table(), float(softening), float(relax_dt), asm


def build_cap_relax(cap_body, impactors, separation):
    """gpu_sph.rs:293-330: the cap where it lies (prov 0) plus each (matter, resolution) impactor
    particalized and placed on +x beyond it: x starts at the cap's largest radius and each impactor of
    radius r sits at x += separation * 2 r + r, then x += r. Returns (particles, eos table, softening,
    relax dt, the assembly), the last two the smallest over the cap and the impactors."""
    asm = SphAssembly()
    asm.add_body(cap_body, 0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    hbs = [particalize(m, res) for m, res in impactors]
    x = max(0.0, float(np.linalg.norm(cap_body.pos, axis=1).max())) if len(cap_body.mass) else 0.0
    for k, hb in enumerate(hbs):
        r = max(0.0, float(np.linalg.norm(hb.pos, axis=1).max()))
        x += separation * 2.0 * r + r
        asm.add_body(hb, k + 1, (x, 0.0, 0.0), (0.0, 0.0, 0.0))
        x += r
    softening = float(np.float32(min([cap_body.softening] + [hb.softening for hb in hbs])))
    relax_dt = float(np.float32(min([cap_body.relax_dt(0.2)] + [hb.relax_dt(0.2) for hb in hbs])))
    return asm.particles, asm.eos_table(), softening, relax_dt, asm


def far_apart_pair(target, impactor, separation_m):
    """gpu_sph.rs:419-441: two particalized bodies recentred, the impactor `separation_m` along +x."""
    tc = (target.pos * target.mass[:, None]).sum(0) / target.mass.sum()
    ic = (impactor.pos * impactor.mass[:, None]).sum(0) / impactor.mass.sum()
    asm = SphAssembly()
    asm.add_body(target, 0, -tc, (0.0, 0.0, 0.0))
    asm.add_body(impactor, 1, -ic + np.array([separation_m, 0.0, 0.0]), (0.0, 0.0, 0.0))
    softening = float(np.float32(min(target.softening, impactor.softening)))
    relax_dt = float(np.float32(min(target.relax_dt(0.2), impactor.relax_dt(0.2))))
    return asm.particles, softening, relax_dt, asm


def body_bulk(particles, prov):
    """gpu_sph.rs:451-503: robust (centre, mass, radius enclosing 98 % of the mass) of source body `prov`,
    clipped twice to 1.5 x the bulk radius so stray particles do not drag it."""
    sel = particles[particles["prov"] == prov]
    pos = sel["pos"].astype(np.float64)
    m = sel["mass"].astype(np.float64)
    mass = float(m.sum())

    def com(mask):
        mm = m[mask].sum()
        return np.zeros(3) if mm <= 0 else (pos[mask] * m[mask, None]).sum(0) / mm

    def bulk_radius(mask, c):
        mm = m[mask].sum()
        r = np.linalg.norm(pos[mask] - c, axis=1)
        order = np.argsort(r, kind="stable")
        cum = np.cumsum(m[mask][order])
        k = np.searchsorted(cum >= 0.98 * mm, True)
        k = min(k, len(order) - 1)
        return float(r[order][k]) if len(order) else 0.0

    allm = np.ones(len(m), bool)
    c = com(allm)
    r = bulk_radius(allm, c)
    for _ in range(2):
        keep = np.linalg.norm(pos - c, axis=1) <= 1.5 * r
        if not keep.any():
            break
        c = com(keep)
        r = bulk_radius(keep, c)
    return c, mass, r


@dataclass
class BodyPlacement:
    """gpu_sph.rs:505-510: where source body `prov` goes, its bulk velocity, its angular velocity vector."""
    offset: tuple = (0.0, 0.0, 0.0)
    vel: tuple = (0.0, 0.0, 0.0)
    spin: tuple = (0.0, 0.0, 0.0)


def assemble_from_relaxed_n(particles, placements):
    """gpu_sph.rs:516-549: recentre each source body on its robust centre, place it at its offset moving at
    vel + spin x (r - centre); particles of unplaced sources are dropped. Returns (particles, softening,
    the shock-safe dt against the fastest approach)."""
    coms = [body_bulk(particles, k)[0] for k in range(len(placements))]
    keep = particles["prov"] < len(placements)
    out = particles[keep].copy()
    prov = out["prov"].astype(np.int64)
    local = out["pos"].astype(np.float64) - np.array(coms)[prov]
    off = np.array([p.offset for p in placements], np.float64)[prov]
    vel = np.array([p.vel for p in placements], np.float64)[prov]
    spin = np.array([p.spin for p in placements], np.float64)[prov]
    out["pos"] = (local + off).astype(np.float32)
    out["vel"] = (vel + np.cross(spin, local)).astype(np.float32)
    min_h = np.min(out["h"]) if len(out) else np.float32(np.inf)
    softening = float(np.float32(0.25) * min_h)
    max_speed = max((float(np.linalg.norm(p.vel)) for p in placements), default=0.0)
    dt = float(np.float32(0.01 * float(min_h) / (20000.0 + max_speed)))
    return out, softening, dt


def assemble_from_relaxed_at(particles, target_spin, impactor_offset, impactor_vel, impactor_spin):
    """gpu_sph.rs:602-628: the two-body case — target (prov 0) at the origin spun about +z, impactor
    (prov 1) at the given offset and velocity."""
    pl = [BodyPlacement((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, float(target_spin))),
          BodyPlacement(tuple(impactor_offset), tuple(impactor_vel), (0.0, 0.0, float(impactor_spin)))]
    return assemble_from_relaxed_n(particles, pl)


@dataclass
class ImpactDef:
    """terra/world_def.rs:317-420 with its defaults; the Birth-of-the-Moon world declares exactly these."""
    target: str = "proto-earth"
    impactor: str = "theia"
    v_esc_multiple: float = 1.15
    start_separation: float = 1.6
    impact_parameter: float = 1.0
    target_spin_rad_s: float = 4.0e-4
    relax_separation: float = 40.0


def assemble_from_relaxed_with(d, particles):
    """gpu_sph.rs:635-653: the declared canonical oblique impact from the relaxed bodies' own radii."""
    _, m_t, r_t = body_bulk(particles, 0)
    _, m_i, r_i = body_bulk(particles, 1)
    contact = r_t + r_i
    v_esc = d.v_esc_multiple * np.sqrt(2.0 * G * (m_t + m_i) / contact)
    d0, b = d.start_separation * contact, d.impact_parameter * r_t
    return assemble_from_relaxed_at(particles, d.target_spin_rad_s, (d0, b, 0.0), (-v_esc, 0.0, 0.0), 0.0)


def assemble_cap_impact(relaxed_impactors, placements, impactor_eos, cap_body):
    """gpu_sph.rs:551-591: the relaxed impactor(s) placed by `assemble_from_relaxed_n`, after a freshly
    assembled cap (prov 0, where it lies). One material table: the cap's materials first, then the
    impactor's (shifted, capped at 16 slots); every impactor particle becomes prov 1. Returns (particles,
    eos table, softening = 0.25 x the smallest h, the impactor placement's shock-safe dt, the f64 rows of
    the table)."""
    imp, _, dt = assemble_from_relaxed_n(relaxed_impactors, placements)
    asm = SphAssembly()
    asm.add_body(cap_body, 0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    cap_mats = len(asm.eos)
    eos = list(asm.eos)
    for e in impactor_eos:
        if len(eos) < MAX_MATERIALS:
            eos.append(e)
    imp["mat"] = np.minimum(imp["mat"].astype(np.int64) + cap_mats, MAX_MATERIALS - 1)
    imp["prov"] = 1
    out = np.concatenate([asm.particles, imp])
    min_h = np.min(out["h"]) if len(out) else np.float32(np.inf)
    table = np.array(eos, dtype=materials.EOS_DTYPE)
    rows = np.array([[float(e[k]) for k in materials.EOS_DTYPE.names[:10]] for e in eos])
    return out, table, float(np.float32(0.25) * min_h), dt, rows


def whole_body_counts(target, impactor, n_target):
    """Equal particle mass across the bodies, the impactor's count from its real mass with the small-body
    floor of 50 (lib.rs:3420-3423)."""
    m_i = target.total_mass() / n_target
    return n_target, int(max(_round_half_away(impactor.total_mass() / m_i), 50.0))


# ---- the scenes, end to end on the device --------------------------------------------------------------
def _temperatures(rec, sh_by_mat):
    return rec["u"].astype(np.float64) / sh_by_mat[rec["mat"]]


def birth_of_the_moon(n_target=2400, relax_steps=2400, sim_seconds=21600.0, cfl=0.3, d=None, device=0,
                      precision=None):
    """The Birth-of-the-Moon world (worlds/birth/world.json): proto-Earth and Theia particalized at equal
    particle mass, relaxed far apart (lib.rs:3341-3366: dt from the staging, damping 0.94, artificial
    viscosity off, 2400 steps), assembled on the declared canonical oblique impact, then stepped at the
    Courant dt with artificial viscosity on to `sim_seconds`. Returns the final read-back, the on-device
    disk view, the disk JSON, the relative total-energy change of the impact phase and the run's
    parameters."""
    from . import isph
    d = d or ImpactDef()
    t_def, i_def = body(d.target), body(d.impactor)
    counts = whole_body_counts(t_def, i_def, n_target)
    staged, eos, soft, relax_dt, asm = build_far_apart_n([(t_def, counts[0]), (i_def, counts[1])],
                                                         d.relax_separation)
    with isph.GpuSph(len(staged), device) as s:
        s.upload(staged, eos, soft)
        s.set_eos_f64(asm.eos_f64())
        if precision is not None:
            s.set_precision(precision)
        s.set_dt(relax_dt, 0.94)
        s.set_av(0.0, 0.0)
        s.encode_relax(relax_steps)
        relaxed = s.readback_sync()
        assembled, soft2, shock_dt = assemble_from_relaxed_with(d, relaxed)
        s.upload(assembled, eos, soft2)
        s.set_eos_f64(asm.eos_f64())
        s.set_dt(shock_dt, 1.0)
        s.set_av(1.0, 2.0)
        d0 = s.diagnostics(with_pe=True)
        t0, _ = s.sim_time()
        steps = 0
        while s.sim_time()[0] - t0 < sim_seconds:
            s.step_cfl(cfl, 20)
            steps += 20
        final = s.readback_sync()
        view = s.disk_view()[0]
        stats = s.disk_stats_json(0)
        t1, _ = s.sim_time()
        d1 = s.diagnostics(with_pe=True)
    e0, e1 = (x["ke"] + x["ie"] + x["pe"] for x in (d0, d1))
    return {"particles": final, "assembled": assembled, "relaxed": relaxed, "view": view, "stats": stats,
            "steps": steps, "sim_seconds": t1 - t0, "counts": counts, "relax_dt": relax_dt, "shock_dt": shock_dt,
            "eos_names": list(asm.names), "dE_rel": (e1 - e0) / abs(e0)}


def moon_drop(n_target=700, relax_steps=800, sim_seconds=6000.0, cfl=0.1, device=0):
    """The dropped Moon on the deformable path (gpu_sph.rs:2272-2497): Earth and Moon staged far apart at
    equal particle mass, relaxed (damping 0.94, artificial viscosity off), the Moon placed just above
    contact (1.35 x) falling at the speed energy conservation gives from the lunar distance, then the
    impact at the Courant dt (cfl 0.1) to `sim_seconds`. Returns the read-backs before and after, the
    per-material specific heats and the on-device disk JSON."""
    from . import isph
    e_def, m_def = body("earth"), body("moon")
    counts = whole_body_counts(e_def, m_def, n_target)
    staged, eos, soft, relax_dt, asm = build_far_apart_n([(e_def, counts[0]), (m_def, counts[1])], 40.0)
    sh_by_mat = np.array([materials.SPECIFIC_HEAT.get(nm, 1000.0) for nm in asm.names])
    with isph.GpuSph(len(staged), device) as s:
        s.upload(staged, eos, soft)
        s.set_eos_f64(asm.eos_f64())
        s.set_dt(relax_dt, 0.94)
        s.set_av(0.0, 0.0)
        s.encode_relax(relax_steps)
        relaxed = s.readback_sync()
        _, me, r_e = body_bulk(relaxed, 0)
        _, mm, r_m = body_bulk(relaxed, 1)
        contact = r_e + r_m
        mu = G * (me + mm)
        d0 = 1.35 * contact
        v0 = np.sqrt(2.0 * mu * (1.0 / d0 - 1.0 / 3.844e8))
        pl = [BodyPlacement(), BodyPlacement((0.0, d0, 0.0), (0.0, -v0, 0.0), (0.0, 0.0, 0.0))]
        assembled, soft2, shock_dt = assemble_from_relaxed_n(relaxed, pl)
        s.upload(assembled, eos, soft2)
        s.set_eos_f64(asm.eos_f64())
        s.set_dt(shock_dt, 1.0)
        s.set_av(1.0, 2.0)
        t0, _ = s.sim_time()
        steps = 0
        while s.sim_time()[0] - t0 < sim_seconds:
            s.step_cfl(cfl, 20)
            steps += 20
        final = s.readback_sync()
        s.disk_view()
        stats = s.disk_stats_json(0)
    return {"assembled": assembled, "final": final, "sh_by_mat": sh_by_mat, "stats": stats, "v0": v0,
            "steps": steps, "counts": counts}


def moon_on_earth_cap(cap_resolution=2000, relax_steps=2400, sim_seconds=3000.0, cfl=0.1, device=0):
    """The Moon dropped on the modern Earth through the engine's resolution-on-demand branch
    (lib.rs:1950-1996 and 3711-3745): a cap of radius min(2 r_moon, 0.55 r_earth) around the impact site
    is particalized at `cap_resolution` (particalize_cap), the Moon whole at 1500 particles (the impactor
    mass over 1500, floor 200); both relax seated on the bulk (Gauss gravity of the planet less the cap
    plus the floor at the cap's inner edge; damping 0.94, artificial viscosity off, build_cap_relax's
    dt). Assembly keeps the cap where it settled and puts the Moon on a radial fall from the lunar
    distance at the hand-off distance max(1 % tidal distance, contact); the impact runs with the bulk at
    the Courant dt to `sim_seconds`. Returns the cap plan, the read-backs and the disk JSON."""
    from . import isph
    e_def, m_def = body("earth"), body("moon")
    r_e, r_m = e_def.radius(), m_def.radius()
    m_e, m_m = e_def.total_mass(), m_def.total_mass()
    site = np.array([0.0, 1.0, 0.0])
    cap_radius = min(2.0 * r_m, 0.55 * r_e)
    cap = particalize_cap(e_def, site, cap_radius, cap_resolution)
    m_i = m_m / 1500.0
    n_imp = int(max(_round_half_away(m_m / m_i), 200.0))
    staged, eos, soft, relax_dt, asm = build_cap_relax(cap.body, [(m_def, n_imp)], 40.0)
    sh_by_mat = np.array([materials.SPECIFIC_HEAT.get(nm, 1000.0) for nm in asm.names])
    with isph.GpuSph(len(staged), device) as s:
        s.upload(staged, eos, soft)
        s.set_eos_f64(asm.eos_f64())
        s.set_dt(relax_dt, 0.94)
        s.set_av(0.0, 0.0)
        s.set_bulk((0.0, 0.0, 0.0), cap.r_core, (0.0, 0.0, 0.0), cap.bulk_mass)
        s.encode_relax(relax_steps)
        relaxed = s.readback_sync()
        cap_com = body_bulk(relaxed, 0)[0]
        d0 = max(resolution_distance(m_e, r_e, m_m, 0.01), r_e + r_m)
        v0 = np.sqrt(2.0 * G * (m_e + m_m) * (1.0 / d0 - 1.0 / 3.844e8))
        pl = [BodyPlacement(tuple(cap_com), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)),
              BodyPlacement(tuple(site * d0), tuple(-site * v0), (0.0, 0.0, 0.0))]
        assembled, soft2, shock_dt = assemble_from_relaxed_n(relaxed, pl)
        s.upload(assembled, eos, soft2)
        s.set_eos_f64(asm.eos_f64())
        s.set_dt(shock_dt, 1.0)
        s.set_av(1.0, 2.0)
        s.set_bulk((0.0, 0.0, 0.0), cap.r_core, (0.0, 0.0, 0.0), cap.bulk_mass)
        t0, _ = s.sim_time()
        steps = 0
        while s.sim_time()[0] - t0 < sim_seconds:
            s.step_cfl(cfl, 20)
            steps += 20
        final = s.readback_sync()
        s.disk_view()
        stats = s.disk_stats_json(0)
    return {"cap": cap, "staged": staged, "relaxed": relaxed, "assembled": assembled, "final": final,
            "sh_by_mat": sh_by_mat, "stats": stats, "v0": v0, "d0": d0, "steps": steps,
            "counts": (len(cap.body.mass), n_imp), "eos_names": list(asm.names)}
