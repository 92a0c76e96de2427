"""Scene builders: particalization of layered bodies and the N-body SPH assembly (host side, no GPU).

Restates the reference's own tests of these builders with their bounds (gpu_sph.rs:1676-1764, 1766-1876,
1878-1954, 1956-2001, 2003-2049, 2051-2136, 2138-2175, 2177-2270), with the bodies built through
`particalize` (the engine's generic path) where a test used the retired iron/basalt builder, plus closed
forms for particalization (layer shells, counts, geotherm) and the host Tillotson sound speed against the
oracle's.
"""
import numpy as np
import pytest

from botheadstudios_amd import bodies as B
from botheadstudios_amd import materials
from oracle import oracle as O

G = B.G


def _com_vel(out, prov):
    ps = out[out["prov"] == prov]
    m = ps["mass"].astype(np.float64)
    c = (ps["pos"].astype(np.float64) * m[:, None]).sum(0) / m.sum()
    v = (ps["vel"].astype(np.float64) * m[:, None]).sum(0) / m.sum()
    return c, v, m.sum()


def test_body_definitions_integrate_to_their_declared_masses():
    """The definitions' own statements: proto-Earth 0.910 M_earth, Theia ~6.5e23 kg, Earth 5.97e24 kg."""
    assert B.body("proto-earth").total_mass() / 5.972e24 == pytest.approx(0.910, abs=0.002)
    assert B.body("theia").total_mass() == pytest.approx(6.5e23, rel=0.01)
    assert B.body("earth").total_mass() == pytest.approx(5.972e24, rel=0.01)
    assert B.body("moon").total_mass() == pytest.approx(7.342e22, rel=0.01)


@pytest.mark.parametrize("name,res", [("earth", 1500), ("moon", 300), ("proto-earth", 2400), ("theia", 286)])
def test_particalize_fills_each_layer_with_its_own_matter(name, res):
    matter = B.body(name)
    hb = B.particalize(matter, res)
    m_i = matter.total_mass() / res
    assert np.all(hb.mass == m_i)
    r = np.linalg.norm(hb.pos, axis=1)
    inner = 0.0
    start = 0
    for layer in matter.layers:
        m_layer = layer.density * 4.0 / 3.0 * np.pi * (layer.outer_r ** 3 - inner ** 3)
        n = int(max(np.floor(m_layer / m_i + 0.5), 1))
        sl = slice(start, start + n)
        assert np.all((r[sl] >= inner * (1 - 1e-12)) & (r[sl] <= layer.outer_r * (1 + 1e-12)))
        assert np.all(np.array(hb.materials[sl]) == layer.material)
        assert np.all(hb.h[sl] == pytest.approx(2.0 * np.cbrt(m_i / layer.density), rel=1e-12))
        assert np.all(hb.rho[sl] == layer.density)
        c = materials.SPECIFIC_HEAT[layer.material]
        t = np.array([matter.temperature_at(x) for x in r[sl]])
        assert np.all(hb.u[sl] == pytest.approx(c * t, rel=1e-12))
        assert np.all((t >= min(layer.t_inner, layer.t_outer) - 1e-9) & (t <= max(layer.t_inner, layer.t_outer) + 1e-9))
        inner = layer.outer_r
        start += n
    assert start == len(hb.mass)
    assert hb.softening == pytest.approx(0.25 * hb.h.min())
    # equal-volume shells: the particles' mean r^3 sits mid-shell
    assert abs(np.mean(r ** 3) / matter.radius() ** 3 - 0.5) < 0.1


def test_host_tillotson_sound_speed_matches_the_oracle():
    for name in ("granite", "basalt", "peridotite", "iron"):
        t = materials.CATALOGUE[name]
        for rho_f, u in ((1.0, 1e3), (1.3, 5e5), (0.7, 2e6), (0.5, 1e7), (0.9, 3e7)):
            rho = rho_f * t[0]
            assert B.till_pressure(t, rho, u) == pytest.approx(O.till_pressure(t, rho, u), rel=1e-13, abs=1e-3)
            assert B.till_sound_speed_sq(t, rho, u) == pytest.approx(O.till_c2(t, rho, u), rel=1e-9, abs=1e-6)


def test_relax_dt_is_the_courant_step_of_the_stiffest_material():
    hb = B.particalize(B.body("earth"), 800)
    c = max(np.sqrt(O.till_c2(materials.CATALOGUE[m], materials.CATALOGUE[m][0], hb.u[0]))
            for m in set(hb.materials))
    assert hb.relax_dt(0.2) == pytest.approx(0.2 * hb.h.min() / c, rel=1e-9)


def test_a_particalized_pair_stages_far_apart_at_real_mass():
    """gpu_sph.rs:1956-2001."""
    e, m = B.body("earth"), B.body("moon")
    n_t, n_i = B.whole_body_counts(e, m, 800)
    sep = 40.0 * (e.radius() + m.radius())
    out, softening, relax_dt, asm = B.far_apart_pair(B.particalize(e, n_t), B.particalize(m, n_i), sep)
    tc, tm, _ = B.body_bulk(out, 0)
    ic, im, _ = B.body_bulk(out, 1)
    assert abs(tm - e.total_mass()) / e.total_mass() < 0.02
    assert abs(im - m.total_mass()) / m.total_mass() < 0.03
    assert abs(np.linalg.norm(ic - tc) - sep) / sep < 0.02
    assert np.all(out["vel"] == 0)
    names = [asm.names[k] for k in set(out["mat"][out["prov"] == 0].tolist())]
    assert "iron" in names and "basalt" in names
    assert softening > 0 and relax_dt > 0


def test_sph_assembly_builds_an_n_material_field_across_bodies():
    """gpu_sph.rs:2003-2049."""
    asm = B.SphAssembly()
    asm.add_body(B.particalize(B.body("earth"), 1500), 0, (0, 0, 0), (0, 0, 0))
    asm.add_body(B.particalize(B.body("moon"), 300), 1, (4.0e7, 0, 0), (0, 0, 0))
    p = asm.particles
    assert 3 <= len(asm.eos) <= B.MAX_MATERIALS
    assert np.all(p["mat"] < len(asm.eos))
    assert (p["prov"] == 0).any() and (p["prov"] == 1).any()
    iron = [k for k, e in enumerate(asm.eos) if abs(e["rho0"] - 7850.0) < 50.0]
    assert iron and (p["mat"] == iron[0]).any()
    real = B.body("earth").total_mass() + B.body("moon").total_mass()
    assert abs(p["mass"].astype(np.float64).sum() - real) / real < 0.03


def test_assemble_n_places_every_body_and_spins_about_any_axis():
    """gpu_sph.rs:2051-2136."""
    e = B.particalize(B.body("earth"), 1200)
    m = B.particalize(B.body("moon"), 250)
    asm = B.SphAssembly()
    asm.add_body(e, 0, (0, 0, 0), (0, 0, 0))
    asm.add_body(m, 1, (4.0e7, 0, 0), (0, 0, 0))
    asm.add_body(m, 2, (-4.0e7, 0, 0), (0, 0, 0))
    pl = [B.BodyPlacement((0, 0, 0), (0, 0, 0), (3.0e-4, 0, 0)),
          B.BodyPlacement((1.5e7, 2.0e6, 0), (-8000.0, 0, 0)),
          B.BodyPlacement((0, 1.6e7, 0), (0, -7000.0, 0))]
    out, _, _ = B.assemble_from_relaxed_n(asm.particles, pl)
    (c0, v0, _), (c1, v1, _), (c2, v2, _) = (_com_vel(out, k) for k in range(3))
    r_e = B.body("earth").radius()
    assert np.linalg.norm(c0) < 0.02 * r_e
    assert np.linalg.norm(c1 - c0 - pl[1].offset) < 0.02 * r_e
    assert np.linalg.norm(c2 - c0 - pl[2].offset) < 0.02 * r_e
    assert np.linalg.norm(v1 - v0 - pl[1].vel) < 1.0
    assert np.linalg.norm(v2 - v0 - pl[2].vel) < 1.0
    assert np.abs(out["vel"][out["prov"] == 0, 2]).max() > 1.0
    assert np.linalg.norm(v0) < 1.0


def test_build_far_apart_n_particalizes_and_separates_the_bodies():
    """gpu_sph.rs:2138-2175."""
    e, m = B.body("earth"), B.body("moon")
    parts, eos, soft, dt, _ = B.build_far_apart_n([(e, 1500), (m, 300)], 40.0)
    assert len(eos) >= 3 and soft > 0 and dt > 0
    sep = np.linalg.norm(_com_vel(parts, 1)[0] - _com_vel(parts, 0)[0])
    assert sep > 10.0 * e.radius()


def test_a_live_drop_assembles_n_materials_on_its_own_geometry():
    """gpu_sph.rs:2177-2270."""
    e, m = B.body("earth"), B.body("moon")
    n_t, n_i = B.whole_body_counts(e, m, 800)
    staged, eos, soft, relax_dt, _ = B.build_far_apart_n([(e, n_t), (m, n_i)], 40.0)
    assert len(eos) >= 3 and soft > 0 and relax_dt > 0
    offset, rel_vel = np.array([1.5e7, 3.0e6, 0.0]), np.array([-9000.0, 500.0, 0.0])
    out, _, _ = B.assemble_from_relaxed_n(staged, [B.BodyPlacement((0, 0, 0), (0, 0, 0), (2.0e-5, 0, 7.0e-5)),
                                                   B.BodyPlacement(tuple(offset), tuple(rel_vel))])
    assert len(out) == len(staged) and np.all(out["mat"] < len(eos))
    tc, tm, _ = B.body_bulk(out, 0)
    ic, im, _ = B.body_bulk(out, 1)
    assert abs(tm - e.total_mass()) / e.total_mass() < 0.02
    assert abs(im - m.total_mass()) / m.total_mass() < 0.03
    assert np.linalg.norm(ic - tc - offset) < 0.02 * e.radius()
    assert np.linalg.norm(_com_vel(out, 1)[1] - _com_vel(out, 0)[1] - rel_vel) < 1.0
    assert np.abs(out["vel"][out["prov"] == 0, 2]).max() > 1.0


def _pair_field(target, impactor, n_target):
    t, i = B.body(target), B.body(impactor)
    n_t, n_i = B.whole_body_counts(t, i, n_target)
    asm = B.SphAssembly()
    asm.add_body(B.particalize(t, n_t), 0, (0, 0, 0), (0, 0, 0))
    asm.add_body(B.particalize(i, n_i), 1, (4.0e7, 0, 0), (0, 0, 0))
    return asm.particles


def test_the_assembled_geometry_hits_even_with_stray_particles():
    """gpu_sph.rs:1676-1764, bodies particalized from their definitions. The start distance is measured by
    the plain mass-weighted centres, which include the stray itself: at 1600 target particles (192 in the
    impactor) one stray moves the impactor's centre by about the reference builder's share."""
    d = B.ImpactDef()
    field = _pair_field(d.target, d.impactor, 1600)
    contact = B.body(d.target).radius() + B.body(d.impactor).radius()

    def closest(particles):
        out, _, _ = B.assemble_from_relaxed_with(d, particles)
        (ec, ev, me), (tc, tv, mt) = _com_vel(out, 0), _com_vel(out, 1)
        q = O.orbit_perigee(tc - ec, tv - ev, G * (me + mt))
        return (q if q is not None else 0.0), np.linalg.norm(tc - ec)

    q_clean, d_clean = closest(field)
    assert q_clean < contact and d_clean < 6.0 * contact
    strayed = field.copy()
    for prov in (0, 1):
        k = np.flatnonzero(strayed["prov"] == prov)[0]
        strayed["pos"][k, 0] += np.float32(20.0 * contact)
    q_stray, d_stray = closest(strayed)
    assert q_stray < contact
    assert abs(d_stray - d_clean) < 0.1 * d_clean


def test_a_moon_drop_builds_and_strikes_through_the_same_assembly():
    """gpu_sph.rs:1766-1876 with the Earth and Moon particalized at their in-situ densities (real masses
    rather than the retired builder's uncompressed ones)."""
    d = B.ImpactDef(target="earth", impactor="moon")
    field = _pair_field("earth", "moon", 1200)
    r_e, r_m = B.body("earth").radius(), B.body("moon").radius()
    # the outer particle tracks the definition radius (measured about the body's own centre, where
    # particalize builds it: its few-particle outer shells put the mass centre a few percent off it)
    for prov, r, at in ((0, r_e, (0.0, 0.0, 0.0)), (1, r_m, (4.0e7, 0.0, 0.0))):
        ps = field[field["prov"] == prov]
        rmax = np.linalg.norm(ps["pos"].astype(np.float64) - np.array(at), axis=1).max()
        assert 0.85 * r <= rmax <= 1.02 * r
    me, mm = _com_vel(field, 0)[2], _com_vel(field, 1)[2]
    assert 40.0 < me / mm < 120.0
    out, _, _ = B.assemble_from_relaxed_with(d, field)
    (ec, ev, me), (mc, mv, mm) = _com_vel(out, 0), _com_vel(out, 1)
    q = O.orbit_perigee(mc - ec, mv - ev, G * (me + mm))
    assert (q if q is not None else 0.0) < r_e + r_m
    v_rel = np.linalg.norm(mv - ev)
    e_impact = 0.5 * (me * mm / (me + mm)) * v_rel * v_rel
    e_bind = 0.6 * G * mm * mm / r_m
    assert e_impact > 10.0 * e_bind


def test_assemble_at_honours_a_given_live_geometry():
    """gpu_sph.rs:1878-1954."""
    field = _pair_field("earth", "moon", 800)
    offset, rel_vel = np.array([1.5e7, 3.0e6, 0.0]), np.array([-9000.0, 500.0, 0.0])
    out, _, _ = B.assemble_from_relaxed_at(field, 0.0, offset, rel_vel, 0.0)
    (ec, ev, _), (mc, mv, _) = _com_vel(out, 0), _com_vel(out, 1)
    r_e = B.body("earth").radius()
    assert np.linalg.norm(ec) < 0.02 * r_e and np.linalg.norm(ev) < 1.0
    assert np.linalg.norm(mc - ec - offset) < 0.02 * r_e
    assert np.linalg.norm(mv - ev - rel_vel) < 1.0


def test_body_bulk_of_a_uniform_ball_and_with_strays():
    """98 % of a uniform ball's mass lies within 0.98^(1/3) R; strays far out move nothing."""
    rng = np.random.default_rng(3)
    n = 20000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    R = 6.0e6
    pos = d * (R * np.cbrt(rng.random(n)))[:, None] + np.array([1e7, -2e6, 3e5])
    from botheadstudios_amd.scene import PARTICLE_DTYPE
    rec = np.zeros(n, dtype=PARTICLE_DTYPE)
    rec["pos"] = pos
    rec["mass"] = 1e20
    c, m, r = B.body_bulk(rec, 0)
    assert np.linalg.norm(c - np.array([1e7, -2e6, 3e5])) < 0.02 * R
    assert r == pytest.approx(R * 0.98 ** (1.0 / 3.0), rel=0.01)
    rec["pos"][:5, 0] += np.float32(20 * R)
    c2, _, r2 = B.body_bulk(rec, 0)
    assert np.linalg.norm(c2 - c) < 0.01 * R and abs(r2 - r) < 0.01 * R
