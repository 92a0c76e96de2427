"""Pins the oracle's measurement restatement: orbit::perigee / is_bound, DiskView::of and
accretion::find_clumps.

Checks: the reference's own tests with their scenarios and bounds (orbit.rs:405-428, 908-940;
gpu_sph.rs:2504-2541; accretion.rs:573-662, 981-1230), closed forms (the periapsis of a constructed
ellipse and hyperbola, a uniform ball's binding energy), brute force (connected components of the
linking graph, the mass quantile of the remnant) and invariants (mass partition, dead particles ignored).
"""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from oracle import oracle as O

from measure_scenes import G, cold_blob, disk_field, records, spinning_blob

MU_EM = G * (5.972e24 + 7.342e22)


# ---- orbit.rs ----
def test_perigee_tracks_how_hard_the_moon_is_braked():
    """orbit.rs:405-428."""
    r = 3.844e8
    vc = np.sqrt(MU_EM / r)
    rp = O.orbit_perigee([r, 0, 0], [0, vc, 0], MU_EM)
    assert abs(rp - r) / r < 1e-3
    rp_half = O.orbit_perigee([r, 0, 0], [0, 0.5 * vc, 0], MU_EM)
    assert abs(rp_half - 0.1429 * r) / (0.1429 * r) < 0.02 and rp_half > 6.371e6
    assert O.orbit_perigee([r, 0, 0], [0, 0, 0], MU_EM) < 1.0


def test_an_unbound_approach_still_has_a_closest_approach():
    """orbit.rs:908-940."""
    mu = G * 5.972e24
    r = 7.0e6
    q = O.orbit_perigee([r, 0, 0], [0, np.sqrt(mu / r), 0], mu)
    assert abs(q - r) < 1.0
    d, v, b = 1.0e8, 11000.0, 2.0e7
    q = O.orbit_perigee([d, b, 0], [-v, 0, 0], mu)
    assert q is not None and 0 < q < b * 0.995
    assert O.orbit_perigee([d, 0, 0], [-v, 0, 0], mu) is None


@pytest.mark.parametrize("e", [0.0, 0.3, 0.8])
def test_perigee_of_a_constructed_ellipse(e):
    """Released at apoapsis r_a = a (1 + e) with the vis-viva speed: periapsis a (1 - e). (e comes from the
    square root of a cancelling sum, so a circular orbit carries ~sqrt(eps) relative error.)"""
    mu, a = G * 6e24, 4.0e7
    ra = a * (1 + e)
    va = np.sqrt(mu * (2 / ra - 1 / a))
    th = 0.7
    p = ra * np.array([np.cos(th), np.sin(th), 0.3])
    p = p / np.linalg.norm(p) * ra
    t = np.cross([0.1, 0.2, 1.0], p)
    v = va * t / np.linalg.norm(t)
    q = O.orbit_perigee(p, v, mu)
    assert O.orbit_is_bound(p, v, mu)
    assert abs(q - a * (1 - e)) <= 1e-7 * a


def test_perigee_of_a_constructed_hyperbola():
    """At periapsis q with speed above escape: unbound, and its closest approach is q itself."""
    mu, q = G * 6e24, 9.0e6
    v = 1.7 * np.sqrt(2 * mu / q)
    p, w = [q, 0, 0], [0, v, 0]
    assert not O.orbit_is_bound(p, w, mu)
    assert abs(O.orbit_perigee(p, w, mu) - q) <= 1e-9 * q


def test_is_bound_needs_negative_energy_and_a_distance():
    mu = G * 6e24
    r = 1.0e7
    ve = np.sqrt(2 * mu / r)
    assert O.orbit_is_bound([r, 0, 0], [0, 0.999 * ve, 0], mu)
    assert not O.orbit_is_bound([r, 0, 0], [0, 1.001 * ve, 0], mu)
    assert not O.orbit_is_bound([0, 0, 0], [0, 0, 0], mu)


# ---- DiskView::of ----
def _three(m_c=6.0e24, r=1.0e8):
    mu = G * float(np.float32(m_c))
    v_circ = np.float32(np.sqrt(mu / r))
    v_esc = np.float32(np.sqrt(2 * mu / r))
    pos = np.array([[0, 0, 0], [r, 0, 0], [0, r, 0]], np.float32)
    vel = np.array([[0, 0, 0], [0, v_circ, 0], [0, 7 * v_esc, 0]], np.float32)
    return records(pos, vel, np.array([m_c, 1e20, 1e20], np.float32), np.ones(3), np.full(3, 3000.0),
                   np.array([0, 1, 1], np.uint32))


def test_escaping_matter_is_not_counted_as_disk():
    """gpu_sph.rs:2504-2541: one bound orbiter is disk, the hyperbolic particle escapes."""
    v, idx = O.disk_view(_three())
    assert v["valid"] == 1 and list(idx) == [1]
    assert abs(v["disk_prov0"] + v["disk_other"] - 1e20) / 1e20 < 1e-6
    assert abs(v["escaped"] - 1e20) / 1e20 < 1e-6
    assert v["disk_prov0"] == 0.0


def test_disk_view_is_none_below_two_live_particles():
    rec = _three()
    rec["mass"][1:] = 0.0
    v, idx = O.disk_view(rec)
    assert v["valid"] == 0 and len(idx) == 0


def test_disk_view_partitions_the_live_mass_and_ignores_dead_particles():
    rec = disk_field(11)
    v, idx = O.disk_view(rec)
    live = (rec["mass"] > 0) & np.isfinite(rec["mass"]) & np.isfinite(rec["pos"]).all(1) & \
        np.isfinite(rec["vel"]).all(1)
    m = rec["mass"][live].astype(np.float64)
    assert v["n_live"] == live.sum()
    total = v["disk_prov0"] + v["disk_other"] + v["escaped"] + v["remnant_prov0"] + v["remnant_other"]
    assert abs(total - m.sum()) <= 1e-12 * m.sum()
    # the disk, the plunging and the escaping matter are all populated
    assert len(idx) > 500 and v["escaped"] > 0 and v["remnant_other"] + v["remnant_prov0"] > 0
    # dead particles are never members; removing them changes nothing
    assert live[idx].all()
    v2, idx2 = O.disk_view(rec[live])
    assert np.array_equal(np.flatnonzero(live)[idx2], idx)
    for k in ("r_remnant", "m_remnant", "disk_prov0", "disk_other", "escaped", "mean_h"):
        assert v2[k] == pytest.approx(v[k], rel=1e-13, abs=0.0)


def test_remnant_radius_is_the_85_percent_mass_quantile():
    rec = disk_field(12)
    v, _ = O.disk_view(rec)
    live = (rec["mass"] > 0) & np.isfinite(rec["mass"]) & np.isfinite(rec["pos"]).all(1) & \
        np.isfinite(rec["vel"]).all(1)
    p = rec["pos"][live].astype(np.float64)
    m = rec["mass"][live].astype(np.float64)
    com = np.array([v["com_x"], v["com_y"], v["com_z"]])
    r = np.linalg.norm(p - com, axis=1)
    inside = m[r <= v["r_remnant"]].sum()
    below = m[r < v["r_remnant"]].sum()
    assert inside >= 0.85 * m.sum() * (1 - 1e-12) and below < 0.85 * m.sum()
    assert v["m_remnant"] == pytest.approx(inside, rel=1e-12)
    assert v["mu"] == pytest.approx(G * v["m_remnant"], rel=1e-15)


# ---- accretion::find_clumps ----
def _accrete(rec, labels):
    """accretion.rs:316-352: the accreting clumps become bodies, the rest stay particles."""
    acc = rec[(rec[:, 1] >= 2) & (rec[:, 17] == 1) & (rec[:, 18] == 1)]
    consumed = np.isin(labels, acc[:, 0].astype(np.uint32))
    return acc, consumed


def test_accretion_conserves_mass_momentum_and_com():
    """accretion.rs:573-662: two cold blobs accrete, five scattered singletons do not."""
    p1, v1, m1, r1 = cold_blob([2e7, 0, 0], [0, 1500, 0], 3e5, 40, 1e19, 3000.0)
    p2, v2, m2, r2 = cold_blob([-1.5e7, 1e7, 0], [-800, 0, 300], 2.5e5, 30, 1e19, 3000.0)
    ps = np.array([[4e7 + k * 5e6, -3e7, 0] for k in range(5)])
    vs = np.array([[0, 0, 500.0 * k] for k in range(5)])
    pos, vel = np.concatenate([p1, p2, ps]), np.concatenate([v1, v2, vs])
    mass, rho = np.concatenate([m1, m2, np.full(5, 1e19)]), np.full(75, 3000.0)
    rec, labels = O.find_clumps(pos, vel, mass, rho, 5e5, G, 1e4, [0, 0, 0], 5e24, 6e6)
    bodies, consumed = _accrete(rec, labels)
    assert len(bodies) == 2 and (~consumed).sum() == 5
    fm = np.concatenate([bodies[:, 2], mass[~consumed]])
    fp = np.concatenate([bodies[:, 3:6], pos[~consumed]])
    fv = np.concatenate([bodies[:, 6:9], vel[~consumed]])
    m0, mom0, com0 = mass.sum(), (vel * mass[:, None]).sum(0), (pos * mass[:, None]).sum(0) / mass.sum()
    assert abs(fm.sum() - m0) / m0 < 1e-12
    assert np.linalg.norm((fv * fm[:, None]).sum(0) - mom0) / np.linalg.norm(mom0) < 1e-12
    assert np.linalg.norm((fp * fm[:, None]).sum(0) / fm.sum() - com0) / np.linalg.norm(com0) < 1e-12


def test_accretion_conserves_angular_momentum_of_a_spinning_clump():
    """accretion.rs:981-1030."""
    pos, vel, mass, rho = spinning_blob([2e7, 0, 0], [0, 1500, 0], [0, 0, 1e-4], 3e5, 60, 1e19, 3000.0)
    rec, labels = O.find_clumps(pos, vel, mass, rho, 5e5, G, 1e4, [0, 0, 0], 5e24, 6e6)
    bodies, _ = _accrete(rec, labels)
    assert len(bodies) == 1
    b = bodies[0]
    l_before = (np.cross(pos, vel) * mass[:, None]).sum(0)
    l_after = b[2] * np.cross(b[3:6], b[6:9]) + b[13:16]
    assert np.linalg.norm(l_after - l_before) / np.linalg.norm(l_before) < 1e-10
    assert np.linalg.norm(b[13:16]) / np.linalg.norm(l_before) > 1e-9
    assert b[16] < 1e-9 * max(0.5 * b[2] * np.dot(b[6:9], b[6:9]), 1.0)


def test_rigid_rotation_is_carried_as_spin_not_as_heat():
    """accretion.rs:1037-1112."""
    pos, vel, mass, rho = spinning_blob([2e7, 0, 0], [0, 0, 0], [0, 0, 1e-4], 3e5, 60, 1e19, 3000.0)
    rec, _ = O.find_clumps(pos, vel, mass, rho, 5e5, G, 1e4, [0, 0, 0], 5e24, 6e6)
    assert len(rec) == 1
    assert rec[0, 11] > 0 and rec[0, 16] / rec[0, 11] < 1e-9
    i = np.arange(60, dtype=np.float64)
    hh = np.modf(np.sin(i * 12.9898) * 43758.5453)[0]
    kk = np.modf(np.sin(i * 78.233) * 27183.2818)[0]
    jitter = np.stack([hh - 0.5, kk - 0.5, np.modf(hh + kk)[0] - 0.5], 1) * 20.0
    rec2, _ = O.find_clumps(pos, vel + jitter, mass, rho, 5e5, G, 1e4, [0, 0, 0], 5e24, 6e6)
    assert rec2[0, 16] / rec2[0, 11] > 0.1
    assert rec2[0, 11] - rec2[0, 16] >= -1e-6 * rec2[0, 11]


def test_roche_gate_blocks_accretion_inside_the_limit():
    """accretion.rs:1114-1160."""
    m_p, r_p, rho_c = 5e24, 6e6, 3000.0
    d_roche = 2.44 * r_p * np.cbrt((m_p / (4.0 / 3.0 * np.pi * r_p ** 3)) / rho_c)
    for frac, want in ((0.6, 0), (2.0, 1)):
        pos, vel, mass, rho = cold_blob([frac * d_roche, 0, 0], [0, 0, 0], 2e5, 30, 1e19, rho_c)
        rec, labels = O.find_clumps(pos, vel, mass, rho, 5e5, G, 1e4, [0, 0, 0], m_p, r_p)
        assert len(_accrete(rec, labels)[0]) == want


def test_unbound_hot_group_does_not_accrete():
    """accretion.rs:1162-1198."""
    pos, vel, mass, rho = cold_blob([3e7, 0, 0], [0, 0, 0], 2e5, 30, 1e15, 3000.0)
    s = np.where(np.arange(30) % 2 == 0, 1.0, -1.0)
    vel = np.stack([5000 * s, -4000 * s, 3000 * s], 1)
    rec, labels = O.find_clumps(pos, vel, mass, rho, 5e5, G, 1e4, [0, 0, 0], 5e24, 6e6)
    big = rec[np.argmax(rec[:, 1])]
    assert big[1] >= 25 and big[17] == 0
    assert len(_accrete(rec, labels)[0]) == 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_clump_partition_is_the_connected_components_of_the_linking_graph(seed):
    rng = np.random.default_rng(seed)
    n = 700
    pos = rng.uniform(0, 1e6, (n, 3))
    ll = 6.0e4
    rec, labels = O.find_clumps(pos, np.zeros((n, 3)), np.ones(n), np.ones(n), ll, G, 1e4, [0, 0, 0], 1.0, 1.0)
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=2)
    i, j = np.nonzero(d <= ll)
    nc, comp = connected_components(coo_matrix((np.ones(len(i)), (i, j)), shape=(n, n)), directed=False)
    assert len(rec) == nc
    # every component is labelled by its smallest member, and the labels induce the same partition
    first = {c: np.flatnonzero(comp == c).min() for c in range(nc)}
    assert np.array_equal(labels, np.array([first[c] for c in comp], np.uint32))
    assert sorted(rec[:, 1].astype(int)) == sorted(np.bincount(comp).tolist())


def test_self_energy_of_a_uniform_ball_and_the_octree_branch():
    """Uniform ball: PE -> -3/5 G M^2 / R. Below 1024 members the reference sum is exact; above it the
    theta = 0.5 octree stays within 1 % of the exact pair sum."""
    rng = np.random.default_rng(5)
    for n in (800, 3000):
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        R, M = 1.0e6, 1.0e21
        pos = d * (R * np.cbrt(rng.random(n)))[:, None]
        mass = np.full(n, M / n)
        args = (pos, np.zeros((n, 3)), mass, np.full(n, 3000.0), 1e7, G, 1.0, [1e9, 0, 0], 1.0, 1.0)
        ref, _ = O.find_clumps(*args)
        ex, _ = O.find_clumps(*args, exact_pe=True)
        assert len(ref) == 1 and len(ex) == 1
        pe_ball = -0.6 * G * M * M / R
        assert abs(ex[0, 12] - pe_ball) / abs(pe_ball) < 0.05
        if n < 1024:
            assert ref[0, 12] == pytest.approx(ex[0, 12], rel=1e-13)
        else:
            assert abs(ref[0, 12] - ex[0, 12]) / abs(ex[0, 12]) < 1e-2
