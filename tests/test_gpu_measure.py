"""On-device measurement against the oracle: DiskView::of (gpu_sph.rs:689-785), DiskView::clumps /
disk_stats_json (:793-851) and accretion::find_clumps (accretion.rs:158-301).

Parity bar: the same disk members, clump partition and gate decisions as the oracle; f64 values within
1e-12 relative for the view (sums in another order) and 1e-10 for clump statistics (a different summation
tree over up to thousands of members); the self energy against the oracle's exact pair sum, and against the
reference's theta = 0.5 octree (used above 1024 members) within 1 %.
"""
import json

import numpy as np
import pytest

from botheadstudios_amd import isph, materials, scene
from oracle import oracle as O

from measure_scenes import G, cold_blob, disk_field, records, spinning_blob

pytestmark = pytest.mark.gpu

VIEW_KEYS = [("r_remnant", "r_remnant"), ("m_remnant", "m_remnant"), ("mu", "mu"), ("disk_prov0", "disk_prov0"),
             ("disk_other", "disk_other"), ("escaped", "escaped"), ("mean_h", "mean_h"),
             ("remnant_prov0", "remnant_prov0"), ("remnant_other", "remnant_other"),
             ("escaped_prov0", "escaped_prov0")]


def _stepper(rec):
    s = isph.GpuSph(max(len(rec), 1))
    s.upload(rec, materials.eos_table(("basalt",)), 1.0e4)
    return s


def _clump_rows(clumps):
    """GPU clumps as rows in the oracle's record order."""
    rows = []
    for c in clumps:
        rows.append([c.id, c.n_members, c.mass, *c.com_pos, *c.com_vel, c.rho, c.radius, c.internal_ke,
                     c.self_pe, *c.ang_mom, c.thermal_ke, float(c.bound), float(c.outside_roche), 0.0])
    return np.array(rows).reshape(-1, 20)


def _assert_clumps_match(gpu_rows, orc_rows, rtol=1e-10):
    """Same ids, sizes and gates; values within rtol. Quantities that are pure round-off in a cold clump
    (internal KE, spin, heat) are compared against a floor far below anything physical: 1e-24 m |V|^2 for
    energies, 1e-12 m R |V| for the spin (V the bulk speed + 1 m/s, R the clump radius)."""
    assert gpu_rows.shape == orc_rows.shape
    assert np.array_equal(gpu_rows[:, [0, 1, 17, 18]], orc_rows[:, [0, 1, 17, 18]])
    m, R = orc_rows[:, 2], orc_rows[:, 10]
    V = np.linalg.norm(orc_rows[:, 6:9], axis=1) + 1.0
    e_floor = 1e-24 * m * V * V
    for col in (2, 9, 10, 12):
        np.testing.assert_allclose(gpu_rows[:, col], orc_rows[:, col], rtol=rtol, atol=0)
    ke = np.abs(orc_rows[:, 11])
    assert np.all(np.abs(gpu_rows[:, 11] - orc_rows[:, 11]) <= rtol * ke + e_floor)
    assert np.all(np.abs(gpu_rows[:, 16] - orc_rows[:, 16]) <= 1e-9 * ke + e_floor)
    for lo, floor in ((3, 0.0), (6, 0.0), (13, 1e-12 * m * R * V)):
        scale = np.linalg.norm(orc_rows[:, lo:lo + 3], axis=1)
        err = np.abs(gpu_rows[:, lo:lo + 3] - orc_rows[:, lo:lo + 3]).max(axis=1)
        assert np.all(err <= rtol * scale + floor), (lo, err, scale)


def _find(s, pos, vel, mass, rho, ll, central=(0.0, 0.0, 0.0), cm=5e24, cr=6e6, soft=1e4, min_members=1):
    clumps, labels = s.find_clumps(pos, vel, mass, rho, ll, G, soft, central, cm, cr, min_members)
    return _clump_rows(clumps), labels


def _fixtures():
    """The scenarios of accretion.rs's tests (573-662, 981-1230)."""
    out = []
    p1, v1, m1, r1 = cold_blob([2e7, 0, 0], [0, 1500, 0], 3e5, 40, 1e19, 3000.0)
    p2, v2, m2, r2 = cold_blob([-1.5e7, 1e7, 0], [-800, 0, 300], 2.5e5, 30, 1e19, 3000.0)
    ps = np.array([[4e7 + k * 5e6, -3e7, 0] for k in range(5)])
    vs = np.array([[0, 0, 500.0 * k] for k in range(5)])
    out.append(("two blobs and singletons", np.concatenate([p1, p2, ps]), np.concatenate([v1, v2, vs]),
                np.concatenate([m1, m2, np.full(5, 1e19)]), np.full(75, 3000.0), 2))
    pos, vel, mass, rho = spinning_blob([2e7, 0, 0], [0, 1500, 0], [0, 0, 1e-4], 3e5, 60, 1e19, 3000.0)
    out.append(("spinning clump", pos, vel, mass, rho, 1))
    pos, vel, mass, rho = spinning_blob([2e7, 0, 0], [0, 0, 0], [0, 0, 1e-4], 3e5, 60, 1e19, 3000.0)
    out.append(("rigid rotator", pos, vel, mass, rho, 1))
    i = np.arange(60, dtype=np.float64)
    hh, kk = np.modf(np.sin(i * 12.9898) * 43758.5453)[0], np.modf(np.sin(i * 78.233) * 27183.2818)[0]
    jit = np.stack([hh - 0.5, kk - 0.5, np.modf(hh + kk)[0] - 0.5], 1) * 20.0
    out.append(("jittered rotator", pos, vel + jit, mass, rho, None))
    d_roche = 2.44 * 6e6 * np.cbrt((5e24 / (4.0 / 3.0 * np.pi * 6e6 ** 3)) / 3000.0)
    for frac, want in ((0.6, 0), (2.0, 1)):
        pos, vel, mass, rho = cold_blob([frac * d_roche, 0, 0], [0, 0, 0], 2e5, 30, 1e19, 3000.0)
        out.append((f"Roche {frac}", pos, vel, mass, rho, want))
    pos, vel, mass, rho = cold_blob([3e7, 0, 0], [0, 0, 0], 2e5, 30, 1e15, 3000.0)
    sg = np.where(np.arange(30) % 2 == 0, 1.0, -1.0)
    out.append(("hot group", pos, np.stack([5000 * sg, -4000 * sg, 3000 * sg], 1), mass, rho, 0))
    return out


@pytest.mark.parametrize("case", range(7))
def test_find_clumps_reference_fixtures(case):
    name, pos, vel, mass, rho, n_accrete = _fixtures()[case]
    print(f"find_clumps fixture: {name}", flush=True)
    with isph.GpuSph(len(mass)) as s:
        rows, labels = _find(s, pos, vel, mass, rho, 5e5)
    orc, olab = O.find_clumps(pos, vel, mass, rho, 5e5, G, 1e4, [0, 0, 0], 5e24, 6e6, exact_pe=True)
    assert np.array_equal(labels, olab)
    _assert_clumps_match(rows, orc)
    accretes = (rows[:, 1] >= 2) & (rows[:, 17] == 1) & (rows[:, 18] == 1)
    if n_accrete is not None:
        assert accretes.sum() == n_accrete
    if name == "jittered rotator":  # accretion.rs:1087-1111
        assert rows[0, 16] / rows[0, 11] > 0.1
    if name == "rigid rotator":
        assert rows[0, 16] / rows[0, 11] < 1e-9
    if name == "spinning clump":
        l_before = (np.cross(pos, vel) * mass[:, None]).sum(0)
        b = rows[0]
        l_after = b[2] * np.cross(b[3:6], b[6:9]) + b[13:16]
        assert np.linalg.norm(l_after - l_before) / np.linalg.norm(l_before) < 1e-10


@pytest.mark.parametrize("seed", [21, 22])
def test_find_clumps_clustered_field(seed):
    rng = np.random.default_rng(seed)
    parts = []
    for k in range(25):
        c = rng.uniform(-5e7, 5e7, 3)
        nk = int(rng.integers(2, 300))
        p, v, m, r = cold_blob(c, rng.normal(scale=300, size=3), rng.uniform(5e4, 4e5), nk, 1e19 * (1 + k % 3),
                               3000.0)
        parts.append((p + rng.normal(scale=2e4, size=p.shape), v + rng.normal(scale=5.0, size=v.shape), m, r))
    noise = rng.uniform(-6e7, 6e7, (1500, 3))
    parts.append((noise, rng.normal(scale=100, size=(1500, 3)), np.full(1500, 1e18), np.full(1500, 2500.0)))
    pos, vel, mass, rho = (np.concatenate([q[i] for q in parts]) for i in range(4))
    perm = rng.permutation(len(mass))
    pos, vel, mass, rho = pos[perm], vel[perm], mass[perm], rho[perm]
    print(f"clustered field seed {seed}: n={len(mass)}", flush=True)
    with isph.GpuSph(len(mass)) as s:
        rows, labels = _find(s, pos, vel, mass, rho, 1.2e5)
        rows2, _ = _find(s, pos, vel, mass, rho, 1.2e5, min_members=2)
    orc, olab = O.find_clumps(pos, vel, mass, rho, 1.2e5, G, 1e4, [0, 0, 0], 5e24, 6e6, exact_pe=True)
    assert np.array_equal(labels, olab)
    _assert_clumps_match(rows, orc)
    assert np.array_equal(rows2, rows[rows[:, 1] >= 2])


def test_find_clumps_large_clump_self_energy():
    """A 3000-member clump: exact pair sum on the device vs the oracle's exact sum and the reference's
    theta = 0.5 octree branch."""
    rng = np.random.default_rng(7)
    n = 3000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pos = d * (1e6 * np.cbrt(rng.random(n)))[:, None] + np.array([3e7, 0, 0])
    vel = rng.normal(scale=1.0, size=(n, 3))
    mass, rho = np.full(n, 1e18), np.full(n, 3000.0)
    print("large clump n=3000", flush=True)
    with isph.GpuSph(n) as s:
        rows, labels = _find(s, pos, vel, mass, rho, 5e6)
    ex, olab = O.find_clumps(pos, vel, mass, rho, 5e6, G, 1e4, [0, 0, 0], 5e24, 6e6, exact_pe=True)
    ref, _ = O.find_clumps(pos, vel, mass, rho, 5e6, G, 1e4, [0, 0, 0], 5e24, 6e6)
    assert len(rows) == 1 and np.array_equal(labels, olab)
    _assert_clumps_match(rows, ex)
    assert abs(rows[0, 12] - ref[0, 12]) / abs(ref[0, 12]) < 1e-2


def _three(m_c=6.0e24, r=1.0e8):
    mu = G * float(np.float32(m_c))
    v_circ, v_esc = np.float32(np.sqrt(mu / r)), np.float32(np.sqrt(2 * mu / r))
    pos = np.array([[0, 0, 0], [r, 0, 0], [0, r, 0]], np.float32)
    vel = np.array([[0, 0, 0], [0, v_circ, 0], [0, 7 * v_esc, 0]], np.float32)
    return records(pos, vel, np.array([m_c, 1e20, 1e20], np.float32), np.ones(3), np.full(3, 3000.0),
                   np.array([0, 1, 1], np.uint32))


def test_escaping_matter_is_not_counted_as_disk():
    """gpu_sph.rs:2504-2541 through the stepper."""
    with _stepper(_three()) as s:
        (d,) = s.disk_view()
        idx, cid = s.disk_members(0)
        assert s.disk_clumps(0) == []
        js = json.loads(s.disk_stats_json(0))
    assert d.valid and d.n_disk == 1 and list(idx) == [1]
    assert abs(d.disk_mass() - 1e20) / 1e20 < 1e-6 and abs(d.escaped - 1e20) / 1e20 < 1e-6
    assert d.disk_prov0 == 0.0 and js["moon"] == 0.0


def test_disk_view_none_and_empty():
    rec = _three()
    rec["mass"][1:] = 0.0
    with _stepper(rec) as s:
        (d,) = s.disk_view()
        assert not d.valid and d.n_live == 1
        assert s.disk_stats_json(0) == "null"
        assert s.disk_clumps(0) == [] and len(s.disk_members(0)[0]) == 0


def _check_view(d, v, idx_gpu, idx_orc):
    assert d.valid == int(v["valid"]) and d.n_live == int(v["n_live"]) and d.n_disk == int(v["n_disk"])
    assert np.array_equal(idx_gpu, idx_orc)
    com = np.array([v["com_x"], v["com_y"], v["com_z"]])
    vcom = np.array([v["vcom_x"], v["vcom_y"], v["vcom_z"]])
    sc = np.linalg.norm(com) + v["r_remnant"]
    assert np.all(np.abs(np.array(d.com) - com) <= 1e-12 * sc)
    assert np.all(np.abs(np.array(d.v_com) - vcom) <= 1e-12 * (np.linalg.norm(vcom) + np.sqrt(v["mu"] / sc)))
    for a, b in VIEW_KEYS:
        assert getattr(d, a) == pytest.approx(v[b], rel=1e-12, abs=1e-300), a


def _oracle_clumps(rec, v, idx, soft=1e4, link=2.0):
    """DiskView::clumps restated over the oracle's view (gpu_sph.rs:793-809)."""
    if len(idx) < 2:
        return np.zeros((0, 20)), np.zeros(0, np.uint32)
    p = rec["pos"][idx].astype(np.float64)
    w = rec["vel"][idx].astype(np.float64)
    m = rec["mass"][idx].astype(np.float64)
    r = np.maximum(rec["rho"][idx], np.float32(1.0)).astype(np.float64)
    return O.find_clumps(p, w, m, r, link * v["mean_h"], G, soft, [v["com_x"], v["com_y"], v["com_z"]],
                         v["m_remnant"], v["r_remnant"], exact_pe=True)


def _stats_json(v, orc):
    m_moon = 7.342e22
    disk = v["disk_prov0"] + v["disk_other"]
    pct = 100.0 * v["disk_prov0"] / disk if disk > 0 else 0.0
    acc = orc[(orc[:, 1] >= 2) & (orc[:, 17] == 1) & (orc[:, 18] == 1)]
    moon = acc[:, 2].max() if len(acc) else 0.0
    return '{"disk":%.3f,"earth_pct":%.0f,"remnant_km":%.0f,"escaped":%.3f,"moon":%.3f}' % (
        disk / m_moon, pct, v["r_remnant"] / 1e3, v["escaped"] / m_moon, moon / m_moon)


@pytest.mark.parametrize("seed", [31, 32, 33])
def test_disk_view_matches_oracle(seed):
    rec = disk_field(seed)
    print(f"disk field seed {seed}: n={len(rec)}", flush=True)
    v, idx = O.disk_view(rec)
    orc, olab = _oracle_clumps(rec, v, idx)
    with _stepper(rec) as s:
        (d,) = s.disk_view()
        gidx, cid = s.disk_members(0)
        rows = _clump_rows(s.disk_clumps(0))
        js = s.disk_stats_json(0)
    _check_view(d, v, gidx, idx)
    assert np.array_equal(cid, olab)
    _assert_clumps_match(rows, orc)
    assert d.n_clumps == int((orc[:, 1] >= 2).sum()) and d.n_bound == int(((orc[:, 1] >= 2) & (orc[:, 17] == 1)).sum())
    assert d.n_clumps >= 6 and (orc[:, 17] == 1).sum() >= 1
    assert js == _stats_json(v, orc)


def test_disk_view_after_an_impact():
    """The 8192-particle two-body impact after 300 Courant steps, measured on the device and by the oracle
    on the read-back state."""
    sc = scene.two_body_8k()
    print("two-body 8k impact: 300 steps", flush=True)
    with isph.GpuSph(len(sc.particles)) as s:
        s.upload(sc.particles, sc.eos, sc.softening)
        s.set_eos_f64(sc.eos64)
        s.step_cfl(0.3, 300)
        rec = s.readback_sync()
        (d,) = s.disk_view()
        gidx, cid = s.disk_members(0)
        rows = _clump_rows(s.disk_clumps(0))
    v, idx = O.disk_view(rec)
    _check_view(d, v, gidx, idx)
    orc, olab = _oracle_clumps(rec, v, idx)
    assert np.array_equal(cid, olab)
    _assert_clumps_match(rows, orc)


def test_ensemble_disk_view_equals_each_instance_alone():
    """Four instances measured together return, bit for bit, what each returns alone."""
    recs = [disk_field(40 + k, n_core=1500 + 700 * k, offset=(1e7 * k, -3e6 * k, 0), drift=(100.0 * k, 0, 0))
            for k in range(4)]
    print("ensemble disk view: 4 instances", flush=True)
    with isph.GpuSph(sum(len(r) for r in recs)) as s:
        s.upload_ensemble(recs, materials.eos_table(("basalt",)), 1.0e4)
        views = s.disk_view()
        together = [(views[k], s.disk_members(k), _clump_rows(s.disk_clumps(k))) for k in range(4)]
        js = [s.disk_stats_json(k) for k in range(4)]
    assert len(views) == 4
    for k, rec in enumerate(recs):
        with _stepper(rec) as s1:
            (d1,) = s1.disk_view()
            m1 = s1.disk_members(0)
            c1 = _clump_rows(s1.disk_clumps(0))
            j1 = s1.disk_stats_json(0)
        d, m, c = together[k]
        assert bytes(d) == bytes(d1)
        assert np.array_equal(m[0], m1[0]) and np.array_equal(m[1], m1[1])
        assert np.array_equal(c, c1) and js[k] == j1


def test_measurement_is_reproducible_and_partitions_mass():
    """At 262,144 particles: two measurements agree bit for bit; the categories partition the live mass;
    the clumps partition the disk."""
    sc = scene.two_body(262144)
    print("two-body 262k: 20 steps, measured twice", flush=True)
    with isph.GpuSph(len(sc.particles)) as s:
        s.upload(sc.particles, sc.eos, sc.softening)
        s.set_eos_f64(sc.eos64)
        s.step_cfl(0.3, 20)
        (a,) = s.disk_view()
        ca = _clump_rows(s.disk_clumps(0))
        (b,) = s.disk_view()
        cb = _clump_rows(s.disk_clumps(0))
        idx, cid = s.disk_members(0)
        rec = s.readback_sync()
    assert bytes(a) == bytes(b) and np.array_equal(ca, cb)
    live = rec["mass"] > 0
    m = rec["mass"][live].astype(np.float64).sum()
    total = a.disk_prov0 + a.disk_other + a.escaped + a.remnant_prov0 + a.remnant_other
    assert abs(total - m) <= 1e-12 * m
    if a.n_disk >= 2:
        assert abs(ca[:, 2].sum() - a.disk_mass()) <= 1e-12 * a.disk_mass()
        assert int(ca[:, 1].sum()) == a.n_disk == len(idx)
