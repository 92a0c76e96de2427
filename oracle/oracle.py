"""ctypes front end of the CPU oracle (oracle/liboracle.so). TEST INFRASTRUCTURE ONLY.

Imported by tests/, by `__graft_entry__.smoke()` and by bench.py's CPU-baseline leg, as the checker.
The product library never loads it. Function-by-function citations are in isph_oracle.cpp.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_D = ctypes.c_double
_SIGS = {
    "orc_bbox_key": (ctypes.c_uint32, [ctypes.c_float]),
    "orc_bbox_decode": (ctypes.c_float, [ctypes.c_uint32]),
    "orc_expand_bits": (ctypes.c_uint32, [ctypes.c_uint32]),
    "orc_bbox_f32": (None, [_P, _U64, _P, _P]),
    "orc_morton_f32": (None, [_P, _U64, _P, _P, _P]),
    "orc_sort_pairs": (None, [_P, _U64, _P, _P]),
    "orc_cluster_codes": (None, [_P, _U64, ctypes.c_uint32, _P]),
    "orc_karras": (None, [_P, ctypes.c_uint32, _P, _P, _P]),
    "orc_karras_variant": (None, [_P, ctypes.c_uint32, _P, _P, _P, ctypes.c_int]),
    "orc_upsweep_f32": (None, [_P, _U64, ctypes.c_uint32, ctypes.c_uint32, _P, _P, _P, _P, _P]),
    "orc_lbvh_gravity_f32": (None, [_P, _U64, ctypes.c_uint32, ctypes.c_uint32, _P, _P, _P, _P, _P,
                                    ctypes.c_float, ctypes.c_float, _P, _P]),
    "orc_lbvh_walk_counts": (None, [_P, _U64, ctypes.c_uint32, ctypes.c_uint32, _P, _P, _P, _P, _P,
                                    ctypes.c_float, ctypes.c_float, _P, _U64, _P]),
    "orc_set_threads": (None, [ctypes.c_int]),
    "orc_get_threads": (ctypes.c_int, []),
    "orc_direct_gravity_f64": (None, [_P, _P, _U64, _D, _P]),
    "orc_lbvh_accel_f64": (None, [_P, _P, _U64, _D, _D, _P]),
    "orc_dual_gradient": (None, [_U64, _P, _P, _P, _P, _P, _P, ctypes.c_uint32, _P, _D, ctypes.c_int, _P,
                                 ctypes.c_uint32, _P, _P, _P, ctypes.c_uint32, _P, _P, _P]),
    "orc_octree_accel": (None, [_P, _P, _U64, _D, _D, _P]),
    "orc_octree_potential": (None, [_P, _P, _U64, _D, _D, _P]),
    "orc_grid_pairs": (_U64, [_P, _U64, _D, _P, _U64]),
    "orc_sph_w": (_D, [_D, _D]),
    "orc_sph_dw": (_D, [_D, _D]),
    "orc_till_pressure": (_D, [_P, _D, _D]),
    "orc_till_sound_speed_sq": (_D, [_P, _D, _D]),
    "orc_hydro_new": (_P, [_U64, _P, _P, _P, _P, _P, _P, ctypes.c_uint32, _P, _D]),
    "orc_hydro_free": (None, [_P]),
    "orc_hydro_set": (None, [_P, _D, _D, _D, ctypes.c_int]),
    "orc_hydro_compute_density": (None, [_P]),
    "orc_hydro_forces": (None, [_P, _P, _P]),
    "orc_hydro_step": (None, [_P, _D]),
    "orc_hydro_relax": (None, [_P, _D, _D, ctypes.c_uint32]),
    "orc_hydro_phase_ms": (None, [_P, _P]),
    "orc_hydro_set_state": (None, [_P, _P, _P]),
    "orc_hydro_courant_dt": (_D, [_P, _D]),
    "orc_hydro_get": (None, [_P, _P, _P, _P, _P]),
    "orc_hydro_eos": (None, [_P, _P, _P]),
    "orc_brute_sph": (None, [_U64, _P, _P, _P, _P, _P, _P, ctypes.c_uint32, _P, _D, _D, _P, _P, _P]),
    "orc_total_energy": (None, [_U64, _P, _P, _P, _P, _D, _P]),
    "orc_sum_oi": (_D, [_P, _U64]),
    "orc_angular_momentum": (None, [_U64, _P, _P, _P, _P]),
    "orc_orbit_is_bound": (ctypes.c_bool, [_P, _P, _D]),
    "orc_orbit_perigee": (ctypes.c_int, [_P, _P, _D, _P]),
    "orc_disk_view": (None, [_P, _U64, _P, _P]),
    "orc_find_clumps": (_U64, [_P, _P, _P, _P, _U64, _D, _D, _D, _P, _D, _D, ctypes.c_int, _P, _P]),
}

# orc_disk_view record fields
DISK_FIELDS = ("valid", "n_live", "com_x", "com_y", "com_z", "vcom_x", "vcom_y", "vcom_z", "r_remnant",
               "m_remnant", "mu", "disk_prov0", "disk_other", "escaped", "mean_h", "remnant_prov0",
               "remnant_other", "escaped_prov0", "n_disk", "m_total")
# orc_find_clumps record fields
CLUMP_FIELDS = ("id", "n_members", "mass", "com_x", "com_y", "com_z", "vel_x", "vel_y", "vel_z", "rho", "radius",
                "internal_ke", "self_pe", "L_x", "L_y", "L_z", "thermal_ke", "bound", "outside_roche", "_r")

_lib = None


def build():
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        l = ctypes.CDLL(LIB)
        for k, (r, a) in _SIGS.items():
            f = getattr(l, k)
            f.restype = r
            f.argtypes = a
        _lib = l
    return _lib


def set_threads(t):
    """Threads of the per-target gravity loops (results do not depend on it)."""
    lib().orc_set_threads(int(t))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def c64(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ---- LBVH leg (f32 / integer) ----
def bbox_f32(pos):
    p = c64(pos, np.float32).reshape(-1, 3)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    lib().orc_bbox_f32(_p(p), p.shape[0], _p(lo), _p(hi))
    return lo, hi


def morton_f32(pos, lo, hi):
    p = c64(pos, np.float32).reshape(-1, 3)
    codes = np.zeros(p.shape[0], np.uint32)
    lib().orc_morton_f32(_p(p), p.shape[0], _p(c64(lo, np.float32)), _p(c64(hi, np.float32)), _p(codes))
    return codes


def sort_pairs(codes):
    c = c64(codes, np.uint32)
    sc = np.zeros_like(c)
    order = np.zeros_like(c)
    lib().orc_sort_pairs(_p(c), c.shape[0], _p(sc), _p(order))
    return sc, order


def cluster_codes(sorted_codes, K):
    n = sorted_codes.shape[0]
    L = (n + K - 1) // K
    out = np.zeros(L, np.uint32)
    lib().orc_cluster_codes(_p(c64(sorted_codes, np.uint32)), n, K, _p(out))
    return out


def karras(leaf_codes, reference_split=False):
    """Karras build; reference_split=True keeps the shader's midpoint split for all-equal-code ranges."""
    c = c64(leaf_codes, np.uint32)
    L = c.shape[0]
    left = np.zeros(max(L - 1, 1), np.uint32)
    right = np.zeros(max(L - 1, 1), np.uint32)
    parent = np.zeros(2 * L - 1, np.uint32)
    lib().orc_karras_variant(_p(c), L, _p(left), _p(right), _p(parent), int(bool(reference_split)))
    return left[:L - 1], right[:L - 1], parent


def upsweep_f32(sbodies, K, left, right):
    sb = c64(sbodies, np.float32).reshape(-1, 4)
    n = sb.shape[0]
    L = (n + K - 1) // K
    nodes = 2 * L - 1
    com = np.zeros((nodes, 4), np.float32)
    bmin = np.zeros((nodes, 3), np.float32)
    bmax = np.zeros((nodes, 3), np.float32)
    l = c64(left, np.uint32) if len(left) else np.zeros(1, np.uint32)
    r = c64(right, np.uint32) if len(right) else np.zeros(1, np.uint32)
    lib().orc_upsweep_f32(_p(sb), n, K, L, _p(l), _p(r), _p(com), _p(bmin), _p(bmax))
    return com, bmin, bmax


def lbvh_gravity_f32(pos, mass, soft, theta=0.5, K=1):
    """The full reference LBVH leg: bbox -> Morton -> sort -> Karras -> moments -> cs_gravity_bh."""
    p = c64(pos, np.float32).reshape(-1, 3)
    m = c64(mass, np.float32)
    lo, hi = bbox_f32(p)
    codes = morton_f32(p, lo, hi)
    sc, order = sort_pairs(codes)
    lc = cluster_codes(sc, K)
    left, right, _ = karras(lc)
    sb = np.concatenate([p[order], m[order, None]], axis=1).astype(np.float32)
    com, bmin, bmax = upsweep_f32(sb, K, left, right)
    n = p.shape[0]
    L = lc.shape[0]
    acc = np.zeros((n, 3), np.float32)
    l = c64(left, np.uint32) if len(left) else np.zeros(1, np.uint32)
    r = c64(right, np.uint32) if len(right) else np.zeros(1, np.uint32)
    lib().orc_lbvh_gravity_f32(_p(sb), n, K, L, _p(l), _p(r), _p(com), _p(c64(bmin, np.float32)),
                               _p(c64(bmax, np.float32)), float(theta), float(np.float32(soft) * np.float32(soft)),
                               _p(c64(order, np.uint32)), _p(acc))
    return acc


def lbvh_walk_counts(pos, mass, soft, targets, theta=0.5):
    """The reference walk's interactions for the given targets (original indices): (cells, particle terms)
    per target, from the same LBVH leg as lbvh_gravity_f32 (leaf size 1)."""
    p = c64(pos, np.float32).reshape(-1, 3)
    m = c64(mass, np.float32)
    lo, hi = bbox_f32(p)
    sc, order = sort_pairs(morton_f32(p, lo, hi))
    lc = cluster_codes(sc, 1)
    left, right, _ = karras(lc)
    sb = np.concatenate([p[order], m[order, None]], axis=1).astype(np.float32)
    com, bmin, bmax = upsweep_f32(sb, 1, left, right)
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order), dtype=order.dtype)
    tgt = c64(inv[np.asarray(targets)], np.uint32)
    counts = np.zeros((len(tgt), 2), np.uint64)
    l = c64(left, np.uint32) if len(left) else np.zeros(1, np.uint32)
    r = c64(right, np.uint32) if len(right) else np.zeros(1, np.uint32)
    lib().orc_lbvh_walk_counts(_p(sb), p.shape[0], 1, lc.shape[0], _p(l), _p(r), _p(com), _p(c64(bmin, np.float32)),
                               _p(c64(bmax, np.float32)), float(theta), float(np.float32(soft) * np.float32(soft)),
                               _p(tgt), len(tgt), _p(counts))
    return counts


def lbvh_accel(pos, mass, soft, theta=0.5):
    """lbvh_gravity_f32 as one C call (f64 in and out), the gravity of the CPU step's "lbvh" mode."""
    p = c64(pos).reshape(-1, 3)
    acc = np.zeros_like(p)
    lib().orc_lbvh_accel_f64(_p(p), _p(c64(mass)), p.shape[0], float(soft), float(theta), _p(acc))
    return acc


def dual_gradient(pos, vel, mass, u, h, mat, eos64, soft, dts, tangents, eos_mat, loss_params,
                  gravity="direct"):
    """Loss after len(dts) steps of the f64 CPU step and its exact derivative (forward-mode duals) along the
    three state tangents [(dpos, dvel, dmass)] and material eos_mat's Tillotson a, b, A, B. Returns
    (loss, grad[7])."""
    n = len(mass)
    tp = np.ascontiguousarray(np.stack([t[0] for t in tangents]), np.float64)
    tv = np.ascontiguousarray(np.stack([t[1] for t in tangents]), np.float64)
    tm = np.ascontiguousarray(np.stack([t[2] for t in tangents]), np.float64)
    eos = c64(eos64).reshape(-1, 10)
    d = c64(dts)
    lp = c64(loss_params)
    out = np.zeros(1)
    grad = np.zeros(7)
    lib().orc_dual_gradient(n, _p(c64(pos)), _p(c64(vel)), _p(c64(mass)), _p(c64(u)), _p(c64(h)),
                            _p(np.ascontiguousarray(mat, np.uint32)), eos.shape[0], _p(eos), float(soft),
                            {"direct": 1, "none": 2}[gravity], _p(d), d.shape[0], _p(tp), _p(tv), _p(tm),
                            int(eos_mat), _p(lp), _p(out), _p(grad))
    return out[0], grad


# ---- f64 physics ----
def direct_gravity(pos, mass, soft):
    p = c64(pos).reshape(-1, 3)
    acc = np.zeros_like(p)
    lib().orc_direct_gravity_f64(_p(p), _p(c64(mass)), p.shape[0], float(soft), _p(acc))
    return acc


def octree_accel(pos, mass, theta, soft):
    p = c64(pos).reshape(-1, 3)
    acc = np.zeros_like(p)
    lib().orc_octree_accel(_p(p), _p(c64(mass)), p.shape[0], float(theta), float(soft), _p(acc))
    return acc


def octree_potential(pos, mass, theta, soft):
    p = c64(pos).reshape(-1, 3)
    phi = np.zeros(p.shape[0])
    lib().orc_octree_potential(_p(p), _p(c64(mass)), p.shape[0], float(theta), float(soft), _p(phi))
    return phi


def grid_pairs(pos, cell, cap=None):
    p = c64(pos).reshape(-1, 3)
    n = p.shape[0]
    cap = cap or n * 64
    buf = np.zeros((cap, 2), np.uint32)
    cnt = lib().orc_grid_pairs(_p(p), n, float(cell), _p(buf), cap)
    if cnt > cap:
        return grid_pairs(pos, cell, int(cnt))
    return buf[:cnt]


def sph_w(r, h):
    return lib().orc_sph_w(float(r), float(h))


def sph_dw(r, h):
    return lib().orc_sph_dw(float(r), float(h))


def till_pressure(params, rho, u):
    return lib().orc_till_pressure(_p(c64(params)), float(rho), float(u))


def till_c2(params, rho, u):
    return lib().orc_till_sound_speed_sq(_p(c64(params)), float(rho), float(u))


class Hydro:
    """HydroBody (hydrostatic.rs:42) over f64 state."""

    def __init__(self, pos, vel, mass, u, h, mat, eos64, softening):
        self.n = len(mass)
        self._keep = [c64(pos).reshape(-1, 3), c64(vel).reshape(-1, 3), c64(mass), c64(u), c64(h),
                      c64(mat, np.uint32), c64(eos64).reshape(-1, 10)]
        a = self._keep
        self.h = lib().orc_hydro_new(self.n, _p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), _p(a[4]), _p(a[5]),
                                     a[6].shape[0], _p(a[6]), float(softening))

    def close(self):
        if self.h:
            lib().orc_hydro_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def configure(self, theta=0.5, av_alpha=1.0, av_beta=2.0, gravity="octree"):
        mode = {"octree": 0, "direct": 1, "none": 2, "lbvh": 3}[gravity]
        lib().orc_hydro_set(self.h, theta, av_alpha, av_beta, mode)

    def compute_density(self):
        lib().orc_hydro_compute_density(self.h)

    def forces(self):
        acc = np.zeros((self.n, 3))
        du = np.zeros(self.n)
        lib().orc_hydro_forces(self.h, _p(acc), _p(du))
        return acc, du

    def step(self, dt):
        lib().orc_hydro_step(self.h, float(dt))

    def courant_dt(self, cfl):
        return lib().orc_hydro_courant_dt(self.h, float(cfl))

    def phase_ms(self):
        """Accumulated wall ms of the CPU step's phases: density, gravity, SPH force + du/dt."""
        out = np.zeros(3)
        lib().orc_hydro_phase_ms(self.h, _p(out))
        return {"density": out[0], "gravity": out[1], "sph_force": out[2]}

    def relax(self, dt, damp, steps=1):
        """`steps` damped relaxation steps (HydroBody::relax_step, hydrostatic.rs:386-393)."""
        lib().orc_hydro_relax(self.h, float(dt), float(damp), int(steps))

    def set_state(self, pos, vel):
        p = c64(pos).reshape(-1, 3)
        v = c64(vel).reshape(-1, 3)
        assert p.shape[0] == self.n and v.shape[0] == self.n
        lib().orc_hydro_set_state(self.h, _p(p), _p(v))

    def state(self):
        pos = np.zeros((self.n, 3))
        vel = np.zeros((self.n, 3))
        u = np.zeros(self.n)
        rho = np.zeros(self.n)
        lib().orc_hydro_get(self.h, _p(pos), _p(vel), _p(u), _p(rho))
        return pos, vel, u, rho

    def eos(self):
        p = np.zeros(self.n)
        c = np.zeros(self.n)
        lib().orc_hydro_eos(self.h, _p(p), _p(c))
        return p, c


def brute_sph(pos, vel, mass, u, h, mat, eos64, av_alpha=1.0, av_beta=2.0):
    n = len(mass)
    rho = np.zeros(n)
    acc = np.zeros((n, 3))
    du = np.zeros(n)
    e = c64(eos64).reshape(-1, 10)
    lib().orc_brute_sph(n, _p(c64(pos).reshape(-1, 3)), _p(c64(vel).reshape(-1, 3)), _p(c64(mass)), _p(c64(u)),
                        _p(c64(h)), _p(c64(mat, np.uint32)), e.shape[0], _p(e), av_alpha, av_beta, _p(rho),
                        _p(acc), _p(du))
    return rho, acc, du


def total_energy(pos, vel, mass, u, softening):
    out = np.zeros(3)
    lib().orc_total_energy(len(mass), _p(c64(pos).reshape(-1, 3)), _p(c64(vel).reshape(-1, 3)), _p(c64(mass)),
                           _p(c64(u)), float(softening), _p(out))
    return out


def angular_momentum(pos, vel, mass):
    out = np.zeros(3)
    lib().orc_angular_momentum(len(mass), _p(c64(pos).reshape(-1, 3)), _p(c64(vel).reshape(-1, 3)),
                               _p(c64(mass)), _p(out))
    return out


# ---- measurement: orbit.rs:91-117, gpu_sph.rs:689-785, accretion.rs:158-301 ----
def orbit_is_bound(p, v, mu):
    return bool(lib().orc_orbit_is_bound(_p(c64(p)), _p(c64(v)), float(mu)))


def orbit_perigee(p, v, mu):
    """orbit::perigee: the periapsis distance, or None."""
    q = np.zeros(1)
    ok = lib().orc_orbit_perigee(_p(c64(p)), _p(c64(v)), float(mu), _p(q))
    return float(q[0]) if ok else None


def disk_view(particles):
    """DiskView::of over a PARTICLE_DTYPE record array: (dict of DISK_FIELDS, disk member indices)."""
    a = np.ascontiguousarray(particles)
    assert a.dtype.itemsize == 48
    view = np.zeros(len(DISK_FIELDS))
    idx = np.zeros(max(len(a), 1), np.uint32)
    lib().orc_disk_view(_p(a), len(a), _p(view), _p(idx))
    return dict(zip(DISK_FIELDS, view.tolist())), idx[:int(view[18])].copy()


def find_clumps(pos, vel, mass, rho, linking_length, g, softening, central_pos, central_mass, central_radius,
                exact_pe=False):
    """accretion::find_clumps: (records (k, 20) heaviest first, labels). exact_pe uses the exact pair sum
    at every clump size instead of the reference's octree above 1024 members."""
    p = c64(pos).reshape(-1, 3)
    n = p.shape[0]
    rec = np.zeros((max(n, 1), len(CLUMP_FIELDS)))
    labels = np.zeros(max(n, 1), np.uint32)
    k = lib().orc_find_clumps(_p(p), _p(c64(vel).reshape(-1, 3)), _p(c64(mass)), _p(c64(rho)), n,
                              float(linking_length), float(g), float(softening), _p(c64(central_pos)),
                              float(central_mass), float(central_radius), 1 if exact_pe else 0, _p(rec),
                              _p(labels))
    return rec[:k].copy(), labels[:n].copy()
