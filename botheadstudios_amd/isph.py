"""Python mirror of the engine's `GpuSph` / `GpuGravity` host API over the C-ABI library (libisph.so).

Method names, argument order and clamp / `None` semantics follow `crates/engine/src/gpu_sph.rs:1115-1505`
and `crates/engine/src/gpu_gravity.rs:92-230`; the wgpu device / queue / encoder arguments are replaced
by the handle, which owns one HIP stream. All compute runs in the HIP library: importing this module on a
machine without the built library raises, and there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from .materials import EOS_DTYPE
from .scene import PARTICLE_DTYPE

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ISPH_LIBRARY", os.path.join(_HERE, "libisph.so"))

OK = 0
EAGAIN = -11
EINVAL = -22
GRAVITY_TREE, GRAVITY_DIRECT, GRAVITY_NONE = 0, 1, 2
CACHE_ALL = 0xFFFFFFFF
PRECISION_FP32, PRECISION_MIXED, PRECISION_F64 = 0, 1, 2
PHASES = ("sort", "tree", "upsweep", "gravity", "density", "force", "integrate")

# Every symbol include/isph.h declares, with its ctypes signature.
_P = ctypes.c_void_p
_SIGS = {
    "isph_create": (_P, [ctypes.c_uint32, ctypes.c_int]),
    "isph_destroy": (None, [_P]),
    "isph_last_error": (ctypes.c_char_p, [_P]),
    "isph_upload": (ctypes.c_int, [_P, _P, ctypes.c_size_t, _P, ctypes.c_size_t, ctypes.c_float]),
    "isph_set_eos_f64": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "isph_count": (ctypes.c_uint32, [_P]),
    "isph_set_dt": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_float]),
    "isph_set_av": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_float]),
    "isph_set_gravity": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_float, ctypes.c_uint32, ctypes.c_int]),
    "isph_set_precision": (ctypes.c_int, [_P, ctypes.c_int]),
    "isph_set_walk": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_int, ctypes.c_int]),
    "isph_set_candidate_cache": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "isph_gravity_direct_noalloc": (ctypes.c_int, [_P, _P, ctypes.c_size_t, ctypes.c_double, _P, _P]),
    "isph_set_reuse": (ctypes.c_int, [_P, ctypes.c_int]),
    "isph_set_graph": (ctypes.c_int, [_P, ctypes.c_int]),
    "isph_set_bulk": (ctypes.c_int, [_P] + [ctypes.c_double] * 8),
    "isph_set_bulk_planar": (ctypes.c_int, [_P] + [ctypes.c_double] * 7),
    "isph_set_external_masses": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "isph_set_merge_budget": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "isph_encode_merge": (ctypes.c_int, [_P]),
    "isph_encode_relax": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "isph_set_omega": (ctypes.c_int, [_P, ctypes.c_float]),
    "isph_autotune": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_uint32, ctypes.c_char_p, _P]),
    "isph_encode_kdk": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "isph_step_cfl": (ctypes.c_int, [_P, ctypes.c_float, ctypes.c_uint32]),
    "isph_sim_time": (ctypes.c_int, [_P, _P, _P]),
    "isph_begin_readback": (ctypes.c_int, [_P]),
    "isph_take_readback": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "isph_readback_sync": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "isph_synchronize": (ctypes.c_int, [_P]),
    "isph_particle_buffer": (ctypes.c_int, [_P, _P, _P]),
    "isph_stream": (_P, [_P]),
    "isph_force_eval": (ctypes.c_int, [_P, _P, _P]),
    "isph_gravity_accelerations": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, ctypes.c_double, _P, ctypes.c_int]),
    "isph_diagnostics": (ctypes.c_int, [_P, ctypes.c_int, _P]),
    "isph_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "isph_phase_times": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "isph_kernel_stats": (ctypes.c_int, [_P, _P]),
    "isph_reset_stats": (ctypes.c_int, [_P]),
    "isph_walk_stats": (ctypes.c_int, [_P, _P]),
    "isph_probe_morton": (ctypes.c_int, [_P, _P, _P, _P]),
    "isph_probe_sort": (ctypes.c_int, [_P, _P, _P]),
    "isph_probe_tree": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "isph_upload_ensemble": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _P, ctypes.c_size_t, ctypes.c_float]),
    "isph_ensemble_times": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32]),
    "isph_loss_value": (ctypes.c_int, [_P, _P, _P]),
    "isph_set_state_f64": (ctypes.c_int, [_P, _P, _P]),
    "isph_state_f64": (ctypes.c_int, [_P, _P, _P]),
    "isph_clamp_energy": (ctypes.c_int, [_P, _P]),
    "isph_vjp_run": (ctypes.c_int, [_P, ctypes.c_float, _P, ctypes.c_uint32, ctypes.c_uint32, _P, _P, _P, _P, _P,
                                    _P, _P, _P, _P]),
    "isph_set_measure": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double]),
    "isph_disk_view": (ctypes.c_int, [_P, _P, ctypes.c_uint32]),
    "isph_disk_members": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P, ctypes.c_uint32]),
    "isph_disk_clumps": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, _P, ctypes.c_uint32]),
    "isph_disk_stats_json": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_size_t]),
    "isph_find_clumps": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_size_t, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, _P, ctypes.c_double, ctypes.c_double, ctypes.c_uint32, _P,
                                        ctypes.c_uint32, _P]),
}


class Tuning(ctypes.Structure):
    """`isph_tuning`: the walk shape chosen by the autotuner."""
    _fields_ = [("leaf_size", ctypes.c_uint32), ("group_extent", ctypes.c_float),
                ("grav_particles_per_lane", ctypes.c_int), ("sph_expand", ctypes.c_int),
                ("ms_per_step", ctypes.c_double)]


class Loss(ctypes.Structure):
    """`isph_loss`: smoothed mass outside r_cut radii whose specific angular momentum exceeds the circular
    value (SI units)."""
    _fields_ = [("radius", ctypes.c_double), ("r_cut", ctypes.c_double), ("k_r", ctypes.c_double),
                ("k_j", ctypes.c_double), ("axis", ctypes.c_double * 3)]

    @classmethod
    def make(cls, radius, r_cut=2.9, k_r=20.0, k_j=10.0, axis=(0.0, 0.0, 1.0)):
        return cls(float(radius), float(r_cut), float(k_r), float(k_j), (ctypes.c_double * 3)(*axis))

class Disk(ctypes.Structure):
    """`isph_disk`: DiskView (gpu_sph.rs:661-686) of one instance plus impact-run's per-source split."""
    _fields_ = [("valid", ctypes.c_int32), ("n_live", ctypes.c_uint32), ("n_disk", ctypes.c_uint32),
                ("n_clumps", ctypes.c_uint32), ("n_bound", ctypes.c_uint32), ("_pad", ctypes.c_uint32),
                ("com", ctypes.c_double * 3), ("v_com", ctypes.c_double * 3), ("r_remnant", ctypes.c_double),
                ("m_remnant", ctypes.c_double), ("mu", ctypes.c_double), ("disk_prov0", ctypes.c_double),
                ("disk_other", ctypes.c_double), ("escaped", ctypes.c_double), ("mean_h", ctypes.c_double),
                ("remnant_prov0", ctypes.c_double), ("remnant_other", ctypes.c_double),
                ("escaped_prov0", ctypes.c_double), ("moon_mass", ctypes.c_double)]

    def disk_mass(self):
        """gpu_sph.rs:788-790."""
        return self.disk_prov0 + self.disk_other


class Clump(ctypes.Structure):
    """`isph_clump`: accretion.rs:37-51 without the member list."""
    _fields_ = [("id", ctypes.c_uint32), ("n_members", ctypes.c_uint32), ("bound", ctypes.c_int32),
                ("outside_roche", ctypes.c_int32), ("mass", ctypes.c_double), ("com_pos", ctypes.c_double * 3),
                ("com_vel", ctypes.c_double * 3), ("rho", ctypes.c_double), ("radius", ctypes.c_double),
                ("internal_ke", ctypes.c_double), ("self_pe", ctypes.c_double), ("ang_mom", ctypes.c_double * 3),
                ("thermal_ke", ctypes.c_double)]

    def accretes(self):
        """accretion.rs:55-57."""
        return self.n_members >= 2 and bool(self.bound) and bool(self.outside_roche)


_lib = None
_warned_runtimes = False


def hip_runtimes(maps_text=None):
    """The HIP runtime libraries (libamdhip64) mapped into this process, as distinct file paths.

    libisph.so binds to the HIP runtime already loaded under the soname libamdhip64.so.7, else to the
    system's. A Python package that ships its own runtime and asks for it by another name (torch's wheel loads
    its bundled libamdhip64.so this way) then adds a second runtime, with its own HSA runtime, when it is
    imported after this library. Streams and device pointers of one runtime mean nothing to the other."""
    if maps_text is None:
        try:
            with open("/proc/self/maps") as f:
                maps_text = f.read()
        except OSError:
            return []
    found = set()
    for line in maps_text.splitlines():
        parts = line.split(None, 5)
        if len(parts) == 6 and os.path.basename(parts[5]).startswith("libamdhip64.so"):
            found.add(os.path.realpath(parts[5]))
    return sorted(found)


def _warn_if_two_runtimes():
    global _warned_runtimes
    if _warned_runtimes:
        return
    libs = hip_runtimes()
    if len(libs) > 1:
        import warnings
        _warned_runtimes = True
        warnings.warn("two HIP runtimes are loaded in this process (" + ", ".join(libs) + "): device pointers "
                      "and streams of one are invalid in the other. Import torch before this library when both "
                      "are used.", RuntimeWarning, stacklevel=3)


def load_library(path=LIB_PATH):
    """Load libisph.so (raises OSError when it is not built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OSError(f"{path} is not built; run __graft_entry__.build() or botheadstudios_amd/build.py")
        lib = ctypes.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class IsphError(RuntimeError):
    pass


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class GpuSph:
    """The GPU SPH stepper (gpu_sph.rs:1086). One handle = one HIP stream on one device."""

    def __init__(self, capacity, device=0):
        self._lib = load_library()
        _warn_if_two_runtimes()
        self._h = self._lib.isph_create(int(capacity), int(device))
        if not self._h:
            raise IsphError("isph_create failed: " + self._lib.isph_last_error(None).decode())
        self.capacity = max(int(capacity), 1)

    # -- lifetime --
    def close(self):
        if getattr(self, "_h", None):
            self._lib.isph_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc < 0 and rc != EAGAIN:
            raise IsphError(f"isph error {rc}: {self._lib.isph_last_error(self._h).decode()}")
        return rc

    # -- GpuSph surface --
    def upload(self, particles, eos, softening):
        """gpu_sph.rs:1323: returns the count kept (clamped to capacity)."""
        p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        e = np.ascontiguousarray(eos, dtype=EOS_DTYPE)
        return self._check(self._lib.isph_upload(self._h, _ptr(p), p.shape[0], _ptr(e), e.shape[0],
                                                 float(softening)))

    def set_eos_f64(self, params):
        a = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, 10)
        self._check(self._lib.isph_set_eos_f64(self._h, _ptr(a), a.shape[0]))

    def count(self):
        return int(self._lib.isph_count(self._h))

    def set_dt(self, dt, damp=1.0):
        self._check(self._lib.isph_set_dt(self._h, float(dt), float(damp)))

    def set_av(self, alpha, beta):
        self._check(self._lib.isph_set_av(self._h, float(alpha), float(beta)))

    def set_gravity(self, mode=GRAVITY_TREE, theta=0.5, leaf_size=1, quadrupole=False):
        self._check(self._lib.isph_set_gravity(self._h, int(mode), float(theta), int(leaf_size), int(bool(quadrupole))))

    def set_walk(self, group_extent=4.0, grav_particles_per_lane=1, sph_expand=4):
        self._check(self._lib.isph_set_walk(self._h, float(group_extent), int(grav_particles_per_lane),
                                            int(sph_expand)))

    def set_candidate_cache(self, max_blocks=CACHE_ALL):
        """Cap the SPH candidate-cache blocks in use (0: the force pass walks every group again)."""
        self._check(self._lib.isph_set_candidate_cache(self._h, int(max_blocks)))

    def set_precision(self, mode):
        self._check(self._lib.isph_set_precision(self._h, int(mode)))

    def set_bulk(self, center, r_core, vel, mass):
        self._check(self._lib.isph_set_bulk(self._h, *[float(x) for x in center], float(r_core),
                                            *[float(x) for x in vel], float(mass)))

    def set_bulk_planar(self, plane_point, up, g):
        self._check(self._lib.isph_set_bulk_planar(self._h, *[float(x) for x in plane_point],
                                                   *[float(x) for x in up], float(g)))

    def set_external_masses(self, bodies):
        """bodies: sequence of (position (3,), mass); returns how many were kept (at most 32)."""
        arr = np.ascontiguousarray([[p[0], p[1], p[2], m] for p, m in bodies], np.float64).reshape(-1, 4)
        rc = self._lib.isph_set_external_masses(self._h, _ptr(arr), arr.shape[0])
        self._check(rc)
        return rc

    def set_merge_budget(self, budget):
        self._check(self._lib.isph_set_merge_budget(self._h, int(budget)))

    def encode_merge(self):
        self._check(self._lib.isph_encode_merge(self._h))

    def encode_relax(self, steps):
        self._check(self._lib.isph_encode_relax(self._h, int(steps)))

    def set_omega(self, omega):
        self._check(self._lib.isph_set_omega(self._h, float(omega)))

    def set_graph(self, enable):
        self._check(self._lib.isph_set_graph(self._h, int(bool(enable))))

    def autotune(self, cfl=0.3, trial_steps=3, cache_path=None):
        """Tune the walk shape on the current state; returns (from_cache, Tuning)."""
        t = Tuning()
        rc = self._lib.isph_autotune(self._h, float(cfl), int(trial_steps),
                                     None if cache_path is None else str(cache_path).encode(), ctypes.byref(t))
        self._check(rc)
        return rc == 1, t

    def set_reuse(self, enable):
        self._check(self._lib.isph_set_reuse(self._h, int(bool(enable))))

    def encode_kdk(self, substeps):
        """gpu_sph.rs:1484: enqueue `substeps` KDK steps at the set_dt dt."""
        self._check(self._lib.isph_encode_kdk(self._h, int(substeps)))

    def step_cfl(self, cfl, steps):
        """Adaptive-dt KDK steps (impact-run main.rs:394-434) with the Courant min on the device."""
        self._check(self._lib.isph_step_cfl(self._h, float(cfl), int(steps)))

    def sim_time(self):
        t = np.zeros(1, np.float64)
        dt = np.zeros(1, np.float32)
        self._check(self._lib.isph_sim_time(self._h, _ptr(t), _ptr(dt)))
        return float(t[0]), float(dt[0])

    def begin_readback(self):
        self._check(self._lib.isph_begin_readback(self._h))

    def take_readback(self):
        """gpu_sph.rs:1502: the snapshot, or None while pending / idle."""
        out = np.zeros(self.capacity, dtype=PARTICLE_DTYPE)
        rc = self._check(self._lib.isph_take_readback(self._h, _ptr(out), out.shape[0]))
        if rc == EAGAIN:
            return None
        return out[:rc]

    def readback_sync(self):
        out = np.zeros(self.capacity, dtype=PARTICLE_DTYPE)
        rc = self._check(self._lib.isph_readback_sync(self._h, _ptr(out), out.shape[0]))
        return out[:rc]

    def particle_buffer(self):
        """gpu_sph.rs:1452-1456: (device address, count) of the packed 48-byte particle records."""
        ptr = ctypes.c_void_p(0)
        cnt = ctypes.c_uint32(0)
        self._check(self._lib.isph_particle_buffer(self._h, ctypes.byref(ptr), ctypes.byref(cnt)))
        return ptr.value, cnt.value

    def stream(self):
        """The handle's hipStream_t as an integer."""
        return self._lib.isph_stream(self._h)

    def synchronize(self):
        self._check(self._lib.isph_synchronize(self._h))

    def force_eval(self):
        """One evaluation: returns (acc (N,3) f32, dudt (N,) f32, rho (N,) f32) in upload order."""
        n = self.count()
        acc = np.zeros((n, 4), np.float32)
        rho = np.zeros(n, np.float32)
        self._check(self._lib.isph_force_eval(self._h, _ptr(acc), _ptr(rho)))
        return acc[:, :3].copy(), acc[:, 3].copy(), rho

    # ---- ensembles ----
    def upload_ensemble(self, instances, eos, softening):
        """Upload a list of record arrays as independent instances of one ensemble; returns the offsets."""
        recs = [np.ascontiguousarray(r, dtype=PARTICLE_DTYPE) for r in instances]
        offsets = np.zeros(len(recs) + 1, np.uint32)
        offsets[1:] = np.cumsum([len(r) for r in recs])
        allp = np.concatenate(recs)
        e = np.ascontiguousarray(eos, dtype=EOS_DTYPE)
        rc = self._lib.isph_upload_ensemble(self._h, _ptr(allp), _ptr(offsets), len(recs), _ptr(e), len(e),
                                            float(softening))
        self._check(rc)
        self._n_inst = len(recs)
        return offsets

    def ensemble_times(self):
        k = getattr(self, "_n_inst", 1)
        t = np.zeros(k, np.float64)
        dt = np.zeros(k, np.float32)
        self._check(self._lib.isph_ensemble_times(self._h, _ptr(t), _ptr(dt), k))
        return t, dt

    # ---- adjoint ----
    def loss_value(self, loss):
        out = ctypes.c_double(0.0)
        self._check(self._lib.isph_loss_value(self._h, ctypes.byref(loss), ctypes.byref(out)))
        return out.value

    def set_state_f64(self, pos, vel):
        p = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
        self._check(self._lib.isph_set_state_f64(self._h, _ptr(p), _ptr(v)))

    def state_f64(self):
        """The f64 position and velocity state (either f64 form), upload order."""
        n = self.count()
        p = np.zeros((n, 3), np.float64)
        v = np.zeros((n, 3), np.float64)
        self._check(self._lib.isph_state_f64(self._h, _ptr(p), _ptr(v)))
        return p, v

    def vjp_run(self, loss, steps, checkpoint_every=20, cfl=0.3, dts=None):
        """Reverse-mode gradient of `loss` after `steps` KDK steps from the current state. Returns a dict:
        loss, dpos (n,3), dvel (n,3), du (n,), dmass (n,), deos (16 materials x (a, b, A, B)), dts, ms
        (forward, reverse)."""
        n = self.count()
        out = {"dpos": np.zeros((n, 3)), "dvel": np.zeros((n, 3)), "du": np.zeros(n), "dmass": np.zeros(n),
               "deos": np.zeros((16, 4)), "dts": np.zeros(steps, np.float32), "ms": np.zeros(2)}
        lv = ctypes.c_double(0.0)
        d_in = None if dts is None else np.ascontiguousarray(dts, np.float32)
        self._check(self._lib.isph_vjp_run(self._h, float(cfl), None if d_in is None else _ptr(d_in), int(steps),
                                           int(checkpoint_every), ctypes.byref(loss), ctypes.byref(lv),
                                           _ptr(out["dpos"]), _ptr(out["dvel"]), _ptr(out["du"]),
                                           _ptr(out["dmass"]), _ptr(out["deos"]), _ptr(out["dts"]),
                                           _ptr(out["ms"])))
        out["loss"] = lv.value
        return out

    def diagnostics(self, with_pe=True):
        out = np.zeros(7, np.float64)
        self._check(self._lib.isph_diagnostics(self._h, int(bool(with_pe)), _ptr(out)))
        return {"ke": out[0], "ie": out[1], "pe": out[2], "L": out[3:6].copy(), "mass": out[6]}

    def set_timing(self, enable):
        self._check(self._lib.isph_set_timing(self._h, int(bool(enable))))

    def phase_times(self):
        ms = np.zeros(len(PHASES), np.float64)
        self._check(self._lib.isph_phase_times(self._h, _ptr(ms), ms.shape[0]))
        return dict(zip(PHASES, ms.tolist()))

    def kernel_stats(self):
        out = np.zeros(8, np.float64)
        self._check(self._lib.isph_kernel_stats(self._h, _ptr(out)))
        keys = ("gravity_launches", "gravity_ms", "cell_interactions", "particle_interactions",
                "quad_interactions", "density_pairs_tested", "force_pairs_tested", "gravity_flops")
        return dict(zip(keys, out.tolist()))

    def walk_stats(self):
        out = np.zeros(9, np.float64)
        self._check(self._lib.isph_walk_stats(self._h, _ptr(out)))
        keys = [f"{w}_{k}" for w in ("gravity", "density", "force") for k in ("rounds", "popped", "max_stack")]
        return dict(zip(keys, out.tolist()))

    def clamp_energy(self):
        """Energy (J) the u >= 0 clamp of the half-kicks added since the last reset_stats."""
        out = np.zeros(1, np.float64)
        self._check(self._lib.isph_clamp_energy(self._h, _ptr(out)))
        return float(out[0])

    def reset_stats(self):
        self._check(self._lib.isph_reset_stats(self._h))

    # ---- measurement (DiskView, accretion::find_clumps) ----
    def set_measure(self, clump_softening=1.0e4, link_factor=2.0):
        self._check(self._lib.isph_set_measure(self._h, float(clump_softening), float(link_factor)))

    def disk_view(self):
        """DiskView::of (gpu_sph.rs:689-785) of every instance, measured on the device: a list of Disk."""
        k = 256  # the most instances an ensemble holds
        out = (Disk * k)()
        rc = self._lib.isph_disk_view(self._h, out, k)
        self._check(rc)
        return list(out)[:rc]

    def disk_members(self, inst=0):
        """DiskView's idx (indices into the instance's particles) and each member's clump id."""
        n = self._lib.isph_disk_members(self._h, int(inst), None, None, 0)
        self._check(n)
        idx = np.zeros(n, np.uint32)
        cid = np.zeros(n, np.uint32)
        self._check(self._lib.isph_disk_members(self._h, int(inst), _ptr(idx), _ptr(cid), n))
        return idx, cid

    def disk_clumps(self, inst=0, min_members=1):
        """DiskView::clumps (gpu_sph.rs:793-809), heaviest first; min_members filters by size."""
        n = self._lib.isph_disk_clumps(self._h, int(inst), int(min_members), None, 0)
        self._check(n)
        out = (Clump * max(n, 1))()
        self._check(self._lib.isph_disk_clumps(self._h, int(inst), int(min_members), out, n))
        return list(out)[:n]

    def disk_stats_json(self, inst=0):
        """disk_stats_json (gpu_sph.rs:825-851)."""
        n = self._lib.isph_disk_stats_json(self._h, int(inst), None, 0)
        self._check(n)
        buf = ctypes.create_string_buffer(n + 1)
        self._check(self._lib.isph_disk_stats_json(self._h, int(inst), buf, n + 1))
        return buf.value.decode()

    def find_clumps(self, pos, vel, mass, rho, linking_length, g, softening, central_pos, central_mass,
                    central_radius, min_members=1):
        """accretion::find_clumps (accretion.rs:158-301) on the device: (clumps heaviest first, labels)."""
        p = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
        m = np.ascontiguousarray(mass, np.float64)
        r = np.ascontiguousarray(rho, np.float64)
        c = np.ascontiguousarray(central_pos, np.float64).reshape(3)
        n = p.shape[0]
        labels = np.zeros(n, np.uint32)
        cap = n
        out = (Clump * max(cap, 1))()
        rc = self._lib.isph_find_clumps(self._h, _ptr(p), _ptr(v), _ptr(m), _ptr(r), n, float(linking_length),
                                        float(g), float(softening), _ptr(c), float(central_mass),
                                        float(central_radius), int(min_members), out, cap, _ptr(labels))
        self._check(rc)
        return list(out)[:rc], labels

    # -- gravity (GpuGravity) --
    def gravity_accelerations(self, pos, mass, softening, mode=GRAVITY_DIRECT):
        """gpu_gravity.rs:187 (direct) / :207 (tree): f64 in, f32 on the device, f64 out."""
        p = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        m = np.ascontiguousarray(mass, np.float64)
        out = np.zeros_like(p)
        self._check(self._lib.isph_gravity_accelerations(self._h, _ptr(p), _ptr(m), p.shape[0], float(softening),
                                                         _ptr(out), int(mode)))
        return out

    # -- stage probes --
    def probe_morton(self):
        n = self.count()
        codes = np.zeros(n, np.uint32)
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        self._check(self._lib.isph_probe_morton(self._h, _ptr(codes), _ptr(lo), _ptr(hi)))
        return codes, lo, hi

    def probe_sort(self):
        n = self.count()
        codes = np.zeros(n, np.uint32)
        order = np.zeros(n, np.uint32)
        self._check(self._lib.isph_probe_sort(self._h, _ptr(codes), _ptr(order)))
        return codes, order

    def probe_tree(self, leaf_size=1):
        n = self.count()
        L = (n + leaf_size - 1) // leaf_size
        nodes = 2 * L - 1
        left = np.zeros(max(L - 1, 1), np.uint32)
        right = np.zeros(max(L - 1, 1), np.uint32)
        parent = np.zeros(nodes, np.uint32)
        com = np.zeros((nodes, 4), np.float32)
        bmin = np.zeros((nodes, 4), np.float32)
        bmax = np.zeros((nodes, 4), np.float32)
        rc = self._check(self._lib.isph_probe_tree(self._h, _ptr(left), _ptr(right), _ptr(parent), _ptr(com),
                                                   _ptr(bmin), _ptr(bmax)))
        assert rc == L, (rc, L)
        return {"L": L, "left": left[:L - 1], "right": right[:L - 1], "parent": parent, "com": com,
                "bmin": bmin, "bmax": bmax}
