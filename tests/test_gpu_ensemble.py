"""Ensembles (the batching form): independent systems advanced by the same launches.

Every instance has its own Morton box, its own subtree of the ensemble tree, its own walks and its own
Courant dt, so an ensemble step must reproduce, bit for bit, each instance stepped alone.
"""
import numpy as np
import pytest

from botheadstudios_amd import isph, materials, scene

pytestmark = pytest.mark.gpu

EOS = materials.eos_table()
EOS64 = materials.eos_f64()


def _scenes():
    return [scene.two_body(1536, seed=0xE1), scene.two_body(3072, seed=0xE2), scene.two_body(800, seed=0xE3)]


@pytest.mark.parametrize("precision", [isph.PRECISION_FP32, isph.PRECISION_MIXED])
def test_ensemble_reproduces_standalone_runs_bitwise(precision):
    scs = _scenes()
    soft = scs[0].softening
    alone = []
    for sc in scs:
        with isph.GpuSph(len(sc.particles)) as s:
            s.upload(sc.particles, EOS, soft)
            s.set_eos_f64(EOS64)
            s.set_precision(precision)
            s.step_cfl(0.3, 5)
            alone.append((s.readback_sync(), s.sim_time()))
    total = sum(len(sc.particles) for sc in scs)
    with isph.GpuSph(total) as s:
        off = s.upload_ensemble([sc.particles for sc in scs], EOS, soft)
        s.set_eos_f64(EOS64)
        s.set_precision(precision)
        s.step_cfl(0.3, 5)
        out = s.readback_sync()
        t, dt = s.ensemble_times()
    for k, (rec, (t_alone, dt_alone)) in enumerate(alone):
        part = out[off[k]:off[k + 1]]
        for f in ("pos", "vel", "u", "rho"):
            assert np.array_equal(part[f], rec[f]), (k, f)
        assert t[k] == t_alone and dt[k] == dt_alone


def test_ensemble_rejects_what_would_couple_instances():
    scs = _scenes()
    with isph.GpuSph(10000) as s:
        s.upload_ensemble([sc.particles for sc in scs], EOS, scs[0].softening)
        with pytest.raises(isph.IsphError):
            s.set_gravity(isph.GRAVITY_DIRECT, 0.5, 1)
        with pytest.raises(isph.IsphError):
            s.set_gravity(isph.GRAVITY_TREE, 0.5, 4)
        with pytest.raises(isph.IsphError):
            s.diagnostics(with_pe=True)
