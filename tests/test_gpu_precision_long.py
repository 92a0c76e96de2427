"""Precision form at scale: the device step against the f64 CPU step (the oracle's HydroBody restatement)
on the 65,536-particle two-body impact over 500 KDK steps at the CPU path's Courant dt sequence.

Long (minutes of CPU time); runs only with ISPH_LONG=1. It states, per precision form, the per-particle
error of density, acceleration and du/dt at the start state (SPH part, exact f64 reference; gravity part
against the f64 direct sum) and the relative difference of the energy drift and angular-momentum change
after 500 steps (tree gravity on the device, octree on the CPU, both theta 0.5).
"""
import os

import numpy as np
import pytest

from botheadstudios_amd import isph, materials, scene
from oracle import oracle

from helpers import hydro_from, rel_err_floor

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("ISPH_LONG") != "1", reason="long study: set ISPH_LONG=1")]

N, STEPS = 65536, 500
EOS = materials.eos_table()
EOS64 = materials.eos_f64()


def _maxrms(x, ref):
    e = rel_err_floor(x, ref)
    return float(e.max()), float(np.sqrt(np.mean(e ** 2)))


@pytest.fixture(scope="module")
def setup():
    sc = scene.two_body(N, seed=0xC3)
    return sc


@pytest.mark.timeout(3000)
def test_precision_form_errors_at_65k(setup):
    sc = setup
    rec = sc.particles
    soft = sc.softening
    # per-quantity errors at the start state (SPH only, then gravity alone)
    hb = hydro_from(rec, EOS64, soft, gravity="none")
    hb.compute_density()
    acc_ref, dudt_ref = hb.forces()
    _, _, _, rho_ref = hb.state()
    pos, vel, mass, u, h, mat = scene.f64_state(rec)
    g_ref = oracle.direct_gravity(pos, mass, soft)
    out = {}
    for name, prec in (("fp32", isph.PRECISION_FP32), ("mixed", isph.PRECISION_MIXED)):
        with isph.GpuSph(N) as s:
            s.upload(rec, EOS, soft)
            s.set_eos_f64(EOS64)
            s.set_precision(prec)
            s.set_gravity(isph.GRAVITY_NONE, 0.5, 1)
            acc, dudt, rho = s.force_eval()
            g = s.gravity_accelerations(pos, mass, soft, isph.GRAVITY_DIRECT)
        out[name] = {"rho": _maxrms(rho, rho_ref), "acc": _maxrms(acc, acc_ref), "dudt": _maxrms(dudt, dudt_ref),
                     "gravity": _maxrms(g, g_ref)}
    # 500 steps: CPU f64 step with octree gravity and its Courant dt sequence, then the device with those dts
    hb = hydro_from(rec, EOS64, soft, gravity="octree")
    hb.compute_density()
    ke, ie, pe = oracle.total_energy(pos, vel, mass, u, soft)
    e0 = ke + ie + pe
    l0 = oracle.angular_momentum(pos, vel, mass)
    dts = []
    for k in range(STEPS):
        dt = float(np.float32(hb.courant_dt(0.3)))
        dts.append(dt)
        hb.step(dt)
        if k % 25 == 0:
            print(f"cpu step {k}", flush=True)
    p1, v1, u1, _ = hb.state()
    ke, ie, pe = oracle.total_energy(p1, v1, mass, u1, soft)
    de_cpu = (ke + ie + pe - e0) / abs(e0)
    dl_cpu = (oracle.angular_momentum(p1, v1, mass) - l0) / np.linalg.norm(l0)
    for name, prec in (("fp32", isph.PRECISION_FP32), ("mixed", isph.PRECISION_MIXED)):
        with isph.GpuSph(N) as s:
            s.upload(rec, EOS, soft)
            s.set_eos_f64(EOS64)
            s.set_precision(prec)
            d0 = s.diagnostics(with_pe=True)
            for dt in dts:
                s.set_dt(dt, 1.0)
                s.encode_kdk(1)
            d1 = s.diagnostics(with_pe=True)
        e0g = d0["ke"] + d0["ie"] + d0["pe"]
        de = (d1["ke"] + d1["ie"] + d1["pe"] - e0g) / abs(e0g)
        dl = (d1["L"] - d0["L"]) / np.linalg.norm(d0["L"])
        out[name].update({"dE_gpu": de, "dE_cpu": de_cpu, "dE_diff": abs(de - de_cpu),
                          "dL_diff": float(np.linalg.norm(dl - dl_cpu))})
    for name, r in out.items():
        print(name, {k: (f"{v[0]:.2e}/{v[1]:.2e}" if isinstance(v, tuple) else f"{v:.3e}") for k, v in r.items()})
    assert out["mixed"]["rho"][0] < 1e-5 and out["mixed"]["dudt"][0] < 1e-5
    assert out["mixed"]["dE_diff"] < 1e-3 and out["fp32"]["dE_diff"] < 1e-3
