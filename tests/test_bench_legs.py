"""The benchmark's single JSON line survives a failing optional leg: the headline is printed, the leg that
raised carries its error, and the other legs still run. GPU work is replaced by stand-ins here."""
import json
import sys
import types

import numpy as np
import pytest

import bench


def _stats():
    return {"gravity_launches": 2.0, "gravity_flops": 4.0e12, "gravity_ms": 10.0, "cell_interactions": 1.0e6,
            "particle_interactions": 2.0e6, "quad_interactions": 0.0}


def _fake_run_gpu(sc, args, precision, steps, warmup, reuse, measure=False):
    out = {"elapsed": 1.0, "stats": _stats(), "phases": {"gravity": 0.5}, "sim_time": 1.0,
           "dE_rel": 1e-6, "dL_rel": 1e-7, "clamp_energy_rel": 0.0}
    if measure:
        out["measure"] = {"workload": "stand-in"}
        out["measure_rec"] = None
    return out


def _boom(*a):
    raise RuntimeError("leg failed")


def test_leg_returns_the_error_of_a_leg_that_raises(capsys):
    assert bench._leg(lambda x: {"v": x}, 3) == {"v": 3}
    r = bench._leg(_boom)
    assert r == {"error": "RuntimeError: leg failed"}
    assert "RuntimeError" in capsys.readouterr().err


def test_the_headline_line_is_printed_when_optional_legs_fail(monkeypatch, capsys):
    fake_scene = types.SimpleNamespace(particles=[0] * 1024, n_target=900, n_impactor=124)
    monkeypatch.setitem(sys.modules, "botheadstudios_amd.scene",
                        types.SimpleNamespace(two_body=lambda n: fake_scene, two_body_1m=lambda: fake_scene))
    import botheadstudios_amd
    monkeypatch.setattr(botheadstudios_amd, "scene", sys.modules["botheadstudios_amd.scene"], raising=False)
    monkeypatch.setattr(bench, "run_gpu", _fake_run_gpu)
    monkeypatch.setattr(bench, "run_debris", _boom)
    monkeypatch.setattr(bench, "run_adjoint", lambda args: {"ratio": 2.0})
    monkeypatch.setattr(bench, "run_adjoint_outcome", _boom)
    monkeypatch.setattr(bench, "run_batch", lambda args: {"value": 1.0})
    monkeypatch.setattr(bench, "run_4m", _boom)
    monkeypatch.setattr(bench, "run_precision_study", lambda args, sc: {"rows": []})
    monkeypatch.setattr(bench, "cpu_baseline", _boom)
    monkeypatch.setattr(bench, "measure_cpu", lambda rec: 1.0)
    monkeypatch.setattr(sys, "argv", ["bench.py", "--n", "1024", "--steps", "3", "--warmup", "1"])
    bench.main()
    out = capsys.readouterr().out.strip().splitlines()
    assert len(out) == 1
    line = json.loads(out[0])
    assert line["value"] == pytest.approx(1024 * 3 / 1.0)
    assert line["unit"] == "particle-steps/s" and line["n_gpus"] == 1
    assert line["measure"]["debris_field"]["error"] == "RuntimeError: leg failed"
    assert line["adjoint"]["ratio"] == 2.0 and "error" in line["adjoint"]["outcome"]
    assert line["batch"] == {"value": 1.0}
    assert "error" in line["precision_4m"] and line["precision_study"] == {"rows": []}
    assert "error" in line["cpu_baseline"] and "frac_reference_walk" not in line["roofline"]
    assert line["mixed"]["value"] == pytest.approx(3072.0) and line["reuse"]["ratio_to_value"] == 1.0


def test_more_than_one_gpu_is_refused_with_a_message():
    import subprocess
    r = subprocess.run([sys.executable, bench.__file__, "--gpus", "2", "--steps", "1"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and r.stdout == ""
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr


class _FakeSph:
    """Stands in for isph.GpuSph: every call succeeds with fixed figures."""

    def __init__(self, n, device=0):
        self.n = n

    def __getattr__(self, name):
        return lambda *a, **k: None

    def diagnostics(self, with_pe=False):
        return {"ke": 1.0, "ie": 2.0, "pe": -4.0, "L": np.array([0.0, 0.0, 1.0])}

    def sim_time(self):
        return 0.0, 1.0

    def phase_times(self):
        return {"gravity": 1.0}

    def kernel_stats(self):
        return _stats()

    def walk_stats(self):
        return {}

    def clamp_energy(self):
        return 0.0


def test_a_failing_disk_measurement_keeps_the_timed_run(monkeypatch):
    import argparse
    from botheadstudios_amd import isph
    monkeypatch.setattr(isph, "GpuSph", _FakeSph)
    monkeypatch.setattr(bench, "time_disk_view", _boom)
    sc = types.SimpleNamespace(particles=[0] * 64, eos=None, eos64=None, softening=1.0)
    args = argparse.Namespace(gravity="tree", theta=0.5, leaf=1, quadrupole=False, group_extent=4.0, ppl=1,
                              expand=4, cfl=0.3, no_energy=False)
    out = bench.run_gpu(sc, args, "fp32", 2, 1, False, measure=True)
    assert out["measure"] == {"error": "RuntimeError: leg failed"} and out["measure_rec"] is None
    assert out["elapsed"] > 0 and out["dE_rel"] == 0.0
