"""The C++ mirror of the engine types (include/isph.hpp) compiles against isph.h with warnings as errors, links
against the built library, and turns the library's failures into isph_engine::Error with the library's
status and message. Run here without a device: construction reports the missing device, argument errors are
raised before any device work, and the empty gravity field needs no handle."""
import os
import shutil
import subprocess

import pytest

from botheadstudios_amd import isph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENODEV, EINVAL = -19, -22  # ISPH_ENODEV, ISPH_EINVAL (include/isph.h)

PROGRAM = r"""
#include "isph.hpp"
#include <cstdio>
#include <type_traits>

static_assert(std::is_same<isph_engine::SphParticle, isph_particle>::value, "record type");
static_assert(!std::is_copy_constructible<isph_engine::GpuSph>::value, "one owner per handle");
static_assert(std::is_move_constructible<isph_engine::GpuSph>::value, "movable");
// every method of the engine's GpuSph (gpu_sph.rs:1115-1505) and GpuGravity (gpu_gravity.rs:187, 207), by name
#define HAS(C, m) static_assert(std::is_member_function_pointer<decltype(&isph_engine::C::m)>::value, #m)
HAS(GpuSph, upload); HAS(GpuSph, set_dt); HAS(GpuSph, set_av); HAS(GpuSph, set_bulk); HAS(GpuSph, set_bulk_planar);
HAS(GpuSph, set_external_masses); HAS(GpuSph, set_merge_budget); HAS(GpuSph, encode_merge); HAS(GpuSph, count);
HAS(GpuSph, particle_buffer); HAS(GpuSph, encode_relax); HAS(GpuSph, encode_kdk); HAS(GpuSph, begin_readback);
HAS(GpuSph, take_readback); HAS(GpuGravity, accelerations); HAS(GpuGravity, accelerations_tree);
HAS(GravityField, accelerations);
// the extensions: measurement (DiskView), differentiation, batching
HAS(GpuSph, disk_view); HAS(GpuSph, disk_clumps); HAS(GpuSph, moonlets); HAS(GpuSph, disk_members);
HAS(GpuSph, disk_stats_json); HAS(GpuSph, find_clumps); HAS(GpuSph, vjp_run); HAS(GpuSph, loss_value); HAS(GpuSph, set_state_f64);
HAS(GpuSph, upload_ensemble); HAS(GpuSph, autotune); HAS(GpuSph, set_graph); HAS(GpuSph, step_cfl);
static_assert(std::is_same<decltype(std::declval<isph_engine::GpuSph&>().take_readback()),
                           std::optional<std::vector<isph_engine::SphParticle>>>::value, "Option<Vec<SphParticle>>");
static_assert(std::is_same<decltype(std::declval<isph_engine::GravityField&>().accelerations({}, {}, 0.0)),
                           std::optional<std::vector<isph_engine::DVec3>>>::value, "Option<Vec<DVec3>>");

int main() {
    try {
        isph_engine::GpuSph s(1024);
        std::printf("created %u\n", s.count());
    } catch (const isph_engine::Error& e) {
        std::printf("create %d %s\n", e.status(), e.what());
    }
    isph_engine::GravityField field;
    auto none = field.accelerations({}, {}, 0.01);
    std::printf("empty %d %zu\n", none.has_value() ? 1 : 0, none ? none->size() : 99);
    isph_engine::GpuGravity g;
    try {
        g.accelerations({{0.0, 0.0, 0.0}}, {1.0, 2.0}, 0.01);
    } catch (const isph_engine::Error& e) {
        std::printf("lengths %d\n", e.status());
    }
    try {
        g.accelerations_tree({{0.0, 0.0, 0.0}, {1.0, 0.0, 0.0}}, {1.0, 2.0}, 0.01);
        std::printf("tree ran\n");
    } catch (const isph_engine::Error& e) {
        std::printf("tree %d\n", e.status());
    }
    return 0;
}
"""


def test_cpp_mirror_compiles_links_and_reports_errors(tmp_path):
    if not os.path.exists(isph.LIB_PATH):
        pytest.skip("library not built")
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "mirror.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "mirror"
    libdir = os.path.dirname(isph.LIB_PATH)
    cmd = [gxx, "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
           str(src), "-o", str(exe), "-L", libdir, "-l:" + os.path.basename(isph.LIB_PATH),
           "-Wl,-rpath," + libdir, "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300,
                       env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    assert r.returncode == 0, r.stderr
    out = r.stdout.splitlines()
    assert out[0].startswith(f"create {ENODEV} isph_create: "), out
    assert out[1] == "empty 1 0", out
    assert out[2] == f"lengths {EINVAL}", out
    assert out[3] == f"tree {ENODEV}", out
