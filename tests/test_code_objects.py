"""Resource use of every gfx950 kernel in the built library, read from the AMDGPU metadata of its code objects:
no kernel uses scratch (no private segment, no VGPR spills, no dynamic stack), and the static LDS of
each fits one CU. A kernel that spills gets a per-lane private segment that the runtime must back with
scratch memory at dispatch; the walks are sized to run without it."""
import os
import re
import struct
import subprocess

import pytest

from botheadstudios_amd import isph

READELF = "/opt/rocm/llvm/bin/llvm-readelf"
LDS_PER_CU = 160 * 1024


def _section(path, name):
    """Bytes of an ELF64 section by name (little-endian)."""
    with open(path, "rb") as f:
        data = f.read()
    shoff, = struct.unpack_from("<Q", data, 0x28)
    shentsize, shnum, shstrndx = struct.unpack_from("<HHH", data, 0x3A)
    sh = [struct.unpack_from("<IIQQQQIIQQ", data, shoff + i * shentsize) for i in range(shnum)]
    stroff = sh[shstrndx][4]
    for s in sh:
        nm = data[stroff + s[0]:data.index(b"\0", stroff + s[0])].decode()
        if nm == name:
            return data[s[4]:s[4] + s[5]]
    return None


def _code_objects(fatbin, arch="gfx950"):
    """The device code objects of every clang offload bundle in a .hip_fatbin section."""
    out = []
    for m in re.finditer(b"__CLANG_OFFLOAD_BUNDLE__", fatbin):
        b = m.start()
        p = b + 24
        n, = struct.unpack_from("<Q", fatbin, p)
        p += 8
        for _ in range(n):
            off, size, tl = struct.unpack_from("<QQQ", fatbin, p)
            p += 24
            triple = fatbin[p:p + tl].decode()
            p += tl
            if arch in triple:
                out.append(fatbin[b + off:b + off + size])
    return out


def _kernels(tmp_path):
    if not os.path.exists(isph.LIB_PATH):
        pytest.skip("library not built")
    if not os.path.exists(READELF):
        pytest.skip("llvm-readelf not available")
    fat = _section(isph.LIB_PATH, ".hip_fatbin")
    assert fat, "no device code in the library"
    kernels = []
    for i, co in enumerate(_code_objects(fat)):
        path = tmp_path / f"co{i}.o"
        path.write_bytes(co)
        text = subprocess.run([READELF, "--notes", str(path)], capture_output=True, text=True, check=True).stdout
        for block in text.split("\n  - ")[1:]:
            fields = dict(re.findall(r"^\s*\.(\w+):\s+(\S+)\s*$", block, re.M))
            if "symbol" in fields:
                kernels.append(fields)
    return kernels


def test_no_kernel_uses_scratch_and_lds_fits(tmp_path):
    ks = _kernels(tmp_path)
    names = [k["name"] for k in ks]
    assert len(ks) > 50
    for needed in ("k_gravity_tree", "k_density", "k_force", "k_upsweep_lvk", "k_radix_scatter"):
        assert any(needed in nm for nm in names), needed
    # SGPR spills are allowed: without a private segment they go to lanes of a VGPR, not to memory
    bad = [(k["name"], k["private_segment_fixed_size"], k["vgpr_spill_count"], k["uses_dynamic_stack"])
           for k in ks
           if int(k["private_segment_fixed_size"]) or int(k["vgpr_spill_count"]) or k["uses_dynamic_stack"] != "false"]
    assert not bad, bad
    big = [(k["name"], k["group_segment_fixed_size"]) for k in ks if int(k["group_segment_fixed_size"]) > LDS_PER_CU]
    assert not big, big


def test_walk_kernels_fit_their_occupancy(tmp_path):
    # wave64 on gfx950: 512 VGPRs per lane per SIMD, so 8 resident waves need <= 64 VGPRs and 6 need <= 80
    # (granule 8); the plain gravity and density walks are declared for 8
    ks = {k["name"]: k for k in _kernels(tmp_path)}
    grav = [k for nm, k in ks.items() if "k_gravity_tree" in nm and "ILi1ELb0ELb0ELb0E" in nm]
    dens = [k for nm, k in ks.items() if "k_density" in nm and "ILb0ELb0E" in nm]
    assert grav and dens
    for k in grav + dens:
        assert int(k["vgpr_count"]) <= 64, (k["name"], k["vgpr_count"])
