"""The library warns when the process holds two HIP runtimes (this library's and a second one that a package
such as torch loads under another name after it), since streams and pointers do not cross between them."""
import os
import subprocess
import sys

import pytest

from botheadstudios_amd import isph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_runtimes_are_read_from_the_mapped_files():
    maps = "\n".join([
        "7f00-7f01 r-xp 00000000 08:01 11 /opt/rocm-7.2.0/lib/libamdhip64.so.7.2.70200",
        "7f01-7f02 r--p 00001000 08:01 11 /opt/rocm-7.2.0/lib/libamdhip64.so.7.2.70200",
        "7f02-7f03 r-xp 00000000 08:01 12 /site/torch/lib/libamdhip64.so",
        "7f03-7f04 r-xp 00000000 08:01 13 /opt/rocm-7.2.0/lib/libhsa-runtime64.so.1",
        "7f04-7f05 rw-p 00000000 00:00 0",
        "7f05-7f06 r-xp 00000000 08:01 14 /tmp/not libamdhip64.so",
    ])
    assert isph.hip_runtimes(maps) == sorted([os.path.realpath("/opt/rocm-7.2.0/lib/libamdhip64.so.7.2.70200"),
                                              os.path.realpath("/site/torch/lib/libamdhip64.so")])
    assert isph.hip_runtimes("") == []


_PROBE = r"""
import sys, warnings
sys.path.insert(0, sys.argv[1])
from botheadstudios_amd import isph
isph.load_library()
import torch  # after the library: torch adds its bundled runtime
n = len(isph.hip_runtimes())
with warnings.catch_warnings(record=True) as w:
    warnings.simplefilter("always")
    try:
        isph.GpuSph(16)
    except isph.IsphError:
        pass  # no device here
print(n, sum("two HIP runtimes" in str(x.message) for x in w))
"""


def test_a_second_runtime_loaded_after_the_library_is_reported():
    pytest.importorskip("torch")
    if not os.path.exists(isph.LIB_PATH):
        pytest.skip("library not built")
    r = subprocess.run([sys.executable, "-c", _PROBE, ROOT], capture_output=True, text=True, timeout=600,
                       env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    assert r.returncode == 0, r.stderr
    n, warned = map(int, r.stdout.split()[-2:])
    if n < 2:
        pytest.skip("torch here shares the library's runtime")
    assert warned == 1
