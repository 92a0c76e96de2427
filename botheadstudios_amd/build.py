"""Builds the HIP library botheadstudios_amd/libisph.so for gfx950 (hipcc, in-tree).

Every HIP translation unit is compiled with packed-FP32 instruction selection disabled
(`-Xclang -target-feature -Xclang -packed-fp32-ops`); the host pass warns that it does not know the
feature, which is harmless.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libisph.so")
SOURCES = ["sort.hip", "tree.hip", "gravity.hip", "sph.hip", "adjoint.hip", "ensemble.hip", "merge.hip", "measure.hip", "noalloc.hip", "isph.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Werror=return-type", "-Werror=uninitialized",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src, defines=(), tag=""):
    obj = os.path.join(CSRC, "_build" + tag, src + ".o")
    os.makedirs(os.path.dirname(obj), exist_ok=True)
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + FLAGS + [f"-D{d}" for d in defines] + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False, out=OUT, defines=()):
    """Compile and link; `defines` are extra -D macros (tuning variants), `out` the library path."""
    tag = "_" + "_".join(d.replace("=", "") for d in defines) if defines else ""
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(lambda s: _compile(s, defines, tag), SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for (o, err), src in zip(results, SOURCES):
            lines = [l for l in err.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if lines:
                print(f"[{src}]\n" + "\n".join(lines))
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
