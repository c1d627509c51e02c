"""Builds the HIP library nmslib_amd/libnmslib_amd.so for gfx950 with hipcc.

All device code is compiled with -ffp-contract=off (FP32 distances follow the
reference's separate multiply and add roundings; fused paths call fmaf
explicitly) and with packed-FP32 instructions disabled."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libnmslib_amd.so")
SOURCES = ["search_f32.hip", "search_bf16.hip", "search_fp8.hip", "launch.hip",
           "lowp_convert.hip", "bruteforce.hip", "bruteforce_exact.hip", "nmslib_amd.cpp", "hnsw_build.cpp"]
# host-only sources: the index builder (x86 SSE, no device code)
HOST_SOURCES = {"hnsw_build.cpp"}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", "-Wall"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function"]


def _stale(src: str, obj: str) -> bool:
    """The object is older than its source or any header of csrc/ and include/."""
    if not os.path.exists(obj):
        return True
    deps = [os.path.join(CSRC, src)]
    deps += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    inc = os.path.join(os.path.dirname(HERE), "include")
    deps += [os.path.join(inc, f) for f in os.listdir(inc)]
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compiles every source (incremental=True: only those whose object is
    out of date) and links the library."""
    objs = []
    builddir = os.path.join(HERE, "_build")
    os.makedirs(builddir, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(builddir, src + ".o")
        if incremental and not _stale(src, obj):
            objs.append(obj)
            continue
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if src in HOST_SOURCES:
            cmd = [CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        elif src.endswith(".cpp"):
            cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(text)
        elif verbose and text.strip():
            sys.stderr.write(text)
    if failed:
        raise RuntimeError(f"hipcc failed on {failed}")
    tmp = OUT + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs],
                   check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
