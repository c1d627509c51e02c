"""Builds the HIP library nmslib_amd/libnmslib_amd.so for gfx950 with hipcc.

All device code is compiled with -ffp-contract=off (FP32 distances follow the
reference's separate multiply and add roundings; fused paths call fmaf
explicitly) and with packed-FP32 instructions disabled."""
from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libnmslib_amd.so")
# (source, object name, extra defines): search_inst.hip is compiled once per
# (vector storage, kernel mode) slice so the slices build in parallel
SOURCES = ([("search_inst.hip", f"search_st{st}_m{m}", [f"-DNMS_ST={st}", f"-DNMS_MODE={m}"])
            for st in range(3) for m in range(6)]
           + [(f, f, []) for f in ("launch.hip", "lowp_convert.hip", "bruteforce.hip",
                                   "bruteforce_exact.hip", "nmslib_amd.cpp", "hnsw_build.cpp")])
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
    jobs = []
    for src, name, defines in SOURCES:
        obj = os.path.join(builddir, name + ".o")
        objs.append(obj)
        if incremental and not _stale(src, obj):
            continue
        cmd = [HIPCC, *FLAGS, *defines, "-c", os.path.join(CSRC, src), "-o", obj]
        if src in HOST_SOURCES:
            cmd = [CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        elif src.endswith(".cpp"):
            cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        jobs.append((name, cmd))
    # at most one compiler per core (MAX_JOBS when set), largest slices first
    limit = max(1, int(os.environ.get("MAX_JOBS", 0)) or os.cpu_count() or 4)
    failed, running = [], []

    def reap(block):
        for name, p in list(running):
            if block or p.poll() is not None:
                out, _ = p.communicate()
                running.remove((name, p))
                text = out.decode(errors="replace")
                if p.returncode != 0:
                    failed.append(name)
                    sys.stderr.write(text)
                elif verbose and text.strip():
                    sys.stderr.write(text)
                if block:
                    return

    for name, cmd in jobs:
        while len(running) >= limit:
            reap(False)
            if len(running) >= limit:
                time.sleep(0.2)
        if verbose:
            print(" ".join(cmd), flush=True)
        running.append((name, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                               stderr=subprocess.STDOUT)))
    while running:
        reap(True)
    if failed:
        raise RuntimeError(f"hipcc failed on {failed}")
    tmp = OUT + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs],
                   check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
