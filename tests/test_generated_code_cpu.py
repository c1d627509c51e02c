"""A property of the generated gfx950 code of the two-wave beam kernels
(hnsw_search_kernel_pair): every workgroup barrier is reached with no LDS
access of the wave still in flight.

The two waves of a query hand data over through LDS around s_barrier
(pair_sync in hnsw_search_impl.h).  With only a workgroup release fence, the
compiler left the second wave's hand-off write (ds_write, then s_barrier on
its loop's back edge) without an s_waitcnt lgkmcnt(0) in some instantiations
(the FP32 and BF16 ones of dimension 100 among them); under the LDS traffic of
other blocks on the CU the first wave then read stale hand-off words and
items.  pair_sync now waits explicitly.  This test disassembles the built
objects of the two-wave slices (search_inst.hip with NMS_MODE=3) and checks,
for every s_barrier, that the nearest preceding LDS-related instruction on
its straight-line path is an s_waitcnt lgkmcnt(0), with no branch target in
between."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "nmslib_amd", "_build")
BUNDLER = "/opt/rocm/llvm/bin/clang-offload-bundler"
OBJDUMP = "/opt/rocm/llvm/bin/llvm-objdump"
OBJCOPY = "/opt/rocm/llvm/bin/llvm-objcopy"


def _disassemble(obj, tmp):
    """The gfx950 code object of a built object file: its .hip_fatbin section
    is an offload bundle holding it."""
    fb = os.path.join(tmp, os.path.basename(obj) + ".fatbin")
    co = os.path.join(tmp, os.path.basename(obj) + ".co")
    subprocess.run([OBJCOPY, f"--dump-section=.hip_fatbin={fb}", obj,
                    os.path.join(tmp, "scratch.o")], check=True)
    subprocess.run([BUNDLER, "--type=o", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                    f"--input={fb}", f"--output={co}", "--unbundle"], check=True)
    return subprocess.run([OBJDUMP, "-d", "--mcpu=gfx950", co], check=True, capture_output=True,
                          text=True).stdout


def unguarded_barriers(text):
    """(function, address) of each s_barrier whose straight-line predecessors
    reach an LDS instruction, a branch target or the function start before an
    s_waitcnt lgkmcnt(0)."""
    funcs = re.split(r"\n(?=[0-9a-f]+ <[^>]+>:\n)", text)
    bad = []
    for f in funcs:
        head = re.match(r"([0-9a-f]+) <([^>]+)>:", f)
        if not head:
            continue
        name = head.group(2)
        insts = []
        targets = set()
        for line in f.split("\n")[1:]:
            m = re.match(r"\s*(\S.*?)\s*//\s*([0-9A-F]+):", line)
            if not m:
                continue
            insts.append((int(m.group(2), 16), m.group(1)))
            t = re.search(r"<[^>+]+\+0x([0-9a-f]+)>", line)
            if m.group(1).startswith(("s_branch", "s_cbranch")) and t:
                targets.add(int(t.group(1), 16) + int(head.group(1), 16))
        for i, (addr, ins) in enumerate(insts):
            if not ins.startswith("s_barrier"):
                continue
            ok = False
            for j in range(i - 1, -1, -1):
                a2, p = insts[j]
                if p.startswith("s_waitcnt") and "lgkmcnt(0)" in p:
                    ok = True
                    break
                if p.startswith("ds_") or insts[j + 1][0] in targets:
                    break
            if not ok:
                bad.append((name, hex(addr)))
    return bad


@pytest.mark.skipif(not all(os.path.exists(t) for t in (BUNDLER, OBJDUMP, OBJCOPY)),
                    reason="ROCm LLVM tools not present")
def test_two_wave_barriers_wait_for_lds(tmp_path):
    objs = [os.path.join(BUILD, f"search_st{st}_m3.o") for st in range(3)]
    if not all(os.path.exists(o) for o in objs):
        pytest.skip("build objects not present (run the build first)")
    total = 0
    for obj in objs:
        text = _disassemble(obj, str(tmp_path))
        total += text.count("s_barrier")
        bad = unguarded_barriers(text)
        assert not bad, f"{os.path.basename(obj)}: {bad[:5]}"
    assert total > 100
