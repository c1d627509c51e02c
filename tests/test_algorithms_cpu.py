"""CPU checks of the order-exact building blocks the device search uses:

* nmslib_amd/csrc/sort_emul.h against libstdc++'s std::sort and
  std::priority_queue on tie-heavy inputs (the reference sorts its
  per-expansion buffer with std::sort, hnsw_distfunc_opt.cc:251, and keeps
  SearchOld's queues in priority_queue, :99-102);
* the register-resident replay of push_or_replace_non_empty_exp
  (sort_arr_bi.h:159-199) used when new keys tie existing ones, against the
  plain sequential insertion, including truncation at capacity."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")


@pytest.mark.parametrize("src", ["sort_emul_check.cc", "insert_replay_check.cc"])
def test_cpp_check(tmp_path, src):
    exe = str(tmp_path / "check")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, src), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert " 0 mismatch" in out.stdout or ": 0 / " in out.stdout or "0 mismatches" in out.stdout
