"""The reference-side plugin of the repository's integration notes (an
Index<float> method that forwards to the C ABI,
similarity_search/include/index.h:50-109) compiles
against the reference's own headers and include/nmslib_amd.h.  Runs where the
reference sources are present; the code blocks are taken from the document
itself, so what it shows is what is compiled."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INC = "/root/reference/similarity_search/include"


def _blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", text, re.S)


def test_integration_document_holds_the_plugin():
    blocks = _blocks()
    assert any("class HnswGpu : public Index<float>" in b for b in blocks)
    assert any("CreateHnswGpu" in b for b in blocks)


@pytest.mark.skipif(not os.path.isdir(REF_INC), reason="reference sources not present")
def test_plugin_compiles_against_the_reference_headers(tmp_path):
    blocks = _blocks()
    plugin = next(b for b in blocks if "class HnswGpu" in b)
    factory = next(b for b in blocks if "CreateHnswGpu" in b)
    (tmp_path / "method").mkdir()
    (tmp_path / "method" / "hnsw_gpu.h").write_text("#pragma once\n" + plugin)
    src = tmp_path / "use.cc"
    src.write_text(factory + '''
namespace similarity {
// every virtual of Index<float> is overridden: the class is instantiable
void use(const ObjectVector& data, Space<float>& sp) {
  std::unique_ptr<Index<float>> idx(CreateHnswGpu(false, "l2", sp, data));
  (void)idx->StrDesc();
}
}  // namespace similarity
''')
    cmd = ["g++", "-std=c++14", "-fsyntax-only", "-Wall", "-Wno-unused-function",
           "-DHAVE_CXX0X", f"-I{tmp_path}", f"-I{REF_INC}", f"-I{os.path.join(ROOT, 'include')}",
           str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
