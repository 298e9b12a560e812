"""The C++ integration snippets compile against the reference's own headers.

Each ```cpp block of INTEGRATION.md is a whole translation unit that uses the
reference's types (`SDFNodeShared`, `ProgramBuffer`, `OpcodeT`, `Drawable`,
the Lua helpers of lua_sdf.h / lua_vec.h) and this library's header.  The test
runs `g++ -std=c++20 -fsyntax-only` on each with the reference's include
directories (tangerine/sdf_evaluator.h:83-163 for the program words,
tangerine/lua_sdf.h and lua_vec.h for the Lua surface, tangerine/sdf_model.h
for the Sodapop drawable).  Skipped where the reference sources are absent.
"""

from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference")
INCLUDES = ["tangerine", "third_party", "third_party/glm-0.9.9.8", "third_party/fmt/include",
            "third_party/naive-surface-nets/include", "third_party/glad/include", "third_party/sdl2/include",
            "third_party/lua-5.4.4"]


def _snippets() -> list[str]:
    text = (ROOT / "INTEGRATION.md").read_text()
    return re.findall(r"^```cpp\n(.*?)^```", text, flags=re.S | re.M)


def test_snippets_present():
    blocks = _snippets()
    assert len(blocks) >= 3
    joined = "\n".join(blocks)
    # the reference's ProgramBuffer surface, with the Stop word SDFInterpreter appends
    assert "Program.Words.data()" in joined and "Program.Size()" in joined
    assert "Program.Push(OpcodeT::Stop)" in joined
    for name in ("LuaEvalGPU", "LuaGradientGPU", "LuaRayCastGPU", "tg_mesher_export_sodapop"):
        assert name in joined


@pytest.mark.skipif(not (REF / "tangerine" / "sdf_evaluator.h").exists(), reason="reference sources absent")
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
@pytest.mark.parametrize("index", range(3))
def test_snippet_compiles(index, tmp_path):
    blocks = _snippets()
    src = tmp_path / f"snippet{index}.cpp"
    src.write_text(blocks[index])
    cmd = ["g++", "-std=c++20", "-fsyntax-only", *(f"-I{REF / d}" for d in INCLUDES), f"-I{ROOT / 'include'}",
           str(src)]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert proc.returncode == 0, proc.stdout[-4000:]
