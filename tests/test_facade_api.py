"""The C++ facade compiles against code written the way the reference's
BUSTED drives the engine (tests/cpp/busted_tile_loop.cpp: Matrix::Zero, +=,
scalar products, noalias, MG94Matrix::transition_matrix and spectral
fields, workspace init() and ws.root_log_L(b)); host compiler only, no GPU."""
import os
import shutil
import subprocess

import pytest

from helpers import ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_busted_style_caller_compiles(tmp_path):
    src = os.path.join(ROOT, "tests", "cpp", "busted_tile_loop.cpp")
    obj = tmp_path / "busted_tile_loop.o"
    r = subprocess.run(["g++", "-std=c++20", "-O0", "-c", src, "-o", str(obj), "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "hyphy-3_amd", "cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
