"""C-ABI library checks that need no GPU."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "hyphy_mi355x.h")).read()
    return sorted(set(re.findall(r"^(?:int|void|double|size_t|const char \*)\s*\*?\s*(hy_\w+)\s*\(", text, re.M)))


def test_header_declares_expected_entry_points():
    from hyphy3 import engine
    assert set(declared_symbols()) == set(engine.EXPORTS)


def test_library_exports_every_declared_symbol():
    from hyphy3 import engine
    L = engine.load_library()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_create_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from hyphy3.engine import Engine, HyError
    with pytest.raises(HyError):
        Engine(0)
