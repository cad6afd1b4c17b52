import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "hyphy-3_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def _ensure_built():
    """Build the FP64 oracle and the engine library in-tree when a fresh
    checkout lacks them (built artefacts are not committed)."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "libhyoracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, stdout=subprocess.DEVNULL)
    pkg = os.path.join(ROOT, "hyphy-3_amd")
    if not os.path.exists(os.path.join(pkg, "libhyphy_mi355x.so")):
        subprocess.run(["make", "-j8", "-C", pkg], check=True, stdout=subprocess.DEVNULL)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")
    _ensure_built()
