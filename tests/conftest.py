"""Shared test set-up: import paths, the gpu marker, the CPU checkers."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "mpi-les_amd", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU runs")
    lib = ROOT / "oracle" / "liblesoracle.so"
    if not lib.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "liblesoracle.so"], check=True,
                       capture_output=True)


@pytest.fixture(scope="session")
def ref_small():
    from oracle import oracle
    if not oracle.ref_available("small"):
        pytest.skip("reference build oracle/_ref/small absent (make -C oracle ref)")
    return "small"
