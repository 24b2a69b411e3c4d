"""The engine driven from C through the C-ABI alone: tests/c_abi/lake_at_rest.c (built next to the
library by __graft_entry__.build()) generates a mesh, creates an FP64 plain, an FP64 fast and an
FP32 context, steps a lake at rest 20 steps and checks it stays at rest (ze change below 1e-12
in FP64, 1e-5 in FP32; momentum below 1e-9 / 1e-3)."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "c_abi" / "lake_at_rest"


def test_c_program_lake_at_rest():
    if not EXE.exists():
        pytest.fail("tests/c_abi/lake_at_rest is not built: run __graft_entry__.build()")
    print("[gpu-c-abi] running", EXE, flush=True)
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    for name in ("fp64 plain", "fp64 fast", "fp32 bulk"):
        assert name + ":" in r.stdout
