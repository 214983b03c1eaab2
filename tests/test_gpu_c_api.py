"""A C program written against the libCEED C API (tests/c/ex-hex.c, the kind
of program examples/ceed/ex1-volume.c and ex2-surface.c are) runs on
/gpu/hip/mi355: exact volume and energy of an affine hex mesh, and the fused
gallery operator equals the staged user-QFunction operator to 1e-12."""

import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
EXE = Path(__file__).resolve().parent / "c" / "bin" / "ex-hex"


@pytest.mark.parametrize("p,n", [(2, 3), (4, 4), (6, 2)])
def test_ex_hex(cuda, p, n):
    assert EXE.exists(), "build the C programs with python -m libceed_amd.build"
    res = subprocess.run([str(EXE), "/gpu/hip/mi355", str(p), str(n)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0 and " ok" in res.stdout, res.stdout + res.stderr
