"""C programs written against the libCEED C API run on /gpu/hip/mi355:
tests/c/ex-hex.c (the kind of program examples/ceed/ex1-volume.c and
ex2-surface.c are: exact volume and energy of an affine hex mesh, and the
fused gallery operator equals the staged user-QFunction operator to 1e-12)
and tests/c/bp-apply.c (CeedOperatorApply throughput of the gallery Poisson
operators, scalar and 3-component, on the perturbed trilinear mesh, timed by
one synchronised loop; A 1 = 0 checked)."""

import json
import math
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


@pytest.mark.parametrize("p,n,nc", [(4, 4, 1), (3, 3, 3), (8, 2, 1)])
def test_bp_apply(cuda, p, n, nc):
    exe = EXE.parent / "bp-apply"
    assert exe.exists(), "build the C programs with python -m libceed_amd.build"
    res = subprocess.run([str(exe), "/gpu/hip/mi355", str(p), str(n), "3", str(nc)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
    r = json.loads(res.stdout.strip().splitlines()[-1])
    assert r["dofs"] == nc * (n * p + 1) ** 3
    assert math.isfinite(r["ms_per_apply"]) and r["ms_per_apply"] > 0
    assert r["ones_max_abs"] < 1e-11  # pure Neumann diffusion: A 1 = 0 on the perturbed trilinear mesh
