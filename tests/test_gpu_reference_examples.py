"""The reference's own examples on /gpu/hip/mi355: examples/ceed/ex1-volume.c and
ex2-surface.c, unchanged, compiled against include/ and linked to
libceed_amd/lib/libceed-mi355.so by `make -C oracle ref` (oracle/_ref/examples,
built where the reference sources are present and carried with the tree).
Each case's computed volume or surface area (the examples print 14 digits) must
equal what the same example computes on the reference's /cpu/self/ref/serial
backend (tests/golden/reference_examples.json, written by
tests/golden/make_example_golden.py) to 1e-12 relative.  -g takes the gallery
QFunctions, whose sources the library holds (the examples' own QFunction
headers stay with the reference)."""

import json
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
EXAMPLES = ROOT / "oracle" / "_ref" / "examples"
GOLDEN = json.loads((ROOT / "tests" / "golden" / "reference_examples.json").read_text())["cases"]
PATTERN = {"ex1-volume": r"Computed mesh volume\s*:\s*(\S+)", "ex2-surface": r"Computed mesh surface area\s*:\s*(\S+)"}

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", GOLDEN, ids=[f"{c['example']} {' '.join(c['args'])}" for c in GOLDEN])
def test_reference_example_on_gpu(cuda, case):
    exe = EXAMPLES / case["example"]
    if not exe.exists():
        pytest.skip("reference examples not built (oracle/_ref/examples)")
    cmd = [str(exe), "-ceed", "/gpu/hip/mi355", "-g", *case["args"]]
    print("running", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    assert res.returncode == 0, res.stdout + res.stderr
    got = float(re.search(PATTERN[case["example"]], res.stdout).group(1))
    assert abs(got - case["value"]) <= 1e-12 * abs(case["value"]), (got, case["value"])


@pytest.mark.parametrize("example,args", [("ex1-volume", ["-d", "3"]), ("ex2-surface", ["-d", "3"])])
def test_reference_example_test_mode_on_gpu(cuda, example, args):
    """The examples' own test mode (-t) prints a line only when the result misses the exact value."""
    exe = EXAMPLES / example
    if not exe.exists():
        pytest.skip("reference examples not built (oracle/_ref/examples)")
    res = subprocess.run([str(exe), "-ceed", "/gpu/hip/mi355", "-t", "-g", *args], capture_output=True, text=True, timeout=240)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "error" not in res.stdout.lower(), res.stdout
