"""Writes tests/golden/reference_examples.json: the volume and surface area the
reference's examples/ceed/ex1-volume.c and ex2-surface.c compute on the
reference's /cpu/self/ref/serial backend, for the cases the GPU test runs on
/gpu/hip/mi355 (tests/test_gpu_reference_examples.py).

Needs the reference sources (/root/reference) and its checker build
oracle/_ref/libceed.so (`make -C oracle ref`).  The examples are compiled
against the reference's own headers and library into a temporary folder; only
their printed results are kept.
"""

import json
import re
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
REF = Path("/root/reference")
LIB = ROOT / "oracle" / "_ref"
OUT = Path(__file__).resolve().parent / "reference_examples.json"

# (example, arguments after -ceed <resource> -g); -g selects the gallery QFunctions
CASES = [("ex1-volume", ["-d", "3"]), ("ex1-volume", ["-d", "3", "-s", "200000"]), ("ex1-volume", ["-d", "3", "-m", "2", "-p", "6", "-s", "50000"]),
         ("ex1-volume", ["-d", "3", "-m", "3", "-p", "2", "-q", "5", "-s", "30000"]), ("ex2-surface", ["-d", "3"]),
         ("ex2-surface", ["-d", "3", "-s", "200000"]), ("ex2-surface", ["-d", "3", "-m", "3", "-p", "5", "-s", "50000"])]
PATTERN = {"ex1-volume": r"Computed mesh volume\s*:\s*(\S+)", "ex2-surface": r"Computed mesh surface area\s*:\s*(\S+)"}
EXACT = {"ex1-volume": r"Exact mesh volume\s*:\s*(\S+)", "ex2-surface": r"Exact mesh surface area\s*:\s*(\S+)"}


def main() -> int:
    if not (REF / "examples" / "ceed").is_dir() or not (LIB / "libceed.so").exists():
        print("needs /root/reference and oracle/_ref/libceed.so (make -C oracle ref)")
        return 1
    out = []
    with tempfile.TemporaryDirectory() as d:
        for ex in sorted({c[0] for c in CASES}):
            subprocess.run(["gcc", "-O2", "-std=c11", f"-I{REF / 'include'}", str(REF / "examples" / "ceed" / f"{ex}.c"), f"-L{LIB}", "-lceed",
                            f"-Wl,-rpath,{LIB}", "-lm", "-o", str(Path(d) / ex)], check=True)
        for ex, args in CASES:
            res = subprocess.run([str(Path(d) / ex), "-ceed", "/cpu/self/ref/serial", "-g", *args], capture_output=True, text=True, check=True)
            value = float(re.search(PATTERN[ex], res.stdout).group(1))
            exact = float(re.search(EXACT[ex], res.stdout).group(1))
            out.append({"example": ex, "args": args, "value": value, "exact": exact})
            print(ex, " ".join(args), value, exact)
    OUT.write_text(json.dumps({"resource": "/cpu/self/ref/serial", "cases": out}, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
