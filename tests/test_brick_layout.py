"""CPU check of the brick E^T layout (mi355::brick_layout, the host half of
CEED_MI355_PLAN_BRICK plans): tests/c/brick-layout-check.cpp builds structured
hex meshes (odd element counts, shuffled elements, nodes renumbered into an
L-vector with unused entries), runs the layout and checks the packed order, the
node lists against the offsets, the completeness bits, the gather map, and
E^T of random element values through the layout against a direct sum."""

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "libceed_amd" / "lib"
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not (LIB / "libceed-mi355.so").exists() or shutil.which(HIPCC) is None:
        pytest.skip("library or hipcc not present")
    exe = tmp_path_factory.mktemp("brick") / "brick-layout-check"
    res = subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O1", "-std=c++17", f"-I{ROOT / 'include'}",
                          f"-I{ROOT / 'libceed_amd' / 'csrc'}", str(ROOT / "tests" / "c" / "brick-layout-check.cpp"), f"-L{LIB}", "-lceed-mi355",
                          f"-Wl,-rpath,{LIB}", "-o", str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


# n elements per direction, p, brick shape, shuffled elements, renumbered nodes
@pytest.mark.parametrize("n,p,shape,shuffle,renumber", [(4, 4, (2, 2, 2), 0, 0), (5, 4, (2, 2, 2), 0, 0), (5, 4, (2, 2, 2), 1, 1),
                                                        (6, 3, (2, 2, 1), 1, 0), (3, 6, (2, 1, 1), 0, 1), (7, 2, (2, 2, 2), 1, 1),
                                                        (1, 4, (2, 2, 2), 0, 0), (4, 1, (2, 2, 2), 1, 1), (3, 5, (1, 2, 2), 1, 1),
                                                        (4, 8, (1, 1, 1), 0, 0)])
def test_brick_layout(checker, n, p, shape, shuffle, renumber):
    res = subprocess.run([str(checker), str(n), str(p), *map(str, shape), str(shuffle), str(renumber)], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0 and res.stdout.startswith("ok"), res.stdout + res.stderr
    fill = float(res.stdout.split("fill=")[1].split()[0])
    if all(n % s == 0 for s in shape):  # a structured mesh tiles exactly, whatever the numbering
        assert fill == 1.0
