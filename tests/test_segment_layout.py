"""CPU check of the segment E^T layout (mi355::segment_layout, the host half of
CEED_MI355_PLAN_SEGMENT plans): tests/c/segment-layout-check.cpp builds
structured hex meshes (shuffled elements, nodes renumbered into an L-vector with
unused entries), runs the layout and checks the packed order, the chained faces,
that segments of one colour share no node, the balanced cut of every x-row, and
E^T of random element values the segment kernel's way against a direct sum."""

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
    exe = tmp_path_factory.mktemp("segment") / "segment-layout-check"
    res = subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O1", "-std=c++17", f"-I{ROOT / 'include'}",
                          f"-I{ROOT / 'libceed_amd' / 'csrc'}", str(ROOT / "tests" / "c" / "segment-layout-check.cpp"), f"-L{LIB}",
                          "-lceed-mi355", f"-Wl,-rpath,{LIB}", "-o", str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


# n elements per direction, p, longest segment, shuffled elements, renumbered nodes
@pytest.mark.parametrize("n,p,max_len,shuffle,renumber", [(6, 4, 7, 0, 0), (10, 4, 7, 1, 1), (9, 2, 4, 1, 0), (5, 3, 16, 0, 1),
                                                           (7, 1, 28, 1, 1), (4, 6, 5, 1, 1), (3, 8, 2, 0, 0), (1, 4, 7, 0, 0),
                                                           (8, 3, 1, 1, 1)])
def test_segment_layout(checker, n, p, max_len, shuffle, renumber):
    res = subprocess.run([str(checker), str(n), str(p), str(max_len), str(shuffle), str(renumber)], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0 and res.stdout.startswith("ok"), res.stdout + res.stderr
    colors = int(res.stdout.split("colors=")[1].split()[0])
    # segments coloured by their parities along x, y, z: 8 colours once every direction has two of them
    k_row = -(-n // max_len)
    expected = (2 if k_row > 1 else 1) * (2 if n > 1 else 1) ** 2
    assert colors == expected
