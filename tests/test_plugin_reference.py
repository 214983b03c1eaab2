"""The /gpu/hip/mi355 backend inside the reference's own libCEED, on the CPU.

The backend sources (libceed_amd/host/ceed-mi355-{objects,operator}.c) are
compiled against the reference's headers and linked with the reference's
interface and CPU backends built from its sources (oracle/_ref/libceed.so),
registered as a runtime plugin (tests/c/plugin/mi355-plugin.c: CeedRegister from a
constructor).  Their calls into the kernel ABI (include/ceed-mi355-kernels.h)
go to a CPU stand-in (tests/c/plugin/mi355-cpu-standin.c) whose arithmetic is the
oracle's, so the whole libCEED call sequence runs up to the device boundary:

  * tests/c/plugin-sequence.c drives vectors (validity, reads, norms, BLAS-1,
    borrowed arrays), QFunction context data, restriction offsets and apply,
    and CeedQFunctionApply on gallery QFunctions, checking closed forms;
  * the reference's examples/ceed/ex1-volume.c, ex2-surface.c and
    ex3-volume.c, unchanged, run in their test mode (-t) with user and
    gallery QFunctions and check their own volumes and surface areas.

Also: the same examples compile and link unchanged against this repository's
include/ and libceed_amd/lib/libceed-mi355.so (no GPU needed to link).

MI355_SANITIZE=1 builds the backend and the programs with AddressSanitizer and
UndefinedBehaviorSanitizer (leak detection on), so the same tests check the
backend's host code for memory errors and leaks."""

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")
REF_LIB = ROOT / "oracle" / "_ref" / "libceed.so"
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"
HOST = ROOT / "libceed_amd" / "host"
C = ROOT / "tests" / "c"
PLUGIN = C / "plugin"

needs_ref = pytest.mark.skipif(not (REF / "include").is_dir() or not REF_LIB.exists() or not ORACLE_LIB.exists(),
                               reason="reference sources or oracle/_ref build not present")


def run(cmd, **kw):
    res = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert res.returncode == 0, f"{' '.join(map(str, cmd))}\n{res.stdout}\n{res.stderr}"
    return res


@pytest.fixture(scope="module")
def plugin(tmp_path_factory):
    if not ((REF / "include").is_dir() and REF_LIB.exists() and ORACLE_LIB.exists()):
        pytest.skip("reference sources or oracle/_ref build not present")
    d = tmp_path_factory.mktemp("plugin")
    inc = d / "inc"  # only the backend's own headers: <ceed/*.h> must resolve to the reference's
    inc.mkdir()
    for h in ("ceed-mi355.h", "ceed-mi355-kernels.h"):
        shutil.copy(ROOT / "include" / h, inc / h)
    so = d / "libmi355-plugin.so"
    rpath = f"-Wl,-rpath,{REF_LIB.parent}:{ORACLE_LIB.parent}:{d}"
    san = ["-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"] if os.environ.get("MI355_SANITIZE") else []
    run(["gcc", "-shared", "-fPIC", "-O1", *san, "-std=c11", "-Wall", "-Werror", f"-I{REF / 'include'}", f"-I{inc}", f"-I{HOST}", f"-I{ROOT / 'oracle'}",
         str(HOST / "ceed-mi355-objects.c"), str(HOST / "ceed-mi355-operator.c"), str(PLUGIN / "mi355-plugin.c"), str(PLUGIN / "mi355-cpu-standin.c"),
         f"-L{REF_LIB.parent}", "-lceed", f"-L{ORACLE_LIB.parent}", "-loracle", "-ldl", "-lm", rpath, "-o", str(so)])

    def build(src: Path, name: str) -> Path:
        exe = d / name
        # --no-as-needed: nothing in the program names a plugin symbol, the
        # library is there for its constructor
        run(["gcc", "-O1", *san, "-std=c11", f"-I{REF / 'include'}", str(src), f"-L{REF_LIB.parent}", "-lceed", "-Wl,--no-as-needed", str(so),
             "-Wl,--as-needed", "-lm", rpath, "-o", str(exe)])
        return exe

    env = dict(os.environ, STANDIN_CEED_INCLUDE=str(REF / "include"), STANDIN_LOG=str(d / "calls.log"))
    return build, env, d


@needs_ref
def test_plugin_call_sequence(plugin):
    build, env, _ = plugin
    exe = build(C / "plugin-sequence.c", "plugin-sequence")
    res = run([str(exe), "/gpu/hip/mi355"], env=env, timeout=300)
    assert res.stdout.strip().endswith("ok"), res.stdout


@needs_ref
@pytest.mark.parametrize("example,args", [("ex1-volume", ["-d", "3", "-t"]), ("ex1-volume", ["-d", "3", "-t", "-g"]),
                                          ("ex1-volume", ["-d", "2", "-t"]), ("ex1-volume", ["-d", "1", "-t", "-g"]),
                                          ("ex2-surface", ["-d", "3", "-t"]), ("ex2-surface", ["-d", "3", "-t", "-g"]),
                                          ("ex2-surface", ["-d", "2", "-t", "-g"]), ("ex3-volume", ["-d", "3", "-t"])])
def test_reference_example_on_plugin(plugin, example, args):
    """The example's own test mode prints an error line when its volume or surface area is off."""
    build, env, d = plugin
    exe = d / example
    if not exe.exists():
        build(REF / "examples" / "ceed" / f"{example}.c", example)
    log = d / "calls.log"
    log.unlink(missing_ok=True)
    res = run([str(exe), "-ceed", "/gpu/hip/mi355", *args], env=env, timeout=600)
    assert "error" not in res.stdout.lower(), res.stdout
    calls = log.read_text().split()
    assert "DeviceInit" in calls  # the backend, not a delegate, ran it
    if "-g" in args and args[1] == "3":  # gallery apply operators of 3D hexes take the fused plan
        assert "OperatorApply" in calls or "OperatorApplyAdd" in calls, sorted(set(calls))


@pytest.mark.skipif(not (REF / "examples" / "ceed").is_dir(), reason="reference examples not present")
@pytest.mark.parametrize("example", ["ex1-volume", "ex2-surface", "ex3-volume"])
def test_reference_examples_link_against_library(tmp_path, example):
    """The reference's examples compile and link unchanged against include/ and libceed-mi355.so."""
    lib = ROOT / "libceed_amd" / "lib" / "libceed-mi355.so"
    if not lib.exists():
        pytest.skip("library not built")
    run(["gcc", "-O1", "-std=c11", f"-I{ROOT / 'include'}", str(REF / "examples" / "ceed" / f"{example}.c"), f"-L{lib.parent}", "-lceed-mi355",
         f"-Wl,-rpath,{lib.parent}", "-lm", "-o", str(tmp_path / example)])
