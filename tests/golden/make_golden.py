"""Write the golden fixtures of tests/golden/*.npz from the reference library.

Runs only in the build container (the reference sources are needed): builds
oracle/_ref/libceed.so (oracle/Makefile) and the driver gen_ref_vectors.c
against it, feeds it seeded inputs and stores inputs + outputs as .npz
(numpy, no pickles).  The fixtures are data: /cpu/self/ref/serial outputs on
small cases.  Usage: python tests/golden/make_golden.py
"""

from __future__ import annotations

import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
from libceed_amd.mesh import hex_mesh  # noqa: E402

REF = Path("/root/reference")
REF_OUT = ROOT / "oracle" / "_ref"
DRIVER = REF_OUT / "gen_ref_vectors"


def build_driver() -> None:
    # oracle/Makefile builds the reference library and this driver (gen_ref_vectors.c) into oracle/_ref
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref", "-j8"], check=True)


def run(*args) -> None:
    subprocess.run([str(DRIVER), *map(str, args)], check=True)


def golden_basis(tmp: Path) -> dict:
    out = {}
    for P in range(2, 11):
        for Q in sorted({max(1, P - 1), P, P + 1, P + 2} - {11, 12}):
            for m in (0, 1):
                if m == 1 and Q < 2:
                    continue
                f = tmp / "b.bin"
                run("basis", P, Q, m, f)
                out[f"P{P}_Q{Q}_m{m}"] = np.fromfile(f, dtype=np.float64)
    return out


def golden_bapply(tmp: Path) -> dict:
    rng = np.random.default_rng(300)
    out = {}
    for dim in (1, 2, 3):
        for nc in (1, 3):
            for P, Q in ((2, 3), (3, 4), (4, 4), (4, 3), (5, 6)):
                for tm in (0, 1):
                    for em in (1, 2, 16):
                        if em == 16 and (tm == 1 or nc != 1):
                            continue
                        nn, nq = P ** dim, Q ** dim
                        qsize = dim * nc * nq if em == 2 else (nq if em == 16 else nc * nq)
                        in_n = nc * nn if tm == 0 else qsize
                        u = rng.uniform(-1, 1, in_n)
                        u.tofile(tmp / "u.bin")
                        run("bapply", dim, nc, P, Q, tm, em, tmp / "u.bin", tmp / "v.bin")
                        key = f"d{dim}_c{nc}_P{P}_Q{Q}_t{tm}_e{em}"
                        out[key + "_u"] = u
                        out[key + "_v"] = np.fromfile(tmp / "v.bin", dtype=np.float64)
    return out


def golden_rstr(tmp: Path) -> dict:
    rng = np.random.default_rng(200)
    out = {}
    cases = [
        # (name, E, es, nc, cs, lsize, offsets)
        ("t200", 3, 2, 1, 1, 4, np.array([0, 1, 1, 2, 2, 3], dtype=np.int32)),
    ]
    for k, (E, es, nc, nnodes) in enumerate([(5, 4, 1, 12), (7, 8, 3, 20), (16, 27, 2, 64), (1, 1, 1, 1)]):
        off = rng.integers(0, nnodes, size=E * es).astype(np.int32)
        cases.append((f"rand{k}", E, es, nc, nnodes, nc * nnodes, off))
    for name, E, es, nc, cs, lsize, off in cases:
        u = rng.uniform(-1, 1, lsize)
        v0 = rng.uniform(-1, 1, lsize)
        off.tofile(tmp / "o.bin")
        u.tofile(tmp / "u.bin")
        v0.tofile(tmp / "v0.bin")
        run("rstr", E, es, nc, cs, lsize, tmp / "o.bin", tmp / "u.bin", tmp / "v0.bin", tmp / "e.bin", tmp / "v.bin")
        out[f"{name}_meta"] = np.array([E, es, nc, cs, lsize], dtype=np.int64)
        out[f"{name}_off"] = off
        out[f"{name}_u"] = u
        out[f"{name}_v0"] = v0
        out[f"{name}_e"] = np.fromfile(tmp / "e.bin", dtype=np.float64)
        out[f"{name}_v"] = np.fromfile(tmp / "v.bin", dtype=np.float64)
    return out


def golden_op(tmp: Path) -> dict:
    out = {}
    cases = [(kind, 1, 2, p, p + 2) for kind in (0, 1) for p in range(1, 9)]
    cases += [(1, 3, 2, 2, 4), (0, 3, 2, 4, 6), (1, 3, 2, 4, 6), (1, 1, 2, 4, 5), (1, 1, 3, 3, 5), (0, 1, 3, 3, 5)]
    for kind, nc, n, p, Q in cases:
        m = hex_mesh(n, p, seed=1000 + 10 * n + p)
        rng = np.random.default_rng(77 + p + 10 * kind + 100 * nc)
        u = rng.uniform(-1, 1, nc * m.num_nodes)
        pre = tmp / "op"
        m.offsets.tofile(str(pre) + ".off")
        m.mesh_offsets.tofile(str(pre) + ".moff")
        m.coords.tofile(str(pre) + ".coords")
        u.tofile(str(pre) + ".u")
        run("op", kind, nc, m.num_elem, p + 1, Q, m.num_vertices, m.num_nodes, pre)
        key = f"k{kind}_c{nc}_n{n}_p{p}_Q{Q}"
        out[key + "_meta"] = np.array([kind, nc, n, p, Q, 1000 + 10 * n + p], dtype=np.int64)
        out[key + "_u"] = u
        out[key + "_v"] = np.fromfile(str(pre) + ".v", dtype=np.float64)
        qd = np.fromfile(str(pre) + ".qdata", dtype=np.float64)
        # geometric factors are kept only where they are small
        out[key + "_qdata"] = qd if qd.size <= 20000 else qd[:20000]
    return out


def main() -> None:
    build_driver()
    with tempfile.TemporaryDirectory() as d:
        tmp = Path(d)
        for name, fn in (("basis", golden_basis), ("bapply", golden_bapply), ("rstr", golden_rstr), ("op", golden_op)):
            data = fn(tmp)
            np.savez_compressed(HERE / f"golden_{name}.npz", **data)
            print(name, len(data), "arrays")


if __name__ == "__main__":
    main()
