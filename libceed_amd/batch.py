"""Many independent hex-mesh systems solved together by the batched CG.

Systems are grouped by (pointwise operator, P, Q); each group becomes one
fused plan over the concatenation of its meshes (element offsets and
geometry shifted into the group's node and vertex ranges), so an iteration
costs one fused launch per group whatever the number of meshes.  The batch
vector is ordered group by group, mesh by mesh.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, ops
from .mesh import HexMesh, hex_mesh


@dataclass
class System:
    kind: int          # ops.QF_MASS or ops.QF_DIFFUSION
    mesh: HexMesh


def random_systems(count: int = 256, seed: int = 0, max_n: int = 24, p_range=(2, 7)) -> list[System]:
    """Meshes with element counts log-uniform in [1, max_n^3] (cubes of n^3 elements), orders uniform in
    p_range, mass or diffusion with equal probability."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        elems = np.exp(rng.uniform(0.0, np.log(max_n ** 3)))
        n = int(min(max_n, max(1, round(elems ** (1.0 / 3.0)))))
        p = int(rng.integers(p_range[0], p_range[1] + 1))
        kind = ops.QF_MASS if rng.random() < 0.5 else ops.QF_DIFFUSION
        out.append(System(kind, hex_mesh(n, p, seed=seed * 100003 + i)))
    return out


class BatchCG:
    """Batched CG over a list of systems; autotune=True tunes each group's thread layout and elements per workgroup."""

    def __init__(self, systems: list[System], scatter: str = "auto", autotune: bool = True):
        self.systems = systems
        groups: dict[tuple, list[int]] = {}
        for i, s in enumerate(systems):
            groups.setdefault((s.kind, s.mesh.P, s.mesh.P + 1), []).append(i)
        self.order: list[int] = []          # batch position -> system index
        self.ops: list[ops.HexOperator] = []
        self.tuning: dict[tuple, tuple[int, int, float]] = {}  # (kind, P, Q) -> (thread layout, elements per workgroup, ms)
        op_begin, sys_begin = [], [0]
        at = 0
        for (kind, P, Q), members in sorted(groups.items()):
            offs, moffs, coords = [], [], [[], [], []]
            node_shift = vert_shift = 0
            for i in members:
                m = systems[i].mesh
                offs.append(m.offsets + node_shift)
                moffs.append(m.mesh_offsets + vert_shift)
                nv = m.num_vertices
                for d in range(3):
                    coords[d].append(m.coords[d * nv:(d + 1) * nv])
                node_shift += m.num_nodes
                vert_shift += nv
                self.order.append(i)
                sys_begin.append(sys_begin[-1] + m.num_nodes)
            off = np.concatenate([o.reshape(-1, P ** 3) for o in offs])
            op = ops.HexOperator(kind, off, off.shape[0], P, Q, node_shift, scatter=scatter)
            xyz = torch.from_numpy(np.concatenate([np.concatenate(c) for c in coords])).cuda()
            op.build_qdata(np.concatenate(moffs), xyz, vert_shift, mesh_P=2)
            if autotune:
                _, ms = op.autotune(5)
                layout, epb = op.launch()
                self.tuning[(kind, P, Q)] = (layout, epb, ms)
            self.ops.append(op)
            op_begin.append(at)
            at += node_shift
        self.n = at
        self.sys_begin = np.array(sys_begin, dtype=np.int64)
        self.position = np.empty(len(systems), dtype=np.int64)  # system index -> batch position
        self.position[self.order] = np.arange(len(systems))
        handles = (C.c_void_p * len(self.ops))(*[o._h.value for o in self.ops])
        ob = np.array(op_begin, dtype=np.int64)
        h = C.c_void_p()
        _lib.call("CeedMi355BatchCGCreate", len(self.ops), handles, ob.ctypes.data, len(systems), self.sys_begin.ctypes.data, C.byref(h))
        self._h = h

    def range(self, i: int) -> slice:
        """Slice of the batch vector holding system i."""
        k = self.position[i]
        return slice(int(self.sys_begin[k]), int(self.sys_begin[k + 1]))

    def solve(self, b: torch.Tensor, x: torch.Tensor, max_iter: int = 5000, rtol: float = 1e-8, check_every: int = 25,
              use_graph: bool = True) -> dict:
        ns = len(self.systems)
        it = np.zeros(ns, dtype=np.int32)
        rr = np.zeros(ns)
        conv = np.zeros(ns, dtype=np.int32)
        checks = C.c_int32()
        _lib.call("CeedMi355BatchCGSolve", self._h, b.data_ptr(), x.data_ptr(), int(max_iter), float(rtol), int(check_every), int(use_graph),
                  ops._stream(), it.ctypes.data, rr.ctypes.data, conv.ctypes.data, C.byref(checks))
        # reorder from batch positions to system indices
        return {"iterations": it[self.position], "rel_residual": rr[self.position], "converged_at_check": conv[self.position],
                "num_checks": checks.value}

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None and _lib._lib is not None:
            _lib.lib().CeedMi355BatchCGDestroy(C.byref(self._h))
