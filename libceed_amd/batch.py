"""Many independent hex-mesh systems solved together.

Systems are grouped by (pointwise operator, P, Q); each group becomes one
fused plan over the concatenation of its meshes (element offsets and
geometry shifted into the group's node and vertex ranges).  Two solvers share
that layout:

* the group-batched CG (``CeedMi355BatchCG*``): every system iterates in
  lockstep, an iteration costs one fused launch per group plus the vector
  kernels over the whole batch, captured in a hipGraph;
* the persistent small-system CG (``CeedMi355BatchSmall*``): one launch in
  which each workgroup takes a whole system off a work list (largest first)
  and runs its entire CG, so a small system's latency is its own solve time.
  It is currently switched off (its creation raises): its first GPU test run
  ended in a GPU memory fault whose cause has not been found.

``BatchCG(systems, small_dofs=N)`` sends the systems of at most N DoFs to the
persistent launch and the others to the group-batched CG; the two run
concurrently on two streams.  The batch vector holds the group-batched
systems first (group by group, mesh by mesh), then the small ones.
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


def _group_plans(systems: list[System], members: list[int], scatter: str, autotune: bool, tuning: dict):
    """One fused plan per (operator, P, Q) group of the given systems, with geometric factors built on the GPU.
    Returns (plans, plan begins, order of the systems in the vector, system begins), all relative to 0."""
    groups: dict[tuple, list[int]] = {}
    for i in members:
        s = systems[i]
        groups.setdefault((s.kind, s.mesh.P, s.mesh.P + 1), []).append(i)
    plans, op_begin, order, sys_begin = [], [], [], [0]
    at = 0
    for (kind, P, Q), group in sorted(groups.items()):
        offs, moffs, coords = [], [], [[], [], []]
        node_shift = vert_shift = 0
        for i in group:
            m = systems[i].mesh
            offs.append(m.offsets + node_shift)
            moffs.append(m.mesh_offsets + vert_shift)
            nv = m.num_vertices
            for d in range(3):
                coords[d].append(m.coords[d * nv:(d + 1) * nv])
            node_shift += m.num_nodes
            vert_shift += nv
            order.append(i)
            sys_begin.append(sys_begin[-1] + m.num_nodes)
        off = np.concatenate([o.reshape(-1, P ** 3) for o in offs])
        op = ops.HexOperator(kind, off, off.shape[0], P, Q, node_shift, scatter=scatter)
        xyz = torch.from_numpy(np.concatenate([np.concatenate(c) for c in coords])).cuda()
        op.build_qdata(np.concatenate(moffs), xyz, vert_shift, mesh_P=2)
        if autotune:
            _, ms = op.autotune(5)
            layout, epb = op.launch()
            tuning[(kind, P, Q)] = (layout, epb, ms)
        plans.append(op)
        op_begin.append(at)
        at += node_shift
    return plans, op_begin, order, sys_begin


class BatchCG:
    """Batched CG over a list of systems; autotune=True tunes each group's thread layout and elements per workgroup.
    Systems of at most small_dofs DoFs are solved by the persistent small-system launch (0: none)."""

    def __init__(self, systems: list[System], scatter: str = "auto", autotune: bool = True, small_dofs: int = 0):
        self.systems = systems
        self.tuning: dict[tuple, tuple[int, int, float]] = {}  # (kind, P, Q) -> (thread layout, elements per workgroup, ms)
        small = [i for i, s in enumerate(systems) if s.mesh.num_nodes <= small_dofs]
        large = [i for i, s in enumerate(systems) if s.mesh.num_nodes > small_dofs]
        self._h = self._hs = None
        self.ops, self.small_ops = [], []
        order, sys_begin = [], [0]
        self.n_large = 0
        if large:
            self.ops, op_begin, order, sys_begin = _group_plans(systems, large, scatter, autotune, self.tuning)
            self.n_large = sys_begin[-1]
            handles = (C.c_void_p * len(self.ops))(*[o._h.value for o in self.ops])
            ob = np.array(op_begin, dtype=np.int64)
            sb = np.array(sys_begin, dtype=np.int64)
            h = C.c_void_p()
            _lib.call("CeedMi355BatchCGCreate", len(self.ops), handles, ob.ctypes.data, len(large), sb.ctypes.data, C.byref(h))
            self._h = h
        if small:
            # colour plans: the persistent launch applies each system colour by colour inside one workgroup
            self.small_ops, sob, sorder, ssb = _group_plans(systems, small, "color", False, {})
            handles = (C.c_void_p * len(self.small_ops))(*[o._h.value for o in self.small_ops])
            ob = np.array(sob, dtype=np.int64)
            sb = np.array(ssb, dtype=np.int64)
            h = C.c_void_p()
            _lib.call("CeedMi355BatchSmallCreate", len(self.small_ops), handles, ob.ctypes.data, len(small), sb.ctypes.data, C.byref(h))
            self._hs = h
            order = order + sorder
            sys_begin = sys_begin + [self.n_large + b for b in ssb[1:]]
        self.order = order                                  # batch position -> system index
        self.num_large = len(large)
        self.n = sys_begin[-1]
        self.sys_begin = np.array(sys_begin, dtype=np.int64)
        self.position = np.empty(len(systems), dtype=np.int64)  # system index -> batch position
        self.position[self.order] = np.arange(len(systems))

    def range(self, i: int) -> slice:
        """Slice of the batch vector holding system i."""
        k = self.position[i]
        return slice(int(self.sys_begin[k]), int(self.sys_begin[k + 1]))

    def is_small(self, i: int) -> bool:
        return self.position[i] >= self.num_large

    def solve(self, b: torch.Tensor, x: torch.Tensor, max_iter: int = 5000, rtol: float = 1e-8, check_every: int = 25,
              use_graph: bool = True) -> dict:
        """Solve every system; returns per system (indexed like ``systems``): iterations, true relative residual,
        the check at which a group-batched system was first seen converged (-1 for the small systems), and the ms
        from the call's start to the system's completion (group-batched: the host time of that check; small systems:
        the device clock from their launch's start; NaN where a system never converged)."""
        ns, nl = len(self.systems), self.num_large
        it = np.zeros(ns, dtype=np.int32)
        rr = np.zeros(ns)
        conv = np.full(ns, -1, dtype=np.int32)
        done = np.full(ns, np.nan)
        checks = C.c_int32(0)
        side = th = None
        err = []
        if self._hs is not None:  # the persistent launch on its own stream, from its own host thread, beside the group-batched solve
            import threading

            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            its, rrs, dms = np.zeros(ns - nl, dtype=np.int32), np.zeros(ns - nl), np.zeros(ns - nl)
            off = self.n_large * 8
            args = (self._hs, b.data_ptr() + off, x.data_ptr() + off, int(max_iter), float(rtol), C.c_void_p(side.cuda_stream),
                    its.ctypes.data, rrs.ctypes.data, dms.ctypes.data)

            def run_small():  # returns when the launch has ended (ctypes releases the GIL meanwhile)
                try:
                    torch.cuda.set_device(b.device)
                    _lib.call("CeedMi355BatchSmallSolve", *args)
                except Exception as ex:  # re-raised in the caller
                    err.append(ex)

            th = threading.Thread(target=run_small)
            th.start()
        if self._h is not None:
            itl, rrl, convl, msl = np.zeros(nl, dtype=np.int32), np.zeros(nl), np.zeros(nl, dtype=np.int32), np.zeros(nl)
            _lib.call("CeedMi355BatchCGSolve", self._h, b.data_ptr(), x.data_ptr(), int(max_iter), float(rtol), int(check_every),
                      int(use_graph), ops._stream(), itl.ctypes.data, rrl.ctypes.data, convl.ctypes.data, C.byref(checks), msl.ctypes.data)
            it[:nl], rr[:nl], conv[:nl], done[:nl] = itl, rrl, convl, msl
        if th is not None:
            th.join()
            if err:
                raise err[0]
            torch.cuda.current_stream().wait_stream(side)
            it[nl:], rr[nl:], done[nl:] = its, rrs, dms
        # reorder from batch positions to system indices
        return {"iterations": it[self.position], "rel_residual": rr[self.position], "converged_at_check": conv[self.position],
                "done_ms": done[self.position], "num_checks": checks.value}

    def __del__(self):
        if _lib is None or _lib._lib is None:
            return
        if getattr(self, "_h", None):
            _lib.lib().CeedMi355BatchCGDestroy(C.byref(self._h))
        if getattr(self, "_hs", None):
            _lib.lib().CeedMi355BatchSmallDestroy(C.byref(self._hs))
