"""Benchmark: 3D diffusion (Poisson) operator apply on one MI355X.

Headline (`value`): FP64 diffusion apply v = A u, p = 4 (P = 5 GLL nodes,
Q = 6 Gauss points per direction), structured hex mesh of ~8M DoFs
(n = 50, 8,120,601 DoFs), perturbed trilinear geometry, u ~ U[-1, 1]; in
GDoF/s = 1e-9 * DoFs * applies / second.  One step = one CeedOperatorApply
(zero v, then the fused E^T B^T D B E kernels over all element colours), with
inputs resident in HBM.

Also reported: a sweep over p = 1..8 at the same size in FP64 and FP32; CG
(200 FP64 iterations at p = 4, 8M DoFs; and time to a 1e-10 relative
residual at 16M DoFs, p = 4 and 6, FP64 CG against FP32 CG with the residual
replaced by the FP64 one every 25 iterations); 256 batched uneven solves
against one CG per mesh; the roofline of the fused kernel (algorithmic bytes per apply / kernel time,
against 8 TB/s), and a CPU baseline: the reference's /cpu/self/avx/blocked
(built from its sources, oracle/_ref) timed on this host, one process per core.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--p 4] [--dofs 8e6] [--no-sweep] [--no-cpu] [--no-cg] [--no-batch]
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, vendor datasheet
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector/matrix, vendor datasheet


def alg_bytes(N: int, E: int, P: int, Q: int, ncomp: int = 1, qcomp: int = 6, s: int = 8) -> int:
    """Read u + write v + read q_data + read int32 offsets (each stream counted once)."""
    return s * ncomp * N * 2 + s * qcomp * E * Q ** 3 + 4 * E * P ** 3


def alg_flops(E: int, P: int, Q: int, ncomp: int = 1) -> int:
    """Sum-factorised interp + collocated grad, forward and transpose, plus 15 flops/qpt of Poisson3DApply."""
    return E * ncomp * (2 * (2 * (P ** 3 * Q + P ** 2 * Q ** 2 + P * Q ** 3) + 6 * Q ** 4) + 15 * Q ** 3 + 2 * Q ** 3)


def time_gpu(fn, steps: int, warmup: int):
    """Average ms per call over `steps`, by HIP events on torch's current stream (where the kernels run) and by wall clock."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return s.elapsed_time(e) / steps, wall


LAYOUT_NAMES = ["columns", "lines", "points"]  # the fused kernel's thread layouts, by their index in the library
KERNEL_OF_LAYOUT = ["hex_apply_kernel", "hex_apply_lines_kernel", "hex_apply_points_kernel"]


def run_case(p: int, dofs: float, dtype, steps: int, warmup: int, seed: int = 0, scatter: str = "auto", epb: int = 0, num_comp: int = 1,
             layout: int = 1):
    import torch

    from libceed_amd import ops
    from libceed_amd.mesh import hex_mesh, mesh_size_for_dofs

    n = mesh_size_for_dofs(p, dofs)
    m = hex_mesh(n, p, seed=seed)
    P, Q = p + 1, p + 2
    op = ops.HexOperator.from_mesh(ops.QF_DIFFUSION, m, Q=Q, num_comp=num_comp, dtype=dtype, scatter=scatter)
    g = torch.Generator(device="cuda").manual_seed(seed)
    u = (torch.rand(num_comp * m.num_nodes, dtype=torch.float64, device="cuda", generator=g) * 2 - 1).to(dtype)
    v = torch.zeros_like(u)
    if epb:
        op.set_layout(layout, epb)
    else:
        op.autotune(5)  # the library's (thread layout, elements per workgroup) tuner; results are bitwise the same for all
    layout, epb = op.launch()
    ms_apply, wall = time_gpu(lambda: op.apply(u, v), steps, warmup)
    ms_kernel, _ = time_gpu(lambda: op.apply_add(u, v), steps, warmup)
    s = 8 if dtype == torch.float64 else 4
    B = alg_bytes(m.num_nodes, m.num_elem, P, Q, ncomp=num_comp, s=s)
    mode = {"tmap": "transpose map", "color": "element colours", "split": "multiplicity split", "atomic": "atomics",
            "brick": "element bricks", "segment": "element segments"}[op.scatter_mode()]
    F = alg_flops(m.num_elem, P, Q, ncomp=num_comp)
    kname = KERNEL_OF_LAYOUT[layout]
    res = {
        "p": p, "n": n, "dofs": num_comp * m.num_nodes, "num_comp": num_comp, "elements": m.num_elem,
        "scatter": mode, "layout": LAYOUT_NAMES[layout], "elems_per_block": epb,
        "kernels": {"transpose map": f"one apply = {kname} + transpose_gather_kernel (E^T by the transpose offset map)",
                    "multiplicity split": f"one apply = zero fill + {kname} + split_gather_kernel (E^T: atomic adds for entries of "
                                          "at most two contributions, the ordered transpose map for the rest)",
                    "atomics": f"one apply = zero fill + {kname} (E^T by atomic adds)",
                    "element bricks": "one apply = hex_apply_brick_kernel (a workgroup per brick of elements, shared nodes summed "
                                      "in LDS) + brick_gather_kernel (brick-surface nodes, partial sums in brick order)",
                    "element segments": f"one apply = {len(op.colors())} {kname} launches (one per colour of x-segments: a workgroup "
                                        "per segment of face-sharing elements, the faces inside it summed in LDS)"}.get(
                        mode, f"one apply = {len(op.colors())} {kname} launches (one per element colour)"),
        "ms_per_apply": ms_apply, "wall_ms_per_apply": wall, "gdofs": num_comp * m.num_nodes / (ms_apply * 1e-3) / 1e9,
        # the roofline figures use Apply's kernel time (the operator's kernels of one apply, HIP events on
        # their stream); ApplyAdd's is reported beside it (the split E^T adds a scratch pass there)
        "kernel_ms": ms_apply, "apply_add_ms": ms_kernel, "alg_bytes": B, "alg_flops": F,
        "kernel_GBs": B / (ms_apply * 1e-3) / 1e9, "kernel_TFLOPs": F / (ms_apply * 1e-3) / 1e12,
    }
    res["hbm_frac"] = res["kernel_GBs"] / HBM_PEAK_GBS
    del op, u, v
    torch.cuda.empty_cache()
    return res


def _cg_problem(p: int, dofs: float, seed: int = 0):
    """Diffusion operator (FP64 and FP32 plans) and rhs = A u_exact, u_exact = sin(pi x) sin(pi y) sin(pi z) at the nodes."""
    import torch

    from libceed_amd import ops
    from libceed_amd.mesh import hex_mesh, mesh_size_for_dofs, node_coordinates

    m = hex_mesh(mesh_size_for_dofs(p, dofs), p, seed=seed)
    Q = p + 2
    op = ops.HexOperator.from_mesh(ops.QF_DIFFUSION, m, Q=Q)
    gll = np.array(ops.lobatto_quadrature(p + 1)[0])
    xyz = node_coordinates(m, gll)
    u = torch.from_numpy(np.sin(np.pi * xyz[0]) * np.sin(np.pi * xyz[1]) * np.sin(np.pi * xyz[2])).cuda()
    b = torch.zeros_like(u)
    op.apply(u, b)
    return m, Q, op, b


def run_cg_fixed(p: int, dofs: float, iterations: int = 200) -> dict:
    """FP64 CG, a fixed number of iterations (the benchmark problems' "DoFs/s in CG")."""
    import torch

    from libceed_amd import ops

    m, Q, op, b = _cg_problem(p, dofs)
    cg = ops.CG(op)
    x = torch.zeros_like(b)
    cg.solve(b, x, max_iter=10, rtol=0.0, check_every=10)  # warm-up: graph capture
    x.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = cg.solve(b, x, max_iter=iterations, rtol=0.0, check_every=50)
    sec = time.perf_counter() - t0
    return {"p": p, "dofs": m.num_nodes, "iterations": st.iterations, "seconds": round(sec, 4),
            "ms_per_iteration": round(1e3 * sec / st.iterations, 4), "gdofs_in_cg": round(m.num_nodes * st.iterations / sec / 1e9, 3),
            "rel_residual": st.rel_residual}


def run_cg_mixed(p: int, dofs: float, ks=(10, 25, 50), rtol: float = 1e-10) -> dict:
    """Time to 1e-10: FP64 CG against FP32 CG with FP64 residual replacement every k iterations, for each k in ks."""
    import torch

    from libceed_amd import ops

    m, Q, op, b = _cg_problem(p, dofs)
    op32 = ops.HexOperator.from_mesh(ops.QF_DIFFUSION, m, Q=Q, dtype=torch.float32)
    out = {"p": p, "dofs": m.num_nodes, "rtol": rtol}

    def timed(solver, every):
        x = torch.zeros_like(b)
        solver.solve(b, x, max_iter=2 * every, rtol=0.0, check_every=every)  # warm-up: graph capture
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solver.solve(b, x, max_iter=20000, rtol=rtol, check_every=every)
        sec = time.perf_counter() - t0
        return x, st, {"seconds": round(sec, 4), "iterations": st.iterations, "rel_residual": st.rel_residual,
                       "ms_per_iteration": round(1e3 * sec / max(1, st.iterations), 4)}

    x64, _, out["fp64"] = timed(ops.CG(op), 50)
    xc = x64 - x64.mean()  # the operator is pure Neumann: compare up to constants
    mixed = {}
    for k in ks:
        x, st, r = timed(ops.CG(op32, op), k)
        d = x - x64
        r["residual_replacements"] = st.replacements
        r["rel_error_vs_fp64"] = (torch.linalg.norm(d - d.mean()) / torch.linalg.norm(xc)).item()
        r["speedup"] = round(out["fp64"]["seconds"] / r["seconds"], 3)
        mixed[str(k)] = r
    out["fp32_refine_every_k"] = mixed
    best = max(mixed, key=lambda k: mixed[k]["speedup"])
    out["best_k"] = int(best)
    out["speedup"] = mixed[best]["speedup"]
    # Jacobi preconditioning with the assembled diagonal (FP64 and mixed at the best k)
    diag = torch.zeros_like(b)
    ms_diag, _ = time_gpu(lambda: op.assemble_diagonal(diag), 5, 1)
    jac = {"ms_assemble_diagonal": round(ms_diag, 4)}
    for name, solver, every in (("fp64", ops.CG(op), 50), ("fp32_refine", ops.CG(op32, op), int(best))):
        solver.set_jacobi(diag)
        x, st, r = timed(solver, every)
        d = x - x64
        r["rel_error_vs_fp64"] = (torch.linalg.norm(d - d.mean()) / torch.linalg.norm(xc)).item()
        r["speedup_vs_plain_fp64_cg"] = round(out["fp64"]["seconds"] / r["seconds"], 3)
        jac[name] = r
    out["jacobi"] = jac
    return out


def run_adjoint(dofs: float = 4e6, p: int = 4, seed: int = 0, directions: int = 64, h: float = 1e-6) -> dict:
    """Variable-coefficient diffusion -div(kappa grad u), p=4, Q=5, FP64, kappa = exp(g) at the quadrature points with g a
    smooth random field: forward apply, fused adjoint (A^T lambda and dL/dkappa in one kernel), and central finite
    differences of L = 1/2 ||A(kappa) u - f||^2 along random kappa directions."""
    import torch

    from libceed_amd import ops
    from libceed_amd.mesh import hex_mesh, mesh_size_for_dofs, quadrature_points, smooth_random_field

    m = hex_mesh(mesh_size_for_dofs(p, dofs), p, seed=seed)
    P, Q = p + 1, p + 1
    op = ops.HexOperator.from_mesh(ops.QF_DIFFUSION_KAPPA, m, Q=Q)
    gq = np.array(ops.gauss_quadrature(Q)[0])
    kappa = torch.from_numpy(np.exp(smooth_random_field(quadrature_points(m, gq), seed=seed + 5))).cuda()
    op.set_kappa(kappa)
    rng = np.random.default_rng(seed)
    u = torch.from_numpy(rng.uniform(-1, 1, m.num_nodes)).cuda()
    f = torch.from_numpy(rng.uniform(-1, 1, m.num_nodes)).cuda()
    v = torch.zeros_like(u)
    lam = torch.zeros_like(u)
    dk = torch.zeros_like(kappa)
    ms_fwd, _ = time_gpu(lambda: op.apply(u, v), 20, 3)
    op.apply(u, v)
    r = v - f  # dL/d(Au)
    ms_adj, _ = time_gpu(lambda: op.apply_adjoint(u, r, lam, dk), 20, 3)
    op.apply_adjoint(u, r, lam, dk)

    def loss(k):
        op.set_kappa(k)
        op.apply(u, v)
        return 0.5 * float(torch.dot(v - f, v - f))

    # L is quadratic in kappa (A is linear in it), so central differences are exact up to roundoff for any step:
    # h = 1e-6 shows the roundoff of the loss divided by 2h, the larger step the agreement itself
    worst = {h: 0.0, 1e-2: 0.0}
    g = torch.Generator(device="cuda").manual_seed(seed + 1)
    for _ in range(directions):
        d = torch.rand(kappa.shape, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        an = float(torch.dot(dk, d))
        for step in worst:
            fd = (loss(kappa + step * d) - loss(kappa - step * d)) / (2 * step)
            worst[step] = max(worst[step], abs(fd - an) / max(abs(an), 1e-300))
    op.set_kappa(kappa)
    return {"p": p, "Q": Q, "dofs": m.num_nodes, "elements": m.num_elem, "ms_forward": round(ms_fwd, 4),
            "ms_adjoint_v_and_dkappa": round(ms_adj, 4), "forward_gdofs": round(m.num_nodes / (ms_fwd * 1e-3) / 1e9, 3),
            "adjoint_gdofs": round(m.num_nodes / (ms_adj * 1e-3) / 1e9, 3), "fd_directions": directions,
            "fd_max_rel_error": {f"h={k:g}": v for k, v in worst.items()}}


def run_drop_in(p: int = 4, n: int = 50, reps: int = 50, num_comp: int = 1, user: bool = False, fuse: bool = True):
    """CeedOperatorApply through the libCEED C API on /gpu/hip/mi355 (tests/c/bp-apply.c: gallery Poisson3DBuild +
    Poisson3DApply, or the user QFunction of ex-hex compiled at run time, fused or staged; library defaults, no
    tuning), in a child process."""
    exe = ROOT / "tests" / "c" / "bin" / "bp-apply"
    if not exe.exists():
        return None
    env = dict(os.environ, CEED_MI355_FUSE_QFUNCTIONS="1" if fuse else "0")
    res = subprocess.run([str(exe), "/gpu/hip/mi355", str(p), str(n), str(reps), str(num_comp)] + (["user"] if user else []),
                         capture_output=True, text=True, timeout=600, env=env)
    if res.returncode != 0:
        return {"error": res.stderr.strip()[-300:]}
    return json.loads(res.stdout.strip().splitlines()[-1])


def run_batch(count: int = 256, rtol: float = 1e-8, seed: int = 0) -> dict:
    """Batched uneven solves: the group-batched CG (one fused plan per (operator, P, Q) group, hipGraph-captured
    iterations) against one CG per mesh solved one after the other.  Latency is a system's completion time counted
    from the moment all systems are handed over (the batch's start): for the batched CG the check at which the
    system was first seen converged; for one solve per mesh the solves before it plus its own, in the given order
    and, as a schedule computed from the same measured solve times, shortest solve first."""
    import torch

    from libceed_amd import ops
    from libceed_amd.batch import BatchCG, random_systems

    systems = random_systems(count, seed=seed)
    rng = np.random.default_rng(seed)
    rhs = []
    for s in systems:
        bi = rng.uniform(-1, 1, s.mesh.num_nodes)
        if s.kind == ops.QF_DIFFUSION:  # pure Neumann: project b onto the range of A
            bi -= bi.mean()
        rhs.append(bi)

    def timed(bcg) -> dict:
        b = np.zeros(bcg.n)
        for i in range(count):
            b[bcg.range(i)] = rhs[i]
        bd = torch.from_numpy(b).cuda()
        x = torch.zeros_like(bd)
        bcg.solve(bd, x, max_iter=50, rtol=0.0, check_every=25)  # warm-up: graph capture, code loading
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = bcg.solve(bd, x, max_iter=20000, rtol=rtol, check_every=25)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        # the host time of the check at which each system was first seen converged, from the solve's start
        # (a system that never converged counts as done at the end)
        lat = np.where(np.isnan(st["done_ms"]), sec, 1e-3 * st["done_ms"])
        return {"seconds": round(sec, 4), "solves_per_s": round(count / sec, 2), "iterations_max": int(st["iterations"].max()),
                "latency_p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3), "latency_p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
                "max_rel_residual": float(st["rel_residual"].max())}, bd

    t0 = time.perf_counter()
    bcg = BatchCG(systems)
    setup = time.perf_counter() - t0
    grouped, bd = timed(bcg)
    out = {"systems": count, "dofs_total": int(bcg.n), "groups": len(bcg.ops), "rtol": rtol, "setup_seconds": round(setup, 2),
           "autotuned_launch": {f"{'mass' if k[0] == 0 else 'diff'}_P{k[1]}_Q{k[2]}": [LAYOUT_NAMES[v[0]], v[1]]
                                for k, v in bcg.tuning.items()},
           "batched": grouped}
    del bcg
    torch.cuda.empty_cache()
    # the same systems in 4 size classes (equal counts by DoFs), a group-batched CG each (with its own stream):
    # solved one class after the other, smallest first (a system's latency: the classes before it plus its own
    # class's solve up to its convergence), and all four at once from four host threads
    import threading

    dofs = np.array([s.mesh.num_nodes for s in systems])
    classes = np.array_split(np.argsort(dofs, kind="stable"), 4)
    subs = []
    for members in classes:
        sub = BatchCG([systems[i] for i in members])
        b = np.zeros(sub.n)
        for j, i in enumerate(members):
            b[sub.range(j)] = rhs[i]
        bd = torch.from_numpy(b).cuda()
        x = torch.zeros_like(bd)
        sub.solve(bd, x, max_iter=50, rtol=0.0, check_every=25)  # warm-up: graph capture, code loading
        subs.append((members, sub, bd, x))
    torch.cuda.synchronize()

    def summary(lat, tot, worst, iters, **extra):
        return {"classes": 4, **extra, "seconds": round(tot, 4), "solves_per_s": round(count / tot, 2), "iterations_max_per_class": iters,
                "latency_p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3), "latency_p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
                "max_rel_residual": worst}

    lat_c, tot_c, worst_c, iters_c = np.zeros(count), 0.0, 0.0, []
    for members, sub, bd, x in subs:
        x.zero_()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        st = sub.solve(bd, x, max_iter=20000, rtol=rtol, check_every=25)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t1
        lat_c[members] = tot_c + np.where(np.isnan(st["done_ms"]), sec, 1e-3 * st["done_ms"])
        tot_c += sec
        worst_c = max(worst_c, float(st["rel_residual"].max()))
        iters_c.append(int(st["iterations"].max()))
    out["batched_by_size_class"] = summary(lat_c, tot_c, worst_c, iters_c, order="smallest class first")
    # all classes at once: each thread's solve blocks on its own stream's checks (ctypes releases the GIL)
    for _, _, _, x in subs:
        x.zero_()
    torch.cuda.synchronize()
    results, errors = [None] * 4, []
    t0 = time.perf_counter()

    def run(k):
        try:
            torch.cuda.set_device(0)
            members, sub, bd, x = subs[k]
            with torch.cuda.stream(torch.cuda.Stream()):  # a stream of this thread's own (default priority)
                t1 = time.perf_counter()
                st = sub.solve(bd, x, max_iter=20000, rtol=rtol, check_every=25)
                results[k] = (t1 - t0, time.perf_counter() - t0, st)
        except Exception as ex:  # re-raised below
            errors.append(ex)

    threads = [threading.Thread(target=run, args=(k,)) for k in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    torch.cuda.synchronize()
    if errors:
        raise errors[0]
    lat_a, worst_a, iters_a = np.zeros(count), 0.0, []
    for (members, _, _, _), (begin, end, st) in zip(subs, results):
        lat_a[members] = np.where(np.isnan(st["done_ms"]), end, begin + 1e-3 * st["done_ms"])
        worst_a = max(worst_a, float(st["rel_residual"].max()))
        iters_a.append(int(st["iterations"].max()))
    out["batched_by_size_class_concurrent"] = summary(lat_a, max(r[1] for r in results), worst_a, iters_a, order="all classes at once")
    del subs
    torch.cuda.empty_cache()
    # one CG per mesh, sequentially
    lat1 = []
    worst = 0.0
    for i, s in enumerate(systems):
        op = ops.HexOperator.from_mesh(s.kind, s.mesh, Q=s.mesh.P + 1)
        cg = ops.CG(op)
        bi = torch.from_numpy(rhs[i]).cuda()
        xi = torch.zeros_like(bi)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sti = cg.solve(bi, xi, max_iter=20000, rtol=rtol, check_every=25)
        lat1.append(time.perf_counter() - t1)
        worst = max(worst, sti.rel_residual)
        del cg, op
    tot = float(np.sum(lat1))
    done = np.cumsum(lat1)                 # completion times in the given order
    done_sjf = np.cumsum(np.sort(lat1))    # the same solves, shortest first
    pct = lambda a, q: round(1e3 * float(np.percentile(a, q)), 3)
    out["one_solve_per_mesh"] = {"seconds": round(tot, 4), "solves_per_s": round(count / tot, 2),
                                 "latency_p50_ms": pct(done, 50), "latency_p99_ms": pct(done, 99),
                                 "latency_shortest_first_p50_ms": pct(done_sjf, 50), "latency_shortest_first_p99_ms": pct(done_sjf, 99),
                                 "solve_ms_p50": pct(lat1, 50), "solve_ms_p99": pct(lat1, 99), "max_rel_residual": worst}
    out["speedup_solves_per_s"] = round(out["batched"]["solves_per_s"] / out["one_solve_per_mesh"]["solves_per_s"], 3)
    for key in ("batched_by_size_class", "batched_by_size_class_concurrent"):
        out[f"speedup_solves_per_s_{key[8:]}"] = round(out[key]["solves_per_s"] / out["one_solve_per_mesh"]["solves_per_s"], 3)
    return out


def physical_cores() -> tuple[int, int]:
    """(physical cores, logical CPUs) this process may run on: distinct (package, core) pairs of the affinity set,
    capped by OMP_NUM_THREADS where the host sets the job's CPU share there."""
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        t = Path(f"/sys/devices/system/cpu/cpu{c}/topology")
        try:
            cores.add(((t / "physical_package_id").read_text().strip(), (t / "core_id").read_text().strip()))
        except OSError:
            cores.add(("?", str(c)))
    n = len(cores)
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n), len(cpus)


def cpu_model() -> tuple[str, bool]:
    """Model name and whether the CPU has the x86-64-v4 (AVX-512 F/CD/BW/DQ/VL) feature set."""
    model, flags = "", set()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and not model:
                model = line.split(":", 1)[1].strip()
            elif line.startswith("flags") and not flags:
                flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return model, {"avx512f", "avx512cd", "avx512bw", "avx512dq", "avx512vl"} <= flags


def cpu_baseline(p: int = 4, n: int = 50, seconds: float = 10.0) -> dict | None:
    """Reference /cpu/self/avx/blocked (built from its sources, oracle/_ref) on this host: one single-threaded process
    per physical core, each on an equal slab of the GPU workload's n^3-element mesh (n x n x n/cores elements)."""
    exes = {"x86-64-v3": ROOT / "oracle" / "_ref" / "ref_cpu_bench", "x86-64-v4": ROOT / "oracle" / "_ref" / "ref_cpu_bench4"}
    model, has_v4 = cpu_model()
    if not has_v4:
        exes.pop("x86-64-v4")
    exes = {k: v for k, v in exes.items() if v.exists()}
    if exes:
        cores, logical = physical_cores()
        nz = max(1, round(n / cores))
        backend = "/cpu/self/avx/blocked"

        def launch(exe, secs, count):
            procs = [subprocess.Popen([str(exe), backend, str(p), str(n), str(n), str(nz), str(secs)], stdout=subprocess.PIPE, text=True)
                     for _ in range(count)]
            rates, dofs = [], 0
            for pr in procs:
                out, _ = pr.communicate(timeout=secs * 6 + 120)
                if pr.returncode != 0:
                    return None, 0
                r = json.loads(out.strip().splitlines()[-1])
                rates.append(r["dofs"] * r["applies"] / r["seconds"])
                dofs = r["dofs"]
            return sum(rates), dofs

        # the faster build on this CPU (AVX2 or AVX-512 code generation; the reference's default is -march=native)
        trial = {k: launch(v, 2.0, 1)[0] or 0.0 for k, v in exes.items()}
        march = max(trial, key=trial.get)
        total, dofs = launch(exes[march], seconds, cores)
        if total is None:
            return None
        return {"value": total / 1e9, "unit": "GDoF/s", "cores": cores, "kind": "reference",
                "sample": f"{backend} (reference sources, -O3 -march={march} -ffp-contract=fast), 3D Poisson p={p} Q={p + 2}, "
                          f"{cores} single-threaded processes, each on a {n}x{n}x{nz}-element slab ({dofs} DoFs) of the {n}^3 mesh, "
                          f"{seconds:.0f} s",
                "cpu": model, "logical_cpus": logical, "march_trial_gdofs_1core": {k: round(v / 1e9, 4) for k, v in trial.items()}}
    # fallback: the oracle restatement of /cpu/self/ref/serial, one core
    import oracle
    from libceed_amd.mesh import hex_mesh

    m = hex_mesh(8, p, seed=0)
    qd = oracle.hex_build_qdata(1, m, p + 2)
    u = np.random.default_rng(0).uniform(-1, 1, m.num_nodes)
    k, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        oracle.hex_apply(1, 1, m.num_elem, p + 1, p + 2, m.offsets, m.num_nodes, qd, u)
        k += 1
    dt = time.perf_counter() - t0
    return {"value": m.num_nodes * k / dt / 1e9, "unit": "GDoF/s", "cores": 1, "kind": "port",
            "sample": f"oracle (ref/serial restatement), p={p}, 8^3 elements, {dt:.0f} s"}


def measured_traffic(head: dict, dtype: str, num_comp: int):
    """HBM bytes per apply from the committed rocprofv3 counter summary of this workload (same order, size,
    arithmetic, components and E^T mode), and where it comes from.  A summary of another launch of the same
    plan (layout, elements per workgroup) is used when none of this launch exists, and says so: the launch
    changes the order of the memory accesses, not the bytes the plan reads and writes."""
    best = None
    for f in sorted((ROOT / "profiles").glob("*_traffic.json")):
        try:
            d = json.loads(f.read_text())
        except (OSError, ValueError):
            continue
        w = d.get("workload", {})
        if (w.get("p"), w.get("dofs"), w.get("dtype"), w.get("scatter"), w.get("num_comp", 1)) != (head["p"], head["dofs"], dtype, head["scatter"], num_comp):
            continue
        same = (w.get("layout"), w.get("elems_per_block")) == (head["layout"], head["elems_per_block"])
        src = d.get("source", f.name) + ("" if same else f" (launch {w.get('layout')} x {w.get('elems_per_block')}; this run's launch "
                                                           f"{head['layout']} x {head['elems_per_block']})")
        if best is None or same:
            best = (d["traffic_bytes_per_apply"], src)
        if same:
            break
    return best if best else (None, None)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--dofs", type=float, default=8e6)
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--scatter", choices=["auto", "tmap", "color", "split", "atomic", "brick", "segment"], default="auto",
                    help="E^T mode of the headline plan (auto is the library default; the others are for profiling)")
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64",
                    help="arithmetic of the headline apply (the metric is FP64; f32 is for profiling)")
    ap.add_argument("--ncomp", type=int, default=1, choices=[1, 3],
                    help="components of the headline operator (the metric is scalar; 3 is the vector variant, for profiling)")
    ap.add_argument("--epb", type=int, default=0,
                    help="elements per workgroup of the headline plan (default: autotuned; fixed values are for profiling)")
    ap.add_argument("--layout", choices=LAYOUT_NAMES, default="lines",
                    help="thread layout of the headline plan when --epb is given")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="time a plain 1 GiB device copy (HBM reference for bandwidth and counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; runs over several GPUs are not supported (the operator apply is per-GPU; sharding is by element partition with a halo exchange)"}))
        sys.exit(2)

    import torch

    from libceed_amd import ops

    ops.device_init(0)
    if args.copy_calibration:
        nbytes = 1 << 30
        a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        ms, _ = time_gpu(lambda: b.copy_(a), args.steps, args.warmup)
        big = torch.ones(4 * nbytes // 8, dtype=torch.float64, device="cuda")
        ms_read, _ = time_gpu(lambda: big.sum(), args.steps, args.warmup)
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "ms_per_copy": ms,
                          "GBs_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9,
                          "sum_bytes_read": 4 * nbytes, "ms_per_sum": ms_read, "GBs_read": 4 * nbytes / (ms_read * 1e-3) / 1e9}), flush=True)
        return
    head = run_case(args.p, args.dofs, torch.float64 if args.dtype == "f64" else torch.float32, args.steps, args.warmup, scatter=args.scatter,
                    epb=args.epb, layout=LAYOUT_NAMES.index(args.layout), num_comp=args.ncomp)
    sweep = {}
    if not args.no_sweep:
        for p in range(1, 9):
            r64 = head if p == args.p else run_case(p, args.dofs, torch.float64, max(10, args.steps // 4), 3)
            r32 = run_case(p, args.dofs, torch.float32, max(10, args.steps // 4), 3)
            sweep[str(p)] = {"dofs": r64["dofs"], "fp64_gdofs": round(r64["gdofs"], 3), "fp64_hbm_frac": round(r64["hbm_frac"], 4),
                             "fp64_launch": [r64["layout"], r64["elems_per_block"]], "fp32_gdofs": round(r32["gdofs"], 3),
                             "fp32_hbm_frac": round(r32["hbm_frac"], 4), "fp32_launch": [r32["layout"], r32["elems_per_block"]]}
            print(f"# p={p}: {sweep[str(p)]}", file=sys.stderr, flush=True)
    sizes = {}
    if not args.no_sweep:
        # the other sizes of the sweep's workload (2M and 32M DoFs) and the 3-component operator (8M nodes), p = 1..8
        for p, label, d, nc in [(p, f"p{p}_{tag}", d, nc) for p in range(1, 9) for tag, d, nc in (("2M", 2e6, 1), ("32M", 32e6, 1),
                                                                                                    ("8M_vector3", 8e6, 3))]:
            r = run_case(p, d, torch.float64, max(10, args.steps // 4), 3, num_comp=nc)
            sizes[label] = {"dofs": r["dofs"], "num_comp": nc, "fp64_gdofs": round(r["gdofs"], 3), "fp64_hbm_frac": round(r["hbm_frac"], 4),
                            "launch": [r["layout"], r["elems_per_block"]]}
            print(f"# {label}: {sizes[label]}", file=sys.stderr, flush=True)
    drop_in = {}
    if not args.no_sweep:
        for nc in (1, 3):
            drop_in[f"p4_n50_ncomp{nc}"] = run_drop_in(num_comp=nc)
            print(f"# drop-in C API: {drop_in[f'p4_n50_ncomp{nc}']}", file=sys.stderr, flush=True)
        for fuse in (True, False):
            key = f"p4_n50_user_{'fused' if fuse else 'staged'}"
            drop_in[key] = run_drop_in(user=True, fuse=fuse)
            print(f"# drop-in C API: {drop_in[key]}", file=sys.stderr, flush=True)
    adjoint = None
    if not args.no_sweep:
        adjoint = run_adjoint()
        print(f"# adjoint: {adjoint}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    cg = {}
    if not args.no_cg:
        cg["fp64_200_iterations_p4_8M"] = run_cg_fixed(4, 8e6)
        print(f"# cg: {cg['fp64_200_iterations_p4_8M']}", file=sys.stderr, flush=True)
        for p in (4, 6):
            cg[f"to_1e-10_p{p}_16M"] = run_cg_mixed(p, 16e6)
            print(f"# cg p={p}: {cg[f'to_1e-10_p{p}_16M']}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    batch = None
    if not args.no_batch:
        batch = run_batch()
        print(f"# batch: {batch}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    cpu = None if args.no_cpu else cpu_baseline(args.p, head["n"])
    traffic = measured_traffic(head, args.dtype, args.ncomp)
    line = {
        "metric": "3D diffusion apply GDoF/s p=1..8 (FP64 vs FP32+refine @1e-10), % HBM peak",
        "value": round(head["gdofs"], 4),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_apply"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic: seeded perturbed structured hex mesh, u ~ U[-1,1]",
        "config": {"workload": f"3D {'3-component ' if args.ncomp == 3 else ''}Poisson operator apply, {args.dtype.upper()}, p={args.p} (P={args.p + 1}, Q={args.p + 2}), "
                               f"{head['n']}^3 hex elements, {head['dofs']} DoFs",
                   "p": args.p, "dofs": head["dofs"], "elements": head["elements"], "scatter": head["scatter"],
                   "layout": head["layout"], "elems_per_block": head["elems_per_block"], "num_comp": args.ncomp},
        "roofline": {"bound": "hbm", "achieved": round(head["kernel_GBs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(head["hbm_frac"], 4), "traffic": traffic[0], "traffic_source": traffic[1],
                     "kernel": head["kernels"],
                     "alg_bytes_per_apply": head["alg_bytes"], "kernel_ms_per_apply": round(head["kernel_ms"], 5),
                     "apply_add_ms": round(head["apply_add_ms"], 5),
                     "fp64_tflops": round(head["kernel_TFLOPs"], 2), "fp64_peak_tflops": FP64_PEAK_TFLOPS},
        "cpu_baseline": cpu,
        "sweep": sweep,
        "sizes": sizes,
        "drop_in_c_api": drop_in,
        "adjoint": adjoint,
        "cg": cg,
        "batched_solves": batch,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
