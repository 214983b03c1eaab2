"""Batched CG over independent meshes of uneven size and order (mass and
diffusion), one fused plan per (operator, P, Q) group: every system converges
to the tolerance, each solution equals the one the single-system CG gives, and
graph replay gives the same bits as the eager iteration.  The autotuner's
choice is admissible and recorded.
"""

import numpy as np
import pytest
import torch

import oracle
from libceed_amd import ops
from libceed_amd.batch import BatchCG, System, random_systems
from libceed_amd.mesh import hex_mesh

pytestmark = pytest.mark.gpu


def batch_rhs(bcg, seed=0):
    rng = np.random.default_rng(seed)
    b = np.zeros(bcg.n)
    for i, s in enumerate(bcg.systems):
        r = bcg.range(i)
        bi = rng.uniform(-1, 1, r.stop - r.start)
        if s.kind == ops.QF_DIFFUSION:  # pure Neumann: keep b in the range of A (orthogonal to constants)
            bi -= bi.mean()
        b[r] = bi
    return torch.from_numpy(b).cuda()


@pytest.fixture(scope="module")
def batch(cuda):
    systems = random_systems(14, seed=3, max_n=4, p_range=(2, 4))
    return BatchCG(systems)


def test_batch_converges_and_matches_single_solves(batch):
    b = batch_rhs(batch)
    x = torch.zeros_like(b)
    st = batch.solve(b, x, max_iter=3000, rtol=1e-10, check_every=10)
    assert np.all(st["rel_residual"] <= 1e-10), st["rel_residual"]
    assert np.all(st["converged_at_check"] >= 0)
    # completion times: finite, positive, ordered like the checks at which the systems converged
    done, at = st["done_ms"], st["converged_at_check"]
    assert np.all(np.isfinite(done)) and np.all(done > 0)
    order = np.argsort(at, kind="stable")
    assert np.all(np.diff(done[order]) >= 0)
    for (kind, P, Q), (layout, epb, ms) in batch.tuning.items():
        assert layout in (ops.LAYOUT_COLUMNS, ops.LAYOUT_LINES, ops.LAYOUT_POINTS) and ms > 0
        assert 1 <= epb and epb * (Q ** 3 if layout == ops.LAYOUT_POINTS else Q * Q) <= 256
    for i, s in enumerate(batch.systems):
        r = batch.range(i)
        op = ops.HexOperator.from_mesh(s.kind, s.mesh, Q=s.mesh.P + 1)
        xi = torch.zeros(r.stop - r.start, dtype=torch.float64, device="cuda")
        ops.CG(op).solve(b[r].contiguous(), xi, max_iter=3000, rtol=1e-12, check_every=10)
        d = x[r] - xi
        if s.kind == ops.QF_DIFFUSION:
            d = d - d.mean()
        assert (torch.linalg.norm(d) / torch.linalg.norm(xi)).item() < 1e-8, (i, s.kind, s.mesh.n, s.mesh.p)


def test_batch_graph_matches_eager(batch):
    b = batch_rhs(batch, seed=1)
    xs = []
    for use_graph in (False, True, True):
        x = torch.zeros_like(b)
        st = batch.solve(b, x, max_iter=60, rtol=1e-14, check_every=20, use_graph=use_graph)
        torch.cuda.synchronize()
        xs.append((x.cpu().numpy(), st["iterations"].copy()))
    for x, it in xs[1:]:
        assert np.array_equal(x, xs[0][0]) and np.array_equal(it, xs[0][1])


@pytest.mark.parametrize("scatter", ["color", "split"])
def test_batch_scatter_modes(cuda, scatter):
    """The batched solve with colour or multiplicity-split plans (graph-captured, frozen systems skipped)
    converges to the solution of the transpose-map batch."""
    systems = random_systems(8, seed=5, max_n=4, p_range=(2, 4))
    ref, bcg = BatchCG(systems, scatter="tmap"), BatchCG(systems, scatter=scatter)
    b = batch_rhs(ref, seed=2)
    x_ref, x = torch.zeros_like(b), torch.zeros_like(b)
    ref.solve(b, x_ref, max_iter=3000, rtol=1e-11, check_every=10)
    st = bcg.solve(b, x, max_iter=3000, rtol=1e-11, check_every=10)
    assert np.all(st["rel_residual"] <= 1e-11), st["rel_residual"]
    for i, s in enumerate(systems):
        r = bcg.range(i)
        d = x[r] - x_ref[r]
        if s.kind == ops.QF_DIFFUSION:
            d = d - d.mean()
        assert (torch.linalg.norm(d) / torch.linalg.norm(x_ref[r])).item() < 1e-8, (scatter, i)


def dense_operator(kind, mesh):
    """The oracle's operator as a dense matrix: column j = A e_j (oracle q_data, Q = P + 1)."""
    Q = mesh.P + 1
    qd = oracle.hex_build_qdata(0 if kind == ops.QF_MASS else 1, mesh, Q)
    N = mesh.num_nodes
    A = np.zeros((N, N))
    e = np.zeros(N)
    for j in range(N):
        e[j] = 1.0
        A[:, j] = oracle.hex_apply(0 if kind == ops.QF_MASS else 1, 1, mesh.num_elem, mesh.P, Q, mesh.offsets, N, qd, e)
        e[j] = 0.0
    return A


def oracle_solution(kind, mesh, b):
    """Mass: A^-1 b.  Diffusion (pure Neumann, null space = constants, b orthogonal to them): the
    minimum-norm solution, (A + 11^T/N)^-1 b, which is orthogonal to the constants."""
    A = dense_operator(kind, mesh)
    if kind == ops.QF_DIFFUSION:
        A = A + 1.0 / mesh.num_nodes
    return np.linalg.solve(A, b)


def test_batch_every_group_vs_oracle(cuda):
    """A batch holding every (mass or diffusion, p = 2..7) group of the batched workload, two meshes of
    different size per group (one of them a single element): every system's batched solution against the
    oracle's dense solve, at 1e-8 (diffusion compared modulo constants)."""
    systems = every_group_systems()
    bcg = BatchCG(systems)
    assert {(s.kind, s.mesh.P) for s in systems} == {(k, p + 1) for k in (ops.QF_MASS, ops.QF_DIFFUSION) for p in range(2, 8)}
    b = batch_rhs(bcg, seed=4)
    x = torch.zeros_like(b)
    st = bcg.solve(b, x, max_iter=5000, rtol=1e-12, check_every=10)
    assert np.all(st["rel_residual"] <= 1e-12), st["rel_residual"]
    xh, bh = x.cpu().numpy(), b.cpu().numpy()
    for i, s in enumerate(systems):
        r = bcg.range(i)
        print("batch vs oracle", i, s.kind, s.mesh.n, s.mesh.p, flush=True)
        ref = oracle_solution(s.kind, s.mesh, bh[r])
        d = xh[r] - ref
        if s.kind == ops.QF_DIFFUSION:
            d = d - d.mean()
        assert np.linalg.norm(d) / np.linalg.norm(ref) < 1e-8, (i, s.kind, s.mesh.n, s.mesh.p)


def every_group_systems():
    systems = []
    for p in range(2, 8):
        for kind in (ops.QF_MASS, ops.QF_DIFFUSION):
            for n in (1, 2 if p >= 5 else 3):
                systems.append(System(kind, hex_mesh(n, p, seed=97 * p + 13 * n + kind)))
    return systems


@pytest.mark.skip(reason="persistent small-system CG switched off: its first GPU run ended in a GPU memory fault whose cause has not been found")
def test_small_batch_every_group_vs_oracle(cuda):
    """The persistent small-system launch (each workgroup solves whole systems off a work list) on every
    (mass or diffusion, p = 2..7) group: each solution against the oracle's dense solve at 1e-8 (diffusion modulo
    constants), true residuals at the tolerance, and a second solve gives identical bits."""
    systems = every_group_systems()
    bcg = BatchCG(systems, small_dofs=10 ** 9)
    assert all(bcg.is_small(i) for i in range(len(systems)))
    b = batch_rhs(bcg, seed=5)
    xs = []
    for _ in range(2):
        x = torch.full_like(b, 3.0)  # the solver starts from x = 0 whatever x holds
        st = bcg.solve(b, x, max_iter=5000, rtol=1e-12)
        torch.cuda.synchronize()
        xs.append(x.cpu().numpy())
        assert np.all(st["rel_residual"] <= 1e-11), st["rel_residual"]
        assert np.all(st["iterations"] > 0) and np.all(np.isfinite(st["done_ms"]))
    assert np.array_equal(xs[0], xs[1])
    bh = b.cpu().numpy()
    for i, s in enumerate(systems):
        r = bcg.range(i)
        print("small batch vs oracle", i, s.kind, s.mesh.n, s.mesh.p, flush=True)
        ref = oracle_solution(s.kind, s.mesh, bh[r])
        d = xs[0][r] - ref
        if s.kind == ops.QF_DIFFUSION:
            d = d - d.mean()
        assert np.linalg.norm(d) / np.linalg.norm(ref) < 1e-8, (i, s.kind, s.mesh.n, s.mesh.p)


@pytest.mark.skip(reason="persistent small-system CG switched off: its first GPU run ended in a GPU memory fault whose cause has not been found")
def test_hybrid_batch_matches_group_batched(cuda):
    """Small systems on the persistent launch while the others run the group-batched CG on another stream:
    every system converges, and each solution equals the group-batched solution of the same system."""
    systems = random_systems(20, seed=9, max_n=5, p_range=(2, 5))
    sizes = sorted(s.mesh.num_nodes for s in systems)
    hyb = BatchCG(systems, small_dofs=sizes[len(sizes) // 2])
    ref = BatchCG(systems)
    assert 0 < hyb.num_large < len(systems)
    b_ref = batch_rhs(ref, seed=6)
    b = torch.empty_like(b_ref)
    for i in range(len(systems)):  # the same right-hand sides in the hybrid's vector order
        b[hyb.range(i)] = b_ref[ref.range(i)]
    x, x_ref = torch.zeros_like(b), torch.zeros_like(b)
    st = hyb.solve(b, x, max_iter=5000, rtol=1e-11, check_every=10)
    ref.solve(b_ref, x_ref, max_iter=5000, rtol=1e-11, check_every=10)
    assert np.all(st["rel_residual"] <= 1e-11), st["rel_residual"]
    for i, s in enumerate(systems):
        d = x[hyb.range(i)] - x_ref[ref.range(i)]
        if s.kind == ops.QF_DIFFUSION:
            d = d - d.mean()
        assert (torch.linalg.norm(d) / torch.linalg.norm(x_ref[ref.range(i)])).item() < 1e-8, (i, hyb.is_small(i))
