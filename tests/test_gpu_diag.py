"""Diagonal assembly of the fused plans (Jacobi preconditioning input),
against the diagonal of the dense operator the oracle assembles column by
column; every E^T mode, scalar and vector, mass, diffusion and the
kappa-weighted diffusion, FP64 (and FP32 within 2e-6)."""

import numpy as np
import pytest
import torch

import oracle
from libceed_amd import ops
from libceed_amd.mesh import hex_mesh

pytestmark = pytest.mark.gpu


def oracle_diag(kind, nc, m, Q, qd, kappa=None):
    N = m.num_nodes
    d = np.zeros(nc * N)
    for j in range(nc * N):
        e = np.zeros(nc * N)
        e[j] = 1.0
        if kind == ops.QF_DIFFUSION_KAPPA:
            col = oracle.hex_apply_kappa(m.num_elem, m.P, Q, m.offsets, qd, kappa, e)
        else:
            col = oracle.hex_apply(kind, nc, m.num_elem, m.P, Q, m.offsets, N, qd, e)
        d[j] = col[j]
    return d


@pytest.mark.parametrize("kind,nc,p,Q", [(0, 1, 2, 3), (1, 1, 3, 4), (1, 3, 2, 3), (0, 3, 1, 2), (2, 1, 2, 3), (1, 1, 4, 5)])
def test_diagonal_vs_dense(cuda, kind, nc, p, Q):
    m = hex_mesh(2, p, seed=p + kind)
    qd = oracle.hex_build_qdata(0 if kind == 0 else 1, m, Q)
    kappa = np.exp(np.random.default_rng(1).uniform(-0.5, 0.5, m.num_elem * Q ** 3)) if kind == 2 else None
    ref = oracle_diag(kind, nc, m, Q, qd, kappa)
    for scatter in ("tmap", "color", "split", "atomic", "brick"):
        if scatter == "brick" and kind == 2:  # no brick E^T for the kappa-weighted plan
            continue
        op = ops.HexOperator.from_mesh(kind, m, Q=Q, num_comp=nc, scatter=scatter)
        if kind == 2:
            op.set_kappa(torch.from_numpy(kappa).cuda())
        d = torch.full((nc * m.num_nodes,), 9.0, dtype=torch.float64, device="cuda")
        op.assemble_diagonal(d)
        got = d.cpu().numpy()
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12, scatter
        op.assemble_diagonal(d, add=True)
        assert np.linalg.norm(d.cpu().numpy() - 2 * ref) / np.linalg.norm(ref) < 1e-12, scatter
    op32 = ops.HexOperator.from_mesh(kind, m, Q=Q, num_comp=nc, dtype=torch.float32)
    if kind == 2:
        op32.set_kappa(torch.from_numpy(kappa).cuda())
    d32 = torch.zeros(nc * m.num_nodes, dtype=torch.float32, device="cuda")
    op32.assemble_diagonal(d32)
    assert np.linalg.norm(d32.double().cpu().numpy() - ref) / np.linalg.norm(ref) < 2e-6
