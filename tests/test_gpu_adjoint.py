"""Variable-coefficient diffusion A(kappa) u = -div(kappa grad u) and its fused
adjoint kernel: v = A(kappa)^T lambda and dkappa = grad(lambda) . S grad(u) at
the quadrature points, checked against the oracle (the CPU transpose apply and
the oracle's kappa gradient) and, through torch.autograd, against central
finite differences of L = 1/2 ||A(kappa) u - f||^2 (h = 1e-6, random directions).
"""

import numpy as np
import pytest
import torch

import oracle
from libceed_amd import ops
from libceed_amd.autograd import kappa_diffusion
from libceed_amd.mesh import hex_mesh, quadrature_points, smooth_random_field

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def problem(cuda):
    m = hex_mesh(3, 4, seed=21)
    P, Q = 5, 5
    qd = oracle.hex_build_qdata(1, m, Q)
    kappa = np.exp(smooth_random_field(quadrature_points(m, oracle.gauss(Q)[0]), seed=5))
    return m, P, Q, qd, kappa


@pytest.mark.parametrize("layout", [ops.LAYOUT_LINES, ops.LAYOUT_COLUMNS, ops.LAYOUT_POINTS])
@pytest.mark.parametrize("scatter", ["tmap", "color", "split", "atomic"])
def test_forward_and_adjoint_vs_oracle(problem, scatter, layout):
    m, P, Q, qd, kappa = problem
    op = ops.HexOperator.from_mesh(ops.QF_DIFFUSION_KAPPA, m, Q=Q, scatter=scatter)
    op.set_layout(layout)
    rng = np.random.default_rng(4)
    u, lam = rng.uniform(-1, 1, (2, m.num_nodes))
    kd, ud, ld = (torch.from_numpy(a).cuda() for a in (kappa, u, lam))
    op.set_kappa(kd)
    v = torch.zeros_like(ud)
    op.apply(ud, v)
    assert rel_l2(v.cpu().numpy(), oracle.hex_apply_kappa(m.num_elem, P, Q, m.offsets, qd, kappa, u)) < RTOL
    w = torch.full_like(ud, 5.0)  # overwritten
    dk = torch.zeros(m.num_elem * Q ** 3, dtype=torch.float64, device="cuda")
    op.apply_adjoint(ud, ld, w, dk)
    # the CPU reference transpose apply (A is symmetric: the same operator applied to lambda)
    assert rel_l2(w.cpu().numpy(), oracle.hex_apply_kappa(m.num_elem, P, Q, m.offsets, qd, kappa, lam)) < RTOL
    # the oracle's gradient is -d(lam^T A u)/dkappa
    assert rel_l2(dk.cpu().numpy(), -oracle.hex_kappa_gradient(m.num_elem, P, Q, m.offsets, qd, u, lam)) < RTOL


def test_autograd_finite_differences(problem):
    m, P, Q, qd, kappa = problem
    op = ops.HexOperator.from_mesh(ops.QF_DIFFUSION_KAPPA, m, Q=Q)
    rng = np.random.default_rng(9)
    u0 = torch.from_numpy(rng.uniform(-1, 1, m.num_nodes)).cuda()
    f = torch.from_numpy(rng.uniform(-1, 1, m.num_nodes)).cuda()
    k0 = torch.from_numpy(kappa).cuda()

    def loss(u, k):
        r = kappa_diffusion(u, k, op) - f
        return 0.5 * torch.dot(r, r)

    u = u0.clone().requires_grad_(True)
    k = k0.clone().requires_grad_(True)
    loss(u, k).backward()
    gu, gk = u.grad.detach(), k.grad.detach()
    h = 1e-6
    with torch.no_grad():
        for i in range(64):  # kappa directions
            d = torch.from_numpy(np.random.default_rng(100 + i).standard_normal(k0.numel())).cuda()
            fd = (loss(u0, k0 + h * d) - loss(u0, k0 - h * d)).item() / (2 * h)
            an = torch.dot(gk, d).item()
            assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3 * torch.linalg.norm(gk).item()), (i, fd, an)
        for i in range(16):  # u directions
            d = torch.from_numpy(np.random.default_rng(200 + i).standard_normal(u0.numel())).cuda()
            fd = (loss(u0 + h * d, k0) - loss(u0 - h * d, k0)).item() / (2 * h)
            an = torch.dot(gu, d).item()
            assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3 * torch.linalg.norm(gu).item()), (i, fd, an)


def test_autograd_matches_oracle_gradients(problem):
    """dL/du = A^T (A u - f) and dL/dkappa from the oracle's transpose apply and kappa gradient."""
    m, P, Q, qd, kappa = problem
    op = ops.HexOperator.from_mesh(ops.QF_DIFFUSION_KAPPA, m, Q=Q)
    rng = np.random.default_rng(13)
    u, f = rng.uniform(-1, 1, (2, m.num_nodes))
    ut = torch.from_numpy(u).cuda().requires_grad_(True)
    kt = torch.from_numpy(kappa).cuda().requires_grad_(True)
    r = kappa_diffusion(ut, kt, op) - torch.from_numpy(f).cuda()
    (0.5 * torch.dot(r, r)).backward()
    r_ref = oracle.hex_apply_kappa(m.num_elem, P, Q, m.offsets, qd, kappa, u) - f
    assert rel_l2(ut.grad.cpu().numpy(), oracle.hex_apply_kappa(m.num_elem, P, Q, m.offsets, qd, kappa, r_ref)) < 1e-11
    assert rel_l2(kt.grad.cpu().numpy(), -oracle.hex_kappa_gradient(m.num_elem, P, Q, m.offsets, qd, u, r_ref)) < 1e-11
