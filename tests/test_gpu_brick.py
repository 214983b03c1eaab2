"""GPU parity of the brick E^T (CEED_MI355_PLAN_BRICK) against the oracle:
a workgroup per brick of face-sharing elements, shared nodes summed in LDS,
brick-surface nodes finished by an ordered gather of partial sums.

Bar: FP64 within 1e-12 relative L2 of the oracle (the brick sums nodes in a
different order than the reference's ascending (elem, comp, node) order);
FP32 within 2e-5; bitwise repeatable; identical bits with and without the
next-element prefetch (same arithmetic, same order)."""

import numpy as np
import pytest
import torch

import oracle
from libceed_amd import ops
from libceed_amd.mesh import hex_mesh

pytestmark = pytest.mark.gpu
RTOL64 = 1e-12
RTOL32 = 2e-5


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def host(t):
    torch.cuda.synchronize()
    return t.double().cpu().numpy()


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def reference(kind, nc, m, Q, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1, 1, nc * m.num_nodes)
    qd = oracle.hex_build_qdata(kind, m, Q)
    return u, qd, oracle.hex_apply(kind, nc, m.num_elem, m.P, Q, m.offsets, m.num_nodes, qd, u)


# (kind, num_comp, elements per direction, p, Q): even and odd element counts (odd ones leave partial
# bricks), mass and diffusion, scalar and 3 components, p = 1..8
CASES = [(1, 1, 4, 4, 6), (1, 1, 5, 4, 6), (0, 1, 4, 3, 5), (1, 3, 3, 4, 6), (1, 1, 3, 6, 8), (1, 1, 3, 5, 7), (1, 1, 2, 8, 10),
         (1, 1, 3, 7, 9), (0, 3, 3, 2, 3), (1, 1, 5, 1, 2), (0, 1, 3, 6, 7), (1, 1, 4, 4, 5)]


@pytest.mark.parametrize("kind,nc,n,p,Q", CASES)
def test_brick_vs_oracle(cuda, kind, nc, n, p, Q):
    print("brick vs oracle", kind, nc, n, p, Q, flush=True)
    m = hex_mesh(n, p, seed=n + 7 * p)
    u, qd, v_ref = reference(kind, nc, m, Q, seed=p + nc)
    op = ops.HexOperator.from_mesh(kind, m, Q=Q, num_comp=nc, scatter="brick")
    assert op.scatter_mode() == "brick"
    assert rel_l2(host(op.get_qdata()), qd) < RTOL64
    v = torch.full_like(dev(u), 7.0)  # Apply overwrites
    op.apply(dev(u), v)
    assert rel_l2(host(v), v_ref) < RTOL64
    # geometric factors handed over in the CEED_STRIDES_BACKEND layout, and ApplyAdd
    op2 = ops.HexOperator.from_mesh(kind, m, Q=Q, num_comp=nc, build=False, scatter="brick")
    op2.set_qdata(dev(qd))
    w0 = np.random.default_rng(3).uniform(-1, 1, u.size)
    w = dev(w0)
    op2.apply_add(dev(u), w)
    assert rel_l2(host(w), w0 + v_ref) < RTOL64
    # bitwise repeatable
    v2 = torch.zeros_like(v)
    op.apply(dev(u), v2)
    assert torch.equal(v, v2)


@pytest.mark.parametrize("shape", ["1,1,1", "2,1,1", "1,2,1", "2,2,1", "1,2,2", "2,2,2"])
def test_brick_shapes(cuda, monkeypatch, shape):
    """Every brick shape (slots per direction), on a mesh with an odd element count: oracle parity."""
    monkeypatch.setenv("CEED_MI355_BRICK", shape)
    m = hex_mesh(5, 3, seed=2)
    u, qd, v_ref = reference(1, 1, m, 5, seed=9)
    op = ops.HexOperator.from_mesh(ops.QF_DIFFUSION, m, Q=5, scatter="brick")
    v = torch.zeros_like(dev(u))
    op.apply(dev(u), v)
    assert rel_l2(host(v), v_ref) < RTOL64, shape


@pytest.mark.parametrize("kind,nc,n,p", [(1, 1, 2, 1), (1, 1, 2, 2), (0, 1, 2, 3), (1, 1, 2, 4), (1, 3, 2, 4), (1, 1, 2, 5), (1, 1, 2, 6),
                                          (1, 1, 1, 7), (1, 1, 1, 8), (0, 1, 1, 8)])
def test_brick_single_brick_bitwise(cuda, monkeypatch, kind, nc, n, p):
    """A mesh that is one brick (2 x 2 x 2 elements, or one element at p = 7, 8): every node's sum runs over
    the brick's slots in the mesh's element order, which is the transpose map's ascending element order, and
    the per-element arithmetic is the line kernel's -- so the two E^T modes give identical bits."""
    if n == 2:
        monkeypatch.setenv("CEED_MI355_BRICK", "2,2,2")
    m = hex_mesh(n, p, seed=3 + p)
    u = dev(np.random.default_rng(p).uniform(-1, 1, nc * m.num_nodes))
    out = []
    for scatter in ("brick", "tmap"):
        op = ops.HexOperator.from_mesh(kind, m, Q=p + 2, num_comp=nc, scatter=scatter)
        v = torch.zeros_like(u)
        op.apply(u, v)
        out.append(v)
    assert torch.equal(out[0], out[1])


@pytest.mark.parametrize("kind,p", [(1, 4), (1, 6), (0, 4)])
def test_brick_fp32(cuda, kind, p):
    m = hex_mesh(4, p, seed=11)
    Q = p + 2
    u, qd, v_ref = reference(kind, 1, m, Q, seed=0)
    op = ops.HexOperator.from_mesh(kind, m, Q=Q, dtype=torch.float32, scatter="brick")
    v = torch.zeros(m.num_nodes, dtype=torch.float32, device="cuda")
    op.apply(dev(u, torch.float32), v)
    assert rel_l2(host(v), v_ref) < RTOL32


def test_brick_unsupported_kind(cuda):
    """The kappa-weighted plan has no brick E^T: asking for it by flag is an error."""
    m = hex_mesh(2, 2, seed=0)
    with pytest.raises(Exception, match="brick"):
        ops.HexOperator.from_mesh(ops.QF_DIFFUSION_KAPPA, m, Q=3, scatter="brick")
