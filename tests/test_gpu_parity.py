"""GPU parity of the HIP kernels against the oracle (/cpu/self/ref/serial
restatement) and the reference's golden vectors.

Bar: restriction E and E^T bit-exact; FP64 basis and operator within
1e-12 relative L2 (FP64 rounding differences from FMA and summation order);
FP32 operator within 2e-5 relative L2 of the FP64 oracle.
"""

from pathlib import Path

import numpy as np
import pytest
import torch

import oracle
from libceed_amd import ops
from libceed_amd.mesh import hex_mesh, node_coordinates

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
RTOL64 = 1e-12
RTOL32 = 2e-5


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def host(t):
    torch.cuda.synchronize()
    return t.double().cpu().numpy()


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# ------------------------------------------------------------ restriction --
def test_restriction_golden_bitexact(cuda):
    g = np.load(GOLD / "golden_rstr.npz", allow_pickle=False)
    for name in sorted({k.rsplit("_", 1)[0] for k in g.files}):
        E, es, nc, cs, ls = (int(x) for x in g[name + "_meta"])
        print("restriction golden", name, flush=True)
        r = ops.Restriction(E, es, nc, ls, offsets=g[name + "_off"], comp_stride=cs)
        e = torch.zeros(E * es * nc, dtype=torch.float64, device="cuda")
        r.apply(ops.NOTRANSPOSE, dev(g[name + "_u"]), e)
        assert np.array_equal(host(e), g[name + "_e"]), name
        v = dev(g[name + "_v0"])
        r.apply(ops.TRANSPOSE, e, v)
        assert np.array_equal(host(v), g[name + "_v"]), name


@pytest.mark.parametrize("nc", [1, 3])
def test_restriction_mesh_bitexact(cuda, nc):
    m = hex_mesh(6, 4, seed=2)
    N = m.num_nodes
    rng = np.random.default_rng(nc)
    u, v0 = rng.uniform(-1, 1, (2, nc * N))
    r = ops.Restriction(m.num_elem, m.P ** 3, nc, nc * N, offsets=m.offsets, comp_stride=N)
    e = torch.zeros(r.e_size, dtype=torch.float64, device="cuda")
    r.apply(ops.NOTRANSPOSE, dev(u), e)
    e_ref = np.zeros(r.e_size)
    oracle.restriction_offset(m.num_elem, m.P ** 3, nc, N, m.offsets, 0, u, e_ref)
    assert np.array_equal(host(e), e_ref)
    ev = rng.uniform(-1, 1, r.e_size)          # arbitrary E-vector, nonzero start vector
    v = dev(v0)
    r.apply(ops.TRANSPOSE, dev(ev), v)
    v_ref = v0.copy()
    oracle.restriction_offset(m.num_elem, m.P ** 3, nc, N, m.offsets, 1, ev, v_ref)
    assert np.array_equal(host(v), v_ref)


def test_restriction_strided_and_empty(cuda):
    E, es, nc = 5, 8, 2
    strides = np.array([1, es, es * nc], dtype=np.int32)
    rng = np.random.default_rng(4)
    u = rng.uniform(-1, 1, E * es * nc)
    r = ops.Restriction(E, es, nc, E * es * nc, strides=strides)
    e = torch.zeros(E * es * nc, dtype=torch.float64, device="cuda")
    r.apply(ops.NOTRANSPOSE, dev(u), e)
    e_ref = np.zeros_like(u)
    oracle.restriction_strided(E, es, nc, strides, 0, u, e_ref)
    assert np.array_equal(host(e), e_ref)
    v0 = rng.uniform(-1, 1, u.size)
    v = dev(v0)
    r.apply(ops.TRANSPOSE, e, v)
    oracle.restriction_strided(E, es, nc, strides, 1, e_ref, v0)
    assert np.array_equal(host(v), v0)
    empty = ops.Restriction(0, 4, 1, 3, offsets=np.zeros(0, dtype=np.int32))
    z = torch.zeros(3, dtype=torch.float64, device="cuda")
    empty.apply(ops.TRANSPOSE, torch.zeros(1, dtype=torch.float64, device="cuda"), z)
    assert not host(z).any()


# ------------------------------------------------------------------ basis --
def test_basis_golden(cuda):
    g = np.load(GOLD / "golden_bapply.npz", allow_pickle=False)
    n = 0
    for k in sorted(g.files):
        if not k.endswith("_u"):
            continue
        key = k[:-2]
        dim, nc, P, Q, tm, em = (int(s[1:]) for s in key.split("_"))
        b = ops.Basis(dim, nc, P, Q)
        ref = g[key + "_v"]
        v = torch.zeros(ref.size, dtype=torch.float64, device="cuda")
        b.apply(1, tm, em, None if em == 16 else dev(g[k]), v)
        err = rel_l2(host(v), ref)
        assert err < RTOL64, (key, err)
        n += 1
    assert n > 100


@pytest.mark.parametrize("dim,P,Q", [(3, 5, 6), (3, 4, 3), (2, 3, 3), (1, 6, 8)])
def test_basis_many_elements(cuda, dim, P, Q):
    """Element-major layout over several elements against the oracle one element at a time."""
    nc, E = 2, 7
    B, G, _, w = oracle.lagrange_1d(P, Q)
    rng = np.random.default_rng(dim * 100 + P)
    nn, nq = P ** dim, Q ** dim
    b = ops.Basis(dim, nc, P, Q)
    u = rng.uniform(-1, 1, E * nc * nn)
    for em in (ops.EVAL_INTERP, ops.EVAL_GRAD):
        qs = nc * nq * (dim if em == ops.EVAL_GRAD else 1)
        v = torch.zeros(E * qs, dtype=torch.float64, device="cuda")
        b.apply(E, ops.NOTRANSPOSE, em, dev(u), v)
        ref = np.concatenate([oracle.basis_apply(dim, nc, B, G, w, 1, 0, em, u[e * nc * nn:(e + 1) * nc * nn]) for e in range(E)])
        assert rel_l2(host(v), ref) < RTOL64
        r = rng.uniform(-1, 1, E * qs)
        vt0 = rng.uniform(-1, 1, E * nc * nn)
        vt = dev(vt0)
        b.apply(E, ops.TRANSPOSE, em, dev(r), vt, add=True)
        reft = vt0 + np.concatenate([oracle.basis_apply(dim, nc, B, G, w, 1, 1, em, r[e * qs:(e + 1) * qs]) for e in range(E)])
        assert rel_l2(host(vt), reft) < RTOL64


# --------------------------------------------------------- fused operator --
def golden_op_cases():
    g = np.load(GOLD / "golden_op.npz", allow_pickle=False)
    return sorted(k[:-5] for k in g.files if k.endswith("_meta"))


@pytest.mark.parametrize("key", golden_op_cases())
def test_operator_golden(cuda, key):
    g = np.load(GOLD / "golden_op.npz", allow_pickle=False)
    kind, nc, n, p, Q, seed = (int(x) for x in g[key + "_meta"])
    print("operator golden", key, flush=True)
    m = hex_mesh(n, p, seed=seed)
    op = ops.HexOperator.from_mesh(kind, m, Q=Q, num_comp=nc)
    qd = host(op.get_qdata())
    gq = g[key + "_qdata"]
    assert rel_l2(qd[:gq.size], gq) < RTOL64
    v = torch.zeros(nc * m.num_nodes, dtype=torch.float64, device="cuda")
    op.apply(dev(g[key + "_u"]), v)
    assert rel_l2(host(v), g[key + "_v"]) < RTOL64


@pytest.mark.parametrize("kind,nc,n,p,Q", [(1, 1, 5, 4, 6), (1, 3, 4, 3, 5), (0, 1, 5, 6, 8), (1, 1, 3, 8, 10), (1, 1, 4, 4, 5),
                                           (0, 3, 3, 2, 3), (1, 1, 6, 1, 2)])
def test_operator_vs_oracle(cuda, kind, nc, n, p, Q):
    m = hex_mesh(n, p, seed=n + p)
    rng = np.random.default_rng(p)
    u = rng.uniform(-1, 1, nc * m.num_nodes)
    qd_ref = oracle.hex_build_qdata(kind, m, Q)
    v_ref = oracle.hex_apply(kind, nc, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qd_ref, u)
    for det in ("tmap", "color", "atomic", "brick", "segment"):  # the E^T modes of the fused plan
        # geometric factors built on the GPU
        op = ops.HexOperator.from_mesh(kind, m, Q=Q, num_comp=nc, scatter=det)
        assert rel_l2(host(op.get_qdata()), qd_ref) < RTOL64
        v = torch.full_like(dev(u), 7.0)  # Apply overwrites
        op.apply(dev(u), v)
        assert rel_l2(host(v), v_ref) < RTOL64, det
        # geometric factors handed over in the CEED_STRIDES_BACKEND layout
        op2 = ops.HexOperator.from_mesh(kind, m, Q=Q, num_comp=nc, build=False, scatter=det)
        op2.set_qdata(dev(qd_ref))
        v2 = torch.zeros_like(v)
        op2.apply(dev(u), v2)
        assert rel_l2(host(v2), v_ref) < RTOL64, det
        # ApplyAdd sums into the output
        w0 = rng.uniform(-1, 1, u.size)
        w = dev(w0)
        op2.apply_add(dev(u), w)
        assert rel_l2(host(w), w0 + v_ref) < RTOL64, det


@pytest.mark.parametrize("kind,nc,n,p,Q,fp32", [(1, 1, 4, 4, 6, False), (1, 3, 3, 3, 5, False), (0, 1, 4, 6, 8, False), (1, 1, 2, 8, 10, False),
                                                 (0, 3, 3, 2, 3, False), (1, 1, 5, 1, 2, False), (1, 1, 3, 5, 6, True), (0, 1, 3, 4, 5, True),
                                                 (2, 1, 3, 4, 5, False), (2, 1, 2, 6, 8, False)])
def test_thread_layouts_identical(cuda, kind, nc, n, p, Q, fp32):
    """Every thread layout of the fused kernel (z-columns, 1D lines, and a thread per point where Q^3 <= 256), every
    admissible elements-per-workgroup count and every deterministic E^T mode: bitwise identical results across the
    launches (so the autotuner's choice never changes them) and oracle parity."""
    from libceed_amd.mesh import quadrature_points, smooth_random_field

    m = hex_mesh(n, p, seed=3 * n + p)
    rng = np.random.default_rng(p + nc)
    u = rng.uniform(-1, 1, nc * m.num_nodes)
    qd_ref = oracle.hex_build_qdata(min(kind, 1), m, Q)
    if kind == ops.QF_DIFFUSION_KAPPA:
        kappa = np.exp(smooth_random_field(quadrature_points(m, oracle.gauss(Q)[0]), seed=p))
        v_ref = oracle.hex_apply_kappa(m.num_elem, p + 1, Q, m.offsets, qd_ref, kappa, u)
    else:
        v_ref = oracle.hex_apply(kind, nc, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qd_ref, u)
    dtype = torch.float32 if fp32 else torch.float64
    for det in ("tmap", "color", "split") if nc == 1 else ("tmap", "color"):
        op = ops.HexOperator.from_mesh(kind, m, Q=Q, num_comp=nc, dtype=dtype, scatter=det)
        if kind == ops.QF_DIFFUSION_KAPPA:
            op.set_kappa(dev(kappa))
        outs = []
        for layout in (ops.LAYOUT_COLUMNS, ops.LAYOUT_LINES, ops.LAYOUT_POINTS):
            for epb in range(1, 256 // (Q ** 3 if layout == ops.LAYOUT_POINTS else Q * Q) + 1):
                try:
                    op.set_layout(layout, epb)
                except ops.CeedMi355Error:
                    continue  # does not fit the LDS budget
                print("layout", layout, "epb", epb, det, flush=True)
                v = torch.full((nc * m.num_nodes,), 3.0, dtype=dtype, device="cuda")
                op.apply(dev(u, dtype), v)
                outs.append((layout, epb, host(v)))
        assert {o[0] for o in outs} == ({ops.LAYOUT_COLUMNS, ops.LAYOUT_LINES} | ({ops.LAYOUT_POINTS} if Q ** 3 <= 256 else set()))
        for layout, epb, o in outs[1:]:
            assert np.array_equal(o, outs[0][2]), (det, layout, epb)
        assert rel_l2(outs[0][2], v_ref) < (RTOL32 if fp32 else RTOL64), det


@pytest.mark.parametrize("kind,nc,n,p,Q,fp32", [(1, 1, 5, 4, 6, False), (0, 1, 4, 3, 5, False), (1, 3, 3, 3, 5, False), (2, 1, 4, 4, 5, False),
                                                 (1, 1, 3, 6, 8, True), (1, 1, 2, 8, 10, False), (0, 3, 3, 2, 3, True), (1, 1, 6, 1, 2, False)])
@pytest.mark.parametrize("layout", [0, 1])
def test_split_scatter(cuda, kind, nc, n, p, Q, fp32, layout):
    """The multiplicity-split E^T (atomic adds onto a zeroed target for entries of at most two contributions,
    the ordered transpose map for the rest) sums every entry in the reference's order: Apply is bitwise equal to
    the transpose-map mode, repeats bit for bit, and ApplyAdd is repeatable and matches v0 + A u."""
    from libceed_amd.mesh import quadrature_points, smooth_random_field

    m = hex_mesh(n, p, seed=7 * n + p)
    rng = np.random.default_rng(3 * p + nc)
    u = rng.uniform(-1, 1, nc * m.num_nodes)
    w0 = rng.uniform(-1, 1, nc * m.num_nodes)
    qd_ref = oracle.hex_build_qdata(min(kind, 1), m, Q)
    kappa = None
    if kind == ops.QF_DIFFUSION_KAPPA:
        kappa = np.exp(smooth_random_field(quadrature_points(m, oracle.gauss(Q)[0]), seed=p))
        v_ref = oracle.hex_apply_kappa(m.num_elem, p + 1, Q, m.offsets, qd_ref, kappa, u)
    else:
        v_ref = oracle.hex_apply(kind, nc, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qd_ref, u)
    dtype = torch.float32 if fp32 else torch.float64
    tol = RTOL32 if fp32 else RTOL64
    out = {}
    for det in ("tmap", "split"):
        op = ops.HexOperator.from_mesh(kind, m, Q=Q, num_comp=nc, dtype=dtype, scatter=det)
        assert op.scatter_mode() == det
        if kappa is not None:
            op.set_kappa(dev(kappa))
        op.set_layout(layout, 0)
        print(det, "kind", kind, "nc", nc, "p", p, "layout", layout, flush=True)
        v = torch.full((nc * m.num_nodes,), 3.0, dtype=dtype, device="cuda")
        op.apply(dev(u, dtype), v)
        v2 = torch.full_like(v, -1.0)
        op.apply(dev(u, dtype), v2)
        w, w2 = dev(w0, dtype), dev(w0, dtype)
        op.apply_add(dev(u, dtype), w)
        op.apply_add(dev(u, dtype), w2)
        out[det] = (host(v), host(v2), host(w), host(w2))
    assert np.array_equal(out["split"][0], out["tmap"][0])
    assert np.array_equal(out["split"][1], out["split"][0])
    assert np.array_equal(out["split"][3], out["split"][2])
    assert rel_l2(out["split"][0], v_ref) < tol
    assert rel_l2(out["split"][2], w0 + v_ref) < tol


@pytest.mark.parametrize("kind,nc,n,p,Q,fp32", [(1, 1, 9, 4, 6, False), (0, 1, 6, 3, 5, False), (1, 3, 5, 3, 5, False), (1, 1, 4, 6, 8, True),
                                                 (1, 1, 3, 8, 10, False), (0, 3, 4, 2, 3, True), (1, 1, 8, 1, 2, False), (1, 1, 15, 4, 6, False)])
def test_segment_scatter(cuda, kind, nc, n, p, Q, fp32):
    """Segments of face-sharing elements along x, a workgroup each, the faces inside a segment summed in LDS and
    the segments coloured: Apply and ApplyAdd against the oracle, bitwise repeatable; the plan's one launch (line
    layout, a slot per element of the longest segment: every x-row of n elements cut into ceil(n / (256 / Q^2))
    balanced pieces), 8 colours on a structured mesh, the autotuner keeping that launch, other launches and the
    diagonal refused."""
    m = hex_mesh(n, p, seed=5 * n + p)
    rng = np.random.default_rng(7 * p + nc)
    u = rng.uniform(-1, 1, nc * m.num_nodes)
    qd_ref = oracle.hex_build_qdata(kind, m, Q)
    v_ref = oracle.hex_apply(kind, nc, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qd_ref, u)
    dtype = torch.float32 if fp32 else torch.float64
    tol = RTOL32 if fp32 else RTOL64
    op = ops.HexOperator.from_mesh(kind, m, Q=Q, num_comp=nc, dtype=dtype, scatter="segment")
    assert op.scatter_mode() == "segment"
    k = -(-n // (256 // (Q * Q)))
    assert op.launch() == (ops.LAYOUT_LINES, -(-n // k))
    assert len(op.colors()) == (8 if k > 1 else 4)
    v = torch.full((nc * m.num_nodes,), 3.0, dtype=dtype, device="cuda")
    op.apply(dev(u, dtype), v)
    first = v.clone()
    op.apply(dev(u, dtype), v)
    assert torch.equal(first, v)
    assert rel_l2(host(v), v_ref) < tol
    w0 = rng.uniform(-1, 1, u.size)
    w = dev(w0, dtype)
    op.apply_add(dev(u, dtype), w)
    assert rel_l2(host(w), w0 + v_ref) < tol
    launch = op.launch()
    op.autotune(2)
    assert op.launch() == launch
    with pytest.raises(ops.CeedMi355Error):
        op.set_layout(ops.LAYOUT_COLUMNS, 1)
    with pytest.raises(ops.CeedMi355Error):
        op.set_launch(launch[1] + 1)
    if nc == 1:
        with pytest.raises(ops.CeedMi355Error):
            op.assemble_diagonal(torch.zeros(m.num_nodes, dtype=dtype, device="cuda"))


def test_mass_config_repeated_apply(cuda):
    """Mass operator, 16^3 elements, p = 3, Q = 5: v = M u applied 100 times, bitwise repeatable, oracle parity."""
    m = hex_mesh(16, 3, seed=1)
    u = np.random.default_rng(3).uniform(-1, 1, m.num_nodes)
    qd = oracle.hex_build_qdata(0, m, 5)
    v_ref = oracle.hex_apply(0, 1, m.num_elem, 4, 5, m.offsets, m.num_nodes, qd, u)
    ud, v = dev(u), torch.zeros(m.num_nodes, dtype=torch.float64, device="cuda")
    for scatter in ("tmap", "color", "brick", "segment"):  # the atomic-free E^T modes: bitwise repeatable
        op = ops.HexOperator.from_mesh(ops.QF_MASS, m, Q=5, scatter=scatter)
        op.apply(ud, v)
        first = v.clone()
        for _ in range(99):
            op.apply(ud, v)
        assert torch.equal(first, v)
        assert rel_l2(host(v), v_ref) < RTOL64
    # hardware atomics: every one of 100 applies within the bar
    op = ops.HexOperator.from_mesh(ops.QF_MASS, m, Q=5, scatter="atomic")
    for _ in range(100):
        op.apply(ud, v)
        assert rel_l2(host(v), v_ref) < RTOL64


@pytest.mark.parametrize("kind,p", [(1, 4), (1, 6), (0, 4)])
def test_operator_fp32(cuda, kind, p):
    m = hex_mesh(4, p, seed=11)
    Q = p + 2
    u = np.random.default_rng(0).uniform(-1, 1, m.num_nodes)
    v_ref = oracle.hex_apply(kind, 1, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, oracle.hex_build_qdata(kind, m, Q), u)
    op = ops.HexOperator.from_mesh(kind, m, Q=Q, dtype=torch.float32)
    v = torch.zeros(m.num_nodes, dtype=torch.float32, device="cuda")
    op.apply(dev(u, torch.float32), v)
    assert rel_l2(host(v), v_ref) < RTOL32


@pytest.mark.parametrize("p", [2, 4, 7])
def test_operator_invariants_large(cuda, p):
    """Size-independent properties at ~0.5-1M DoFs: volume, A 1 = 0, int |grad(x+y+z)|^2 = 3, symmetry."""
    n = max(2, int(round((8e5 ** (1 / 3) - 1) / p)))
    m = hex_mesh(n, p, seed=21)
    Q = p + 2
    A = ops.HexOperator.from_mesh(ops.QF_DIFFUSION, m, Q=Q)
    M = ops.HexOperator.from_mesh(ops.QF_MASS, m, Q=Q)
    one = torch.ones(m.num_nodes, dtype=torch.float64, device="cuda")
    v = torch.empty_like(one)
    M.apply(one, v)
    assert abs(v.sum().item() - 1.0) < 1e-11
    A.apply(one, v)
    assert v.abs().max().item() < 1e-11
    gll, _ = oracle.lobatto(p + 1)
    u = dev(node_coordinates(m, gll).sum(0))
    A.apply(u, v)
    assert abs(torch.dot(u, v).item() - 3.0) < 1e-9
    g = torch.Generator(device="cuda").manual_seed(p)
    a = torch.rand(m.num_nodes, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    b = torch.rand(m.num_nodes, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    Aa, Ab = torch.empty_like(a), torch.empty_like(a)
    A.apply(a, Aa)
    A.apply(b, Ab)
    assert abs(torch.dot(a, Ab).item() - torch.dot(b, Aa).item()) < 1e-10 * abs(torch.dot(a, Ab).item())


@pytest.mark.parametrize("p,num_comp,mode", [(4, 1, "color"), (5, 1, "tmap"), (6, 1, "split"), (7, 1, "split"), (8, 1, "tmap"),
                                              (4, 3, "color"), (6, 3, "tmap"), (7, 3, "tmap")])
def test_default_mode_choice(cuda, p, num_comp, mode):
    """The default E^T of FP64 diffusion plans with more than 10M E-vector entries, by order and components (the
    multiplicity split only for one component at P = 7, 8; three components keep the transpose map there)."""
    from libceed_amd.mesh import mesh_size_for_dofs

    m = hex_mesh(mesh_size_for_dofs(p, 12e6 / num_comp), p, seed=3)
    assert m.num_elem * m.P ** 3 * num_comp >= 10_000_000
    A = ops.HexOperator.from_mesh(ops.QF_DIFFUSION, m, Q=p + 2, num_comp=num_comp, build=False)
    assert A.scatter_mode() == mode


@pytest.mark.parametrize("p,mode", [(4, "color"), (6, "split"), (7, "split")])
def test_default_modes_full_size(cuda, p, mode):
    """At ~8M DoFs the size-based default E^T is colours (p <= 4) or the multiplicity split (p = 6, 7): the
    split's Apply is bitwise equal to the transpose map's there as on small meshes, colours agree with it to
    rounding, and A 1 = 0 and symmetry hold for the default plan."""
    from libceed_amd.mesh import mesh_size_for_dofs

    m = hex_mesh(mesh_size_for_dofs(p, 8e6), p, seed=31)
    Q = p + 2
    A = ops.HexOperator.from_mesh(ops.QF_DIFFUSION, m, Q=Q)
    assert A.scatter_mode() == mode
    T = ops.HexOperator.from_mesh(ops.QF_DIFFUSION, m, Q=Q, scatter="tmap")
    g = torch.Generator(device="cuda").manual_seed(p)
    a = torch.rand(m.num_nodes, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    b = torch.rand(m.num_nodes, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    Aa, Ta, Ab = torch.empty_like(a), torch.empty_like(a), torch.empty_like(a)
    A.apply(a, Aa)
    T.apply(a, Ta)
    print("p", p, mode, "dofs", m.num_nodes, flush=True)
    if mode == "split":
        assert torch.equal(Aa, Ta)
    else:
        assert (torch.linalg.norm(Aa - Ta) / torch.linalg.norm(Ta)).item() < 1e-13
    A.apply(b, Ab)
    assert abs(torch.dot(a, Ab).item() - torch.dot(b, Aa).item()) < 1e-10 * abs(torch.dot(a, Ab).item())
    one = torch.ones_like(a)
    A.apply(one, Ab)
    assert Ab.abs().max().item() < 1e-10


@pytest.mark.parametrize("scatter", ["tmap", "color", "split", "atomic", "brick", "segment"])
@pytest.mark.parametrize("kind,nc,p", [(1, 1, 3), (0, 1, 4), (1, 3, 2)])
def test_operator_unstructured_numbering(cuda, scatter, kind, nc, p):
    """Nodes renumbered by a random permutation into an L-vector with unused entries, elements in random order
    (so colours, maps and multiplicities come from a numbering with no structure): Apply matches the oracle and
    leaves the unreached entries 0; ApplyAdd leaves them as they were."""
    m = hex_mesh(3, p, seed=13 + p)
    rng = np.random.default_rng(100 + p)
    P, Q, E = p + 1, p + 2, m.num_elem
    stride = m.num_nodes + 11                                    # per-component range with 11 unused entries
    perm = rng.permutation(stride)[: m.num_nodes].astype(np.int32)
    order = rng.permutation(E)
    offsets = perm[m.offsets.reshape(E, -1)[order]].reshape(-1)
    qc = 1 if kind == 0 else 6
    qd = oracle.hex_build_qdata(kind, m, Q).reshape(E, qc * Q ** 3)[order].reshape(-1)
    l_size = nc * stride
    u = rng.uniform(-1, 1, l_size)
    v_ref = oracle.hex_apply(kind, nc, E, P, Q, offsets, stride, qd, u)
    unused = np.ones(l_size, dtype=bool)
    for k in range(nc):
        unused[perm + k * stride] = False
    op = ops.HexOperator(kind, offsets, E, P, Q, l_size, num_comp=nc, comp_stride=stride, scatter=scatter)
    assert op.scatter_mode() == scatter
    op.set_qdata(dev(qd))
    v = torch.full((l_size,), 3.0, dtype=torch.float64, device="cuda")
    op.apply(dev(u), v)
    vh = host(v)
    assert np.all(vh[unused] == 0.0)
    assert rel_l2(vh, v_ref) < RTOL64
    w0 = rng.uniform(-1, 1, l_size)
    w = dev(w0)
    op.apply_add(dev(u), w)
    wh = host(w)
    assert np.array_equal(wh[unused], w0[unused])
    assert rel_l2(wh, w0 + v_ref) < RTOL64


@pytest.mark.parametrize("scatter", ["tmap", "color", "split", "atomic", "brick", "segment"])
def test_operator_no_elements(cuda, scatter):
    """A plan with no elements: Apply gives zeros, ApplyAdd leaves v unchanged."""
    op = ops.HexOperator(ops.QF_DIFFUSION, np.zeros(0, dtype=np.int32), 0, 3, 4, 10, scatter=scatter)
    v = torch.full((10,), 3.0, dtype=torch.float64, device="cuda")
    op.apply(torch.ones(10, dtype=torch.float64, device="cuda"), v)
    assert torch.count_nonzero(v).item() == 0
    w = torch.full((10,), 2.0, dtype=torch.float64, device="cuda")
    op.apply_add(torch.ones(10, dtype=torch.float64, device="cuda"), w)
    assert torch.all(w == 2.0).item()


_HEADLINE = {}


def headline_reference(p):
    """The ~8M-DoF diffusion workload (p=4: 50^3 elements, 8.12M DoFs; p=6: 33^3, 7.88M), Q = p + 2: mesh,
    seeded u, the oracle's q_data and the oracle's v = A u (about half a minute of host time, done once)."""
    if p not in _HEADLINE:
        from libceed_amd.mesh import mesh_size_for_dofs

        m = hex_mesh(mesh_size_for_dofs(p, 8e6), p, seed=41)
        Q = p + 2
        u = np.random.default_rng(p).uniform(-1, 1, m.num_nodes)
        qd = oracle.hex_build_qdata(1, m, Q)
        _HEADLINE[p] = (m, Q, u, qd, oracle.hex_apply(1, 1, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qd, u))
    return _HEADLINE[p]


@pytest.mark.parametrize("p,scatter", [(4, "auto"), (4, "brick"), (4, "segment"), (6, "auto"), (6, "brick")])
def test_headline_workload_vs_oracle(cuda, p, scatter):
    """The benchmarked FP64 apply at full size against the oracle at 1e-12: q_data built on the GPU, then
    Apply and ApplyAdd with the library's default E^T (colours at p=4, the multiplicity split at p=6) and
    with bricks and segments."""
    m, Q, u, qd, v_ref = headline_reference(p)
    print("headline vs oracle p", p, scatter, "dofs", m.num_nodes, flush=True)
    op = ops.HexOperator.from_mesh(ops.QF_DIFFUSION, m, Q=Q, scatter=scatter)
    assert rel_l2(host(op.get_qdata()), qd) < RTOL64
    v = torch.full((m.num_nodes,), 5.0, dtype=torch.float64, device="cuda")
    ud = dev(u)
    op.apply(ud, v)
    assert rel_l2(host(v), v_ref) < RTOL64
    op.apply_add(ud, v)
    assert rel_l2(host(v), 2.0 * v_ref) < RTOL64


def test_headline_workload_fp32_vs_oracle(cuda):
    """The FP32 operator (FP32 q_data, vectors and accumulation) at 8.12M DoFs, p=4, against the FP64 oracle."""
    m, Q, u, qd, v_ref = headline_reference(4)
    op = ops.HexOperator.from_mesh(ops.QF_DIFFUSION, m, Q=Q, dtype=torch.float32)
    v = torch.zeros(m.num_nodes, dtype=torch.float32, device="cuda")
    op.apply(dev(u, torch.float32), v)
    assert rel_l2(host(v), v_ref) < RTOL32


@pytest.mark.parametrize("p,n,mode", [(3, 38, "color"), (6, 22, "tmap")])
def test_vector_default_vs_oracle(cuda, p, n, mode):
    """The 3-component diffusion operator with the library's default E^T at a size where the size-based default
    applies (more than 10M E-vector entries; p=3 colours, p=6 the transpose map): q_data built on the GPU, then
    Apply and ApplyAdd against the oracle at 1e-12."""
    m = hex_mesh(n, p, seed=43)
    Q = p + 2
    assert m.num_elem * m.P ** 3 * 3 >= 10_000_000
    u = np.random.default_rng(p + 100).uniform(-1, 1, 3 * m.num_nodes)
    qd = oracle.hex_build_qdata(1, m, Q)
    v_ref = oracle.hex_apply(1, 3, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qd, u)
    print("vector vs oracle p", p, "dofs", 3 * m.num_nodes, flush=True)
    op = ops.HexOperator.from_mesh(ops.QF_DIFFUSION, m, Q=Q, num_comp=3)
    assert op.scatter_mode() == mode
    assert rel_l2(host(op.get_qdata()), qd) < RTOL64
    v = torch.full((3 * m.num_nodes,), 5.0, dtype=torch.float64, device="cuda")
    ud = dev(u)
    op.apply(ud, v)
    assert rel_l2(host(v), v_ref) < RTOL64
    op.apply_add(ud, v)
    assert rel_l2(host(v), 2.0 * v_ref) < RTOL64
