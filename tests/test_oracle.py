"""Pins the CPU oracle (oracle/ceed_oracle.c) before it is trusted as the checker.

1. Golden vectors: outputs of the reference library (/cpu/self/ref/serial,
   built from the reference sources by oracle/Makefile) stored in
   tests/golden/*.npz -- the oracle must reproduce them bit for bit.
2. The reference's own known answers: tests/output/t300-basis.out (1D bases at
   8 digits) and t200 (restriction).
3. Mathematics the reference does not pin: quadrature exactness, polynomial
   reproduction, brute-force Kronecker products, operator invariants
   (volume, zero row sums, energy of linear fields, symmetry), and finite
   differences for the kappa gradient (no reference counterpart).
4. When oracle/_ref is built (build container only), a live comparison at the
   full mass configuration (16^3 elements, p = 3, Q = 5).
"""

import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from libceed_amd.mesh import hex_mesh, node_coordinates

GOLD = Path(__file__).resolve().parent / "golden"


def _g(name):
    return np.load(GOLD / f"golden_{name}.npz", allow_pickle=False)


# ---------------------------------------------------------------- golden ----
def test_golden_basis_bitexact():
    g = _g("basis")
    for k in g.files:
        P, Q, m = (int(s[1:]) for s in k.split("_"))
        B, G, x, w = oracle.lagrange_1d(P, Q, m)
        D = oracle.collocated_grad(B, G).ravel() if Q >= P else np.zeros(Q * Q)
        got = np.concatenate([B.ravel(), G.ravel(), x, w, D])
        assert np.array_equal(got, g[k]), k


def test_golden_basis_apply_bitexact():
    g = _g("bapply")
    n = 0
    for k in g.files:
        if not k.endswith("_u"):
            continue
        key = k[:-2]
        dim, nc, P, Q, tm, em = (int(s[1:]) for s in key.split("_"))
        B, G, _, w = oracle.lagrange_1d(P, Q, oracle.GAUSS)
        v = oracle.basis_apply(dim, nc, B, G, w, 1, tm, em, None if em == 16 else g[k])
        assert np.array_equal(v, g[key + "_v"]), key
        n += 1
    assert n > 100


def test_golden_restriction_bitexact():
    g = _g("rstr")
    for name in sorted({k.rsplit("_", 1)[0] for k in g.files}):
        E, es, nc, cs, ls = (int(x) for x in g[name + "_meta"])
        e = np.zeros(E * es * nc)
        oracle.restriction_offset(E, es, nc, cs, g[name + "_off"], 0, g[name + "_u"], e)
        v = g[name + "_v0"].copy()
        oracle.restriction_offset(E, es, nc, cs, g[name + "_off"], 1, e, v)
        assert np.array_equal(e, g[name + "_e"]), name
        assert np.array_equal(v, g[name + "_v"]), name


def test_golden_operator_bitexact():
    g = _g("op")
    for k in sorted(g.files):
        if not k.endswith("_meta"):
            continue
        key = k[:-5]
        kind, nc, n, p, Q, seed = (int(x) for x in g[k])
        m = hex_mesh(n, p, seed=seed)
        qd = oracle.hex_build_qdata(kind, m, Q)
        gq = g[key + "_qdata"]
        assert np.array_equal(qd[:gq.size], gq), key
        v = oracle.hex_apply(kind, nc, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qd, g[key + "_u"])
        assert np.array_equal(v, g[key + "_v"]), key


# ----------------------------------------------- reference known answers ----
def test_t300_known_answer():
    """tests/output/t300-basis.out: P = Q = 4 on GLL and Gauss points, 8 digits."""
    B, G, x, w = oracle.lagrange_1d(4, 4, oracle.GAUSS_LOBATTO)
    assert np.allclose(x, [-1, -0.44721360, 0.44721360, 1], atol=5e-9)
    assert np.allclose(w, [0.16666667, 0.83333333, 0.83333333, 0.16666667], atol=5e-9)
    assert np.allclose(B, np.eye(4), atol=1e-15)
    assert np.allclose(G[0], [-3.0, 4.04508497, -1.54508497, 0.5], atol=5e-9)
    assert np.allclose(G[1], [-0.80901699, 0.0, 1.11803399, -0.30901699], atol=5e-9)
    B, G, x, w = oracle.lagrange_1d(4, 4, oracle.GAUSS)
    assert np.allclose(x, [-0.86113631, -0.33998104, 0.33998104, 0.86113631], atol=5e-9)
    assert np.allclose(w, [0.34785485, 0.65214515, 0.65214515, 0.34785485], atol=5e-9)
    assert np.allclose(B[0], [0.62994317, 0.47255875, -0.14950343, 0.04700152], atol=5e-9)
    assert np.allclose(G[1], [-0.51670214, -0.48795249, 1.33790510, -0.33325047], atol=5e-9)


def test_t200_known_answer():
    """t200-elemrestriction.c:36: y[i] = 10 + (i+1)/2 for 3 two-node elements."""
    off = np.array([0, 1, 1, 2, 2, 3], dtype=np.int32)
    x = 10.0 + np.arange(4)
    y = np.zeros(6)
    oracle.restriction_offset(3, 2, 1, 1, off, 0, x, y)
    assert np.array_equal(y, 10 + (np.arange(6) + 1) // 2)
    # t209-style multiplicity: E^T of ones counts element memberships
    mult = np.zeros(4)
    oracle.restriction_offset(3, 2, 1, 1, off, 1, np.ones(6), mult)
    assert np.array_equal(mult, [1, 2, 2, 1])


# ----------------------------------------------------------- mathematics ----
@pytest.mark.parametrize("Q", range(1, 12))
def test_gauss_exactness(Q):
    x, w = oracle.gauss(Q)
    for deg in range(2 * Q):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.dot(w, x ** deg) - exact) < 2e-14
    if Q >= 2:
        x, w = oracle.lobatto(Q)
        assert x[0] == -1 and x[-1] == 1
        for deg in range(2 * Q - 2):
            exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
            assert abs(np.dot(w, x ** deg) - exact) < 2e-14


@pytest.mark.parametrize("P", range(2, 11))
def test_lagrange_properties(P):
    for Q in (P - 1, P, P + 1, P + 2):
        if Q < 1:
            continue
        B, G, x, w = oracle.lagrange_1d(P, Q, oracle.GAUSS)
        assert np.allclose(B.sum(1), 1, atol=1e-13)
        assert np.allclose(G.sum(1), 0, atol=1e-11)
        nodes, _ = oracle.lobatto(P)
        # polynomial reproduction: interpolate x^k from the nodes
        for k in range(P):
            assert np.allclose(B @ nodes ** k, x ** k, atol=1e-12)
            assert np.allclose(G @ nodes ** k, k * x ** max(k - 1, 0) if k else 0 * x, atol=1e-10)
        if Q >= P:
            D = oracle.collocated_grad(B, G)
            assert np.allclose(D @ B, G, atol=1e-11)
            Bp = oracle.pseudoinverse(B)
            assert np.allclose(Bp @ B, np.eye(P), atol=1e-12)


def _kron3(a, b, c):
    return np.kron(np.kron(a, b), c)


@pytest.mark.parametrize("P,Q", [(2, 3), (3, 4), (4, 4), (4, 3), (5, 6)])
def test_basis_apply_bruteforce_3d(P, Q):
    """Sum factorisation equals explicit Kronecker products (x fastest)."""
    rng = np.random.default_rng(P * 10 + Q)
    B, G, _, w = oracle.lagrange_1d(P, Q, oracle.GAUSS)
    u = rng.uniform(-1, 1, P ** 3)
    I3 = _kron3(B, B, B)                   # [z][y][x] ordering: kron(Bz, By, Bx)
    assert np.allclose(oracle.basis_apply(3, 1, B, G, w, 1, 0, oracle.EVAL_INTERP, u), I3 @ u, atol=1e-13)
    gx, gy, gz = _kron3(B, B, G), _kron3(B, G, B), _kron3(G, B, B)
    g = oracle.basis_apply(3, 1, B, G, w, 1, 0, oracle.EVAL_GRAD, u)
    assert np.allclose(g, np.concatenate([gx @ u, gy @ u, gz @ u]), atol=1e-11)
    r = rng.uniform(-1, 1, 3 * Q ** 3)
    gt = oracle.basis_apply(3, 1, B, G, w, 1, 1, oracle.EVAL_GRAD, r)
    assert np.allclose(gt, gx.T @ r[:Q ** 3] + gy.T @ r[Q ** 3:2 * Q ** 3] + gz.T @ r[2 * Q ** 3:], atol=1e-11)
    wt = oracle.basis_apply(3, 1, B, G, w, 1, 0, oracle.EVAL_WEIGHT, None)
    assert np.allclose(wt, _kron3(w, w, w), atol=1e-15)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 6])
def test_operator_invariants(p):
    n, Q = 3, p + 2
    m = hex_mesh(n, p, seed=5)
    qm = oracle.hex_build_qdata(0, m, Q)
    qd = oracle.hex_build_qdata(1, m, Q)
    one = np.ones(m.num_nodes)
    mass1 = oracle.hex_apply(0, 1, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qm, one)
    assert abs(mass1.sum() - 1.0) < 1e-13                                  # volume (t501, ex1)
    diff1 = oracle.hex_apply(1, 1, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qd, one)
    assert np.abs(diff1).max() < 1e-12                                    # A 1 = 0 (t522)
    _, _, _, _ = oracle.lagrange_1d(p + 1, Q)
    gll, _ = oracle.lobatto(p + 1)
    X = node_coordinates(m, gll)
    u = X.sum(0)                                                          # x + y + z is in the space
    Au = oracle.hex_apply(1, 1, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qd, u)
    assert abs(u @ Au - 3.0) < 1e-11                                      # int |grad u|^2 = 3 |Omega|
    interior = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=0)
    assert np.abs(Au[interior]).max() < 1e-11                             # ex2: interior rows vanish
    rng = np.random.default_rng(p)
    a, b = rng.uniform(-1, 1, (2, m.num_nodes))
    Ab = oracle.hex_apply(1, 1, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qd, b)
    Aa = oracle.hex_apply(1, 1, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qd, a)
    assert abs(a @ Ab - b @ Aa) < 1e-11 * abs(a @ Ab)                     # symmetry
    assert a @ Aa > 0


def test_kappa_operator_and_gradient():
    """kappa = 1 reproduces Poisson3DApply bitwise; dk = -d(lam^T A(kappa) u)/dkappa by finite differences."""
    n, p, Q = 2, 2, 3
    m = hex_mesh(n, p, seed=9)
    qd = oracle.hex_build_qdata(1, m, Q)
    rng = np.random.default_rng(1)
    u, lam = rng.uniform(-1, 1, (2, m.num_nodes))
    nq = m.num_elem * Q ** 3
    plain = oracle.hex_apply(1, 1, m.num_elem, p + 1, Q, m.offsets, m.num_nodes, qd, u)
    k1 = oracle.hex_apply_kappa(m.num_elem, p + 1, Q, m.offsets, qd, np.ones(nq), u)
    assert np.array_equal(plain, k1)
    kappa = np.exp(rng.uniform(-0.5, 0.5, nq))
    dk = oracle.hex_kappa_gradient(m.num_elem, p + 1, Q, m.offsets, qd, u, lam)
    h = 1e-6
    for q in rng.choice(nq, 12, replace=False):
        kp, km = kappa.copy(), kappa.copy()
        kp[q] += h
        km[q] -= h
        fp = lam @ oracle.hex_apply_kappa(m.num_elem, p + 1, Q, m.offsets, qd, kp, u)
        fm = lam @ oracle.hex_apply_kappa(m.num_elem, p + 1, Q, m.offsets, qd, km, u)
        assert abs(-(fp - fm) / (2 * h) - dk[q]) < 1e-7 * max(1.0, abs(dk[q]))


# --------------------------------------------------- live reference build ----
REF_DRIVER = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "gen_ref_vectors"


@pytest.mark.skipif(not REF_DRIVER.exists(), reason="oracle/_ref not built (reference sources absent)")
def test_live_reference_mass_config(tmp_path):
    """Full mass configuration (16^3 elements, p = 3, Q = 5, 117,649 DoFs) against the reference build."""
    m = hex_mesh(16, 3, seed=1)
    u = np.random.default_rng(3).uniform(-1, 1, m.num_nodes)
    pre = tmp_path / "op"
    m.offsets.tofile(str(pre) + ".off")
    m.mesh_offsets.tofile(str(pre) + ".moff")
    m.coords.tofile(str(pre) + ".coords")
    u.tofile(str(pre) + ".u")
    subprocess.run([str(REF_DRIVER), "op", "0", "1", str(m.num_elem), "4", "5", str(m.num_vertices), str(m.num_nodes), str(pre)],
                   check=True)
    qd = oracle.hex_build_qdata(0, m, 5)
    v = oracle.hex_apply(0, 1, m.num_elem, 4, 5, m.offsets, m.num_nodes, qd, u)
    assert np.array_equal(qd, np.fromfile(str(pre) + ".qdata"))
    assert np.array_equal(v, np.fromfile(str(pre) + ".v"))


def test_pointwise_qfunctions_closed_forms():
    """The oracle's gallery QFunctions against their mathematics at random points: Mass3DBuild = w det J,
    Poisson3DBuild = w det J J^-1 J^-T (J = dx/dX, stored transposed) in Voigt order {xx, yy, zz, yz, xz, xy}, the apply functions as
    the products they define (a wrong Voigt index or a transposed J fails here)."""
    rng = np.random.default_rng(11)
    Q = 7
    J = np.eye(3)[:, :, None] + 0.4 * rng.uniform(-1, 1, (3, 3, Q))
    w = rng.uniform(0.1, 1, Q)
    Jl = [J[:, :, i] for i in range(Q)]
    (qm,) = oracle.qf_apply("Mass3DBuild", Q, [J.reshape(-1), w])
    assert np.allclose(qm, [np.linalg.det(a) * wi for a, wi in zip(Jl, w)], rtol=1e-13, atol=0)
    (qp,) = oracle.qf_apply("Poisson3DBuild", Q, [J.reshape(-1), w])
    voigt = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    for i, (a, wi) in enumerate(zip(Jl, w)):
        inv = np.linalg.inv(a)  # a[d][c] = dx_c/dX_d (GRAD layout [dir][comp]): the Jacobian is a^T
        S = wi * np.linalg.det(a) * inv.T @ inv
        assert np.allclose(qp.reshape(6, Q)[:, i], [S[r, c] for r, c in voigt], rtol=1e-12, atol=1e-14)
    # applies: S from the Voigt entries, ug stored [dir][Q] (scalar) or [dir][comp][Q] (vector)
    S = np.zeros((Q, 3, 3))
    for k, (r, c) in enumerate(voigt):
        S[:, r, c] = S[:, c, r] = qp.reshape(6, Q)[k]
    ug = rng.uniform(-1, 1, (3, Q))
    (vg,) = oracle.qf_apply("Poisson3DApply", Q, [ug.reshape(-1), qp])
    assert np.allclose(vg.reshape(3, Q), np.einsum("qij,iq->jq", S, ug), rtol=1e-13, atol=1e-15)
    ug3 = rng.uniform(-1, 1, (3, 3, Q))
    (vg3,) = oracle.qf_apply("Vector3Poisson3DApply", Q, [ug3.reshape(-1), qp])
    assert np.allclose(vg3.reshape(3, 3, Q), np.einsum("qij,icq->jcq", S, ug3), rtol=1e-13, atol=1e-15)
    u, u3 = rng.uniform(-1, 1, Q), rng.uniform(-1, 1, (3, Q))
    assert np.array_equal(oracle.qf_apply("MassApply", Q, [u, qm])[0], u * qm)
    assert np.array_equal(oracle.qf_apply("Vector3MassApply", Q, [u3.reshape(-1), qm])[0], (u3 * qm).reshape(-1))
