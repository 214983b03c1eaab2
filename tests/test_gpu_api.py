"""The libCEED C API on /gpu/hip/mi355, driven the way the reference's Python
tests drive libCEED (python/tests/test-1-vector.py, test-2-elemrestriction.py,
test-3-basis.py, test-5-operator.py) and examples/ceed/ex1-volume.c,
ex2-surface.c build their operators, checked against the oracle.

Bars: restriction bit-exact; FP64 basis, QFunction and operator results
within 1e-12 relative L2 of the oracle; vector reductions within 1e-14.
"""

from pathlib import Path

import numpy as np
import pytest

import oracle
from libceed_amd import ceed as libceed
from libceed_amd.mesh import hex_mesh, node_coordinates

pytestmark = pytest.mark.gpu
QF_FILE = Path(__file__).resolve().parent / "qfunctions" / "t-api.h"
RTOL = 1e-12


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ceed(cuda):
    return libceed.Ceed("/gpu/hip/mi355")


# ------------------------------------------------------------------ Ceed --
def test_ceed_resource(ceed):
    assert ceed.get_resource() == "/gpu/hip/mi355"
    assert ceed.get_preferred_memtype() == libceed.MEM_DEVICE
    assert ceed.is_deterministic()  # every E^T sums in a fixed order, without atomics
    x, w = ceed.gauss_quadrature(5)
    assert abs(w.sum() - 2.0) < 1e-14 and np.allclose(x, -x[::-1], atol=1e-15)


def test_ceed_device_id_and_bad_resource(cuda):
    c = libceed.Ceed("/gpu/hip/mi355:device_id=0")
    assert c.get_resource().startswith("/gpu/hip/mi355")
    with pytest.raises(libceed.CeedError):
        libceed.Ceed("/gpu/cuda/gen")


# ---------------------------------------------------------------- Vector --
def test_vector_ops(ceed):
    rng = np.random.default_rng(7)
    n = 100_003
    a, b, c = rng.uniform(-1, 1, (3, n))
    x, y, z = (ceed.Vector(n) for _ in range(3))
    x.set_array(a)
    y.set_array(b)
    assert abs(x.norm(libceed.NORM_2) - np.linalg.norm(a)) < 1e-14 * np.linalg.norm(a)
    assert abs(x.norm(libceed.NORM_1) - np.abs(a).sum()) < 1e-14 * np.abs(a).sum()
    assert x.norm(libceed.NORM_MAX) == np.abs(a).max()
    y.axpy(0.5, x)
    ref_y = b + 0.5 * a
    assert rel_l2(y.to_numpy(), ref_y) < 1e-15
    y.axpby(2.0, -1.0, x)
    ref_y = 2.0 * a - ref_y
    assert rel_l2(y.to_numpy(), ref_y) < 1e-15
    z.set_array(c)
    z.pointwise_mult(x, y)
    assert np.array_equal(z.to_numpy(), a * ref_y)
    z.scale(-3.0)
    assert np.array_equal(z.to_numpy(), -3.0 * (a * ref_y))
    z.set_array(c)
    z.reciprocal()
    assert np.array_equal(z.to_numpy(), 1.0 / c)
    z.set_value(4.25)
    assert np.all(z.to_numpy() == 4.25)
    # host writes after device work are seen by the next device read
    with x.array() as h:
        h[:] = 1.0
    assert abs(x.norm(libceed.NORM_1) - n) < 1e-9
    w = ceed.Vector(n)
    w.copy_from(x)
    assert np.all(w.to_numpy() == 1.0)


def test_vector_access_lock(ceed):
    v = ceed.Vector(10)
    v.set_value(1.0)
    with v.array_read():
        with pytest.raises(libceed.CeedError):
            v.set_value(2.0)
    v.set_value(2.0)


# ------------------------------------------------------- ElemRestriction --
@pytest.mark.parametrize("nc", [1, 3])
def test_restriction_bitexact(ceed, nc):
    m = hex_mesh(4, 3, seed=3)
    N, es = m.num_nodes, m.P ** 3
    rng = np.random.default_rng(nc)
    u, v0 = rng.uniform(-1, 1, (2, nc * N))
    r = ceed.ElemRestriction(m.num_elem, es, nc, N, nc * N, m.offsets)
    lv, ev = r.create_vector()
    lv.set_array(u)
    r.apply(lv, ev)
    e_ref = np.zeros(m.num_elem * es * nc)
    oracle.restriction_offset(m.num_elem, es, nc, N, m.offsets, 0, u, e_ref)
    assert np.array_equal(ev.to_numpy(), e_ref)
    lv.set_array(v0)
    r.apply(ev, lv, tmode=libceed.TRANSPOSE)
    v_ref = v0.copy()
    oracle.restriction_offset(m.num_elem, es, nc, N, m.offsets, 1, e_ref, v_ref)
    assert np.array_equal(lv.to_numpy(), v_ref)
    mult = r.get_multiplicity().to_numpy()
    assert np.array_equal(mult, np.tile(np.bincount(m.offsets.reshape(-1), minlength=N), nc).astype(float))


def test_strided_restriction(ceed):
    E, es, nc = 5, 8, 2
    u = np.arange(E * es * nc, dtype=float)
    r = ceed.StridedElemRestriction(E, es, nc, E * es * nc, libceed.STRIDES_BACKEND)
    lv, ev = r.create_vector()
    lv.set_array(u)
    r.apply(lv, ev)
    assert np.array_equal(ev.to_numpy(), u)  # backend strides: E-layout is the L-layout
    strides = [1, E * es, es]  # node, comp, elem
    r2 = ceed.StridedElemRestriction(E, es, nc, E * es * nc, strides)
    r2.apply(lv, ev)
    ref = np.zeros(E * es * nc)
    oracle.restriction_strided(E, es, nc, np.array(strides, dtype=np.int32), 0, u, ref)
    assert np.array_equal(ev.to_numpy(), ref)


# ----------------------------------------------------------------- Basis --
@pytest.mark.parametrize("dim,nc,P,Q", [(3, 1, 5, 6), (3, 3, 4, 4), (2, 2, 3, 5), (1, 1, 6, 8)])
def test_basis_apply(ceed, dim, nc, P, Q):
    E = 3
    B, G, _, w = oracle.lagrange_1d(P, Q)
    b = ceed.BasisTensorH1Lagrange(dim, nc, P, Q, libceed.GAUSS)
    assert np.array_equal(b.get_interp_1d(), B.reshape(Q, P)) and np.array_equal(b.get_grad_1d(), G.reshape(Q, P))
    nn, nq = P ** dim, Q ** dim
    rng = np.random.default_rng(P * 10 + Q)
    u = rng.uniform(-1, 1, E * nc * nn)
    uv = ceed.Vector(u.size)
    uv.set_array(u)
    for em in (libceed.EVAL_INTERP, libceed.EVAL_GRAD):
        qs = nc * nq * (dim if em == libceed.EVAL_GRAD else 1)
        qv = ceed.Vector(E * qs)
        b.apply(E, em, uv, qv)
        ref = np.concatenate([oracle.basis_apply(dim, nc, B, G, w, 1, 0, em, u[e * nc * nn:(e + 1) * nc * nn]) for e in range(E)])
        assert rel_l2(qv.to_numpy(), ref) < RTOL
        tv = ceed.Vector(E * nc * nn)
        t0 = rng.uniform(-1, 1, E * nc * nn)
        tv.set_array(t0)
        b.apply_add(E, em, qv, tv, tmode=libceed.TRANSPOSE)
        reft = t0 + np.concatenate([oracle.basis_apply(dim, nc, B, G, w, 1, 1, em, ref[e * qs:(e + 1) * qs]) for e in range(E)])
        assert rel_l2(tv.to_numpy(), reft) < RTOL
    wv = ceed.Vector(E * nq)
    b.apply(E, libceed.EVAL_WEIGHT, None, wv)
    wref = np.tile(oracle.basis_apply(dim, 1, B, G, w, 1, 0, 16, None), E)
    assert rel_l2(wv.to_numpy(), wref) < 1e-15


# -------------------------------------------------------------- Operators --
def _mesh_objects(ceed, m, Q, kind):
    """Geometry objects and the qdata set-up operator, as ex1-volume.c/ex2-surface.c build them."""
    E, nv = m.num_elem, m.num_vertices
    qc = 1 if kind == 0 else 6
    mesh_basis = ceed.BasisTensorH1Lagrange(3, 3, 2, Q, libceed.GAUSS)
    mesh_rstr = ceed.ElemRestriction(E, 8, 3, nv, 3 * nv, m.mesh_offsets)
    qd_rstr = ceed.StridedElemRestriction(E, Q ** 3, qc, E * Q ** 3 * qc, libceed.STRIDES_BACKEND)
    x = ceed.Vector(3 * nv)
    x.set_array(m.coords)
    return mesh_basis, mesh_rstr, qd_rstr, x


def build_qdata(ceed, m, Q, kind, qf=None):
    mesh_basis, mesh_rstr, qd_rstr, x = _mesh_objects(ceed, m, Q, kind)
    qc = 1 if kind == 0 else 6
    qf = qf or ceed.QFunctionByName("Mass3DBuild" if kind == 0 else "Poisson3DBuild")
    op = ceed.Operator(qf)
    op.set_field("dx", mesh_rstr, mesh_basis, libceed.VECTOR_ACTIVE)
    op.set_field("weights", libceed.ELEMRESTRICTION_NONE, mesh_basis, libceed.VECTOR_NONE)
    op.set_field("qdata", qd_rstr, libceed.BASIS_NONE, libceed.VECTOR_ACTIVE)
    qdata = ceed.Vector(m.num_elem * Q ** 3 * qc)
    op.apply(x, qdata)
    return qdata, qd_rstr


def apply_operator(ceed, m, P, Q, kind, nc, qf, qdata, qd_rstr, u, v0=None):
    N = m.num_nodes
    basis = ceed.BasisTensorH1Lagrange(3, nc, P, Q, libceed.GAUSS)
    rstr = ceed.ElemRestriction(m.num_elem, P ** 3, nc, N, nc * N, m.offsets)
    op = ceed.Operator(qf)
    names = ("u", "qdata", "v") if kind == 0 else ("du", "qdata", "dv")
    op.set_field(names[0], rstr, basis, libceed.VECTOR_ACTIVE)
    op.set_field(names[1], qd_rstr, libceed.BASIS_NONE, qdata)
    op.set_field(names[2], rstr, basis, libceed.VECTOR_ACTIVE)
    uv, vv = ceed.Vector(nc * N), ceed.Vector(nc * N)
    uv.set_array(u)
    if v0 is None:
        vv.set_value(123.0)  # Apply overwrites
        op.apply(uv, vv)
    else:
        vv.set_array(v0)
        op.apply_add(uv, vv)
    return vv.to_numpy(), op


@pytest.mark.parametrize("kind", [0, 1])
def test_qdata_setup_operator(ceed, kind):
    """Gallery Mass3DBuild / Poisson3DBuild through the generic operator (run-time compiled QFunction)."""
    m = hex_mesh(3, 2, seed=11)
    Q = 4
    qdata, _ = build_qdata(ceed, m, Q, kind)
    ref = oracle.hex_build_qdata(kind, m, Q)
    assert rel_l2(qdata.to_numpy(), ref) < 1e-13


GALLERY = {(0, 1): "MassApply", (0, 3): "Vector3MassApply", (1, 1): "Poisson3DApply", (1, 3): "Vector3Poisson3DApply"}


@pytest.mark.parametrize("scatter", ["", "color", "split", "atomic"])
@pytest.mark.parametrize("kind,nc,p", [(0, 1, 3), (1, 1, 4), (0, 3, 2), (1, 3, 3), (1, 1, 7)])
def test_gallery_operator_fused(ceed, monkeypatch, scatter, kind, nc, p):
    """Gallery apply QFunctions take the fused single-kernel path, in the
    default E^T mode (on meshes this small: the transpose map) and with
    CEED_MI355_SCATTER forcing colours, the multiplicity split or atomics."""
    if scatter:
        monkeypatch.setenv("CEED_MI355_SCATTER", scatter)
    m = hex_mesh(3, p, seed=p)
    P, Q = p + 1, p + 2
    qdata, qd_rstr = build_qdata(ceed, m, Q, kind)
    qref = oracle.hex_build_qdata(kind, m, Q)
    rng = np.random.default_rng(kind * 10 + nc)
    u = rng.uniform(-1, 1, nc * m.num_nodes)
    v, _ = apply_operator(ceed, m, P, Q, kind, nc, ceed.QFunctionByName(GALLERY[kind, nc]), qdata, qd_rstr, u)
    ref = oracle.hex_apply(kind, nc, m.num_elem, P, Q, m.offsets, m.num_nodes, qref, u)
    assert rel_l2(v, ref) < RTOL
    v0 = rng.uniform(-1, 1, u.size)
    v2, _ = apply_operator(ceed, m, P, Q, kind, nc, ceed.QFunctionByName(GALLERY[kind, nc]), qdata, qd_rstr, u, v0=v0)
    assert rel_l2(v2, ref + v0) < RTOL


def test_fused_operator_bitwise_repeatable(cuda):
    """The fused operator gives identical bits on every apply."""
    c = libceed.Ceed("/gpu/hip/mi355")
    m = hex_mesh(5, 3, seed=8)
    qdata, qd_rstr = build_qdata(c, m, 5, 1)
    u = np.random.default_rng(2).uniform(-1, 1, m.num_nodes)
    v1, op = apply_operator(c, m, 4, 5, 1, 1, c.QFunctionByName("Poisson3DApply"), qdata, qd_rstr, u)
    uv, vv = c.Vector(m.num_nodes), c.Vector(m.num_nodes)
    uv.set_array(u)
    for _ in range(20):
        op.apply(uv, vv)
        assert np.array_equal(vv.to_numpy(), v1)
    ref = oracle.hex_apply(1, 1, m.num_elem, 4, 5, m.offsets, m.num_nodes, oracle.hex_build_qdata(1, m, 5), u)
    assert rel_l2(v1, ref) < RTOL


def _user_qf(ceed, kind):
    if kind == 0:
        qf = ceed.QFunction(1, None, f"{QF_FILE}:user_mass")
        qf.add_input("u", 1, libceed.EVAL_INTERP)
        qf.add_input("qdata", 1, libceed.EVAL_NONE)
        qf.add_output("v", 1, libceed.EVAL_INTERP)
    else:
        qf = ceed.QFunction(1, None, f"{QF_FILE}:user_diffusion")
        qf.add_input("du", 3, libceed.EVAL_GRAD)
        qf.add_input("qdata", 6, libceed.EVAL_NONE)
        qf.add_output("dv", 3, libceed.EVAL_GRAD)
    return qf


@pytest.mark.parametrize("fuse", ["1", "0"])
@pytest.mark.parametrize("kind,p", [(0, 3), (1, 4), (1, 2), (1, 5), (0, 7), (1, 8)])
def test_user_qfunction(ceed, monkeypatch, fuse, kind, p):
    """A user QFunction (not in the gallery): fused into the run-time compiled
    line-layout kernel (default; on meshes this small the E^T is the
    transpose map), or the staged E -> B -> D -> B^T -> E^T path with
    CEED_MI355_FUSE_QFUNCTIONS=0.  Apply and apply-add."""
    monkeypatch.setenv("CEED_MI355_FUSE_QFUNCTIONS", fuse)
    m = hex_mesh(3 if p < 7 else 2, p, seed=20 + p)
    P, Q = p + 1, p + 2
    qdata, qd_rstr = build_qdata(ceed, m, Q, kind)
    qref = oracle.hex_build_qdata(kind, m, Q)
    rng = np.random.default_rng(p)
    u = rng.uniform(-1, 1, m.num_nodes)
    v, _ = apply_operator(ceed, m, P, Q, kind, 1, _user_qf(ceed, kind), qdata, qd_rstr, u)
    ref = oracle.hex_apply(kind, 1, m.num_elem, P, Q, m.offsets, m.num_nodes, qref, u)
    assert rel_l2(v, ref) < RTOL
    v0 = rng.uniform(-1, 1, u.size)
    v2, _ = apply_operator(ceed, m, P, Q, kind, 1, _user_qf(ceed, kind), qdata, qd_rstr, u, v0=v0)
    assert rel_l2(v2, ref + v0) < RTOL


@pytest.mark.parametrize("scatter", ["color", "tmap", "split", "atomic"])
@pytest.mark.parametrize("kind,p", [(0, 3), (1, 4), (1, 6)])
def test_user_qfunction_scatter_modes(ceed, monkeypatch, scatter, kind, p):
    """The run-time fused kernel in each E^T mode (CEED_MI355_SCATTER replaces
    the size-based default): colour launches with first-touch overwrite, the
    transpose map, atomics.  Apply and apply-add against the oracle."""
    monkeypatch.setenv("CEED_MI355_FUSE_QFUNCTIONS", "1")
    monkeypatch.setenv("CEED_MI355_SCATTER", scatter)
    m = hex_mesh(4, p, seed=40 + p)
    P, Q = p + 1, p + 2
    qdata, qd_rstr = build_qdata(ceed, m, Q, kind)
    qref = oracle.hex_build_qdata(kind, m, Q)
    rng = np.random.default_rng(p + 7)
    u = rng.uniform(-1, 1, m.num_nodes)
    v, _ = apply_operator(ceed, m, P, Q, kind, 1, _user_qf(ceed, kind), qdata, qd_rstr, u)
    ref = oracle.hex_apply(kind, 1, m.num_elem, P, Q, m.offsets, m.num_nodes, qref, u)
    assert rel_l2(v, ref) < RTOL
    v0 = rng.uniform(-1, 1, u.size)
    v2, _ = apply_operator(ceed, m, P, Q, kind, 1, _user_qf(ceed, kind), qdata, qd_rstr, u, v0=v0)
    assert rel_l2(v2, ref + v0) < RTOL


def test_user_qfunction_fused_matches_staged(ceed, monkeypatch):
    """The run-time fused kernel and the staged path agree to rounding on the
    same user operator; the fused result repeats bit for bit."""
    m = hex_mesh(4, 4, seed=31)
    P, Q = 5, 6
    qdata, qd_rstr = build_qdata(ceed, m, Q, 1)
    u = np.random.default_rng(7).uniform(-1, 1, m.num_nodes)
    out = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("CEED_MI355_FUSE_QFUNCTIONS", fuse)
        out[fuse], op = apply_operator(ceed, m, P, Q, 1, 1, _user_qf(ceed, 1), qdata, qd_rstr, u)
        uv, vv = ceed.Vector(m.num_nodes), ceed.Vector(m.num_nodes)
        uv.set_array(u)
        op.apply(uv, vv)
        assert np.array_equal(vv.to_numpy(), out[fuse])
    assert rel_l2(out["1"], out["0"]) < 1e-14


@pytest.mark.parametrize("fuse", ["1", "0"])
def test_two_field_operator(ceed, monkeypatch, fuse):
    """Two active inputs (INTERP and GRAD of u), two passive inputs and two
    active outputs: mass + diffusion.  Two passive inputs keep it on the
    staged path whatever CEED_MI355_FUSE_QFUNCTIONS says."""
    monkeypatch.setenv("CEED_MI355_FUSE_QFUNCTIONS", fuse)
    m = hex_mesh(3, 3, seed=5)
    P, Q = 4, 5
    N = m.num_nodes
    qm, qm_rstr = build_qdata(ceed, m, Q, 0)
    qd, qd_rstr = build_qdata(ceed, m, Q, 1)
    qf = ceed.QFunction(1, None, f"{QF_FILE}:user_mass_diffusion")
    qf.add_input("u", 1, libceed.EVAL_INTERP)
    qf.add_input("du", 3, libceed.EVAL_GRAD)
    qf.add_input("qm", 1, libceed.EVAL_NONE)
    qf.add_input("qd", 6, libceed.EVAL_NONE)
    qf.add_output("v", 1, libceed.EVAL_INTERP)
    qf.add_output("dv", 3, libceed.EVAL_GRAD)
    basis = ceed.BasisTensorH1Lagrange(3, 1, P, Q, libceed.GAUSS)
    rstr = ceed.ElemRestriction(m.num_elem, P ** 3, 1, N, N, m.offsets)
    op = ceed.Operator(qf)
    op.set_field("u", rstr, basis, libceed.VECTOR_ACTIVE)
    op.set_field("du", rstr, basis, libceed.VECTOR_ACTIVE)
    op.set_field("qm", qm_rstr, libceed.BASIS_NONE, qm)
    op.set_field("qd", qd_rstr, libceed.BASIS_NONE, qd)
    op.set_field("v", rstr, basis, libceed.VECTOR_ACTIVE)
    op.set_field("dv", rstr, basis, libceed.VECTOR_ACTIVE)
    u = np.random.default_rng(1).uniform(-1, 1, N)
    uv, vv = ceed.Vector(N), ceed.Vector(N)
    uv.set_array(u)
    op.apply(uv, vv)
    ref = oracle.hex_apply(0, 1, m.num_elem, P, Q, m.offsets, N, oracle.hex_build_qdata(0, m, Q), u)
    ref = oracle.hex_apply(1, 1, m.num_elem, P, Q, m.offsets, N, oracle.hex_build_qdata(1, m, Q), u, v=ref)
    assert rel_l2(vv.to_numpy(), ref) < RTOL


@pytest.mark.parametrize("fuse", ["1", "0"])
def test_two_field_operator_one_passive(ceed, monkeypatch, fuse):
    """INTERP and GRAD inputs and outputs with one passive field of 7 values
    per point (mass factor, then the 6 diffusion factors): fused at run time
    by default."""
    monkeypatch.setenv("CEED_MI355_FUSE_QFUNCTIONS", fuse)
    m = hex_mesh(3, 4, seed=6)
    P, Q, N = 5, 6, m.num_nodes
    qm, qd = oracle.hex_build_qdata(0, m, Q), oracle.hex_build_qdata(1, m, Q)
    q3 = Q ** 3
    both = np.concatenate([qm.reshape(m.num_elem, 1, q3), qd.reshape(m.num_elem, 6, q3)], axis=1).ravel()
    q_rstr = ceed.StridedElemRestriction(m.num_elem, q3, 7, both.size, libceed.STRIDES_BACKEND)
    qv = ceed.Vector(both.size)
    qv.set_array(both)
    qf = ceed.QFunction(1, None, f"{QF_FILE}:user_mass_diffusion7")
    qf.add_input("u", 1, libceed.EVAL_INTERP)
    qf.add_input("du", 3, libceed.EVAL_GRAD)
    qf.add_input("q", 7, libceed.EVAL_NONE)
    qf.add_output("v", 1, libceed.EVAL_INTERP)
    qf.add_output("dv", 3, libceed.EVAL_GRAD)
    basis = ceed.BasisTensorH1Lagrange(3, 1, P, Q, libceed.GAUSS)
    rstr = ceed.ElemRestriction(m.num_elem, P ** 3, 1, N, N, m.offsets)
    op = ceed.Operator(qf)
    op.set_field("u", rstr, basis, libceed.VECTOR_ACTIVE)
    op.set_field("du", rstr, basis, libceed.VECTOR_ACTIVE)
    op.set_field("q", q_rstr, libceed.BASIS_NONE, qv)
    op.set_field("v", rstr, basis, libceed.VECTOR_ACTIVE)
    op.set_field("dv", rstr, basis, libceed.VECTOR_ACTIVE)
    u = np.random.default_rng(2).uniform(-1, 1, N)
    uv, vv = ceed.Vector(N), ceed.Vector(N)
    uv.set_array(u)
    op.apply(uv, vv)
    ref = oracle.hex_apply(0, 1, m.num_elem, P, Q, m.offsets, N, qm, u)
    ref = oracle.hex_apply(1, 1, m.num_elem, P, Q, m.offsets, N, qd, u, v=ref)
    assert rel_l2(vv.to_numpy(), ref) < RTOL


@pytest.mark.parametrize("fuse", ["1", "0"])
def test_apply_qfunction_context(ceed, monkeypatch, fuse):
    """An apply QFunction reading its context (v = scale * qd * u); a new
    context value is seen by the next apply, fused or staged."""
    monkeypatch.setenv("CEED_MI355_FUSE_QFUNCTIONS", fuse)
    m = hex_mesh(3, 3, seed=12)
    P, Q = 4, 5
    qdata, qd_rstr = build_qdata(ceed, m, Q, 0)
    qf = ceed.QFunction(1, None, f"{QF_FILE}:scaled_mass")
    qf.add_input("u", 1, libceed.EVAL_INTERP)
    qf.add_input("qdata", 1, libceed.EVAL_NONE)
    qf.add_output("v", 1, libceed.EVAL_INTERP)
    ctx = ceed.QFunctionContext()
    ctx.set_data(np.array([3.0]))
    qf.set_context(ctx)
    u = np.random.default_rng(4).uniform(-1, 1, m.num_nodes)
    v, op = apply_operator(ceed, m, P, Q, 0, 1, qf, qdata, qd_rstr, u)
    ref = oracle.hex_apply(0, 1, m.num_elem, P, Q, m.offsets, m.num_nodes, oracle.hex_build_qdata(0, m, Q), u)
    assert rel_l2(v, 3.0 * ref) < RTOL
    with ctx.data() as d:
        d[0] = -0.5
    uv, vv = ceed.Vector(m.num_nodes), ceed.Vector(m.num_nodes)
    uv.set_array(u)
    op.apply(uv, vv)
    assert rel_l2(vv.to_numpy(), -0.5 * ref) < RTOL


def test_qfunction_context(ceed):
    """A context-dependent set-up QFunction; changing the context data is seen by the next apply."""
    m = hex_mesh(2, 2, seed=4)
    Q = 4
    qf = ceed.QFunction(1, None, f"{QF_FILE}:scaled_mass_build")
    qf.add_input("dx", 9, libceed.EVAL_GRAD)
    qf.add_input("weights", 1, libceed.EVAL_WEIGHT)
    qf.add_output("qdata", 1, libceed.EVAL_NONE)
    ctx = ceed.QFunctionContext()
    ctx.set_data(np.array([2.5]))
    qf.set_context(ctx)
    qdata, _ = build_qdata(ceed, m, Q, 0, qf=qf)
    ref = oracle.hex_build_qdata(0, m, Q)
    assert rel_l2(qdata.to_numpy(), 2.5 * ref) < 1e-14
    with ctx.data() as d:
        d[0] = -1.0
    qdata, _ = build_qdata(ceed, m, Q, 0, qf=qf)
    assert rel_l2(qdata.to_numpy(), -ref) < 1e-14


def test_passive_input_state(ceed):
    """Changing the passive qdata vector between applies reloads the fused plan's copy."""
    m = hex_mesh(3, 3, seed=9)
    P, Q = 4, 5
    qdata, qd_rstr = build_qdata(ceed, m, Q, 1)
    u = np.random.default_rng(3).uniform(-1, 1, m.num_nodes)
    v1, op = apply_operator(ceed, m, P, Q, 1, 1, ceed.QFunctionByName("Poisson3DApply"), qdata, qd_rstr, u)
    qdata.scale(2.0)
    uv, vv = ceed.Vector(m.num_nodes), ceed.Vector(m.num_nodes)
    uv.set_array(u)
    op.apply(uv, vv)
    assert rel_l2(vv.to_numpy(), 2.0 * v1) < 1e-15


def test_volume_and_flux(ceed):
    """ex1-volume: 1^T M 1 = |[0,1]^3| = 1.  ex2-surface applies the Laplacian
    to u = x+y+z (a boundary-flux vector); exact identities of that vector on
    the perturbed trilinear mesh: 1^T A u = 0 and x^T A u = int grad x . grad u = 1
    (the integrand is det J, integrated exactly by Q >= 2 Gauss points)."""
    m = hex_mesh(4, 3, seed=1)
    P, Q = 4, 5
    N = m.num_nodes
    qm, qm_rstr = build_qdata(ceed, m, Q, 0)
    v, _ = apply_operator(ceed, m, P, Q, 0, 1, ceed.QFunctionByName("MassApply"), qm, qm_rstr, np.ones(N))
    assert abs(v.sum() - 1.0) < 1e-12
    qd, qd_rstr = build_qdata(ceed, m, Q, 1)
    xyz = node_coordinates(m, oracle.lobatto(P)[0])
    v, _ = apply_operator(ceed, m, P, Q, 1, 1, ceed.QFunctionByName("Poisson3DApply"), qd, qd_rstr, xyz.sum(axis=0))
    assert abs(v.sum()) < 1e-12
    assert abs(xyz[0] @ v - 1.0) < 1e-12
    interior = np.ones(N, dtype=bool)
    interior[np.isclose(xyz, 0.0).any(axis=0) | np.isclose(xyz, 1.0).any(axis=0)] = False
    assert np.abs(v[interior]).max() < 1e-12


def test_field_errors(ceed):
    m = hex_mesh(2, 2)
    basis = ceed.BasisTensorH1Lagrange(3, 1, 3, 4, libceed.GAUSS)
    rstr = ceed.ElemRestriction(m.num_elem, 27, 1, m.num_nodes, m.num_nodes, m.offsets)
    op = ceed.Operator(ceed.QFunctionByName("Poisson3DApply"))
    with pytest.raises(libceed.CeedError, match="no field named"):
        op.set_field("nonexistent", rstr, basis, libceed.VECTOR_ACTIVE)
    vbasis = ceed.BasisTensorH1Lagrange(3, 3, 3, 4, libceed.GAUSS)
    with pytest.raises(libceed.CeedError):
        op.set_field("du", rstr, vbasis, libceed.VECTOR_ACTIVE)  # 9 components vs QFunction size 3
    with pytest.raises(libceed.CeedError, match="not all operator fields"):
        op.check()


def test_composite_operator_and_diagonal(ceed):
    """A composite of mass and diffusion sub-operators applies their sum; its
    assembled diagonal is the sum of the sub-operators' diagonals, each equal to
    the diagonal of the dense operator the oracle assembles."""
    m = hex_mesh(2, 2, seed=6)
    P, Q = 3, 4
    N = m.num_nodes
    qm, qm_rstr = build_qdata(ceed, m, Q, 0)
    qd, qd_rstr = build_qdata(ceed, m, Q, 1)
    _, op_m = apply_operator(ceed, m, P, Q, 0, 1, ceed.QFunctionByName("MassApply"), qm, qm_rstr, np.zeros(N))
    _, op_k = apply_operator(ceed, m, P, Q, 1, 1, ceed.QFunctionByName("Poisson3DApply"), qd, qd_rstr, np.zeros(N))
    comp = ceed.CompositeOperator()
    comp.add_sub(op_m)
    comp.add_sub(op_k)
    u = np.random.default_rng(5).uniform(-1, 1, N)
    uv, vv = ceed.Vector(N), ceed.Vector(N)
    uv.set_array(u)
    comp.apply(uv, vv)
    qm_ref, qd_ref = oracle.hex_build_qdata(0, m, Q), oracle.hex_build_qdata(1, m, Q)
    ref = oracle.hex_apply(0, 1, m.num_elem, P, Q, m.offsets, N, qm_ref, u)
    ref = oracle.hex_apply(1, 1, m.num_elem, P, Q, m.offsets, N, qd_ref, u, v=ref)
    assert rel_l2(vv.to_numpy(), ref) < RTOL
    dm, dk = np.zeros(N), np.zeros(N)
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        dm[j] = oracle.hex_apply(0, 1, m.num_elem, P, Q, m.offsets, N, qm_ref, e)[j]
        dk[j] = oracle.hex_apply(1, 1, m.num_elem, P, Q, m.offsets, N, qd_ref, e)[j]
    dv = ceed.Vector(N)
    dv.set_value(3.0)  # overwritten
    comp.linear_assemble_diagonal(dv)
    assert rel_l2(dv.to_numpy(), dm + dk) < RTOL
    op_k.linear_assemble_add_diagonal(dv)  # adds the diffusion diagonal once more
    assert rel_l2(dv.to_numpy(), dm + 2 * dk) < RTOL
