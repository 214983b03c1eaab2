"""CeedQFunctionApply on /gpu/hip/mi355: the gallery QFunctions by name and a
user QFunction with a context, each evaluated at Q points of device vectors,
the way the reference's tests call them (tests/t410-qfunction.c:39-52,
t400-t415), checked against the oracle's pointwise functions
(oracle/ceed_oracle.c orc_qf_*).

Bar: within 1e-12 relative L2 of the oracle (FMA contraction on the GPU may
change the last bits of the set-up algebra); MassApply, a single product per
point, bitwise.  Point counts cover one point, a partial wavefront, and
several thousand points in a partial last workgroup."""

from pathlib import Path

import numpy as np
import pytest

import oracle
from libceed_amd import ceed as libceed

pytestmark = pytest.mark.gpu
QF_FILE = Path(__file__).resolve().parent / "qfunctions" / "t-api.h"
RTOL = 1e-12


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ceed(cuda):
    return libceed.Ceed("/gpu/hip/mi355")


def jacobians(rng, Q):
    """Q Jacobians near the identity (det J > 0), component (a, b) of point i at (3 a + b) Q + i."""
    J = np.eye(3)[:, :, None] + 0.3 * rng.uniform(-1, 1, (3, 3, Q))
    return J.reshape(-1)


def inputs_for(name, rng, Q):
    _, in_sizes, _ = oracle.QF_GALLERY[name]
    if name.endswith("Build"):
        return [jacobians(rng, Q), rng.uniform(0.1, 1.0, Q)]
    ins = [rng.uniform(-1, 1, s * Q) for s in in_sizes]
    if name.startswith("Poisson") or name.startswith("Vector3Poisson"):  # symmetric positive q_data in Voigt order
        qd = rng.uniform(0.0, 0.2, (6, Q))
        qd[:3] += 1.0
        ins[1] = qd.reshape(-1)
    return ins


def apply_qf(ceed, qf, Q, ins, out_sizes):
    vin = []
    for a in ins:
        x = ceed.Vector(a.size)
        x.set_array(a)
        vin.append(x)
    vout = []
    for s in out_sizes:
        y = ceed.Vector(s * Q)
        y.set_value(0.0)
        vout.append(y)
    qf.apply(Q, vin, vout)
    return [y.to_numpy() for y in vout]


@pytest.mark.parametrize("Q", [1, 37, 5000])
@pytest.mark.parametrize("name", list(oracle.QF_GALLERY))
def test_gallery_qfunction_apply(ceed, name, Q):
    print("CeedQFunctionApply", name, Q, flush=True)
    rng = np.random.default_rng(hash((name, Q)) % 2**32)
    ins = inputs_for(name, rng, Q)
    ref = oracle.qf_apply(name, Q, ins)
    qf = ceed.QFunctionByName(name)
    got = apply_qf(ceed, qf, Q, ins, oracle.QF_GALLERY[name][2])
    for g, r in zip(got, ref):
        if name == "MassApply":
            assert np.array_equal(g, r)
        else:
            assert rel_l2(g, r) < RTOL, (name, Q, rel_l2(g, r))


def test_mass_apply_known_answer(ceed):
    """tests/t410-qfunction.c: MassApply of w and u at 8 points gives w u exactly."""
    q = 8
    x = 2.0 * np.arange(q) / (q - 1) - 1
    w, u = 1 - x * x, 2 + 3 * x + 5 * x * x
    (v,) = apply_qf(ceed, ceed.QFunctionByName("MassApply"), q, [w, u], (1,))
    assert np.array_equal(v, w * u)


@pytest.mark.parametrize("Q", [17, 3000])
def test_user_qfunction_apply_with_context(ceed, Q):
    """A user QFunction compiled at run time (scaled_mass_build: scale * w * det J, scale from the
    context); a changed context value is seen by the next apply."""
    rng = np.random.default_rng(Q)
    ins = [jacobians(rng, Q), rng.uniform(0.1, 1.0, Q)]
    (base,) = oracle.qf_apply("Mass3DBuild", Q, ins)
    qf = ceed.QFunction(1, None, f"{QF_FILE}:scaled_mass_build")
    qf.add_input("dx", 9, libceed.EVAL_GRAD)
    qf.add_input("weights", 1, libceed.EVAL_WEIGHT)
    qf.add_output("qdata", 1, libceed.EVAL_NONE)
    ctx = ceed.QFunctionContext()
    ctx.set_data(np.array([2.5]))
    qf.set_context(ctx)
    (got,) = apply_qf(ceed, qf, Q, ins, (1,))
    assert rel_l2(got, 2.5 * base) < RTOL
    with ctx.data() as d:
        d[0] = -0.75
    (got,) = apply_qf(ceed, qf, Q, ins, (1,))
    assert rel_l2(got, -0.75 * base) < RTOL
