"""CPU oracle for the hex operator path -- TEST INFRASTRUCTURE ONLY.

A ctypes wrapper around ``liboracle.so`` (``ceed_oracle.c``, a plain-C
restatement of the reference CPU backend /cpu/self/ref/serial).  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may use it, as
the checker; the product (``libceed_amd``) never does.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libceed.so"

GAUSS, GAUSS_LOBATTO = 0, 1
NOTRANSPOSE, TRANSPOSE = 0, 1
EVAL_NONE, EVAL_INTERP, EVAL_GRAD, EVAL_WEIGHT = 0, 1, 2, 16

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE), "oracle"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        vp, i = C.c_void_p, C.c_int
        sig = {
            "orc_gauss_quadrature": [i, vp, vp],
            "orc_lobatto_quadrature": [i, vp, vp],
            "orc_lagrange_1d": [i, i, i, vp, vp, vp, vp],
            "orc_pseudoinverse": [vp, i, i, vp],
            "orc_collocated_grad": [i, i, vp, vp, vp],
            "orc_basis_apply": [i, i, i, i, vp, vp, vp, i, i, i, i, vp, vp],
            "orc_restriction_offset": [i, i, i, i, vp, i, vp, vp],
            "orc_restriction_strided": [i, i, i, vp, i, vp, vp],
            "orc_hex_build_qdata": [i, i, i, i, i, vp, i, vp, vp],
            "orc_hex_apply": [i, i, i, i, i, i, vp, i, vp, vp, vp],
            "orc_hex_apply_kappa": [i, i, i, i, vp, vp, vp, vp, vp],
            "orc_hex_kappa_gradient": [i, i, i, i, vp, vp, vp, vp, vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int if name not in ("orc_restriction_offset", "orc_restriction_strided") else None
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return a.ctypes.data


def _ok(ierr):
    if ierr:
        raise RuntimeError(f"oracle call failed ({ierr})")


def gauss(Q):
    x, w = np.zeros(Q), np.zeros(Q)
    _ok(lib().orc_gauss_quadrature(Q, _p(x), _p(w)))
    return x, w


def lobatto(Q):
    x, w = np.zeros(Q), np.zeros(Q)
    _ok(lib().orc_lobatto_quadrature(Q, _p(x), _p(w)))
    return x, w


def lagrange_1d(P, Q, quad_mode=GAUSS):
    B, G, x, w = np.zeros((Q, P)), np.zeros((Q, P)), np.zeros(Q), np.zeros(Q)
    _ok(lib().orc_lagrange_1d(P, Q, quad_mode, _p(B), _p(G), _p(x), _p(w)))
    return B, G, x, w


def pseudoinverse(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    m, n = A.shape
    out = np.zeros((n, m))
    _ok(lib().orc_pseudoinverse(_p(A), m, n, _p(out)))
    return out


def collocated_grad(B, G):
    Q, P = B.shape
    D = np.zeros((Q, Q))
    _ok(lib().orc_collocated_grad(P, Q, _p(np.ascontiguousarray(B)), _p(np.ascontiguousarray(G)), _p(D)))
    return D


def basis_apply(dim, num_comp, B, G, w, num_elem, t_mode, eval_mode, u, v=None, add=False):
    """Reference-layout basis apply (element index fastest)."""
    Q, P = B.shape
    nn, nq = P ** dim, Q ** dim
    if v is None:
        if t_mode == NOTRANSPOSE:
            size = nq * num_elem * (dim * num_comp if eval_mode == EVAL_GRAD else (1 if eval_mode == EVAL_WEIGHT else num_comp))
        else:
            size = nn * num_comp * num_elem
        v = np.zeros(size)
    u = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
    _ok(lib().orc_basis_apply(dim, num_comp, P, Q, _p(np.ascontiguousarray(B)), _p(np.ascontiguousarray(G)),
                              _p(np.ascontiguousarray(w)), num_elem, t_mode, eval_mode, int(add), _p(u), _p(v)))
    return v


def restriction_offset(num_elem, elem_size, num_comp, comp_stride, offsets, t_mode, u, v):
    off = np.ascontiguousarray(offsets, dtype=np.int32)
    lib().orc_restriction_offset(num_elem, elem_size, num_comp, comp_stride, _p(off), t_mode, _p(u), _p(v))
    return v


def restriction_strided(num_elem, elem_size, num_comp, strides, t_mode, u, v):
    st = np.ascontiguousarray(strides, dtype=np.int32)
    lib().orc_restriction_strided(num_elem, elem_size, num_comp, _p(st), t_mode, _p(u), _p(v))
    return v


def hex_build_qdata(kind, mesh, Q, quad_mode=GAUSS, P_mesh=2):
    qc = 1 if kind == 0 else 6
    qd = np.zeros(mesh.num_elem * qc * Q ** 3)
    _ok(lib().orc_hex_build_qdata(kind, mesh.num_elem, P_mesh, Q, quad_mode, _p(np.ascontiguousarray(mesh.mesh_offsets)),
                                  mesh.num_vertices, _p(mesh.coords), _p(qd)))
    return qd


def hex_apply(kind, num_comp, num_elem, P, Q, offsets, comp_stride, q_data, u, v=None, quad_mode=GAUSS):
    """v += A u (v defaults to zeros)."""
    if v is None:
        v = np.zeros_like(u)
    _ok(lib().orc_hex_apply(kind, num_comp, num_elem, P, Q, quad_mode, _p(np.ascontiguousarray(offsets, dtype=np.int32)),
                            comp_stride, _p(q_data), _p(u), _p(v)))
    return v


def hex_apply_kappa(num_elem, P, Q, offsets, q_data, kappa, u, v=None, quad_mode=GAUSS):
    if v is None:
        v = np.zeros_like(u)
    _ok(lib().orc_hex_apply_kappa(num_elem, P, Q, quad_mode, _p(np.ascontiguousarray(offsets, dtype=np.int32)), _p(q_data),
                                  _p(kappa), _p(u), _p(v)))
    return v


def hex_kappa_gradient(num_elem, P, Q, offsets, q_data, u, lam, quad_mode=GAUSS):
    dk = np.zeros(num_elem * Q ** 3)
    _ok(lib().orc_hex_kappa_gradient(num_elem, P, Q, quad_mode, _p(np.ascontiguousarray(offsets, dtype=np.int32)), _p(q_data),
                                     _p(u), _p(lam), _p(dk)))
    return dk


# gallery name -> (oracle function, input sizes per point, output sizes per point)
QF_GALLERY = {
    "Mass3DBuild": ("orc_qf_mass3d_build", (9, 1), (1,)),
    "MassApply": ("orc_qf_mass_apply", (1, 1), (1,)),
    "Poisson3DBuild": ("orc_qf_poisson3d_build", (9, 1), (6,)),
    "Poisson3DApply": ("orc_qf_poisson3d_apply", (3, 6), (3,)),
    "Vector3MassApply": ("orc_qf_vector3_mass_apply", (3, 1), (3,)),
    "Vector3Poisson3DApply": ("orc_qf_vector3_poisson3d_apply", (9, 6), (9,)),
}


def qf_apply(name, Q, inputs):
    """The oracle's pointwise gallery QFunction ``name`` over Q points.

    ``inputs`` are arrays of size * Q values each (field component k of point
    i at k * Q + i, the layout CeedQFunctionApply hands a QFunction); returns
    the list of output arrays in the same layout."""
    fname, in_sizes, out_sizes = QF_GALLERY[name]
    ins = [np.ascontiguousarray(a, dtype=np.float64) for a in inputs]
    assert [a.size for a in ins] == [s * Q for s in in_sizes], "input sizes"
    outs = [np.zeros(s * Q) for s in out_sizes]
    f = getattr(lib(), fname)
    f.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    in_ptrs = (C.c_void_p * len(ins))(*[a.ctypes.data for a in ins])
    out_ptrs = (C.c_void_p * len(outs))(*[a.ctypes.data for a in outs])
    _ok(f(None, Q, in_ptrs, out_ptrs))
    return outs
