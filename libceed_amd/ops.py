"""Python objects over the C ABI: restriction, tensor basis, fused hex operator.

Device arrays are torch CUDA (HIP) tensors; every call is enqueued on torch's
current stream.  These wrappers are what the tests, the benchmark and the
autograd bindings use; the computation always runs in the HIP kernels of
``libceed-mi355.so``.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import CeedMi355Error, call  # noqa: F401  (CeedMi355Error re-exported)

NOTRANSPOSE, TRANSPOSE = 0, 1
EVAL_NONE, EVAL_INTERP, EVAL_GRAD, EVAL_WEIGHT = 0, 1, 2, 16
QF_MASS, QF_DIFFUSION, QF_DIFFUSION_KAPPA = 0, 1, 2
LAYOUT_COLUMNS, LAYOUT_LINES, LAYOUT_POINTS = 0, 1, 2  # thread layouts of the fused kernel
FP64, FP32 = 0, 1
# E^T of the fused plan: "auto" (colours below P = 6 and the multiplicity split at P = 7, 8 in
# FP64, both on meshes of at least 10M E-vector entries; the transpose map otherwise), transpose offset map, element
# colours, multiplicity split (atomic adds for entries of at most two contributions, the ordered map for the rest),
# hardware atomics, bricks of face-sharing elements, coloured x-segments of face-sharing elements
SCATTER_FLAGS = {"auto": 0, "tmap": 8, "color": 2, "atomic": 4, "split": 16, "brick": 32, "segment": 64}
GAUSS, GAUSS_LOBATTO = 0, 1


def _ptr(a) -> C.c_void_p:
    if a is None:
        return C.c_void_p(0)
    if isinstance(a, torch.Tensor):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return C.c_void_p(a.data_ptr())
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return C.c_void_p(a.ctypes.data)
    raise TypeError(type(a))


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _device_tensor(a, dtype) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        if a.device.type != "cuda":
            raise ValueError("expected a device tensor")
        if a.dtype != dtype:
            raise TypeError(f"expected {dtype}, got {a.dtype}")
        return a
    raise TypeError("expected a torch CUDA tensor")


def device_init(device_id: int = 0) -> None:
    call("CeedMi355DeviceInit", device_id)


def lagrange_basis_1d(P: int, Q: int, quad_mode: int = GAUSS):
    """(interp, grad, q_ref, q_weight) of the 1D Lagrange basis on P GLL nodes at Q points (host, FP64)."""
    interp = np.zeros((Q, P))
    grad = np.zeros((Q, P))
    qref = np.zeros(Q)
    qw = np.zeros(Q)
    call("CeedMi355LagrangeBasis1D", P, Q, quad_mode, _ptr(interp), _ptr(grad), _ptr(qref), _ptr(qw))
    return interp, grad, qref, qw


def gauss_quadrature(Q: int):
    x, w = np.zeros(Q), np.zeros(Q)
    call("CeedGaussQuadrature", Q, _ptr(x), _ptr(w))
    return x, w


def lobatto_quadrature(Q: int):
    x, w = np.zeros(Q), np.zeros(Q)
    call("CeedLobattoQuadrature", Q, _ptr(x), _ptr(w))
    return x, w


class Restriction:
    """Element restriction E (offsets or strides) with bit-exact transpose."""

    def __init__(self, num_elem: int, elem_size: int, num_comp: int, l_size: int, offsets=None, comp_stride: int = 1,
                 strides=None):
        self.num_elem, self.elem_size, self.num_comp, self.l_size = num_elem, elem_size, num_comp, l_size
        h = C.c_void_p()
        if offsets is not None:
            off = np.ascontiguousarray(offsets, dtype=np.int32).reshape(-1)
            if off.size != num_elem * elem_size:
                raise ValueError("offsets must have num_elem * elem_size entries")
            call("CeedMi355RestrictionCreateOffset", num_elem, elem_size, num_comp, comp_stride, l_size, _ptr(off), C.byref(h))
        else:
            st = np.ascontiguousarray(strides, dtype=np.int32)
            call("CeedMi355RestrictionCreateStrided", num_elem, elem_size, num_comp, l_size, _ptr(st), C.byref(h))
        self._h = h

    @property
    def e_size(self) -> int:
        return self.num_elem * self.elem_size * self.num_comp

    def apply(self, t_mode: int, u: torch.Tensor, v: torch.Tensor) -> None:
        u, v = _device_tensor(u, torch.float64), _device_tensor(v, torch.float64)
        need_u, need_v = (self.l_size, self.e_size) if t_mode == NOTRANSPOSE else (self.e_size, self.l_size)
        if u.numel() < need_u or v.numel() < need_v:
            raise ValueError(f"vector too short for restriction ({u.numel()}, {v.numel()}) vs ({need_u}, {need_v})")
        call("CeedMi355RestrictionApply", self._h, t_mode, _ptr(u), _ptr(v), _stream())

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None and _lib._lib is not None:
            _lib.lib().CeedMi355RestrictionDestroy(C.byref(self._h))


class Basis:
    """Tensor-product H1 basis of dimension 1..3."""

    def __init__(self, dim: int, num_comp: int, P: int, Q: int, interp=None, grad=None, q_weight=None, quad_mode: int = GAUSS):
        if interp is None:
            interp, grad, _, q_weight = lagrange_basis_1d(P, Q, quad_mode)
        self.dim, self.num_comp, self.P, self.Q = dim, num_comp, P, Q
        self.interp = np.ascontiguousarray(interp, dtype=np.float64)
        self.grad = np.ascontiguousarray(grad, dtype=np.float64)
        self.q_weight = np.ascontiguousarray(q_weight, dtype=np.float64)
        h = C.c_void_p()
        call("CeedMi355BasisCreateTensorH1", dim, num_comp, P, Q, _ptr(self.interp), _ptr(self.grad), _ptr(self.q_weight), C.byref(h))
        self._h = h

    def apply(self, num_elem: int, t_mode: int, eval_mode: int, u, v: torch.Tensor, add: bool = False) -> None:
        v = _device_tensor(v, torch.float64)
        if u is not None:
            u = _device_tensor(u, torch.float64)
        call("CeedMi355BasisApply", self._h, num_elem, t_mode, eval_mode, int(add), _ptr(u), _ptr(v), _stream())

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None and _lib._lib is not None:
            _lib.lib().CeedMi355BasisDestroy(C.byref(self._h))


class HexOperator:
    """Fused 3D hex operator v = E^T B^T D B E u with a gallery pointwise D.

    qf: QF_MASS, QF_DIFFUSION or QF_DIFFUSION_KAPPA; dtype torch.float64 or
    torch.float32 (FP32 storage and arithmetic for u, v and q_data).
    """

    def __init__(self, qf: int, offsets: np.ndarray, num_elem: int, P: int, Q: int, l_size: int, num_comp: int = 1,
                 comp_stride: int | None = None, dtype=torch.float64, quad_mode: int = GAUSS, scatter: str = "auto"):
        self.qf, self.P, self.Q, self.num_elem, self.num_comp = qf, P, Q, num_elem, num_comp
        self.dtype = dtype
        self.quad_mode = quad_mode
        self.comp_stride = comp_stride if comp_stride is not None else (l_size // num_comp)
        self.l_size = l_size
        self.interp, self.grad, self.q_ref, self.q_weight = lagrange_basis_1d(P, Q, quad_mode)
        off = np.ascontiguousarray(offsets, dtype=np.int32).reshape(-1)
        h = C.c_void_p()
        st = FP64 if dtype == torch.float64 else FP32
        call("CeedMi355OperatorCreateHex", qf, st, num_comp, num_elem, P, Q, _ptr(self.interp), _ptr(self.grad), _ptr(off),
             self.comp_stride, l_size, SCATTER_FLAGS[scatter], C.byref(h))
        self.scatter = scatter
        self._h = h
        self.q_comp = {QF_MASS: 1, QF_DIFFUSION: 6, QF_DIFFUSION_KAPPA: 7}[qf]

    @classmethod
    def from_mesh(cls, qf: int, mesh, Q: int | None = None, num_comp: int = 1, dtype=torch.float64, build: bool = True,
                  scatter: str = "auto"):
        """Operator on a HexMesh with geometric factors built on the GPU from its coordinates."""
        Q = Q if Q is not None else mesh.P + 1
        op = cls(qf, mesh.offsets, mesh.num_elem, mesh.P, Q, num_comp * mesh.num_nodes, num_comp=num_comp,
                 comp_stride=mesh.num_nodes, dtype=dtype, scatter=scatter)
        if build:
            coords = torch.from_numpy(mesh.coords).to("cuda")
            op.build_qdata(mesh.mesh_offsets, coords, mesh.num_vertices, mesh_P=2)
        return op

    def build_qdata(self, mesh_offsets: np.ndarray, coords: torch.Tensor, mesh_comp_stride: int, mesh_P: int = 2) -> None:
        mi, mg, _, mw = lagrange_basis_1d(mesh_P, self.Q, self.quad_mode)
        mo = np.ascontiguousarray(mesh_offsets, dtype=np.int32).reshape(-1)
        coords = _device_tensor(coords, torch.float64)
        build = 0 if self.qf == QF_MASS else 1
        call("CeedMi355OperatorBuildQData", self._h, build, mesh_P, _ptr(mi), _ptr(mg), _ptr(mw), _ptr(mo), mesh_comp_stride,
             _ptr(coords), _stream())

    def set_qdata(self, q_data: torch.Tensor) -> None:
        call("CeedMi355OperatorSetQData", self._h, _ptr(_device_tensor(q_data, self.dtype)), _stream())

    def get_qdata(self) -> torch.Tensor:
        out = torch.empty(self.num_elem * self.q_comp * self.Q ** 3, dtype=self.dtype, device="cuda")
        call("CeedMi355OperatorGetQData", self._h, _ptr(out), _stream())
        return out

    def set_kappa(self, kappa: torch.Tensor) -> None:
        call("CeedMi355OperatorSetKappa", self._h, _ptr(_device_tensor(kappa, torch.float64)), _stream())

    def apply(self, u: torch.Tensor, v: torch.Tensor) -> None:
        self._check(u, v)
        call("CeedMi355OperatorApply", self._h, _ptr(u), _ptr(v), _stream())

    def apply_add(self, u: torch.Tensor, v: torch.Tensor) -> None:
        self._check(u, v)
        call("CeedMi355OperatorApplyAdd", self._h, _ptr(u), _ptr(v), _stream())

    def apply_adjoint(self, u: torch.Tensor, lam: torch.Tensor, v: torch.Tensor, dkappa: torch.Tensor) -> None:
        """v = A(kappa)^T lam and dkappa[e, q] = grad(lam) . S grad(u) (QF_DIFFUSION_KAPPA, FP64)."""
        self._check(u, v)
        self._check(lam, v)
        _device_tensor(dkappa, torch.float64)
        if dkappa.numel() < self.num_elem * self.Q ** 3:
            raise ValueError("dkappa needs num_elem * Q^3 entries")
        call("CeedMi355OperatorApplyAdjoint", self._h, _ptr(u), _ptr(lam), _ptr(v), _ptr(dkappa), _stream())

    def assemble_diagonal(self, diag: torch.Tensor, add: bool = False) -> None:
        """diag (=|+=) the diagonal of the operator."""
        _device_tensor(diag, self.dtype)
        if diag.numel() < self.l_size:
            raise ValueError("diagonal vector too short")
        call("CeedMi355OperatorAssembleDiagonal", self._h, _ptr(diag), int(add), _stream())

    def set_launch(self, elems_per_block: int) -> None:
        call("CeedMi355OperatorSetLaunch", self._h, int(elems_per_block))

    def set_layout(self, layout: int, elems_per_block: int = 0) -> None:
        """Thread layout of the fused kernel (LAYOUT_COLUMNS or LAYOUT_LINES); elems_per_block 0: the default."""
        call("CeedMi355OperatorSetLayout", self._h, int(layout), int(elems_per_block))

    def launch(self) -> tuple[int, int]:
        """(thread layout, elements per workgroup) of the fused kernel."""
        layout, epb = C.c_int32(), C.c_int32()
        call("CeedMi355OperatorGetLaunch", self._h, C.byref(layout), C.byref(epb))
        return layout.value, epb.value

    def autotune(self, reps: int = 5) -> tuple[int, float]:
        """Pick the fastest (thread layout, elements per workgroup) for this plan; returns (epb, ms per apply)."""
        best, ms = C.c_int32(), C.c_double()
        call("CeedMi355OperatorAutotune", self._h, int(reps), C.byref(best), C.byref(ms), _stream())
        return best.value, ms.value

    def scatter_mode(self) -> str:
        """The plan's E^T mode: "tmap", "color", "split", "atomic", "brick" or "segment"."""
        mode = C.c_int32()
        call("CeedMi355OperatorGetScatter", self._h, C.byref(mode))
        return {v: k for k, v in SCATTER_FLAGS.items() if k != "auto"}[mode.value]

    def colors(self) -> list[int]:
        n = C.c_int32()
        call("CeedMi355OperatorGetColors", self._h, C.byref(n), C.c_void_p(0))
        sizes = np.zeros(n.value, dtype=np.int32)
        call("CeedMi355OperatorGetColors", self._h, C.byref(n), _ptr(sizes))
        return sizes.tolist()

    def _check(self, u, v):
        for a in (u, v):
            _device_tensor(a, self.dtype)
            if a.numel() < self.l_size:
                raise ValueError(f"L-vector has {a.numel()} entries, operator needs {self.l_size}")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None and _lib._lib is not None:
            _lib.lib().CeedMi355OperatorDestroy(C.byref(self._h))


class CGStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("replacements", C.c_int32), ("rel_residual", C.c_double),
                ("rel_residual_recursive", C.c_double)]


class CG:
    """Conjugate gradients on the GPU (CeedMi355CGCreate/Solve).

    CG(op64): plain FP64 CG.  CG(op32, op64): FP32 iteration with the residual
    replaced by the FP64 one every ``check_every`` iterations.  set_jacobi(diag)
    preconditions either with the operator diagonal.
    """

    def __init__(self, op: HexOperator, op64: HexOperator | None = None):
        self.op, self.op64 = op, op64
        h = C.c_void_p()
        call("CeedMi355CGCreate", op._h, op64._h if op64 is not None else None, C.byref(h))
        self._h = h

    def set_jacobi(self, diag: torch.Tensor | None) -> None:
        """Jacobi preconditioning with the operator diagonal ``diag`` (FP64, e.g. HexOperator.assemble_diagonal); None: off."""
        if diag is not None:
            _device_tensor(diag, torch.float64)
        self._diag = diag  # kept alive until the inverse is formed
        call("CeedMi355CGSetJacobi", self._h, _ptr(diag) if diag is not None else None, _stream())

    def solve(self, b: torch.Tensor, x: torch.Tensor, max_iter: int = 1000, rtol: float = 1e-10, check_every: int = 25,
              use_graph: bool = True) -> CGStats:
        for a in (b, x):
            _device_tensor(a, torch.float64)
        st = CGStats()
        call("CeedMi355CGSolve", self._h, _ptr(b), _ptr(x), int(max_iter), float(rtol), int(check_every), int(use_graph), _stream(),
             C.byref(st))
        return st

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None and _lib._lib is not None:
            _lib.lib().CeedMi355CGDestroy(C.byref(self._h))
