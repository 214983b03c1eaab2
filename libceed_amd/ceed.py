"""Python interface to the libCEED C API of ``libceed-mi355.so``.

Class and method names follow the reference's Python bindings
(python/ceed.py:31-642, ceed_vector.py, ceed_elemrestriction.py,
ceed_basis.py, ceed_qfunction.py, ceed_qfunctioncontext.py,
ceed_operator.py) so code and tests written against ``libceed`` read the
same here: ``Ceed(resource)``, ``ceed.Vector(n)``,
``ceed.ElemRestriction(...)``, ``ceed.BasisTensorH1Lagrange(...)``,
``ceed.QFunctionByName(...)``, ``ceed.Operator(qf)``, ``op.set_field(...)``,
``op.apply(u, v)``.  The bindings are ctypes over the C API; every call's
error code is checked and raised with the library's stored message.
"""

from __future__ import annotations

import contextlib
import ctypes as C
import os

import numpy as np

from . import _lib

# ---- constants (include/ceed/types.h enum values) ------------------------
MEM_HOST, MEM_DEVICE = 0, 1
COPY_VALUES, USE_POINTER, OWN_POINTER = 0, 1, 2
NORM_1, NORM_2, NORM_MAX = 0, 1, 2
NOTRANSPOSE, TRANSPOSE = 0, 1
EVAL_NONE, EVAL_INTERP, EVAL_GRAD, EVAL_DIV, EVAL_CURL, EVAL_WEIGHT = 0, 1, 2, 4, 8, 16
GAUSS, GAUSS_LOBATTO = 0, 1
EPSILON = 1e-16

_p = C.c_void_p
_i = C.c_int32
_sz = C.c_int64
_d = C.c_double
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

_API = {
    "CeedInit": (C.c_char_p, C.POINTER(_p)),
    "CeedDestroy": (C.POINTER(_p),),
    "CeedGetResource": (_p, C.POINTER(C.c_char_p)),
    "CeedGetPreferredMemType": (_p, C.POINTER(C.c_int)),
    "CeedSetErrorHandler": (_p, _p),
    "CeedGetErrorMessage": (_p, C.POINTER(C.c_char_p)),
    "CeedResetErrorMessage": (_p, C.POINTER(C.c_char_p)),
    "CeedRegistryGetList": (C.POINTER(C.c_size_t), C.POINTER(C.POINTER(C.c_char_p)), C.POINTER(_ip)),
    "CeedIsDeterministic": (_p, C.POINTER(C.c_bool)),
    "CeedVectorCreate": (_p, _sz, C.POINTER(_p)),
    "CeedVectorDestroy": (C.POINTER(_p),),
    "CeedVectorSetArray": (_p, C.c_int, C.c_int, _p),
    "CeedVectorTakeArray": (_p, C.c_int, C.POINTER(_p)),
    "CeedVectorSyncArray": (_p, C.c_int),
    "CeedVectorSetValue": (_p, _d),
    "CeedVectorGetArray": (_p, C.c_int, C.POINTER(_p)),
    "CeedVectorGetArrayRead": (_p, C.c_int, C.POINTER(_p)),
    "CeedVectorGetArrayWrite": (_p, C.c_int, C.POINTER(_p)),
    "CeedVectorRestoreArray": (_p, C.POINTER(_p)),
    "CeedVectorRestoreArrayRead": (_p, C.POINTER(_p)),
    "CeedVectorGetLength": (_p, C.POINTER(_sz)),
    "CeedVectorNorm": (_p, C.c_int, _dp),
    "CeedVectorScale": (_p, _d),
    "CeedVectorAXPY": (_p, _d, _p),
    "CeedVectorAXPBY": (_p, _d, _d, _p),
    "CeedVectorPointwiseMult": (_p, _p, _p),
    "CeedVectorReciprocal": (_p,),
    "CeedVectorCopy": (_p, _p),
    "CeedElemRestrictionCreate": (_p, _i, _i, _i, _i, _sz, C.c_int, C.c_int, _ip, C.POINTER(_p)),
    "CeedElemRestrictionCreateStrided": (_p, _i, _i, _i, _sz, _ip, C.POINTER(_p)),
    "CeedElemRestrictionApply": (_p, C.c_int, _p, _p, _p),
    "CeedElemRestrictionCreateVector": (_p, C.POINTER(_p), C.POINTER(_p)),
    "CeedElemRestrictionGetMultiplicity": (_p, _p),
    "CeedElemRestrictionDestroy": (C.POINTER(_p),),
    "CeedBasisCreateTensorH1Lagrange": (_p, _i, _i, _i, _i, C.c_int, C.POINTER(_p)),
    "CeedBasisCreateTensorH1": (_p, _i, _i, _i, _i, _dp, _dp, _dp, _dp, C.POINTER(_p)),
    "CeedBasisApply": (_p, _i, C.c_int, C.c_int, _p, _p),
    "CeedBasisApplyAdd": (_p, _i, C.c_int, C.c_int, _p, _p),
    "CeedBasisGetNumNodes": (_p, _ip),
    "CeedBasisGetNumQuadraturePoints": (_p, _ip),
    "CeedBasisGetInterp1D": (_p, C.POINTER(_dp)),
    "CeedBasisGetGrad1D": (_p, C.POINTER(_dp)),
    "CeedBasisGetQRef": (_p, C.POINTER(_dp)),
    "CeedBasisGetQWeights": (_p, C.POINTER(_dp)),
    "CeedBasisGetNumQuadraturePoints1D": (_p, _ip),
    "CeedBasisGetNumNodes1D": (_p, _ip),
    "CeedBasisDestroy": (C.POINTER(_p),),
    "CeedQFunctionCreateInterior": (_p, _i, _p, C.c_char_p, C.POINTER(_p)),
    "CeedQFunctionCreateInteriorByName": (_p, C.c_char_p, C.POINTER(_p)),
    "CeedQFunctionAddInput": (_p, C.c_char_p, _i, C.c_int),
    "CeedQFunctionAddOutput": (_p, C.c_char_p, _i, C.c_int),
    "CeedQFunctionSetContext": (_p, _p),
    "CeedQFunctionApply": (_p, _i, C.POINTER(_p), C.POINTER(_p)),
    "CeedQFunctionDestroy": (C.POINTER(_p),),
    "CeedQFunctionContextCreate": (_p, C.POINTER(_p)),
    "CeedQFunctionContextSetData": (_p, C.c_int, C.c_int, C.c_size_t, _p),
    "CeedQFunctionContextGetData": (_p, C.c_int, C.POINTER(_p)),
    "CeedQFunctionContextRestoreData": (_p, C.POINTER(_p)),
    "CeedQFunctionContextGetContextSize": (_p, C.POINTER(C.c_size_t)),
    "CeedQFunctionContextDestroy": (C.POINTER(_p),),
    "CeedOperatorCreate": (_p, _p, _p, _p, C.POINTER(_p)),
    "CeedOperatorSetField": (_p, C.c_char_p, _p, _p, _p),
    "CeedOperatorCheckReady": (_p,),
    "CeedOperatorApply": (_p, _p, _p, _p),
    "CeedOperatorApplyAdd": (_p, _p, _p, _p),
    "CeedOperatorDestroy": (C.POINTER(_p),),
    "CeedOperatorCreateComposite": (_p, C.POINTER(_p)),
    "CeedOperatorCompositeAddSub": (_p, _p),
    "CeedOperatorLinearAssembleDiagonal": (_p, _p, _p),
    "CeedOperatorLinearAssembleAddDiagonal": (_p, _p, _p),
}


def _api():
    lib = _lib.lib()
    if not getattr(lib, "_ceed_api_bound", False):
        for name, args in _API.items():
            fn = getattr(lib, name)
            fn.argtypes = list(args)
            fn.restype = C.c_int
        lib._ceed_api_bound = True
    return lib


def _special(name: str) -> _p:
    return _p.in_dll(_api(), name)


class CeedError(RuntimeError):
    pass


def _ptr(x) -> _p:
    return _p(x) if isinstance(x, int) else x


# ---------------------------------------------------------------------------
class Ceed:
    """A library context on ``resource`` (python/ceed.py:31)."""

    def __init__(self, resource: str = "/gpu/hip/mi355", on_error: str = "store"):
        lib = _api()
        self._lib = lib
        self._pointer = _p()
        # errors during initialisation are returned, not aborted on (python/ceed.py:37)
        os.environ["CEED_ERROR_HANDLER"] = "return"
        err = lib.CeedInit(resource.encode(), C.byref(self._pointer))
        if err:
            raise CeedError(f"CeedInit({resource!r}) failed with code {err}")
        if on_error == "store":
            lib.CeedSetErrorHandler(self._pointer, C.cast(lib.CeedErrorStore, _p))

    def _check_error(self, err: int) -> None:
        if err:
            msg = C.c_char_p()
            self._lib.CeedGetErrorMessage(self._pointer, C.byref(msg))
            text = (msg.value or b"").decode(errors="replace")
            self._lib.CeedResetErrorMessage(self._pointer, C.byref(msg))
            raise CeedError(text or f"libCEED error {err}")

    def _call(self, name: str, *args) -> None:
        self._check_error(getattr(self._lib, name)(*args))

    def __del__(self):
        if getattr(self, "_pointer", None) and self._pointer.value:
            self._lib.CeedDestroy(C.byref(self._pointer))

    def get_resource(self) -> str:
        r = C.c_char_p()
        self._call("CeedGetResource", self._pointer, C.byref(r))
        return r.value.decode()

    def get_preferred_memtype(self) -> int:
        m = C.c_int()
        self._call("CeedGetPreferredMemType", self._pointer, C.byref(m))
        return m.value

    def is_deterministic(self) -> bool:
        b = C.c_bool()
        self._call("CeedIsDeterministic", self._pointer, C.byref(b))
        return bool(b.value)

    def gauss_quadrature(self, q: int):
        x, w = np.zeros(q), np.zeros(q)
        self._call("CeedGaussQuadrature", q, x.ctypes.data, w.ctypes.data)
        return x, w

    def lobatto_quadrature(self, q: int):
        x, w = np.zeros(q), np.zeros(q)
        self._call("CeedLobattoQuadrature", q, x.ctypes.data, w.ctypes.data)
        return x, w

    # factories (python/ceed.py:184-642)
    def Vector(self, size: int) -> "Vector":
        return Vector(self, size)

    def ElemRestriction(self, nelem, elemsize, ncomp, compstride, lsize, offsets, memtype=MEM_HOST, cmode=COPY_VALUES):
        return ElemRestriction(self, nelem, elemsize, ncomp, compstride, lsize, offsets, memtype, cmode)

    def StridedElemRestriction(self, nelem, elemsize, ncomp, lsize, strides):
        return StridedElemRestriction(self, nelem, elemsize, ncomp, lsize, strides)

    def BasisTensorH1(self, dim, ncomp, P1d, Q1d, interp1d, grad1d, qref1d, qweight1d):
        return BasisTensorH1(self, dim, ncomp, P1d, Q1d, interp1d, grad1d, qref1d, qweight1d)

    def BasisTensorH1Lagrange(self, dim, ncomp, P, Q, qmode):
        return BasisTensorH1Lagrange(self, dim, ncomp, P, Q, qmode)

    def QFunction(self, vlength, f, source):
        return QFunction(self, vlength, f, source)

    def QFunctionByName(self, name):
        return QFunctionByName(self, name)

    def QFunctionContext(self):
        return QFunctionContext(self)

    def Operator(self, qf, dqf=None, dqfT=None):
        return Operator(self, qf, dqf, dqfT)

    def CompositeOperator(self):
        return CompositeOperator(self)


def registry() -> list[str]:
    """Resource prefixes registered in the library (CeedRegistryGetList)."""
    lib = _api()
    n = C.c_size_t()
    names = C.POINTER(C.c_char_p)()
    prio = _ip()
    err = lib.CeedRegistryGetList(C.byref(n), C.byref(names), C.byref(prio))
    if err:
        raise CeedError(f"CeedRegistryGetList failed with code {err}")
    return [names[i].decode() for i in range(n.value)]


# ---------------------------------------------------------------------------
class _Object:
    _destroy = ""

    def __del__(self):
        if getattr(self, "_pointer", None) and self._pointer.value:
            getattr(self._ceed._lib, self._destroy)(C.byref(self._pointer))


class Vector(_Object):
    """CeedVector (python/ceed_vector.py:21)."""

    _destroy = "CeedVectorDestroy"

    def __init__(self, ceed: Ceed, size: int):
        self._ceed = ceed
        self._pointer = _p()
        self._keep = None
        ceed._call("CeedVectorCreate", ceed._pointer, int(size), C.byref(self._pointer))

    def __len__(self) -> int:
        return self.get_length()

    def get_length(self) -> int:
        n = _sz()
        self._ceed._call("CeedVectorGetLength", self._pointer, C.byref(n))
        return n.value

    def set_array(self, array, memtype=MEM_HOST, cmode=COPY_VALUES):
        """Host numpy array, or a device pointer (int) with memtype=MEM_DEVICE."""
        if memtype == MEM_HOST:
            a = np.ascontiguousarray(array, dtype=np.float64)
            if cmode == USE_POINTER:
                self._keep = a
            ptr = a.ctypes.data_as(_p)
        else:
            ptr = _p(int(array))
        self._ceed._call("CeedVectorSetArray", self._pointer, memtype, cmode, ptr)

    def set_value(self, value: float):
        self._ceed._call("CeedVectorSetValue", self._pointer, float(value))

    def sync_array(self, memtype=MEM_HOST):
        self._ceed._call("CeedVectorSyncArray", self._pointer, memtype)

    def copy_from(self, vec_source: "Vector"):
        self._ceed._call("CeedVectorCopy", vec_source._pointer, self._pointer)

    @contextlib.contextmanager
    def array_read(self, *shape, memtype=MEM_HOST):
        p = _p()
        self._ceed._call("CeedVectorGetArrayRead", self._pointer, memtype, C.byref(p))
        try:
            if memtype == MEM_HOST:
                a = np.ctypeslib.as_array(C.cast(p, _dp), shape=(self.get_length(),))
                yield a.reshape(shape) if shape else a
            else:
                yield p.value
        finally:
            self._ceed._call("CeedVectorRestoreArrayRead", self._pointer, C.byref(p))

    @contextlib.contextmanager
    def array(self, *shape, memtype=MEM_HOST):
        p = _p()
        self._ceed._call("CeedVectorGetArray", self._pointer, memtype, C.byref(p))
        try:
            if memtype == MEM_HOST:
                a = np.ctypeslib.as_array(C.cast(p, _dp), shape=(self.get_length(),))
                yield a.reshape(shape) if shape else a
            else:
                yield p.value
        finally:
            self._ceed._call("CeedVectorRestoreArray", self._pointer, C.byref(p))

    @contextlib.contextmanager
    def array_write(self, *shape, memtype=MEM_HOST):
        p = _p()
        self._ceed._call("CeedVectorGetArrayWrite", self._pointer, memtype, C.byref(p))
        try:
            if memtype == MEM_HOST:
                a = np.ctypeslib.as_array(C.cast(p, _dp), shape=(self.get_length(),))
                yield a.reshape(shape) if shape else a
            else:
                yield p.value
        finally:
            self._ceed._call("CeedVectorRestoreArray", self._pointer, C.byref(p))

    def to_numpy(self) -> np.ndarray:
        with self.array_read() as a:
            return np.array(a, copy=True)

    def norm(self, normtype=NORM_2) -> float:
        r = _d()
        self._ceed._call("CeedVectorNorm", self._pointer, normtype, C.byref(r))
        return r.value

    def scale(self, alpha):
        self._ceed._call("CeedVectorScale", self._pointer, float(alpha))

    def axpy(self, alpha, x: "Vector"):
        self._ceed._call("CeedVectorAXPY", self._pointer, float(alpha), x._pointer)

    def axpby(self, alpha, beta, x: "Vector"):
        self._ceed._call("CeedVectorAXPBY", self._pointer, float(alpha), float(beta), x._pointer)

    def pointwise_mult(self, x: "Vector", y: "Vector"):
        self._ceed._call("CeedVectorPointwiseMult", self._pointer, x._pointer, y._pointer)

    def reciprocal(self):
        self._ceed._call("CeedVectorReciprocal", self._pointer)


class _SpecialVector:
    def __init__(self, name):
        self._name = name

    @property
    def _pointer(self):
        return _special(self._name)


VECTOR_ACTIVE = _SpecialVector("CEED_VECTOR_ACTIVE")
VECTOR_NONE = _SpecialVector("CEED_VECTOR_NONE")


class _SpecialHandle:
    def __init__(self, name):
        self._name = name

    @property
    def _pointer(self):
        return _special(self._name)


BASIS_NONE = _SpecialHandle("CEED_BASIS_NONE")
ELEMRESTRICTION_NONE = _SpecialHandle("CEED_ELEMRESTRICTION_NONE")
STRIDES_BACKEND = "backend"


class ElemRestriction(_Object):
    """Offset restriction (python/ceed_elemrestriction.py:176)."""

    _destroy = "CeedElemRestrictionDestroy"

    def __init__(self, ceed, nelem, elemsize, ncomp, compstride, lsize, offsets, memtype=MEM_HOST, cmode=COPY_VALUES):
        self._ceed = ceed
        self._pointer = _p()
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        self._offsets = off
        ceed._call("CeedElemRestrictionCreate", ceed._pointer, nelem, elemsize, ncomp, compstride, int(lsize), memtype, COPY_VALUES,
                   off.ctypes.data_as(_ip), C.byref(self._pointer))

    def apply(self, u: Vector, v: Vector, tmode=NOTRANSPOSE):
        self._ceed._call("CeedElemRestrictionApply", self._pointer, tmode, u._pointer, v._pointer, None)

    def create_vector(self, createLvec=True, createEvec=True):
        lp, ep = _p(), _p()
        self._ceed._call("CeedElemRestrictionCreateVector", self._pointer, C.byref(lp) if createLvec else None,
                         C.byref(ep) if createEvec else None)
        out = []
        for ptr, want in ((lp, createLvec), (ep, createEvec)):
            if want:
                v = Vector.__new__(Vector)
                v._ceed, v._pointer, v._keep = self._ceed, ptr, None
                out.append(v)
            else:
                out.append(None)
        return out

    def get_multiplicity(self) -> Vector:
        lvec, _ = self.create_vector(createEvec=False)
        lvec.set_value(0.0)
        self._ceed._call("CeedElemRestrictionGetMultiplicity", self._pointer, lvec._pointer)
        return lvec


class StridedElemRestriction(ElemRestriction):
    """Strided restriction (python/ceed_elemrestriction.py:288); strides may
    be STRIDES_BACKEND."""

    def __init__(self, ceed, nelem, elemsize, ncomp, lsize, strides):
        self._ceed = ceed
        self._pointer = _p()
        if isinstance(strides, str):  # the library's CEED_STRIDES_BACKEND array itself
            st = C.cast(C.addressof(_i.in_dll(_api(), "CEED_STRIDES_BACKEND")), _ip)
        else:
            self._strides = np.ascontiguousarray(strides, dtype=np.int32)
            st = self._strides.ctypes.data_as(_ip)
        ceed._call("CeedElemRestrictionCreateStrided", ceed._pointer, nelem, elemsize, ncomp, int(lsize), st, C.byref(self._pointer))


class Basis(_Object):
    _destroy = "CeedBasisDestroy"

    def apply(self, nelem, emode, u, v, tmode=NOTRANSPOSE):
        up = VECTOR_NONE._pointer if u is None else u._pointer
        self._ceed._call("CeedBasisApply", self._pointer, nelem, tmode, emode, up, v._pointer)

    def apply_add(self, nelem, emode, u, v, tmode=NOTRANSPOSE):
        up = VECTOR_NONE._pointer if u is None else u._pointer
        self._ceed._call("CeedBasisApplyAdd", self._pointer, nelem, tmode, emode, up, v._pointer)

    def get_num_nodes(self) -> int:
        n = _i()
        self._ceed._call("CeedBasisGetNumNodes", self._pointer, C.byref(n))
        return n.value

    def get_num_quadrature_points(self) -> int:
        n = _i()
        self._ceed._call("CeedBasisGetNumQuadraturePoints", self._pointer, C.byref(n))
        return n.value

    def _get_1d(self, getter, rows_fn):
        p = _dp()
        self._ceed._call(getter, self._pointer, C.byref(p))
        P1, Q1 = _i(), _i()
        self._ceed._call("CeedBasisGetNumNodes1D", self._pointer, C.byref(P1))
        self._ceed._call("CeedBasisGetNumQuadraturePoints1D", self._pointer, C.byref(Q1))
        shape = rows_fn(P1.value, Q1.value)
        return np.array(np.ctypeslib.as_array(p, shape=(int(np.prod(shape)),)).reshape(shape), copy=True)

    def get_interp_1d(self):
        return self._get_1d("CeedBasisGetInterp1D", lambda P, Q: (Q, P))

    def get_grad_1d(self):
        return self._get_1d("CeedBasisGetGrad1D", lambda P, Q: (Q, P))

    def get_q_weight_1d(self):
        return self._get_1d("CeedBasisGetQWeights", lambda P, Q: (Q,))

    def get_q_ref_1d(self):
        return self._get_1d("CeedBasisGetQRef", lambda P, Q: (Q,))


class BasisTensorH1(Basis):
    """python/ceed_basis.py:122"""

    def __init__(self, ceed, dim, ncomp, P1d, Q1d, interp1d, grad1d, qref1d, qweight1d):
        self._ceed = ceed
        self._pointer = _p()
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (interp1d, grad1d, qref1d, qweight1d)]
        self._arrs = arrs
        ceed._call("CeedBasisCreateTensorH1", ceed._pointer, dim, ncomp, P1d, Q1d, *[a.ctypes.data_as(_dp) for a in arrs],
                   C.byref(self._pointer))


class BasisTensorH1Lagrange(Basis):
    """python/ceed_basis.py:286"""

    def __init__(self, ceed, dim, ncomp, P, Q, qmode):
        self._ceed = ceed
        self._pointer = _p()
        ceed._call("CeedBasisCreateTensorH1Lagrange", ceed._pointer, dim, ncomp, P, Q, qmode, C.byref(self._pointer))


class QFunction(_Object):
    """User QFunction from ``source`` = "path/to/file.h:Name"
    (python/ceed_qfunction.py:79).  ``f`` is the host function pointer
    (int or None); the device code is compiled from the source file."""

    _destroy = "CeedQFunctionDestroy"

    def __init__(self, ceed, vlength, f, source):
        self._ceed = ceed
        self._pointer = _p()
        if ceed is not None:
            ceed._call("CeedQFunctionCreateInterior", ceed._pointer, vlength, _p(f) if f else None, source.encode(),
                       C.byref(self._pointer))

    def add_input(self, fieldname, size, emode):
        self._ceed._call("CeedQFunctionAddInput", self._pointer, fieldname.encode(), size, emode)

    def add_output(self, fieldname, size, emode):
        self._ceed._call("CeedQFunctionAddOutput", self._pointer, fieldname.encode(), size, emode)

    def set_context(self, ctx: "QFunctionContext"):
        self._ceed._call("CeedQFunctionSetContext", self._pointer, ctx._pointer)

    def apply(self, q, inputs, outputs):
        ins = (_p * max(1, len(inputs)))(*[v._pointer for v in inputs])
        outs = (_p * max(1, len(outputs)))(*[v._pointer for v in outputs])
        self._ceed._call("CeedQFunctionApply", self._pointer, q, ins, outs)


class QFunctionByName(QFunction):
    """Gallery QFunction (python/ceed_qfunction.py:153)."""

    def __init__(self, ceed, name):
        super().__init__(None, 0, None, "")
        self._ceed = ceed
        ceed._call("CeedQFunctionCreateInteriorByName", ceed._pointer, name.encode(), C.byref(self._pointer))


class QFunctionContext(_Object):
    """python/ceed_qfunctioncontext.py:21"""

    _destroy = "CeedQFunctionContextDestroy"

    def __init__(self, ceed):
        self._ceed = ceed
        self._pointer = _p()
        ceed._call("CeedQFunctionContextCreate", ceed._pointer, C.byref(self._pointer))

    def set_data(self, data, memtype=MEM_HOST, cmode=COPY_VALUES):
        a = np.ascontiguousarray(data)
        self._keep = a
        ceed = self._ceed
        ceed._call("CeedQFunctionContextSetData", self._pointer, memtype, cmode, a.nbytes, a.ctypes.data_as(_p))

    @contextlib.contextmanager
    def data(self, dtype=np.float64):
        p = _p()
        size = C.c_size_t()
        self._ceed._call("CeedQFunctionContextGetContextSize", self._pointer, C.byref(size))
        self._ceed._call("CeedQFunctionContextGetData", self._pointer, MEM_HOST, C.byref(p))
        try:
            yield np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(size.value,)).view(dtype)
        finally:
            self._ceed._call("CeedQFunctionContextRestoreData", self._pointer, C.byref(p))


class Operator(_Object):
    """python/ceed_operator.py:285"""

    _destroy = "CeedOperatorDestroy"

    def __init__(self, ceed, qf, dqf=None, dqfT=None):
        self._ceed = ceed
        self._pointer = _p()
        self._refs = [qf]
        ceed._call("CeedOperatorCreate", ceed._pointer, qf._pointer, None, None, C.byref(self._pointer))

    def set_field(self, fieldname, restriction, basis, vector):
        self._refs += [restriction, basis, vector]
        self._ceed._call("CeedOperatorSetField", self._pointer, fieldname.encode(), restriction._pointer, basis._pointer, vector._pointer)

    def check(self):
        self._ceed._call("CeedOperatorCheckReady", self._pointer)

    def apply(self, u, v):
        up = VECTOR_NONE._pointer if u is None else u._pointer
        self._ceed._call("CeedOperatorApply", self._pointer, up, v._pointer, None)

    def apply_add(self, u, v):
        up = VECTOR_NONE._pointer if u is None else u._pointer
        self._ceed._call("CeedOperatorApplyAdd", self._pointer, up, v._pointer, None)

    def linear_assemble_diagonal(self, d):
        """python/ceed_operator.py:54"""
        self._ceed._call("CeedOperatorLinearAssembleDiagonal", self._pointer, d._pointer, None)

    def linear_assemble_add_diagonal(self, d):
        self._ceed._call("CeedOperatorLinearAssembleAddDiagonal", self._pointer, d._pointer, None)


class CompositeOperator(Operator):
    """Sum of sub-operators (python/ceed_operator.py:327)."""

    def __init__(self, ceed):
        self._ceed = ceed
        self._pointer = _p()
        self._refs = []
        ceed._call("CeedOperatorCreateComposite", ceed._pointer, C.byref(self._pointer))

    def add_sub(self, subop: Operator):
        self._refs.append(subop)
        self._ceed._call("CeedOperatorCompositeAddSub", self._pointer, subop._pointer)
