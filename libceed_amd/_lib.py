"""ctypes binding of ``lib/libceed-mi355.so`` (the C ABI in include/ceed-mi355-kernels.h).

There is no fallback: if the library is missing or a call fails, an exception
is raised.  Build the library with ``python -m libceed_amd.build``.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libceed-mi355.so"

c_int, c_i32, c_i64, c_sz = C.c_int, C.c_int32, C.c_int64, C.c_size_t
vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)

# name -> (restype, argtypes); every function declared in include/ceed-mi355*.h
SIGNATURES = {
    "CeedMi355GetLastError": (C.c_char_p, []),
    "CeedMi355DeviceInit": (c_int, [c_int]),
    "CeedMi355Malloc": (c_int, [c_sz, C.POINTER(vp)]),
    "CeedMi355Free": (c_int, [vp]),
    "CeedMi355Memcpy": (c_int, [vp, vp, c_sz, vp]),
    "CeedMi355Memset": (c_int, [vp, c_int, c_sz, vp]),
    "CeedMi355StreamSynchronize": (c_int, [vp]),
    "CeedMi355RestrictionCreateOffset": (c_int, [c_i32, c_i32, c_i32, c_i32, c_i64, vp, C.POINTER(vp)]),
    "CeedMi355RestrictionCreateStrided": (c_int, [c_i32, c_i32, c_i32, c_i64, vp, C.POINTER(vp)]),
    "CeedMi355RestrictionApply": (c_int, [vp, c_int, vp, vp, vp]),
    "CeedMi355RestrictionDestroy": (c_int, [C.POINTER(vp)]),
    "CeedMi355BasisCreateTensorH1": (c_int, [c_i32, c_i32, c_i32, c_i32, vp, vp, vp, C.POINTER(vp)]),
    "CeedMi355BasisApply": (c_int, [vp, c_i32, c_int, c_int, c_int, vp, vp, vp]),
    "CeedMi355BasisDestroy": (c_int, [C.POINTER(vp)]),
    "CeedMi355OperatorCreateHex": (c_int, [c_int, c_int, c_i32, c_i32, c_i32, c_i32, vp, vp, vp, c_i32, c_i64, c_int, C.POINTER(vp)]),
    "CeedMi355OperatorCreateHexJit": (c_int, [C.c_char_p, C.c_char_p, c_i32, vp, c_i32, vp, c_i32, c_i32, c_i32, c_i32, vp, vp, vp, c_i64, c_int,
                                              C.POINTER(vp)]),
    "CeedMi355OperatorSetContext": (c_int, [vp, vp]),
    "CeedMi355OperatorBuildQData": (c_int, [vp, c_int, c_i32, vp, vp, vp, vp, c_i32, vp, vp]),
    "CeedMi355OperatorSetQData": (c_int, [vp, vp, vp]),
    "CeedMi355OperatorGetQData": (c_int, [vp, vp, vp]),
    "CeedMi355OperatorSetKappa": (c_int, [vp, vp, vp]),
    "CeedMi355OperatorApplyAdd": (c_int, [vp, vp, vp, vp]),
    "CeedMi355OperatorApplyAdjoint": (c_int, [vp, vp, vp, vp, vp, vp]),
    "CeedMi355OperatorAssembleDiagonal": (c_int, [vp, vp, c_int, vp]),
    "CeedMi355OperatorApply": (c_int, [vp, vp, vp, vp]),
    "CeedMi355OperatorGetColors": (c_int, [vp, C.POINTER(c_i32), vp]),
    "CeedMi355OperatorGetScatter": (c_int, [vp, C.POINTER(c_i32)]),
    "CeedMi355OperatorDestroy": (c_int, [C.POINTER(vp)]),
    "CeedGaussQuadrature": (c_int, [c_i32, vp, vp]),
    "CeedLobattoQuadrature": (c_int, [c_i32, vp, vp]),
    "CeedMi355LagrangeBasis1D": (c_int, [c_i32, c_i32, c_i32, vp, vp, vp, vp]),
    "CeedMi355VectorOp": (c_int, [c_int, c_int, c_i64, C.c_double, C.c_double, vp, vp, vp, vp]),
    "CeedMi355VectorReduce": (c_int, [c_int, c_int, c_i64, vp, vp, vp, vp]),
    "CeedMi355VectorReduceNoAlloc": (c_int, [c_int, c_int, c_i64, vp, vp, vp, vp]),
    "CeedMi355VectorConvert": (c_int, [c_i64, c_int, vp, c_int, vp, vp]),
    "CeedMi355QFunctionCompile": (c_int, [C.c_char_p, C.c_char_p, c_i32, vp, c_i32, vp, C.POINTER(vp)]),
    "CeedMi355QFunctionApply": (c_int, [vp, c_i64, c_i32, vp, vp, vp, vp]),
    "CeedMi355QFunctionDestroy": (c_int, [C.POINTER(vp)]),
    "CeedMi355CGCreate": (c_int, [vp, vp, C.POINTER(vp)]),
    "CeedMi355CGSolve": (c_int, [vp, vp, vp, c_i32, C.c_double, c_i32, c_int, vp, vp]),
    "CeedMi355CGSetJacobi": (c_int, [vp, vp, vp]),
    "CeedMi355CGDestroy": (c_int, [C.POINTER(vp)]),
    "CeedMi355BatchCGCreate": (c_int, [c_i32, vp, vp, c_i32, vp, C.POINTER(vp)]),
    "CeedMi355BatchCGSolve": (c_int, [vp, vp, vp, c_i32, C.c_double, c_i32, c_int, vp, vp, vp, vp, vp, vp]),
    "CeedMi355BatchCGDestroy": (c_int, [C.POINTER(vp)]),
    "CeedMi355BatchSmallCreate": (c_int, [c_i32, vp, vp, c_i32, vp, C.POINTER(vp)]),
    "CeedMi355BatchSmallSolve": (c_int, [vp, vp, vp, c_i32, C.c_double, vp, vp, vp, vp]),
    "CeedMi355BatchSmallDestroy": (c_int, [C.POINTER(vp)]),
    "CeedMi355OperatorSetLaunch": (c_int, [vp, c_i32]),
    "CeedMi355OperatorAutotune": (c_int, [vp, c_i32, vp, vp, vp]),
    "CeedMi355OperatorSetLayout": (c_int, [vp, c_i32, c_i32]),
    "CeedMi355OperatorGetLaunch": (c_int, [vp, vp, vp]),
    "CeedRegister_Mi355": (c_int, []),
    "CeedInit_Mi355": (c_int, [C.c_char_p, vp]),
}


class CeedMi355Error(RuntimeError):
    pass


_lib = None


def lib() -> C.CDLL:
    """Load the kernel library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise CeedMi355Error(f"{LIB_PATH} not found: build it with `python -m libceed_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(ierr: int, what: str = "") -> None:
    if ierr != 0:
        msg = lib().CeedMi355GetLastError().decode(errors="replace")
        raise CeedMi355Error(f"{what} failed with code {ierr}: {msg}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
