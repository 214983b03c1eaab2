"""CPU checks of the product library: it loads, exports every function that
include/*.h declares, and its host-side math (quadrature, 1D Lagrange
matrices) matches the pinned oracle bit for bit.  No GPU compute is called."""

import re
from pathlib import Path

import numpy as np
import pytest

import oracle
from libceed_amd import _lib, ops

INCLUDE = Path(__file__).resolve().parent.parent / "include"


def declared_functions():
    names = set()
    for h in INCLUDE.glob("ceed-mi355*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w]*\s*\**\s*(Ceed\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert "CeedMi355OperatorApply" in names and "CeedMi355RestrictionApply" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(declared_functions()) <= set(_lib.SIGNATURES), set(declared_functions()) - set(_lib.SIGNATURES)


@pytest.mark.parametrize("P", range(2, 11))
def test_host_basis_matches_oracle(P):
    for Q in (P - 1, P, P + 1, P + 2):
        for mode in (ops.GAUSS, ops.GAUSS_LOBATTO):
            if Q < 1 or (mode == ops.GAUSS_LOBATTO and Q < 2):
                continue
            for a, b in zip(ops.lagrange_basis_1d(P, Q, mode), oracle.lagrange_1d(P, Q, mode)):
                assert np.array_equal(a, b)


def test_host_quadrature_matches_oracle():
    for Q in range(1, 12):
        assert all(np.array_equal(a, b) for a, b in zip(ops.gauss_quadrature(Q), oracle.gauss(Q)))
    for Q in range(2, 12):
        assert all(np.array_equal(a, b) for a, b in zip(ops.lobatto_quadrature(Q), oracle.lobatto(Q)))


def test_error_reporting():
    with pytest.raises(_lib.CeedMi355Error, match="CeedGaussQuadrature"):
        ops.gauss_quadrature(0)


# ------------------------------------------------------------ libCEED API --
REFERENCE_INCLUDE = Path("/root/reference/include")
HOST = Path(__file__).resolve().parent.parent / "libceed_amd" / "host"


def test_registry_lists_backend():
    from libceed_amd import ceed as libceed

    assert "/gpu/hip/mi355" in libceed.registry()


def test_init_without_gpu_fails_loudly():
    """No CPU fallback: on a machine without a GPU, CeedInit of the backend is an error."""
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    from libceed_amd import ceed as libceed

    with pytest.raises(libceed.CeedError):
        libceed.Ceed("/gpu/hip/mi355")
    with pytest.raises(libceed.CeedError):
        libceed.Ceed("/cpu/self/ref/serial")  # not provided by this library


@pytest.mark.skipif(not REFERENCE_INCLUDE.exists(), reason="reference headers not present")
@pytest.mark.parametrize("src", ["ceed-mi355-objects.c", "ceed-mi355-operator.c"])
def test_backend_compiles_against_reference_headers(tmp_path, src):
    """The backend is written to the reference's backend API: it compiles
    unchanged against the reference's include/ceed/{ceed,backend}.h."""
    import shutil
    import subprocess

    for h in ("ceed-mi355.h", "ceed-mi355-kernels.h"):
        shutil.copy(INCLUDE / h, tmp_path / h)
    res = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-fsyntax-only", f"-I{REFERENCE_INCLUDE}", f"-I{tmp_path}", f"-I{HOST}",
                          str(HOST / src)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
