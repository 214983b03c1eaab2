"""CeedMi355VectorReduceNoAlloc (partials in a fixed device array, no
allocation) against CeedMi355VectorReduce (partials in a buffer the library
allocates), both called through the C ABI: every reduction (dot, 1-norm,
max-norm) in FP64 and FP32 gives identical bits -- the same kernels on the
same grid -- on two calls in a row, the second on a smaller input, and an
empty input gives 0."""

import ctypes as C

import pytest
import torch

from libceed_amd import _lib

pytestmark = pytest.mark.gpu
FP64, FP32 = 0, 1  # CEED_MI355_FP64, CEED_MI355_FP32
OPS = {"dot": 0, "abs_sum": 1, "max_abs": 2}


def reduce(name, op, dtype, x, y, out):
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.call(name, op, FP64 if dtype == torch.float64 else FP32, x.numel(), C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
              C.c_void_p(out.data_ptr()), stream)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("op", list(OPS))
def test_no_alloc_reduce_matches_reduce(cuda, op, dtype):
    g = torch.Generator(device="cuda").manual_seed(OPS[op] + (7 if dtype == torch.float32 else 0))
    for n in (3_000_017, 4_099, 1):  # second and third calls on smaller inputs
        print("no-alloc reduce", op, dtype, n, flush=True)
        x = (torch.rand(n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1).to(dtype)
        y = (torch.rand(n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1).to(dtype)
        ref = torch.full((1,), 123.0, dtype=torch.float64, device="cuda")
        got = torch.full((1,), -5.0, dtype=torch.float64, device="cuda")
        reduce("CeedMi355VectorReduce", OPS[op], dtype, x, y, ref)
        reduce("CeedMi355VectorReduceNoAlloc", OPS[op], dtype, x, y, got)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (op, dtype, n, got.item(), ref.item())
        xd = x.double()
        expect = {"dot": (xd * y.double()).sum(), "abs_sum": xd.abs().sum(), "max_abs": xd.abs().max()}[op].item()
        assert abs(got.item() - expect) <= 1e-12 * max(1.0, abs(expect)) * (1 if n < 10 else n ** 0.5)


def test_no_alloc_reduce_empty(cuda):
    x = torch.zeros(1, dtype=torch.float64, device="cuda")
    out = torch.full((1,), 9.0, dtype=torch.float64, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.call("CeedMi355VectorReduceNoAlloc", 0, FP64, 0, C.c_void_p(x.data_ptr()), C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), stream)
    torch.cuda.synchronize()
    assert out.item() == 0.0
