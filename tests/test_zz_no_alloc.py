"""The allocation-free execute (mofft_execute_async_scratch_*) through the
C ABI, against the workspace execute it replaces for callers that cannot
allocate (mofft_execute_async_workspace_*): same plan, same input, outputs
bit-identical (the kernels are the same; only the workspace - the library's
per-device scratch array instead of a caller buffer - and the chunking
differ).  Two calls in a row on one device, the second with a smaller input,
so the second runs on the scratch contents the first left behind.  A plan
whose single transform needs more workspace than the scratch holds is
rejected with INVALID_ARGUMENT, before anything runs.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20   # MOFFT_SCRATCH_MAX_BYTES (include/mofft_hip.h)


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    return torch


def _scratch_fn(gpu, bits):
    fn = getattr(gpu.lib(), f"mofft_execute_async_scratch_f{bits}")
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return fn


def _run_pair(gpu, torch, n, batch, bits, direction, in_place, seed):
    cdt = torch.complex64 if bits == 32 else torch.complex128
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randn(batch, n, dtype=cdt, device="cuda", generator=g)
    s = torch.cuda.current_stream().cuda_stream
    with gpu.Plan(n, batch, bits, direction, in_place=in_place) as p:
        need = p.workspace_size()
        # the existing entry point, caller workspace
        a_in = x.clone()
        a_out = a_in if in_place else torch.empty_like(x)
        work = torch.empty(max(1, need), dtype=torch.uint8, device="cuda")
        p.execute_async_workspace(a_in, a_out, work if need else None, need, s)
        # the allocation-free one
        b_in = x.clone()
        b_out = b_in if in_place else torch.empty_like(x)
        st = _scratch_fn(gpu, bits)(p.handle, b_in.data_ptr(), b_out.data_ptr(), s)
        torch.cuda.synchronize()
        assert st == gpu.SUCCESS
    return a_out, b_out, need


@pytest.mark.parametrize("first,second", [
    ((1 << 16, 256, 64, -1, True), (1 << 20, 4, 32, 1, True)),
    ((1 << 22, 8, 32, -1, False), (1 << 12, 300, 64, 1, True)),
])
def test_scratch_execute_matches_workspace_execute(gpu, torch_dev, first, second):
    torch = torch_dev
    for case, seed in ((first, 11), (second, 12)):
        n, batch, bits, direction, in_place = case
        a, b, need = _run_pair(gpu, torch, n, batch, bits, direction, in_place, seed)
        assert torch.equal(a, b), (case, need)
    # the first call needed more workspace than the scratch, chunked down
    n, batch, bits, _, _ = first
    assert n * batch * (8 if bits == 32 else 16) > SCRATCH_BYTES


def test_scratch_execute_rejects_oversized_transform(gpu, torch_dev):
    torch = torch_dev
    n = 1 << 23                    # one FP64 transform: 128 MiB of workspace
    x = torch.zeros(n, dtype=torch.complex128, device="cuda")
    with gpu.Plan(n, 1, 64, gpu.FORWARD, in_place=True) as p:
        assert p.workspace_size() > SCRATCH_BYTES
        st = _scratch_fn(gpu, 64)(p.handle, x.data_ptr(), x.data_ptr(),
                                  torch.cuda.current_stream().cuda_stream)
        assert st == gpu.INVALID_ARGUMENT
    with gpu.Plan(1024, 4, 32) as p:        # argument checks
        y = torch.zeros(4, 1024, dtype=torch.complex64, device="cuda")
        assert _scratch_fn(gpu, 32)(p.handle, y.data_ptr(), y.data_ptr(), None) == gpu.INVALID_ARGUMENT
        assert _scratch_fn(gpu, 32)(None, y.data_ptr(), y.data_ptr(), None) == gpu.INVALID_ARGUMENT
