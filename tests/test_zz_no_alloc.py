"""The entry point that allocates no device memory, hip_transpose_0100_noalloc_es, called through
the C ABI: its results must equal hip_transpose_0100_es's on the same inputs, over two calls in a
row (the second smaller, so the shared scratch array holds the first call's bits when the second
starts), and a step whose visited bits exceed the fixed 32 MiB array is refused before anything
runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_pair(hip, orc, es, A, a, B, b, seed):
    import torch
    x = orc.tagged(A * a * B * b, es, seed=seed)
    d1 = torch.from_numpy(x).cuda()
    d2 = d1.clone()
    L = hip.lib()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.hip_transpose_0100_noalloc_es(st, ctypes.c_void_p(d1.data_ptr()), es, A, a, B, b) is False
    route = hip.last_route()
    assert hip.transpose_0100(d2, A, a, B, b, elem_bytes=es) is False
    torch.cuda.synchronize()
    assert torch.equal(d1, d2), (es, A, a, B, b, route)
    assert np.array_equal(d1.cpu().numpy(), orc.op_0100(x, A, a, B, b, es=es))
    return route


def test_noalloc_matches_pool_entry_point(hip, orc):
    # global visited bits (80-byte super-elements, 4 instances), then a smaller step
    r1 = run_pair(hip, orc, 4, 4, 300, 500, 20, seed=1)
    assert r1 == "0100_PTTWAC-", r1
    r2 = run_pair(hip, orc, 2, 2, 129, 77, 33, seed=2)
    assert r2 == "0100_PTTWAC-", r2
    # a single 100 over 4-byte elements: a full transpose on the in-place routes
    run_pair(hip, orc, 4, 1, 1000, 999, 1, seed=3)
    # many small instances: LDS bits, no scratch
    run_pair(hip, orc, 8, 64, 30, 20, 3, seed=4)


def test_noalloc_refuses_oversized_bits(hip):
    import torch
    # 16384 x 16385 super-elements of 64 bytes: 2^28 + 16384 bits > 32 MiB
    R, C = 16384, 16385
    d = torch.empty(R * C * 64, dtype=torch.uint8, device="cuda")
    L = hip.lib()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.hip_transpose_0100_noalloc_es(st, ctypes.c_void_p(d.data_ptr()), 4, 1, R, C, 16) is True
    del d
    torch.cuda.empty_cache()
