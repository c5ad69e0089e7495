"""Full-size in-place transposes that would not fit out of place twice over: 98304 x 131071
float64 (103 GB) and 250007 x 199999 uint8 (50 GB, both dimensions prime).

Size-independent checks: a sample of a million (row, col) positions must land at (col, row),
and transposing back must restore the buffer exactly (compared against a device copy).
"""
import sys
import time

import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.large]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("M,N,es", [(98304, 131071, 8), (250007, 199999, 1)])
def test_full_size_prime_dimension_transpose(hip, M, N, es):
    import torch
    need = 2 * M * N * es + (M * N) // 8 + (16 << 30)
    free, total = torch.cuda.mem_get_info()
    if free < need:
        pytest.skip(f"needs {need / 1e9:.0f} GB of device memory, {free / 1e9:.0f} GB free")
    dt = torch.float64 if es == 8 else torch.uint8
    g = torch.Generator(device="cuda").manual_seed(M ^ N)
    n = M * N
    if es == 8:
        x = torch.empty(n, dtype=dt, device="cuda")
        for s in range(0, n, 1 << 30):  # fill in 8 GiB slices
            e = min(n, s + (1 << 30))
            x[s:e].copy_(torch.rand(e - s, dtype=dt, device="cuda", generator=g))
    else:
        x = torch.empty(n, dtype=dt, device="cuda")
        for s in range(0, n, 1 << 32):
            e = min(n, s + (1 << 32))
            x[s:e].copy_(torch.randint(0, 256, (e - s,), dtype=torch.uint8, device="cuda",
                                       generator=g))
    orig = x.clone()
    r = torch.randint(0, M, (1 << 20,), device="cuda", generator=g)
    c = torch.randint(0, N, (1 << 20,), device="cuda", generator=g)
    before = orig[r * N + c]
    torch.cuda.synchronize()
    say("transpose", M, N, es, hip.plan_transpose(M, N, es))
    t0 = time.time()
    assert hip.transpose_full(x, M, N) is False
    torch.cuda.synchronize()
    t1 = time.time()
    say("route", hip.last_route(), f"{t1 - t0:.2f} s", f"{2 * n * es / (t1 - t0) / 1e9:.0f} GB/s")
    assert torch.equal(x[c * M + r].view(torch.uint8 if es == 1 else torch.int64),
                       before.view(torch.uint8 if es == 1 else torch.int64))
    say("back", N, M)
    assert hip.transpose_full(x, N, M) is False
    torch.cuda.synchronize()
    say("back done", f"{time.time() - t1:.2f} s")
    xb, ob = x.view(torch.uint8), orig.view(torch.uint8)
    step = 1 << 32  # compare in 4 GiB slices: a whole-buffer compare needs another full copy
    for s in range(0, xb.numel(), step):
        assert torch.equal(xb[s:s + step], ob[s:s + step]), f"mismatch in bytes [{s}, {s + step})"
    del x, orig, xb, ob
    torch.cuda.empty_cache()
