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


def _fill(x, g):
    import torch
    for s in range(0, x.numel(), 1 << 32):
        e = min(x.numel(), s + (1 << 32))
        x[s:e].copy_(torch.randint(0, 256, (e - s,), dtype=torch.uint8, device="cuda",
                                   generator=g))


@pytest.mark.parametrize("case", ["transpose_u8", "aos_f32", "soa_u8_s15"])
def test_through_scratch_beyond_4gib(hip, case):
    """The through-scratch forms with offsets past 2^32 bytes: a 4.3 GB u8 transpose with
    misaligned rows (skewed tiles), a 5.1 GB AoS->ASTA in several scratch fills, and a 4.5 GB
    SoA->ASTA over 15-byte super-elements (the gather). Sampled positions must land where
    the index map says, and the inverse must restore the buffer exactly."""
    import torch
    keep = hip.get_tuning("st_scratch_mb")
    if case == "transpose_u8":
        es, M, N = 1, 70001, 61357
        nbytes = M * N
    elif case == "aos_f32":
        es, h, w, t = 4, 64 * 20000, 1001, 64
        nbytes = h * w * es
    else:
        es, h, w, t = 1, 15 * 20000, 15001, 15
        nbytes = h * w * es
    free, _ = torch.cuda.mem_get_info()
    if free < 3 * nbytes + (8 << 30):
        pytest.skip(f"needs {3 * nbytes / 1e9:.0f} GB of device memory")
    try:
        hip.set_tuning("st_scratch_mb", (nbytes >> 20) + 64 if case != "aos_f32" else 2048)
        g = torch.Generator(device="cuda").manual_seed(nbytes)
        x = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        _fill(x, g)
        orig = x.clone()
        k = 1 << 20
        if case == "transpose_u8":
            r = torch.randint(0, M, (k,), device="cuda", generator=g)
            c = torch.randint(0, N, (k,), device="cuda", generator=g)
            src, dst = r * N + c, c * M + r
            fwd = lambda: hip.transpose_full(x, M, N)
            inv = lambda: hip.transpose_full(x, N, M)
            tag = "OOP-"
        elif case == "aos_f32":  # 010(h/t, t, w): [A][a][B] -> [A][B][a], 4-byte units
            A, a, B = h // t, t, w
            i = torch.randint(0, A, (k,), device="cuda", generator=g)
            j = torch.randint(0, a, (k,), device="cuda", generator=g)
            l = torch.randint(0, B, (k,), device="cuda", generator=g)
            src, dst = i * a * B + j * B + l, i * a * B + l * a + j
            fwd = lambda: hip.aos_asta(x, h, w, t, elem_bytes=es)
            inv = lambda: hip.asta_aos(x, h, w, t, elem_bytes=es)
            tag = "010_OOP-"
        else:  # 100(w, h/t, t): [A][B][b] -> [B][A][b], one-byte units
            A, B, b = w, h // t, t
            i = torch.randint(0, A, (k,), device="cuda", generator=g)
            j = torch.randint(0, B, (k,), device="cuda", generator=g)
            l = torch.randint(0, b, (k,), device="cuda", generator=g)
            src, dst = (i * B + j) * b + l, (j * A + i) * b + l
            fwd = lambda: hip.soa_asta(x, h, w, t, elem_bytes=es)
            inv = lambda: hip.asta_soa(x, h, w, t, elem_bytes=es)
            tag = "0100_OOP-"
        if es == 4:  # element index -> its 4 bytes
            src = (src.unsqueeze(1) * 4 + torch.arange(4, device="cuda")).reshape(-1)
            dst = (dst.unsqueeze(1) * 4 + torch.arange(4, device="cuda")).reshape(-1)
        before = orig[src]
        say(case, nbytes)
        assert fwd() is False
        torch.cuda.synchronize()
        say("route", hip.last_route())
        assert hip.last_route() == tag
        assert torch.equal(x[dst], before)
        assert inv() is False
        torch.cuda.synchronize()
        for s in range(0, nbytes, 1 << 32):
            assert torch.equal(x[s:s + (1 << 32)], orig[s:s + (1 << 32)]), f"bytes from {s}"
        del x, orig
        torch.cuda.empty_cache()
    finally:
        hip.set_tuning("st_scratch_mb", keep)
