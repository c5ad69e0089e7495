"""Serving loop: continuous batching with admission by HBM budget.

Each request's result must be bit-identical to the oracle and to the plain entry point run on
it alone, whether requests arrive together, trickle in while others run, or queue behind a
tight budget.
"""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_requests(orc, count, seed):
    import libmarshal_amd as lm
    from libmarshal_amd import workloads as wl
    reqs = wl.request_stream(rate=2000.0, duration=count / 2000.0, seed=seed,
                             min_bytes=4 << 10, max_bytes=2 << 20)
    out = []
    for i, (ta, op, es, rows, cols, tile) in enumerate(reqs[:count]):
        x = orc.tagged(rows * cols, es, seed=seed * 7919 + i)
        if op == lm.OP_TRANSPOSE:
            want = orc.op_transpose(x, rows, cols, es=es)
        elif op == lm.OP_AOS_ASTA:
            want = orc.op_010(x, rows // tile, tile, cols, es=es)
        else:
            want = orc.op_010(x, rows // tile, cols, tile, es=es)
        out.append((ta, op, es, rows, cols, tile, x, want))
    return out


def run_and_check(hip, reqs, server, realtime):
    import torch
    bufs = [torch.from_numpy(r[6].copy()).cuda() for r in reqs]
    torch.cuda.synchronize()
    t0 = server.now() + 0.02
    ids = []
    for (ta, op, es, rows, cols, tile, x, want), b in zip(reqs, bufs):
        if realtime:
            while server.now() < t0 + ta:
                time.sleep(0.0001)
        ids.append(server.submit(b, op, es, rows, cols, tile,
                                 arrival=t0 + ta if realtime else 0.0))
    assert server.drain(120.0) is False
    for rid, b, r in zip(ids, bufs, reqs):
        assert server.status(rid) == 2
        assert np.array_equal(b.cpu().numpy(), r[7]), (rid, r[1:6])
    return ids


@pytest.mark.parametrize("realtime", [False, True])
def test_server_results_match_alone(hip, orc, realtime):
    reqs = small_requests(orc, 400, seed=1 if realtime else 2)
    server = hip.Server(64 << 30, streams=4, max_batch=64)
    run_and_check(hip, reqs, server, realtime)
    st = server.stats()
    say("stats", st)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["batches"] >= 1 and st["p99_ms"] >= st["p50_ms"] > 0
    server.close()
    # alone: plain calls give the same bytes
    import torch
    for (ta, op, es, rows, cols, tile, x, want) in reqs[:60]:
        d = torch.from_numpy(x.copy()).cuda()
        if op == hip.OP_TRANSPOSE:
            assert hip.transpose_full(d, rows, cols, elem_bytes=es) is False
        elif op == hip.OP_AOS_ASTA:
            assert hip.aos_asta(d, rows, cols, tile, elem_bytes=es) is False
        else:
            assert hip.asta_aos(d, rows, cols, tile, elem_bytes=es) is False
        assert np.array_equal(d.cpu().numpy(), want)


def test_server_tight_budget_queues_but_stays_exact(hip, orc):
    """A budget of a few MB forces requests to wait for earlier ones to finish."""
    reqs = small_requests(orc, 150, seed=3)
    biggest = max(r[3] * r[4] * r[2] for r in reqs)
    server = hip.Server(3 * biggest, streams=3, max_batch=8)
    run_and_check(hip, reqs, server, realtime=False)
    st = server.stats()
    say("tight", st)
    assert st["completed"] == len(reqs) and st["batches"] > len(reqs) // 8
    server.close()


def test_server_rejects_oversized_and_invalid(hip):
    import torch
    d = torch.zeros(1 << 20, dtype=torch.uint8, device="cuda")
    server = hip.Server(1 << 16, streams=1, max_batch=4)
    with pytest.raises(ValueError):
        server.submit(d, hip.OP_TRANSPOSE, 4, 512, 512)  # 1 MiB > 64 KiB budget
    with pytest.raises(ValueError):
        server.submit(d, hip.OP_AOS_ASTA, 4, 100, 3, 64)  # 100 rows not a multiple of 64
    server.close()


def test_server_large_requests_run_alone_and_stay_exact(hip, orc):
    """Requests of 64 MiB or more run alone on the plain path (the stripe route for prime
    dimensions) next to small batched ones; every result must match the oracle."""
    import torch
    big = [(4099, 4111, 4, hip.OP_TRANSPOSE, 0), (2053, 6007, 8, hip.OP_TRANSPOSE, 0),
           (1 << 18, 64, 4, hip.OP_AOS_ASTA, 64)]
    reqs = []
    for k, (rows, cols, es, op, tile) in enumerate(big):
        x = orc.tagged(rows * cols, es, seed=500 + k)
        if op == hip.OP_TRANSPOSE:
            want = orc.op_transpose(x, rows, cols, es=es)
        else:
            want = orc.op_010(x, rows // tile, tile, cols, es=es)
        reqs.append((0.0, op, es, rows, cols, tile, x, want))
    small = small_requests(orc, 60, seed=9)
    mixed = [reqs[0]] + small[:30] + [reqs[1]] + small[30:] + [reqs[2]]
    server = hip.Server(64 << 30, streams=4, max_batch=32)
    run_and_check(hip, mixed, server, realtime=False)
    st = server.stats()
    say("large+small", st)
    assert st["completed"] == len(mixed) and st["failed"] == 0
    server.close()


def test_server_counts_scratch_against_the_budget(hip, orc):
    """A request's scratch counts against the HBM budget: a 101 MB transpose whose plain route
    goes through scratch (a second copy of its buffer) is refused by a budget of 1.5x its buffer
    and accepted, on an in-place route, once the per-request scratch cap is 0; under a budget
    of 2.6x the largest buffer, large requests that need scratch queue behind each other and
    every result stays exact."""
    import torch
    M, N, es = 4099, 6151, 4  # 100.9 MB, both dimensions prime
    x = orc.tagged(M * N, es, seed=61)
    want = orc.op_transpose(x, M, N, es=es)
    nbytes = M * N * es
    keep = hip.get_tuning("serve_scratch_mb")
    try:
        server = hip.Server(int(1.5 * nbytes), streams=2, max_batch=8)
        d = torch.from_numpy(x.copy()).cuda()
        with pytest.raises(ValueError):
            server.submit(d, hip.OP_TRANSPOSE, es, M, N)
        server.close()
        hip.set_tuning("serve_scratch_mb", 0)  # in-place routes only: the buffer fits alone
        server = hip.Server(int(1.5 * nbytes), streams=2, max_batch=8)
        rid = server.submit(d, hip.OP_TRANSPOSE, es, M, N)
        assert server.drain(120.0) is False and server.status(rid) == 2
        assert np.array_equal(d.cpu().numpy(), want)
        server.close()
    finally:
        hip.set_tuning("serve_scratch_mb", keep)
    # through-scratch requests under a budget that holds about two of them with scratch
    big = []
    for k, (m, n) in enumerate(((4099, 6151), (4111, 6143), (4093, 6163))):
        y = orc.tagged(m * n, es, seed=70 + k)
        big.append((0.0, hip.OP_TRANSPOSE, es, m, n, 0, y, orc.op_transpose(y, m, n, es=es)))
    small = small_requests(orc, 40, seed=12)
    mixed = [big[0]] + small[:20] + [big[1]] + small[20:] + [big[2]]
    server = hip.Server(int(2.6 * nbytes) + (64 << 20), streams=3, max_batch=8)
    run_and_check(hip, mixed, server, realtime=False)
    st = server.stats()
    say("scratch budget", st)
    assert st["completed"] == len(mixed) and st["failed"] == 0
    server.close()
