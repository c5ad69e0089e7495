"""The batched and served workloads at their own sizes, checked bit for bit.

Batch: the 4096-matrix mix (rows and cols log-uniform in [16, 8192], element sizes 1-16 bytes,
50 % full transpose / 25 % AoS->ASTA / 25 % SoA->ASTA; ~43 GB). Every buffer gets a seeded fill
and a device clone; the batch's captured graph runs twice on the originals, the plain entry
points twice on the clones, and every pair must be equal. Matrices up to 1 MB are also checked
against the oracle on the host.

Serving: 0.2 s of the Poisson stream at 2000 requests/s with sizes log-uniform from 64 KiB to
2 GiB (element sizes 4/8/16, full transpose or AoS<->ASTA with widths 32/64/128) under a 256 GB
budget, submitted at their arrival times. Each request gets a seeded fill and a device clone;
every served result must equal the same request run alone (the plain entry point) on its clone.
"""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fill(t, seed):
    """Seeded bytes on the device (uint8 tensor)."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    step = 1 << 30
    n = t.numel()
    for s in range(0, n, step):
        e = min(n, s + step)
        t[s:e] = torch.randint(0, 256, (e - s,), dtype=torch.uint8, device="cuda", generator=g)


def plain(hip, buf, op, es, rows, cols, tile):
    if op == hip.OP_TRANSPOSE:
        return hip.transpose_full(buf, rows, cols, elem_bytes=es)
    if op == hip.OP_AOS_ASTA:
        return hip.aos_asta(buf, rows, cols, tile, elem_bytes=es)
    if op == hip.OP_SOA_ASTA:
        return hip.soa_asta(buf, rows, cols, tile, elem_bytes=es)
    if op == hip.OP_ASTA_AOS:
        return hip.asta_aos(buf, rows, cols, tile, elem_bytes=es)
    return hip.asta_soa(buf, rows, cols, tile, elem_bytes=es)


def oracle_op(orc, x, op, es, rows, cols, tile):
    from libmarshal_amd import workloads as wl
    if op == wl.OP_TRANSPOSE:
        return orc.op_transpose(x, rows, cols, es=es)
    if op == wl.OP_AOS_ASTA:
        return orc.op_010(x, rows // tile, tile, cols, es=es)
    if op == wl.OP_SOA_ASTA:
        return orc.op_100(x, cols, rows // tile, tile, es=es)
    if op == wl.OP_ASTA_AOS:
        return orc.op_010(x, rows // tile, cols, tile, es=es)
    return orc.op_100(x, rows // tile, cols, tile, es=es)


def test_batch_4096_mixed_workload(hip, orc):
    import torch
    from libmarshal_amd import workloads as wl
    items = wl.batch_mix(4096)
    total = sum(r * c * es for (_, es, r, c, _) in items)
    say("batch mix:", len(items), "matrices,", round(total / 1e9, 2), "GB")
    free, _ = torch.cuda.mem_get_info()
    assert free > 2 * total + (40 << 30), "needs room for the matrices, their clones and scratch"
    bufs, clones, small = [], [], {}
    for i, (op, es, r, c, t) in enumerate(items):
        b = torch.empty(r * c * es, dtype=torch.uint8, device="cuda")
        fill(b, 1000 + i)
        bufs.append(b)
        clones.append(b.clone())
        if r * c * es <= (1 << 20):
            small[i] = b.cpu().numpy()
    batch = hip.Batch([(b, op, es, r, c, t) for b, (op, es, r, c, t) in zip(bufs, items)])
    say("batch: kernels", batch.info("kernels"), "side-stream matrices", batch.info("solo"))
    for rep in range(2):  # the first run captures the graph, the second replays it
        assert batch.run() is False
        torch.cuda.synchronize()
        say("batch run", rep, "done")
    for b, (op, es, r, c, t) in zip(clones, items):
        for _ in range(2):
            assert plain(hip, b, op, es, r, c, t) is False
    torch.cuda.synchronize()
    bad = [i for i in range(len(items)) if not torch.equal(bufs[i], clones[i])]
    assert not bad, f"{len(bad)} matrices differ from the plain path, first {items[bad[0]]}"
    for i, x in small.items():
        op, es, r, c, t = items[i]
        y = oracle_op(orc, oracle_op(orc, x, op, es, r, c, t), op, es, r, c, t)
        assert np.array_equal(bufs[i].cpu().numpy(), y), (i, items[i])
    say("batch: all", len(items), "equal to plain calls;", len(small), "checked by the oracle")
    batch.close()
    del bufs, clones
    torch.cuda.empty_cache()


def test_serve_poisson_2000_full_size_range(hip):
    import torch
    from libmarshal_amd import workloads as wl
    trace = wl.request_stream(2000.0, 0.2, seed=17)
    total = sum(wl.request_bytes(r) for r in trace)
    say("serve:", len(trace), "requests,", round(total / 1e9, 1), "GB, largest",
        round(max(wl.request_bytes(r) for r in trace) / 1e9, 2), "GB")
    free, _ = torch.cuda.mem_get_info()
    assert free > 2 * total + (40 << 30)
    bufs, clones = [], []
    for i, req in enumerate(trace):
        b = torch.empty(wl.request_bytes(req), dtype=torch.uint8, device="cuda")
        fill(b, 5000 + i)
        bufs.append(b)
        clones.append(b.clone())
    torch.cuda.synchronize()
    server = hip.Server(256 << 30, streams=4, max_batch=256)
    t0 = server.now() + 0.02
    ids = []
    for (ta, op, es, rows, cols, tile), b in zip(trace, bufs):
        while server.now() < t0 + ta:
            time.sleep(0.0001)
        ids.append(server.submit(b, op, es, rows, cols, tile, arrival=t0 + ta))
    assert server.drain(300.0) is False
    st = server.stats()
    say("served:", {k: st[k] for k in ("completed", "failed", "batches", "max_batch",
                                       "p50_ms", "p99_ms")})
    assert st["completed"] == len(trace) and st["failed"] == 0
    for rid in ids:
        assert server.status(rid) == 2
    server.close()
    for (ta, op, es, rows, cols, tile), b in zip(trace, clones):  # each request alone
        assert plain(hip, b, op, es, rows, cols, tile) is False
        torch.cuda.synchronize()
    bad = [i for i in range(len(trace)) if not torch.equal(bufs[i], clones[i])]
    assert not bad, f"{len(bad)} requests differ from the same request alone, first {trace[bad[0]]}"
    say("serve: all", len(trace), "results equal to the request alone")
    del bufs, clones
    torch.cuda.empty_cache()
