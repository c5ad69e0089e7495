"""Batched form: many conversions of uneven shapes and element sizes in one graph launch.

Every buffer must end bit-identical to the oracle and to the plain single-matrix entry point
run on it alone; graph replays are checked one by one with eager work in between.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ES = (1, 2, 4, 8, 16)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_mix(orc, count, lo, hi, seed):
    """Descriptors like the batched benchmark's, at smaller sizes: rows/cols log-uniform in
    [lo, hi], element size uniform over 1..16 bytes, 50% transpose / 25% AoS->ASTA /
    25% SoA->ASTA (ASTA tile width: a divisor of rows, so every request is valid)."""
    import libmarshal_amd as lm
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        rows = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        cols = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        es = int(rng.choice(ES))
        u = rng.uniform()
        if u < 0.5:
            op, tile = lm.OP_TRANSPOSE, 0
        else:
            op = lm.OP_AOS_ASTA if u < 0.75 else lm.OP_SOA_ASTA
            tile = next(t for t in (64, 32, 16, 8, 4, 2, 1) if rows % t == 0)
        x = orc.tagged(rows * cols, es, seed=seed * 100003 + i)
        out.append((x, op, es, rows, cols, tile))
    return out


def expected(orc, item):
    import libmarshal_amd as lm
    x, op, es, rows, cols, tile = item
    if op == lm.OP_TRANSPOSE:
        return orc.op_transpose(x, rows, cols, es=es)
    if op == lm.OP_AOS_ASTA:
        return orc.op_010(x, rows // tile, tile, cols, es=es)
    return orc.op_100(x, cols, rows // tile, tile, es=es)


def upload(items):
    import torch
    return [torch.from_numpy(it[0].copy()).cuda() for it in items]


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_oracle_and_plain(hip, orc, use_graph):
    import torch
    items = make_mix(orc, 300, 16, 700, seed=11 if use_graph else 12)
    want = [expected(orc, it) for it in items]
    bufs = upload(items)
    batch = hip.Batch([(b, op, es, r, c, t) for b, (_, op, es, r, c, t) in zip(bufs, items)],
                      use_graph=use_graph)
    say("batch", use_graph, batch.info("levels"), batch.info("kernels"), batch.info("stages"))
    assert batch.info("kernels") * 4 < batch.info("stages")  # grouped, not one launch per stage
    for replay in range(3):
        if replay:
            for b, it in zip(bufs, items):  # eager work between replays
                b.copy_(torch.from_numpy(it[0]).cuda())
            torch.cuda.synchronize()
        assert batch.run() is False
        torch.cuda.synchronize()
        for i, (b, w) in enumerate(zip(bufs, want)):
            got = b.cpu().numpy()
            assert np.array_equal(got, w), (replay, i, items[i][1:], batch.route(i))
    batch.close()
    # alone: the plain entry points on the same inputs give the same bytes
    for i, it in enumerate(items[:80]):
        x, op, es, rows, cols, tile = it
        d = torch.from_numpy(x.copy()).cuda()
        if op == hip.OP_TRANSPOSE:
            r = hip.transpose_full(d, rows, cols, elem_bytes=es)
        elif op == hip.OP_AOS_ASTA:
            r = hip.aos_asta(d, rows, cols, tile, elem_bytes=es)
        else:
            r = hip.soa_asta(d, rows, cols, tile, elem_bytes=es)
        assert r is False
        assert np.array_equal(d.cpu().numpy(), want[i]), (i, it[1:])


def test_batch_inverse_ops_round_trip(hip, orc):
    import torch
    rng = np.random.default_rng(3)
    items, bufs = [], []
    for i in range(60):
        h = int(rng.integers(1, 40)) * 64
        w = int(rng.integers(1, 40))
        es = int(rng.choice(ES))
        x = orc.tagged(h * w, es, seed=i)
        items.append((x, es, h, w))
        bufs.append(torch.from_numpy(x.copy()).cuda())
    fwd = hip.Batch([(b, hip.OP_AOS_ASTA if i % 2 else hip.OP_SOA_ASTA, es, h, w, 64)
                     for i, (b, (x, es, h, w)) in enumerate(zip(bufs, items))])
    inv = hip.Batch([(b, hip.OP_ASTA_AOS if i % 2 else hip.OP_ASTA_SOA, es, h, w, 64)
                     for i, (b, (x, es, h, w)) in enumerate(zip(bufs, items))])
    assert fwd.run() is False
    torch.cuda.synchronize()
    for i, (b, (x, es, h, w)) in enumerate(zip(bufs, items)):
        want = orc.op_010(x, h // 64, 64, w, es=es) if i % 2 else orc.op_100(x, w, h // 64, 64, es=es)
        assert np.array_equal(b.cpu().numpy(), want), i
    assert inv.run() is False
    torch.cuda.synchronize()
    for b, (x, es, h, w) in zip(bufs, items):
        assert np.array_equal(b.cpu().numpy(), x)


def test_batch_rejects_invalid_descriptor(hip):
    import torch
    d = torch.zeros(100, dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError):
        hip.Batch([(d, hip.OP_AOS_ASTA, 4, 10, 2, 3)])  # 10 rows not a multiple of tile 3
    with pytest.raises(ValueError):
        hip.Batch([(d, hip.OP_TRANSPOSE, 3, 10, 2, 0)])  # 3-byte elements


def test_batch_large_members(hip, orc):
    """A few members big enough to span the whole GPU, mixed with tiny ones."""
    import torch
    rng = np.random.default_rng(9)
    items = []
    for i, (r, c, es, op) in enumerate([(4096, 3000, 4, 0), (16, 17, 16, 0), (8192, 64, 8, 1),
                                        (2048, 2048, 2, 0), (640, 1024, 1, 2), (5, 7, 1, 0)]):
        tile = 64 if op else 0
        items.append((orc.tagged(r * c, es, seed=i + 50), op, es, r, c, tile))
    want = [expected(orc, it) for it in items]
    bufs = upload(items)
    batch = hip.Batch([(b, op, es, r, c, t) for b, (_, op, es, r, c, t) in zip(bufs, items)])
    assert batch.run() is False
    torch.cuda.synchronize()
    for i, (b, w) in enumerate(zip(bufs, want)):
        assert np.array_equal(b.cpu().numpy(), w), (i, batch.route(i))


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_with_stripe_route_matrices(hip, orc, use_graph):
    """Transposes whose best route is the stripe route run on the batch's second stream next
    to the grouped small ones; every buffer must still match the oracle, on every replay."""
    import torch
    big = [(4096, 4099, 16), (2053, 6007, 8)]
    small = make_mix(orc, 40, 16, 300, seed=21)
    items = list(small)
    for k, (M, N, es) in enumerate(big):
        items.append((orc.tagged(M * N, es, seed=900 + k), hip.OP_TRANSPOSE, es, M, N, 0))
    want = [expected(orc, it) for it in items]
    bufs = upload(items)
    batch = hip.Batch([(b, op, es, r, c, t) for b, (_, op, es, r, c, t) in zip(bufs, items)],
                      use_graph=use_graph)
    say("solo", batch.info("solo"), [batch.route(len(small) + k) for k in range(len(big))])
    assert batch.info("solo") >= 1
    for replay in range(3):
        if replay:
            for b, it in zip(bufs, items):
                b.copy_(torch.from_numpy(it[0]).cuda())
            torch.cuda.synchronize()
        assert batch.run() is False
        torch.cuda.synchronize()
        for i, (b, w) in enumerate(zip(bufs, want)):
            assert np.array_equal(b.cpu().numpy(), w), (replay, i, items[i][1:], batch.route(i))
    batch.close()


@pytest.mark.parametrize("arena_mb,lanes", [(1024, 4), (1, 4), (1024, 1), (1, 1)])
def test_batch_through_scratch_rounds(hip, orc, arena_mb, lanes):
    """Matrices whose best route is not one streaming pass go through the batch's arena: one
    grouped copy per round, then grouped out-of-place transposes (aligned, misaligned and
    skewed destination rows) or super-element gathers (SoA tiles of 3 to 64 elements) back.
    With a 1 MiB round the same matrices take many rounds (and the larger ones other paths);
    every buffer must match the oracle on every replay, with the groups of a level on branch
    streams and all on one stream."""
    import torch
    shapes = [(0, 1, 333, 517, 0), (0, 2, 1001, 77, 0), (0, 4, 129, 4099, 0), (0, 8, 700, 31, 0),
              (0, 16, 97, 211, 0), (0, 1, 2048, 1500, 0), (0, 2, 17, 3000, 0),
              (2, 4, 96 * 3, 50, 3), (2, 1, 640, 33, 5), (2, 8, 1280, 7, 64), (2, 2, 330, 19, 11),
              (1, 1, 1536, 700, 3), (1, 16, 640, 40, 5), (1, 2, 4096, 300, 64),
              (0, 8, 16, 8192, 0), (0, 1, 5000, 16, 0)]
    items = []
    for k, (op, es, r, c, t) in enumerate(shapes):
        items.append((orc.tagged(r * c, es, seed=700 + k), op, es, r, c, t))
    want = [expected(orc, it) for it in items]
    bufs = upload(items)
    # one buffer 8 bytes past a 16-byte boundary: never taken into the arena
    x = orc.tagged(300 * 257, 8, seed=799)
    holder = torch.zeros(300 * 257 * 8 + 8, dtype=torch.uint8, device="cuda")
    holder[8:].copy_(torch.from_numpy(x.view(np.uint8)).cuda())
    items.append((x, 0, 8, 300, 257, 0))
    want.append(orc.op_transpose(x, 300, 257, es=8))
    bufs.append(holder[8:])
    keep = {k: hip.get_tuning(k) for k in ("batch_arena_mb", "batch_level_lanes")}
    hip.set_tuning("batch_arena_mb", arena_mb)
    hip.set_tuning("batch_level_lanes", lanes)
    try:
        batch = hip.Batch([(b, op, es, r, c, t) for b, (_, op, es, r, c, t) in zip(bufs, items)])
    finally:
        for k, v in keep.items():
            hip.set_tuning(k, v)
    say("arena", arena_mb, batch.info("arena_matrices"), batch.info("arena_rounds"),
        batch.info("arena_bytes"), [batch.route(i) for i in range(len(items))])
    assert batch.info("arena_matrices") >= (8 if arena_mb > 1 else 1)
    assert batch.route(len(items) - 1) != "ARENA-"
    if arena_mb == 1:
        assert batch.info("arena_rounds") >= 2
    for replay in range(3):
        if replay:
            for b, it in zip(bufs, items):
                b.copy_(torch.from_numpy(it[0]).cuda())
            torch.cuda.synchronize()
        assert batch.run() is False
        torch.cuda.synchronize()
        for i, (b, w) in enumerate(zip(bufs, want)):
            assert np.array_equal(b.cpu().numpy(), w), (replay, i, items[i][1:], batch.route(i))
    batch.close()
