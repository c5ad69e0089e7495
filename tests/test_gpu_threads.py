"""Concurrent use: several host threads, each on a stream of its own, convert their own buffers
at the same time.

The calls share the library's per-device state (the scratch pool of visited bits and
through-scratch buffers, the cycle-leader cache keyed by grid, the cluster barrier words kept
per stream, the class-choice table) and the ctypes calls release the GIL, so the host sides
really overlap. Every buffer must end bit-identical to the oracle, as when run alone.
"""
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = 4
ROUNDS = 3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def work_items(orc, seed):
    """(name, run(hip, buf, stream, elapsed), input bytes, expected bytes) for one thread.
    Every thread gets the same shapes (so their cached cycle-leader lists and pool sizes
    collide) with its own data."""
    items = []

    def add(name, run, x, want):
        items.append((name, run, x, want))

    # full transpose, non-square: the through-scratch route (pool-held scratch buffer)
    x = orc.tagged(1536 * 2311, 4, seed=seed + 1)
    add("T 1536x2311 f32", lambda h, b, s, e: h.transpose_full(b, 1536, 2311, elapsed=e,
                                                                elem_bytes=4, stream=s),
        x, orc.op_transpose(x, 1536, 2311, es=4))
    # square transpose: tile-pair swaps
    x = orc.tagged(1024 * 1024, 2, seed=seed + 2)
    add("T 1024x1024 u16", lambda h, b, s, e: h.transpose_full(b, 1024, 1024, elapsed=e,
                                                                elem_bytes=2, stream=s),
        x, orc.op_transpose(x, 1024, 1024, es=2))
    # AoS -> ASTA: the 010 on the 1024 x 64 x 16 grid (cached cycle leaders)
    x = orc.tagged(65536 * 16, 4, seed=seed + 3)
    add("AoS->ASTA 65536x16/64", lambda h, b, s, e: h.aos_asta(b, 65536, 16, 64, elapsed=e,
                                                                elem_bytes=4, stream=s),
        x, orc.op_010(x, 1024, 64, 16, es=4))
    # the cycle-following 010 on 64 x 16 tiles: every thread asks for the same ordered leader
    # list at once on its first call
    x = orc.tagged(2048 * 64 * 16, 4, seed=seed + 7)
    add("PT010 2048x64x16 f32", lambda h, b, s, e: h.transpose_010_pttwac(
        b, 2048, 64, 16, elapsed=e, R=1, P=1, elem_bytes=4, stream=s),
        x, orc.op_010(x, 2048, 64, 16, es=4))
    # SoA -> ASTA: the 100 on the 16 x 1024 grid of 64-element super-elements
    x = orc.tagged(65536 * 16, 4, seed=seed + 4)
    add("SoA->ASTA 65536x16/64", lambda h, b, s, e: h.soa_asta(b, 65536, 16, 64, elapsed=e,
                                                                elem_bytes=4, stream=s),
        x, orc.op_100(x, 16, 1024, 64, es=4))
    # 0100 with global visited bits
    x = orc.tagged(3 * 40 * 257 * 24, 8, seed=seed + 5)
    add("0100 3x40x257x24 f64", lambda h, b, s, e: h.transpose_0100(b, 3, 40, 257, 24, elapsed=e,
                                                                     elem_bytes=8, stream=s),
        x, orc.op_0100(x, 3, 40, 257, 24, es=8))
    # padding of 700 rows of 1001 bytes to a pitch of 1024
    x = np.zeros(700 * 1024, dtype=np.uint8)
    x[:700 * 1001] = orc.tagged(700 * 1001, 1, seed=seed + 6)
    want = x.copy()
    orc.padding(want, 1, 1001, 700, 1024)
    add("PAD 1001->1024 x700 u8", lambda h, b, s, e: h.padding(b, 1001, 700, 1024, elapsed=e,
                                                                elem_bytes=1, stream=s),
        x, want)
    return items


def worker(hip, orc, k, failures, started, profiling=1):
    import torch
    try:
        items = work_items(orc, seed=1000 * (k + 1))
        stream = torch.cuda.Stream()  # default priority
        # odd threads time their calls (each such call synchronises its own stream)
        elapsed = hip.Elapsed() if k % 2 else None
        with torch.cuda.stream(stream):
            bufs = [torch.from_numpy(x.copy()).cuda() for (_, _, x, _) in items]
        stream.synchronize()
        started.wait(timeout=120)
        for r in range(ROUNDS):
            for (name, run, x, want), b in zip(items, bufs):
                if r:
                    with torch.cuda.stream(stream):
                        b.copy_(torch.from_numpy(x).cuda(non_blocking=False))
                assert run(hip, b, stream, elapsed) is False, (k, r, name)
            stream.synchronize()
            for (name, run, x, want), b in zip(items, bufs):
                got = b.cpu().numpy()
                if not np.array_equal(got, want):
                    bad = np.nonzero(got != want)[0]
                    failures.append(f"thread {k} round {r} {name}: {bad.size} bytes differ, "
                                    f"first at {bad[0]}")
        if elapsed is not None and profiling:
            assert elapsed.ns > 0
    except BaseException as e:  # reported by the main thread
        failures.append(f"thread {k}: {type(e).__name__}: {e}")


@pytest.mark.parametrize("profiling", [1, 0])
def test_concurrent_threads_on_own_streams(hip, orc, profiling):
    """With profiling on every call waits for its own work; with it off the calls are only
    stream-ordered, so the streams' kernels overlap on the GPU and the pool's hand-over from
    one stream to the next is what keeps a stage's visited bits and scratch to itself."""
    import torch
    say("threads", THREADS, "rounds", ROUNDS, "profiling", profiling)
    failures = []
    started = threading.Barrier(THREADS)
    ts = [threading.Thread(target=worker, args=(hip, orc, k, failures, started, profiling))
          for k in range(THREADS)]
    hip.set_profiling(profiling)
    try:
        for t in ts:
            t.start()
        for t in ts:
            t.join(timeout=150)
    finally:
        hip.set_profiling(1)
    assert not any(t.is_alive() for t in ts), "a worker thread did not finish"
    torch.cuda.synchronize()
    assert not failures, failures
    assert hip.cluster_aborts() == 0


def test_threads_sharing_one_stream(hip, orc):
    """Threads that enqueue on one shared stream: the calls are ordered by the stream, each
    thread's buffers must still come out right."""
    import torch
    stream = torch.cuda.Stream()
    per = [work_items(orc, seed=7000 + 100 * k)[:4] for k in range(3)]
    bufs = [[torch.from_numpy(x.copy()).cuda() for (_, _, x, _) in it] for it in per]
    torch.cuda.synchronize()
    errors = []

    def go(k):
        try:
            for (name, run, _, _), b in zip(per[k], bufs[k]):
                assert run(hip, b, stream, None) is False, (k, name)
        except BaseException as e:
            errors.append(f"thread {k}: {e}")

    ts = [threading.Thread(target=go, args=(k,)) for k in range(3)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    stream.synchronize()
    assert not errors, errors
    for k in range(3):
        for (name, _, _, want), b in zip(per[k], bufs[k]):
            assert np.array_equal(b.cpu().numpy(), want), (k, name)


@pytest.mark.parametrize("use_graph", [True, False])
def test_batches_built_and_run_beside_plain_calls(hip, orc, use_graph):
    """Three threads build batches (planning, allocation, graph capture on their own streams)
    and replay them while a fourth makes plain calls that grow the shared pool and the
    cycle-leader cache: a capture is local to its thread, so neither side may disturb the
    other. Every buffer of every replay must match the oracle."""
    import torch
    from tests.test_gpu_batch import expected, make_mix
    mixes = [make_mix(orc, 60, 16, 600, seed=300 + k) for k in range(3)]
    wants = [[expected(orc, it) for it in mix] for mix in mixes]
    plain = work_items(orc, seed=9000)
    failures = []
    started = threading.Barrier(4)

    def batch_worker(k):
        try:
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                bufs = [torch.from_numpy(it[0].copy()).cuda() for it in mixes[k]]
            stream.synchronize()
            started.wait(timeout=120)
            batch = hip.Batch([(b, op, es, r, c, t)
                               for b, (_, op, es, r, c, t) in zip(bufs, mixes[k])],
                              use_graph=use_graph)
            for replay in range(2):
                if replay:
                    with torch.cuda.stream(stream):
                        for b, it in zip(bufs, mixes[k]):
                            b.copy_(torch.from_numpy(it[0]).cuda())
                assert batch.run(stream=stream) is False, (k, replay)
                stream.synchronize()
                for i, (b, w) in enumerate(zip(bufs, wants[k])):
                    if not np.array_equal(b.cpu().numpy(), w):
                        failures.append(f"batch {k} replay {replay} item {i} "
                                        f"{mixes[k][i][1:]} {batch.route(i)}")
            batch.close()
        except BaseException as e:
            failures.append(f"batch thread {k}: {type(e).__name__}: {e}")

    def plain_worker():
        try:
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                bufs = [torch.from_numpy(x.copy()).cuda() for (_, _, x, _) in plain]
            stream.synchronize()
            started.wait(timeout=120)
            for (name, run, x, want), b in zip(plain, bufs):
                assert run(hip, b, stream, None) is False, name
            stream.synchronize()
            for (name, run, x, want), b in zip(plain, bufs):
                if not np.array_equal(b.cpu().numpy(), want):
                    failures.append(f"plain {name}")
        except BaseException as e:
            failures.append(f"plain thread: {type(e).__name__}: {e}")

    ts = [threading.Thread(target=batch_worker, args=(k,)) for k in range(3)]
    ts.append(threading.Thread(target=plain_worker))
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=150)
    assert not any(t.is_alive() for t in ts), "a worker thread did not finish"
    torch.cuda.synchronize()
    assert not failures, failures[:10]
    assert hip.cluster_aborts() == 0
