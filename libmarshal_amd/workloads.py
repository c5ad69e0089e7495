"""Synthetic workloads: the mixed batch and the served request stream.

Batch mix (4096 matrices): rows and cols log-uniform in [16, 8192], element size uniform over
{1, 2, 4, 8, 16} bytes, 50% full transpose / 25% AoS->ASTA / 25% SoA->ASTA.

Request stream: Poisson arrivals at a given rate; request sizes log-uniform in [64 KiB, 2 GiB];
element size uniform over {4, 8, 16}; 50% full transpose, 25% AoS->ASTA, 25% ASTA->AoS with
ASTA width uniform over {32, 64, 128}.

ASTA widths must divide the number of structures (rows). For the batch mix the width is the
largest divisor of rows that is at most 64 (1 when rows is prime: AoS->ASTA is then the
identity and SoA->ASTA a full transpose); for requests the number of structures is rounded
down to a multiple of the drawn width.
"""
import math
import time

import numpy as np

OP_TRANSPOSE, OP_AOS_ASTA, OP_SOA_ASTA, OP_ASTA_AOS, OP_ASTA_SOA = 0, 1, 2, 3, 4


def largest_divisor_at_most(n, cap):
    for t in range(min(n, cap), 0, -1):
        if n % t == 0:
            return t
    return 1


def batch_mix(count=4096, lo=16, hi=8192, seed=2024, es_choices=(1, 2, 4, 8, 16)):
    """List of (op, elem_bytes, rows, cols, tile)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        r = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
        c = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
        es = int(rng.choice(es_choices))
        u = rng.uniform()
        if u < 0.5:
            out.append((OP_TRANSPOSE, es, r, c, 0))
        else:
            op = OP_AOS_ASTA if u < 0.75 else OP_SOA_ASTA
            out.append((op, es, r, c, largest_divisor_at_most(r, 64)))
    return out


def request_stream(rate, duration, seed=7, min_bytes=64 << 10, max_bytes=2 << 30):
    """List of (arrival_s, op, elem_bytes, rows, cols, tile) for `duration` seconds."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        size = int(math.exp(rng.uniform(math.log(min_bytes), math.log(max_bytes))))
        es = int(rng.choice((4, 8, 16)))
        n = max(4, size // es)
        u = rng.uniform()
        if u < 0.5:
            s = math.sqrt(n)
            m = int(math.exp(rng.uniform(math.log(max(2.0, s / 8)), math.log(max(2.0, s * 8)))))
            m = max(2, min(m, n // 2))
            out.append((t, OP_TRANSPOSE, es, m, n // m, 0))
        else:
            tile = int(rng.choice((32, 64, 128)))
            w = int(rng.integers(2, 65))
            h = max(tile, (n // w) // tile * tile)
            op = OP_AOS_ASTA if u < 0.75 else OP_ASTA_AOS
            out.append((t, op, es, h, w, tile))
    return out


def request_bytes(req):
    _, _, es, rows, cols, _ = req
    return rows * cols * es


class Ring:
    """Ring allocator over one device arena: a new region never overlaps a region whose
    request has not completed (it waits for those first)."""

    def __init__(self, server, arena_bytes, align=256):
        self.server, self.size, self.align = server, arena_bytes, align
        self.head, self.live = 0, []  # (start, end, request id)

    def alloc(self, nbytes):
        nbytes = (nbytes + self.align - 1) // self.align * self.align
        if nbytes > self.size:
            raise ValueError("request larger than the arena")
        if self.head + nbytes > self.size:
            self.head = 0
        start, end = self.head, self.head + nbytes
        keep = []
        for (s, e, rid) in self.live:
            if s < end and start < e:
                self.server.wait(rid, 600.0)
            elif self.server.status(rid) not in (2, -1):
                keep.append((s, e, rid))
        self.live = keep
        self.head = end
        return start

    def track(self, start, nbytes, rid):
        self.live.append((start, start + nbytes, rid))


def serve_trace(lm, server, arena, trace, realtime=True):
    """Submits every request of `trace` to `server` at its arrival time (or all at once when
    realtime is False), placing buffers in `arena` (a uint8 device tensor). Returns ids."""
    ring = Ring(server, arena.numel())
    base = arena.data_ptr()
    t0 = server.now() + 0.05
    ids = []
    for (ta, op, es, rows, cols, tile) in trace:
        nbytes = rows * cols * es
        off = ring.alloc(nbytes)
        arrival = t0 + ta if realtime else server.now()
        if realtime:
            delay = arrival - server.now()
            if delay > 0:
                time.sleep(delay)
        rid = server.submit(None, op, es, rows, cols, tile, arrival=arrival, ptr=base + off)
        ring.track(off, nbytes, rid)
        ids.append(rid)
    return ids


def plain_trace(lm, arena, trace, realtime=True, now=None):
    """The same trace served by plain calls, one request at a time in arrival order.
    Returns (latencies in seconds, bytes, span in seconds)."""
    import torch
    now = now or time.perf_counter
    lat, nbytes = [], 0
    head = 0
    t0 = now() + 0.05
    first = None
    for (ta, op, es, rows, cols, tile) in trace:
        n = rows * cols * es
        if head + n > arena.numel():
            head = 0
        buf = arena[head:head + n]
        head += (n + 255) // 256 * 256
        arrival = t0 + ta if realtime else now()
        if realtime:
            while now() < arrival:
                time.sleep(max(0.0, arrival - now()))
        first = arrival if first is None else first
        if op == OP_TRANSPOSE:
            e = lm.transpose_full(buf, rows, cols, elem_bytes=es)
        elif op == OP_AOS_ASTA:
            e = lm.aos_asta(buf, rows, cols, tile, elem_bytes=es)
        else:
            e = lm.asta_aos(buf, rows, cols, tile, elem_bytes=es)
        if e:
            raise RuntimeError("plain call failed")
        torch.cuda.synchronize()
        lat.append(now() - arrival)
        nbytes += 2 * n
    return lat, nbytes, now() - first
