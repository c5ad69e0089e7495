"""100 / 0100 by cycle leaders: grids whose transpose map has a small order (every cycle is
short) move each cycle with one walker and no visited bits. Results must equal the oracle's
100 / 0100, for super-elements of several sizes and alignments, several instances, and the
inverse maps."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


CASES = [  # (es, A, B, b): 100(A, B, b) = SoA [A][B][b] -> ASTA [B][A][b]
    (4, 16, 1 << 14, 64),    # 2^24 x 16 f32 SoA shape at 1/16 size: order 9, 256-byte slots
    (4, 1 << 14, 16, 64),    # its inverse
    (1, 64, 4096, 160),      # order 3, 160-byte slots of bytes
    (2, 64, 4096, 100),      # 200-byte slots, order 3
    (16, 16, 16384, 16),     # 256-byte slots of 16-byte elements
    (8, 4, 1 << 16, 40),     # 320-byte slots, R = 4
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "es%d_%dx%d_b%d" % c)
def test_100_cycle_leaders(hip, orc, case):
    es, A, B, b = case
    n, k = hip.cycle_count(A, B)
    assert n > 0 and k <= 64
    say("100 leaders", case, "cycles", n, "order", k)
    x = orc.tagged(A * B * b, es, seed=A + B + b)
    d = dev(x)
    assert hip.transpose_100(d, A, B, b, elem_bytes=es) is False
    assert hip.last_route() == "0100_CYC-", hip.last_route()
    assert np.array_equal(host(d), orc.op_100(x, A, B, b, es=es)), case
    # and back
    assert hip.transpose_100(d, B, A, b, elem_bytes=es) is False
    assert np.array_equal(host(d), x), case


def test_soa_asta_takes_cycle_leaders(hip, orc):
    """SoA -> ASTA of 16 fields, ASTA width 64, f32 (the 2^24-struct conversion at 1/16 size)."""
    h, w, t = 64 << 14, 16, 64
    x = orc.tagged(h * w, 4, seed=77)
    d = dev(x)
    assert hip.soa_asta(d, h, w, t, elem_bytes=4) is False
    assert hip.last_route() == "0100_CYC-", hip.last_route()
    assert np.array_equal(host(d), orc.op_100(x, w, h // t, t, es=4))
    assert hip.asta_soa(d, h, w, t, elem_bytes=4) is False
    assert np.array_equal(host(d), x)


@pytest.mark.parametrize("es", (1, 4, 8))
def test_0100_instances_and_misaligned_base(hip, orc, es):
    """Several instances share one leader list; a base aligned to the element only makes the
    walkers move narrower units. The bytes around the matrix stay put."""
    A, a, B = 3, 16, 4096
    b = 128 // es
    n = A * a * B * b
    x = orc.tagged(n + 2, es, seed=es)
    d = dev(x)
    assert hip.transpose_0100(d[es:es + n * es], A, a, B, b, elem_bytes=es) is False
    assert hip.last_route() == "0100_CYC-", hip.last_route()
    got = host(d)
    assert np.array_equal(got[:es], x[:es]) and np.array_equal(got[es + n * es:], x[es + n * es:])
    assert np.array_equal(got[es:es + n * es], orc.op_0100(x[es:es + n * es], A, a, B, b, es=es))


def test_long_cycles_keep_the_other_forms(hip, orc):
    """A grid whose order is far above the limit is not given to the leader form."""
    es, A, B, b = 4, 1800, 100, 72  # the first 100 of the 1800 x 7200 route
    assert hip.cycle_count(A, B) == (0, 0)
    x = orc.tagged(A * B * b, es, seed=5)
    d = dev(x)
    assert hip.transpose_100(d, A, B, b, elem_bytes=es) is False
    assert hip.last_route() != "0100_CYC-"
    assert np.array_equal(host(d), orc.op_100(x, A, B, b, es=es))


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_side_stream_uses_cycle_leaders(hip, orc, use_graph):
    """A SoA conversion on the leader form inside a batch (a side stream of the captured
    graph): the first capture searches its leaders in the batch's own scratch or reuses the
    cached list; every replay matches the oracle."""
    import torch
    items = [(hip.OP_SOA_ASTA, 4, 64 << 14, 16, 64), (hip.OP_TRANSPOSE, 4, 300, 77, 0),
             (hip.OP_SOA_ASTA, 8, 40 << 16, 4, 40), (hip.OP_AOS_ASTA, 2, 64, 90, 8)]
    xs = [orc.tagged(r * c, es, seed=60 + k) for k, (op, es, r, c, t) in enumerate(items)]

    def expected(x, op, es, r, c, t):
        if op == hip.OP_SOA_ASTA:
            return orc.op_100(x, c, r // t, t, es=es)
        if op == hip.OP_AOS_ASTA:
            return orc.op_010(x, r // t, t, c, es=es)
        return orc.op_transpose(x, r, c, es=es)

    want = [expected(x, *it) for x, it in zip(xs, items)]
    bufs = [dev(x) for x in xs]
    batch = hip.Batch([(b, op, es, r, c, t) for b, (op, es, r, c, t) in zip(bufs, items)],
                      use_graph=use_graph)
    routes = [batch.route(i) for i in range(len(items))]
    say("routes", routes)
    assert "CYC" in routes[0] and "CYC" in routes[2], routes
    for replay in range(3):
        if replay:
            for b, x in zip(bufs, xs):
                b.copy_(torch.from_numpy(x).cuda())
            torch.cuda.synchronize()
        assert batch.run() is False
        torch.cuda.synchronize()
        for i, (b, w) in enumerate(zip(bufs, want)):
            assert np.array_equal(b.cpu().numpy(), w), (replay, i, routes[i])
    batch.close()
