"""Single 010 / 0100 steps through scratch: AoS<->ASTA with tiles too large for LDS,
SoA<->ASTA and 0100 over super-elements of any byte size (odd sizes, misaligned bases),
several scratch fills per call, and the same conversions inside a batch (run on its side
streams). Every result must match the oracle."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ES = (1, 2, 4, 8, 16)
LDS = 160 << 10


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


class knobs:
    """Sets tuning keys for the duration of a with-block."""

    def __init__(self, hip, **kv):
        self.hip, self.kv = hip, kv

    def __enter__(self):
        self.keep = {k: self.hip.get_tuning(k) for k in self.kv}
        for k, v in self.kv.items():
            self.hip.set_tuning(k, v)

    def __exit__(self, *exc):
        for k, v in self.keep.items():
            self.hip.set_tuning(k, v)


@pytest.mark.parametrize("es", ES)
def test_aos_asta_large_tiles_through_scratch(hip, orc, es):
    """Tiles of t x w elements beyond LDS go through scratch in both directions; ragged and
    prime widths, tiles whose rows are not 16-byte multiples."""
    for (h, w, t) in ((640, 4096, 64), (98, 2021, 49), (234, 5894, 39), (256, 4099, 64),
                      (45, 7785, 45)):
        if h * w * es > 48 << 20:
            continue
        say("aos oop", es, h, w, t)
        x = orc.tagged(h * w, es, seed=h * 7 + w + t)
        want = orc.op_010(x, h // t, t, w, es=es)
        d = dev(x)
        assert hip.aos_asta(d, h, w, t, elem_bytes=es) is False
        route = hip.last_route()
        assert np.array_equal(host(d), want), (h, w, t, route)
        if t * w * es > LDS:
            assert route == "010_OOP-", route
        assert hip.asta_aos(d, h, w, t, elem_bytes=es) is False
        assert np.array_equal(host(d), x), (h, w, t, hip.last_route())


SOA_CASES = [  # (es, h, w, t): super-elements of t * es bytes
    (1, 15 * 97, 331, 15), (1, 6 * 301, 557, 6), (2, 11 * 120, 401, 11), (4, 3 * 500, 299, 3),
    (8, 3 * 211, 407, 3), (16, 3 * 130, 257, 3), (1, 62 * 40, 777, 62), (16, 64 * 9, 301, 64),
    (1, 1 * 1201, 1999, 1), (2, 8 * 150, 613, 8), (1, 5 * 40, 3001, 5), (4, 7 * 3, 5000, 7),
    (1, 3001 * 2, 37, 3001),
]


@pytest.mark.parametrize("case", SOA_CASES, ids=lambda c: "es%d_t%d" % (c[0], c[3]))
def test_soa_asta_any_super_element_size(hip, orc, case):
    es, h, w, t = case
    say("soa", case)
    x = orc.tagged(h * w, es, seed=h + w * 3 + t)
    want = orc.op_100(x, w, h // t, t, es=es)
    d = dev(x)
    assert hip.soa_asta(d, h, w, t, elem_bytes=es) is False
    route = hip.last_route()
    assert np.array_equal(host(d), want), (case, route)
    S = t * es
    if S in (1, 2, 4, 8, 16):
        assert route.startswith("0100_FULL:"), route
    else:  # through scratch, on chip in a cluster of workgroups (grids of a few MB), or for
        # long super-elements the one-pass segmented walk
        assert route in ("0100_OOP-", "0100_SEG-", "0100_CL-"), route
    assert hip.asta_soa(d, h, w, t, elem_bytes=es) is False
    assert np.array_equal(host(d), x), (case, hip.last_route())


@pytest.mark.parametrize("es", (1, 2, 4, 8))
def test_100_misaligned_base_gathers_in_small_units(hip, orc, es):
    """A base aligned to the element size only: super-elements of 16 bytes (and odd sizes)
    are gathered in units of the base's alignment; the bytes around the matrix stay put."""
    for b in (16 // es, 3, 5):
        A, B = 301, 203
        n = A * B * b
        x = orc.tagged(n + 2, es, seed=b + es)
        d = dev(x)
        say("100 misaligned", es, b)
        assert hip.transpose_100(d[es:es + n * es], A, B, b, elem_bytes=es) is False
        got = host(d)
        assert np.array_equal(got[:es], x[:es]) and np.array_equal(got[es + n * es:],
                                                                      x[es + n * es:])
        assert np.array_equal(got[es:es + n * es], orc.op_100(x[es:es + n * es], A, B, b, es=es))


@pytest.mark.parametrize("es", ES)
def test_0100_instances_through_scratch(hip, orc, es):
    for (A, a, B, b) in ((3, 211, 307, 3), (5, 64, 1000, 1), (2, 999, 101, 16 // es), (7, 50, 41, 6)):
        if A * a * B * b * es > 48 << 20:
            continue
        say("0100 oop", es, A, a, B, b)
        x = orc.tagged(A * a * B * b, es, seed=A + a + B + b)
        d = dev(x)
        assert hip.transpose_0100(d, A, a, B, b, elem_bytes=es) is False
        route = hip.last_route()
        assert np.array_equal(host(d), orc.op_0100(x, A, a, B, b, es=es)), (A, a, B, b, route)


def test_several_scratch_fills_per_call(hip, orc):
    """A 1 MB scratch cap: blocks are moved a group at a time (one block per fill when a block
    is close to the cap), for the tiled and the gather kernels."""
    with knobs(hip, st_scratch_mb=1):
        for (es, h, w, t) in ((4, 640, 4096, 64), (1, 39 * 60, 5894, 39), (16, 200, 3001, 20)):
            say("fills aos", es, h, w, t)
            x = orc.tagged(h * w, es, seed=h + w)
            d = dev(x)
            assert hip.aos_asta(d, h, w, t, elem_bytes=es) is False
            assert hip.last_route() == "010_OOP-", hip.last_route()
            assert np.array_equal(host(d), orc.op_010(x, h // t, t, w, es=es)), (es, h, w, t)
        for (es, A, a, B, b) in ((1, 40, 211, 307, 3), (2, 9, 300, 301, 5)):
            say("fills 0100", es, A, a, B, b)
            x = orc.tagged(A * a * B * b, es, seed=A * a)
            d = dev(x)
            assert hip.transpose_0100(d, A, a, B, b, elem_bytes=es) is False
            assert hip.last_route() == "0100_OOP-", hip.last_route()
            assert np.array_equal(host(d), orc.op_0100(x, A, a, B, b, es=es)), (es, A, a, B, b)


@pytest.mark.parametrize("arena_mb", [0, 1024])
def test_batch_runs_scratch_steps_on_side_streams(hip, orc, arena_mb):
    """Large AoS / SoA conversions whose plain route goes through scratch run beside the
    grouped small ones: on side streams with the batch's through-scratch rounds off, in those
    rounds with them on (the four are at most 10 MB); every buffer matches the oracle on every
    replay."""
    import torch
    items = []
    for k, (op, es, r, c, t) in enumerate([(hip.OP_AOS_ASTA, 4, 640, 4096, 64),
                                           (hip.OP_AOS_ASTA, 16, 98, 2021, 49),
                                           (hip.OP_SOA_ASTA, 1, 15 * 400, 801, 15),
                                           (hip.OP_SOA_ASTA, 2, 6 * 300, 1001, 6),
                                           (hip.OP_AOS_ASTA, 2, 64, 90, 8),
                                           (hip.OP_SOA_ASTA, 4, 33, 70, 3),
                                           (hip.OP_TRANSPOSE, 8, 100, 37, 0)]):
        items.append((orc.tagged(r * c, es, seed=700 + k), op, es, r, c, t))

    def expected(it):
        x, op, es, r, c, t = it
        if op == hip.OP_AOS_ASTA:
            return orc.op_010(x, r // t, t, c, es=es)
        if op == hip.OP_SOA_ASTA:
            return orc.op_100(x, c, r // t, t, es=es)
        return orc.op_transpose(x, r, c, es=es)

    want = [expected(it) for it in items]
    bufs = [dev(it[0]) for it in items]
    keep = hip.get_tuning("batch_arena_mb")
    hip.set_tuning("batch_arena_mb", arena_mb)
    try:
        batch = hip.Batch([(b, op, es, r, c, t) for b, (_, op, es, r, c, t) in zip(bufs, items)],
                          use_graph=True)
    finally:
        hip.set_tuning("batch_arena_mb", keep)
    routes = [batch.route(i) for i in range(len(items))]
    say("routes", routes, "solo", batch.info("solo"), "arena", batch.info("arena_matrices"))
    if arena_mb:
        assert all(routes[i] == "ARENA-" for i in range(4)), routes
    else:
        assert batch.info("solo") >= 4
        assert all("OOP" in routes[i] for i in range(4)), routes
    for replay in range(3):
        if replay:
            for b, it in zip(bufs, items):
                b.copy_(torch.from_numpy(it[0]).cuda())
            torch.cuda.synchronize()
        assert batch.run() is False
        torch.cuda.synchronize()
        for i, (b, w) in enumerate(zip(bufs, want)):
            assert np.array_equal(b.cpu().numpy(), w), (replay, i, routes[i])
    batch.close()
