"""Degenerate and boundary shapes on every entry point: vectors (M or N = 1), single elements,
unit tiles, no-op padding, square matrices on every route, stage codes outside the
composer's list. Every result must match the oracle (identity where the permutation is
trivial)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ES = (1, 2, 4, 8, 16)


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


@pytest.mark.parametrize("es", ES)
def test_vectors_and_single_elements(hip, orc, es):
    for (M, N) in ((1, 1), (1, 777), (777, 1), (2, 1), (1, 2)):
        x = orc.tagged(M * N, es, seed=M * 13 + N)
        d = dev(x)
        assert hip.transpose_full(d, M, N, elem_bytes=es) is False
        assert np.array_equal(host(d), orc.op_transpose(x, M, N, es=es)), (M, N)
        for stages in (2, 22, 3, 32, 4, 5, 6, 7, -1):
            d = dev(x)
            assert hip.transpose(d, 1, M, 1, N, 1, stages, elem_bytes=es) is False
            assert np.array_equal(host(d), x), (M, N, stages)  # a vector is its own transpose


@pytest.mark.parametrize("es", ES)
def test_unit_tiles_and_widths(hip, orc, es):
    cases = [(64, 13, 1), (64, 1, 64), (128, 1, 8), (1, 5, 1), (96, 7, 96)]
    for (h, w, t) in cases:
        x = orc.tagged(h * w, es, seed=h + w * 7 + t)
        d = dev(x)
        assert hip.aos_asta(d, h, w, t, elem_bytes=es) is False
        assert np.array_equal(host(d), orc.op_010(x, h // t, t, w, es=es)), (h, w, t)
        d = dev(x)
        assert hip.soa_asta(d, h, w, t, elem_bytes=es) is False
        assert np.array_equal(host(d), orc.op_100(x, w, h // t, t, es=es)), (h, w, t)


@pytest.mark.parametrize("es", ES)
def test_elementary_stages_with_unit_dimensions(hip, orc, es):
    x = orc.tagged(7 * 5 * 3, es, seed=5)
    for (A, a, B) in ((7, 1, 15), (7, 15, 1), (1, 7, 15)):
        d = dev(x)
        assert hip.transpose_010_bs(d, A, a, B, elem_bytes=es) is False
        assert np.array_equal(host(d), orc.op_010(x, A, a, B, es=es)), (A, a, B)
    for (A, B, b) in ((1, 105, 1), (105, 1, 1), (7, 15, 1), (7, 1, 15)):
        d = dev(x)
        assert hip.transpose_100(d, A, B, b, elem_bytes=es) is False
        assert np.array_equal(host(d), orc.op_100(x, A, B, b, es=es)), (A, B, b)
    for (A, a, B, b) in ((1, 7, 15, 1), (7, 1, 15, 1), (7, 15, 1, 1), (3, 5, 7, 1)):
        d = dev(x)
        assert hip.transpose_0100(d, A, a, B, b, elem_bytes=es) is False
        assert np.array_equal(host(d), orc.op_0100(x, A, a, B, b, es=es)), (A, a, B, b)


@pytest.mark.parametrize("es", (1, 4, 16))
def test_padding_no_op_and_single_row(hip, orc, es):
    y, x = 37, 53
    buf = orc.tagged(y * x, es, seed=9)
    d = dev(buf)
    assert hip.padding(d, x, y, x, elem_bytes=es) is False    # pad == x: unchanged
    assert np.array_equal(host(d), buf)
    assert hip.unpadding(d, x, y, x, elem_bytes=es) is False
    assert np.array_equal(host(d), buf)
    one = orc.tagged(64, es, seed=10)
    d = dev(one)
    assert hip.padding(d, 50, 1, 64, elem_bytes=es) is False  # one row: tail zeroed
    got = host(d).view(np.uint8).reshape(64, es)
    assert np.array_equal(got[:50], one.view(np.uint8).reshape(64, es)[:50])
    assert not got[50:].any()


@pytest.mark.parametrize("es", ES)
def test_square_matrices_every_route(hip, orc, es):
    for n in (2, 64, 96, 257, 1000):
        if n * n * es > 32 << 20:
            continue
        x = orc.tagged(n * n, es, seed=n)
        want = orc.op_transpose(x, n, n, es=es)
        d = dev(x)
        assert hip.transpose_full(d, n, n, elem_bytes=es) is False
        assert np.array_equal(host(d), want), (n, "planner", hip.last_route())
        for stages in (8, 5, 6, 7, -1):
            d = dev(x)
            assert hip.transpose(d, 1, n, 1, n, 1, stages, elem_bytes=es) is False
            assert np.array_equal(host(d), want), (n, stages, hip.last_route())
            if stages == 8:
                assert hip.last_route() == "SQ-"
    # rows of n * es bytes (not a multiple of 16 for es < 16) and a base one element past an
    # allocation's start take the element form; the element before the matrix is untouched
    n = 333
    x = orc.tagged(n * n + 1, es, seed=11)  # flat bytes: one extra element in front
    d = dev(x)
    assert hip.transpose(d[es:], 1, n, 1, n, 1, 8, elem_bytes=es) is False
    got = host(d)
    assert np.array_equal(got[:es], x[:es])
    assert np.array_equal(got[es:], orc.op_transpose(x[es:], n, n, es=es))


def test_unknown_stage_code_takes_a_valid_route(hip, orc):
    M, N, es = 120, 77, 4
    x = orc.tagged(M * N, es, seed=3)
    d = dev(x)
    assert hip.transpose(d, 1, M, 1, N, 1, 99, elem_bytes=es) is False
    say("stage 99 ->", hip.last_route())
    assert np.array_equal(host(d), orc.op_transpose(x, M, N, es=es))
