"""The CPU oracle, pinned against the reference's known answers, its own planner header
(oracle/_ref), numpy's textbook definitions (golden vectors) and brute force."""
import itertools

import numpy as np
import pytest

from tests.golden import vectors

ES = (1, 2, 4, 8, 16)


def brute_cycles(m, n):
    """Cycle decomposition of i -> i*m mod (mn-1) by marking, independent of the oracle."""
    N = m * n
    if N <= 2:
        return 0, 1
    seen = np.zeros(N, bool)
    count, longest = 0, 1
    for i in range(1, N - 1):
        if seen[i]:
            continue
        j, length = i, 0
        while True:
            seen[j] = True
            j = (j * m) % (N - 1)
            length += 1
            if j == i:
                break
        if length > 1:
            count += 1
            longest = max(longest, length)
    return count, longest


def test_cycle_kat(orc):
    # libmarshal_plan_test.cycle, test/cl_unittest.cc:705-709
    assert orc.num_cycles(2, 3) == 1
    assert orc.num_cycles(3, 5) == 2


@pytest.mark.parametrize("m,n", [(2, 3), (3, 5), (4, 4), (7, 9), (16, 5), (1, 8), (8, 1),
                                 (12, 18), (25, 13)])
def test_cycles_brute_force(orc, m, n):
    c, longest = brute_cycles(m, n)
    assert orc.num_cycles(m, n) == c
    assert orc.longest_cycle(m, n) == longest


def test_next_closed_form(orc):
    for m, n in [(3, 5), (18, 72), (100, 100), (7, 1), (1, 7), (131, 97)]:
        N = m * n
        for i in range(0, N, max(1, N // 500)):
            want = i if i == N - 1 else (i * m) % (N - 1) if N > 1 else 0
            assert orc.next_index(i, m, n) == want
            # the map sends row-major (r, c) of m x n to row-major (c, r) of n x m
            r, c = divmod(i, n)
            assert orc.next_index(i, m, n) == c * m + r


def test_next_large_no_overflow(orc):
    m, n = 250007, 199999
    N = m * n
    for i in [1, 2, N // 3, N - 2, N - 1, 123456789012]:
        assert orc.next_index(i, m, n) == (i if i == N - 1 else (i * m) % (N - 1))


def test_against_reference_plan_header(orc):
    ref = orc.ref_lib()
    if ref is None:
        pytest.skip("reference planner not built here (oracle/_ref needs the reference tree)")
    for m, n in [(2, 3), (3, 5), (18, 72), (17, 31), (64, 16), (100, 37)]:
        assert ref.ref_num_cycles(m, n) == orc.num_cycles(m, n)
        for i in range(m * n):
            assert ref.ref_next(i, m, n) == orc.next_index(i, m, n)
    for n in [2, 12, 97, 1800, 7200, 98304, 1024, 36]:
        buf = (__import__("ctypes").c_int * 4096)()
        c = ref.ref_tile_sizes(n, buf, 4096)
        assert sorted(buf[:c]) == orc.tile_sizes(n)


def test_tile_sizes_brute(orc):
    for n in [1, 2, 3, 12, 97, 360, 1800, 131071, 98304]:
        want = [d for d in range(2, n + 1) if n % d == 0] if n < 200000 else None
        got = orc.tile_sizes(n)
        if want is not None:
            assert got == want
        assert all(n % d == 0 for d in got)


def test_choose_tiles_alg5(orc):
    # Alg. 5 of test/cl_unittest.cc:396-467 with the SP limits min 24, max sqrt(MAX_MEM)=110
    aa, bb, ph, pw = orc.choose_tiles(1800, 7200, 24, 110)
    assert (aa, bb, ph, pw) == (100, 100, 0, 0)
    # a prime dimension gets padded until a divisor in range appears
    aa, bb, ph, pw = orc.choose_tiles(1009, 1000, 24, 110)
    assert pw == 0 and bb == 100
    assert ph > 0 and (1009 + ph) % aa == 0 and 24 <= aa <= 110
    for h in range(1009, 1009 + ph):
        assert not any(24 <= d <= 110 for d in orc.tile_sizes(h))


@pytest.mark.parametrize("op,es,dims,x,y", list(vectors.cases()),
                         ids=lambda v: str(v) if isinstance(v, (str, int)) else None)
def test_golden(orc, op, es, dims, x, y):
    x = x.copy()
    if op == "010":
        A, a, B = dims
        assert np.array_equal(orc.op_010(x, A, a, B, es=es), y)
        z = x.copy(); assert orc.inplace_010_bs(z, A, a, B, es=es) == 0
        assert np.array_equal(z, y)
        z = x.copy(); assert orc.inplace_010_pttwac(z, A, a, B, 2, 1, es=es) == 0
        assert np.array_equal(z, y)
    elif op == "100":
        A, B, b = dims
        assert np.array_equal(orc.op_100(x, A, B, b, es=es), y)
        z = x.copy(); assert orc.inplace_0100_pttwac(z, 1, A, B, b, es=es) == 0
        assert np.array_equal(z, y)
    elif op == "0100":
        A, a, B, b = dims
        assert np.array_equal(orc.op_0100(x, A, a, B, b, es=es), y)
        z = x.copy(); assert orc.inplace_0100_pttwac(z, A, a, B, b, es=es) == 0
        assert np.array_equal(z, y)
    elif op == "T":
        M, N = dims
        assert np.array_equal(orc.op_transpose(x, M, N, es=es), y)
        z = x.copy(); r, route = orc.inplace_transpose(z, 1, M, 1, N, 0, es=es)
        assert r == 0 and route == "Fallback-" and np.array_equal(z, y)
    elif op == "PAD":
        xs, ys, pad = dims
        z = x.copy(); orc.padding(z, es, xs, ys, pad)
        assert np.array_equal(z, y)
    elif op == "UNPAD":
        xs, ys, pad = dims
        z = x.copy(); orc.unpadding(z, es, xs, ys, pad)
        assert np.array_equal(z, y)


def _divs(n):
    return [d for d in range(1, n + 1) if n % d == 0]


@pytest.mark.parametrize("M,N", [(6, 10), (12, 8), (9, 15), (16, 16), (7, 12), (20, 6)])
@pytest.mark.parametrize("stages", [2, 22, 3, 32, 4, 1])
def test_composer_all_factorisations(orc, M, N, stages):
    """Every stage sequence of cl_transpose composes to the transpose, for every
    factorisation M = A*a, N = B*b (brute force), with 1-, 4- and 16-byte elements."""
    for es in (1, 4, 16):
        x = orc.tagged(M * N, es, seed=M * 131 + N)
        want = x.reshape(M, N, es).transpose(1, 0, 2).reshape(-1)
        for a, b in itertools.product(_divs(M), _divs(N)):
            for max_mem in (12288, 16):  # 16 forces the LDS-flag 010 and the fallback
                z = x.copy()
                r, route = orc.inplace_transpose(z, M // a, a, N // b, b, stages,
                                                 max_mem=max_mem, es=es)
                assert r == 0
                assert np.array_equal(z, want), (es, a, b, stages, max_mem, route)


def test_inplace_stage_forms_agree(orc):
    rng = np.random.default_rng(1)
    for _ in range(40):
        A, a, B = (int(v) for v in rng.integers(1, 20, 3))
        es = int(rng.choice(ES))
        x = orc.tagged(A * a * B, es, seed=int(rng.integers(1 << 30)))
        want = orc.op_010(x, A, a, B, es=es)
        z1, z2, z3 = x.copy(), x.copy(), x.copy()
        orc.inplace_010_bs(z1, A, a, B, es=es)
        orc.inplace_010_pttwac(z2, A, a, B, 1 << int(rng.integers(0, 5)), int(rng.integers(0, 2)), es=es)
        orc.par_010_bs(z3, A, a, B, nthreads=3, es=es)
        assert np.array_equal(z1, want) and np.array_equal(z2, want) and np.array_equal(z3, want)
        b = int(rng.integers(1, 9))
        x = orc.tagged(A * a * B * b, es, seed=int(rng.integers(1 << 30)))
        want = orc.op_0100(x, A, a, B, b, es=es)
        z1, z2 = x.copy(), x.copy()
        orc.inplace_0100_pttwac(z1, A, a, B, b, es=es)
        orc.par_0100_pttwac(z2, A, a, B, b, nthreads=4, es=es)
        assert np.array_equal(z1, want) and np.array_equal(z2, want)


def test_parallel_three_stage_config1_shape(orc):
    # 1800 x 7200 int32 with a = 18, b = 72 (the 3-stage route of src/cl_marshal.cc:636-677)
    M, N = 1800, 7200
    x = orc.tagged(M * N, 4, seed=7).view(np.uint32)
    want = x.reshape(M, N).T.reshape(-1).copy()
    assert orc.par_transpose_3(x, M // 18, 18, N // 72, 72, nthreads=0) == 0
    assert np.array_equal(x, want)


def test_oracle_detects_planted_index_error(orc):
    """A planted mistake (an off-by-one tile width, standing for a dropped term in the index
    map) must fail the golden comparison."""
    x = orc.tagged(3 * 4 * 5, 4, seed=5)
    good = orc.op_010(x, 3, 4, 5, es=4)
    bad = orc.op_010(x, 3, 5, 4, es=4)  # a and B swapped: a transposed operand
    assert not np.array_equal(good, bad)
    assert np.array_equal(good, x.reshape(3, 4, 5, 4).transpose(0, 2, 1, 3).reshape(-1))


@pytest.mark.parametrize("M,N,es", [(37, 53, 8), (1, 9, 4), (64, 3, 1), (30, 30, 16)])
def test_golden_transpose_is_the_transpose(orc, M, N, es):
    """The reference test's two-call CPU golden (a 010 with unit tiles, then a 100), which the
    benchmark times as its CPU baseline, is the plain transpose (numpy)."""
    x = orc.tagged(M * N, es, seed=M * N)
    y = x.copy()
    orc.golden_transpose(y, M, N, es=es)
    want = x.reshape(M, N, es).transpose(1, 0, 2).reshape(-1)
    assert np.array_equal(y, want)
