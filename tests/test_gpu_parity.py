"""Parity of the HIP path (called through the C ABI) with the CPU oracle and the golden vectors.

Shapes follow the reference's own suites: test/cl_unittest.cc (bug533/536/537, test_0100,
full) and test/unittest.cu (bug523-538). Every comparison is bit-exact on tagged elements.
"""
import itertools
import sys

import numpy as np
import pytest

from tests.golden import vectors

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


def check(fn, x, want, *args, **kw):
    t = dev(x)
    r = fn(t, *args, **kw)
    got = host(t)
    assert r is False, "library returned an error"
    if not np.array_equal(got, want):
        bad = np.nonzero(got != want)[0]
        raise AssertionError(f"{bad.size} bytes differ, first at byte {bad[0]}")


@pytest.mark.parametrize("op,es,dims,x,y", list(vectors.cases()),
                         ids=lambda v: str(v) if isinstance(v, (str, int)) else None)
def test_golden_vectors(hip, op, es, dims, x, y):
    say("golden", op, es, dims)
    if op == "010":
        A, a, B = dims
        check(hip.transpose_010_bs, x, y, A, a, B, elem_bytes=es)
        check(hip.transpose_010_pttwac, x, y, A, a, B, R=2, P=1, elem_bytes=es)
    elif op == "100":
        A, B, b = dims
        check(hip.transpose_100, x, y, A, B, b, elem_bytes=es)
    elif op == "0100":
        A, a, B, b = dims
        check(hip.transpose_0100, x, y, A, a, B, b, elem_bytes=es)
    elif op == "T":
        M, N = dims
        check(hip.transpose_full, x, y, M, N, elem_bytes=es)
        check(hip.transpose, x, y, 1, M, 1, N, 1, 1, elem_bytes=es)  # fallback route
    elif op == "PAD":
        xs, ys, pad = dims
        check(hip.padding, x, y, xs, ys, pad, elem_bytes=es)
    elif op == "UNPAD":
        xs, ys, pad = dims
        check(hip.unpadding, x, y, xs, ys, pad, elem_bytes=es)


def test_010_bs_bug533_shapes(hip, orc):
    # test/cl_unittest.cc:295-325: w = 3..384 doubling, t = 1..8, h = (500/w+1)*ceil(13000/t)*t
    for w in (3, 6, 12, 24, 48, 96, 192, 384):
        for t in range(1, 9):
            h = (500 // w + 1) * ((100 * 130 + t - 1) // t) * t
            es = (4, 1, 2, 8, 16)[(w + t) % 5]
            say("bug533", w, t, h, es)
            x = orc.tagged(h * w, es, seed=w * 10 + t)
            want = orc.op_010(x, h // t, t, w, es=es)
            check(hip.transpose_010_bs, x, want, h // t, t, w, elem_bytes=es)


def test_010_bs_subdword_units(hip, orc):
    # 1- and 2-byte elements move in dword units; tiles whose rows and columns are whole
    # units take carry-free LDS offsets, the others step element by element.
    for es in (1, 2):
        for a, B in ((64, 16), (4, 4), (8, 12), (12, 8), (64, 6), (6, 64), (7, 16), (16, 5),
                     (2, 2), (4, 2), (128, 36)):
            for A in (1, 3, 1001):
                say("bs subdword", es, A, a, B)
                x = orc.tagged(A * a * B, es, seed=A + a * 7 + B * 131 + es)
                check(hip.transpose_010_bs, x, orc.op_010(x, A, a, B, es=es), A, a, B,
                      elem_bytes=es)


def test_010_pttwac_bug536_shapes(hip, orc):
    # test/cl_unittest.cc:226-292: w in {197, 215}, t in {16, 32, 64}, R = 1..16, P = 1
    for w, h0 in ((197, 35588), (215, 44609)):
        for t in (16, 32, 64):
            h = (h0 + t - 1) // t * t
            x = orc.tagged(h * w, 4, seed=w + t)
            want = orc.op_010(x, h // t, t, w, es=4)
            for R in (1, 2, 4, 8, 16):
                say("bug536", w, t, R)
                check(hip.transpose_010_pttwac, x, want, h // t, t, w, R=R, P=1, elem_bytes=4)


@pytest.mark.parametrize("mode", ["lead", "lead_search", "lead_cached", "wave", "wg"])
@pytest.mark.parametrize("es", [1, 4, 8, 16])
def test_010_pttwac_element_sizes(hip, orc, es, mode):
    """Every form of the cycle-following 010: leader-only when all cycles are short ("lead",
    the default where it applies, with its leaders from the list ordered for few sectors per
    instruction; searched by every workgroup, "lead_search"; or from the cached list in search
    order, "lead_cached"), visited bits with one tile per wave ("wave") and visited bits with
    one tile per 256-lane workgroup ("wg"), forced with the knobs "pt010_lead_order",
    "pt010_lead_cached" and "pt010_wave"."""
    keep = {k: hip.get_tuning(k) for k in ("pt010_wave", "pt010_lead_order", "pt010_lead_cached")}
    hip.set_tuning("pt010_lead_order", 64 if mode.startswith("lead") else 0)
    hip.set_tuning("pt010_wave", 0 if mode == "wg" else 1)
    if mode == "lead_search":
        hip.set_tuning("pt010_lead_cached", 0)
    elif mode == "lead_cached":
        hip.set_tuning("pt010_lead_cached", 1)
    try:
        for (A, a, B) in ((50, 33, 47), (7, 64, 100), (3, 300, 211), (1000, 2, 3), (4099, 64, 16),
                          (257, 16, 64), (9, 8, 8), (5, 2, 1000)):
            say("pttwac", es, mode, A, a, B)
            x = orc.tagged(A * a * B, es, seed=A + a + B)
            for R, P in ((1, 1), (4, 0)):
                check(hip.transpose_010_pttwac, x, orc.op_010(x, A, a, B, es=es), A, a, B, R=R,
                      P=P, elem_bytes=es)
    finally:
        for k, v in keep.items():
            hip.set_tuning(k, v)


def test_010_pttwac_random_tiles(hip, orc):
    """Random tile shapes and element sizes through the cycle-following 010 with its default
    leader source (lists ordered per tile shape and element size, any start slot per cycle,
    padded rounds): every result bit-identical to the oracle, then the inverse restores."""
    rng = np.random.default_rng(2024)
    for k in range(40):
        a, B = (int(v) for v in rng.integers(2, 160, size=2))
        es = int(rng.choice([1, 2, 4, 8, 16]))
        A = int(rng.integers(1, 6))
        say("pttwac random", k, A, a, B, es)
        x = orc.tagged(A * a * B, es, seed=k + 77)
        want = orc.op_010(x, A, a, B, es=es)
        check(hip.transpose_010_pttwac, x, want, A, a, B, R=1, P=1, elem_bytes=es)
        check(hip.transpose_010_pttwac, want, x, A, B, a, R=1, P=1, elem_bytes=es)
    # power-of-two sides: the tile map is a rotation of the slot's bits, every cycle short, so
    # these all take the leader-only form
    for k in range(24):
        a, B = (1 << int(v) for v in rng.integers(1, 9, size=2))
        es = int(rng.choice([1, 2, 4, 8, 16]))
        A = int(rng.integers(1, 9))
        say("pttwac pow2", k, A, a, B, es)
        x = orc.tagged(A * a * B, es, seed=k + 177)
        want = orc.op_010(x, A, a, B, es=es)
        check(hip.transpose_010_pttwac, x, want, A, a, B, R=1, P=1, elem_bytes=es)
        check(hip.transpose_010_pttwac, want, x, A, B, a, R=1, P=1, elem_bytes=es)


def test_100_bug537_shapes(hip, orc):
    # test/cl_unittest.cc:171-223: 100(w, h/t, t), w in {197, 215}, t = 1..1024 powers of two
    for w, h0 in ((197, 35588), (215, 44609)):
        t = 1
        while t <= 1024:
            h = (h0 + t - 1) // t * t
            say("bug537", w, t)
            x = orc.tagged(h * w, 4, seed=w * t)
            check(hip.transpose_100, x, orc.op_100(x, w, h // t, t, es=4), w, h // t, t,
                  elem_bytes=4)
            t *= 2


@pytest.mark.parametrize("es", ES)
def test_100_element_sizes(hip, orc, es):
    for (A, B, b) in ((16, 4096, 64), (37, 53, 3), (5, 7, 1000), (1000, 3, 1), (2, 2, 5000)):
        say("100", es, A, B, b)
        x = orc.tagged(A * B * b, es, seed=A * B + b)
        check(hip.transpose_100, x, orc.op_100(x, A, B, b, es=es), A, B, b, elem_bytes=es)


def test_0100_reference_case(hip, orc):
    # test_0100, test/cl_unittest.cc:666-703: A=128, a=62, B=57, b=256
    A, a, B, b = 128, 62, 57, 256
    say("test_0100")
    x = orc.tagged(A * a * B * b, 4, seed=3)
    check(hip.transpose_0100, x, orc.op_0100(x, A, a, B, b, es=4), A, a, B, b, elem_bytes=4)


def test_marshal_h_twins(hip, orc):
    """The include/marshal.h entry points on the reference's CUDA suite shapes
    (test/unittest.cu:101-246)."""
    import ctypes
    import torch
    L = hip.lib()

    def run(fn, x, h, w, t):
        d = dev(x)
        r = fn(ctypes.c_void_p(d.data_ptr()), h, w, t, None)
        out = host(d)
        assert r is False
        return out

    # bug538: soa, h = 65538, t = 2, w = 2
    h, t, w = 65538, 2, 2
    x = orc.tagged(h * w, 4, seed=538)
    assert np.array_equal(run(L.gpu_soa_asta_pttwac, x, h, w, t), orc.op_100(x, w, h // t, t, es=4))
    # bug528: soa, h = 1024, t = 64, w = 1..99
    for w in range(1, 100, 7):
        say("bug528", w)
        h, t = 1024, 64
        x = orc.tagged(h * w, 4, seed=w)
        assert np.array_equal(run(L.gpu_soa_asta_pttwac, x, h, w, t),
                              orc.op_100(x, w, h // t, t, es=4))
    # bug525 / bug524: aos pttwac, h = 16384, t = 16, w = 1..99 and 66
    for w in list(range(1, 100, 7)) + [66]:
        say("bug525", w)
        h, t = 16 * 1024, 16
        x = orc.tagged(h * w, 4, seed=w)
        assert np.array_equal(run(L.gpu_aos_asta_pttwac, x, h, w, t),
                              orc.op_010(x, h // t, t, w, es=4))
    # bug523 / bug523d: BS float and double, w = 6, t = 16
    h, t, w = 16 * 1024, 16, 6
    x = orc.tagged(h * w, 4, seed=523)
    assert np.array_equal(run(L.gpu_aos_asta_bs_float, x, h, w, t), orc.op_010(x, h // t, t, w, es=4))
    x = orc.tagged(h * w, 8, seed=524)
    assert np.array_equal(run(L.gpu_aos_asta_bs_double, x, h, w, t), orc.op_010(x, h // t, t, w, es=8))
    # errors: height not a multiple of the tile
    d = torch.zeros(100, dtype=torch.float32, device="cuda")
    assert L.gpu_aos_asta_bs_float(ctypes.c_void_p(d.data_ptr()), 10, 10, 3, None) is True


def _divisors(n):
    return [d for d in range(1, n + 1) if n % d == 0]


@pytest.mark.parametrize("stages", [2, 22, 3, 32, 4, 1])
def test_transpose_routes_all_factorisations(hip, orc, stages):
    """cl_transpose's stage sequences (src/cl_marshal.cc:561-776) on every factorisation of
    small shapes, for 1-, 4- and 16-byte elements."""
    for (M, N) in ((12, 18), (30, 8), (7, 20)):
        for es in (1, 4, 16):
            x = orc.tagged(M * N, es, seed=M * N + es)
            want = orc.op_transpose(x, M, N, es=es)
            for a, b in itertools.product(_divisors(M), _divisors(N)):
                check(hip.transpose, x, want, M // a, a, N // b, b, 1, stages, elem_bytes=es)


@pytest.mark.parametrize("hw", [(1800, 7200), (2500, 5100), (3200, 4000), (3900, 3300),
                                (5100, 2500), (7200, 1800)])
def test_transpose_ppopp_sizes(hip, orc, hw):
    """The PPoPP'14 sizes of the reference's `full` test (test/cl_unittest.cc:386-387) with the
    tile factors of its Alg. 5 and the 3-stage routes of its Alg. 3, plus the planner route."""
    h, w = hw
    aa, bb, ph, pw = orc.choose_tiles(h, w, 24, 110)
    assert ph == 0 and pw == 0
    x = orc.tagged(h * w, 4, seed=h)
    want = orc.op_transpose(x, h, w, es=4)
    say("ppopp", h, w, aa, bb)
    check(hip.transpose, x, want, h // aa, aa, w // bb, bb, 1, 3 if aa <= bb else 32,
          elem_bytes=4)
    check(hip.transpose_full, x, want, h, w, elem_bytes=4)


def test_config1_three_stage_exact(hip, orc):
    # 1800 x 7200 int32, a = 18, b = 72 (A = B = 100), stages = 3
    M, N = 1800, 7200
    x = (orc.splitmix64(2024, M * N) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    want = x.reshape(M, N).T.copy()
    t = dev(x)
    et = hip.Elapsed()
    assert hip.transpose(t, M // 18, 18, N // 72, 72, 1, 3, elapsed=et) is False
    assert hip.last_route().endswith("3.1-")
    assert np.array_equal(host(t).reshape(N, M), want)
    assert et.ns > 0


def test_full_flow_with_padding(hip, orc):
    """The reference's `full` flow for a dimension with no divisor in range: pad the columns,
    transpose, unpad (test/cl_unittest.cc:526-543 / 607-628)."""
    h0, w0 = 1009, 1013  # both prime
    aa, bb, ph, pw = orc.choose_tiles(h0, w0, 24, 110)
    h, w = h0 + ph, w0 + pw
    x = orc.tagged(h * w, 4, seed=9)
    src = x.view(np.uint32)[: h0 * w0].reshape(h0, w0).copy()
    t = dev(x)
    assert hip.padding(t, w0, h0, w, elem_bytes=4) is False
    assert hip.transpose(t, h // aa, aa, w // bb, bb, 1, 3 if aa <= bb else 32,
                         elem_bytes=4) is False
    assert hip.unpadding(t, h0, w0, h, elem_bytes=4) is False
    got = host(t).view(np.uint32)[: h0 * w0].reshape(w0, h0)
    assert np.array_equal(got, src.T)


@pytest.mark.parametrize("es", ES)
def test_forced_routes_match(hip, orc, es):
    """Shrinking the LDS budget forces the LDS-flag and global-flag forms of 010; every route
    must give the same bytes (fallbacks forced from the host, as a device that cannot run the
    LDS form would force them, src/cl_marshal.cc:814-818)."""
    A, a, B = 40, 37, 29
    x = orc.tagged(A * a * B, es, seed=es)
    want = orc.op_010(x, A, a, B, es=es)
    try:
        for budget in (0, 2048, 64):
            hip.set_lds_budget(budget)
            say("forced", es, budget)
            check(hip.aos_asta, x, want, A * a, B, a, 0, elem_bytes=es)
        hip.set_lds_budget(64)
        assert hip.aos_asta(dev(x), A * a, B, a, 1, elem_bytes=es) is True  # BS cannot run
    finally:
        hip.set_lds_budget(0)


@pytest.mark.parametrize("es", ES)
def test_aos_soa_round_trips(hip, orc, es):
    h, w, t = 4096, 13, 64
    x = orc.tagged(h * w, es, seed=77)
    d = dev(x)
    assert hip.aos_asta(d, h, w, t, elem_bytes=es) is False
    assert np.array_equal(host(d), orc.op_010(x, h // t, t, w, es=es))
    assert hip.asta_aos(d, h, w, t, elem_bytes=es) is False
    assert np.array_equal(host(d), x)
    assert hip.soa_asta(d, h, w, t, elem_bytes=es) is False
    assert np.array_equal(host(d), orc.op_100(x, w, h // t, t, es=es))
    assert hip.asta_soa(d, h, w, t, elem_bytes=es) is False
    assert np.array_equal(host(d), x)


def test_config2_full_size_round_trip(hip, orc):
    """2^24 structs of 16 f32 fields (1 GiB), ASTA width 64: AoS->ASTA with the LDS-staged and
    the cycle-following kernel, back, and SoA->ASTA and back; bit-exact."""
    import torch
    n, w, t = 1 << 24, 16, 64
    g = torch.Generator(device="cuda").manual_seed(5)
    d = torch.rand(n * w, device="cuda", generator=g) * 2 - 1
    orig = d.clone()
    sample = torch.randint(0, n * w, (1 << 16,), device="cuda", generator=g)
    for method in (1, 2):
        say("config2 aos method", method)
        assert hip.aos_asta(d, n, w, t, method=method) is False
        # element (r, c) of AoS lands at ASTA (r // t, c, r % t)
        r, c = sample // w, sample % w
        dst = (r // t) * (w * t) + c * t + (r % t)
        assert torch.equal(d[dst].view(torch.int32), orig[sample].view(torch.int32))
        assert hip.asta_aos(d, n, w, t, method=method) is False
        assert torch.equal(d.view(torch.int32), orig.view(torch.int32))
    say("config2 soa")
    assert hip.soa_asta(d, n, w, t) is False
    # SoA element (c, r) (field-major) lands at ASTA (r // t, c, r % t)
    c, r = sample // n, sample % n
    dst = (r // t) * (w * t) + c * t + (r % t)
    assert torch.equal(d[dst].view(torch.int32), orig[sample].view(torch.int32))
    assert hip.asta_soa(d, n, w, t) is False
    assert torch.equal(d.view(torch.int32), orig.view(torch.int32))


@pytest.mark.parametrize("es", ES)
def test_0100_lds_and_global_bit_forms_agree(hip, orc, es):
    """The LDS-bit (one workgroup per instance) and global-bit forms of 0100 give the same
    bytes, including instances whose super-elements are split into pieces."""
    cases = ((100, 100, 72, 18), (40, 7, 9, 1), (3, 33, 17, 700), (64, 16, 64, 1),
             (5, 2, 3, 4096))
    try:
        for (A, a, B, b) in cases:
            x = orc.tagged(A * a * B * b, es, seed=A * a + B * b + es)
            want = orc.op_0100(x, A, a, B, b, es=es)
            for min_inst in (1, 1 << 20):
                hip.set_tuning("lds0100_min_inst", min_inst)
                for vw in (64, 8, 1):
                    hip.set_tuning("pt100_vw_max", vw)
                    say("0100 forms", es, (A, a, B, b), min_inst, vw)
                    check(hip.transpose_0100, x, want, A, a, B, b, elem_bytes=es)
    finally:
        hip.set_tuning("lds0100_min_inst", 32)
        hip.set_tuning("pt100_vw_max", 8)


@pytest.mark.parametrize("es", ES)
def test_row_column_decomposition_route(hip, orc, es):
    """stages = 5: the LDS-staged row/column decomposition, on coprime and non-coprime shapes,
    primes included (no factorisation needed)."""
    shapes = [(2, 3), (3, 2), (17, 19), (12, 18), (64, 48), (1000, 999), (997, 1009),
              (4096, 3001), (256, 4096), (5, 1024), (1024, 5), (2, 6000), (6000, 2)]
    for (M, N) in shapes:
        if M * N * es > 200 << 20 or max(M, N) * es > 160 * 1024:
            continue
        say("rc", es, M, N)
        x = orc.tagged(M * N, es, seed=M * 7 + N)
        want = orc.op_transpose(x, M, N, es=es)
        t = dev(x)
        assert hip.transpose(t, 1, M, 1, N, 1, 5, elem_bytes=es) is False
        assert hip.last_route() == "RC-"
        assert np.array_equal(host(t), want)


@pytest.mark.parametrize("es", [1, 4, 16])
def test_soa_small_tiles_take_row_column_route(hip, orc, es):
    """SoA->ASTA with a small tile is a transpose of b*es-byte super-elements; with the
    through-scratch forms switched off (scratch cap 0) odd super-element sizes are routed
    through the in-place row/column decomposition and must match the oracle's 100."""
    keep = hip.get_tuning("st_scratch_mb")
    hip.set_tuning("st_scratch_mb", 0)
    try:
        for (h, w, t) in ((999, 37, 1), (1002, 13, 3), (4096, 7, 2), (35, 4096, 5)):
            say("soa small tile", es, h, w, t)
            x = orc.tagged(h * w, es, seed=h + w + t)
            check(hip.soa_asta, x, orc.op_100(x, w, h // t, t, es=es), h, w, t, elem_bytes=es)
            route = hip.last_route()
            assert "OOP" not in route and "ST_" not in route, route
            if t * es not in (1, 2, 4, 8, 16) and t * es < hip.get_tuning("rc_max_super_bytes"):
                assert route == "0100_RC-", route
    finally:
        hip.set_tuning("st_scratch_mb", keep)


@pytest.mark.parametrize("es", ES)
def test_stripe_route(hip, orc, es):
    """stages = 6 / 7: stripes of rows / of columns through bounded scratch, with and without a
    ragged remainder (prime dimensions included), several scratch fills per call."""
    shapes = [(768, 131), (131, 768), (1031, 997), (640, 4099), (4099, 640), (4096, 3001)]
    keep = {k: hip.get_tuning(k)
            for k in ("st_super_bytes", "st_scratch_mb", "st_rotate", "st_max_height")}
    try:
        # small scratch: rotated stripes, or (rotation off) several scratch fills; large
        # scratch: one fill
        for super_bytes, scratch_mb, rotate, hmax in ((64, 1, 1, 1024), (64, 1, 0, 1024),
                                                      (2048, 256, 1, 1024), (8192, 256, 1, 4096)):
            hip.set_tuning("st_super_bytes", super_bytes)
            hip.set_tuning("st_scratch_mb", scratch_mb)
            hip.set_tuning("st_rotate", rotate)
            hip.set_tuning("st_max_height", hmax)
            for (M, N) in shapes + ([(9001, 1031), (1031, 9001)] if hmax > 1024 else []):
                if M * N * es > 100 << 20:
                    continue
                x = orc.tagged(M * N, es, seed=M * 3 + N + es)
                want = orc.op_transpose(x, M, N, es=es)
                for stages, tag in ((6, "ST_R-"), (7, "ST_C-")):
                    say("stripes", es, M, N, stages, super_bytes, scratch_mb, rotate)
                    t = dev(x)
                    assert hip.transpose(t, 1, M, 1, N, 1, stages, elem_bytes=es) is False
                    route = hip.last_route()
                    got = host(t)
                    if route != tag:  # too few stripes for this shape: another route ran
                        assert route in ("RC-", "Fallback-"), route
                    assert np.array_equal(got, want), (M, N, stages, route)
    finally:
        for k, v in keep.items():
            hip.set_tuning(k, v)


@pytest.mark.parametrize("es", ES)
def test_aos_asta_large_tiles_split(hip, orc, es):
    """AoS->ASTA tiles too large for LDS (t x w elements) with the through-scratch form
    switched off (scratch cap 0) take the in-place split route 0100 + 010 (or the row/column
    decomposition); both directions must match the oracle's 010, with a ragged and a prime
    tile width."""
    keep = hip.get_tuning("st_scratch_mb")
    hip.set_tuning("st_scratch_mb", 0)
    try:
        for (h, w, t) in ((640, 4096, 64), (128, 3000, 64), (192, 1024, 32), (256, 4099, 64)):
            if h * w * es > 64 << 20:
                continue
            say("aos large tile", es, h, w, t)
            x = orc.tagged(h * w, es, seed=h * 31 + w + t)
            check(hip.aos_asta, x, orc.op_010(x, h // t, t, w, es=es), h, w, t, elem_bytes=es)
            assert "OOP" not in hip.last_route()
            y = orc.op_010(x, h // t, t, w, es=es)
            check(hip.asta_aos, y, x, h, w, t, elem_bytes=es)
    finally:
        hip.set_tuning("st_scratch_mb", keep)


def test_autotune_restores_buffer_and_sets_plan(hip, orc):
    """The measured autotuner times candidate routes in place: the buffer must come back
    bit-identical, the chosen stage code becomes the plan, and the planned transpose matches
    the oracle."""
    for (M, N, es) in ((1000, 1000, 8), (1031, 997, 1), (1800, 7200, 4), (333, 4099, 16)):
        say("autotune", M, N, es)
        x = orc.tagged(M * N, es, seed=M + N + es)
        d = dev(x)
        stages = hip.autotune(d, M, N, elem_bytes=es)
        assert np.array_equal(host(d), x)
        assert hip.plan_transpose(M, N, es)[2] == stages
        assert hip.transpose_full(d, M, N, elem_bytes=es) is False
        assert np.array_equal(host(d), orc.op_transpose(x, M, N, es=es))


@pytest.mark.parametrize("es", ES)
def test_through_scratch_route(hip, orc, es):
    """stages = 9: a matrix within the scratch cap is copied to scratch and transposed back;
    rows of any byte alignment, primes and vectors' neighbours included."""
    for (M, N) in ((2, 3), (17, 19), (300, 200), (1031, 997), (64, 4099), (4099, 64)):
        if M * N * es > 64 << 20:
            continue
        say("oop", es, M, N)
        x = orc.tagged(M * N, es, seed=M * 5 + N)
        d = dev(x)
        assert hip.transpose(d, 1, M, 1, N, 1, 9, elem_bytes=es) is False
        assert hip.last_route() == "OOP-"
        assert np.array_equal(host(d), orc.op_transpose(x, M, N, es=es))


@pytest.mark.parametrize("es", ES)
def test_through_scratch_tile_rows(hip, orc, es):
    """Tile shapes of the out-of-place transpose: no minimum, the default, and 1024 rows
    (few columns) give the same bytes on aligned and misaligned rows."""
    for min_rows in (0, 1024):
        for (M, N) in ((1031, 997), (512, 3000), (4099, 64), (300, 2048)):
            if M * N * es > 32 << 20:
                continue
            say("tile rows", min_rows, es, M, N)
            x = orc.tagged(M * N, es, seed=M + N + min_rows)
            d = dev(x)
            keep = hip.get_tuning("oop_min_rows")
            try:
                hip.set_tuning("oop_min_rows", min_rows)
                assert hip.transpose(d, 1, M, 1, N, 1, 9, elem_bytes=es) is False
            finally:
                hip.set_tuning("oop_min_rows", keep)
            assert np.array_equal(host(d), orc.op_transpose(x, M, N, es=es)), (min_rows, M, N)


@pytest.mark.parametrize("es", ES)
def test_through_scratch_pitched_form(hip, orc, es):
    """Destination rows that are not 16-byte aligned take the pitched form (transpose into
    scratch rows of a 16-byte multiple pitch, then one contiguous copy back in aligned chunks,
    rows meeting inside chunks). Bases offset by 0..15 elements, rows just above the length
    threshold, and a 1 MB cap; the plain form gives the same bytes, and the bytes around the
    matrix stay put."""
    cases = ((257, 301), (1031, 997), (300, 200), (97, 4099), (4099, 65), (16 * 17 + 1, 129))
    for pitched in (2, 0):
        for cap in (0, 1):
            for k, (M, N) in enumerate(cases):
                if M * N * es > (1 << 20 if cap else 32 << 20) or M * es < 256:
                    continue
                off = (k * 5 + es) % 16  # elements before the matrix
                n = M * N
                x = orc.tagged(n + off + 3, es, seed=M + N * 7 + off)
                d = dev(x)
                sub = d[off * es:(off + n) * es]
                say("pitched", pitched, cap, es, M, N, off)
                keep = {kk: hip.get_tuning(kk) for kk in ("oop_pitched", "st_scratch_mb")}
                try:
                    hip.set_tuning("oop_pitched", pitched)
                    if cap:
                        hip.set_tuning("st_scratch_mb", 1)
                    assert hip.transpose(sub, 1, M, 1, N, 1, 9, elem_bytes=es) is False
                    assert hip.last_route() == "OOP-"
                finally:
                    for kk, v in keep.items():
                        hip.set_tuning(kk, v)
                got = host(d)
                lo, hi = off * es, (off + n) * es
                assert np.array_equal(got[:lo], x[:lo]) and np.array_equal(got[hi:], x[hi:])
                assert np.array_equal(got[lo:hi], orc.op_transpose(x[lo:hi], M, N, es=es)), \
                    (pitched, cap, es, M, N, off)
