"""The int-typed twins of include/cl_marshal.h, called the way a ported caller binds them
(ctypes, 32-bit int dimensions, cl_command_queue -> hipStream_t, cl_mem -> void*,
cl_ulong* -> uint64_t*), on the shapes of the reference's own suite (test/cl_unittest.cc:
bug537 171-224, bug536 226-292, bug533 295-342, full 344-664, test_0100 666-703), checked
bit-exactly against the oracle.

Also the reference's conventions on those twins (include/cl_profile.h:23-56,
src/cl_marshal.cc:269-275): a non-NULL elapsed counter is accumulated with +=, a NULL one makes
the call print "[Profile] <name> <ns>ns; <GB/s> GB/s" on stderr; false = success, true = error.
And the two forced forms: the element-granular "Fallback-" route (src/cl_marshal.cc:769-775)
and cl_aos_asta's cycle-following retry when the barrier-sync form cannot run
(src/cl_marshal.cc:814-818), both forced by lowering the LDS budget from the host.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Dev:
    """A device copy of a host byte array and the raw pointer a C caller would pass."""

    def __init__(self, x):
        import torch
        self.t = torch.from_numpy(np.ascontiguousarray(x)).cuda()
        self.p = ctypes.c_void_p(self.t.data_ptr())

    def get(self):
        import torch
        torch.cuda.synchronize()
        return self.t.cpu().numpy()


def stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def counter(v=0):
    return ctypes.c_uint64(v)


@pytest.fixture(autouse=True)
def _profiling_on(hip):
    hip.set_profiling(True)  # the drop-in's default: elapsed times are measured
    yield
    hip.set_lds_budget(0)


def test_aos_asta_bs_bug533(hip, orc):
    """cl_aos_asta_bs(queue, mem, h, w, t): test/cl_unittest.cc:295-325."""
    L = hip.lib()
    for w in (3, 24, 384):
        for t in (1, 5, 8):
            h = (500 // w + 1) * ((100 * 130 + t - 1) // t) * t
            say("bug533 twin", w, t, h)
            x = orc.tagged(h * w, 4, seed=w * 7 + t)
            d = Dev(x)
            assert L.hip_aos_asta_bs(stream(), d.p, h, w, t) is False
            assert np.array_equal(d.get(), orc.op_010(x, h // t, t, w, es=4))
    # a height that is not a multiple of the tile: like the reference (height / tile_size tiles,
    # src/cl_marshal.cc:792-799) the whole tiles are converted and the tail rows are left alone
    x = orc.tagged(100, 4, seed=1)
    d = Dev(x)
    assert L.hip_aos_asta_bs(stream(), d.p, 10, 10, 3) is False
    got = d.get()
    assert np.array_equal(got[:360], orc.op_010(x[:360], 3, 3, 10, es=4))
    assert np.array_equal(got[360:], x[360:])


def test_aos_asta_pttwac_bug536(hip, orc):
    """cl_aos_asta_pttwac(queue, mem, h, w, t, R, P) and cl_transpose_010_pttwac(queue, mem, A,
    a, B, &et, R, P): test/cl_unittest.cc:226-292 (w = 197, 215; t = 16, 32, 64; R = 1..32
    flag spreading, P = 0/1 padding)."""
    L = hip.lib()
    for w, h0 in ((197, 35588), (215, 44609)):
        for t in (16, 32, 64):
            h = (h0 + t - 1) // t * t
            x = orc.tagged(h * w, 4, seed=w + t)
            want = orc.op_010(x, h // t, t, w, es=4)
            for R, P in ((1, 0), (4, 1), (32, 1)):
                say("bug536 twin", w, t, R, P)
                d = Dev(x)
                assert L.hip_aos_asta_pttwac(stream(), d.p, h, w, t, R, P) is False
                assert np.array_equal(d.get(), want)
            d = Dev(x)
            et = counter()
            assert L.hip_transpose_010_pttwac(stream(), d.p, h // t, t, w, ctypes.byref(et),
                                              2, 1) is False
            assert np.array_equal(d.get(), want)
            assert et.value > 0
    d = Dev(orc.tagged(64, 4))
    assert L.hip_aos_asta_pttwac(stream(), d.p, 8, 8, 3, 1, 1) is True  # 8 % 3 != 0
    assert L.hip_transpose_010_pttwac(stream(), d.p, 2, 4, 8, None, 0, 1) is True  # R < 1


def test_aos_asta_retries_cycle_following(hip, orc):
    """cl_aos_asta(queue, mem, h, w, t, R): the barrier-sync form first, the cycle-following
    form when it fails (src/cl_marshal.cc:814-818). With a 2 KiB LDS budget a 64 x 197 tile
    cannot be staged, but its visited bits fit."""
    L = hip.lib()
    h, w, t = 64 * 50, 197, 64
    x = orc.tagged(h * w, 4, seed=5)
    want = orc.op_010(x, h // t, t, w, es=4)
    d = Dev(x)
    assert L.hip_aos_asta(stream(), d.p, h, w, t, 1) is False
    assert hip.last_route() == "010_BS-"
    assert np.array_equal(d.get(), want)
    hip.set_lds_budget(2048)
    d = Dev(x)
    assert L.hip_aos_asta_bs(stream(), d.p, h, w, t) is True  # the staged form cannot run
    assert L.hip_aos_asta(stream(), d.p, h, w, t, 1) is False
    assert hip.last_route() == "010_PTTWAC-"
    assert np.array_equal(d.get(), want)


def test_transpose_010_bs_and_accumulation(hip, orc):
    """cl_transpose_010_bs(queue, mem, A, a, B, &et): two calls on one counter add up."""
    L = hip.lib()
    A, a, B = 4096, 16, 24
    x = orc.tagged(A * a * B, 4, seed=11)
    d = Dev(x)
    et = counter()
    assert L.hip_transpose_010_bs(stream(), d.p, A, a, B, ctypes.byref(et)) is False
    first = et.value
    assert first > 0
    assert np.array_equal(d.get(), orc.op_010(x, A, a, B, es=4))
    # the inverse on the same counter: the counter holds the sum of both calls
    assert L.hip_transpose_010_bs(stream(), d.p, A, B, a, ctypes.byref(et)) is False
    assert et.value > first
    assert np.array_equal(d.get(), x)
    # accumulation from a caller-chosen start value (the reference's callers zero it)
    big = 1 << 40
    et2 = counter(big)
    assert L.hip_transpose_010_bs(stream(), d.p, A, a, B, ctypes.byref(et2)) is False
    assert et2.value > big


def test_transpose_100_bug537(hip, orc):
    """cl_transpose_100(queue, mem, w, h/t, t, &et): test/cl_unittest.cc:171-224."""
    L = hip.lib()
    for w, h0 in ((197, 35588), (215, 44609)):
        for t in (1, 4, 32, 256, 1024):
            h = (h0 + t - 1) // t * t
            say("bug537 twin", w, t)
            x = orc.tagged(h * w, 4, seed=w * 3 + t)
            d = Dev(x)
            et = counter()
            assert L.hip_transpose_100(stream(), d.p, w, h // t, t, ctypes.byref(et)) is False
            assert np.array_equal(d.get(), orc.op_100(x, w, h // t, t, es=4))
            assert et.value > 0


def test_transpose_0100_reference_case(hip, orc):
    """cl_transpose_0100(queue, mem, A, a, B, b, NULL): test/cl_unittest.cc:666-703, whose
    expected result is A instances of cpu_soa_asta (test/cl_unittest.cc:693-696)."""
    L = hip.lib()
    A, a, B, b = 128, 62, 57, 256
    x = orc.tagged(A * a * B * b, 4, seed=666)
    d = Dev(x)
    assert L.hip_transpose_0100(stream(), d.p, A, a, B, b, ctypes.byref(counter())) is False
    got = d.get()
    want = np.concatenate([orc.op_100(x[i * a * B * b * 4:(i + 1) * a * B * b * 4], a, B, b, es=4)
                           for i in range(A)])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("stages", [2, 22, 3, 32, 4])
def test_transpose_twin_paper_sizes(hip, orc, stages):
    """cl_transpose(queue, mem, A, a, B, b, R, stages, &et) with the PPoPP'14 sizes
    (test/cl_unittest.cc:386-387) and tile factors from the reference's Alg. 5; two calls (the
    transpose and its inverse) accumulate into one counter."""
    L = hip.lib()
    for (h, w) in ((1800, 7200), (3900, 3300)):
        aa, bb, ph, pw = orc.choose_tiles(h, w, 24, 110)
        assert ph == 0 and pw == 0
        say("cl_transpose twin", h, w, aa, bb, stages)
        x = orc.tagged(h * w, 4, seed=h + stages)
        d = Dev(x)
        et = counter()
        assert L.hip_transpose(stream(), d.p, h // aa, aa, w // bb, bb, 1, stages,
                               ctypes.byref(et)) is False
        route = hip.last_route()
        tag = {2: "2.1-", 22: "2.2-", 3: "3.1-", 32: "3.2-", 4: "4-"}[stages]
        assert route.endswith(tag), route
        assert np.array_equal(d.get(), orc.op_transpose(x, h, w, es=4))
        one = et.value
        assert one > 0
        assert L.hip_transpose(stream(), d.p, w // bb, bb, h // aa, aa, 1, stages,
                               ctypes.byref(et)) is False
        assert et.value > one
        assert np.array_equal(d.get(), x)


def test_full_flow_padding_twins(hip, orc):
    """cl_padding / cl_transpose / cl_unpadding as the reference's `full` test chains them
    (test/cl_unittest.cc:561-583: pad the columns, 22-route transpose, unpad the rows)."""
    L = hip.lib()
    h0, w0 = 1009, 1013  # both prime: no divisor in the tile range
    aa, bb, ph, pw = orc.choose_tiles(h0, w0, 24, 110)
    h, w = h0 + ph, w0 + pw
    x = orc.tagged(h * w, 4, seed=44)
    src = x.view(np.uint32)[: h0 * w0].reshape(h0, w0).copy()
    d = Dev(x)
    e_pad, e_tr, e_unpad = counter(), counter(), counter()
    if pw:
        assert L.hip_padding(stream(), d.p, w0, h0, w, ctypes.byref(e_pad)) is False
        assert e_pad.value > 0
    assert L.hip_transpose(stream(), d.p, h // aa, aa, w // bb, bb, 1, 3 if aa <= bb else 32,
                           ctypes.byref(e_tr)) is False
    if ph:
        assert L.hip_unpadding(stream(), d.p, h0, w0, h, ctypes.byref(e_unpad)) is False
        assert e_unpad.value > 0
    got = d.get().view(np.uint32)[: h0 * w0].reshape(w0, h0)
    assert np.array_equal(got, src.T)
    # the padding twins alone, against the oracle
    xs, ys, pad = 1000, 333, 1024
    x = orc.tagged(ys * pad, 4, seed=45)
    d = Dev(x)
    assert L.hip_padding(stream(), d.p, xs, ys, pad, ctypes.byref(counter())) is False
    want = x.copy()
    orc.padding(want, 4, xs, ys, pad)
    assert np.array_equal(d.get(), want)
    orc.unpadding(want, 4, xs, ys, pad)
    assert L.hip_unpadding(stream(), d.p, xs, ys, pad, ctypes.byref(counter())) is False
    assert np.array_equal(d.get(), want)
    assert L.hip_padding(stream(), d.p, 10, 5, 9, None) is True  # x_size > pad_size


def test_null_elapsed_prints_profile_line(hip, orc, capfd):
    """A NULL elapsed pointer prints the reference's profile line (include/cl_profile.h:43-56)."""
    L = hip.lib()
    A, a, B = 256, 16, 24
    d = Dev(orc.tagged(A * a * B, 4, seed=3))
    capfd.readouterr()
    assert L.hip_transpose_010_bs(stream(), d.p, A, a, B, None) is False
    assert L.hip_transpose_100(stream(), d.p, A, a * B // 8, 8, None) is False
    err = capfd.readouterr().err
    lines = [ln for ln in err.splitlines() if ln.startswith("[Profile]")]
    assert len(lines) == 2, err
    for ln in lines:
        assert "ns" in ln and ln.rstrip().endswith("GB/s"), ln


def test_forced_fallback_route(hip, orc):
    """The element-granular fallback, cl_transpose's last resort (src/cl_marshal.cc:769-775):
    with a 64-byte LDS budget neither the row/column route nor any staged form can run, so an
    unknown stage code lands on "Fallback-" (a 100 over single elements with global bits)."""
    L = hip.lib()
    hip.set_lds_budget(64)
    for (M, N, es) in ((97, 131, 4), (67, 90, 1), (33, 70, 16)):  # rows and columns > 64 B
        x = orc.tagged(M * N, es, seed=M + N)
        d = Dev(x)
        assert hip.transpose(d.t, 1, M, 1, N, 1, 1, elem_bytes=es) is False
        assert hip.last_route() == "Fallback-", hip.last_route()
        assert np.array_equal(d.get(), orc.op_transpose(x, M, N, es=es))
    x = orc.tagged(97 * 131, 4, seed=9)
    d = Dev(x)
    assert L.hip_transpose(stream(), d.p, 1, 97, 1, 131, 1, 1, None) is False
    assert hip.last_route() == "Fallback-"
    assert np.array_equal(d.get(), orc.op_transpose(x, 97, 131, es=4))
