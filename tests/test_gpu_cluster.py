"""The cluster-staged form of 100 / 0100 steps (route tag "0100_CL-"): G workgroups hold one
instance on chip (LDS and registers), meet at a barrier and write it back transposed. Checked
bit-exactly against the oracle for every piece width (16-, 8- and 4-byte pieces), clusters of
1 to 16 workgroups, more workgroups than the device holds at once (clusters formed over
several rounds), a 4-byte-aligned base, repeated calls (the barrier words are cleared by each
launch) and the inverse step restoring the input."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture
def anywhere(hip):
    """The cluster form wherever it can run (by default only many instances of 256 KB or more
    take it)."""
    keep = {k: hip.get_tuning(k) for k in ("cl_min_bytes", "cl_min_inst")}
    hip.set_tuning("cl_min_bytes", 0)
    hip.set_tuning("cl_min_inst", 1)
    yield
    for k, v in keep.items():
        hip.set_tuning(k, v)


CASES = [  # (instances, R, C, b, elem_bytes, base offset in elements)
    (100, 100, 72, 18, 4, 0),    # 72-byte super-elements, 518 KB instances (config 1 stage 3)
    (3, 97, 131, 5, 16, 0),      # 80-byte, 16-byte pieces, 1 MB instances
    (2, 64, 33, 200, 8, 0),      # 1600-byte, 3.4 MB instances
    (1, 30, 40, 1024, 4, 0),     # 4 KiB super-elements, 4.9 MB: 14 workgroups
    (1, 128, 100, 45, 8, 0),     # 360-byte, 4.6 MB: 16 workgroups of 8-byte pieces
    (5, 41, 29, 25, 4, 0),       # 100-byte: 4-byte pieces, one workgroup per instance
    (600, 17, 23, 16, 4, 0),     # 600 small instances: more workgroups than CUs
    (4, 60, 25, 24, 4, 1),       # 96-byte super-elements at a 4-byte-aligned base
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c)))
def test_cluster_0100_matches_oracle(hip, orc, anywhere, case):
    import torch
    inst, R, C, b, es, off = case
    n = inst * R * C * b
    x = orc.tagged(n, es, seed=R * C + b + inst)
    want = orc.op_0100(x, inst, R, C, b, es=es)
    buf = torch.zeros((n + off) * es, dtype=torch.uint8, device="cuda")
    buf[off * es:] = torch.from_numpy(x.view(np.uint8)).cuda()
    d = buf[off * es:]
    for rep in range(2):  # forward, back, forward again: the barrier words start from zero
        assert hip.transpose_0100(d, inst, R, C, b, elem_bytes=es) is False
        assert hip.last_route() == "0100_CL-", (case, hip.last_route())
        got = buf.cpu().numpy()
        assert np.array_equal(got[off * es:], want.view(np.uint8)), (case, rep)
        assert not got[:off * es].any()
        assert hip.transpose_0100(d, inst, C, R, b, elem_bytes=es) is False
        assert hip.last_route() == "0100_CL-", (case, hip.last_route())
        assert np.array_equal(buf.cpu().numpy()[off * es:], x.view(np.uint8)), (case, rep)
    say("cluster", case, "ok")


def test_cluster_limits_fall_back(hip, orc, anywhere):
    """Instances beyond the cluster cap, super-elements under 64 bytes or not a multiple of 4
    take the other forms, and the knob "cl_max_wgs" = 0 switches the form off."""
    import torch
    keep = hip.get_tuning("cl_max_wgs")
    try:
        for case, knob in (((1, 60, 40, 1024, 4), 16),   # 9.8 MB: more than 16 workgroups
                           ((3, 40, 30, 15, 4), 16),     # 60-byte super-elements
                           ((2, 37, 45, 65, 1), 16),     # 65-byte: no 4-byte pieces
                           ((100, 100, 72, 18, 4), 0)):  # switched off
            hip.set_tuning("cl_max_wgs", knob)
            inst, R, C, b, es = case
            n = inst * R * C * b
            x = orc.tagged(n, es, seed=n)
            d = torch.from_numpy(x.view(np.uint8).copy()).cuda()
            assert hip.transpose_0100(d, inst, R, C, b, elem_bytes=es) is False
            assert hip.last_route() != "0100_CL-", case
            assert np.array_equal(d.cpu().numpy(), orc.op_0100(x, inst, R, C, b, es=es).view(np.uint8))
    finally:
        hip.set_tuning("cl_max_wgs", keep)


def test_default_takes_cluster_form_for_many_mid_sized_instances(hip, orc):
    """By default the form is taken for config 1's stage-3 shape (100 instances of 518 KB) and
    not for a single 4 MB grid, where the segmented and cycle-leader forms measured faster."""
    import torch
    for case, want in (((100, 100, 72, 18, 4), True), ((1, 16, 1024, 64, 4), False),
                       ((16, 64, 64, 64, 4), False)):
        inst, R, C, b, es = case
        n = inst * R * C * b
        x = orc.tagged(n, es, seed=n + 1)
        d = torch.from_numpy(x.view(np.uint8).copy()).cuda()
        assert hip.transpose_0100(d, inst, R, C, b, elem_bytes=es) is False
        assert (hip.last_route() == "0100_CL-") == want, (case, hip.last_route())
        assert np.array_equal(d.cpu().numpy(), orc.op_0100(x, inst, R, C, b, es=es).view(np.uint8))


@pytest.mark.parametrize("use_graph", [True, False])
def test_cluster_step_in_a_batch(hip, orc, anywhere, use_graph):
    """A 4 MB SoA -> ASTA conversion runs on a batch's side stream in the cluster form; under
    graph capture its barrier words come from the graph's scratch and are cleared by the
    launch on every replay. Each replay applies the step again: checked after every one."""
    import torch
    from libmarshal_amd import workloads as wl
    es, h, w, t = 4, 64 * 256, 64, 64  # 100 over a 64 x 256 grid of 256-byte super-elements
    x = orc.tagged(h * w, es, seed=31)
    d = torch.from_numpy(x.view(np.uint8).copy()).cuda()
    small = orc.tagged(32 * 48, es, seed=32)  # a grouped item beside it
    ds = torch.from_numpy(small.view(np.uint8).copy()).cuda()
    batch = hip.Batch([(d, wl.OP_SOA_ASTA, es, h, w, t), (ds, wl.OP_TRANSPOSE, es, 32, 48, 0)],
                      use_graph=use_graph)
    try:
        assert batch.route(0).endswith("0100_CL-"), batch.route(0)
        want, want_s = x, small
        shape = (32, 48)
        for rep in range(3):
            assert batch.run() is False
            torch.cuda.synchronize()
            want = orc.op_100(want, w, h // t, t, es=es)
            assert np.array_equal(d.cpu().numpy(), want.view(np.uint8)), rep
            want_s = orc.op_transpose(want_s, shape[0], shape[1], es=es)  # same descriptor
            assert np.array_equal(ds.cpu().numpy(), want_s.view(np.uint8)), rep
    finally:
        batch.close()


FUSED = [  # (M, N, a, b, elem_bytes): the 3-stage route; its stages 2 and 3 fuse
    (1800, 7200, 18, 72, 4),     # config 1: 72-byte super-elements, 16-byte tile-row vectors
    (7200, 1800, 72, 18, 4),     # and back: 288-byte super-elements, 8-byte vectors
    (1800, 7200, 180, 180, 4),   # one 130 KB tile per workgroup
    (1200, 2400, 24, 48, 8),     # 8-byte elements
    (960, 1920, 16, 24, 8),      # 128-byte super-elements, 192-byte tile rows
]


@pytest.mark.parametrize("case", FUSED, ids=lambda c: "x".join(map(str, c)))
def test_three_stage_route_fuses_stages_2_and_3(hip, orc, case):
    """The 3-stage route's 010 (stage 2) and 0100 (stage 3) run as one cluster-staged pass:
    no 010 kernel is launched, and the result equals the oracle's transpose; with the fusion
    switched off the same route runs both stages and gives the same bytes."""
    import torch
    M, N, a, b, es = case
    x = orc.tagged(M * N, es, seed=M + N + a)
    want = orc.op_transpose(x, M, N, es=es)
    keep = hip.get_tuning("cl_fuse")
    try:
        for fuse in (1, 0):
            hip.set_tuning("cl_fuse", fuse)
            d = torch.from_numpy(x.view(np.uint8).copy()).cuda()
            torch.cuda.synchronize()
            hip.set_kernel_timing(True)
            assert hip.transpose(d, M // a, a, N // b, b, 1, 3, elem_bytes=es) is False
            names = {k["name"] for k in hip.kernel_times()}
            hip.set_kernel_timing(False)
            assert hip.last_route().endswith("3.1-"), hip.last_route()
            assert np.array_equal(d.cpu().numpy(), want.view(np.uint8)), (case, fuse, names)
            if fuse:
                assert "k_100_cluster" in names and "k_010_bs" not in names, names
            else:
                assert "k_010_bs" in names, names
            say("fused", case, fuse, sorted(names))
    finally:
        hip.set_tuning("cl_fuse", keep)


@pytest.mark.parametrize("case", [
    (1800, 7200, 18, 72, 4, 1),   # base 4-byte aligned only: no 8-byte vectors, not fused
    (600, 1200, 6, 12, 4, 0),     # 24-byte super-elements (< 64): not fused
    (1024, 2048, 8, 16, 8, 0),    # 64-byte super-elements, 128-byte tile rows
    (2000, 3000, 40, 60, 4, 2),   # base 8-byte aligned: 8-byte vectors
])
def test_three_stage_route_fusion_edges(hip, orc, case):
    """Bases of any element alignment and super-elements around the 64-byte limit: the route
    fuses where it can and every result equals the oracle's transpose, with the bytes around
    the matrix untouched."""
    import torch
    M, N, a, b, es, off = case
    x = orc.tagged(M * N, es, seed=M * 3 + b)
    want = orc.op_transpose(x, M, N, es=es)
    buf = torch.zeros((M * N + off + 1) * es, dtype=torch.uint8, device="cuda")
    buf[off * es:(off + M * N) * es] = torch.from_numpy(x.view(np.uint8)).cuda()
    d = buf[off * es:(off + M * N) * es]
    assert hip.transpose(d, M // a, a, N // b, b, 1, 3, elem_bytes=es) is False
    got = buf.cpu().numpy()
    assert np.array_equal(got[off * es:(off + M * N) * es], want.view(np.uint8)), case
    assert not got[:off * es].any() and not got[(off + M * N) * es:].any()
    say("fusion edge", case, hip.last_route())


def test_cluster_that_gives_up_is_reported(hip, orc, anywhere):
    """A cluster whose workgroups cannot all arrive within the barrier's bounded wait gives up
    and stores nothing, so its instance stays unconverted; that must be reported, never
    silent. Forced from the host with a one-poll bound ("cl_spin" = 1: every wait still
    ends): a synchronizing call returns an error exactly when an instance was left untouched,
    and asynchronous calls report the clusters that gave up through cluster_aborts (one
    cluster per instance here, so the count equals the untouched instances). Every instance
    is either converted or untouched."""
    import torch
    inst, R, C, b, es = 100, 100, 72, 18, 4  # 518 KB instances, four workgroups each
    n = inst * R * C * b
    x = orc.tagged(n, es, seed=31)
    want = orc.op_0100(x, inst, R, C, b, es=es).view(np.uint8).reshape(inst, -1)
    orig = x.view(np.uint8).reshape(inst, -1)
    assert hip.cluster_aborts() == 0

    def outcome(d):
        got = d.cpu().numpy().reshape(inst, -1)
        done = [np.array_equal(got[k], want[k]) for k in range(inst)]
        kept = [np.array_equal(got[k], orig[k]) for k in range(inst)]
        assert all(a or k for a, k in zip(done, kept)), "an instance was partly written"
        return sum(1 for a in done if not a)

    keep = hip.get_tuning("cl_spin")
    hip.set_tuning("cl_spin", 1)
    try:
        d = torch.from_numpy(x.copy()).cuda()
        et = hip.Elapsed()
        r = hip.transpose_0100(d, inst, R, C, b, elem_bytes=es, elapsed=et)  # synchronizing
        assert hip.last_route() == "0100_CL-"
        untouched = outcome(d)
        say("synchronizing call:", r, "untouched", untouched)
        assert r is (untouched > 0)
        d = torch.from_numpy(x.copy()).cuda()
        hip.set_profiling(False)
        try:
            assert hip.transpose_0100(d, inst, R, C, b, elem_bytes=es) is False
            aborted = hip.cluster_aborts()
        finally:
            hip.set_profiling(True)
        untouched = outcome(d)
        say("asynchronous call: clusters that gave up", aborted, "untouched", untouched)
        assert aborted == untouched
    finally:
        hip.set_tuning("cl_spin", keep)
    # with the default bound the same step converts every instance again
    d = torch.from_numpy(x.copy()).cuda()
    assert hip.transpose_0100(d, inst, R, C, b, elem_bytes=es) is False
    assert outcome(d) == 0 and hip.cluster_aborts() == 0
