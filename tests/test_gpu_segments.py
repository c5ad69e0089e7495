"""The segmented cycle walk for 100 / 0100 steps (route tag "0100_SEG-"): one pass with no
visited bits, for grids whose cycles are few and long. Checked bit-exactly against the oracle
over super-element sizes that give every unit width (1-16 bytes), single and split pieces,
several instance counts, a misaligned base, and segment lengths from 5 to 100; plus the
reference's 3-stage route on config 1's shape, whose 100 / 0100 stages take this form."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture
def seg_only(hip):
    """The segmented walk wherever it can run (no through-scratch or cycle-leader forms)."""
    keep = {k: hip.get_tuning(k) for k in ("st_scratch_mb", "pl_max_order", "seg_len",
                                             "seg_force", "cl_max_wgs")}
    hip.set_tuning("st_scratch_mb", 0)
    hip.set_tuning("cl_max_wgs", 0)
    hip.set_tuning("pl_max_order", 1)
    hip.set_tuning("seg_force", 1)  # also where the model prefers another form (narrow units)
    yield
    for k, v in keep.items():
        hip.set_tuning(k, v)


CASES = [  # (instances, R, C, b, elem_bytes, base offset in elements)
    (1, 1800, 100, 72, 4, 0),    # 288-byte super-elements, 14 cycles up to 12857 long
    (100, 100, 72, 18, 4, 0),    # 72-byte (8-byte units), 100 instances
    (3, 97, 131, 5, 16, 0),      # 80-byte
    (2, 64, 33, 200, 8, 0),      # 1600-byte: 100 units of 16 bytes, two per lane
    (1, 50, 40, 1024, 4, 0),     # 4 KiB: split into pieces
    (5, 41, 29, 33, 2, 0),       # 66-byte: 2-byte units
    (2, 37, 45, 65, 1, 0),       # 65-byte: single bytes
    (4, 60, 25, 24, 4, 1),       # 96-byte super-elements at a 4-byte-aligned base
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c)))
def test_segmented_0100_matches_oracle(hip, orc, seg_only, case):
    import torch
    inst, R, C, b, es, off = case
    n = inst * R * C * b
    x = orc.tagged(n, es, seed=R * C + b)
    want = orc.op_0100(x, inst, R, C, b, es=es)
    for L in (32, 5, 7, 100):
        hip.set_tuning("seg_len", L)
        buf = torch.zeros((n + off) * es, dtype=torch.uint8, device="cuda")
        buf[off * es:] = torch.from_numpy(x).cuda()
        d = buf[off * es:]
        assert hip.transpose_0100(d, inst, R, C, b, elem_bytes=es) is False
        assert hip.last_route() == "0100_SEG-", (case, L, hip.last_route())
        got = buf.cpu().numpy()
        assert np.array_equal(got[off * es:], want), (case, L)
        assert not got[:off * es].any()
    say("seg", case, "ok")


def test_config1_three_stage_route_segments(hip, orc):
    """1800 x 7200 int32 with a = 18, b = 72 and back by the 3-stage route: its 100 stage
    (288-byte super-elements, one 51.84 MB grid) takes the segmented walk, its 0100 stage
    (100 instances of 72-byte super-elements) the cluster-staged form."""
    import torch
    M, N = 1800, 7200
    x = (orc.splitmix64(77, M * N) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    d = torch.from_numpy(x.view(np.int32).copy()).cuda()
    assert hip.transpose(d, M // 18, 18, N // 72, 72, 1, 3) is False
    assert np.array_equal(d.cpu().numpy().view(np.uint32).reshape(N, M), x.reshape(M, N).T)
    assert hip.transpose(d, N // 72, 72, M // 18, 18, 1, 3) is False
    assert np.array_equal(d.cpu().numpy().view(np.uint32), x)
