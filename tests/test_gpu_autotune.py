"""The measured autotuner per shape class (hip_marshal_autotune_desc) and its use by batches.

Each tuned conversion must leave its buffer bit-identical; the class's choice must then be the
form later plain calls of the class take (on another shape of the same class too), and every
result stays equal to the oracle. A batch created with the knob "autotune" tunes each class of
its descriptors once and stays bit-exact on every replay.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _forget_choices(hip):
    """Measured choices apply to whole classes: forget them afterwards, so later tests see the
    cost model's routes."""
    yield
    import torch
    torch.cuda.synchronize()
    hip.finalize()


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


@pytest.mark.parametrize("case", [
    (4, 64 * 4000, 16, 64),     # AoS -> ASTA, 64 x 16 f32 tiles (LDS-staged or cycle following)
    (8, 32 * 3000, 37, 32),     # odd width
    (2, 128 * 900, 300, 128),   # tiles beyond the budget of one workgroup's share
])
def test_autotune_aos_class(hip, orc, case):
    es, h, w, t = case
    x = orc.tagged(h * w, es, seed=h + w)
    d = dev(x)
    form = hip.autotune_desc(d, hip.OP_AOS_ASTA, es, h, w, t)
    say("aos class", case, form, hip.FORM_010.get(form))
    assert form in hip.FORM_010
    assert np.array_equal(host(d), x)  # the tuner restores the buffer
    # a plain call of the class takes the chosen form (for a shape in the same class too)
    for hh in (h, h - 2 * t):
        y = orc.tagged(hh * w, es, seed=hh)
        dy = dev(y)
        assert hip.aos_asta(dy, hh, w, t, elem_bytes=es) is False
        assert hip.last_route() == hip.FORM_010[form], (hh, hip.last_route())
        assert np.array_equal(host(dy), orc.op_010(y, hh // t, t, w, es=es))


@pytest.mark.parametrize("case", [
    (4, 64 * 3000, 16, 64),     # SoA -> ASTA, 256-byte super-elements (short cycles)
    (4, 18 * 4000, 101, 18),    # 72-byte super-elements, long cycles
    (1, 48 * 2000, 77, 48),     # 48-byte super-elements
])
def test_autotune_soa_class(hip, orc, case):
    es, h, w, t = case
    x = orc.tagged(h * w, es, seed=h * 3 + w)
    d = dev(x)
    form = hip.autotune_desc(d, hip.OP_SOA_ASTA, es, h, w, t)
    say("soa class", case, form, hip.FORM_0100.get(form))
    assert form in hip.FORM_0100
    assert np.array_equal(host(d), x)
    assert hip.soa_asta(d, h, w, t, elem_bytes=es) is False
    assert hip.last_route() == hip.FORM_0100[form], hip.last_route()
    assert np.array_equal(host(d), orc.op_100(x, w, h // t, t, es=es))


def test_autotune_transpose_desc(hip, orc):
    M, N, es = 1009, 2003, 4
    x = orc.tagged(M * N, es, seed=7)
    d = dev(x)
    stages = hip.autotune_desc(d, hip.OP_TRANSPOSE, es, M, N)
    assert np.array_equal(host(d), x)
    assert hip.plan_transpose(M, N, es)[2] == stages


def test_batch_with_autotune_is_exact(hip, orc):
    import torch
    from libmarshal_amd import workloads as wl
    items = wl.batch_mix(120, lo=16, hi=1500, seed=99)
    bufs, wants = [], []
    for i, (op, es, r, c, t) in enumerate(items):
        x = orc.tagged(r * c, es, seed=500 + i)
        bufs.append(dev(x))
        if op == wl.OP_TRANSPOSE:
            wants.append(orc.op_transpose(x, r, c, es=es))
        elif op == wl.OP_AOS_ASTA:
            wants.append(orc.op_010(x, r // t, t, c, es=es))
        else:
            wants.append(orc.op_100(x, c, r // t, t, es=es))
    keep = hip.get_tuning("autotune")
    hip.set_tuning("autotune", 1)
    try:
        batch = hip.Batch([(b, op, es, r, c, t) for b, (op, es, r, c, t) in zip(bufs, items)])
    finally:
        hip.set_tuning("autotune", keep)
    torch.cuda.synchronize()
    # tuning left every buffer as it was
    for b, (op, es, r, c, t), i in zip(bufs, items, range(len(items))):
        assert np.array_equal(host(b), orc.tagged(r * c, es, seed=500 + i)), i
    assert batch.run() is False
    for b, want, i in zip(bufs, wants, range(len(items))):
        assert np.array_equal(host(b), want), (i, items[i], batch.route(i))
    batch.close()
