"""Host check of the segmented cycle walk used for 100 / 0100 steps (csrc/kernels_seg.hip).

The walk is replayed in Python exactly as the two launches order it: the segment table is built
by walking every cycle of the R x C grid's map s -> s*R mod (R*C - 1) and cutting it every
`L` slots; launch 1 saves each segment's start; launch 2 walks each segment a window of W slots
at a time (all loads of a window, then its stores). Every grid up to 24 x 24, several segment
lengths and window sizes, must give the out-of-place transpose (numpy), and every slot must be
stored exactly once.
"""
import numpy as np
import pytest


def seg_table(R, C, L):
    RC = R * C
    m = RC - 1
    seen = np.zeros(RC, dtype=bool)
    table = []
    for s in range(1, m):
        if seen[s]:
            continue
        first, t, n = len(table), s, 0
        while True:
            seen[t] = True
            if n % L == 0:
                table.append([t, 0])
            n += 1
            t = (t % C) * R + t // C
            if t == s:
                break
        if n < 2:
            del table[first:]
            continue
        for j, k in enumerate(range(first, len(table))):
            table[k][1] = min(L, n - j * L)
    return table


def seg_walk(x, R, C, L, W):
    """x: [R*C] array of super-elements (any dtype); returns the walked copy and store counts."""
    mem = x.copy()
    stores = np.zeros(len(x), dtype=int)
    table = seg_table(R, C, L)
    save = [mem[s].copy() for s, _ in table]  # launch 1
    for g, (t, n) in enumerate(table):  # launch 2, one walker per segment (any order)
        carry = save[g]
        i = 0
        while i < n:
            w = min(W, n - i)
            q, d = [], []
            for j in range(w):
                t = (t % C) * R + t // C
                q.append(t)
                d.append(mem[t].copy() if i + j + 1 < n else None)
            for j in range(w):
                mem[q[j]] = carry if j == 0 else d[j - 1]
                stores[q[j]] += 1
            if w == W:
                carry = d[W - 1]
            i += w
    return mem, stores


@pytest.mark.parametrize("L,W", [(2, 1), (3, 2), (5, 16), (8, 4), (32, 16), (7, 7)])
def test_segmented_walk_is_the_transpose(L, W):
    for R in range(2, 25):
        for C in range(2, 25):
            x = np.arange(R * C, dtype=np.int64)
            got, stores = seg_walk(x, R, C, L, W)
            want = x.reshape(R, C).T.reshape(-1)
            assert np.array_equal(got, want), (R, C, L, W)
            moved = want != x
            assert np.all(stores[moved] == 1) and np.all(stores[~moved] == 0), (R, C, L, W)


def test_segment_table_config1_grids():
    """The grids of the reference's 3-stage route on 1800 x 7200 (a = 18, b = 72) and back:
    segments of at most L slots that together cover every slot the map moves."""
    for (R, C) in ((1800, 100), (100, 72), (100, 18), (7200, 100)):
        t = seg_table(R, C, 32)
        assert all(1 <= n <= 32 for _, n in t)
        assert sum(n for _, n in t) == sum(1 for s in range(1, R * C - 1)
                                           if (s % C) * R + s // C != s)
        assert len(t) >= (R * C - 2 + 31) // 32
