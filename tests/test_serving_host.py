"""Host logic of the serving loop: the arena allocator (admission pool) and the step
planner (continuous batching), without a GPU."""

import numpy as np

from voxneus_amd.serving import Arena, Result, _Live, plan_step


def test_arena_matches_a_bitmap_model():
    rng = np.random.default_rng(0)
    rows = 1000
    a = Arena(rows)
    used = np.zeros(rows, bool)
    held = []
    for _ in range(2000):
        if held and rng.random() < 0.45:
            off, n = held.pop(int(rng.integers(len(held))))
            assert used[off:off + n].all()
            used[off:off + n] = False
            a.release(off, n)
        else:
            n = int(rng.integers(1, 120))
            off = a.alloc(n)
            # first fit: the lowest offset with n free rows, if any
            free = ~used
            fits = [o for o in range(rows - n + 1) if free[o:o + n].all()]
            if off is None:
                # no free range in the free list: the bitmap must agree unless free
                # ranges are adjacent (never: the free list merges neighbours)
                assert not fits
                continue
            assert fits and off == fits[0]
            used[off:off + n] = True
            held.append((off, n))
        assert a.free_rows == int((~used).sum())
        # the free list is sorted, disjoint and merged
        for (o1, l1), (o2, l2) in zip(a.free, a.free[1:]):
            assert o1 + l1 < o2


def _live(i, rows):
    return _Live(Result(i, "render", 0, rows, 0.0), off=1000 * i)


def test_plan_step_fills_oldest_first_and_advances_cursors():
    live = [_live(0, 50), _live(1, 300), _live(2, 10)]
    seg = plan_step(live, 200, 8)
    assert [(lv.res.index, first, n) for lv, first, n in seg] == [(0, 0, 50), (1, 0, 150)]
    seg = plan_step(live, 200, 8)
    assert [(lv.res.index, first, n) for lv, first, n in seg] == [(1, 150, 150), (2, 0, 10)]
    assert plan_step(live, 200, 8) == []
    live = [_live(i, 5) for i in range(10)]
    seg = plan_step(live, 1000, 4)
    assert len(seg) == 4 and sum(n for *_, n in seg) == 20


def test_plan_step_covers_every_row_once():
    rng = np.random.default_rng(1)
    live = [_live(i, int(rng.integers(1, 5000))) for i in range(30)]
    seen = {i: np.zeros(lv.res.rows, int) for i, lv in enumerate(live)}
    while True:
        seg = plan_step(live, 1024, 7)
        if not seg:
            break
        assert sum(n for *_, n in seg) <= 1024 and len(seg) <= 7
        for lv, first, n in seg:
            seen[lv.res.index][first:first + n] += 1
    assert all((v == 1).all() for v in seen.values())
