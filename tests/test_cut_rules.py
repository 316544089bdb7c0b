"""CPU: the long-run cutting protocol of the stack kernels, emulated step by step.

Restates the control flow of smm_mfma / smm_batch (owner continues past its end, cut checked at each
segment boundary it reaches) and of smm_stack / smm_lanes (piece end found by a scan), with the shared
predicate lxs_cut (libxstream_amd/csrc/lxs_mfma.h): boundary entry i starts a new piece when entry
i - T has the same C block. Over random stacks (sorted and unsorted, bs 1..70, T 0..80) every entry
must be summed exactly once, and in a sorted stack every C block split over several pieces must be
flushed only atomically (the flag each kernel computes from the piece's neighbours).
"""
import numpy as np


def cut(c, i, ci, T):
    return 0 < T <= i and c[i - T] == ci


def owner_walk(c, w, bs, T):
    """smm_mfma / smm_batch: pieces of work-group w as (c, entries, atomic)."""
    S = len(c)
    begin, end = w * bs, min(w * bs + bs, S)
    h = next((i for i in range(begin, end) if i == 0 or c[i] != c[i - 1] or (i == begin and cut(c, i, c[i], T))), end)
    if h >= end:
        return []
    nbd = [end]

    def cut_at(i, ci):
        if i != nbd[0]:
            return False
        nbd[0] += bs
        return i < S and cut(c, i, ci, T)

    e, c0, pieces = h, c[h], []
    have1, c1 = e + 1 < S, (c[e + 1] if e + 1 < S else -1)
    have2, c2 = e + 2 < S, (c[e + 2] if e + 2 < S else -1)
    cut1, cut2 = cut_at(e + 1, c1), cut_at(e + 2, c2)
    split = e > 0 and c[e - 1] == c0
    cur = [e]
    while True:
        while True:
            run_ends = (not have1) or c1 != c0 or cut1
            own_next = have1 and ((c1 == c0 and not cut1) or e + 1 < end)
            if run_ends or not own_next:
                break
            e, c1, have1, cut1 = e + 1, c2, have2, cut2
            have2, c2 = e + 2 < S, (c[e + 2] if e + 2 < S else -1)
            cut2 = cut_at(e + 2, c2)
            cur.append(e)
        pieces.append((c0, cur, split or (have1 and c1 == c0)))
        if not own_next:
            return pieces
        e, c0, c1, have1, cut1, split = e + 1, c1, c2, have2, cut2, False
        have2, c2 = e + 2 < S, (c[e + 2] if e + 2 < S else -1)
        cut2 = cut_at(e + 2, c2)
        cur = [e]


def scan_walk(c, wbeg, wend, T, boundary):
    """smm_stack (W = 1) / smm_lanes (W wavefront segments): pieces of the range [wbeg, wend)."""
    S = len(c)
    h = next((x for x in range(wbeg, wend) if x == 0 or c[x] != c[x - 1] or (x == wbeg and cut(c, x, c[x], T))), wend)
    pieces = []
    while h < wend:
        c0, r = c[h], h + 1
        while r < S and not (c[r] != c0 or (wend <= r and boundary(r) and cut(c, r, c0, T))):
            r += 1
        pieces.append((c0, list(range(h, r)), (h > 0 and c[h - 1] == c0) or (r < S and c[r] == c0)))
        h = r
    return pieces


def check(c, pieces, sorted_stack):
    seen = np.zeros(len(c), dtype=int)
    flags = {}
    for c0, entries, atomic in pieces:
        assert all(c[x] == c0 for x in entries)
        seen[entries] += 1
        flags.setdefault(c0, []).append(atomic)
    assert (seen == 1).all()
    if sorted_stack:
        assert all(all(f) for f in flags.values() if len(f) > 1)


def stacks(rng, count):
    for t in range(count):
        S = int(rng.integers(1, 400))
        lens = rng.integers(1, (5, 60, 400)[t % 3], size=S)
        c = np.repeat(np.arange(len(lens)), lens)[:S]
        if t % 7 == 0:
            yield rng.integers(0, 5, size=S), False
        else:
            yield c, True


def test_owner_walk_exactly_once():
    rng = np.random.default_rng(1)
    for c, srt in stacks(rng, 1500):
        bs, T = int(rng.integers(1, 40)), int(rng.integers(0, 80))
        pieces = [p for w in range((len(c) + bs - 1) // bs) for p in owner_walk(c, w, bs, T)]
        check(c, pieces, srt)


def test_scan_walk_exactly_once():
    rng = np.random.default_rng(2)
    for c, srt in stacks(rng, 1500):
        S, bs, T, W = len(c), int(rng.integers(1, 70)), int(rng.integers(0, 80)), int(rng.integers(1, 5))
        bsw = (bs + W - 1) // W
        pieces = []
        for q in range((S + bs - 1) // bs):
            for w in range(W):
                wbeg = q * bs + w * bsw
                wend = min(wbeg + bsw, (q + 1) * bs, S)
                if wbeg < wend:
                    pieces += scan_walk(c, wbeg, wend, T, lambda x: (x % bs) % bsw == 0)
        check(c, pieces, srt)


def test_long_tail_is_spread():
    """acc_bench's INIT_STACK gives the last C block the remainder (about 950 entries at S = 30000):
    with T = 2 bs no work-group sums more than a few segments of it."""
    c = np.concatenate([np.repeat(np.arange(1874), 15), np.full(30000 - 1874 * 15, 1874)])
    for T, bound in ((0, 1890), (32, 34 + 16)):
        most = max(sum(len(p[1]) for p in owner_walk(c, w, 16, T)) for w in range(30000 // 16))
        assert most <= bound


def ep_walk(c, w, bs, T, EP):
    """smm_mfma_ep: the pieces of work-group w, each with the entries every wavefront sums. Follows the
    kernel's per-wavefront state: the entry whose first k-stage is in flight (e), prefetched at the end of a
    run as r + q when the work-group owns the next run; an entry is summed from the fragments in flight only
    when they belong to it (checked), and every stack index the kernel reads is in bounds (checked)."""
    S = len(c)

    def rd(i):
        assert 0 <= i < S, "stack read out of bounds: %d" % i
        return c[i]

    begin, end = w * bs, min(w * bs + bs, S)
    h = next((i for i in range(begin, end) if i == 0 or rd(i) != rd(i - 1) or (i == begin and cut(c, i, c[i], T))), end)
    if h >= end:
        return []
    inflight = [None] * EP  # entry whose fragments each wavefront holds
    pieces = []
    while True:
        c0 = rd(h)
        r = h + 1
        while r < S and not (rd(r) != c0 or (end <= r and (r - end) % bs == 0 and cut(c, r, c0, T))):
            r += 1
        atomic = (h > 0 and rd(h - 1) == c0) or (r < S and rd(r) == c0)
        own_next = r < end
        per_wave = []
        for q in range(EP):
            mine = []
            cur = h + q
            if cur < r:
                if inflight[q] != cur:
                    inflight[q] = cur  # fresh load of the first k-stage
                while True:
                    assert inflight[q] == cur, "fragments of entry %s used for entry %d" % (inflight[q], cur)
                    rd(cur)
                    mine.append(cur)
                    nx = cur + EP
                    in_run = nx < r
                    cand = r + q
                    pre = (not in_run) and own_next and cand < S
                    nxt = nx if in_run else (cand if pre else -1)
                    if nxt >= 0:
                        rd(nxt)
                    inflight[q] = nxt if nxt >= 0 else None
                    if not in_run:
                        break
                    cur = nx
            per_wave.append(mine)
        entries = sorted(x for mine in per_wave for x in mine)
        assert entries == list(range(h, r))
        pieces.append((c0, entries, atomic))
        if not own_next:
            return pieces
        h = r


def test_ep_walk_exactly_once():
    rng = np.random.default_rng(3)
    for c, srt in stacks(rng, 1500):
        bs, T, EP = int(rng.integers(1, 70)), int(rng.integers(0, 80)), int(rng.choice([2, 4, 8]))
        pieces = [p for w in range((len(c) + bs - 1) // bs) for p in ep_walk(c, w, bs, T, EP)]
        check(c, pieces, srt)
