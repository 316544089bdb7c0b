"""Learned stack-kernel predictor: training data from exhaustive sweeps, training, evaluation.

The sweeps are the JSON lines tune_multiply writes (one per measured configuration, plus a "done"
line per shape). For every (dtype, m, n, k) the targets of the network's heads are soft class
distributions over the configurations within `within` of the best throughput, weighted by
exp((gflops/best - 1) / temperature), and log2 of the best GFLOP/s. Training and inference run on
the GPU through libsmm_acc_mlp_* (hand-written HIP kernels); this module only prepares arrays.

Evaluation follows the exhaustive autotuner as the judge: the network's choice for a held-out
shape is looked up in that shape's sweep (every configuration was measured there), and compared
with the best measured one; the method without the network (the built-in default heuristic, or
the reference's nearest tuned triplet, samples/smm/smm_kernel.c:116-148) is scored the same way.
"""
import json
from collections import defaultdict

import numpy as np

import libxstream_amd as lx

HEAD_KEYS = ("variant", "tm", "tn", "nt", "kt", "bs", "flush")


def triplets(count=500, seed=25071975, lo=4, hi=80):
    """seeded (m, n, k) triplets, uniform integers in [lo, hi], distinct"""
    rng = np.random.default_rng(seed)
    seen, out = set(), []
    while len(out) < count:
        t = tuple(int(x) for x in rng.integers(lo, hi + 1, 3))
        if t not in seen:
            seen.add(t)
            out.append(t)
    return out


def split(trips, ntrain=400):
    return trips[:ntrain], trips[ntrain:]


def load_sweep(paths):
    """{(tid, m, n, k): [row, ...]} of measured configurations (rows are dicts)"""
    rows = defaultdict(list)
    for path in ([paths] if isinstance(paths, str) else paths):
        with open(path) as f:
            for line in f:
                line = line.strip()
                if not line:
                    continue
                r = json.loads(line)
                if r.get("done") or r.get("gflops", 0) <= 0:
                    continue
                rows[(r["dtype"], r["m"], r["n"], r["k"])].append(r)
    return dict(rows)


def head_value(row, key):
    """class value of a sweep row for a head: the vector kernel (variant 0) has no compiled kt"""
    if key == "kt" and row["variant"] == 0:
        return 0
    return row[key]


def targets(rows, seg, values, within=0.05, temperature=0.02):
    """soft per-head targets and log2 of the best GFLOP/s for one shape"""
    nout = int(seg[-1])
    t = np.zeros(nout, dtype=np.float32)
    best = max(r["gflops"] for r in rows)
    good = [r for r in rows if r["gflops"] >= (1.0 - within) * best]
    wts = np.array([np.exp((r["gflops"] / best - 1.0) / temperature) for r in good])
    wts /= wts.sum()
    for h, key in enumerate(HEAD_KEYS):
        c0, c1 = int(seg[h]), int(seg[h + 1])
        cls = {int(v): c0 + i for i, v in enumerate(values[c0:c1])}
        for r, w in zip(good, wts):
            c = cls.get(int(head_value(r, key)))
            if c is not None:
                t[c] += w
        s = t[c0:c1].sum()
        if s > 0:
            t[c0:c1] /= s
    t[int(seg[len(HEAD_KEYS)])] = np.log2(best)
    return t


def dataset(sweep, shapes, within=0.05, temperature=0.02):
    """features [N][8] and targets [N][nout] for the (tid, m, n, k) keys of `shapes` present in the sweep"""
    seg, values = lx.mlp_heads()
    keys = [s for s in shapes if s in sweep]
    x = np.stack([lx.mlp_features(*key) for key in keys]) if keys else np.zeros((0, 8), np.float32)
    t = np.stack([targets(sweep[key], seg, values, within, temperature) for key in keys]) if keys else None
    return keys, x, t


def config_key(cfg):
    """(variant, tm, tn, nt, kt class, bs, flush) of a sweep row (dict) or an SmmConfig"""
    get = cfg.get if isinstance(cfg, dict) else (lambda k: getattr(cfg, k))
    return tuple(int(head_value({"variant": get("variant"), key: get(key)}, key)) for key in HEAD_KEYS)


def lookup(rows, cfg):
    """measured GFLOP/s of configuration cfg in a shape's sweep rows, or None when it was not measured"""
    want = config_key(cfg)
    for r in rows:
        if config_key(r) == want:
            return r["gflops"]
    return None


def evaluate(model, sweep, keys):
    """per held-out shape: (key, chosen GFLOP/s, best GFLOP/s, default GFLOP/s, confidence)"""
    out = []
    for key in keys:
        rows = sweep[key]
        best = max(r["gflops"] for r in rows)
        cfg, conf = model.predict_config(*key)
        got = lookup(rows, cfg)
        out.append((key, got, best, conf))
    return out


def nearest_tuned(key, tuned):
    """the reference's method: parameters of the tuned triplet nearest to (m, n, k) (same dtype)"""
    tid, m, n, k = key
    cands = [t for t in tuned if t[0] == tid]
    if not cands:
        return None
    d = [(t[1] - m) ** 2 + (t[2] - n) ** 2 + (t[3] - k) ** 2 for t in cands]
    return cands[int(np.argmin(d))]


def write_triplets(path, count=500, seed=25071975):
    """the sweep's triplet file (one 'm n k' per line, tune_multiply -f): 400 training, then 100 held out"""
    with open(path, "w") as f:
        for m, n, k in triplets(count, seed):
            f.write("%d %d %d\n" % (m, n, k))


if __name__ == "__main__":
    import sys

    if len(sys.argv) == 3 and sys.argv[1] == "triplets":
        write_triplets(sys.argv[2])
    else:
        print("usage: python -m libxstream_amd.learned triplets FILE", file=sys.stderr)
        sys.exit(2)
