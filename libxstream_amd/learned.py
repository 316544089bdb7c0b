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
import gzip
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
        with (gzip.open(path, "rt") if str(path).endswith(".gz") else open(path)) as f:
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
    """class value of a sweep row for a head: the vector kernel (variant 0) has no compiled kt, and
    flush 0 (automatic) launches the atomic flush (2)"""
    if key == "kt" and row["variant"] == 0:
        return 0
    if key == "flush" and row[key] == 0:
        return 2
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


def best_row(rows):
    return max(rows, key=lambda r: r["gflops"])


def nearest_tuned(key, tuned_keys):
    """the reference's method (samples/smm/smm_kernel.c:116-148): the tuned triplet of the same type
    nearest to (m, n, k)"""
    tid, m, n, k = key
    cands = [t for t in tuned_keys if t[0] == tid]
    if not cands:
        return None
    d = [(t[1] - m) ** 2 + (t[2] - n) ** 2 + (t[3] - k) ** 2 for t in cands]
    return cands[int(np.argmin(d))]


def _config(row_or_cfg):
    cfg = lx.SmmConfig()
    get = row_or_cfg.get if isinstance(row_or_cfg, dict) else (lambda k: getattr(row_or_cfg, k))
    for key in ("variant", "tm", "tn", "nt", "bs", "kt", "flush"):
        setattr(cfg, key, int(get(key)))
    return cfg


def evaluate(model, sweep, keys, train_keys):
    """per held-out shape: measured GFLOP/s of the network's choice, of the default heuristic, of the
    nearest tuned triplet's configuration, and the best measured (the exhaustive autotuner's)"""
    out = []
    for key in keys:
        rows = sweep[key]
        best = best_row(rows)["gflops"]
        cfg, conf = model.predict_config(*key)
        learned = lookup(rows, cfg)
        default = lookup(rows, lx.default_config(*key))
        near = nearest_tuned(key, train_keys)
        nearest = None
        if near is not None:
            ncfg = _config(best_row(sweep[near]))
            if lx.config_valid(key[0], key[1], key[2], key[3], ncfg):
                nearest = lookup(rows, ncfg)
            if nearest is None:  # not valid for this shape: the reference falls back to its default
                nearest = default
        out.append({"key": list(key), "best": best, "learned": learned, "confidence": conf, "default": default,
                    "nearest": nearest})
    return out


def summary(results, field):
    ratios = np.array([(r[field] or 0.0) / r["best"] for r in results])
    return {"mean": round(float(ratios.mean()), 4), "median": round(float(np.median(ratios)), 4),
            "min": round(float(ratios.min()), 4), "within_10pct": round(float((ratios >= 0.9).mean()), 4),
            "geomean": round(float(np.exp(np.log(np.maximum(ratios, 1e-6)).mean())), 4)}


def train(sweep, train_keys, hidden=64, nhidden=3, seed=1, epochs=(3000, 3000), lrs=(3e-3, 1e-3), weight_decay=1e-4,
          within=0.05, temperature=0.02):
    """train a network on the training shapes; returns (model, final loss)"""
    keys, x, t = dataset(sweep, train_keys, within, temperature)
    hw = np.ones(len(HEAD_KEYS) + 1, dtype=np.float32)
    hw[-1] = 0.5  # log2 GFLOP/s regression
    model = lx.Mlp(hidden, nhidden, seed=seed)
    loss = None
    for ep, lr in zip(epochs, lrs):
        loss = model.train(x, t, ep, lr=lr, weight_decay=weight_decay, head_weights=hw)
    return model, loss


def write_triplets(path, count=500, seed=25071975):
    """the sweep's triplet file (one 'm n k' per line, tune_multiply -f): 400 training, then 100 held out"""
    with open(path, "w") as f:
        for m, n, k in triplets(count, seed):
            f.write("%d %d %d\n" % (m, n, k))


def main(argv):
    import argparse

    p = argparse.ArgumentParser(prog="python -m libxstream_amd.learned")
    sub = p.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("triplets", help="write the sweep's triplet file")
    t.add_argument("path")
    r = sub.add_parser("train", help="train on the sweeps' training shapes, evaluate on the held-out ones")
    r.add_argument("sweeps", nargs="+")
    r.add_argument("--model", required=True, help="output model file")
    r.add_argument("--report", help="output JSON report")
    r.add_argument("--seeds", type=int, default=4, help="networks trained (lowest training loss kept)")
    args = p.parse_args(argv)
    if args.cmd == "triplets":
        write_triplets(args.path)
        return 0
    lx.init()
    sweep = load_sweep(args.sweeps)
    trips = triplets()
    train_t, held_t = split(trips)
    tids = sorted({key[0] for key in sweep})
    train_keys = [(tid,) + t for tid in tids for t in train_t if (tid,) + t in sweep]
    held_keys = [(tid,) + t for tid in tids for t in held_t if (tid,) + t in sweep]
    best = None
    for seed in range(1, args.seeds + 1):
        model, loss = train(sweep, train_keys, seed=seed)
        print("seed %d: training loss %.4f" % (seed, loss), flush=True)
        if best is None or loss < best[1]:
            if best is not None:
                best[0].close()
            best = (model, loss, seed)
        else:
            model.close()
    model, loss, seed = best
    model.save(args.model)
    results = evaluate(model, sweep, held_keys, train_keys)
    report = {"train_shapes": len(train_keys), "heldout_shapes": len(held_keys), "seed": seed, "loss": loss,
              "learned": summary(results, "learned"), "default": summary(results, "default"),
              "nearest_tuned": summary(results, "nearest")}
    for tid in tids:
        sub_r = [x for x in results if x["key"][0] == tid]
        report["type%d" % tid] = {f: summary(sub_r, f) for f in ("learned", "default", "nearest")}
    conf = np.array([x["confidence"] for x in results])
    ok = np.array([(x["learned"] or 0) >= 0.9 * x["best"] for x in results])
    report["confidence_quantiles"] = [round(float(q), 4) for q in np.quantile(conf, [0.1, 0.25, 0.5, 0.75, 0.9])]
    report["within_10pct_by_confidence"] = {str(q): round(float(ok[conf >= q].mean()), 4) if (conf >= q).any() else None
                                           for q in (0.0, 0.05, 0.1, 0.2, 0.3, 0.5)}
    report["results"] = results
    print(json.dumps({k: v for k, v in report.items() if k != "results"}, indent=1))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f)
    model.close()
    lx.finalize()
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(main(sys.argv[1:]))
