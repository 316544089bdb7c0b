"""Learned stack-kernel predictor: training data from exhaustive sweeps, training, evaluation.

The sweeps are the JSON lines tune_multiply writes (one per measured configuration, plus a "done"
line per shape). Every measured (shape, configuration) is one training row: the network's inputs are
libsmm_acc_mlp_features (8 shape features and the configuration's one-hot parameters), its targets
the configuration's throughput relative to the shape's best and log2 of its GFLOP/s. Training and
inference run on the GPU through libsmm_acc_mlp_* (hand-written HIP kernels); this module only
prepares arrays.

The exhaustive autotuner is the measure of every choice: the network's choice for a held-out shape
is looked up in that shape's sweep (every configuration was measured there) and compared with the
best measured one. The learned-guided autotuner (tune_multiply -l model -k K) measures only the
network's top K; its result is the best of those K measured configurations, looked up the same way. The methods without the network are scored the same way: the built-in default
heuristic, and the reference's nearest tuned triplet (samples/smm/smm_kernel.c:116-148).
"""
import gzip
import json
import os
from collections import defaultdict

import numpy as np

import libxstream_amd as lx

PARAM_KEYS = ("variant", "tm", "tn", "nt", "kt", "bs", "flush")
OUT_WEIGHTS = (1.0, 0.01)  # loss weights of the outputs: relative throughput, log2 GFLOP/s
TOPK = (2, 4, 8)  # learned-guided tuning: configurations measured per shape


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
    """the first ntrain triplets train, the rest are held out"""
    return trips[:ntrain], trips[ntrain:]


def load_sweep(paths):
    """{(tid, m, n, k): [row, ...]} of measured configurations (rows are dicts); .gz files are read too"""
    rows_by_key = defaultdict(list)
    for path in ([paths] if isinstance(paths, str) else paths):
        with (gzip.open(path, "rt") if str(path).endswith(".gz") else open(path)) as f:
            for line in f:
                line = line.strip()
                if not line:
                    continue
                r = json.loads(line)
                if r.get("done") or r.get("gflops", 0) <= 0:
                    continue
                rows_by_key[(r["dtype"], r["m"], r["n"], r["k"])].append(r)
    return dict(rows_by_key)


def param_value(cfg, key):
    """a configuration parameter as the network encodes it: the vector kernel (variant 0) has no
    compiled kt, and flush 0 (automatic) launches the atomic flush (2)"""
    get = cfg.get if isinstance(cfg, dict) else (lambda k: getattr(cfg, k))
    if key == "kt" and get("variant") == 0:
        return 0
    if key == "flush" and get("flush") == 0:
        return 2
    return int(get(key))


def config_key(cfg):
    """(variant, tm, tn, nt, kt, bs, flush) of a sweep row (dict) or an SmmConfig, as encoded"""
    return tuple(param_value(cfg, key) for key in PARAM_KEYS)


def to_config(row):
    """SmmConfig of a sweep row"""
    cfg = lx.SmmConfig()
    for key in ("variant", "tm", "tn", "nt", "bs", "kt", "flush"):
        setattr(cfg, key, int(row[key]))
    return cfg


def grouped():
    """LIBXSTREAM_SMM_GROUPED: stacks are declared grouped by C, so configurations with flush 1 (no atomics)
    launch as measured; otherwise they launch with the atomic flush and are not candidates"""
    return os.environ.get("LIBXSTREAM_SMM_GROUPED", "0") not in ("", "0")


def runnable(rows):
    """the measured configurations that libsmm_acc_process can launch as measured"""
    return rows if grouped() else [r for r in rows if param_value(r, "flush") != 1]


def best_row(rows):
    return max(rows, key=lambda r: r["gflops"])


def lookup(rows, cfg):
    """measured GFLOP/s of configuration cfg in a shape's sweep rows, or None when it was not measured"""
    want = config_key(cfg)
    for r in rows:
        if config_key(r) == want:
            return r["gflops"]
    return None


def dataset(sweep, keys):
    """inputs [R][nin], targets [R][2] and the shape index of every measured configuration of `keys`"""
    xs, ts, gs = [], [], []
    for i, key in enumerate(keys):
        best = best_row(sweep[key])["gflops"]
        for r in sweep[key]:
            xs.append(lx.mlp_features(*key, to_config(r)))
            ts.append((r["gflops"] / best, np.log2(r["gflops"])))
            gs.append(i)
    return np.array(xs, dtype=np.float32), np.array(ts, dtype=np.float32), np.array(gs)


def train(sweep, train_keys, nnets=8, hidden=64, nhidden=3, seed=1, epochs=4000, lr=3e-3, weight_decay=1e-5):
    """an ensemble trained on every measured configuration of the training shapes: (model, mean loss)"""
    x, t, _ = dataset(sweep, train_keys)
    model = lx.Mlp(nnets, hidden, nhidden, seed=seed)
    loss = model.train(x, t, epochs, lr=lr, weight_decay=weight_decay, out_weights=OUT_WEIGHTS)
    return model, loss


def nearest_tuned(key, tuned_keys):
    """the reference's method (samples/smm/smm_kernel.c:116-148): the tuned triplet of the same type
    nearest to (m, n, k)"""
    tid, m, n, k = key
    cands = [t for t in tuned_keys if t[0] == tid]
    if not cands:
        return None
    d = [(t[1] - m) ** 2 + (t[2] - n) ** 2 + (t[3] - k) ** 2 for t in cands]
    return cands[int(np.argmin(d))]


def evaluate(model, sweep, keys, tuned_keys):
    """per held-out shape: measured GFLOP/s of the network's choice, of the default heuristic, of the
    nearest tuned triplet's configuration, and the best measured one (the exhaustive autotuner's)"""
    out = []
    for key in keys:
        rows = runnable(sweep[key])
        best = best_row(rows)["gflops"]
        cfg, conf = model.predict_config(*key)
        default = lookup(rows, lx.default_config(*key))
        near = nearest_tuned(key, tuned_keys)
        nearest = None
        if near is not None:
            ncfg = to_config(best_row(runnable(sweep[near])))
            if lx.config_valid(key[0], key[1], key[2], key[3], ncfg):
                nearest = lookup(rows, ncfg)
        if nearest is None:  # not valid for this shape: the reference falls back to its default
            nearest = default
        entry = {"key": list(key), "best": best, "learned": lookup(rows, cfg), "confidence": conf,
                 "predicted_gflops": cfg.gflops, "default": default, "nearest": nearest}
        # learned-guided tuning: the best measured among the network's top K (K measurements per shape)
        ranked = model.predict_topk(*key, max(TOPK))
        for kk in TOPK:
            got = [lookup(rows, c) for c, _ in ranked[:kk]]
            entry["learned_top%d" % kk] = max([g for g in got if g is not None], default=None)
        out.append(entry)
    return out


def summary(results, field):
    """fraction of the exhaustive best reached by `field` over the results"""
    ratios = np.array([(r[field] or 0.0) / r["best"] for r in results])
    return {"mean": round(float(ratios.mean()), 4), "median": round(float(np.median(ratios)), 4),
            "min": round(float(ratios.min()), 4), "within_10pct": round(float((ratios >= 0.9).mean()), 4),
            "geomean": round(float(np.exp(np.log(np.maximum(ratios, 1e-6)).mean())), 4)}


def report(results, tids):
    fields = ("learned",) + tuple("learned_top%d" % kk for kk in TOPK) + ("default", "nearest")
    rep = {"heldout_shapes": len(results)}
    for f in fields:
        rep[f] = summary(results, f)
    for tid in tids:
        sub = [x for x in results if x["key"][0] == tid]
        if sub:
            rep["type%d" % tid] = {f: summary(sub, f) for f in fields}
    return rep


def _time_configs(configs, w, rounds=3, reps=10, stream=None):
    """median ms per launch of every configuration on workload w: `rounds` rounds that each time every
    configuration once (reps launches back to back between two HIP events), interleaved so that drift in
    clock or temperature hits all configurations alike"""
    tid = lx.typeid(w.dtype)
    a, b, c = (lx.DeviceBuffer.from_array(v) for v in (w.a, w.b_device_layout(), w.c))
    st = lx.DeviceBuffer.from_array(w.stack)
    e0, e1 = lx.Event(), lx.Event()
    times = [[] for _ in configs]
    try:
        for _ in range(rounds):
            for i, cfg in enumerate(configs):
                lx.process_config(cfg, st, w.stack_size, tid, a, b, c, w.m, w.n, w.k, stream)
                e0.record(stream)
                for _ in range(reps):
                    lx.process_config(cfg, st, w.stack_size, tid, a, b, c, w.m, w.n, w.k, stream)
                e1.record(stream)
                stream.sync()
                times[i].append(e0.elapsed_ms(e1) / reps)
    finally:
        e0.close()
        e1.close()
        for x in (a, b, c, st):
            x.free()
    return [float(np.median(t)) for t in times]


def evaluate_live(model, sweep, keys, stack_size=30000, nbest=3, topk=4, stream=None, log=None):
    """Held-out shapes timed now: the network's choice, its top-K and the sweep's nbest fastest runnable
    configurations are timed together (interleaved rounds, medians) on a fresh S-entry stack of the shape.
    The exhaustive best is the fastest of the sweep's nbest re-timed (one noisy sweep sample no longer
    decides it); each method's fraction is t_best / t_method."""
    from libxstream_amd.workload import make_stack_workload

    out = []
    for key in keys:
        tid, m, n, k = key
        rows = sorted(runnable(sweep[key]), key=lambda r: -r["gflops"])[:nbest]
        cands, index = [], {}

        def add(cfg):
            ck = config_key(cfg)
            if ck not in index:
                index[ck] = len(cands)
                cands.append(cfg)
            return index[ck]

        best_ix = [add(to_config(r)) for r in rows]
        ranked = model.predict_topk(*key, topk)
        top_ix = [add(c) for c, _ in ranked]
        w = make_stack_workload(m, n, k, stack_size, dtype=lx.nptype(tid), seed=m * 7919 + n * 104729 + k)
        t = _time_configs(cands, w, stream=stream)
        t_best = min(t[i] for i in best_ix)
        entry = {"key": list(key), "best_ms": t_best, "learned_ms": t[top_ix[0]],
                 "learned": t_best / t[top_ix[0]], "learned_top%d" % topk: t_best / min(t[i] for i in top_ix),
                 "confidence": ranked[0][1], "margin": ranked[0][1] - (ranked[1][1] if len(ranked) > 1 else 0.0)}
        out.append(entry)
        if log is not None:
            log("%s learned %.3f top%d %.3f" % (key, entry["learned"], topk, entry["learned_top%d" % topk]))
    return out


def summary_live(results, field):
    r = np.array([x[field] for x in results])
    return {"mean": round(float(r.mean()), 4), "median": round(float(np.median(r)), 4), "min": round(float(r.min()), 4),
            "within_10pct": round(float((r >= 0.9).mean()), 4), "geomean": round(float(np.exp(np.log(r).mean())), 4)}


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
    r.add_argument("--epochs", type=int, default=4000)
    r.add_argument("--nets", type=int, default=8)
    e = sub.add_parser("evaluate", help="evaluate a model on the held-out shapes of the sweeps")
    e.add_argument("sweeps", nargs="+")
    e.add_argument("--model", required=True)
    v = sub.add_parser("live", help="time the held-out choices now against the sweeps' fastest (GPU)")
    v.add_argument("sweeps", nargs="+")
    v.add_argument("--model", required=True)
    v.add_argument("--report", help="output JSON report")
    args = p.parse_args(argv)
    if args.cmd == "triplets":
        write_triplets(args.path)
        return 0
    lx.init()
    sweep = load_sweep(args.sweeps)
    train_t, held_t = split(triplets())
    tids = sorted({key[0] for key in sweep})
    train_keys = [(tid,) + t for tid in tids for t in train_t if (tid,) + t in sweep]
    held_keys = [(tid,) + t for tid in tids for t in held_t if (tid,) + t in sweep]
    if args.cmd == "live":
        model = lx.Mlp(path=args.model)
        stream = lx.Stream("live")
        results = evaluate_live(model, sweep, held_keys, stream=stream, log=lambda x: print(x, flush=True))
        rep = {"heldout_shapes": len(results)}
        for f in ("learned", "learned_top4"):
            rep[f] = summary_live(results, f)
            for tid in tids:
                rep.setdefault("type%d" % tid, {})[f] = summary_live([x for x in results if x["key"][0] == tid], f)
        print(json.dumps(rep, indent=1))
        if args.report:
            rep["results"] = results
            with open(args.report, "w") as f:
                json.dump(rep, f)
        stream.close()
        model.close()
        lx.finalize()
        return 0
    if args.cmd == "train":
        model, loss = train(sweep, train_keys, nnets=args.nets, epochs=args.epochs)
        model.save(args.model)
        print("trained %d networks on %d shapes: loss %.5f" % (args.nets, len(train_keys), loss), flush=True)
    else:
        model, loss = lx.Mlp(path=args.model), None
    results = evaluate(model, sweep, held_keys, train_keys)
    rep = report(results, tids)
    rep["train_shapes"], rep["loss"] = len(train_keys), loss
    print(json.dumps(rep, indent=1))
    if getattr(args, "report", None):
        rep["results"] = results
        with open(args.report, "w") as f:
            json.dump(rep, f)
    model.close()
    lx.finalize()
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(main(sys.argv[1:]))
