"""Learned front ordering: training data from the exact GPU traversal, the
scorer's training on the GPU, and its evaluation.

The data come from the GPU traversal itself: every CA step after the first
replays the previous step's traversal, and each replayed pair whose
replayed subtree holds a leaf pair becomes a sample -- its 32 features
(c2a_amd/csrc/scorer.h) and the label log(closest replayed triangle distance
in that subtree + 1e-4).  Training has two stages: a first scorer learns
from the plain (reference-order) traversal's samples; then the traversal
runs in that scorer's order, and the final scorer learns from both sets, so
it also sees the pairs its own order visits.  Splits are by query, by
hashing the query index with a seed, so no query contributes to both sides.
Training runs the project's own matrix-core forward/backward kernels
(c2a_scorer_train).  The scorer only orders siblings; it never prunes a
pair, so every result stays a conservative-advancement result.

    python -m c2a_amd.learn --out c2a_amd/scorer_knot_uneven.c2as
"""
import argparse
import os
import time

import numpy as np

import c2a_amd as G
from c2a_amd import workloads as W

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_SCORER = os.path.join(_HERE, "scorer_knot_uneven.c2as")


def collect_samples(meshes, queries, tol_d, tol_t, rate, cap, scorer=None):
    """One batch run with sample logging, in the plain order or in `scorer`'s
    learned order: (X, y, keys, selected)."""
    kw = dict(order=G.ORDER_LEARNED, scorer=scorer) if scorer is not None else {}
    p = G.default_params(tol_d=tol_d, tol_t=tol_t, with_contacts=0, sample_rate=rate,
                         sample_cap=cap, **kw)
    b = G.Batch(meshes, queries, p)
    b.run()
    return b.samples(cap)


def query_of(keys):
    return keys >> 34


def split_mask(keys, train_frac, seed):
    """True for samples of training queries (a seeded hash of the query index)."""
    q = query_of(keys).astype(np.uint64)
    with np.errstate(over="ignore"):
        h = (q + np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)) * np.uint64(0xBF58476D1CE4E5B9)
        h ^= h >> np.uint64(31)
    return (h % np.uint64(1000)) < np.uint64(int(train_frac * 1000))


def query_split(nq, train_frac, seed):
    """The same split over query indices 0..nq-1."""
    return split_mask(np.arange(nq, dtype=np.int64) << 34, train_frac, seed)


def sibling_accuracy(keys, labels, scores):
    """Fraction of sibling pairs (records 2e+1, 2e+2 of one step, both
    sampled, with different labels) that `scores` order the way the labels
    do (lower first), and the number of such pairs.  Passing feature 18
    (log RSS distance) as the scores measures the plain order."""
    o = np.argsort(keys, kind="stable")
    k, y, s = keys[o], labels[o], scores[o]
    first = (k & 1) == 1            # record 2e+1
    nxt = np.zeros_like(first)
    nxt[:-1] = k[1:] == k[:-1] + 1
    i = np.nonzero(first & nxt)[0]
    if len(i) == 0:
        return float("nan"), 0
    j = i + 1
    differ = y[i] != y[j]
    i, j = i[differ], j[differ]
    good = (s[j] < s[i]) == (y[j] < y[i])
    return float(good.mean()), len(i)


def train_scorer(X, y, epochs=8, batch=2048, lr=2e-3, seed=0, log=print):
    s = G.Scorer(seed)
    t0 = time.perf_counter()
    losses = s.train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    log("trained on %d samples, %d epochs in %.1f s; loss per epoch %s"
        % (len(y), epochs, time.perf_counter() - t0, np.array2string(losses, precision=4)))
    return s


def knot_batch(nq, seed):
    tris, idx = W.torus_knot()
    t9 = tris.reshape(-1, 9)
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    P = W.knot_queries(nq, seed, collide_frac=0.3)
    return meshes, P, None


def uneven_batch(nq, seed):
    tab = W.uneven_meshes()
    meshes = [G.Mesh(t, vidx=i) for t, i, _ in tab]
    P00, P01, P10, P11, mp = W.uneven_queries(nq, seed, [len(t) for t, _, _ in tab],
                                              [r for _, _, r in tab])
    return meshes, (P00, P01, P10, P11), mp


def compare(plain, learned, mask=None):
    """Work and TOC agreement of a learned-order run against the plain run."""
    if mask is not None:
        plain, learned = plain[mask], learned[mask]
    dt = learned["toc"] - plain["toc"]
    return {
        "queries": int(len(plain)),
        "bv_tests_plain": int(plain["bv_tests"].sum()),
        "bv_tests_learned": int(learned["bv_tests"].sum()),
        "bv_ratio": float(plain["bv_tests"].sum() / max(1, learned["bv_tests"].sum())),
        "toc_err_max": float(np.abs(dt).max()) if len(dt) else 0.0,
        "toc_err_mean": float(np.abs(dt).mean()) if len(dt) else 0.0,
        "toc_within_1e-4": float((np.abs(dt) <= 1e-4).mean()) if len(dt) else 1.0,
        "collision_flag_mismatches": int((plain["collisionfree"] != learned["collisionfree"]).sum()),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_SCORER)
    ap.add_argument("--knot-nq", type=int, default=4096)
    ap.add_argument("--uneven-nq", type=int, default=16384)
    ap.add_argument("--rate", type=float, default=0.1)
    ap.add_argument("--cap", type=int, default=12_000_000)
    ap.add_argument("--epochs", type=int, default=12, help="first-stage epochs")
    ap.add_argument("--epochs2", type=int, default=16, help="second-stage epochs (0: one stage)")
    ap.add_argument("--split-seed", type=int, default=7)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args()
    Xs, ys, runs = [], [], {}
    for name, (meshes, P, mp), cfg, nq in (
            ("knot", knot_batch(a.knot_nq, W.CONFIG_KNOT_BATCH["seed"]), W.CONFIG_KNOT_BATCH,
             a.knot_nq),
            ("uneven", uneven_batch(a.uneven_nq, W.CONFIG_UNEVEN["seed"]), W.CONFIG_UNEVEN,
             a.uneven_nq)):
        q = G.make_queries(*P, mesh_pairs=mp)
        X, y, k, n = collect_samples(meshes, q, cfg["tol_d"], cfg["tol_t"], a.rate, a.cap)
        if n > a.cap:
            raise SystemExit("%s: %d samples selected, above the cap %d (lower --rate)"
                             % (name, n, a.cap))
        m = split_mask(k, 0.8, a.split_seed)
        acc, npair = sibling_accuracy(k[~m], y[~m], X[~m, 18])
        print("%s: %d samples, %d for training; held-out sibling pairs %d, plain-order "
              "accuracy %.4f" % (name, n, m.sum(), npair, acc), flush=True)
        Xs.append(X[m])
        ys.append(y[m])
        runs[name] = (meshes, q, cfg, nq, X[~m], y[~m], k[~m])
    X, y = np.concatenate(Xs), np.concatenate(ys)
    s = train_scorer(X, y, epochs=a.epochs, seed=a.seed)
    if a.epochs2 > 0:
        # second stage: samples of the training queries in the first scorer's order
        for name, (meshes, q, cfg, nq, _, _, _) in runs.items():
            X2, y2, k2, n2 = collect_samples(meshes, q, cfg["tol_d"], cfg["tol_t"], a.rate, a.cap,
                                             scorer=s)
            if n2 > a.cap:
                raise SystemExit("%s: %d samples selected, above the cap %d (lower --rate)"
                                 % (name, n2, a.cap))
            m2 = split_mask(k2, 0.8, a.split_seed)
            print("%s, learned order: %d samples, %d for training" % (name, n2, m2.sum()),
                  flush=True)
            Xs.append(X2[m2])
            ys.append(y2[m2])
        X, y = np.concatenate(Xs), np.concatenate(ys)
        s = train_scorer(X, y, epochs=a.epochs2, seed=a.seed)
    s.save(a.out)
    print("saved", a.out)
    for name, (meshes, q, cfg, nq, Xv, yv, kv) in runs.items():
        sc = s.infer(Xv)
        acc, npair = sibling_accuracy(kv, yv, sc)
        print("%s held-out: mse %.4f, learned sibling accuracy %.4f over %d pairs"
              % (name, float(np.mean((sc - yv) ** 2)), acc, npair))
        res = {}
        for order in ("plain", "learned"):
            kw = dict(order=G.ORDER_LEARNED, scorer=s) if order == "learned" else {}
            b = G.Batch(meshes, q, G.default_params(tol_d=cfg["tol_d"], tol_t=cfg["tol_t"],
                                                    with_contacts=0, **kw))
            b.run()
            b.run()
            res[order] = (b.fetch(), b.last_ms())
        held = ~query_split(nq, 0.8, a.split_seed)
        c = compare(res["plain"][0], res["learned"][0], held)
        c.update(plain_ms=res["plain"][1], learned_ms=res["learned"][1])
        print("%s held-out queries: %s" % (name, c), flush=True)


if __name__ == "__main__":
    main()
