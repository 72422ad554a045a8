"""Autotuner of the CCD kernel's traversal knobs per mesh-size bucket.

A query of a batch belongs to the size bucket of its larger mesh.  For each
bucket the tuner takes the queries of the uneven batch (16,384 queries over
icospheres of 80..81,920 triangles, a torus and the knot) that fall in it --
and, for the bucket above 131,072 triangles, the grazing batch between two
262,144-triangle spheres -- and times one launch per candidate setting of

  spec_depth  levels of the speculative subtree one round computes (<= 6)
  miss_depth  depth of the first round below a replayed pair the step must
              expand further than the last one
  replay_log  expansions logged per CA step for the next step's replay
  ext_levels  levels the replay computes ahead below replayed pairs that
              newly pass the descend test

(depth and miss depth first at the bucket's default log without extension,
then the log at the best pair, then the extension levels at the best
triple), median of three kernel timings after a warm-up.  A larger log or
more extension levels cost arena memory and traffic, so they are taken only
when at least MARGIN faster than the best smaller setting.  Every
setting gives the same results bit for bit; only the time differs.  The
chosen knobs go to tuning.json, which c2a_amd/build.py compiles into the
library: c2a_batch_create applies them per query when spec_depth is 0.

    python -m c2a_amd.autotune --out c2a_amd/tuning.json
"""
import argparse
import json
import os
import time

import numpy as np

import c2a_amd as G
from c2a_amd import workloads as W

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(_HERE, "tuning.json")
BUCKETS = [128, 512, 2048, 8192, 32768, 131072, 1 << 30]
MARGIN = 0.02   # relative gain a larger replay log or extension must bring


def bucket_of(ntri):
    for b, m in enumerate(BUCKETS):
        if ntri <= m:
            return b
    return len(BUCKETS) - 1


def default_log(max_ntri):
    cap = 8192
    while cap < max_ntri // 8 and cap < (1 << 20):
        cap <<= 1
    return cap


def time_config(meshes, queries, tol_d, tol_t, depth, miss, log, ext=0, reps=3):
    p = G.default_params(tol_d=tol_d, tol_t=tol_t, with_contacts=0, spec_depth=depth,
                         miss_depth=miss, replay_log=log, ext_levels=ext)
    b = G.Batch(meshes, queries, p)
    b.run()
    ms = []
    for _ in range(reps):
        b.run()
        ms.append(b.last_ms())
    return float(np.median(ms)), b.fetch()


def tune_bucket(meshes, queries, tol_d, tol_t, max_ntri, depths, misses, logs, exts, log):
    base_log = default_log(max_ntri)
    trials = []
    best = None
    ref = None
    for d in depths:
        for m in misses:
            if m >= d:
                continue
            ms, res = time_config(meshes, queries, tol_d, tol_t, d, m, base_log)
            if ref is None:
                ref = res
            elif not np.array_equal(ref["toc"], res["toc"]):
                raise SystemExit("autotune: results differ between settings")
            trials.append({"spec_depth": d, "miss_depth": m, "replay_log": base_log,
                           "ext_levels": 0, "ms": ms})
            log("  depth %d miss %d log %d: %.2f ms" % (d, m, base_log, ms))
            if best is None or ms < best["ms"]:
                best = trials[-1]
    # logs in increasing size: a larger one must beat the best so far by MARGIN
    by_log = [best]
    for lg in logs:
        if lg == base_log:
            continue
        ms, res = time_config(meshes, queries, tol_d, tol_t, best["spec_depth"],
                              best["miss_depth"], lg)
        if not np.array_equal(ref["toc"], res["toc"]):
            raise SystemExit("autotune: results differ between settings")
        trials.append({"spec_depth": best["spec_depth"], "miss_depth": best["miss_depth"],
                       "replay_log": lg, "ext_levels": 0, "ms": ms})
        log("  depth %d miss %d log %d: %.2f ms" % (best["spec_depth"], best["miss_depth"], lg, ms))
        by_log.append(trials[-1])
    pick = None
    for t in sorted(by_log, key=lambda t: t["replay_log"]):
        if pick is None or t["ms"] < pick["ms"] * (1 - MARGIN):
            pick = t
    best = pick
    for ext in sorted(exts):
        if ext == 0:
            continue
        bd, bm, bl = best["spec_depth"], best["miss_depth"], best["replay_log"]
        ms, res = time_config(meshes, queries, tol_d, tol_t, bd, bm, bl, ext)
        if not np.array_equal(ref["toc"], res["toc"]):
            raise SystemExit("autotune: results differ between settings")
        trials.append({"spec_depth": bd, "miss_depth": bm, "replay_log": bl, "ext_levels": ext,
                       "ms": ms})
        log("  depth %d miss %d log %d ext %d: %.2f ms" % (bd, bm, bl, ext, ms))
        if ms < best["ms"] * (1 - MARGIN):
            best = trials[-1]
    return best, trials


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--depths", default="3,4,5,6")
    ap.add_argument("--misses", default="1,2,3")
    ap.add_argument("--logs", default="2048,8192,32768")
    ap.add_argument("--exts", default="0,1,2")
    ap.add_argument("--no-grazing", action="store_true")
    a = ap.parse_args()
    depths = [int(v) for v in a.depths.split(",")]
    misses = [int(v) for v in a.misses.split(",")]
    logs = [int(v) for v in a.logs.split(",")]
    exts = [int(v) for v in a.exts.split(",")]
    t0 = time.time()

    def log(msg):
        print("[%6.1fs] %s" % (time.time() - t0, msg), flush=True)

    cfg = W.CONFIG_UNEVEN
    tab = W.uneven_meshes()
    sizes = [len(t) for t, _, _ in tab]
    P00, P01, P10, P11, mp = W.uneven_queries(cfg["nq"], cfg["seed"], sizes, [r for _, _, r in tab])
    meshes = [G.Mesh(t, vidx=i) for t, i, _ in tab]
    big = np.maximum(np.array(sizes)[mp[:, 0]], np.array(sizes)[mp[:, 1]])
    qb = np.array([bucket_of(n) for n in big])
    rows, measured = [], {}
    for b, max_ntri in enumerate(BUCKETS):
        sel = np.nonzero(qb == b)[0]
        wl = None
        if len(sel) >= 32:
            q = G.make_queries(P00[sel], P01[sel], P10[sel], P11[sel], mesh_pairs=mp[sel])
            wl = ("uneven", meshes, q, len(sel), int(big[sel].max()), cfg["tol_d"], cfg["tol_t"])
        elif b == len(BUCKETS) - 1 and not a.no_grazing:
            g = W.CONFIG_GRAZING
            s, idx = W.uv_sphere(g["n_lon"], g["n_lat"], g["noise"], 3)
            t9 = s.reshape(-1, 9)
            gm = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
            P = W.grazing_queries(g["nq"], g["seed"])
            wl = ("grazing", gm, G.make_queries(*P), g["nq"], len(t9), g["tol_d"], g["tol_t"])
        if wl is None:
            rows.append({"max_ntri": max_ntri, "spec_depth": 5, "miss_depth": 2,
                         "replay_log": default_log(max_ntri), "ext_levels": 0})
            log("bucket <= %d: no queries, defaults" % max_ntri)
            continue
        name, ms_, q, n, mx, td, tt = wl
        log("bucket <= %d: %d %s queries (largest mesh %d triangles)" % (max_ntri, n, name, mx))
        best, trials = tune_bucket(ms_, q, td, tt, mx, depths, misses, logs, exts, log)
        rows.append({"max_ntri": max_ntri, "spec_depth": best["spec_depth"],
                     "miss_depth": best["miss_depth"], "replay_log": best["replay_log"],
                     "ext_levels": best["ext_levels"]})
        measured[str(max_ntri)] = {"workload": name, "queries": n, "best_ms": best["ms"],
                                   "trials": trials}
        log("  -> depth %d miss %d log %d ext %d (%.2f ms)"
            % (best["spec_depth"], best["miss_depth"], best["replay_log"], best["ext_levels"],
               best["ms"]))
    doc = json.load(open(DEFAULT_TABLE)) if os.path.exists(DEFAULT_TABLE) else {}
    doc["buckets"] = rows
    doc["measured"] = measured
    json.dump(doc, open(a.out, "w"), indent=1)
    log("wrote %s" % a.out)


if __name__ == "__main__":
    main()
