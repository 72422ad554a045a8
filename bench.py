"""Benchmark: batched conservative-advancement CCD on one MI355X.

One step = one launch of the CCD kernel over the whole query batch (every
query runs its full CA loop: the reference's C2A_QueryTimeOfContact).  The
workload is the knot-pair batch: two (2,3) torus-knot tubes of 8192 FP64
triangles each, 4096 independent queries with seeded poses (translation
U[-2,2]^3, uniform random rotation), about 30% of them colliding (end poses
pulled toward collision by a seeded lerp; c2a_amd.workloads.
knot_batch_queries), tolerance_d = 1e-3, tolerance_t = 1e-4.  Inputs are
uploaded once; the timed region covers the kernel launches only.

Prints one JSON line.  `value` is the plain (reference-order, bit-exact)
throughput; the `learned` object times the same batch with the shipped
learned front ordering (c2a_amd/scorer_knot_uneven.c2as) and reports its BV
tests, TOC error against the plain order and a conservativeness count (the
colliding queries whose TOC pose is already in contact), for both orders;
the `strict` object times both orders in strict (conservative) mode.  The
CPU baseline is the FP64 oracle timed on this host's cores over a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CCD queries/s, 1 MI355X: plain vs learned-order BVTT at TOC err<=1e-4"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per unit of work (one FP32-equivalent RSS node record per
# BV test, two FP64 triangles per triangle test, two FP64 poses per CA step).
BYTES_PER_BV_TEST = 68
BYTES_PER_TRI_TEST = 144
BYTES_PER_CA_STEP = 192


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nq", type=int, default=0, help="queries (default: the workload's)")
    ap.add_argument("--workload", choices=["knot", "uneven", "grazing", "demo"], default="knot",
                    help="knot: the headline batch; uneven / grazing: the other configurations; "
                         "demo: the reference demo's 303 bunny-vs-bunny queries")
    ap.add_argument("--spec-depth", type=int, default=0)
    ap.add_argument("--order", choices=["both", "plain", "learned"], default="both",
                    help="time the plain order, the learned order, or both (value = plain)")
    ap.add_argument("--scorer", default="", help="scorer file (default: the shipped one)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-strict", action="store_true",
                    help="skip the strict (conservative) mode figures")
    ap.add_argument("--no-distance", action="store_true",
                    help="skip the static distance-query figure (knot workload)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain 256 MiB device copy (to scale HBM counters)")
    return ap.parse_args()


def copy_calibration():
    """A plain device-to-device copy of known size: the FETCH_SIZE /
    WRITE_SIZE counters of this dispatch give the scale factor between the
    counters and bytes moved."""
    import torch
    n = 256 * 1024 * 1024 // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 8, "launches": 3}), flush=True)


def cpu_baseline(W, nq_sample, runs=5, single_sample=512):
    """Oracle (reference-faithful FP64 restatement, built with -O3) timed on
    this host: all cores over the first nq_sample queries of the same
    workload, median of `runs` runs after one warm-up run, and one core over
    the first single_sample queries (median of 3)."""
    import numpy as np

    from oracle import oracle as O
    threads = min(64, os.cpu_count() or 1)
    mp_all = W["pairs"]
    used = sorted(set(mp_all[:nq_sample].reshape(-1).tolist()))
    models = {i: O.OModel(W["tris"][i]) for i in used}
    twins = {i: O.OModel(W["tris"][i]) for i in used}   # the second object of a self pair

    def once(n, nthreads):
        P = [p[:n] for p in W["P"]]
        mp = mp_all[:n]
        t0 = time.perf_counter()
        for a, b in sorted(set(map(tuple, mp))):
            sel = np.nonzero((mp[:, 0] == a) & (mp[:, 1] == b))[0]
            mb = models[b] if a != b else twins[b]
            O.batch_toc(models[a], mb, *[p[sel] for p in P], W["tol_d"], W["tol_t"],
                        seeds=np.zeros((len(sel), 2), np.int32), nthreads=nthreads)
        return time.perf_counter() - t0

    once(nq_sample, threads)                              # warm-up
    walls = sorted(once(nq_sample, threads) for _ in range(runs))
    dt = walls[len(walls) // 2]
    n1 = min(single_sample, nq_sample)
    walls1 = sorted(once(n1, 1) for _ in range(3))
    dt1 = walls1[1]
    return {"value": nq_sample / dt, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": "first %d queries of the same workload, %d threads, median of %d runs "
                      "after a warm-up: %.3f s (runs %s)"
                      % (nq_sample, threads, runs, dt, ", ".join("%.3f" % w for w in walls)),
            "single_core": {"value": n1 / dt1, "unit": "queries/s", "cores": 1,
                            "sample": "first %d queries, median of 3: %.3f s" % (n1, dt1)}}


def source_digest():
    """sha256 of the library's sources (c2a_amd/csrc and include/): ties a
    committed profile to the code it measured."""
    import hashlib
    h = hashlib.sha256()
    files = []
    for d in ("c2a_amd/csrc", "include"):
        for dp, _, fs in os.walk(os.path.join(ROOT, d)):
            files += [os.path.join(dp, f) for f in fs]
    for f in sorted(files, key=lambda x: os.path.relpath(x, ROOT)):
        h.update(os.path.relpath(f, ROOT).encode())
        h.update(open(f, "rb").read())
    return h.hexdigest()[:16]


def distance_bench(meshes, tris, nq=65536, sample=4096, reps=3):
    """Static distance queries (C2A_Distance, zero error bounds) between the
    two knots at the start poses of the knot workload's motions: GPU
    queries/s through the one-shot C ABI call (host arrays in and out, so the
    figure includes the PCIe copies and the call's device allocations), and
    the oracle on all host cores over the first `sample` queries."""
    import numpy as np

    import c2a_amd as G
    from c2a_amd import workloads as Wm
    from oracle import oracle as O
    P = Wm.knot_queries(nq, 2024, collide_frac=0.3)
    poses = np.ascontiguousarray(np.concatenate([P[0], P[2]], axis=1))
    pairs = np.tile(np.array([0, 1], np.int32), (nq, 1))
    ref = G.distance(meshes, poses, pairs)   # warm-up
    t0 = time.perf_counter()
    for _ in range(reps):
        got = G.distance(meshes, poses, pairs)
    gpu_s = (time.perf_counter() - t0) / reps
    if not np.array_equal(ref["distance"], got["distance"]):
        raise SystemExit("bench.py: distance results changed between calls")
    threads = min(64, os.cpu_count() or 1)
    m1, m2 = O.OModel(tris), O.OModel(tris)
    t0 = time.perf_counter()
    cpu = O.distance_batch(m1, m2, poses[:sample], nthreads=threads)
    cpu_s = time.perf_counter() - t0
    same = all(c.distance == g and c.bv_tests == b
               for c, g, b in zip(cpu, got["distance"][:sample], got["bv_tests"][:sample]))
    return {"value": nq / gpu_s, "unit": "queries/s", "nq": nq,
            "timing": "host to host through c2a_distance (PCIe copies and allocations included)",
            "bv_tests_mean": float(got["bv_tests"].mean()),
            "zero_distance_fraction": float((got["distance"] == 0).mean()),
            "matches_oracle_on_sample": bool(same),
            "cpu_baseline": {"value": sample / cpu_s, "unit": "queries/s", "cores": threads,
                             "kind": "port",
                             "sample": "first %d queries, %d threads, %.2f s wall"
                                       % (sample, threads, cpu_s)}}


def overshoots(r):
    """Colliding results whose TOC pose is already in contact: toc > 0 and the
    final step's distance, which is exact there, is 0 (the CA loop stepped
    past the first contact)."""
    return int(((r["collisionfree"] == 0) & (r["toc"] > 0) & (r["distance"] == 0)).sum())


def time_batch(batch, steps, warmup):
    """Warm-up launches, then `steps` timed launches; (results, wall ms per
    launch, mean kernel ms)."""
    import numpy as np
    import torch
    for _ in range(warmup):
        batch.run()
    torch.cuda.synchronize()
    ref = batch.fetch() if warmup > 0 else None
    kernel_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run()
        kernel_ms.append(batch.last_ms())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    res = batch.fetch()
    for f in ("toc", "distance", "numCA", "bv_tests", "tri_tests"):
        if ref is not None and not np.array_equal(ref[f], res[f]):
            raise SystemExit("bench.py: results changed between launches (%s)" % f)
    return res, wall * 1000.0 / steps, float(np.mean(kernel_ms))


def load_workload(name, nq):
    """Synthetic inputs of one configuration (host arrays)."""
    import numpy as np

    from c2a_amd import workloads as W
    if name == "knot":
        cfg = W.CONFIG_KNOT_BATCH
        nq = min(nq or cfg["nq"], cfg["nq"])
        tris, idx = W.torus_knot()
        P = W.knot_batch_queries(nq)
        return dict(tris=[tris.reshape(-1, 9)] * 2, vidx=[idx] * 2, P=P,
                    pairs=np.tile(np.array([0, 1], np.int32), (nq, 1)), tol_d=cfg["tol_d"],
                    tol_t=cfg["tol_t"], nq=nq, sample=nq,
                    desc="knot-pair batch: %d CCD queries between two (2,3) torus-knot tubes "
                         "(8192 triangles each), translation U[-2,2]^3, uniform rotations, "
                         "tol_d 1e-3, tol_t 1e-4, about 30%% colliding (seeded candidate "
                         "streams 20261/20271, mask 20281: tests/golden/knot_batch_4096.npz)"
                         % nq)
    if name == "uneven":
        cfg = W.CONFIG_UNEVEN
        nq = nq or cfg["nq"]
        tab = W.uneven_meshes()
        sizes = [len(t) for t, _, _ in tab]
        P00, P01, P10, P11, mp = W.uneven_queries(nq, cfg["seed"], sizes, [r for _, _, r in tab])
        return dict(tris=[t for t, _, _ in tab], vidx=[i for _, i, _ in tab],
                    P=(P00, P01, P10, P11), pairs=mp, tol_d=cfg["tol_d"], tol_t=cfg["tol_t"],
                    nq=nq, sample=min(nq, 4096),
                    desc="uneven batch: %d queries over icospheres of 80..81,920 triangles "
                         "(value noise 0.1), a 64x32 torus and the knot; log-uniform mesh "
                         "sizes, rigid motions spanning 0.5-4 body radii" % nq)
    if name == "demo":
        t9, idx = W.bunny()
        P = W.demo_queries()
        n = len(P[0])
        return dict(tris=[t9, t9], vidx=[idx, idx], P=P,
                    pairs=np.tile(np.array([0, 1], np.int32), (n, 1)), tol_d=1e-4, tol_t=1e-4,
                    nq=n, sample=n,
                    desc="reference demo: %d C2A_Solve queries, two 69,664-triangle bunnies "
                         "(tri_models/bunny_noholes.tri) moved by the torus-knot animations, "
                         "tol_d = tol_t = 1e-4" % n)
    cfg = W.CONFIG_GRAZING
    nq = nq or cfg["nq"]
    s, idx = W.uv_sphere(cfg["n_lon"], cfg["n_lat"], cfg["noise"], 3)
    P = W.grazing_queries(nq, cfg["seed"])
    return dict(tris=[s.reshape(-1, 9)] * 2, vidx=[idx] * 2, P=P,
                pairs=np.tile(np.array([0, 1], np.int32), (nq, 1)), tol_d=cfg["tol_d"],
                tol_t=cfg["tol_t"], nq=nq, sample=nq,
                desc="grazing contacts: %d near-tangential passes between two noisy 512x256 UV "
                     "spheres (262,144 triangles each), closest approach within 5e-3 radii" % nq)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py: measures one GPU only (the CCD batch has no "
              "cross-GPU step); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import c2a_amd as G

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    wl = load_workload(args.workload, args.nq)
    nq = wl["nq"]
    if args.workload == "knot":
        t9, idx = wl["tris"][0], wl["vidx"][0]
        meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    else:
        meshes = [G.Mesh(t, vidx=i) for t, i in zip(wl["tris"], wl["vidx"])]
    queries = G.make_queries(*wl["P"], mesh_pairs=wl["pairs"])
    res = learned = None
    if args.order in ("both", "plain"):
        params = G.default_params(tol_d=wl["tol_d"], tol_t=wl["tol_t"], with_contacts=0,
                                  spec_depth=args.spec_depth)
        res, ms_per_step, kms = time_batch(G.Batch(meshes, queries, params), args.steps,
                                           args.warmup)
    if args.order in ("both", "learned"):
        from c2a_amd import learn as Lm
        scorer = G.Scorer(path=args.scorer or Lm.DEFAULT_SCORER)
        lp = G.default_params(tol_d=wl["tol_d"], tol_t=wl["tol_t"], with_contacts=0,
                              order=G.ORDER_LEARNED, scorer=scorer)
        lres, lms_step, lkms = time_batch(G.Batch(meshes, queries, lp, scorer=scorer),
                                          args.steps, args.warmup)
        learned = {"value": nq / (lms_step / 1000.0), "ms_per_step": lms_step, "kernel_ms": lkms,
                   "bv_tests": int(lres["bv_tests"].sum()),
                   "scorer": os.path.basename(args.scorer or Lm.DEFAULT_SCORER)}
        if res is not None:
            dt = lres["toc"] - res["toc"]
            learned.update({
                "speedup_vs_plain": ms_per_step / lms_step,
                "bv_ratio_vs_plain": float(res["bv_tests"].sum() / max(1, lres["bv_tests"].sum())),
                "toc_err_max": float(np.abs(dt).max()),
                "toc_err_mean": float(np.abs(dt).mean()),
                "toc_within_1e-4": float((np.abs(dt) <= 1e-4).mean()),
                "meets_toc_err_1e-4": bool((np.abs(dt) <= 1e-4).all()),
                "toc_err_note": "a different sibling order gives a different but valid mint "
                                "at each CA step, so the two runs stop at different poses "
                                "within the last step (about tol_d / speed): only the "
                                "reference order reproduces the oracle's TOC",
                "collision_flag_mismatches": int((lres["collisionfree"] != res["collisionfree"]).sum()),
                "conservativeness": {
                    "colliding": int((res["collisionfree"] == 0).sum()),
                    "overshoots_plain": overshoots(res),
                    "overshoots_learned": overshoots(lres),
                    "definition": "colliding queries whose TOC pose is in contact (toc > 0, "
                                  "exact final distance 0); the reference order's count is "
                                  "the plain figure"}})
        if res is None:
            res, ms_per_step, kms = lres, lms_step, lkms
    strict = None
    if not args.no_strict and args.order == "both":
        # strict (conservative) mode: world-frame rotation axis in the motion
        # bounds, every tested leaf bounds mint; zero TOC poses in contact
        strict = {}
        for name in ("plain", "learned"):
            kw = {}
            if name == "learned":
                kw = dict(order=G.ORDER_LEARNED, scorer=scorer)
            sp = G.default_params(tol_d=wl["tol_d"], tol_t=wl["tol_t"], with_contacts=0,
                                  strict=1, **kw)
            sres, sms, skms = time_batch(G.Batch(meshes, queries, sp, scorer=kw.get("scorer")),
                                         min(args.steps, 5), 1)
            strict[name] = {"value": nq / (sms / 1000.0), "ms_per_step": sms,
                            "kernel_ms": skms, "bv_tests": int(sres["bv_tests"].sum()),
                            "colliding": int((sres["collisionfree"] == 0).sum()),
                            "toc_pose_in_contact": overshoots(sres)}
            if res is not None:
                dt = sres["toc"] - res["toc"]
                strict[name]["toc_diff_vs_reference_order_max"] = float(np.abs(dt).max())
                strict[name]["collision_flag_mismatches_vs_reference_order"] = int(
                    (sres["collisionfree"] != res["collisionfree"]).sum())
    value = nq / (ms_per_step / 1000.0)
    algo_bytes = (BYTES_PER_BV_TEST * int(res["bv_tests"].sum()) +
                  BYTES_PER_TRI_TEST * int(res["tri_tests"].sum()) +
                  BYTES_PER_CA_STEP * int(res["numCA"].sum()))
    achieved = algo_bytes / (kms / 1000.0) / 1e9
    # measured HBM bytes per launch, only from a counter profile of these
    # very sources (the profile records the source digest it was taken at)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ccd_kernel_knot4096_traffic.json")
    if args.workload == "knot" and nq == 4096 and args.order != "learned" and \
            os.path.exists(tpath):
        try:
            prof = json.load(open(tpath))
            if prof.get("source_digest") == source_digest():
                traffic = prof.get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": wl["desc"],
            "order": ("plain (reference DFS order, bit-exact)" if args.order != "learned"
                      else "learned (scorer-ordered siblings)"),
            "nq": nq,
            "spec_depth": args.spec_depth or "autotuned per mesh-size bucket (c2a_amd/tuning.json)",
            "bv_tests": int(res["bv_tests"].sum()),
            "tri_tests": int(res["tri_tests"].sum()),
            "ca_steps": int(res["numCA"].sum()),
            "collision_fraction": float(1.0 - res["collisionfree"].mean()),
            "kernel_ms": kms,
            "bv_tests_per_s": float(res["bv_tests"].sum()) / (kms / 1000.0),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
    }
    if learned is not None and args.order == "both":
        out["learned"] = learned
    if strict is not None:
        out["strict"] = strict
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, wl["sample"])
    if args.workload == "knot" and not args.no_distance:
        out["distance_query"] = distance_bench(meshes, wl["tris"][0])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
