"""Benchmark: batched conservative-advancement CCD on one MI355X.

One step = one launch of the CCD kernel over the whole query batch (every
query runs its full CA loop: the reference's C2A_QueryTimeOfContact).  The
workload is the knot-pair batch: two (2,3) torus-knot tubes of 8192 FP64
triangles each, 4096 independent queries with seeded poses (translation
U[-2,2]^3, uniform random rotation; 30% of end poses pulled toward
collision), tolerance_d = 1e-3, tolerance_t = 1e-4.  Inputs are uploaded
once; the timed region covers the kernel launches only.

Prints one JSON line (see README of the driver contract).  The CPU baseline is
the FP64 oracle timed on this host's cores over the same workload.
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
    ap.add_argument("--nq", type=int, default=4096)
    ap.add_argument("--spec-depth", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
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


def cpu_baseline(P, tris, nq_sample):
    """Oracle (reference-faithful FP64 restatement) on all host cores."""
    import numpy as np

    from oracle import oracle as O
    threads = min(64, os.cpu_count() or 1)
    t9 = tris.reshape(-1, 9)
    m1, m2 = O.OModel(t9), O.OModel(t9)
    sel = [p[:nq_sample] for p in P]
    seeds = np.zeros((nq_sample, 2), np.int32)
    t0 = time.perf_counter()
    O.batch_toc(m1, m2, *sel, 1e-3, 1e-4, seeds=seeds, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": nq_sample / dt, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": "first %d queries of the same knot-pair batch, %d threads, %.2f s wall"
                      % (nq_sample, threads, dt)}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the CCD batch has no "
              "cross-GPU step); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import c2a_amd as G
    from c2a_amd import workloads as W

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    cfg = W.CONFIG_KNOT_BATCH
    nq = args.nq
    tris, idx = W.torus_knot()
    t9 = tris.reshape(-1, 9)
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    P = W.knot_queries(nq, cfg["seed"], collide_frac=cfg["collide_frac"])
    params = G.default_params(tol_d=cfg["tol_d"], tol_t=cfg["tol_t"], with_contacts=0,
                              spec_depth=args.spec_depth)
    batch = G.Batch(meshes, G.make_queries(*P), params)

    for _ in range(args.warmup):
        batch.run()
    torch.cuda.synchronize()
    ref = batch.fetch() if args.warmup > 0 else None
    kernel_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
        kernel_ms.append(batch.last_ms())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    res = batch.fetch()
    for f in ("toc", "distance", "numCA", "bv_tests", "tri_tests"):
        if ref is not None and not np.array_equal(ref[f], res[f]):
            raise SystemExit("bench.py: results changed between launches (%s)" % f)

    ms_per_step = wall * 1000.0 / args.steps
    value = nq / (ms_per_step / 1000.0)
    algo_bytes = (BYTES_PER_BV_TEST * int(res["bv_tests"].sum()) +
                  BYTES_PER_TRI_TEST * int(res["tri_tests"].sum()) +
                  BYTES_PER_CA_STEP * int(res["numCA"].sum()))
    kms = float(np.mean(kernel_ms))
    achieved = algo_bytes / (kms / 1000.0) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ccd_kernel_knot4096_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
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
            "workload": "knot-pair batch: 4096 CCD queries between two (2,3) torus-knot tubes "
                        "(8192 triangles each), tol_d 1e-3, tol_t 1e-4, 30% pulled to collide",
            "order": "plain (reference DFS order, bit-exact)",
            "nq": nq,
            "spec_depth": args.spec_depth or 5,
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
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(P, tris, nq)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
