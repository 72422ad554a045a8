"""Writes tests/golden/knot_batch_4096.npz: which seeded candidate queries
make up the 4096-query knot batch, so that about 30% of its queries collide.

Candidates come from c2a_amd.workloads.knot_queries (translation U[-2,2]^3,
uniform random rotations): a stream of plain candidates (seed SEED_FREE) and
a stream whose end pose of object 2 is pulled toward object 1's end pose by
the seeded lerp (collide_frac 1, seed SEED_HIT).  With these poses most
unpulled queries collide too (two knots of radius 1 within 2 of each other),
so each candidate is classified with the FP64 oracle (reference semantics,
tol_d 1e-3, tol_t 1e-4, seed pair (0, 0)): query i of the batch is the next
colliding pulled candidate when a seeded hash of i falls below 0.3, and the
next collision-free plain candidate otherwise.  The file holds the streams'
seeds and the chosen candidate indices; workloads.knot_batch_queries
rebuilds the poses from them.

    python tests/golden/make_knot_batch.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from c2a_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

NQ = 4096
SEED_FREE, SEED_HIT, SEED_MASK = 20261, 20271, 20281
HIT_FRAC = 0.3


def classify(m1, m2, P):
    r = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=np.zeros((len(P[0]), 2), np.int32),
                    nthreads=os.cpu_count() or 4)
    return r["collisionfree"] == 0


def main():
    is_hit = W.Rng(SEED_MASK, 9).uniform(NQ) < HIT_FRAC
    nhit = int(is_hit.sum())
    tris, _ = W.torus_knot()
    t9 = tris.reshape(-1, 9)
    m1, m2 = O.OModel(t9), O.OModel(t9)
    n_free_cand, n_hit_cand = 3 * (NQ - nhit), 2 * nhit
    free_ok = ~classify(m1, m2, W.knot_queries(n_free_cand, SEED_FREE, 0.0))
    hit_ok = classify(m1, m2, W.knot_queries(n_hit_cand, SEED_HIT, 1.0))
    free_idx = np.nonzero(free_ok)[0][:NQ - nhit].astype(np.int32)
    hit_idx = np.nonzero(hit_ok)[0][:nhit].astype(np.int32)
    assert len(free_idx) == NQ - nhit and len(hit_idx) == nhit, "too few candidates"
    np.savez_compressed(os.path.join(HERE, "knot_batch_4096.npz"),
                        seeds=np.array([SEED_FREE, SEED_HIT, SEED_MASK], np.int64),
                        n_cand=np.array([n_free_cand, n_hit_cand], np.int64),
                        is_hit=is_hit, free_idx=free_idx, hit_idx=hit_idx)
    print("knot batch: %d queries, %d colliding (%.3f); plain candidates used %d of %d "
          "(%.3f collision-free), pulled %d of %d (%.3f colliding)"
          % (NQ, nhit, nhit / NQ, free_idx[-1] + 1, n_free_cand, free_ok.mean(), hit_idx[-1] + 1,
             n_hit_cand, hit_ok.mean()))


if __name__ == "__main__":
    main()
