"""The tail pass on the GPU: queries that reach the BV-test budget at the end
of a CA step are handed from the one-wavefront launch to a second launch
with four wavefronts per query, which continues them from their saved
CA-loop state (TailRec).  Every field must equal the oracle's and the
one-launch run's, bit for bit; the CPU harness checks the same hand-over
with the kernel's code (tests/test_kernel_logic_cpu.py)."""
import numpy as np
import pytest
import torch   # before the library: torch's HIP runtime must be the first one loaded

import c2a_amd as G
from c2a_amd import workloads as W
from oracle import oracle as O
from tests.test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu

ALL = ("toc", "distance", "mint", "numCA", "collisionfree", "bv_tests", "tri_tests",
       "num_contact", "flags", "last_tri", "R", "T", "R_toc", "T_toc", "p1", "p2",
       "abs_err", "rel_err", "UpboundTOC")


def same_all(a, b, label):
    for f in ALL:
        assert np.array_equal(a[f], b[f]), (label, f)


@pytest.fixture(scope="module")
def knot_setup(knot_tris):
    t9, idx = knot_tris
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    P = W.knot_batch_queries(512)
    return t9, meshes, P


@pytest.mark.parametrize("budget", [2000, 20000])
def test_tail_pass_bit_exact(knot_setup, budget):
    t9, meshes, P = knot_setup
    print("tail pass: knot batch 512, budget %d" % budget, flush=True)
    q = G.make_queries(*P)
    b = G.Batch(meshes, q, G.default_params(tol_d=1e-3, tol_t=1e-4, tail_budget=budget))
    b.run()
    got = b.fetch()
    handed = b.tail_count()
    one = G.ccd_batch(meshes, q, G.default_params(tol_d=1e-3, tol_t=1e-4, tail_budget=0))
    ref = O.batch_toc(O.OModel(t9), O.OModel(t9), *P, 1e-3, 1e-4,
                      seeds=np.zeros((len(P[0]), 2), np.int32), with_contacts=True, nthreads=8)
    assert_same(ref, got, "tail budget %d" % budget)
    same_all(one, got, "tail budget %d against one launch" % budget)
    # only queries that made at least `budget` tests can have been handed over
    assert 0 < handed <= int((ref["num_bv_tests"] >= budget).sum())


def test_tail_pass_strict_and_tuned(knot_setup):
    """Strict mode and the autotuned knobs (depth 4 in the first launch,
    template depth 5 in the tail pass) give the one-launch results."""
    _, meshes, P = knot_setup
    q = G.make_queries(*P)
    for kw in (dict(strict=1), dict(spec_depth=0), dict(spec_depth=3, miss_depth=1)):
        print("tail pass: %s" % kw, flush=True)
        a = G.ccd_batch(meshes, q, G.default_params(tol_d=1e-3, tol_t=1e-4, tail_budget=3000,
                                                    **kw))
        b = G.ccd_batch(meshes, q, G.default_params(tol_d=1e-3, tol_t=1e-4, tail_budget=0, **kw))
        same_all(b, a, str(kw))


def test_tail_pass_graph_replay(knot_setup):
    """Both launches in one captured graph: every replay equals the eager run,
    with eager runs in between."""
    _, meshes, P = knot_setup
    q = G.make_queries(*P)
    b = G.Batch(meshes, q, G.default_params(tol_d=1e-3, tol_t=1e-4, tail_budget=5000))
    b.run()
    eager = b.fetch()
    assert b.tail_count() > 0
    for k in range(3):
        b.run_graph()
        torch.cuda.synchronize()
        same_all(eager, b.fetch(), "replay %d" % k)
        assert b.tail_count() > 0
        b.run()
        same_all(eager, b.fetch(), "eager %d" % k)


def test_tail_pass_off_for_learned_order(knot_setup):
    _, meshes, P = knot_setup
    sc = G.Scorer()
    b = G.Batch(meshes, G.make_queries(*[p[:64] for p in P]),
                G.default_params(tol_d=1e-3, tol_t=1e-4, tail_budget=100,
                                 order=G.ORDER_LEARNED, scorer=sc), scorer=sc)
    b.run()
    b.fetch()
    assert b.tail_count() == 0
