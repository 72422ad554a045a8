"""GPU parity: the HIP path through the C ABI against the FP64 oracle.

The GPU reproduces the reference's DFS decisions in reference order with the
same FP64 operation order (both sides built with -ffp-contract=off), so every
field is compared for exact equality: time of contact, distance, mint,
numCA, collision flag, BV/triangle test counts, contact count and the
updated last_tri pair.
"""
import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIELDS = [("toc", "toc"), ("distance", "distance"), ("mint", "mint"), ("numCA", "numCA"),
          ("collisionfree", "collisionfree"), ("num_bv_tests", "bv_tests"),
          ("num_tri_tests", "tri_tests"), ("num_contact", "num_contact"), ("flags", "flags")]


def assert_same(ref, got, label):
    for f, g in FIELDS:
        a, b = ref[f], got[g]
        if not np.array_equal(a, b):
            bad = np.nonzero(a != b)[0]
            raise AssertionError("%s: field %s differs at %d queries, first %s: oracle %s gpu %s"
                                 % (label, f, len(bad), bad[:5], a[bad[:5]], b[bad[:5]]))
    assert np.array_equal(ref["out_last_triA"], got["last_tri"][:, 0]), label
    assert np.array_equal(ref["out_last_triB"], got["last_tri"][:, 1]), label
    assert np.array_equal(ref["R"].reshape(-1, 9), got["R"]), label
    assert np.array_equal(ref["T"], got["T"]), label
    for f in ("abs_err", "rel_err", "UpboundTOC"):
        assert np.array_equal(ref[f], got[f]), (label, f)
    nc = ref["collisionfree"] == 0
    assert np.array_equal(ref["R_toc"].reshape(-1, 9)[nc], got["R_toc"][nc]), label
    assert np.array_equal(ref["T_toc"][nc], got["T_toc"][nc]), label


@pytest.fixture(scope="module")
def knot_pair(knot_tris):
    t9, idx = knot_tris
    return t9, [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)], O.OModel(t9), O.OModel(t9)


@pytest.mark.parametrize("spec_depth", [5, 1, 2, 3, 4, 6])
def test_knot_batch_bit_exact(knot_pair, spec_depth):
    t9, meshes, m1, m2 = knot_pair
    nq = 256 if spec_depth == 5 else 96
    print("knot batch nq=%d spec_depth=%d" % (nq, spec_depth), flush=True)
    P = W.knot_queries(nq, 20261 + spec_depth, collide_frac=0.3)
    got = G.ccd_batch(meshes, G.make_queries(*P),
                      G.default_params(tol_d=1e-3, tol_t=1e-4, spec_depth=spec_depth))
    ref = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=np.zeros((nq, 2), np.int32),
                      with_contacts=True, nthreads=8)
    assert_same(ref, got, "spec_depth=%d" % spec_depth)
    assert (got["collisionfree"] == 0).any() and (got["collisionfree"] == 1).any()


def test_solve_tolerances_bit_exact(knot_pair):
    """C2A_Solve's hard-coded 1e-4/1e-4 and its contact count (C2A.cpp:2384-2435)."""
    t9, meshes, m1, m2 = knot_pair
    nq = 128
    P = W.knot_queries(nq, 99, collide_frac=0.5)
    seeds = np.stack([np.arange(nq) % len(t9), (7 * np.arange(nq)) % len(t9)], 1).astype(np.int32)
    got = G.ccd_batch(meshes, G.make_queries(*P, seeds=seeds), G.default_params())
    ref = O.batch_toc(m1, m2, *P, 1e-4, 1e-4, seeds=seeds, with_contacts=True, nthreads=8)
    assert_same(ref, got, "solve")


@pytest.mark.parametrize("knob", ["ext_levels=2", "miss_depth=5", "replay_log=0"])
def test_traversal_knobs_bit_exact(knot_pair, knob):
    """Speculation and replay settings change the work, never the result."""
    t9, meshes, m1, m2 = knot_pair
    nq = 128
    name, val = knob.split("=")
    print("knot batch nq=%d %s" % (nq, knob), flush=True)
    P = W.knot_queries(nq, 4242, collide_frac=0.3)
    got = G.ccd_batch(meshes, G.make_queries(*P),
                      G.default_params(tol_d=1e-3, tol_t=1e-4, **{name: int(val)}))
    ref = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=np.zeros((nq, 2), np.int32),
                      with_contacts=True, nthreads=8)
    assert_same(ref, got, knob)


def test_knot_batch_4096_full_bit_exact(knot_pair):
    """The whole 4096-query knot batch (about 30% colliding) against the
    oracle, every field, with C2A_Solve's contact count."""
    t9, meshes, m1, m2 = knot_pair
    P = W.knot_batch_queries()
    nq = len(P[0])
    print("knot batch 4096 full", flush=True)
    got = G.ccd_batch(meshes, G.make_queries(*P), G.default_params(tol_d=1e-3, tol_t=1e-4))
    ref = O.batch_toc(m1, m2, *P, 1e-3, 1e-4, seeds=np.zeros((nq, 2), np.int32),
                      with_contacts=True, nthreads=16)
    assert_same(ref, got, "knot 4096")
    frac = float((got["collisionfree"] == 0).mean())
    assert 0.25 < frac < 0.35, frac


def test_deep_hierarchy_pair_bit_exact():
    """Hierarchy pairs deeper than 96 levels run on the 192-entry stack."""
    from tests.test_kernel_logic_cpu import chain_queries, factorial_chain
    t1, t2 = factorial_chain(100, 1), factorial_chain(100, 2, mirror=True)
    meshes = [G.Mesh(t1), G.Mesh(t2)]
    assert meshes[0].max_depth + meshes[1].max_depth + 2 > 96
    nq = 64
    P = chain_queries(nq, 808)
    seeds = np.full((nq, 2), 99, np.int32)
    print("deep hierarchy pair nq=%d" % nq, flush=True)
    got = G.ccd_batch(meshes, G.make_queries(*P, mesh_pairs=np.tile([0, 1], (nq, 1)), seeds=seeds),
                      G.default_params(tol_d=1e-6, tol_t=1e-6))
    ref = O.batch_toc(O.OModel(t1), O.OModel(t2), *P, 1e-6, 1e-6, seeds=seeds,
                      with_contacts=True, nthreads=8)
    assert_same(ref, got, "deep")
