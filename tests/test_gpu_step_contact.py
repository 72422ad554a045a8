"""GPU parity of the explicit-step and contact entry points (C ABI) against
the oracle: c2a_ccd_step backs C2A_TimeOfContactStep (C2A.cpp:1778-1931),
c2a_contact_count backs C2A_QueryContact (C2A.cpp:1937-1966, 1525-1775).
Every field is compared for equality."""
import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def knot(knot_tris):
    t9, idx = knot_tris
    return t9, [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]


def test_ccd_step_bit_exact(knot):
    t9, meshes = knot
    nq = 160
    rng = np.random.default_rng(1234)
    P = W.knot_queries(nq, 777, collide_frac=0.5)
    steps = (G.StepQuery * nq)()
    ref = []
    m1, m2 = O.OModel(t9), O.OModel(t9)
    for i in range(nq):
        s = steps[i]
        t = float(rng.uniform(0.0, 1.0))
        mo = [G.motion_prepare(P[0][i], P[1][i]), G.motion_prepare(P[2][i], P[3][i])]
        p1, _ = G.motion_pose(mo[0], t)
        p2, _ = G.motion_pose(mo[1], t)
        s.R1[:], s.T1[:], s.R2[:], s.T2[:] = p1[:9], p1[9:], p2[:9], p2[9:]
        s.pose_s[0][:], s.pose_e[0][:] = P[0][i], P[1][i]
        s.pose_s[1][:], s.pose_e[1][:] = P[2][i], P[3][i]
        s.mesh[:] = [0, 1]
        seeds = rng.integers(0, len(t9), 2)
        last = rng.integers(0, len(t9), 2)
        s.seed_tri[:] = seeds
        s.last_tri[:] = last
        s.numCA = int(rng.choice([0, 1, 2, 3, 6, 10]))
        s.mint_prev = float(rng.choice([1.0, 0.02, 0.004]))
        s.UpboundTOC = 1.0 - t
        om1 = O.OMotion(P[0][i][:9].reshape(3, 3), P[0][i][9:], P[1][i][:9].reshape(3, 3), P[1][i][9:])
        om2 = O.OMotion(P[2][i][:9].reshape(3, 3), P[2][i][9:], P[3][i][:9].reshape(3, 3), P[3][i][9:])
        m1.last_tri, m2.last_tri = int(last[0]), int(last[1])
        r = O.toc_step(m1, m2, om1, om2, p1[:9].reshape(3, 3), p1[9:], p2[:9].reshape(3, 3), p2[9:],
                       seeds, s.numCA, s.mint_prev, s.UpboundTOC)
        ref.append((r.distance, r.mint, r.abs_err, r.rel_err, tuple(r.p1), tuple(r.p2),
                    r.num_bv_tests, r.num_tri_tests, m1.last_tri, m2.last_tri,
                    np.array(r.R).ravel().copy(), np.array(r.T).copy()))
    print("ccd_step: %d steps" % nq, flush=True)
    got = G.ccd_step(meshes, steps)
    for i in range(nq):
        g = got[i]
        e = ref[i]
        assert (g.distance, g.mint, g.abs_err, g.rel_err) == e[:4], i
        assert tuple(g.p1) == e[4] and tuple(g.p2) == e[5], i
        assert (g.bv_tests, g.tri_tests) == e[6:8], i
        assert (g.last_tri[0], g.last_tri[1]) == e[8:10], i
        assert np.array_equal(np.array(g.R), e[10]) and np.array_equal(np.array(g.T), e[11]), i


def test_contact_count_bit_exact(knot):
    t9, meshes = knot
    nq = 96
    rng = np.random.default_rng(99)
    P = W.knot_queries(nq, 4242, collide_frac=1.0)
    poses = np.zeros((nq, 24))
    thr = rng.uniform(0.0, 0.05, nq)
    m1, m2 = O.OModel(t9), O.OModel(t9)
    ref = np.zeros(nq, np.int64)
    for i in range(nq):
        t = float(rng.uniform(0.3, 1.0))
        a, _ = G.motion_pose(G.motion_prepare(P[0][i], P[1][i]), t)
        b, _ = G.motion_pose(G.motion_prepare(P[2][i], P[3][i]), t)
        poses[i, :12], poses[i, 12:] = a, b
        ref[i] = O.query_contact(m1, m2, a[:9].reshape(3, 3), a[9:], b[:9].reshape(3, 3), b[9:],
                                 thr[i])[0]
    print("contact_count: %d queries" % nq, flush=True)
    got = G.contact_count(meshes, poses, np.tile([0, 1], (nq, 1)), thr)
    assert np.array_equal(got, ref)
    assert (ref > 0).any()


def test_contact_list_bit_exact(knot):
    """c2a_contact_list: the records of C2A_QueryContact's cont_l (triangle
    ids, feature types, vertex ids of the features, closest points in each
    model's frame, distance) equal the oracle's, in visiting order; a cap
    smaller than the count keeps the first records and the full count."""
    t9, idx = W.torus_knot()
    t9 = t9.reshape(-1, 9)
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    nq = 64
    rng = np.random.default_rng(7)
    P = W.knot_queries(nq, 4343, collide_frac=1.0)
    poses = np.zeros((nq, 24))
    thr = rng.uniform(0.0, 0.03, nq)
    m1, m2 = O.OModel(t9, idx=idx), O.OModel(t9, idx=idx)
    refs = []
    for i in range(nq):
        t = float(rng.uniform(0.3, 1.0))
        a, _ = G.motion_pose(G.motion_prepare(P[0][i], P[1][i]), t)
        b, _ = G.motion_pose(G.motion_prepare(P[2][i], P[3][i]), t)
        poses[i, :12], poses[i, 12:] = a, b
        refs.append(O.query_contact(m1, m2, a[:9].reshape(3, 3), a[9:], b[:9].reshape(3, 3), b[9:],
                                    thr[i], cap=1 << 15))
    cap = max(1, max(n for n, _, _ in refs))
    print("contact_list: %d queries, cap %d" % (nq, cap), flush=True)
    counts, recs = G.contact_list(meshes, poses, np.tile([0, 1], (nq, 1)), thr, cap)
    assert sum(counts) > 20
    for i, (n, lst, _) in enumerate(refs):
        assert counts[i] == n and len(lst) == n, i
        # the oracle lists them as the reference's push_front list does
        for k, c in enumerate(reversed(lst)):
            r = recs[i, k]
            assert (r["tri"][0], r["tri"][1]) == (c.triA, c.triB), (i, k)
            assert (r["type"][0], r["type"][1]) == (c.typeA, c.typeB), (i, k)
            assert list(r["fid"][0]) == list(c.fidA) and list(r["fid"][1]) == list(c.fidB), (i, k)
            assert np.array_equal(r["pa"], np.array(c.PA)) and np.array_equal(r["pb"], np.array(c.PB))
            assert r["distance"] == c.distance
    small_counts, small = G.contact_list(meshes, poses, np.tile([0, 1], (nq, 1)), thr, 2)
    assert np.array_equal(small_counts, counts)
    for i in range(nq):
        k = min(2, counts[i])
        assert np.array_equal(small[i, :k], recs[i, :k])
