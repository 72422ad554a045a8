"""Pinning of the oracle's static queries beyond C2A_Distance's depth-first
form: the best-first distance search (qsize > 2, C2A_PQP.cpp:624-786),
C2A_Distance_At_ContactSpace (:1282-1496), C2A_Tolerance (:1498-1845),
PQP's TriContact and the PQP_CollideResult form of C2A_Collide (:798-968).

Each is checked against independent brute force over all triangle pairs
(tests/brute.py) and against properties the search fixes:
  - best-first distance at zero error bounds is the true minimum, for every
    queue size, including sizes small enough that nested queues run; with
    error bounds it stays inside the envelope the stop rule implies;
  - contact-space distance finds a pair iff some pair is closer than the
    bound, and the pair it reports is closer;
  - tolerance reports closer iff the minimum distance is within tolerance;
  - TriContact agrees with the triangle distance being zero away from
    touching, and on constructed piercing / coplanar / separated pairs;
  - the collide pair list is exactly the set of intersecting triangle pairs
    (every such pair lies under overlapping volumes), each listed once, and
    the first-contact flag stops after one pair.
The order in which PQP's BVTQ releases tests of equal distance is not pinned
(PQP is not vendored); the counters of the best-first searches rest on the
binary-heap restatement."""
import numpy as np
import pytest

from c2a_amd import workloads as W
from oracle import oracle as O
from tests import brute as B


@pytest.fixture(scope="module")
def meshes():
    t1, _ = W.icosphere(1, 0.15, 41)
    t2, _ = W.icosphere(1, 0.15, 42)
    t1, t2 = t1.reshape(-1, 9), t2.reshape(-1, 9)
    return t1, t2, O.OModel(t1), O.OModel(t2)


def cases(seed, n, seps):
    rng = np.random.default_rng(seed)
    P = np.zeros((n, 24))
    for i in range(n):
        R1, R2 = B.random_rotation(rng), B.random_rotation(rng)
        T1 = rng.normal(size=3) * 0.1
        d = rng.normal(size=3)
        T2 = T1 + seps[i % len(seps)] * d / np.linalg.norm(d)
        P[i, :9], P[i, 9:12], P[i, 12:21], P[i, 21:] = R1.ravel(), T1, R2.ravel(), T2
    return P


def unpack(p):
    return p[:9].reshape(3, 3), p[9:12], p[12:21].reshape(3, 3), p[21:]


def world_tris(tris, R, T):
    return tris.reshape(-1, 3, 3) @ np.asarray(R).T + np.asarray(T)


def brute_pairs_within(t1, t2, p, thr):
    """{(i, j): distance(i, j) <= thr} in caller ids, with the distances."""
    R1, T1, R2, T2 = unpack(p)
    A, Bt = world_tris(t1, R1, T1), world_tris(t2, R2, T2)
    ca, cb = A.mean(1), Bt.mean(1)
    ra = np.linalg.norm(A - ca[:, None], axis=2).max(1)
    rb = np.linalg.norm(Bt - cb[:, None], axis=2).max(1)
    dc = np.linalg.norm(ca[:, None] - cb[None], axis=2) - ra[:, None] - rb[None]
    out = {}
    for i, j in zip(*np.nonzero(dc <= thr)):
        d = B.poly_poly(A[i], Bt[j])
        if d <= thr:
            out[(int(i), int(j))] = d
    return out


def brute_min(t1, t2, p):
    R1, T1, R2, T2 = unpack(p)
    return B.mesh_min_distance(t1, t2, R1, T1, R2, T2)[0]


SEPS = [0.9, 1.0, 1.1, 1.25, 1.6, 2.4]


@pytest.mark.parametrize("qsize", [3, 4, 6, 16, 200])
def test_distance_queue_exact_minimum(meshes, qsize):
    t1, t2, m1, m2 = meshes
    P = cases(5 + qsize, 10, SEPS)
    out = O.distance_queue_batch(m1, m2, P, qsize)
    depth_first = O.distance_batch(m1, m2, P)
    for i in range(len(P)):
        ref = brute_min(t1, t2, P[i])
        assert abs(out[i].distance - ref) <= 1e-12 * max(1.0, ref), (qsize, i, out[i].distance, ref)
        # both searches are exact at zero bounds: the same minimum, bit for bit
        assert out[i].distance == depth_first[i].distance, (qsize, i)


def test_distance_queue_nested_queues_run(meshes):
    """qsize 3 fills after two expansions, so the search nests queues; the
    BV test count then exceeds what one queue of that size could hold."""
    t1, t2, m1, m2 = meshes
    P = cases(77, 4, [1.05])
    small = O.distance_queue_batch(m1, m2, P, 3)
    big = O.distance_queue_batch(m1, m2, P, 100000)
    for s, b in zip(small, big):
        assert s.bv_tests > 2 and b.bv_tests > 2
        assert s.distance == b.distance


@pytest.mark.parametrize("rel,abs_", [(0.1, 0.0), (0.0, 0.02), (0.05, 0.01)])
def test_distance_queue_error_envelope(meshes, rel, abs_):
    t1, t2, m1, m2 = meshes
    P = cases(11, 10, SEPS)
    out = O.distance_queue_batch(m1, m2, P, 8, rel_err=rel, abs_err=abs_)
    for i in range(len(P)):
        ref = brute_min(t1, t2, P[i])
        d = out[i].distance
        assert d >= ref - 1e-12
        assert d <= max(ref + abs_, ref * (1 + rel)) + 1e-12, (i, d, ref)


def test_distance_queue_closest_points(meshes):
    t1, t2, m1, m2 = meshes
    P = cases(12, 8, SEPS[2:])
    for r, p in zip(O.distance_queue_batch(m1, m2, P, 5), P):
        if r.distance == 0.0:   # intersecting: the points are not a closest pair
            continue
        R1, T1, R2, T2 = unpack(p)
        w1 = R1 @ np.array(r.p1) + T1
        w2 = R2 @ np.array(r.p2) + T2
        assert abs(np.linalg.norm(w1 - w2) - r.distance) <= 1e-9


def test_distance_contact_space(meshes):
    t1, t2, m1, m2 = meshes
    P = cases(21, 12, SEPS)
    mins = np.array([brute_min(t1, t2, p) for p in P])
    for scale in (0.9, 1.1, 3.0):
        upper = mins * scale
        res, found = O.distance_cs_batch(m1, m2, P, upper)
        for i in range(len(P)):
            assert found[i] == (mins[i] < upper[i]), (scale, i)
            if found[i]:
                assert mins[i] - 1e-12 <= res[i].distance < upper[i]
                R1, T1, R2, T2 = unpack(P[i])
                a, b = res[i].last_tri
                A = world_tris(m1.tri_xyz(), R1, T1)[a]
                Bt = world_tris(m2.tri_xyz(), R2, T2)[b]
                assert abs(B.poly_poly(A, Bt) - res[i].distance) <= 1e-9
            else:
                assert res[i].distance == upper[i]
                assert res[i].tri_tests == 0 or mins[i] >= upper[i]


def test_distance_contact_space_first_leaf(meshes):
    """With an infinite bound the first leaf pair reached improves on it, and
    every later call returns at once: exactly one triangle test."""
    t1, t2, m1, m2 = meshes
    P = cases(22, 6, SEPS)
    res, found = O.distance_cs_batch(m1, m2, P, np.inf)
    assert found.all()
    assert all(r.tri_tests == 1 for r in res)


@pytest.mark.parametrize("qsize", [2, 3, 7, 1000])
def test_tolerance_against_brute_force(meshes, qsize):
    t1, t2, m1, m2 = meshes
    P = cases(31 + qsize, 12, SEPS)
    mins = np.array([brute_min(t1, t2, p) for p in P])
    for scale in (0.999, 1.001, 2.0):
        tol = mins * scale
        out = O.tolerance_batch(m1, m2, P, tol, qsize)
        for i in range(len(P)):
            assert out[i].closer == (mins[i] <= tol[i]), (qsize, scale, i)
            if out[i].closer:
                assert out[i].distance <= tol[i]
                R1, T1, R2, T2 = unpack(P[i])
                a, b = out[i].tri
                A = world_tris(m1.tri_xyz(), R1, T1)[a]
                Bt = world_tris(m2.tri_xyz(), R2, T2)[b]
                assert abs(B.poly_poly(A, Bt) - out[i].distance) <= 1e-9
    neg = O.tolerance_batch(m1, m2, P[:2], -1.0, qsize)
    assert all(r.tolerance == 0.0 for r in neg)


def test_tolerance_trivial_reject(meshes):
    """Far apart: the root pair's distance rejects before any test is counted."""
    t1, t2, m1, m2 = meshes
    P = cases(40, 3, [10.0])
    for r in O.tolerance_batch(m1, m2, P, 0.01, 2) + O.tolerance_batch(m1, m2, P, 0.01, 9):
        assert r.closer == 0 and r.bv_tests == 0 and r.tri_tests == 0


def test_tri_contact_constructed():
    A = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    pierce = np.array([[0.2, 0.2, -1], [0.3, 0.2, 1], [0.2, 0.3, 1]], np.float64)
    coplanar_overlap = np.array([[0.2, 0.2, 0], [1.2, 0.2, 0], [0.2, 1.2, 0]], np.float64)
    coplanar_apart = np.array([[1, 1, 0], [2, 1, 0], [1, 2, 0]], np.float64)
    parallel = A + [0, 0, 1e-3]
    shared_vertex = np.array([[1, 0, 0], [2, 0, 1], [2, 1, -1]], np.float64)
    above = np.array([[0.2, 0.2, 0.5], [0.3, 0.2, 1], [0.2, 0.3, 1]], np.float64)
    assert O.tri_contact(A, pierce) == 1
    assert O.tri_contact(A, coplanar_overlap) == 1
    assert O.tri_contact(A, coplanar_apart) == 0
    assert O.tri_contact(A, parallel) == 0
    assert O.tri_contact(A, shared_vertex) == 1
    assert O.tri_contact(A, above) == 0
    # symmetric in the two triangles and in vertex order
    assert O.tri_contact(pierce, A) == 1
    assert O.tri_contact(A[[1, 2, 0]], pierce[[2, 0, 1]]) == 1


def test_tri_contact_random_agrees_with_distance():
    rng = np.random.default_rng(5)
    agree = 0
    for _ in range(3000):
        A = rng.normal(size=(3, 3))
        Bt = rng.normal(size=(3, 3)) * rng.uniform(0.3, 1.5) + rng.normal(size=3) * 0.8
        d = B.poly_poly(A, Bt)
        if 0 < d < 1e-9:
            continue
        assert O.tri_contact(A, Bt) == (d == 0.0), d
        agree += 1
    assert agree > 2900


@pytest.mark.parametrize("sep", [0.8, 0.95, 1.05])
def test_collide_pairs_are_the_intersecting_pairs(meshes, sep):
    t1, t2, m1, m2 = meshes
    P = cases(50 + int(sep * 100), 6, [sep])
    res, pairs = O.collide_pairs_batch(m1, m2, P, flag=1, cap=4096)
    ids1 = m1.tris()["id"]
    ids2 = m2.tris()["id"]
    any_contact = False
    for i in range(len(P)):
        n = res[i].num_pairs
        got = [(int(ids1[a]), int(ids2[b])) for a, b in pairs[i, :n]]
        assert len(set(got)) == len(got)
        want = set(brute_pairs_within(t1, t2, P[i], 0.0))
        assert set(got) == want, (sep, i, len(got), len(want))
        assert res[i].tri_tests >= n
        any_contact |= n > 0
        first, fp = O.collide_pairs_batch(m1, m2, P[i:i + 1], flag=2, cap=4)
        assert first[0].num_pairs == min(n, 1)
        if n:
            assert (int(fp[0, 0, 0]), int(fp[0, 0, 1])) == tuple(pairs[i, 0])
            assert first[0].bv_tests <= res[i].bv_tests
    assert any_contact


def test_collide_pairs_single_triangles():
    """Known answer: two one-triangle models are one root test and, when
    the volumes overlap, one triangle test."""
    A = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float64)
    m1, m2 = O.OModel(A), O.OModel(A)
    I = np.eye(3)
    P = np.zeros((2, 24))
    P[:, :9] = I.ravel()
    P[:, 12:21] = B.rodrigues([1, 0, 0], np.pi / 2).ravel()
    P[0, 21:] = [0.2, 0.2, -0.5]   # pierces the first triangle
    P[1, 21:] = [0.2, 0.2, 5.0]    # far above it
    res, pairs = O.collide_pairs_batch(m1, m2, P, flag=1, cap=4)
    assert (res[0].bv_tests, res[0].tri_tests, res[0].num_pairs) == (1, 1, 1)
    assert tuple(pairs[0, 0]) == (0, 0)
    assert (res[1].bv_tests, res[1].tri_tests, res[1].num_pairs) == (1, 0, 0)
