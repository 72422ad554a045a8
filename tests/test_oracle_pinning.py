"""Pins the FP64 oracle (oracle/c2a_oracle.c) to mathematics it must satisfy.

The reference ships no tests or golden outputs and its PQP dependency is not
vendored, so each oracle routine is checked against closed forms, brute force
(tests/brute.py, an independent construction) and invariants chosen so that a
dropped term, a wrong sign, a swapped operand or a wrong branch fails.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import brute as B


def rng(seed=0):
    return np.random.default_rng(seed)


# ---------------------------------------------------------------------------
# rectangle distance (C2A_RectDist.h:157-934)
# ---------------------------------------------------------------------------
def test_rect_dist_matches_brute_force():
    r = rng(1)
    worst = 0.0
    for k in range(400):
        a = r.uniform(0.0, 2.0, 2)
        b = r.uniform(0.0, 2.0, 2)
        if k % 10 == 0:
            a[r.integers(2)] = 0.0          # degenerate: segment
        R = B.random_rotation(r)
        if k % 7 == 0:                      # parallel / coplanar configurations
            R = B.rodrigues([0, 0, 1], r.uniform(0, 2 * np.pi))
        T = r.normal(size=3) * r.choice([0.3, 1.0, 3.0])
        d, P, Q, S = O.rect_dist(R, T, a, b)
        ref = B.poly_poly(B.rect_corners(a[0], a[1]), B.rect_corners(b[0], b[1], R, T))
        worst = max(worst, abs(d - ref))
        assert abs(d - ref) <= 1e-9 * (1 + ref), (k, d, ref)
        if ref > 1e-6:
            assert abs(np.linalg.norm(S) - d) <= 1e-9 * (1 + d)
    assert worst < 1e-9


def test_rect_dist_closed_forms():
    I = np.eye(3)
    # parallel stacked rectangles: distance is the z gap
    d, _, _, S = O.rect_dist(I, [0.2, 0.1, 0.7], [1, 1], [0.5, 0.5])
    assert d == pytest.approx(0.7) and S[2] == pytest.approx(0.7)
    # side by side in the plane: gap along x
    d, _, _, _ = O.rect_dist(I, [1.5, 0.0, 0.0], [1, 1], [1, 1])
    assert d == pytest.approx(0.5)
    # corner to corner diagonal gap
    d, _, _, _ = O.rect_dist(I, [2.0, 2.0, 0.0], [1, 1], [1, 1])
    assert d == pytest.approx(math.sqrt(2.0))
    # crossing rectangles intersect
    Rx = B.rodrigues([1, 0, 0], np.pi / 2)
    d, _, _, _ = O.rect_dist(Rx, [0.2, 0.5, -0.5], [1, 1], [0.5, 1.0])
    assert d == 0.0


def test_rect_dist_symmetry():
    r = rng(2)
    for _ in range(200):
        a, b = r.uniform(0.1, 2, 2), r.uniform(0.1, 2, 2)
        R, T = B.random_rotation(r), r.normal(size=3)
        d1 = O.rect_dist(R, T, a, b)[0]
        d2 = O.rect_dist(R.T, -R.T @ T, b, a)[0]
        assert abs(d1 - d2) <= 1e-9 * (1 + d1)


# ---------------------------------------------------------------------------
# triangle distance (C2A.cpp:59-405)
# ---------------------------------------------------------------------------
def test_tri_dist_matches_brute_force():
    r = rng(3)
    for k in range(600):
        S = r.normal(size=(3, 3))
        T = r.normal(size=(3, 3)) + r.normal(size=3) * r.choice([0.2, 1.0, 3.0])
        if k % 25 == 0:
            T = S + np.array([0, 0, 1e-3]) * r.uniform(0, 1)   # nearly parallel copies
        if k % 33 == 0:
            T[2] = T[1]                                         # degenerate triangle
        d, P, Q, feat = O.tri_dist(S, T)
        ref = B.poly_poly(S, T)
        assert abs(d - ref) <= 1e-9 * (1 + ref), (k, d, ref)
        if d > 1e-9:
            assert abs(np.linalg.norm(Q - P) - d) <= 1e-9 * (1 + d)
            # closest points lie on their triangles
            assert B.pt_poly(P, S) <= 1e-9 * (1 + d)
            assert B.pt_poly(Q, T) <= 1e-9 * (1 + d)
            assert feat[0] in (0, 1, 2) and feat[2] in (0, 1, 2)


def test_tri_dist_intersecting_is_zero():
    S = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0]], dtype=float)
    T = np.array([[0.5, 0.5, -1], [0.5, 0.5, 1], [1.5, 0.2, 0]], dtype=float)
    assert O.tri_dist(S, T)[0] == 0.0


def test_tri_dist_vertex_face_closed_form():
    S = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0]], dtype=float)
    T = np.array([[0.3, 0.4, 0.25], [0.3, 0.4, 2.0], [1.0, 1.0, 3.0]], dtype=float)
    d, P, Q, feat = O.tri_dist(S, T)
    assert d == pytest.approx(0.25)
    assert np.allclose(P, [0.3, 0.4, 0.0]) and np.allclose(Q, [0.3, 0.4, 0.25])
    assert feat[0] == 2 and feat[2] == 0     # face of S, vertex of T


# ---------------------------------------------------------------------------
# rotation conventions and sin/cos (LinearMath.h)
# ---------------------------------------------------------------------------
def test_quaternion_matrix_roundtrip():
    r = rng(4)
    for _ in range(300):
        R = B.random_rotation(r)
        q = O.quat_from_matrix(R)
        assert abs(np.linalg.norm(q) - 1) < 1e-12
        assert np.allclose(O.matrix_from_quat(q), R, atol=1e-12)
    # known value: 90 degrees about z -> (0, 0, sin45, cos45)
    q = O.quat_from_matrix(B.rodrigues([0, 0, 1], np.pi / 2))
    assert np.allclose(np.abs(q), [0, 0, math.sqrt(0.5), math.sqrt(0.5)])


def test_sincos_against_libm():
    xs = np.concatenate([np.linspace(0, math.pi, 20001), np.linspace(-20, 20, 20001)])
    ulp = 0
    for x in xs:
        s, c = O.sincos(x)
        ulp = max(ulp, abs(s - math.sin(x)) / math.ulp(max(abs(math.sin(x)), 1e-300)) if abs(math.sin(x)) > 1e-300 else 0,
                  abs(c - math.cos(x)) / math.ulp(max(abs(math.cos(x)), 1e-300)) if abs(math.cos(x)) > 1e-300 else 0)
        assert abs(s - math.sin(x)) <= 4e-16 * max(1.0, abs(x)) and abs(c - math.cos(x)) <= 4e-16 * max(1.0, abs(x))
    # near-integral ulp error on the interpolation range
    assert ulp < 4


# ---------------------------------------------------------------------------
# linear motion (InterpMotion.cpp:228-287, 486-568)
# ---------------------------------------------------------------------------
def test_motion_endpoints_and_axis():
    r = rng(5)
    for _ in range(200):
        R0, R1 = B.random_rotation(r), B.random_rotation(r)
        T0, T1 = r.normal(size=3), r.normal(size=3)
        m = O.OMotion(R0, T0, R1, T1)
        Ra, Ta, _ = m.integrate(0.0)
        assert np.allclose(Ra, R0, atol=1e-13) and np.allclose(Ta, T0, atol=0)
        Rb, Tb, _ = m.integrate(1.0)
        assert np.allclose(Rb, R1, atol=1e-10) and np.allclose(Tb, T1, atol=1e-14)
        # clamp: t > 1 behaves like t = 1
        assert np.array_equal(m.integrate(3.0)[0], Rb)
        # R(t) = R0 * Rot(axis, angVel * t) with the body-frame axis
        for t in (0.25, 0.5, 0.8):
            Rt, Tt, _ = m.integrate(t)
            assert np.allclose(Rt, R0 @ B.rodrigues(m.axis, m.angVel * t), atol=1e-10)
            assert np.allclose(Tt, T0 + t * (T1 - T0), atol=1e-14)
        assert 0 <= m.angVel <= math.pi + 1e-12


def test_motion_pure_translation_axis_default():
    m = O.OMotion(np.eye(3), [0, 0, 0], np.eye(3), [1, 2, 3])
    assert m.angVel == 0.0 and np.array_equal(m.axis, [1.0, 0.0, 0.0])


def test_motion_bounds_closed_form_and_sampling():
    r = rng(6)
    for _ in range(100):
        R1 = B.random_rotation(r)
        T1 = r.normal(size=3)
        m = O.OMotion(np.eye(3), np.zeros(3), R1, T1)   # R0 = I: body axis == world axis
        N = r.normal(size=3)
        n = N / np.linalg.norm(N)
        angR = r.uniform(0.1, 2.0)
        w = angR * np.linalg.norm(np.cross(m.axis, n)) * m.angVel
        v = max(0.0, float(m.cv @ n))
        assert m.bound(angR, N) == pytest.approx(v + w, rel=1e-12)
        # sampled points within angR of the origin never move along n faster
        # than the bound (directional conservativeness for R0 = I)
        pts = r.normal(size=(64, 3))
        pts = pts / np.linalg.norm(pts, axis=1)[:, None] * r.uniform(0, angR, (64, 1))
        bound = m.bound(angR, N)
        for t in (0.1, 0.5, 1.0):
            Rt, Tt, _ = m.integrate(t)
            disp = (pts @ Rt.T + Tt - pts) @ n
            assert disp.max() <= bound * t + 1e-9
    # leaf bound with zero rotation is the approach speed only
    m = O.OMotion(np.eye(3), [0, 0, 0], np.eye(3), [0, 0, -2])
    assert m.bound_leaf(5.0, [0, 0, -1]) == pytest.approx(2.0)
    assert m.bound_leaf(5.0, [0, 0, 1]) == 1e-30


# ---------------------------------------------------------------------------
# Jacobi eigen-solver and hierarchy build (C2A_Build.cpp)
# ---------------------------------------------------------------------------
def test_jacobi_eigen_against_numpy():
    r = rng(7)
    for _ in range(200):
        A = r.normal(size=(3, 3))
        A = A + A.T
        V, d = O.jacobi_eigen(A)
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-12)
        assert np.allclose(A @ V, V * d, atol=1e-10)
        assert np.allclose(np.sort(d), np.linalg.eigvalsh(A), atol=1e-10)


def _world_frames(bv):
    """Accumulate parent-relative (R, Tr) into model-frame transforms."""
    n = len(bv)
    Rw, Tw = [None] * n, [None] * n
    Rw[0], Tw[0] = np.array(bv[0]["R"]), np.array(bv[0]["Tr"])
    stack = [0]
    while stack:
        i = stack.pop()
        fc = bv[i]["first_child"]
        if fc >= 0:
            for c in (fc, fc + 1):
                Rw[c] = Rw[i] @ np.array(bv[c]["R"])
                Tw[c] = Rw[i] @ np.array(bv[c]["Tr"]) + Tw[i]
                stack.append(c)
    return Rw, Tw


@pytest.mark.parametrize("shape", ["ico2", "torus_small", "knot_small"])
def test_bvh_invariants(shape):
    from c2a_amd import workloads as W
    if shape == "ico2":
        tris, idx = W.icosphere(2, 0.1, 3)
    elif shape == "torus_small":
        tris, idx = W.torus(16, 8)
    else:
        tris, idx = W.torus_knot(n_seg=48, n_ring=6)
    n = len(tris)
    m = O.OModel(tris.reshape(-1, 9), idx=idx)
    bv = m.bv_arrays()
    assert m.nbv == 2 * n - 1
    ordered = m.tri_xyz()
    # the build only reorders triangles
    key = lambda a: sorted(map(tuple, np.round(a.reshape(-1, 9), 12)))
    assert key(ordered) == key(tris)
    Rw, Tw = _world_frames(bv)
    leaves = []
    # every vertex under a node lies inside its rectangle-swept sphere
    def tris_under(i):
        fc = bv[i]["first_child"]
        if fc < 0:
            return [-fc - 1]
        return tris_under(fc) + tris_under(fc + 1)
    for i in range(m.nbv):
        ids = tris_under(i)
        if bv[i]["first_child"] < 0:
            leaves.append(ids[0])
        pts = ordered[ids].reshape(-1, 3)
        loc = (pts - Tw[i]) @ Rw[i]
        l0, l1, rr = bv[i]["l"][0], bv[i]["l"][1], bv[i]["r"]
        cx = np.clip(loc[:, 0], 0, l0)
        cy = np.clip(loc[:, 1], 0, l1)
        dist = np.sqrt((loc[:, 0] - cx) ** 2 + (loc[:, 1] - cy) ** 2 + loc[:, 2] ** 2)
        assert dist.max() <= rr + 1e-9 * (1 + rr), (i, dist.max(), rr)
        # angular radius = max vertex norm about the model origin
        assert bv[i]["angR"] == pytest.approx(np.linalg.norm(pts, axis=1).max(), rel=1e-15)
        assert bv[i]["size"] == pytest.approx(math.hypot(l0, l1) + 2 * rr, rel=1e-15)
        assert np.allclose(Rw[i], np.array(bv[i]["R_loc"]), atol=1e-12)
    assert sorted(leaves) == list(range(n))


def test_empty_model_rejected():
    with pytest.raises(ValueError):
        O.OModel(np.zeros((0, 9)))


# ---------------------------------------------------------------------------
# traversal, CA loop and contacts on tiny meshes
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def spheres():
    from c2a_amd import workloads as W
    t1, _ = W.icosphere(1, 0.1, 11)
    t2, _ = W.icosphere(1, 0.1, 12)
    return t1.reshape(-1, 9), t2.reshape(-1, 9)


def test_exact_step_distance_is_all_pairs_minimum(spheres):
    """With the error schedule at zero (numCA > 5), the step's distance is the
    brute-force minimum over all triangle pairs whenever no pair is pruned by
    time (UpboundTOC large)."""
    t1, t2 = spheres
    r = rng(8)
    m1, m2 = O.OModel(t1), O.OModel(t2)
    for k in range(10):
        R1, R2 = B.random_rotation(r), B.random_rotation(r)
        T1, T2 = np.zeros(3), r.normal(size=3) * 1.5 + np.array([2.6, 0, 0])
        # motions with no velocity: mint is irrelevant, distance must be exact
        mo1 = O.OMotion(R1, T1, R1, T1)
        mo2 = O.OMotion(R2, T2, R2, T2)
        res = O.Result()
        res.last_triA, res.last_triB, res.numCA, res.mint, res.UpboundTOC = 0, 0, 6, 1.0, 1e300
        O.lib().c2o_toc_step(O.C.byref(m1.m), O.C.byref(m2.m), O.C.byref(mo1.m),
                             O.C.byref(mo2.m), O._dp(O.f64(R1)), O._dp(O.f64(T1)),
                             O._dp(O.f64(R2)), O._dp(O.f64(T2)), O.C.byref(m1.m),
                             O.C.byref(m2.m), O.C.byref(res))
        ref, _ = B.mesh_min_distance(m1.tri_xyz(), m2.tri_xyz(), R1, T1, R2, T2)
        assert abs(res.distance - ref) <= 1e-9 * (1 + ref), (k, res.distance, ref)
        assert res.num_tri_tests >= 1 and res.num_bv_tests % 2 == 0


def test_ca_loop_conservative_and_flags(spheres):
    """Reported TOC never exceeds the first sampled contact; collision-free
    results only when the sampled motion never touches."""
    t1, t2 = spheres
    m1, m2 = O.OModel(t1), O.OModel(t2)
    r = rng(9)
    seen_hit = seen_free = 0
    for k in range(8):
        R00, R01, R10, R11 = (B.random_rotation(r) for _ in range(4))
        T00 = np.zeros(3)
        T01 = r.normal(size=3) * 0.3
        T10 = np.array([3.0, 0, 0]) + r.normal(size=3) * 0.3
        T11 = np.array([0.6, 0, 0]) + r.normal(size=3) * (0.3 if k % 3 else 2.5)
        res = O.query_toc(m1, m2, (R00, T00), (R01, T01), (R10, T10), (R11, T11), 1e-3, 1e-4)
        mo1, mo2 = O.OMotion(R00, T00, R01, T01), O.OMotion(R10, T10, R11, T11)
        t_hit = B.first_touch_time(m1.tri_xyz(), m2.tri_xyz(), lambda t: mo1.integrate(t)[:2],
                                   lambda t: mo2.integrate(t)[:2], n=100)
        if res.collisionfree:
            seen_free += 1
            assert t_hit is None, (k, t_hit)
            assert res.toc == 0.0
        else:
            seen_hit += 1
            if t_hit is not None:
                assert res.toc <= t_hit + 1e-12, (k, res.toc, t_hit)
            assert res.distance <= 1e-3 or res.flags & O.FLAG_ITER_CAP or res.mint < 1e-4
    assert seen_hit and seen_free


@pytest.fixture()
def rotation_path():
    """Pure translations through the rotation path, as a direct
    C2A_QueryTimeOfContact call takes them (C2A_Solve alone switches to the
    translation-only branch)."""
    O.set_solve_translation(0)
    yield
    O.set_solve_translation(1)


def test_ca_loop_immediate_contact_returns_first_mint(spheres, rotation_path):
    """If the first step already has dist <= tol_d the loop does not run and
    the reported TOC is the first step's mint (C2A.cpp:2072, 2131): two
    spheres 0.5 * tol_d apart (placed by bisection on the exact distance),
    approaching, report toc == mint > 0 after one step."""
    t1, _ = spheres
    m1, m2 = O.OModel(t1), O.OModel(t1)
    I = np.eye(3)

    def gap(x):
        return O.distance(m1, m2, I, np.zeros(3), I, np.array([x, 0, 0]), seeds=(0, 0)).distance
    lo, hi = 1.0, 3.0            # touching at lo, apart at hi
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if gap(mid) < 5e-4:
            lo = mid
        else:
            hi = mid
    x0 = hi
    assert 4e-4 < gap(x0) <= 6e-4
    res = O.query_toc(m1, m2, (I, np.zeros(3)), (I, np.zeros(3)), (I, np.array([x0, 0, 0])),
                      (I, np.array([x0 - 0.5, 0, 0])), 1e-3, 1e-4)
    assert res.numCA == 1 and res.collisionfree == 0
    assert res.distance <= 1e-3
    assert res.mint > 0 and res.toc == res.mint


def test_separating_motion_is_collision_free(spheres, rotation_path):
    t1, t2 = spheres
    m1, m2 = O.OModel(t1), O.OModel(t2)
    I = np.eye(3)
    res = O.query_toc(m1, m2, (I, np.zeros(3)), (I, np.array([-1.0, 0, 0])),
                      (I, np.array([3.0, 0, 0])), (I, np.array([5.0, 0, 0])), 1e-3, 1e-4)
    assert res.collisionfree == 1 and res.toc == 0.0 and res.numCA == 1


def test_head_on_translation_toc_close_to_truth(spheres):
    """Two unit icospheres approaching head-on: TOC within tol_d / v of the
    sampled contact time and never after it."""
    t1, _ = spheres
    m1, m2 = O.OModel(t1), O.OModel(t1)
    I = np.eye(3)
    Rs = B.rodrigues([0, 0, 1], 1e-3)   # a tiny rotation keeps the rotation path
    res = O.query_toc(m1, m2, (I, np.zeros(3)), (Rs, np.zeros(3)),
                      (I, np.array([4.0, 0, 0])), (I, np.array([0.0, 0, 0])), 1e-3, 1e-4)
    mo1, mo2 = O.OMotion(I, np.zeros(3), Rs, np.zeros(3)), O.OMotion(I, [4.0, 0, 0], I, [0, 0, 0])
    lo, hi = 0.0, 1.0
    for _ in range(24):     # bisection on brute-force contact
        mid = 0.5 * (lo + hi)
        R1, T1, _ = mo1.integrate(mid)
        R2, T2, _ = mo2.integrate(mid)
        hit = B.touching(m1.tri_xyz(), m2.tri_xyz(), R1, T1, R2, T2)
        lo, hi = (lo, mid) if hit else (mid, hi)
    assert res.collisionfree == 0
    assert res.toc <= hi + 1e-12
    assert hi - res.toc <= 1e-3 / 4.0 + 1e-4     # approach speed 4


def test_iteration_cap_flag():
    """max_iter bounds the loop and raises the cap flag (C2A.cpp:2079)."""
    from c2a_amd import workloads as W
    tris, _ = W.torus_knot(n_seg=64, n_ring=8)
    m1, m2 = O.OModel(tris.reshape(-1, 9)), O.OModel(tris.reshape(-1, 9))
    P = W.knot_queries(64, 1234, collide_frac=0.5)
    res = O.batch_toc(m1, m2, *P, 1e-6, 1e-9, seeds=np.zeros((64, 2), np.int32), max_iter=3)
    assert (res["flags"] & O.FLAG_ITER_CAP).any()
    assert res["numCA"].max() <= 4


def test_contact_count_matches_brute_force(spheres):
    t1, t2 = spheres
    m1, m2 = O.OModel(t1), O.OModel(t2)
    r = rng(10)
    for k in range(6):
        R1, R2 = B.random_rotation(r), B.random_rotation(r)
        T1, T2 = np.zeros(3), np.array([2.05, 0, 0]) + r.normal(size=3) * 0.05
        thr = [0.02, 0.1, 0.3][k % 3]
        n, recs, _ = O.query_contact(m1, m2, R1, T1, R2, T2, thr, cap=4096)
        ref = B.count_pairs_within(m1.tri_xyz(), m2.tri_xyz(), R1, T1, R2, T2, thr)
        assert n == ref, (k, n, ref)
        for c in recs:
            assert c.distance <= thr


def test_solve_reference_semantics(spheres):
    """C2A_Solve: 1e-4 tolerances, contact count at the TOC pose, last_tri
    maintenance on the models (C2A.cpp:2315-2444)."""
    t1, t2 = spheres
    m1, m2 = O.OModel(t1), O.OModel(t2)
    I = np.eye(3)
    Rz = B.rodrigues([0, 0, 1], 0.3)
    res, (R0o, T0o), (R1o, T1o) = O.solve(m1, m2, (I, np.zeros(3)), (Rz, np.zeros(3)),
                                         (I, np.array([3.0, 0.1, 0])), (I, np.array([0.0, 0.1, 0])))
    assert res.collisionfree == 0 and 0 < res.toc < 1
    assert res.distance <= 1e-4 or res.mint < 1e-4
    assert res.num_contact >= 1
    mo1 = O.OMotion(I, np.zeros(3), Rz, np.zeros(3))
    assert np.allclose(R0o, mo1.integrate(res.toc)[0], atol=1e-15)
    assert (m1.last_tri, m2.last_tri) == (res.out_last_triA, res.out_last_triB)


def test_demo_fixtures_are_the_reference_data():
    """The committed demo fixtures (tests/golden/make_fixtures.py): the
    bunny's counts as the .tri header states them, a closed mesh (every edge
    shared by two triangles), and rotation matrices in every animation frame."""
    from c2a_amd import workloads as W
    t9, f = W.bunny()
    assert t9.shape == (69664, 9) and f.max() == 34833
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert (cnt == 2).all()
    P = W.demo_queries()
    assert P[0].shape == (303, 12)
    for p in P:
        R = p[:, :9].reshape(-1, 3, 3)
        assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=2e-5)
