"""The reference's C++ CCD interface (include/c2a/) over the engine.

The host test builds a model and motions through the source-compatible
classes (tests/compat/host_check.cpp) and compares them with the oracle.  The
GPU test runs the reference demo's CCD loop (tests/compat/solve_demo.cpp:
C2A_Solve per query, last_tri carried between queries as in
CCDDemo/mainTorusknot.cpp:312-321) and compares every output with the
oracle's C2A_Solve run in the same order.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from c2a_amd import workloads as W
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
COMPAT = os.path.join(HERE, "compat")


def _binary(name):
    path = os.path.join(COMPAT, name)
    if not os.path.exists(path):
        import __graft_entry__
        __graft_entry__._build_compat_demo()
    return path


def _run(name, payload, tmp_path, *extra):
    inp, out = tmp_path / (name + ".in"), tmp_path / (name + ".out")
    inp.write_bytes(payload)
    subprocess.run([_binary(name), str(inp), str(out)] + list(extra), check=True, timeout=300)
    return out.read_bytes()


def _i32(x):
    return struct.pack("<i", int(x))


def test_model_and_motion_match_oracle(tmp_path):
    tris, _ = W.torus_knot(n_seg=32, n_ring=6)
    t9 = np.ascontiguousarray(tris.reshape(-1, 9))
    n = len(t9)
    P = W.knot_queries(6, 404, collide_frac=0.0)
    motions = np.concatenate([P[0], P[1]], axis=1)   # object 1: start, end
    ts = np.array([0.0, 0.125, 0.5, 0.9, 1.0, 1.5])
    payload = (_i32(n) + t9.tobytes() + _i32(len(motions)) + motions.tobytes() +
               _i32(len(ts)) + ts.tobytes())
    b = _run("host_check", payload, tmp_path)
    off = 0
    (nt,) = struct.unpack_from("<i", b, off); off += 4
    assert nt == n
    got_tris = np.frombuffer(b, np.float64, n * 9, off).reshape(n, 9); off += n * 72
    got_ids = np.frombuffer(b, np.int32, n, off); off += 4 * n
    (last,) = struct.unpack_from("<i", b, off); off += 4
    m = O.OModel(t9, ids=1000 + np.arange(n))
    ref = m.tris()
    assert np.array_equal(got_tris, np.array(ref["p"]).reshape(n, 9))
    assert np.array_equal(got_ids, np.array(ref["id"]))
    assert last == 0                       # last_tri = tris after EndModel
    for k in range(len(motions)):
        a = motions[k]
        om = O.OMotion(a[:9].reshape(3, 3), a[9:12], a[12:21].reshape(3, 3), a[21:24])
        ax = np.frombuffer(b, np.float64, 4, off); off += 32
        assert np.array_equal(ax[:3], om.axis) and ax[3] == om.angVel
        for t in ts:
            R, T, q7 = om.integrate(t)
            rec = np.frombuffer(b, np.float64, 19, off); off += 19 * 8
            assert np.array_equal(rec[:9], R.reshape(-1)), (k, t)
            assert np.array_equal(rec[9:12], T), (k, t)
            assert np.array_equal(rec[12:19], q7), (k, t)
    assert off == len(b)


@pytest.mark.gpu
def test_solve_demo_matches_oracle(tmp_path):
    """Configuration 0: two 8192-triangle knots, queries in order with the
    demo's last_tri protocol, C2A_Solve's 1e-4 tolerances."""
    tris, _ = W.torus_knot()
    t9 = np.ascontiguousarray(tris.reshape(-1, 9))
    cfg = W.CONFIG_KNOT_SOLVE
    nq = cfg["nq"]   # all 1000 queries
    P = [p[:nq] for p in W.knot_queries(cfg["nq"], cfg["seed"], collide_frac=0.0)]
    poses = np.stack(P, axis=1)            # nq x 4 x 12
    payload = (_i32(len(t9)) + t9.tobytes() + _i32(len(t9)) + t9.tobytes() + _i32(nq) +
               np.ascontiguousarray(poses).tobytes())
    print("solve_demo: %d queries" % nq, flush=True)
    b = _run("solve_demo", payload, tmp_path)
    rec = np.dtype([("td", "<f8", 2), ("i5", "<i4", 7), ("tr", "<f8", 24)])
    got = np.frombuffer(b, rec)
    assert len(got) == nq
    m1, m2 = O.OModel(t9), O.OModel(t9)
    for i in range(nq):
        p = poses[i]
        res, (R0, T0), (R1, T1) = O.solve(m1, m2, (p[0, :9].reshape(3, 3), p[0, 9:]),
                                          (p[1, :9].reshape(3, 3), p[1, 9:]),
                                          (p[2, :9].reshape(3, 3), p[2, 9:]),
                                          (p[3, :9].reshape(3, 3), p[3, 9:]))
        g = got[i]
        assert g["td"][0] == res.toc and g["td"][1] == res.distance, i
        assert g["i5"][0] == res.numCA and g["i5"][2] == res.collisionfree, i
        assert g["i5"][3] == m1.m.last_tri and g["i5"][4] == m2.m.last_tri, i
        if not res.collisionfree:
            assert g["i5"][1] == res.num_contact, i
            assert np.array_equal(g["tr"][:9], R0.reshape(-1)) and np.array_equal(g["tr"][9:12], T0), i
            assert np.array_equal(g["tr"][12:21], R1.reshape(-1)) and np.array_equal(g["tr"][21:], T1), i


@pytest.mark.gpu
def test_reference_demo_bunny_matches_oracle(tmp_path):
    """The reference's own demo data: two 69,664-triangle bunnies
    (tri_models/bunny_noholes.tri) moved by the torus-knot animations
    (models/torusknot{1,2}.ani), the demo's 303 C2A_Solve calls in order with
    its last_tri protocol (CCDDemo/mainTorusknot.cpp:216-321)."""
    t9, _ = W.bunny()
    t9 = np.ascontiguousarray(t9)
    P = W.demo_queries()
    nq = len(P[0])
    poses = np.stack(P, axis=1)
    payload = (_i32(len(t9)) + t9.tobytes() + _i32(len(t9)) + t9.tobytes() + _i32(nq) +
               np.ascontiguousarray(poses).tobytes())
    print("solve_demo on the bunny demo: %d queries" % nq, flush=True)
    b = _run("solve_demo", payload, tmp_path)
    rec = np.dtype([("td", "<f8", 2), ("i5", "<i4", 7), ("tr", "<f8", 24)])
    got = np.frombuffer(b, rec)
    assert len(got) == nq
    m1, m2 = O.OModel(t9), O.OModel(t9)
    hits = 0
    for i in range(nq):
        p = poses[i]
        res, (R0, T0), (R1, T1) = O.solve(m1, m2, (p[0, :9].reshape(3, 3), p[0, 9:]),
                                          (p[1, :9].reshape(3, 3), p[1, 9:]),
                                          (p[2, :9].reshape(3, 3), p[2, 9:]),
                                          (p[3, :9].reshape(3, 3), p[3, 9:]))
        g = got[i]
        assert g["td"][0] == res.toc and g["td"][1] == res.distance, i
        assert g["i5"][0] == res.numCA and g["i5"][2] == res.collisionfree, i
        assert g["i5"][3] == m1.m.last_tri and g["i5"][4] == m2.m.last_tri, i
        if not res.collisionfree:
            hits += 1
            assert g["i5"][1] == res.num_contact, i
            assert np.array_equal(g["tr"][:9], R0.reshape(-1)) and np.array_equal(g["tr"][9:12], T0), i
            assert np.array_equal(g["tr"][12:21], R1.reshape(-1)) and np.array_equal(g["tr"][21:], T1), i
    assert hits > 100


@pytest.mark.gpu
def test_distance_sequence_matches_oracle(tmp_path):
    """C2A_Distance through the C++ layer: 64 queries of the knot pair in
    order, the models' last_tri carried between calls, every output equal to
    the oracle's run of the same sequence."""
    tris, _ = W.torus_knot()
    t9 = np.ascontiguousarray(tris.reshape(-1, 9))
    nq, rel, abs_ = 64, 0.05, 0.01
    P = W.knot_queries(nq, 515, collide_frac=0.5)
    poses = np.ascontiguousarray(np.concatenate([P[0], P[2]], axis=1))   # nq x 24
    payload = (_i32(len(t9)) + t9.tobytes() + _i32(len(t9)) + t9.tobytes() +
               np.array([rel, abs_]).tobytes() + _i32(nq) + poses.tobytes())
    print("distance_demo: %d queries" % nq, flush=True)
    b = _run("distance_demo", payload, tmp_path)
    rec = np.dtype([("d7", "<f8", 7), ("i6", "<i4", 6), ("rt", "<f8", 12)])
    got = np.frombuffer(b, rec)
    assert len(got) == nq
    m1, m2 = O.OModel(t9), O.OModel(t9)
    ids = np.array(m1.tris()["id"])
    for i in range(nq):
        p = poses[i]
        R1, T1, R2, T2 = p[:9].reshape(3, 3), p[9:12], p[12:21].reshape(3, 3), p[21:]
        r = O.distance(m1, m2, R1, T1, R2, T2, rel, abs_)
        g = got[i]
        assert g["d7"][0] == r.distance, i
        assert tuple(g["d7"][1:4]) == tuple(r.p1) and tuple(g["d7"][4:]) == tuple(r.p2), i
        assert (g["i6"][0], g["i6"][1]) == (r.bv_tests, r.tri_tests), i
        assert (g["i6"][4], g["i6"][5]) == (r.last_tri[0], r.last_tri[1]), i
        assert (g["i6"][2], g["i6"][3]) == (ids[r.last_tri[0]], ids[r.last_tri[1]]), i
        Rrel = R1.T @ R2
        assert np.allclose(g["rt"][:9], Rrel.reshape(-1), atol=1e-12), i
        assert np.allclose(g["rt"][9:], R1.T @ (T2 - T1), atol=1e-12), i


@pytest.mark.gpu
def test_solve_reused_result_translation_only(tmp_path):
    """C2A_Solve with one C2A_TimeOfContactResult reused across calls on
    translation-only motions: the translation branch leaves the models'
    last_tri alone and moves dres.last_triA/B to its last improving pair
    (C2A.cpp:1413-1414), which seeds the next call."""
    tris, _ = W.torus_knot(n_seg=96, n_ring=10)
    t9 = np.ascontiguousarray(tris.reshape(-1, 9))
    nq = 40
    P = [p.copy() for p in W.knot_queries(nq, 909, collide_frac=0.5)]
    P[1][:, :9] = P[0][:, :9]      # end rotation = start rotation: pure translation
    P[3][:, :9] = P[2][:, :9]
    poses = np.stack(P, axis=1)
    payload = (_i32(len(t9)) + t9.tobytes() + _i32(len(t9)) + t9.tobytes() + _i32(nq) +
               np.ascontiguousarray(poses).tobytes())
    b = _run("solve_demo", payload, tmp_path, "reuse")
    rec = np.dtype([("td", "<f8", 2), ("i5", "<i4", 7), ("tr", "<f8", 24)])
    got = np.frombuffer(b, rec)
    m1, m2 = O.OModel(t9), O.OModel(t9)
    seed = (0, 0)
    for i in range(nq):
        p = poses[i]
        res = O.query_toc(m1, m2, (p[0, :9].reshape(3, 3), p[0, 9:]),
                          (p[1, :9].reshape(3, 3), p[1, 9:]),
                          (p[2, :9].reshape(3, 3), p[2, 9:]),
                          (p[3, :9].reshape(3, 3), p[3, 9:]), 1e-4, 1e-4, seed=seed)
        g = got[i]
        assert res.flags & O.FLAG_TRANSLATION_ONLY, i
        assert g["td"][0] == res.toc, i
        assert (g["i5"][3], g["i5"][4]) == (0, 0), i          # models untouched
        assert (g["i5"][5], g["i5"][6]) == (res.out_last_triA, res.out_last_triB), i
        seed = (res.out_last_triA, res.out_last_triB)


def _static_payload(t1, t2, kind, qf, param, poses):
    return (_i32(len(t1)) + t1.tobytes() + _i32(len(t2)) + t2.tobytes() + _i32(kind) + _i32(qf) +
            np.array([param], np.float64).tobytes() + _i32(len(poses)) +
            np.ascontiguousarray(poses).tobytes())


@pytest.mark.gpu
def test_static_queries_through_compat(tmp_path):
    """C2A_Distance with qsize 6, C2A_Distance_At_ContactSpace, C2A_Tolerance
    (depth and best first) and C2A_Collide's PQP_CollideResult form through
    the C++ layer, the models' last_tri carried between calls: every output
    equal to the oracle's run of the same sequence (caller ids are
    1000 + the AddTri order)."""
    tris, _ = W.torus_knot(n_seg=96, n_ring=10)
    t9 = np.ascontiguousarray(tris.reshape(-1, 9))
    nq = 24
    P = W.knot_queries(nq, 717, collide_frac=0.5)
    poses = np.ascontiguousarray(np.concatenate([P[0], P[2]], axis=1))
    m1, m2 = O.OModel(t9), O.OModel(t9)
    ids = np.array(m1.tris()["id"]) + 1000

    # C2A_Distance, qsize 6 (best first), then the contact-space distance
    for kind, qf, param in [(0, 6, 0.0), (1, 0, 0.05)]:
        print("static_demo kind %d: %d queries" % (kind, nq), flush=True)
        b = _run("static_demo", _static_payload(t9, t9, kind, qf, param, poses), tmp_path)
        got = np.frombuffer(b, np.dtype([("d7", "<f8", 7), ("i6", "<i4", 6)]))
        assert len(got) == nq
        m1.last_tri, m2.last_tri = 0, 0
        for i in range(nq):
            seeds = np.array([[m1.last_tri, m2.last_tri]], np.int32)
            if kind == 0:
                r = O.distance_queue_batch(m1, m2, poses[i:i + 1], qf, seeds=seeds)[0]
                found = True
            else:
                rr, fd = O.distance_cs_batch(m1, m2, poses[i:i + 1], param, seeds=seeds)
                r, found = rr[0], bool(fd[0])
            m1.last_tri, m2.last_tri = r.last_tri[0], r.last_tri[1]
            g = got[i]
            assert g["d7"][0] == r.distance, (kind, i)
            if found:
                assert tuple(g["d7"][1:4]) == tuple(r.p1), (kind, i)
                assert tuple(g["d7"][4:]) == tuple(r.p2), (kind, i)
                assert (g["i6"][2], g["i6"][3]) == (ids[r.last_tri[0]], ids[r.last_tri[1]])
            assert (g["i6"][0], g["i6"][1]) == (r.bv_tests, r.tri_tests), (kind, i)
            assert (g["i6"][4], g["i6"][5]) == (r.last_tri[0], r.last_tri[1]), (kind, i)

    # C2A_Tolerance, depth first and best first
    for qs in (2, 5):
        print("static_demo tolerance qsize %d" % qs, flush=True)
        b = _run("static_demo", _static_payload(t9, t9, 2, qs, 0.02, poses), tmp_path)
        got = np.frombuffer(b, np.dtype([("i3", "<i4", 3), ("d7", "<f8", 7)]))
        ref = O.tolerance_batch(m1, m2, poses, 0.02, qs)
        for i, r in enumerate(ref):
            g = got[i]
            assert tuple(g["i3"]) == (r.closer, r.bv_tests, r.tri_tests), (qs, i)
            if r.closer:
                assert g["d7"][0] == r.distance
                assert tuple(g["d7"][1:4]) == tuple(r.p1) and tuple(g["d7"][4:]) == tuple(r.p2)

    # C2A_Collide, all contacts then the first contact
    for flag in (1, 2):
        print("static_demo collide flag %d" % flag, flush=True)
        b = _run("static_demo", _static_payload(t9, t9, 3, flag, 0.0, poses), tmp_path)
        ref, pairs = O.collide_pairs_batch(m1, m2, poses, flag=flag, cap=1 << 14)
        words = np.frombuffer(b, "<i4")
        k = 0
        hits = 0
        for i, r in enumerate(ref):
            n, bv, tri = words[k:k + 3]
            k += 3
            assert (n, bv, tri) == (r.num_pairs, r.bv_tests, r.tri_tests), (flag, i)
            got_pairs = words[k:k + 2 * n].reshape(-1, 2)
            k += 2 * n
            want = ids[pairs[i, :n]]
            assert np.array_equal(got_pairs, want), (flag, i)
            hits += n > 0
        assert k == len(words) and hits > 0
