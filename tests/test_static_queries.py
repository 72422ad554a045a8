"""Static queries on the one-thread-per-query kernel (static_query.h): the
best-first distance search (C2A_Distance, qsize > 2, C2A_PQP.cpp:624-786),
C2A_Distance_At_ContactSpace (:1282-1496), C2A_Tolerance (:1498-1845) and
C2A_Collide's PQP_CollideResult form (:798-968).

Every field is compared for exact equality with the FP64 oracle (pinned in
tests/test_oracle_static.py): on the CPU through the host build of the
kernel's per-query code, on the GPU through the C ABI.  The GPU tests add
the knot pair (8192 triangles each) at the knot workload's start poses,
checked against the oracle and through properties that do not depend on
size: every best-first distance equals the depth-first one at zero error
bounds, tolerance at the distance found says closer, contact-space search
at the distance found finds nothing closer."""
import ctypes as C
import os

import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import workloads as W
from oracle import oracle as O
from tests import brute as B

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu", "libccd_emu.so")

SQ_DISTANCE_QUEUE, SQ_DISTANCE_CS, SQ_TOLERANCE, SQ_COLLIDE = 0, 1, 2, 3
STATIC_OUT = np.dtype([("distance", "f8"), ("p1", "f8", 3), ("p2", "f8", 3),
                       ("bv_tests", "i8"), ("tri_tests", "i8"), ("tri", "i4", 2),
                       ("found", "i4"), ("flags", "i4")])


@pytest.fixture(scope="module")
def small():
    t1, _ = W.icosphere(2, 0.15, 51)
    t2, _ = W.icosphere(1, 0.15, 52)
    t1, t2 = t1.reshape(-1, 9), t2.reshape(-1, 9)
    return t1, t2, O.OModel(t1), O.OModel(t2)


def cases(seed, n, seps=(0.85, 1.0, 1.1, 1.3, 2.0)):
    rng = np.random.default_rng(seed)
    P = np.zeros((n, 24))
    for i in range(n):
        R1, R2 = B.random_rotation(rng), B.random_rotation(rng)
        T1 = rng.normal(size=3) * 0.1
        d = rng.normal(size=3)
        T2 = T1 + seps[i % len(seps)] * d / np.linalg.norm(d)
        P[i, :9], P[i, 9:12], P[i, 12:21], P[i, 21:] = R1.ravel(), T1, R2.ravel(), T2
    return P


def emu():
    L = C.CDLL(EMU)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)
    L.emu_static.argtypes = [dp, C.c_int64, dp, C.c_int64, C.c_int, dp, ip, dp, C.c_double,
                             C.c_double, C.c_int, C.c_int, C.c_int64, C.c_int, ip, C.c_void_p]
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def run_emu(t1, t2, kind, P, seeds=None, params=None, rel=0.0, abs_=0.0, qsize=2, flag=1,
            cap=0):
    nq = len(P)
    out = np.zeros(nq, STATIC_OUT)
    pairs = np.full((nq, max(cap, 1), 2), -1, np.int32)
    sd = None if seeds is None else np.ascontiguousarray(seeds, np.int32)
    pr = None if params is None else np.ascontiguousarray(params, np.float64)
    rc = emu().emu_static(_dp(t1), len(t1), _dp(t2), len(t2), kind, _dp(np.ascontiguousarray(P)),
                          None if sd is None else _ip(sd), None if pr is None else _dp(pr),
                          rel, abs_, qsize, flag, nq, cap, _ip(pairs), out.ctypes.data)
    assert rc == 0
    return out, pairs


def same_distance(ref, got, label):
    for i, r in enumerate(ref):
        g = got[i]
        assert r.distance == g["distance"], (label, i, r.distance, g["distance"])
        assert tuple(r.p1) == tuple(g["p1"]) and tuple(r.p2) == tuple(g["p2"]), (label, i)
        assert (r.bv_tests, r.tri_tests) == (g["bv_tests"], g["tri_tests"]), (label, i)
        last = g["last_tri"] if "last_tri" in g.dtype.names else g["tri"]
        assert tuple(r.last_tri) == tuple(last), (label, i)
        assert g["flags"] == 0


def same_tolerance(ref, got, label):
    for i, r in enumerate(ref):
        g = got[i]
        assert r.closer == g["found"], (label, i)
        assert (r.bv_tests, r.tri_tests) == (g["bv_tests"], g["tri_tests"]), (label, i)
        assert r.distance == g["distance"], (label, i)
        assert tuple(r.p1) == tuple(g["p1"]) and tuple(r.p2) == tuple(g["p2"]), (label, i)
        assert tuple(r.tri) == tuple(g["tri"]), (label, i)
        assert g["flags"] == 0


def same_collide(ref, rpairs, got, gpairs, label):
    for i, r in enumerate(ref):
        g = got[i]
        assert r.num_pairs == g["found"], (label, i)
        assert (r.bv_tests, r.tri_tests) == (g["bv_tests"], g["tri_tests"]), (label, i)
        n = min(r.num_pairs, rpairs.shape[1], gpairs.shape[1])
        assert np.array_equal(rpairs[i, :n], gpairs[i, :n]), (label, i)
        assert g["flags"] == 0


# ---- CPU: the kernel's per-query code built for the host -------------------

@pytest.mark.parametrize("qsize", [3, 5, 64])
def test_emu_distance_queue_bit_exact(small, qsize):
    t1, t2, m1, m2 = small
    P = cases(100 + qsize, 12)
    seeds = np.stack([np.arange(12) % len(t1), (5 * np.arange(12)) % len(t2)], 1).astype(np.int32)
    for rel, abs_ in [(0.0, 0.0), (0.05, 0.01)]:
        ref = O.distance_queue_batch(m1, m2, P, qsize, seeds=seeds, rel_err=rel, abs_err=abs_)
        got, _ = run_emu(t1, t2, SQ_DISTANCE_QUEUE, P, seeds=seeds, rel=rel, abs_=abs_,
                         qsize=qsize)
        same_distance(ref, got, "qsize %d rel %g abs %g" % (qsize, rel, abs_))


def test_emu_distance_contact_space_bit_exact(small):
    t1, t2, m1, m2 = small
    P = cases(7, 15)
    ref0 = O.distance_batch(m1, m2, P)
    base = np.array([r.distance for r in ref0])
    for scale in (0.5, 1.0, 1.5, np.inf):
        up = base * scale if np.isfinite(scale) else np.full(len(P), np.inf)
        ref, found = O.distance_cs_batch(m1, m2, P, up)
        got, _ = run_emu(t1, t2, SQ_DISTANCE_CS, P, params=up)
        assert np.array_equal(found, got["found"]), scale
        same_distance(ref, got, "scale %s" % scale)


@pytest.mark.parametrize("qsize", [2, 3, 9])
def test_emu_tolerance_bit_exact(small, qsize):
    t1, t2, m1, m2 = small
    P = cases(200 + qsize, 15)
    base = np.array([r.distance for r in O.distance_batch(m1, m2, P)])
    for scale in (0.99, 1.01, 3.0):
        tol = base * scale
        ref = O.tolerance_batch(m1, m2, P, tol, qsize)
        got, _ = run_emu(t1, t2, SQ_TOLERANCE, P, params=tol, qsize=qsize)
        same_tolerance(ref, got, "qsize %d scale %g" % (qsize, scale))


@pytest.mark.parametrize("flag", [1, 2])
def test_emu_collide_pairs_bit_exact(small, flag):
    t1, t2, m1, m2 = small
    P = cases(300 + flag, 15, seps=(0.7, 0.85, 0.95, 1.05))
    ref, rp = O.collide_pairs_batch(m1, m2, P, flag=flag, cap=2048)
    got, gp = run_emu(t1, t2, SQ_COLLIDE, P, flag=flag, cap=2048)
    same_collide(ref, rp, got, gp, "flag %d" % flag)
    assert any(r.num_pairs > 0 for r in ref)


# ---- GPU: the C ABI ---------------------------------------------------------

@pytest.fixture(scope="module")
def gpu_small(small):
    t1, t2, m1, m2 = small
    return t1, t2, m1, m2, [G.Mesh(t1), G.Mesh(t2)]


@pytest.mark.gpu
@pytest.mark.parametrize("qsize", [3, 5, 64])
def test_gpu_distance_queue_bit_exact(gpu_small, qsize):
    t1, t2, m1, m2, meshes = gpu_small
    nq = 200
    P = cases(400 + qsize, nq)
    seeds = np.stack([np.arange(nq) % len(t1), (3 * np.arange(nq)) % len(t2)], 1).astype(np.int32)
    mp = np.tile([0, 1], (nq, 1))
    print("distance queue qsize=%d nq=%d" % (qsize, nq), flush=True)
    for rel, abs_ in [(0.0, 0.0), (0.05, 0.01)]:
        ref = O.distance_queue_batch(m1, m2, P, qsize, seeds=seeds, rel_err=rel, abs_err=abs_)
        got = G.distance_queue(meshes, P, mp, qsize, seed_tris=seeds, rel_err=rel, abs_err=abs_)
        same_distance(ref, got, "gpu qsize %d" % qsize)


@pytest.mark.gpu
def test_gpu_distance_contact_space_bit_exact(gpu_small):
    t1, t2, m1, m2, meshes = gpu_small
    nq = 200
    P = cases(501, nq)
    mp = np.tile([0, 1], (nq, 1))
    base = np.array([r.distance for r in O.distance_batch(m1, m2, P)])
    print("contact-space distance nq=%d" % nq, flush=True)
    for scale in (0.5, 1.0, 1.5):
        up = base * scale
        ref, found = O.distance_cs_batch(m1, m2, P, up)
        got = G.distance_contact_space(meshes, P, mp, up)
        assert np.array_equal(found, got["found"])
        same_distance(ref, got, "gpu cs %g" % scale)


@pytest.mark.gpu
@pytest.mark.parametrize("qsize", [2, 7])
def test_gpu_tolerance_bit_exact(gpu_small, qsize):
    t1, t2, m1, m2, meshes = gpu_small
    nq = 200
    P = cases(600 + qsize, nq)
    mp = np.tile([0, 1], (nq, 1))
    base = np.array([r.distance for r in O.distance_batch(m1, m2, P)])
    print("tolerance qsize=%d nq=%d" % (qsize, nq), flush=True)
    for scale in (0.99, 1.01, 3.0):
        tol = base * scale
        ref = O.tolerance_batch(m1, m2, P, tol, qsize)
        got = G.tolerance(meshes, P, mp, tol, qsize)
        same_tolerance(ref, got, "gpu tol %d %g" % (qsize, scale))


@pytest.mark.gpu
@pytest.mark.parametrize("flag", [1, 2])
def test_gpu_collide_pairs_bit_exact(gpu_small, flag):
    t1, t2, m1, m2, meshes = gpu_small
    nq = 200
    P = cases(700 + flag, nq, seps=(0.7, 0.85, 0.95, 1.05))
    mp = np.tile([0, 1], (nq, 1))
    print("collide pairs flag=%d nq=%d" % (flag, nq), flush=True)
    ref, rp = O.collide_pairs_batch(m1, m2, P, flag=flag, cap=512)
    got, gp = G.collide_pairs(meshes, P, mp, flag=flag, cap=512)
    same_collide(ref, rp, got, gp, "gpu flag %d" % flag)
    # the list is cut at cap, the count is not
    got_small, gp_small = G.collide_pairs(meshes, P, mp, flag=flag, cap=2)
    assert np.array_equal(got_small["found"], got["found"])
    assert np.array_equal(gp_small, gp[:, :2])


@pytest.mark.gpu
def test_gpu_static_knot_pair(knot_tris):
    """The knot pair at the knot workload's start poses: against the oracle
    on a sample, and the size-independent relations on all of them."""
    t9, idx = knot_tris
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    o1, o2 = O.OModel(t9), O.OModel(t9)
    P00, _, P10, _ = W.knot_batch_queries(512)
    P = np.concatenate([P00, P10], 1)
    nq = len(P)
    mp = np.tile([0, 1], (nq, 1))
    print("static knot nq=%d" % nq, flush=True)
    dfs = G.distance(meshes, P, mp)
    q16 = G.distance_queue(meshes, P, mp, 16)
    assert np.array_equal(dfs["distance"], q16["distance"])
    tol = G.tolerance(meshes, P, mp, dfs["distance"], 2)
    assert tol["found"].all()
    assert (tol["distance"] <= dfs["distance"]).all()
    cs = G.distance_contact_space(meshes, P, mp, dfs["distance"])
    assert not cs["found"].any()
    hits, pairs = G.collide_pairs(meshes, P, mp, flag=1, cap=64)
    assert ((hits["found"] > 0) == (dfs["distance"] == 0.0)).all()
    k = 48
    ref = O.distance_queue_batch(o1, o2, P[:k], 16)
    same_distance(ref, q16[:k], "knot queue")
    rc, rpairs = O.collide_pairs_batch(o1, o2, P[:k], flag=1, cap=64)
    same_collide(rc, rpairs, hits[:k], pairs[:k], "knot collide")
    rt = O.tolerance_batch(o1, o2, P[:k], dfs["distance"][:k], 2)
    same_tolerance(rt, tol[:k], "knot tolerance")


def test_static_entry_points_reject_bad_input(small):
    """Host-side argument checks run before any device call, so they hold
    without a GPU; an empty batch is a no-op."""
    t1, t2, m1, m2 = small
    meshes = [G.Mesh(t1), G.Mesh(t2)]
    P = cases(1, 2)
    mp = np.tile([0, 1], (2, 1))
    with pytest.raises(G.C2AError, match="flag must be 1"):
        G.collide_pairs(meshes, P, mp, flag=3)
    with pytest.raises(G.C2AError, match="qsize above"):
        G.tolerance(meshes, P, mp, 0.1, qsize=(1 << 20) + 1)
    with pytest.raises(G.C2AError, match="error bounds"):
        G.distance_queue(meshes, P, mp, 8, rel_err=-1.0)
    assert len(G.tolerance(meshes, P[:0], mp[:0], 0.1)) == 0
