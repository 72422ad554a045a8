"""Static distance query (C2A_Distance with qsize <= 2, C2A_PQP.cpp:970-1057
and C2ADistanceRecurse :481-620).

The oracle's restatement is pinned against the independent brute-force
minimum over all triangle pairs (tests/brute.py) at zero error bounds, and
against the error-bound envelope the pruning rule implies for nonzero
bounds: a pair is skipped only when d >= D - abs_err and d (1 + rel_err) >=
D, so the reported D satisfies D <= D* + abs_err and D <= D* (1 + rel_err)
for the true minimum D* (and with either bound zero the search is exact).  The kernel's per-query logic is then checked
bit-exact against the oracle on the CPU (host build of the kernel code) and
on the GPU through the C ABI."""
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


@pytest.fixture(scope="module")
def spheres():
    t1, _ = W.icosphere(1, 0.1, 31)
    t2, _ = W.icosphere(1, 0.1, 32)
    return t1.reshape(-1, 9), t2.reshape(-1, 9)


def random_case(rng, sep):
    """Poses of two unit-size meshes with their centres about `sep` apart."""
    R1, R2 = B.random_rotation(rng), B.random_rotation(rng)
    T1 = rng.normal(size=3) * 0.1
    d = rng.normal(size=3)
    T2 = T1 + sep * d / np.linalg.norm(d)
    return R1, T1, R2, T2


def poses24(cases):
    P = np.zeros((len(cases), 24))
    for i, (R1, T1, R2, T2) in enumerate(cases):
        P[i, :9], P[i, 9:12], P[i, 12:21], P[i, 21:] = R1.ravel(), T1, R2.ravel(), T2
    return P


def world(R, T, p):
    return np.asarray(R) @ np.asarray(p) + np.asarray(T)


def oracle_batch(t1, t2, cases, rel, abs_, seeds, overlap_only=False):
    m1, m2 = O.OModel(t1), O.OModel(t2)
    out = []
    for (R1, T1, R2, T2), s in zip(cases, seeds):
        r = O.distance(m1, m2, R1, T1, R2, T2, rel, abs_, seeds=s, overlap_only=overlap_only)
        out.append((r.distance, tuple(r.p1), tuple(r.p2), r.bv_tests, r.tri_tests,
                    r.last_tri[0], r.last_tri[1]))
    return out


def check_same(ref, got):
    for i, e in enumerate(ref):
        g = got[i]
        assert g["distance"] == e[0], i
        assert tuple(g["p1"]) == e[1], i
        assert tuple(g["p2"]) == e[2], i
        assert (g["bv_tests"], g["tri_tests"]) == e[3:5], i
        assert (g["last_tri"][0], g["last_tri"][1]) == e[5:], i
        assert g["flags"] == 0, i


# ---------------------------------------------------------------- oracle pins

@pytest.mark.parametrize("sep", [2.4, 2.1, 3.5])
def test_oracle_exact_distance_is_all_pairs_minimum(spheres, sep):
    t1, t2 = spheres
    rng = np.random.default_rng(int(sep * 10))
    m1, m2 = O.OModel(t1), O.OModel(t2)
    for k in range(6):
        R1, T1, R2, T2 = random_case(rng, sep)
        r = O.distance(m1, m2, R1, T1, R2, T2, 0.0, 0.0)
        want = B.mesh_min_distance(t1, t2, R1, T1, R2, T2)[0]
        assert abs(r.distance - want) <= 1e-10 * max(1.0, want), (k, r.distance, want)
        # p1 (model 1's frame) and p2 (model 2's) realise the distance
        w1, w2 = world(R1, T1, r.p1), world(R2, T2, r.p2)
        assert abs(np.linalg.norm(w1 - w2) - r.distance) <= 1e-9
        # the search visits fewer pairs than the n1 x n2 brute force
        assert 0 < r.tri_tests < len(t1) * len(t2)


def test_oracle_error_bound_envelope(spheres):
    t1, t2 = spheres
    rng = np.random.default_rng(5)
    m1, m2 = O.OModel(t1), O.OModel(t2)
    work = {}
    for k in range(8):
        R1, T1, R2, T2 = random_case(rng, 2.2 + 0.3 * k)
        exact = B.mesh_min_distance(t1, t2, R1, T1, R2, T2)[0]
        for rel, abs_ in [(0.0, 0.0), (0.5, 0.0), (0.0, 0.2), (0.1, 0.05), (3.0, 1e30)]:
            m1.last_tri, m2.last_tri = 0, 0
            r = O.distance(m1, m2, R1, T1, R2, T2, rel, abs_)
            work[rel, abs_] = work.get((rel, abs_), 0) + r.tri_tests
            tol = 1e-10 * max(1.0, exact)
            assert r.distance >= exact - tol
            assert r.distance <= exact + abs_ + tol
            assert r.distance <= exact * (1 + rel) + tol
            w1, w2 = world(R1, T1, r.p1), world(R2, T2, r.p2)
            assert abs(np.linalg.norm(w1 - w2) - r.distance) <= 1e-9
    # a pair is entered when EITHER bound admits it, so one zero bound leaves
    # the exact search; both positive prune
    for (rel, abs_), n in work.items():
        if rel == 0.0 or abs_ == 0.0:
            assert n == work[0.0, 0.0], (rel, abs_)
        else:
            assert n < work[0.0, 0.0], (rel, abs_)


def test_oracle_seed_pair_and_last_tri(spheres):
    """The seed pair gives the first bound; last_tri moves only on a strictly
    closer leaf, so seeding with the closest pair leaves it in place and
    costs no more work than seeding elsewhere."""
    t1, t2 = spheres
    rng = np.random.default_rng(17)
    m1, m2 = O.OModel(t1), O.OModel(t2)
    R1, T1, R2, T2 = random_case(rng, 2.3)
    r0 = O.distance(m1, m2, R1, T1, R2, T2, 0.0, 0.0, seeds=(0, 0))
    best = (r0.last_tri[0], r0.last_tri[1])
    assert (m1.last_tri, m2.last_tri) == best
    r1 = O.distance(m1, m2, R1, T1, R2, T2, 0.0, 0.0, seeds=best)
    assert r1.distance == r0.distance
    assert (r1.last_tri[0], r1.last_tri[1]) == best
    assert r1.tri_tests <= r0.tri_tests
    assert r1.bv_tests <= r0.bv_tests


def test_oracle_overlap_is_zero(spheres):
    t1, t2 = spheres
    m1, m2 = O.OModel(t1), O.OModel(t2)
    I, z = np.eye(3), np.zeros(3)
    r = O.distance(m1, m2, I, z, I, np.array([0.3, 0.0, 0.0]), 0.0, 0.0)
    assert r.distance == 0.0


def test_oracle_collide_form(spheres):
    """C2A_Collide's distance form enters only overlapping RSS pairs: with
    the two spheres far apart nothing is visited and the seed pair's distance
    stands; with them interpenetrating the zero distance is found (every
    ancestor of an intersecting triangle pair overlaps); in between, the
    result is never below the true minimum."""
    t1, t2 = spheres
    rng = np.random.default_rng(23)
    m1, m2 = O.OModel(t1), O.OModel(t2)
    tris1, tris2 = m1.tri_xyz(), m2.tri_xyz()   # build order
    for sep in (6.0, 1.0, 0.3, 2.05):
        R1, T1, R2, T2 = random_case(rng, sep)
        seeds = (int(rng.integers(len(t1))), int(rng.integers(len(t2))))
        r = O.distance(m1, m2, R1, T1, R2, T2, 0.0, 0.0, seeds=seeds, overlap_only=True)
        exact = B.mesh_min_distance(t1, t2, R1, T1, R2, T2)[0]
        assert r.distance >= exact - 1e-10 * max(1.0, exact), sep
        if sep == 6.0:
            A = tris1[seeds[0]].reshape(3, 3) @ R1.T + T1
            Bt = tris2[seeds[1]].reshape(3, 3) @ R2.T + T2
            assert abs(r.distance - B.poly_poly(A, Bt)) <= 1e-10
            assert (r.bv_tests, r.tri_tests) == (0, 0)
            assert (r.last_tri[0], r.last_tri[1]) == seeds
        if exact == 0.0:
            assert r.distance == 0.0, sep


# ------------------------------------------------- kernel logic on the CPU

@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU):
        import __graft_entry__
        __graft_entry__._build_emu()
    return C.CDLL(EMU)


def run_emu(emu, t1, t2, cases, rel, abs_, seeds, overlap_only=0):
    P = poses24(cases)
    s = np.ascontiguousarray(seeds, dtype=np.int32).reshape(-1, 2)
    out = np.zeros(len(cases), dtype=G.DISTANCE_DTYPE)
    a, b = np.ascontiguousarray(t1), np.ascontiguousarray(t2)
    rc = emu.emu_distance(G._dp(a), C.c_int64(len(a)), G._dp(b), C.c_int64(len(b)), G._dp(P),
                          G._ip(s), C.c_double(rel), C.c_double(abs_), C.c_int64(len(cases)),
                          out.ctypes.data_as(C.c_void_p), int(overlap_only))
    assert rc == 0
    return out


@pytest.mark.parametrize("rel,abs_", [(0.0, 0.0), (0.5, 0.02), (0.05, 0.1), (3.0, 1e30)])
def test_kernel_logic_distance_matches_oracle(emu, knot_tris, rel, abs_):
    t9, _ = knot_tris
    rng = np.random.default_rng(int(rel * 10 + 3))
    cases = [random_case(rng, s) for s in rng.uniform(0.0, 1.6, 12)]
    seeds = rng.integers(0, len(t9), (12, 2))
    ref = oracle_batch(t9, t9, cases, rel, abs_, seeds)
    check_same(ref, run_emu(emu, t9, t9, cases, rel, abs_, seeds))


@pytest.mark.parametrize("rel,abs_", [(0.0, 0.0), (0.5, 0.02)])
def test_kernel_logic_collide_form_matches_oracle(emu, knot_tris, rel, abs_):
    t9, _ = knot_tris
    rng = np.random.default_rng(61 + int(rel * 10))
    cases = [random_case(rng, s) for s in rng.uniform(0.0, 2.5, 12)]
    seeds = rng.integers(0, len(t9), (12, 2))
    ref = oracle_batch(t9, t9, cases, rel, abs_, seeds, overlap_only=True)
    check_same(ref, run_emu(emu, t9, t9, cases, rel, abs_, seeds, overlap_only=1))
    assert any(e[3] == 0 for e in ref) and any(e[3] > 0 for e in ref)


def test_kernel_logic_distance_unequal_meshes(emu, spheres, knot_tris):
    t1, _ = spheres
    t9, _ = knot_tris
    rng = np.random.default_rng(44)
    cases = [random_case(rng, s) for s in rng.uniform(0.5, 2.5, 10)]
    seeds = np.stack([rng.integers(0, len(t1), 10), rng.integers(0, len(t9), 10)], 1)
    ref = oracle_batch(t1, t9, cases, 0.0, 0.0, seeds)
    check_same(ref, run_emu(emu, t1, t9, cases, 0.0, 0.0, seeds))


# ---------------------------------------------------------------- the GPU

@pytest.mark.gpu
@pytest.mark.parametrize("rel,abs_", [(0.0, 0.0), (0.1, 0.05), (3.0, 1e30)])
def test_gpu_distance_bit_exact(knot_tris, rel, abs_):
    t9, idx = knot_tris
    nq = 128
    rng = np.random.default_rng(int(rel * 100) + (abs_ > 0))
    cases = [random_case(rng, s) for s in rng.uniform(0.0, 2.0, nq)]
    seeds = rng.integers(0, len(t9), (nq, 2))
    ref = oracle_batch(t9, t9, cases, rel, abs_, seeds)
    print("distance: %d queries rel %g abs %g" % (nq, rel, abs_), flush=True)
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    got = G.distance(meshes, poses24(cases), np.tile([0, 1], (nq, 1)), seeds, rel, abs_)
    check_same(ref, got)
    assert (got["distance"] == 0).any() and (got["distance"] > 0).any()


@pytest.mark.gpu
def test_gpu_collide_form_bit_exact(knot_tris):
    t9, idx = knot_tris
    nq = 128
    rng = np.random.default_rng(71)
    cases = [random_case(rng, s) for s in rng.uniform(0.0, 2.5, nq)]
    seeds = rng.integers(0, len(t9), (nq, 2))
    ref = oracle_batch(t9, t9, cases, 0.0, 0.0, seeds, overlap_only=True)
    print("collide form: %d queries" % nq, flush=True)
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    got = G.distance(meshes, poses24(cases), np.tile([0, 1], (nq, 1)), seeds,
                     overlap_only=True)
    check_same(ref, got)
    assert (got["bv_tests"] == 0).any() and (got["bv_tests"] > 0).any()


@pytest.mark.gpu
def test_gpu_distance_mesh_table(spheres, knot_tris):
    """Several meshes in one table, mixed pairs per query, default seeds."""
    t1, t2 = spheres
    t9, idx = knot_tris
    tab = [t1, t9, t2]
    meshes = [G.Mesh(t1), G.Mesh(t9, vidx=idx), G.Mesh(t2)]
    rng = np.random.default_rng(8)
    nq = 60
    mp = rng.integers(0, 3, (nq, 2))
    cases = [random_case(rng, s) for s in rng.uniform(0.3, 2.5, nq)]
    print("distance: mesh table, %d queries" % nq, flush=True)
    got = G.distance(meshes, poses24(cases), mp)
    for i in range(nq):
        ref = oracle_batch(tab[mp[i, 0]], tab[mp[i, 1]], [cases[i]], 0.0, 0.0, [(0, 0)])
        check_same(ref, got[i:i + 1])
