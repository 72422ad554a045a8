"""C2A_Solve's translation-only branch (both rotations below 1e-8;
C2A.cpp:1362-1521, 1818-1852, 1900-1917, 2032-2049; InterpMotion.cpp:570-743):
the oracle's restatement pinned against the exact first-contact time, the
reference's quirks (object 2's velocity ignored, toc = mint, no CA loop), and
the GPU path through the C ABI against the oracle, bit for bit."""
import numpy as np
import pytest

import c2a_amd as G
from c2a_amd import workloads as W
from oracle import oracle as O


@pytest.fixture(scope="module")
def balls():
    t1, _ = W.icosphere(2, 0.1, 21)
    t2, _ = W.icosphere(2, 0.1, 22)
    return t1.reshape(-1, 9), t2.reshape(-1, 9)


def rot(axis, ang):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def pose(R, T):
    return np.concatenate([np.asarray(R, float).reshape(9), np.asarray(T, float)])


def translation_queries(n, seed, mover=1):
    """Object `mover` translates (fixed rotation) toward the other, which
    rests at the origin; both rotations are constant."""
    rng = np.random.default_rng(seed)
    P = [np.zeros((n, 12)) for _ in range(4)]
    for i in range(n):
        Ra = rot(rng.normal(size=3), rng.uniform(0, np.pi))
        Rb = rot(rng.normal(size=3), rng.uniform(0, np.pi))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        start = d * rng.uniform(2.5, 4.0)
        end = start - d * rng.uniform(2.0, 6.0) + rng.normal(size=3) * 0.3
        moving = (pose(Ra, start), pose(Ra, end))
        still = (pose(Rb, np.zeros(3)), pose(Rb, np.zeros(3)))
        pa, pb = (moving, still) if mover == 1 else (still, moving)
        P[0][i], P[1][i], P[2][i], P[3][i] = pa[0], pa[1], pb[0], pb[1]
    return P


def exact_distance(m1, m2, P, i, t):
    """Closest distance at time t (a CA step in exact mode: numCA > 5)."""
    mo1 = O.OMotion(P[0][i][:9].reshape(3, 3), P[0][i][9:], P[1][i][:9].reshape(3, 3), P[1][i][9:])
    mo2 = O.OMotion(P[2][i][:9].reshape(3, 3), P[2][i][9:], P[3][i][:9].reshape(3, 3), P[3][i][9:])
    R1, T1, _ = mo1.integrate(t)
    R2, T2, _ = mo2.integrate(t)
    return O.toc_step(m1, m2, mo1, mo2, R1, T1, R2, T2, (0, 0), 10, 1.0, 1.0).distance


def first_contact(m1, m2, P, i):
    ts = np.linspace(0.0, 1.0, 401)
    ds = [exact_distance(m1, m2, P, i, t) for t in ts]
    k = next((j for j, d in enumerate(ds) if d <= 0), None)
    if k is None:
        return None
    lo, hi = ts[k - 1], ts[k]
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if exact_distance(m1, m2, P, i, mid) > 0:
            lo = mid
        else:
            hi = mid
    return hi


def test_translation_toc_before_first_contact(balls):
    t1, t2 = balls
    P = translation_queries(12, 5, mover=1)
    m1, m2 = O.OModel(t1), O.OModel(t2)
    r = O.batch_toc(m1, m2, *P, 1e-4, 1e-4, seeds=np.zeros((12, 2), np.int32), nthreads=4)
    assert (r["flags"] & 2).all() and (r["numCA"] == 0).all()
    hits = 0
    for i in range(12):
        tc = first_contact(m1, m2, P, i)
        if tc is None:
            assert r["collisionfree"][i] == 1 and r["toc"][i] >= 1.0
            continue
        hits += 1
        assert r["collisionfree"][i] == 0
        assert r["toc"][i] <= tc
        assert tc - r["toc"][i] < 0.02
        assert exact_distance(m1, m2, P, i, r["toc"][i]) > 0
    assert hits >= 4


def test_translation_ignores_object2_velocity(balls):
    """Only object 1's velocity enters (C2A.cpp:1472-1474): a resting object 1
    hit by a moving object 2 is reported collision-free with toc = mint = 1."""
    t1, t2 = balls
    P = translation_queries(6, 6, mover=2)
    m1, m2 = O.OModel(t1), O.OModel(t2)
    r = O.batch_toc(m1, m2, *P, 1e-4, 1e-4, seeds=np.zeros((6, 2), np.int32), nthreads=4)
    assert (r["collisionfree"] == 1).all() and (r["toc"] == 1.0).all()
    O.set_solve_translation(0)
    try:
        rr = O.batch_toc(m1, m2, *P, 1e-4, 1e-4, seeds=np.zeros((6, 2), np.int32), nthreads=4)
    finally:
        O.set_solve_translation(1)
    assert (rr["collisionfree"] == 0).sum() >= 3   # the rotation path sees the hits


def _both_movers():
    return [np.concatenate([a, b]) for a, b in zip(translation_queries(48, 7, 1),
                                                    translation_queries(16, 8, 2))]


def _same(ref, got):
    for f, g in [("toc", "toc"), ("distance", "distance"), ("mint", "mint"),
                 ("numCA", "numCA"), ("collisionfree", "collisionfree"),
                 ("num_bv_tests", "bv_tests"), ("num_tri_tests", "tri_tests"),
                 ("num_contact", "num_contact"), ("flags", "flags")]:
        assert np.array_equal(ref[f], got[g]), f
    assert np.array_equal(ref["p1"], got["p1"]) and np.array_equal(ref["p2"], got["p2"])
    assert np.array_equal(ref["out_last_triA"], got["last_tri"][:, 0])
    assert np.array_equal(ref["out_last_triB"], got["last_tri"][:, 1])


def test_kernel_logic_translation_branch(balls):
    """The kernel's translation-only query run on the CPU harness."""
    import ctypes as C
    import os
    from tests.test_kernel_logic_cpu import EMU, run_emu
    if not os.path.exists(EMU):
        import __graft_entry__
        __graft_entry__._build_emu()
    emu = C.CDLL(EMU)
    t1, t2 = balls
    P = _both_movers()
    got = run_emu(emu, t1, t2, P, G.default_params(), 5)
    m1, m2 = O.OModel(t1), O.OModel(t2)
    ref = O.batch_toc(m1, m2, *P, 1e-4, 1e-4, seeds=np.zeros((64, 2), np.int32),
                      with_contacts=True, nthreads=4)
    _same(ref, got)
    assert (got["collisionfree"] == 0).sum() >= 10


@pytest.mark.gpu
def test_gpu_translation_branch_bit_exact(balls):
    t1, t2 = balls
    P = _both_movers()
    print("translation-only batch: 64 queries", flush=True)
    meshes = [G.Mesh(t1), G.Mesh(t2)]
    got = G.ccd_batch(meshes, G.make_queries(*P), G.default_params())
    m1, m2 = O.OModel(t1), O.OModel(t2)
    ref = O.batch_toc(m1, m2, *P, 1e-4, 1e-4, seeds=np.zeros((64, 2), np.int32),
                      with_contacts=True, nthreads=8)
    _same(ref, got)
    assert (got["collisionfree"] == 0).sum() >= 10
    # the switch off: the rotation path, as a direct C2A_QueryTimeOfContact
    rot = G.ccd_batch(meshes, G.make_queries(*P), G.default_params(translation=0))
    O.set_solve_translation(0)
    try:
        rref = O.batch_toc(m1, m2, *P, 1e-4, 1e-4, seeds=np.zeros((64, 2), np.int32),
                           with_contacts=True, nthreads=8)
    finally:
        O.set_solve_translation(1)
    _same(rref, rot)
